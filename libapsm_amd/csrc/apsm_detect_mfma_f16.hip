// Matrix-core detect kernels with f16 cross-term inputs, every antenna and user bound.
#include "apsm_detect_mfma.h"

namespace apsm {

APSM_MFMA_ENTRY( mfmaEntryF16 )
{
    return launchMfmaPrec<PREC_F16>( nmax, umax, batch, s, rx, M, Z, Zn, C, ldC, theta, G, U, wL, wG, g2, est, N, d, items, nItems, grid,
                                    shape );
}

} // namespace apsm
