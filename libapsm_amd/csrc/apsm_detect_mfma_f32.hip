// Matrix-core detect kernels with f32 cross-term inputs, every antenna and user bound.
#include "apsm_detect_mfma.h"

namespace apsm {

APSM_MFMA_ENTRY( mfmaEntryF32 )
{
    return launchMfmaPrec<PREC_F32>( nmax, umax, batch, s, rx, M, Z, Zn, C, ldC, theta, G, U, wL, wG, g2, est, N, d, items, nItems, grid );
}

} // namespace apsm
