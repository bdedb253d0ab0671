// Matrix-core detect kernels with f32 cross-term inputs, every antenna and user bound.
#include "apsm_detect_mfma.h"
#include "apsm_entry_mfma.h"

namespace apsm {

APSM_MFMA_ENTRY( mfmaEntryF32 )
{
    return launchMfmaPrec<PREC_F32>( nmax, umax, batch, s, rx, M, Z, Zn, C, ldC, theta, G, U, wL, wG, g2, est, N, d, items, nItems, grid,
                                    shape );
}

hipError_t launchEntryMfma( int nmax, int umax, hipStream_t s, const float* Z, const float* Zn, const float* C, float* part,
                            const EntryArgs& a, int64_t jBeg, int nChunks, int chunk0, int count )
{
    if ( count <= 0 )
        return hipSuccess;
    return launchEntryAnyN( nmax, umax, s, Z, Zn, C, part, a, jBeg, nChunks, chunk0, count );
}

hipError_t launchEntryCombine( hipStream_t s, const float* Z, const float* thF, const float* part, float* P, const EntryArgs& a,
                               int nChunks )
{
    const int n = a.U * a.nNewC;
    hipLaunchKernelGGL( apsm_entry_combine, dim3( ( n + 255 ) / 256 ), dim3( 256 ), 0, s, Z, thF, part, P, a, nChunks );
    return hipGetLastError();
}

hipError_t launchBatchEntryCombine( hipStream_t s, const EntryMfmaDesc* d, int n, int gridX )
{
    hipLaunchKernelGGL( batch_entry_combine, dim3( gridX, n ), dim3( 256 ), 0, s, d );
    return hipGetLastError();
}

hipError_t launchBlockPre( int nmax, int umax, hipStream_t s, const PreDesc& d )
{
    const int grid = d.entryWg + d.gramWg + d.evWg;
    if ( grid <= 0 )
        return hipSuccess;
    return launchPreAnyN( nmax, umax, false, s, &d, 1, grid );
}

hipError_t launchBatchBlockPre( int nmax, int umax, hipStream_t s, const PreDesc* d, int n, int gridX )
{
    if ( gridX <= 0 || n <= 0 )
        return hipSuccess;
    return launchPreAnyN( nmax, umax, true, s, d, n, gridX );
}

} // namespace apsm
