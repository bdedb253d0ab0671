// Matrix-core detect kernels with f32 cross-term inputs, every antenna and user bound.
#include "apsm_detect_mfma.h"
#include "apsm_entry_mfma.h"
#include "apsm_gram_mfma.h"

namespace apsm {

APSM_MFMA_ENTRY( mfmaEntryF32 )
{
    return launchMfmaPrec<PREC_F32>( nmax, umax, batch, s, rx, M, Z, Zn, C, ldC, theta, G, U, wL, wG, g2, est, N, d, items, nItems, grid );
}

hipError_t launchEntryMfma( int nmax, int umax, hipStream_t s, const float* Z, const float* Zn, const float* C, float* part,
                            const EntryArgs& a, int64_t jBeg, int nChunks, int chunk0, int count )
{
    if ( count <= 0 )
        return hipSuccess;
    return launchEntryAnyN( nmax, umax, false, s, Z, Zn, C, part, a, jBeg, nChunks, nullptr, chunk0, count );
}

hipError_t launchEntryCombine( hipStream_t s, const float* Z, const float* thF, const float* part, float* P, const EntryArgs& a,
                               int nChunks )
{
    const int n = a.U * a.nNewC;
    hipLaunchKernelGGL( apsm_entry_combine, dim3( ( n + 255 ) / 256 ), dim3( 256 ), 0, s, Z, thF, part, P, a, nChunks );
    return hipGetLastError();
}

hipError_t launchBatchEntryMfma( int nmax, int umax, hipStream_t s, const EntryMfmaDesc* d, int n, int gridX )
{
    return launchEntryAnyN( nmax, umax, true, s, nullptr, nullptr, nullptr, nullptr, EntryArgs {}, 0, 0, d, n, gridX );
}

hipError_t launchBatchEntryCombine( hipStream_t s, const EntryMfmaDesc* d, int n, int gridX )
{
    hipLaunchKernelGGL( batch_entry_combine, dim3( gridX, n ), dim3( 256 ), 0, s, d );
    return hipGetLastError();
}

hipError_t launchGramMfma( int nmax, hipStream_t s, const float* Z, const float* Zn, float* Gact, const GramArgs& a )
{
    if ( a.nA <= 0 )
        return hipSuccess;
    return launchGramAnyN( nmax, false, s, Z, Zn, Gact, a, nullptr, 0, 0 );
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

hipError_t launchBatchGramMfma( int nmax, hipStream_t s, const GramDesc* d, int n, int gridX )
{
    return launchGramAnyN( nmax, true, s, nullptr, nullptr, nullptr, GramArgs {}, d, n, gridX );
}

} // namespace apsm
