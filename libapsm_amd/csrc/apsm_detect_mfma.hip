// Dispatch of the matrix-core detect launchers to the translation unit of each precision.
#include "apsm_detect_mfma.h"

namespace apsm {

namespace {

hipError_t dispatch( int prec, int nmax, int umax, bool batch, hipStream_t s, const float* rx, int64_t M, const float* Z,
                     const float* Zn, const float* C, int64_t ldC, const float* theta, int64_t G, int U, float wL, float wG,
                     float g2, float* est, int N, const DetDesc* d, const WorkItem* items, int nItems, int grid, int shape )
{
    if ( !mfmaShapeBuilt( shape, prec, nmax, umax ) )
        return hipErrorInvalidValue;
    switch ( prec )
    {
    case PREC_F32: return mfmaEntryF32( nmax, umax, batch, s, rx, M, Z, Zn, C, ldC, theta, G, U, wL, wG, g2, est, N, d, items, nItems, grid, shape );
    case PREC_BF16: return mfmaEntryBF16( nmax, umax, batch, s, rx, M, Z, Zn, C, ldC, theta, G, U, wL, wG, g2, est, N, d, items, nItems, grid, shape );
    case PREC_F16: return mfmaEntryF16( nmax, umax, batch, s, rx, M, Z, Zn, C, ldC, theta, G, U, wL, wG, g2, est, N, d, items, nItems, grid, shape );
    default: return hipErrorInvalidValue;
    }
}

} // namespace

hipError_t launchDetectMfma( int prec, int nmax, int umax, hipStream_t s, const float* rx, int64_t M, const float* Z,
                             const float* Zn, const float* C, int64_t ldC, const float* theta, int64_t G, int U, float wL,
                             float wG, float g2, float* est, int N )
{
    return dispatch( prec, nmax, umax, false, s, rx, M, Z, Zn, C, ldC, theta, G, U, wL, wG, g2, est, N, nullptr, nullptr, 0, 0, 0 );
}

hipError_t launchBatchDetectMfma( int prec, int nmax, int umax, hipStream_t s, const DetDesc* d, const WorkItem* items,
                                  int nItems, int grid, int shape )
{
    return dispatch( prec, nmax, umax, true, s, nullptr, 0, nullptr, nullptr, nullptr, 0, nullptr, 0, 0, 0.f, 0.f, 0.f, nullptr, 0,
                     d, items, nItems, grid, shape );
}

} // namespace apsm
