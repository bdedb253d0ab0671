// Detect with the cross terms on the matrix cores.
//
// For data samples x (columns) and dictionary entries z (rows) the two real cross terms are
// the GEMMs  P^T = Zp X^T  and  S^T = Zs X^T  with K = 2*NMAX (zero padded):
//   Zp row = [Re z | Im z],  Zs row = [Im z | -Re z],  X row = [Re x | Im x],
// so p = a_x.a_z and s = a_x.b_z. One wave computes a 32 entries x 32 samples block of each
// with v_mfma_f32_32x32x2_f32 (PREC_F32) or v_mfma_f32_32x32x16_{bf16,f16} (inputs rounded
// to 16 bits, FP32 accumulation). In the result layout lane l holds sample l % 32 and the 16
// entries 8*(v/4) + 4*(l/32) + v%4 (v = 0..15), so the rest of the pair work stays in the
// lane: squared distances from the FP32 norms |x|^2 + |z|^2 (the norm correction is FP32 for
// every precision), three exponentials, and the contraction with every user's coefficients.
// The two lane halves hold disjoint entries of the same sample and are added at the end.
//
// Workgroup: 4 waves x 32 samples; the dictionary streams through LDS in tiles of 64 entries
// (32 at 64 antennas).
#pragma once

#include "apsm_detect_launch.h"

namespace apsm {


typedef float mf32x16 __attribute__( ( ext_vector_type( 16 ) ) );
typedef __bf16 mbf16x8 __attribute__( ( ext_vector_type( 8 ) ) );
typedef _Float16 mf16x8 __attribute__( ( ext_vector_type( 8 ) ) );

template <int PREC>
struct MfmaIn;
template <>
struct MfmaIn<PREC_BF16>
{
    typedef __bf16 T;
    typedef mbf16x8 V;
    static __device__ __forceinline__ mf32x16 mma( V a, V b, mf32x16 c ) { return __builtin_amdgcn_mfma_f32_32x32x16_bf16( a, b, c, 0, 0, 0 ); }
};
template <>
struct MfmaIn<PREC_F16>
{
    typedef _Float16 T;
    typedef mf16x8 V;
    static __device__ __forceinline__ mf32x16 mma( V a, V b, mf32x16 c ) { return __builtin_amdgcn_mfma_f32_32x32x16_f16( a, b, c, 0, 0, 0 ); }
};

template <int NMAX, int UMAX, int PREC>
__device__ __forceinline__ void detectTileMfma( const float* __restrict__ rx, int64_t M, const float* __restrict__ Z,
                                                const float* __restrict__ Zn, const float* __restrict__ C, int64_t ldC,
                                                const float* __restrict__ theta, int64_t G, int U, float wL, float wG, float g2,
                                                float* __restrict__ est, int N, const int64_t tile )
{
    constexpr int K = 2 * NMAX;
    constexpr int kMfmaTile = NMAX >= 64 ? 32 : 64; // dictionary entries per LDS tile (<= 64 KB of LDS)
    // 16-bit inputs: K padded to a multiple of 16 (one MFMA step), rows padded by 8 elements
    // (16 bytes) against LDS bank conflicts of the 16-byte reads
    constexpr int K16 = K < 16 ? 16 : K;
    constexpr int RowH = K16 + 8;
    constexpr int kTileF = PREC == PREC_F32 ? 2 * K * kMfmaTile : 0; // f32: [2][K][tile] (transposed)
    constexpr int kTileH = PREC == PREC_F32 ? 0 : 2 * kMfmaTile * RowH; // 16-bit: [2][tile][RowH]
    __shared__ float sZf[ kTileF > 0 ? kTileF : 1 ];
    __shared__ __attribute__( ( aligned( 16 ) ) ) uint16_t sZh[ kTileH > 0 ? kTileH : 1 ];
    __shared__ float sZn[ kMfmaTile ];
    __shared__ float2 sC[ UMAX * kMfmaTile ];

    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int col = lane & 31, h = lane >> 5;
    const int64_t m = tile * kMfmaSamples + wave * 32 + col; // this lane's sample
    const bool valid = m < M;
    const int L = 2 * N;

    // This lane's half of the sample in B-operand order: component k of X row = [Re x | Im x]
    // with k = 2t + h (FP32 steps of 2) or k = 16t + 8h + i (16-bit steps of 16). The norm is
    // summed in FP32 over both halves; the linear term reloads x at the end.
    constexpr int NB = PREC == PREC_F32 ? K / 2 : K16 / 2;
    typedef typename MfmaIn<PREC == PREC_F32 ? PREC_BF16 : PREC>::V VH; // 16-bit B operand (unused for FP32)
    typedef typename MfmaIn<PREC == PREC_F32 ? PREC_BF16 : PREC>::T TH;
    float xb[ PREC == PREC_F32 ? NB : 1 ];
    VH xv[ PREC == PREC_F32 ? 1 : NB / 8 ]; // 16-bit inputs are packed right away (half the registers)
    float xn = 0.f;
#pragma unroll
    for ( int j = 0; j < NB; j++ )
    {
        const int k = PREC == PREC_F32 ? 2 * j + h : 16 * ( j / 8 ) + 8 * h + ( j % 8 );
        const int ant = k < NMAX ? k : k - NMAX;
        float v = 0.f;
        if ( valid && k < K && ant < N )
            v = rx[ ( (int64_t)ant * M + m ) * 2 + ( k < NMAX ? 0 : 1 ) ];
        if constexpr ( PREC == PREC_F32 )
            xb[ j ] = v;
        else
            xv[ j / 8 ][ j % 8 ] = (TH)v;
        xn += v * v;
    }
    xn += __shfl_xor( xn, 32 );
    // exponent arguments with the scale folded in: -g2 * (|x|^2 + |z|^2 - 2p) = q + k2 * p
    const float xq = -g2 * xn, k2 = 2.f * g2;

    float fa[ UMAX ], fb[ UMAX ];
#pragma unroll
    for ( int u = 0; u < UMAX; u++ )
        fa[ u ] = fb[ u ] = 0.f;

    for ( int64_t j0 = 0; j0 < G; j0 += kMfmaTile )
    {
        const int nt = ( G - j0 ) < kMfmaTile ? (int)( G - j0 ) : kMfmaTile;
        __syncthreads();
        // tile: Zp and Zs rows of the entries (zeros past the dictionary end and past N)
        if constexpr ( PREC == PREC_F32 )
        {
#pragma unroll 2
            for ( int i = tid; i < K * kMfmaTile; i += kMfmaThreads )
            {
                const int k = i / kMfmaTile, e = i % kMfmaTile; // consecutive threads: consecutive entries
                const int n = k < NMAX ? k : k - NMAX;
                float zr = 0.f, zi = 0.f;
                if ( e < nt && n < N )
                {
                    zr = Z[ ( j0 + e ) * L + n ];
                    zi = Z[ ( j0 + e ) * L + N + n ];
                }
                sZf[ k * kMfmaTile + e ] = k < NMAX ? zr : zi; // Zp = [zr | zi]
                sZf[ K * kMfmaTile + k * kMfmaTile + e ] = k < NMAX ? zi : -zr; // Zs = [zi | -zr]
            }
        }
        else
        {
            typedef typename MfmaIn<PREC>::T T;
#pragma unroll 2
            for ( int i = tid; i < kMfmaTile * K16; i += kMfmaThreads )
            {
                const int e = i / K16, k = i % K16;
                const int n = k < NMAX ? k : k - NMAX;
                float zr = 0.f, zi = 0.f;
                if ( e < nt && n < N && k < K )
                {
                    zr = Z[ ( j0 + e ) * L + n ];
                    zi = Z[ ( j0 + e ) * L + N + n ];
                }
                const T vp = (T)( k < NMAX ? zr : zi ), vs = (T)( k < NMAX ? zi : -zr );
                sZh[ e * RowH + k ] = __builtin_bit_cast( uint16_t, vp );
                sZh[ kMfmaTile * RowH + e * RowH + k ] = __builtin_bit_cast( uint16_t, vs );
            }
        }
        for ( int i = tid; i < kMfmaTile; i += kMfmaThreads )
            sZn[ i ] = i < nt ? -g2 * Zn[ j0 + i ] : 0.f;
        for ( int i = tid; i < UMAX * kMfmaTile; i += kMfmaThreads )
        {
            const int u = i / kMfmaTile, e = i % kMfmaTile;
            sC[ i ] = ( u < U && e < nt ) ? make_float2( C[ u * ldC + 2 * ( j0 + e ) ], C[ u * ldC + 2 * ( j0 + e ) + 1 ] )
                                          : make_float2( 0.f, 0.f );
        }
        __syncthreads();

#pragma unroll 1
        for ( int cb = 0; cb < nt; cb += 32 )
        {
            mf32x16 accP = {}, accS = {};
            if constexpr ( PREC == PREC_F32 )
            {
#pragma unroll
                for ( int t = 0; t < K / 2; t++ )
                {
                    const int k = 2 * t + h;
                    const float ap = sZf[ k * kMfmaTile + cb + col ];
                    const float as = sZf[ K * kMfmaTile + k * kMfmaTile + cb + col ];
                    accP = __builtin_amdgcn_mfma_f32_32x32x2f32( ap, xb[ t ], accP, 0, 0, 0 );
                    accS = __builtin_amdgcn_mfma_f32_32x32x2f32( as, xb[ t ], accS, 0, 0, 0 );
                    if ( ( t & 7 ) == 7 )
                        __builtin_amdgcn_sched_barrier( 0 ); // bound the operand loads in flight (registers)
                }
            }
            else
            {
                typedef typename MfmaIn<PREC>::V V;
#pragma unroll
                for ( int t = 0; t < K16 / 16; t++ )
                {
                    const int k0 = 16 * t + 8 * h;
                    const V b = xv[ t ];
                    const V ap = *reinterpret_cast<const V*>( &sZh[ ( cb + col ) * RowH + k0 ] );
                    const V as = *reinterpret_cast<const V*>( &sZh[ kMfmaTile * RowH + ( cb + col ) * RowH + k0 ] );
                    accP = MfmaIn<PREC>::mma( ap, b, accP );
                    accS = MfmaIn<PREC>::mma( as, b, accS );
                    if ( ( t & 1 ) == 1 )
                        __builtin_amdgcn_sched_barrier( 0 );
                }
            }
#pragma unroll
            for ( int v = 0; v < 16; v++ )
            {
                const int e = cb + 8 * ( v / 4 ) + 4 * h + ( v % 4 ); // entry within the tile
                const float q = xq + sZn[ e ];
                const float p = accP[ v ], s = accS[ v ];
                // exp2 of a clamped non-positive argument (a rounded distance below zero counts as 0)
                const float E0 = __builtin_amdgcn_exp2f( fminf( fmaf( k2, p, q ), 0.f ) );
                const float E1 = __builtin_amdgcn_exp2f( fminf( fmaf( k2, s, q ), 0.f ) );
                const float E2 = __builtin_amdgcn_exp2f( fminf( fmaf( -k2, s, q ), 0.f ) );
#pragma unroll
                for ( int u = 0; u < UMAX; u++ )
                {
                    const float2 c = sC[ u * kMfmaTile + e ];
                    fa[ u ] = fmaf( c.x, E0, fmaf( c.y, E1, fa[ u ] ) );
                    fb[ u ] = fmaf( c.x, E2, fmaf( c.y, E0, fb[ u ] ) );
                }
                // bound the coefficient loads in flight; measured: helps up to 32 antennas, at 64 the
                // scheduler then serialises the loads ahead of the math and spills
                if ( NMAX <= 32 )
                    __builtin_amdgcn_sched_barrier( 0 );
            }
        }
    }

    // the two lane halves saw disjoint entries of the same sample
#pragma unroll
    for ( int u = 0; u < UMAX; u++ )
    {
        fa[ u ] += __shfl_xor( fa[ u ], 32 );
        fb[ u ] += __shfl_xor( fb[ u ], 32 );
    }
    if ( !valid || h != 0 )
        return;
#pragma unroll
    for ( int u = 0; u < UMAX; u++ )
    {
        if ( u < U )
        {
            const float* th = theta + u * L;
            float la = 0.f, lb = 0.f;
            for ( int n = 0; n < N; n++ )
            {
                const float2 x = reinterpret_cast<const float2*>( rx )[ (int64_t)n * M + m ];
                la += th[ n ] * x.x + th[ N + n ] * x.y;
                lb += th[ n ] * x.y - th[ N + n ] * x.x;
            }
            reinterpret_cast<float2*>( est )[ (int64_t)u * M + m ] = make_float2( wL * la + wG * fa[ u ], wL * lb + wG * fb[ u ] );
        }
    }
}

template <int NMAX, int UMAX, int PREC>
__global__ void __launch_bounds__( kMfmaThreads ) apsm_detect_mfma( const float* __restrict__ rx, int64_t M, const float* __restrict__ Z,
                                                                   const float* __restrict__ Zn, const float* __restrict__ C, int64_t ldC,
                                                                   const float* __restrict__ theta, int64_t G, int U, float wL, float wG,
                                                                   float g2, float* __restrict__ est, int N )
{
    detectTileMfma<NMAX, UMAX, PREC>( rx, M, Z, Zn, C, ldC, theta, G, U, wL, wG, g2, est, N, blockIdx.x );
}

// Persistent batched detect over a work list (all items of one launch share NMAX and UMAX).
template <int NMAX, int UMAX, int PREC>
__global__ void __launch_bounds__( kMfmaThreads ) batch_detect_mfma( const DetDesc* __restrict__ d, const WorkItem* __restrict__ items,
                                                                    int nItems )
{
    for ( int w = blockIdx.x; w < nItems; w += gridDim.x )
    {
        const WorkItem it = items[ w ];
        const DetDesc& x = d[ it.desc ];
        detectTileMfma<NMAX, UMAX, PREC>( x.rx, x.M, x.Z, x.Zn, x.C, x.ldC, x.theta, x.G, x.U, x.wL, x.wG, x.g2, x.est, x.N, it.tile );
    }
}

// Launchers for one precision over every (NMAX, UMAX); each precision's translation unit
// instantiates them.
template <int PREC, int NMAX, int UMAX>
hipError_t launchMfmaNU( hipStream_t s, const float* rx, int64_t M, const float* Z, const float* Zn, const float* C, int64_t ldC,
                         const float* theta, int64_t G, int U, float wL, float wG, float g2, float* est, int N )
{
    const int64_t grid = ( M + kMfmaSamples - 1 ) / kMfmaSamples;
    hipLaunchKernelGGL( ( apsm_detect_mfma<NMAX, UMAX, PREC> ), dim3( (unsigned)grid ), dim3( kMfmaThreads ), 0, s, rx, M, Z, Zn, C, ldC,
                        theta, G, U, wL, wG, g2, est, N );
    return hipGetLastError();
}

template <int PREC, int NMAX, int UMAX>
hipError_t launchBatchMfmaNU( hipStream_t s, const DetDesc* d, const WorkItem* items, int nItems, int grid )
{
    hipLaunchKernelGGL( ( batch_detect_mfma<NMAX, UMAX, PREC> ), dim3( grid ), dim3( kMfmaThreads ), 0, s, d, items, nItems );
    return hipGetLastError();
}

template <int PREC, int NMAX>
struct MfmaTable
{
    template <int UMAX>
    static hipError_t one( bool batch, hipStream_t s, const float* rx, int64_t M, const float* Z, const float* Zn, const float* C,
                           int64_t ldC, const float* theta, int64_t G, int U, float wL, float wG, float g2, float* est, int N,
                           const DetDesc* d, const WorkItem* items, int nItems, int grid )
    {
        return batch ? launchBatchMfmaNU<PREC, NMAX, UMAX>( s, d, items, nItems, grid )
                     : launchMfmaNU<PREC, NMAX, UMAX>( s, rx, M, Z, Zn, C, ldC, theta, G, U, wL, wG, g2, est, N );
    }
    static hipError_t byU( int umax, bool batch, hipStream_t s, const float* rx, int64_t M, const float* Z, const float* Zn,
                           const float* C, int64_t ldC, const float* theta, int64_t G, int U, float wL, float wG, float g2, float* est,
                           int N, const DetDesc* d, const WorkItem* items, int nItems, int grid )
    {
        switch ( umax )
        {
        case 1: return one<1>( batch, s, rx, M, Z, Zn, C, ldC, theta, G, U, wL, wG, g2, est, N, d, items, nItems, grid );
        case 2: return one<2>( batch, s, rx, M, Z, Zn, C, ldC, theta, G, U, wL, wG, g2, est, N, d, items, nItems, grid );
        case 4: return one<4>( batch, s, rx, M, Z, Zn, C, ldC, theta, G, U, wL, wG, g2, est, N, d, items, nItems, grid );
        case 8: return one<8>( batch, s, rx, M, Z, Zn, C, ldC, theta, G, U, wL, wG, g2, est, N, d, items, nItems, grid );
        case 12: return one<12>( batch, s, rx, M, Z, Zn, C, ldC, theta, G, U, wL, wG, g2, est, N, d, items, nItems, grid );
        case 16: return one<16>( batch, s, rx, M, Z, Zn, C, ldC, theta, G, U, wL, wG, g2, est, N, d, items, nItems, grid );
        default: return hipErrorInvalidValue;
        }
    }
};

template <int PREC>
hipError_t launchMfmaPrec( int nmax, int umax, bool batch, hipStream_t s, const float* rx, int64_t M, const float* Z, const float* Zn,
                           const float* C, int64_t ldC, const float* theta, int64_t G, int U, float wL, float wG, float g2, float* est,
                           int N, const DetDesc* d, const WorkItem* items, int nItems, int grid )
{
    switch ( nmax )
    {
    case 4: return MfmaTable<PREC, 4>::byU( umax, batch, s, rx, M, Z, Zn, C, ldC, theta, G, U, wL, wG, g2, est, N, d, items, nItems, grid );
    case 8: return MfmaTable<PREC, 8>::byU( umax, batch, s, rx, M, Z, Zn, C, ldC, theta, G, U, wL, wG, g2, est, N, d, items, nItems, grid );
    case 16: return MfmaTable<PREC, 16>::byU( umax, batch, s, rx, M, Z, Zn, C, ldC, theta, G, U, wL, wG, g2, est, N, d, items, nItems, grid );
    case 32: return MfmaTable<PREC, 32>::byU( umax, batch, s, rx, M, Z, Zn, C, ldC, theta, G, U, wL, wG, g2, est, N, d, items, nItems, grid );
    case 64: return MfmaTable<PREC, 64>::byU( umax, batch, s, rx, M, Z, Zn, C, ldC, theta, G, U, wL, wG, g2, est, N, d, items, nItems, grid );
    default: return hipErrorInvalidValue;
    }
}

// Entries of each precision's translation unit (defined there).
#define APSM_MFMA_ENTRY( name )                                                                                                   \
    hipError_t name( int nmax, int umax, bool batch, hipStream_t s, const float* rx, int64_t M, const float* Z, const float* Zn,   \
                     const float* C, int64_t ldC, const float* theta, int64_t G, int U, float wL, float wG, float g2, float* est,  \
                     int N, const DetDesc* d, const WorkItem* items, int nItems, int grid )
APSM_MFMA_ENTRY( mfmaEntryF32 );
APSM_MFMA_ENTRY( mfmaEntryBF16 );
APSM_MFMA_ENTRY( mfmaEntryF16 );

} // namespace apsm
