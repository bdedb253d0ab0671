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

// WAVES: waves per workgroup (32 samples each); ETILE: dictionary entries per LDS tile (0: 64,
// or 32 at 64 antennas).
template <int NMAX, int UMAX, int PREC, int WAVES = 4, int ETILE = 0>
__device__ __forceinline__ void detectTileMfma( const float* __restrict__ rx, int64_t M, const float* __restrict__ Z,
                                                const float* __restrict__ Zn, const float* __restrict__ C, int64_t ldC,
                                                const float* __restrict__ theta, int64_t G, int U, float wL, float wG, float g2,
                                                float* __restrict__ est, int N, const int64_t tile )
{
    constexpr int K = 2 * NMAX;
    constexpr int kMfmaTile = ETILE ? ETILE : ( NMAX >= 64 ? 32 : 64 ); // dictionary entries per LDS tile
    constexpr int kMfmaThreads = 64 * WAVES; // (shadows the single-handle shape)
    constexpr int kMfmaSamples = 32 * WAVES;
    static_assert( kMfmaTile <= kMfmaThreads && kMfmaTile % 32 == 0, "one entry norm per thread, 32-entry MFMA blocks" );
    // 16-bit inputs: K padded to a multiple of 16 (one MFMA step), rows padded by 8 elements
    // (16 bytes) against LDS bank conflicts of the 16-byte reads
    constexpr int K16 = K < 16 ? 16 : K;
    constexpr int RowH = K16 + 8;
    // Only Zp = [Re z | Im z] is staged; Zs = [Im z | -Re z] is the same row shifted by NMAX
    // with the sign of its second half flipped. Two tile buffers: the next tile's loads are in
    // flight (in registers) while the current tile is used, so a tile costs one barrier.
    constexpr int kTileF = PREC == PREC_F32 ? K * kMfmaTile : 0; // f32: [K][tile] (transposed)
    constexpr int kTileH = PREC == PREC_F32 ? 0 : kMfmaTile * RowH; // 16-bit: [tile][RowH]
    constexpr int KL = PREC == PREC_F32 ? K : K16; // staged components per entry
    constexpr int kLoads = ( KL * kMfmaTile + kMfmaThreads - 1 ) / kMfmaThreads; // per thread and tile
    constexpr int kCLoads = ( UMAX * kMfmaTile + kMfmaThreads - 1 ) / kMfmaThreads;
    __shared__ float sZf[ 2 ][ kTileF > 0 ? kTileF : 1 ];
    __shared__ __attribute__( ( aligned( 16 ) ) ) uint16_t sZh[ 2 ][ kTileH > 0 ? kTileH : 1 ];
    __shared__ float sZn[ 2 ][ kMfmaTile ];
    __shared__ float2 sC[ 2 ][ UMAX * kMfmaTile ];

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

    // tile loads into registers: element r of this thread is component k, entry e with
    // r * kMfmaThreads + tid = (f32) k * tile + e, (16-bit) e * K16 + k (zeros past the
    // dictionary end, past N and, 16-bit, past K)
    float zreg[ kLoads ];
    float zn;
    float2 creg[ kCLoads ];
    auto loadTile = [ & ]( int64_t j0, int nt ) {
#pragma unroll
        for ( int r = 0; r < kLoads; r++ )
        {
            const int i = r * kMfmaThreads + tid;
            const int k = PREC == PREC_F32 ? i / kMfmaTile : i % K16;
            const int e = PREC == PREC_F32 ? i % kMfmaTile : i / K16;
            const int n = k < NMAX ? k : k - NMAX;
            float v = 0.f;
            if ( i < KL * kMfmaTile && e < nt && k < K && n < N )
                v = Z[ ( j0 + e ) * L + ( k < NMAX ? n : N + n ) ];
            zreg[ r ] = v;
        }
        zn = tid < nt ? -g2 * Zn[ j0 + tid ] : 0.f;
#pragma unroll
        for ( int r = 0; r < kCLoads; r++ )
        {
            const int i = r * kMfmaThreads + tid;
            const int u = i / kMfmaTile, e = i % kMfmaTile;
            creg[ r ] = ( i < UMAX * kMfmaTile && u < U && e < nt )
                ? make_float2( C[ u * ldC + 2 * ( j0 + e ) ], C[ u * ldC + 2 * ( j0 + e ) + 1 ] )
                : make_float2( 0.f, 0.f );
        }
    };
    auto storeTile = [ & ]( int buf ) {
#pragma unroll
        for ( int r = 0; r < kLoads; r++ )
        {
            const int i = r * kMfmaThreads + tid;
            if ( i < KL * kMfmaTile )
            {
                if constexpr ( PREC == PREC_F32 )
                    sZf[ buf ][ i ] = zreg[ r ];
                else
                {
                    typedef typename MfmaIn<PREC>::T T;
                    sZh[ buf ][ ( i / K16 ) * RowH + i % K16 ] = __builtin_bit_cast( uint16_t, (T)zreg[ r ] );
                }
            }
        }
        if ( tid < kMfmaTile )
            sZn[ buf ][ tid ] = zn;
#pragma unroll
        for ( int r = 0; r < kCLoads; r++ )
        {
            const int i = r * kMfmaThreads + tid;
            if ( i < UMAX * kMfmaTile )
                sC[ buf ][ i ] = creg[ r ];
        }
    };

    const int nTiles = (int)( ( G + kMfmaTile - 1 ) / kMfmaTile );
    if ( nTiles > 0 )
    {
        loadTile( 0, G < kMfmaTile ? (int)G : kMfmaTile );
        storeTile( 0 );
    }
    __syncthreads();
#pragma unroll 1
    for ( int tt = 0; tt < nTiles; tt++ )
    {
        const int buf = tt & 1;
        const int64_t j0 = (int64_t)tt * kMfmaTile;
        const int nt = ( G - j0 ) < kMfmaTile ? (int)( G - j0 ) : kMfmaTile;
        if ( tt + 1 < nTiles )
        {
            const int64_t jn = j0 + kMfmaTile;
            loadTile( jn, ( G - jn ) < kMfmaTile ? (int)( G - jn ) : kMfmaTile );
        }
        const float* zf = sZf[ buf ];
        const uint16_t* zh = sZh[ buf ];
        const float* zq = sZn[ buf ];
        const float2* cc = sC[ buf ];

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
                    const float ap = zf[ k * kMfmaTile + cb + col ];
                    // Zs component k: Im z_k (k < NMAX) or -Re z_(k-NMAX); k < NMAX for t < NMAX/2
                    const float as = t < NMAX / 2 ? zf[ ( k + NMAX ) * kMfmaTile + cb + col ] : -zf[ ( k - NMAX ) * kMfmaTile + cb + col ];
                    accP = __builtin_amdgcn_mfma_f32_32x32x2f32( ap, xb[ t ], accP, 0, 0, 0 );
                    accS = __builtin_amdgcn_mfma_f32_32x32x2f32( as, xb[ t ], accS, 0, 0, 0 );
                    if ( ( t & 7 ) == 7 )
                        __builtin_amdgcn_sched_barrier( 0 ); // bound the operand loads in flight (registers)
                }
            }
            else
            {
                typedef typename MfmaIn<PREC>::V V;
                typedef unsigned int u32x4 __attribute__( ( ext_vector_type( 4 ) ) );
#pragma unroll
                for ( int t = 0; t < K16 / 16; t++ )
                {
                    const int k0 = 16 * t + 8 * h;
                    const V b = xv[ t ];
                    const uint16_t* row = zh + ( cb + col ) * RowH;
                    const V ap = *reinterpret_cast<const V*>( row + k0 );
                    // Zs: the row shifted by NMAX, second half negated (components past K meet
                    // zeros of x, so whatever the shifted read holds there does not count)
                    u32x4 bits;
                    if constexpr ( NMAX % 8 == 0 )
                    {
                        // the lane's 8 components lie in one half: k0 < NMAX gives Im z, else -Re z
                        const bool lowHalf = k0 < NMAX;
                        bits = *reinterpret_cast<const u32x4*>( row + ( lowHalf ? k0 + NMAX : k0 - NMAX ) ); // 16-byte aligned
                        bits ^= lowHalf ? 0u : 0x80008000u;
                    }
                    else
                    {
                        // NMAX = 4: components 0..3 are Im z (row + 4), 4..7 are -Re z (row + 0);
                        // the lane with k0 = 8 is past K
                        typedef unsigned int u32x2 __attribute__( ( ext_vector_type( 2 ) ) );
                        const u32x2 lo = *reinterpret_cast<const u32x2*>( row + k0 + 4 );
                        const u32x2 hi = *reinterpret_cast<const u32x2*>( row + k0 ) ^ 0x80008000u;
                        bits = u32x4 { lo.x, lo.y, hi.x, hi.y };
                    }
                    const V as = __builtin_bit_cast( V, bits );
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
                const float q = xq + zq[ e ];
                const float p = accP[ v ], s = accS[ v ];
                // exp2 of a clamped non-positive argument (a rounded distance below zero counts as 0)
                const float E0 = __builtin_amdgcn_exp2f( fminf( fmaf( k2, p, q ), 0.f ) );
                const float E1 = __builtin_amdgcn_exp2f( fminf( fmaf( k2, s, q ), 0.f ) );
                const float E2 = __builtin_amdgcn_exp2f( fminf( fmaf( -k2, s, q ), 0.f ) );
#pragma unroll
                for ( int u = 0; u < UMAX; u++ )
                {
                    const float2 c = cc[ u * kMfmaTile + e ];
                    fa[ u ] = fmaf( c.x, E0, fmaf( c.y, E1, fa[ u ] ) );
                    fb[ u ] = fmaf( c.x, E2, fmaf( c.y, E0, fb[ u ] ) );
                }
                // bound the coefficient loads in flight; measured: helps up to 32 antennas, at 64 the
                // scheduler then serialises the loads ahead of the math and spills
                if ( NMAX <= 32 )
                    __builtin_amdgcn_sched_barrier( 0 );
            }
        }
        if ( tt + 1 < nTiles )
            storeTile( buf ^ 1 ); // the other buffer was released by the previous iteration's barrier
        __syncthreads();
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

// Persistent batched detect over a work list (all items of one launch share NMAX and UMAX),
// workgroups of WAVES waves, tiles of 32 * WAVES samples.
template <int NMAX, int UMAX, int PREC, int WAVES = 4, int ETILE = 0>
__global__ void __launch_bounds__( 64 * WAVES ) batch_detect_mfma( const DetDesc* __restrict__ d, const WorkItem* __restrict__ items,
                                                                   int nItems )
{
    for ( int w = blockIdx.x; w < nItems; w += gridDim.x )
    {
        const WorkItem it = items[ w ];
        const DetDesc& x = d[ it.desc ];
        detectTileMfma<NMAX, UMAX, PREC, WAVES, ETILE>( x.rx, x.M, x.Z, x.Zn, x.C, x.ldC, x.theta, x.G, x.U, x.wL, x.wG, x.g2, x.est,
                                                         x.N, it.tile );
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

template <int PREC, int NMAX, int UMAX, int WAVES = 4, int ETILE = 0>
hipError_t launchBatchMfmaNU( hipStream_t s, const DetDesc* d, const WorkItem* items, int nItems, int grid )
{
    hipLaunchKernelGGL( ( batch_detect_mfma<NMAX, UMAX, PREC, WAVES, ETILE> ), dim3( grid ), dim3( 64 * WAVES ), 0, s, d, items, nItems );
    return hipGetLastError();
}

// The batched launch in tile shape `shape` (kMfmaShape); shapes other than 0 exist for one-user
// FP32 instances up to 32 antennas (mfmaShapeBuilt).
template <int PREC, int NMAX, int UMAX>
hipError_t launchBatchMfmaShape( int shape, hipStream_t s, const DetDesc* d, const WorkItem* items, int nItems, int grid )
{
    if constexpr ( PREC == PREC_F32 && NMAX <= 32 && UMAX == 1 )
    {
        switch ( shape )
        {
        case 1: return launchBatchMfmaNU<PREC, NMAX, UMAX, 2, 64>( s, d, items, nItems, grid );
        case 2: return launchBatchMfmaNU<PREC, NMAX, UMAX, 4, 32>( s, d, items, nItems, grid );
        case 3: return launchBatchMfmaNU<PREC, NMAX, UMAX, 8, 64>( s, d, items, nItems, grid );
        case 4: return launchBatchMfmaNU<PREC, NMAX, UMAX, 4, 128>( s, d, items, nItems, grid );
        default: break;
        }
    }
    return shape == 0 ? launchBatchMfmaNU<PREC, NMAX, UMAX>( s, d, items, nItems, grid ) : hipErrorInvalidValue;
}

template <int PREC, int NMAX>
struct MfmaTable
{
    template <int UMAX>
    static hipError_t one( bool batch, hipStream_t s, const float* rx, int64_t M, const float* Z, const float* Zn, const float* C,
                           int64_t ldC, const float* theta, int64_t G, int U, float wL, float wG, float g2, float* est, int N,
                           const DetDesc* d, const WorkItem* items, int nItems, int grid, int shape )
    {
        return batch ? launchBatchMfmaShape<PREC, NMAX, UMAX>( shape, s, d, items, nItems, grid )
                     : launchMfmaNU<PREC, NMAX, UMAX>( s, rx, M, Z, Zn, C, ldC, theta, G, U, wL, wG, g2, est, N );
    }
    static hipError_t byU( int umax, bool batch, hipStream_t s, const float* rx, int64_t M, const float* Z, const float* Zn,
                           const float* C, int64_t ldC, const float* theta, int64_t G, int U, float wL, float wG, float g2, float* est,
                           int N, const DetDesc* d, const WorkItem* items, int nItems, int grid, int shape )
    {
        switch ( umax )
        {
        case 1: return one<1>( batch, s, rx, M, Z, Zn, C, ldC, theta, G, U, wL, wG, g2, est, N, d, items, nItems, grid, shape );
        case 2: return one<2>( batch, s, rx, M, Z, Zn, C, ldC, theta, G, U, wL, wG, g2, est, N, d, items, nItems, grid, shape );
        case 4: return one<4>( batch, s, rx, M, Z, Zn, C, ldC, theta, G, U, wL, wG, g2, est, N, d, items, nItems, grid, shape );
        case 8: return one<8>( batch, s, rx, M, Z, Zn, C, ldC, theta, G, U, wL, wG, g2, est, N, d, items, nItems, grid, shape );
        case 12: return one<12>( batch, s, rx, M, Z, Zn, C, ldC, theta, G, U, wL, wG, g2, est, N, d, items, nItems, grid, shape );
        case 16: return one<16>( batch, s, rx, M, Z, Zn, C, ldC, theta, G, U, wL, wG, g2, est, N, d, items, nItems, grid, shape );
        default: return hipErrorInvalidValue;
        }
    }
};

template <int PREC>
hipError_t launchMfmaPrec( int nmax, int umax, bool batch, hipStream_t s, const float* rx, int64_t M, const float* Z, const float* Zn,
                           const float* C, int64_t ldC, const float* theta, int64_t G, int U, float wL, float wG, float g2, float* est,
                           int N, const DetDesc* d, const WorkItem* items, int nItems, int grid, int shape )
{
    switch ( nmax )
    {
    case 4: return MfmaTable<PREC, 4>::byU( umax, batch, s, rx, M, Z, Zn, C, ldC, theta, G, U, wL, wG, g2, est, N, d, items, nItems, grid, shape );
    case 8: return MfmaTable<PREC, 8>::byU( umax, batch, s, rx, M, Z, Zn, C, ldC, theta, G, U, wL, wG, g2, est, N, d, items, nItems, grid, shape );
    case 16: return MfmaTable<PREC, 16>::byU( umax, batch, s, rx, M, Z, Zn, C, ldC, theta, G, U, wL, wG, g2, est, N, d, items, nItems, grid, shape );
    case 32: return MfmaTable<PREC, 32>::byU( umax, batch, s, rx, M, Z, Zn, C, ldC, theta, G, U, wL, wG, g2, est, N, d, items, nItems, grid, shape );
    case 64: return MfmaTable<PREC, 64>::byU( umax, batch, s, rx, M, Z, Zn, C, ldC, theta, G, U, wL, wG, g2, est, N, d, items, nItems, grid, shape );
    default: return hipErrorInvalidValue;
    }
}

// Entries of each precision's translation unit (defined there).
#define APSM_MFMA_ENTRY( name )                                                                                                   \
    hipError_t name( int nmax, int umax, bool batch, hipStream_t s, const float* rx, int64_t M, const float* Z, const float* Zn,   \
                     const float* C, int64_t ldC, const float* theta, int64_t G, int U, float wL, float wG, float g2, float* est,  \
                     int N, const DetDesc* d, const WorkItem* items, int nItems, int grid, int shape )
APSM_MFMA_ENTRY( mfmaEntryF32 );
APSM_MFMA_ENTRY( mfmaEntryBF16 );
APSM_MFMA_ENTRY( mfmaEntryF16 );

} // namespace apsm
