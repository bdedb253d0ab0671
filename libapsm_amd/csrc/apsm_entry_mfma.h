// Frozen-dictionary share of the entering pilots of one train block, with the cross terms on
// the matrix cores (FP32 inputs and accumulation).
//
// For the entering complex samples i < nNewC of a block (rows newC0 + i of the stream buffer)
// and every user u: g_u(i) = sum over frozen dictionary entries j in [j0(i), j1) of
// (c_a E0 + c_b E1, c_a E2 + c_b E0), where j1 is the block's frozen boundary and j0(i) the
// dictionary's lower end when sample i enters (the FIFO may already have dropped entries).
// The frozen range is cut into chunks of kEntryChunk entries counted from its top (chunk 0 holds
// the most recently frozen entries); one workgroup takes one chunk
// for all (at most 128) entering samples, with the layout of the matrix-core detect (4 waves x
// 32 samples, lane l holds sample l % 32 and entries 8(v/4) + 4(l/32) + v%4 of each 32-entry
// block), masks the entries below j0(i), and writes its partial sums. A second kernel adds the
// chunks in order and the linear share: P_u(i) = wL thF_u.x_i + wG g_u(i).
#pragma once

#include "apsm_detect_mfma.h"
#include "apsm_gram_mfma.h"

namespace apsm {

template <int NMAX, int UMAX>
__device__ __forceinline__ void entryChunkMfma( const float* __restrict__ Z, const float* __restrict__ Zn, const float* __restrict__ C,
                                                float* __restrict__ part, const EntryArgs& a, const int64_t jBeg, const int nChunks,
                                                const int work )
{
    // work item = (tile of kMfmaSamples entering samples, chunk of the frozen range)
    const int chunk = work % nChunks, mBase = ( work / nChunks ) * kMfmaSamples;
    constexpr int K = 2 * NMAX;
    constexpr int kTile = NMAX >= 64 ? 32 : 64;
    __shared__ float sZf[ 2 * K * kTile ];
    __shared__ float sZn[ kTile ];
    __shared__ float2 sC[ UMAX * kTile ];

    const int N = a.N, L = a.L, U = a.U;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int col = lane & 31, h = lane >> 5;
    const int m = mBase + wave * 32 + col; // entering sample of this lane
    const bool valid = m < a.nNewC;
    const float g2 = a.g2;

    // dictionary lower end when sample m enters (entries below it are masked)
    const int64_t tc = a.newC0 + m;
    const int64_t rt = a.rowBase + 2 * tc;
    const int64_t se = rt < a.sFirst ? a.sFirst : a.sFirst + a.q * ( ( rt - a.sFirst ) / a.q + 1 );
    const int64_t lo = se > a.H ? se - a.H : 0;
    const int64_t jlo = ( lo - a.rowBase ) / 2;
    const int64_t j1 = ( a.a0 - a.rowBase ) / 2;

    float xb[ K / 2 ];
    float xn = 0.f;
#pragma unroll
    for ( int t = 0; t < K / 2; t++ )
    {
        const int k = 2 * t + h;
        const int ant = k < NMAX ? k : k - NMAX;
        float v = 0.f;
        if ( valid && ant < N )
            v = Z[ tc * L + ( k < NMAX ? ant : N + ant ) ];
        xb[ t ] = v;
        xn += v * v;
    }
    xn += __shfl_xor( xn, 32 );
    const float xq = -g2 * xn, k2 = 2.f * g2;

    float fa[ UMAX ], fb[ UMAX ];
#pragma unroll
    for ( int u = 0; u < UMAX; u++ )
        fa[ u ] = fb[ u ] = 0.f;

    // chunks from the top of the frozen range: chunk 0 holds the most recently frozen entries
    const int64_t c1 = j1 - (int64_t)chunk * kEntryChunk;
    const int64_t c0 = ( c1 - kEntryChunk ) > jBeg ? c1 - kEntryChunk : jBeg;
    for ( int64_t j0 = c0; j0 < c1; j0 += kTile )
    {
        const int nt = ( c1 - j0 ) < kTile ? (int)( c1 - j0 ) : kTile;
        __syncthreads();
#pragma unroll 2
        for ( int i = tid; i < K * kTile; i += kMfmaThreads )
        {
            const int k = i / kTile, e = i % kTile;
            const int n = k < NMAX ? k : k - NMAX;
            float zr = 0.f, zi = 0.f;
            if ( e < nt && n < N )
            {
                zr = Z[ ( j0 + e ) * L + n ];
                zi = Z[ ( j0 + e ) * L + N + n ];
            }
            sZf[ k * kTile + e ] = k < NMAX ? zr : zi;
            sZf[ K * kTile + k * kTile + e ] = k < NMAX ? zi : -zr;
        }
        for ( int i = tid; i < kTile; i += kMfmaThreads )
            sZn[ i ] = i < nt ? -g2 * Zn[ j0 + i ] : 0.f;
        for ( int i = tid; i < UMAX * kTile; i += kMfmaThreads )
        {
            const int u = i / kTile, e = i % kTile;
            sC[ i ] = ( u < U && e < nt ) ? make_float2( C[ u * a.ldYC + 2 * ( j0 + e ) ], C[ u * a.ldYC + 2 * ( j0 + e ) + 1 ] )
                                          : make_float2( 0.f, 0.f );
        }
        __syncthreads();
        // first entry of this tile the lane's sample may use (the FIFO dropped those below)
        const int rel = jlo <= j0 ? 0 : ( jlo - j0 >= kTile ? kTile : (int)( jlo - j0 ) );

#pragma unroll 1
        for ( int cb = 0; cb < nt; cb += 32 )
        {
            mf32x16 accP = {}, accS = {};
#pragma unroll
            for ( int t = 0; t < K / 2; t++ )
            {
                const int k = 2 * t + h;
                const float ap = sZf[ k * kTile + cb + col ];
                const float as = sZf[ K * kTile + k * kTile + cb + col ];
                accP = __builtin_amdgcn_mfma_f32_32x32x2f32( ap, xb[ t ], accP, 0, 0, 0 );
                accS = __builtin_amdgcn_mfma_f32_32x32x2f32( as, xb[ t ], accS, 0, 0, 0 );
                if ( ( t & 7 ) == 7 )
                    __builtin_amdgcn_sched_barrier( 0 );
            }
#pragma unroll
            for ( int v = 0; v < 16; v++ )
            {
                const int e = cb + 8 * ( v / 4 ) + 4 * h + ( v % 4 );
                // entries the FIFO dropped before this sample entered contribute nothing: their
                // exponent argument is pushed to -1e30 (an arithmetic mask; a select here costs
                // the scheduler its freedom and doubles the register count)
                const float q = xq + sZn[ e ] - ( e >= rel ? 0.f : 1.0e30f );
                const float p = accP[ v ], s = accS[ v ];
                const float E0 = __builtin_amdgcn_exp2f( fminf( fmaf( k2, p, q ), 0.f ) );
                const float E1 = __builtin_amdgcn_exp2f( fminf( fmaf( k2, s, q ), 0.f ) );
                const float E2 = __builtin_amdgcn_exp2f( fminf( fmaf( -k2, s, q ), 0.f ) );
#pragma unroll
                for ( int u = 0; u < UMAX; u++ )
                {
                    const float2 c = sC[ u * kTile + e ];
                    fa[ u ] = fmaf( c.x, E0, fmaf( c.y, E1, fa[ u ] ) );
                    fb[ u ] = fmaf( c.x, E2, fmaf( c.y, E0, fb[ u ] ) );
                }
                if ( NMAX <= 32 )
                    __builtin_amdgcn_sched_barrier( 0 );
            }
        }
    }
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
        if ( u < U )
            reinterpret_cast<float2*>( part )[ ( (int64_t)u * nChunks + chunk ) * a.nNewC + m ] = make_float2( fa[ u ], fb[ u ] );
}

// P_u(i) = wL thF_u.x_i + wG (chunk partials added in chunk order); one thread per (u, i).
__device__ __forceinline__ void entryCombineBody( const float* __restrict__ Z, const float* __restrict__ thF, const float* __restrict__ part,
                                                  float* __restrict__ P, const EntryArgs& a, const int nChunks, const int idx )
{
    if ( idx >= a.U * a.nNewC )
        return;
    const int u = idx / a.nNewC, i = idx % a.nNewC;
    const int N = a.N;
    const float* x = Z + ( a.newC0 + i ) * a.L;
    const float* th = thF + (int64_t)u * a.L;
    float la = 0.f, lb = 0.f;
    for ( int n = 0; n < N; n++ )
    {
        const float xr = x[ n ], xi = x[ N + n ];
        la += th[ n ] * xr + th[ N + n ] * xi;
        lb += th[ n ] * xi - th[ N + n ] * xr;
    }
    float ga = 0.f, gb = 0.f;
    const float2* pp = reinterpret_cast<const float2*>( part ) + (int64_t)u * nChunks * a.nNewC + i;
#pragma unroll 8
    for ( int c = 0; c < nChunks; c++ )
    {
        const float2 v = pp[ (int64_t)c * a.nNewC ];
        ga += v.x;
        gb += v.y;
    }
    P[ ( (int64_t)u * a.nNewC + i ) * 2 + 0 ] = a.wL * la + a.wG * ga;
    P[ ( (int64_t)u * a.nNewC + i ) * 2 + 1 ] = a.wL * lb + a.wG * gb;
}

template <int NMAX, int UMAX>
__global__ void __launch_bounds__( kMfmaThreads ) apsm_entry_mfma( const float* __restrict__ Z, const float* __restrict__ Zn,
                                                                  const float* __restrict__ C, float* __restrict__ part, const EntryArgs a,
                                                                  int64_t jBeg, int nChunks, int chunk0 )
{
    entryChunkMfma<NMAX, UMAX>( Z, Zn, C, part, a, jBeg, nChunks, chunk0 + blockIdx.x );
}

__global__ void apsm_entry_combine( const float* __restrict__ Z, const float* __restrict__ thF, const float* __restrict__ part,
                                    float* __restrict__ P, const EntryArgs a, int nChunks )
{
    entryCombineBody( Z, thF, part, P, a, nChunks, blockIdx.x * blockDim.x + threadIdx.x );
}

__global__ void batch_entry_combine( const EntryMfmaDesc* __restrict__ d )
{
    const EntryMfmaDesc& e = d[ blockIdx.y ];
    entryCombineBody( e.Z, e.thF, e.part, e.P, e.a, e.nChunks, blockIdx.x * blockDim.x + threadIdx.x );
}

// One block's precompute, workgroup b: entry work item (the first entryWg workgroups), Gram
// blocks (the next gramWg) or eviction shares (the last evWg); batched: gridX = the largest sum.
template <int NMAX, int UMAX>
__device__ __forceinline__ void blockPreBody( const PreDesc& d, int b )
{
    if ( b < d.entryWg )
    {
        entryChunkMfma<NMAX, UMAX>( d.e.Z, d.e.Zn, d.e.C, d.e.part, d.e.a, d.e.jBeg, d.e.nChunks, b );
        return;
    }
    b -= d.entryWg;
    if ( b < d.gramWg )
    {
        const int64_t blk = (int64_t)b * ( kMfmaThreads / 64 ) + ( threadIdx.x >> 6 );
        if ( blk < gramBlocks( d.g.a.nA ) )
            gramBlockMfma<NMAX>( d.g.Z, d.g.Zn, d.g.Gact, d.g.a, blk );
        return;
    }
    b -= d.gramWg;
    if ( b < d.evWg )
        evictCorrBody( d.v.Z, d.v.Zn, d.v.C, d.v.EvF, d.v.a, (int64_t)b * kMfmaThreads + threadIdx.x );
}

template <int NMAX, int UMAX>
__global__ void __launch_bounds__( kMfmaThreads ) apsm_block_pre( const PreDesc d )
{
    blockPreBody<NMAX, UMAX>( d, blockIdx.x );
}

template <int NMAX, int UMAX>
__global__ void __launch_bounds__( kMfmaThreads ) batch_block_pre( const PreDesc* __restrict__ d )
{
    const PreDesc& x = d[ blockIdx.y ];
    if ( (int)blockIdx.x >= x.entryWg + x.gramWg + x.evWg )
        return;
    blockPreBody<NMAX, UMAX>( x, blockIdx.x );
}

template <int NMAX, int UMAX>
hipError_t launchPreNU( bool batch, hipStream_t s, const PreDesc* d, int n, int gridX )
{
    if ( batch )
        hipLaunchKernelGGL( ( batch_block_pre<NMAX, UMAX> ), dim3( gridX, n ), dim3( kMfmaThreads ), 0, s, d );
    else
        hipLaunchKernelGGL( ( apsm_block_pre<NMAX, UMAX> ), dim3( gridX ), dim3( kMfmaThreads ), 0, s, *d );
    return hipGetLastError();
}

template <int NMAX>
hipError_t launchPreN( int umax, bool batch, hipStream_t s, const PreDesc* d, int n, int gridX )
{
    switch ( umax )
    {
    case 1: return launchPreNU<NMAX, 1>( batch, s, d, n, gridX );
    case 2: return launchPreNU<NMAX, 2>( batch, s, d, n, gridX );
    case 4: return launchPreNU<NMAX, 4>( batch, s, d, n, gridX );
    case 8: return launchPreNU<NMAX, 8>( batch, s, d, n, gridX );
    case 12: return launchPreNU<NMAX, 12>( batch, s, d, n, gridX );
    case 16: return launchPreNU<NMAX, 16>( batch, s, d, n, gridX );
    default: return hipErrorInvalidValue;
    }
}

inline hipError_t launchPreAnyN( int nmax, int umax, bool batch, hipStream_t s, const PreDesc* d, int n, int gridX )
{
    switch ( nmax )
    {
    case 4: return launchPreN<4>( umax, batch, s, d, n, gridX );
    case 8: return launchPreN<8>( umax, batch, s, d, n, gridX );
    case 16: return launchPreN<16>( umax, batch, s, d, n, gridX );
    case 32: return launchPreN<32>( umax, batch, s, d, n, gridX );
    case 64: return launchPreN<64>( umax, batch, s, d, n, gridX );
    default: return hipErrorInvalidValue;
    }
}

// Entry work items [chunk0, chunk0 + count) of one block of a single handle.
template <int NMAX, int UMAX>
hipError_t launchEntryNU( hipStream_t s, const float* Z, const float* Zn, const float* C, float* part, const EntryArgs& a, int64_t jBeg,
                          int nChunks, int chunk0, int count )
{
    hipLaunchKernelGGL( ( apsm_entry_mfma<NMAX, UMAX> ), dim3( count ), dim3( kMfmaThreads ), 0, s, Z, Zn, C, part, a, jBeg, nChunks, chunk0 );
    return hipGetLastError();
}

template <int NMAX>
hipError_t launchEntryN( int umax, hipStream_t s, const float* Z, const float* Zn, const float* C, float* part, const EntryArgs& a,
                         int64_t jBeg, int nChunks, int chunk0, int count )
{
    switch ( umax )
    {
    case 1: return launchEntryNU<NMAX, 1>( s, Z, Zn, C, part, a, jBeg, nChunks, chunk0, count );
    case 2: return launchEntryNU<NMAX, 2>( s, Z, Zn, C, part, a, jBeg, nChunks, chunk0, count );
    case 4: return launchEntryNU<NMAX, 4>( s, Z, Zn, C, part, a, jBeg, nChunks, chunk0, count );
    case 8: return launchEntryNU<NMAX, 8>( s, Z, Zn, C, part, a, jBeg, nChunks, chunk0, count );
    case 12: return launchEntryNU<NMAX, 12>( s, Z, Zn, C, part, a, jBeg, nChunks, chunk0, count );
    case 16: return launchEntryNU<NMAX, 16>( s, Z, Zn, C, part, a, jBeg, nChunks, chunk0, count );
    default: return hipErrorInvalidValue;
    }
}

inline hipError_t launchEntryAnyN( int nmax, int umax, hipStream_t s, const float* Z, const float* Zn, const float* C, float* part,
                                   const EntryArgs& a, int64_t jBeg, int nChunks, int chunk0, int count )
{
    switch ( nmax )
    {
    case 4: return launchEntryN<4>( umax, s, Z, Zn, C, part, a, jBeg, nChunks, chunk0, count );
    case 8: return launchEntryN<8>( umax, s, Z, Zn, C, part, a, jBeg, nChunks, chunk0, count );
    case 16: return launchEntryN<16>( umax, s, Z, Zn, C, part, a, jBeg, nChunks, chunk0, count );
    case 32: return launchEntryN<32>( umax, s, Z, Zn, C, part, a, jBeg, nChunks, chunk0, count );
    case 64: return launchEntryN<64>( umax, s, Z, Zn, C, part, a, jBeg, nChunks, chunk0, count );
    default: return hipErrorInvalidValue;
    }
}

} // namespace apsm
