// Device helpers shared by every translation unit of the library (no kernels here, so that
// each unit may include it): launch constants, the Gaussian of a squared distance, wave sums,
// the complex-pair cross terms, and the eviction shares of the blocked train.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace apsm {

constexpr int kWave = 64;
constexpr int kSeqThreads = 1024; // train sequencer workgroup
constexpr int kSeqWaves = kSeqThreads / kWave;
constexpr int kNewMax = 4; // complex samples entering per chunk inside one step
constexpr int kDetTile = 64; // dictionary entries per LDS tile in detect
constexpr int kDetThreads = 256;

// exp(-gamma*d) for d >= 0 via the hardware base-2 exponential; g2 = gamma*log2(e).
__device__ __forceinline__ float gaussFromDist( float d, float g2 )
{
    return __builtin_amdgcn_exp2f( -g2 * fmaxf( d, 0.0f ) );
}

// Sum over the 64 lanes of a wave, returned in every lane. Within each row of 16 lanes the
// DPP steps quad_perm [1,0,3,2], quad_perm [2,3,0,1], row_ror:4 and row_ror:8 leave the row
// sum in all 16 lanes; the four row sums are then read from lanes 0, 16, 32 and 48.
__device__ __forceinline__ float waveSum( float v )
{
    int x = __float_as_int( v );
    v += __int_as_float( __builtin_amdgcn_update_dpp( 0, x, 0xB1, 0xF, 0xF, false ) );
    x = __float_as_int( v );
    v += __int_as_float( __builtin_amdgcn_update_dpp( 0, x, 0x4E, 0xF, 0xF, false ) );
    x = __float_as_int( v );
    v += __int_as_float( __builtin_amdgcn_update_dpp( 0, x, 0x124, 0xF, 0xF, false ) );
    x = __float_as_int( v );
    v += __int_as_float( __builtin_amdgcn_update_dpp( 0, x, 0x128, 0xF, 0xF, false ) );
    x = __float_as_int( v );
    const float r0 = __int_as_float( __builtin_amdgcn_readlane( x, 0 ) );
    const float r1 = __int_as_float( __builtin_amdgcn_readlane( x, 16 ) );
    const float r2 = __int_as_float( __builtin_amdgcn_readlane( x, 32 ) );
    const float r3 = __int_as_float( __builtin_amdgcn_readlane( x, 48 ) );
    return ( r0 + r1 ) + ( r2 + r3 );
}

// Complex-pair kernel values for (x, z): p = a_x.a_z, s = a_x.b_z.
__device__ __forceinline__ void pairDots( const float* __restrict__ x, const float* __restrict__ z, int N, float& p, float& sp )
{
    p = 0.f;
    sp = 0.f;
    for ( int n = 0; n < N; n++ )
    {
        float xr = x[ n ], xi = x[ N + n ], zr = z[ n ], zi = z[ N + n ];
        p = fmaf( xr, zr, p );
        p = fmaf( xi, zi, p );
        sp = fmaf( xr, zi, sp );
        sp = fmaf( -xi, zr, sp );
    }
}

struct EvArgs
{
    int N, L, U, q, w2, nSteps;
    int64_t rowBase, sPrev, sFirst, H, ldYC;
    float wG, g2;
};

// EvF[u][k][2*pc + part]: Gaussian share of the rows leaving the dictionary at step k, for
// the window row at position 2*pc + part (row s_k - W_k + 2*pc + part) if it entered before
// step k; zero otherwise. One thread per (step, complex window position); the kernel values
// of the (at most kNewMax) leaving samples are computed once and contracted for every user.
__device__ __forceinline__ void evictCorrBody( const float* __restrict__ Z, const float* __restrict__ Zn, const float* __restrict__ C,
                                               float* __restrict__ EvF, const EvArgs& a, const int64_t idx )
{
    const int Wc = a.w2 / 2;
    if ( idx >= (int64_t)a.nSteps * Wc )
        return;
    const int kk = (int)( idx / Wc ), pc = (int)( idx % Wc );
    const int64_t s = a.sFirst + (int64_t)kk * a.q;
    const int64_t sp = kk == 0 ? a.sPrev : s - a.q;
    const int64_t W = s < a.w2 ? s : a.w2;
    const int64_t lo = s > a.H ? s - a.H : 0, loP = sp > a.H ? sp - a.H : 0;
    const int64_t r = s - W + 2 * pc; // absolute real row (a-row of the complex sample)
    const bool live = ( 2 * pc < W ) && ( r < sp ) && ( lo > loP );
    const int64_t e0 = ( loP - a.rowBase ) / 2;
    const int ne = live ? (int)( ( lo - loP ) / 2 ) : 0; // <= kNewMax
    float K0[ kNewMax ], K1[ kNewMax ], K2[ kNewMax ];
    if ( live )
    {
        const int64_t vv = ( r - a.rowBase ) / 2;
#pragma unroll
        for ( int i = 0; i < kNewMax; i++ )
        {
            K0[ i ] = K1[ i ] = K2[ i ] = 0.f;
            if ( i < ne )
            {
                float p, sp2;
                pairDots( Z + vv * a.L, Z + ( e0 + i ) * a.L, a.N, p, sp2 );
                const float nn = Zn[ vv ] + Zn[ e0 + i ];
                K0[ i ] = gaussFromDist( nn - 2.f * p, a.g2 );
                K1[ i ] = gaussFromDist( nn - 2.f * sp2, a.g2 );
                K2[ i ] = gaussFromDist( nn + 2.f * sp2, a.g2 );
            }
        }
    }
    for ( int u = 0; u < a.U; u++ )
    {
        float dFa = 0.f, dFb = 0.f;
#pragma unroll
        for ( int i = 0; i < kNewMax; i++ )
        {
            if ( i < ne )
            {
                const float ca = C[ u * a.ldYC + 2 * ( e0 + i ) ], cb = C[ u * a.ldYC + 2 * ( e0 + i ) + 1 ];
                dFa += ca * K0[ i ] + cb * K1[ i ];
                dFb += ca * K2[ i ] + cb * K0[ i ];
            }
        }
        float* o = EvF + ( (int64_t)u * a.nSteps + kk ) * a.w2;
        o[ 2 * pc ] = a.wG * dFa;
        o[ 2 * pc + 1 ] = a.wG * dFb;
    }
}

struct EvDesc
{
    const float *Z, *Zn, *C;
    float* EvF;
    EvArgs a;
};

} // namespace apsm
