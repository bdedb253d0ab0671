// Device code of libapsm_amd: APSM train / detect and LLS kernels for gfx950 (CDNA4).
//
// Data layout in HBM (one handle):
//   stream buffer  Z  [K][2N] float  complex pilots, row k = [Re x_k (N) | Im x_k (N)],
//                                    oldest first; the dictionary is its last <= Dc rows
//                  Zn [K]            |x_k|^2
//                  Y  [U][2K]        real-row targets (Re s_k, Im s_k interleaved)
//                  C  [U][2K]        real-row Gaussian coefficients (row 2k: a_k, 2k+1: b_k)
//   window state   F  [U][W2]        f of each window row under the current state, by slot
//                  G  [U][W2][W2]    window Gram matrix kappa(t,u) by slot
//   linear part    thOld [U][2N]     theta_0 + sum of evicted rows' c_j x_j
//                  theta [U][2N]     full linear weights (detect)
// where a real row with absolute stream index r sits in slot r % W2 (W2 = 2*windowSize).
//
// Real embedding (cpp/lib/cuda/cuda_matrix_device.cuh:180-188,288-296): complex x gives the
// rows a = [Re x; Im x] and b = [Im x; -Re x]. For a pair (x, z) with p = a_x.a_z,
// s = a_x.b_z and n = |x|^2 + |z|^2 the four real squared distances are
//   |a_x-a_z|^2 = |b_x-b_z|^2 = n - 2p,  |a_x-b_z|^2 = n - 2s,  |b_x-a_z|^2 = n + 2s,
// so one complex pair needs 2N complex multiply-adds and three exponentials.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "apsm_common.h"

namespace apsm {

// ---------------------------------------------------------------------------------------
// Staging: host layout rx [N][S][2] (+ optional permutation) -> stream rows Z [S][2N] and
// norms; tx [U][S][2] -> Y [U][2S]; C of the new rows cleared.
// ---------------------------------------------------------------------------------------
__device__ __forceinline__ void stagePilotsBody( const float* __restrict__ rx, const float* __restrict__ tx, const int32_t* __restrict__ perm,
                                                 int N, int U, int64_t S, float* __restrict__ Z, float* __restrict__ Zn,
                                                 float* __restrict__ Y, float* __restrict__ C, int64_t ldYC, const int64_t k )
{
    if ( k >= S )
        return;
    const int64_t src = perm ? perm[ k ] : k;
    const int L = 2 * N;
    float nrm = 0.f;
    for ( int n = 0; n < N; n++ )
    {
        float2 v = reinterpret_cast<const float2*>( rx )[ (int64_t)n * S + src ];
        Z[ k * L + n ] = v.x;
        Z[ k * L + N + n ] = v.y;
        nrm += v.x * v.x + v.y * v.y;
    }
    Zn[ k ] = nrm;
    for ( int u = 0; u < U; u++ )
    {
        float2 t = reinterpret_cast<const float2*>( tx )[ (int64_t)u * S + src ];
        Y[ u * ldYC + 2 * k ] = t.x;
        Y[ u * ldYC + 2 * k + 1 ] = t.y;
        C[ u * ldYC + 2 * k ] = 0.f;
        C[ u * ldYC + 2 * k + 1 ] = 0.f;
    }
}

__global__ void stage_pilots( const float* __restrict__ rx, const float* __restrict__ tx, const int32_t* __restrict__ perm, int N, int U,
                              int64_t S, float* __restrict__ Z, float* __restrict__ Zn, float* __restrict__ Y, float* __restrict__ C,
                              int64_t ldYC )
{
    stagePilotsBody( rx, tx, perm, N, U, S, Z, Zn, Y, C, ldYC, (int64_t)blockIdx.x * blockDim.x + threadIdx.x );
}

// ---------------------------------------------------------------------------------------
// Train sequencer. One workgroup per detected user runs every projection step of one
// train call (the reference's host loop, cpp/lib/apsm/apsm_train.cu:107-160, moved onto
// the device). Per step (apsm_train.cu:333-428, same pre-step-state semantics):
//   1. rows leaving the dictionary subtract their Gaussian share from window rows' f;
//      their c_j x_j moves into thOld;
//   2. entering rows get f = wL*thOld.x + sum_j c_j kappa(j, x) over the dictionary
//      (rows not yet projected have c = 0); their Gram entries with the window are stored;
//   3. beta_t = soft(y_t - f_t, eB) / (W*(wL|x_t|^2 + wG)) for the window rows; c_t += beta_t;
//   4. f_u += sum_t kappa(t,u) beta_t for every window row (the window Gram product).
// ---------------------------------------------------------------------------------------
struct SeqArgs
{
    int N, L, U, w2, q, passes, fullStart, pad;
    int64_t H; // real dictionary capacity 2*Dc
    int64_t usedOld; // absolute real rows before this call
    int64_t R2; // real rows in this call
    int64_t rowBase; // absolute real row of buffer row 0 (even)
    int64_t ldYC; // row stride of Y and C (2K)
    float wL, wG, g2, eB;
};

__device__ __forceinline__ int slotOf( int64_t absRow, int w2 ) { return (int)( ( absRow % w2 + w2 ) % w2 ); }

__global__ void __launch_bounds__( kSeqThreads ) apsm_train_seq( const float* __restrict__ Z, const float* __restrict__ Zn,
                                                               const float* __restrict__ Y, float* __restrict__ C,
                                                               float* __restrict__ Fg, float* __restrict__ Gram,
                                                               float* __restrict__ thOld, const SeqArgs a )
{
    const int u = blockIdx.x;
    const int tid = threadIdx.x;
    const int lane = tid & ( kWave - 1 );
    const int wave = tid / kWave;
    const int N = a.N, L = a.L, w2 = a.w2;

    extern __shared__ float sm[];
    float* sF = sm; // [w2]
    float* sB = sF + w2; // [w2]
    float* sTh = sB + w2; // [L]
    float* sX = sTh + L; // [kNewMax][L]
    float* sXn = sX + kNewMax * L; // [kNewMax]
    float* sRed = sXn + kNewMax; // [kSeqWaves][2*kNewMax]
    float* sPart = sRed + kSeqWaves * 2 * kNewMax; // [kSeqWaves][kWave]

    const float* Yu = Y + u * a.ldYC;
    float* Cu = C + u * a.ldYC;
    float* Gu = Gram + (int64_t)u * w2 * w2;
    float* Fu = Fg + (int64_t)u * w2;
    const float wL = a.wL, wG = a.wG, g2 = a.g2;

    for ( int i = tid; i < w2; i += kSeqThreads )
        sF[ i ] = Fu[ i ];
    for ( int i = tid; i < L; i += kSeqThreads )
        sTh[ i ] = thOld[ u * L + i ];
    __syncthreads();

    // F-update work split: waves cover (u-slot block of 64) x (t-slot part)
    const int uBlocks = ( w2 + kWave - 1 ) / kWave;
    const int tParts = kSeqWaves / uBlocks; // >= 1 because w2 <= 1024
    const int myUB = wave % uBlocks;
    const int myTP = wave / uBlocks;
    const int tPer = ( w2 + tParts - 1 ) / tParts;

    int64_t sPrev = a.usedOld;
    const int64_t first = ( a.fullStart && a.usedOld == 0 ) ? (int64_t)w2 : (int64_t)a.q;
    const int64_t sFinal = a.usedOld + ( a.R2 >= first ? first + ( ( a.R2 - first ) / a.q ) * a.q : 0 );

    for ( int pass = 0; pass < a.passes; pass++ )
    {
        for ( int64_t sl = first; sl <= a.R2; sl += a.q )
        {
            const bool moved = ( pass == 0 );
            const int64_t s = moved ? a.usedOld + sl : sFinal;
            const int64_t W = s < (int64_t)w2 ? s : (int64_t)w2;
            const int64_t nb = moved ? W : ( W < sl ? W : sl );
            const int64_t winLo = s - W; // absolute

            if ( moved )
            {
                const int64_t lo = s > a.H ? s - a.H : 0;
                const int64_t loPrev = sPrev > a.H ? sPrev - a.H : 0;
                // ---- 1. evictions (complex entries [loPrev/2, lo/2)) ----
                if ( lo > loPrev )
                {
                    const int64_t e0 = ( loPrev - a.rowBase ) / 2, e1 = ( lo - a.rowBase ) / 2;
                    // thOld += c_a a_e + c_b b_e
                    for ( int d = tid; d < L; d += kSeqThreads )
                    {
                        float acc = 0.f;
                        for ( int64_t e = e0; e < e1; e++ )
                        {
                            const float* ze = Z + e * L;
                            float bd = d < N ? ze[ N + d ] : -ze[ d - N ];
                            acc += Cu[ 2 * e ] * ze[ d ] + Cu[ 2 * e + 1 ] * bd;
                        }
                        sTh[ d ] += acc;
                    }
                    // window rows that were already in the window lose the Gaussian share
                    const int64_t v0 = ( winLo - a.rowBase ) / 2, v1 = ( sPrev - a.rowBase ) / 2;
                    for ( int64_t v = v0 + tid; v < v1; v += kSeqThreads )
                    {
                        const float* xv = Z + v * L;
                        float dFa = 0.f, dFb = 0.f;
                        for ( int64_t e = e0; e < e1; e++ )
                        {
                            const float* ze = Z + e * L;
                            float p = 0.f, sp = 0.f;
                            for ( int n = 0; n < N; n++ )
                            {
                                float xr = xv[ n ], xi = xv[ N + n ], zr = ze[ n ], zi = ze[ N + n ];
                                p += xr * zr + xi * zi;
                                sp += xr * zi - xi * zr;
                            }
                            float nn = Zn[ v ] + Zn[ e ];
                            float E0 = gaussFromDist( nn - 2.f * p, g2 );
                            float E1 = gaussFromDist( nn - 2.f * sp, g2 );
                            float E2 = gaussFromDist( nn + 2.f * sp, g2 );
                            float ca = Cu[ 2 * e ], cb = Cu[ 2 * e + 1 ];
                            dFa += ca * E0 + cb * E1;
                            dFb += ca * E2 + cb * E0;
                        }
                        const int64_t rv = a.rowBase + 2 * v;
                        sF[ slotOf( rv, w2 ) ] -= wG * dFa;
                        sF[ slotOf( rv + 1, w2 ) ] -= wG * dFb;
                    }
                    __syncthreads();
                }
                // ---- 2. entering rows, in chunks of kNewMax complex samples ----
                const int64_t t0all = ( sPrev - a.rowBase ) / 2, t1all = ( s - a.rowBase ) / 2;
                const int64_t jLo = ( lo - a.rowBase ) / 2;
                const int64_t winLoC = ( winLo - a.rowBase ) / 2;
                for ( int64_t tc = t0all; tc < t1all; tc += kNewMax )
                {
                    const int nNew = (int)( ( t1all - tc ) < kNewMax ? ( t1all - tc ) : kNewMax );
                    for ( int i = tid; i < nNew * L; i += kSeqThreads )
                        sX[ i ] = Z[ tc * L + i ];
                    for ( int i = tid; i < nNew; i += kSeqThreads )
                        sXn[ i ] = Zn[ tc + i ];
                    __syncthreads();
                    float pa[ kNewMax ], pb[ kNewMax ];
#pragma unroll
                    for ( int t = 0; t < kNewMax; t++ )
                        pa[ t ] = pb[ t ] = 0.f;
                    const int64_t jHi = tc + nNew; // includes the chunk itself (c = 0 there)
                    for ( int64_t j = jLo + tid; j < jHi; j += kSeqThreads )
                    {
                        const float* zj = Z + j * L;
                        const float znj = Zn[ j ];
                        const float ca = Cu[ 2 * j ], cb = Cu[ 2 * j + 1 ];
                        const bool inWin = j >= winLoC;
                        const int64_t rj = a.rowBase + 2 * j;
                        const int sja = slotOf( rj, w2 ), sjb = slotOf( rj + 1, w2 );
                        for ( int t = 0; t < nNew; t++ )
                        {
                            const float* xt = sX + t * L;
                            float p = 0.f, sp = 0.f;
                            for ( int n = 0; n < N; n++ )
                            {
                                float xr = xt[ n ], xi = xt[ N + n ], zr = zj[ n ], zi = zj[ N + n ];
                                p += xr * zr + xi * zi;
                                sp += xr * zi - xi * zr;
                            }
                            float nn = sXn[ t ] + znj;
                            float E0 = gaussFromDist( nn - 2.f * p, g2 );
                            float E1 = gaussFromDist( nn - 2.f * sp, g2 );
                            float E2 = gaussFromDist( nn + 2.f * sp, g2 );
                            float kaa = wL * p + wG * E0; // kappa(a_j,a_t) = kappa(b_j,b_t)
                            float kba = wL * sp + wG * E1; // kappa(b_j,a_t)
                            float kab = -wL * sp + wG * E2; // kappa(a_j,b_t)
                            pa[ t ] += ca * kaa + cb * kba;
                            pb[ t ] += ca * kab + cb * kaa;
                            if ( inWin )
                            {
                                const int64_t rt = a.rowBase + 2 * ( tc + t );
                                const int sta = slotOf( rt, w2 ), stb = slotOf( rt + 1, w2 );
                                Gu[ sta * w2 + sja ] = kaa;
                                Gu[ sja * w2 + sta ] = kaa;
                                Gu[ stb * w2 + sjb ] = kaa;
                                Gu[ sjb * w2 + stb ] = kaa;
                                Gu[ sta * w2 + sjb ] = kba;
                                Gu[ sjb * w2 + sta ] = kba;
                                Gu[ stb * w2 + sja ] = kab;
                                Gu[ sja * w2 + stb ] = kab;
                            }
                        }
                    }
                    // block reduction of the partial f values
#pragma unroll
                    for ( int t = 0; t < kNewMax; t++ )
                    {
                        float va = waveSum( pa[ t ] ), vb = waveSum( pb[ t ] );
                        if ( lane == 0 )
                        {
                            sRed[ wave * 2 * kNewMax + 2 * t ] = va;
                            sRed[ wave * 2 * kNewMax + 2 * t + 1 ] = vb;
                        }
                    }
                    __syncthreads();
                    if ( tid < 2 * nNew )
                    {
                        const int t = tid >> 1, part = tid & 1;
                        float acc = 0.f;
                        for ( int w = 0; w < kSeqWaves; w++ )
                            acc += sRed[ w * 2 * kNewMax + tid ];
                        const float* xt = sX + t * L;
                        float lin = 0.f;
                        for ( int n = 0; n < N; n++ )
                        {
                            float xr = xt[ n ], xi = xt[ N + n ];
                            // a = [xr; xi], b = [xi; -xr]
                            lin += part == 0 ? ( sTh[ n ] * xr + sTh[ N + n ] * xi ) : ( sTh[ n ] * xi - sTh[ N + n ] * xr );
                        }
                        const int64_t r = a.rowBase + 2 * ( tc + t ) + part;
                        sF[ slotOf( r, w2 ) ] = wL * lin + acc;
                    }
                    __syncthreads();
                }
            }
            // ---- 3. projection weights over the window ----
            for ( int i = tid; i < w2; i += kSeqThreads )
            {
                const int64_t r = s - w2 + i; // the last w2 absolute rows cover every slot once
                const int sl_ = slotOf( r, w2 );
                float beta = 0.f;
                if ( r >= winLo && r - winLo < nb )
                {
                    const int64_t rl = r - a.rowBase;
                    float res = Yu[ rl ] - sF[ sl_ ];
                    float b = 0.f;
                    if ( res > a.eB )
                        b = res - a.eB;
                    else if ( res < -a.eB )
                        b = res + a.eB;
                    b /= (float)W;
                    b /= ( wL * Zn[ rl >> 1 ] + wG );
                    beta = b;
                    Cu[ rl ] += b;
                }
                sB[ sl_ ] = beta;
            }
            __syncthreads();
            // ---- 4. f_u += sum_t kappa(t,u) beta_t ----
            {
                float acc = 0.f;
                const int us = myUB * kWave + lane;
                if ( myTP < tParts && us < w2 )
                {
                    const int tb = myTP * tPer;
                    const int te = tb + tPer < w2 ? tb + tPer : w2;
                    for ( int t = tb; t < te; t++ )
                    {
                        float bt = sB[ t ];
                        if ( bt != 0.f )
                            acc += Gu[ t * w2 + us ] * bt;
                    }
                }
                sPart[ wave * kWave + lane ] = acc;
                __syncthreads();
                for ( int us2 = tid; us2 < w2; us2 += kSeqThreads )
                {
                    const int ub = us2 / kWave, ln = us2 % kWave;
                    float sum = 0.f;
                    for ( int tp = 0; tp < tParts; tp++ )
                        sum += sPart[ ( tp * uBlocks + ub ) * kWave + ln ];
                    sF[ us2 ] += sum;
                }
                __syncthreads();
            }
            sPrev = s;
        }
    }
    for ( int i = tid; i < w2; i += kSeqThreads )
        Fu[ i ] = sF[ i ];
    for ( int i = tid; i < L; i += kSeqThreads )
        thOld[ u * L + i ] = sTh[ i ];
}

// theta = thOld + sum over dictionary rows of c_j x_j (one workgroup of 256 threads per user u).
__device__ __forceinline__ void thetaFullBody( const float* __restrict__ Z, const float* __restrict__ C, const float* __restrict__ thOld,
                                               float* __restrict__ theta, int N, int64_t j0, int64_t j1, int64_t ldC, const int u )
{
    const int L = 2 * N;
    __shared__ float red[ 4 ][ 256 ];
    const float* Cu = C + u * ldC;
    for ( int d = 0; d < L; d++ )
    {
        float acc = 0.f;
        for ( int64_t j = j0 + threadIdx.x; j < j1; j += blockDim.x )
        {
            const float* zj = Z + j * L;
            float bd = d < N ? zj[ N + d ] : -zj[ d - N ];
            acc += Cu[ 2 * j ] * zj[ d ] + Cu[ 2 * j + 1 ] * bd;
        }
        acc = waveSum( acc );
        if ( ( threadIdx.x & 63 ) == 0 )
            red[ threadIdx.x / 64 ][ d & 255 ] = acc;
        __syncthreads();
        if ( threadIdx.x == 0 )
        {
            float s = thOld[ u * L + d ];
            for ( int w = 0; w < (int)( blockDim.x / 64 ); w++ )
                s += red[ w ][ d & 255 ];
            theta[ u * L + d ] = s;
        }
        __syncthreads();
    }
}

__global__ void apsm_theta_full( const float* __restrict__ Z, const float* __restrict__ C, const float* __restrict__ thOld,
                                 float* __restrict__ theta, int N, int64_t j0, int64_t j1, int64_t ldC )
{
    thetaFullBody( Z, C, thOld, theta, N, j0, j1, ldC, blockIdx.x );
}

// ---------------------------------------------------------------------------------------
// Detect (cpp/lib/apsm/apsm_detect.cu:674-876 computes the same f per real row).
// One thread per complex data sample, reading rx directly in the host layout [N][M][2];
// dictionary tiles of kDetTile complex entries (rows, norms, both coefficient rows of
// every user) are staged in LDS and read as wave-wide broadcasts.
// est[u][m] = ( wL th_u.a_m + wG sum_j(c_2j E0 + c_2j+1 E1),
//               wL th_u.b_m + wG sum_j(c_2j E2 + c_2j+1 E0) )
// ---------------------------------------------------------------------------------------
// One tile of THREADS samples m = tile*THREADS + threadIdx.x (outputs only for m < M). Every
// thread of the workgroup takes part in the tile loads and barriers.
template <int NMAX, int UMAX, int THREADS>
__device__ __forceinline__ void detectTile( const float* __restrict__ rx, int64_t M, const float* __restrict__ Z,
                                            const float* __restrict__ Zn, const float* __restrict__ C, int64_t ldC,
                                            const float* __restrict__ theta, int64_t G, int U, float wL, float wG, float g2,
                                            float* __restrict__ est, int N, const int64_t tile )
{
    // tiles are stored with the antenna count padded to NMAX (zeros), so the inner loops
    // carry no bounds checks for N < NMAX
    __shared__ float sZr[ kDetTile * NMAX ];
    __shared__ float sZi[ kDetTile * NMAX ];
    __shared__ float sZn[ kDetTile ];
    __shared__ float sC[ UMAX * 2 * kDetTile ];
    const int L = 2 * N;

    const int64_t m = tile * THREADS + threadIdx.x;
    const bool valid = m < M;
    float xr[ NMAX ], xi[ NMAX ];
    float xn = 0.f;
#pragma unroll
    for ( int n = 0; n < NMAX; n++ )
    {
        float2 v = ( valid && n < N ) ? reinterpret_cast<const float2*>( rx )[ (int64_t)n * M + m ] : make_float2( 0.f, 0.f );
        xr[ n ] = v.x;
        xi[ n ] = v.y;
        xn += v.x * v.x + v.y * v.y;
    }
    float fa[ UMAX ], fb[ UMAX ];
#pragma unroll
    for ( int uu = 0; uu < UMAX; uu++ )
        fa[ uu ] = fb[ uu ] = 0.f;

    for ( int64_t j0 = 0; j0 < G; j0 += kDetTile )
    {
        const int nt = ( G - j0 ) < kDetTile ? (int)( G - j0 ) : kDetTile;
        __syncthreads();
        for ( int i = threadIdx.x; i < kDetTile * NMAX; i += THREADS )
        {
            const int j = i / NMAX, n = i % NMAX;
            const bool ok = j < nt && n < N;
            sZr[ i ] = ok ? Z[ ( j0 + j ) * L + n ] : 0.f;
            sZi[ i ] = ok ? Z[ ( j0 + j ) * L + N + n ] : 0.f;
        }
        for ( int i = threadIdx.x; i < kDetTile; i += THREADS )
            sZn[ i ] = i < nt ? Zn[ j0 + i ] : 0.f;
        for ( int i = threadIdx.x; i < UMAX * 2 * kDetTile; i += THREADS )
        {
            int uu = i / ( 2 * kDetTile ), jj = i % ( 2 * kDetTile );
            sC[ i ] = ( uu < U && jj < 2 * nt ) ? C[ uu * ldC + 2 * j0 + jj ] : 0.f;
        }
        __syncthreads();
        for ( int j = 0; j < nt; j++ )
        {
            const float* zr = sZr + j * NMAX;
            const float* zi = sZi + j * NMAX;
            float p = 0.f, sp = 0.f;
#pragma unroll
            for ( int n = 0; n < NMAX; n++ )
            {
                p = fmaf( xr[ n ], zr[ n ], p );
                p = fmaf( xi[ n ], zi[ n ], p );
                sp = fmaf( xr[ n ], zi[ n ], sp );
                sp = fmaf( -xi[ n ], zr[ n ], sp );
            }
            float nn = xn + sZn[ j ];
            float E0 = gaussFromDist( nn - 2.f * p, g2 );
            float E1 = gaussFromDist( nn - 2.f * sp, g2 );
            float E2 = gaussFromDist( nn + 2.f * sp, g2 );
#pragma unroll
            for ( int uu = 0; uu < UMAX; uu++ )
            {
                float ca = sC[ uu * 2 * kDetTile + 2 * j ], cb = sC[ uu * 2 * kDetTile + 2 * j + 1 ];
                fa[ uu ] = fmaf( ca, E0, fmaf( cb, E1, fa[ uu ] ) );
                fb[ uu ] = fmaf( ca, E2, fmaf( cb, E0, fb[ uu ] ) );
            }
        }
    }
    if ( valid )
    {
#pragma unroll
        for ( int uu = 0; uu < UMAX; uu++ )
        {
            if ( uu < U )
            {
                const float* th = theta + uu * L;
                float la = 0.f, lb = 0.f;
#pragma unroll
                for ( int n = 0; n < NMAX; n++ )
                {
                    if ( n < N )
                    {
                        la += th[ n ] * xr[ n ] + th[ N + n ] * xi[ n ];
                        lb += th[ n ] * xi[ n ] - th[ N + n ] * xr[ n ];
                    }
                }
                float2 o = make_float2( wL * la + wG * fa[ uu ], wL * lb + wG * fb[ uu ] );
                reinterpret_cast<float2*>( est )[ (int64_t)uu * M + m ] = o;
            }
        }
    }
}

template <int NMAX, int UMAX>
__global__ void __launch_bounds__( kDetThreads ) apsm_detect_kernel( const float* __restrict__ rx, int64_t M,
                                                                    const float* __restrict__ Z, const float* __restrict__ Zn,
                                                                    const float* __restrict__ C, int64_t ldC,
                                                                    const float* __restrict__ theta, int64_t G, int U,
                                                                    float wL, float wG, float g2, float* __restrict__ est, int N )
{
    detectTile<NMAX, UMAX, kDetThreads>( rx, M, Z, Zn, C, ldC, theta, G, U, wL, wG, g2, est, N, blockIdx.x );
}

// ---------------------------------------------------------------------------------------
// LLS (cpp/lib/noma/lls_detector.cu:136-218): min_w |A w - y| over the real embedding.
// Normal equations A^T A w = A^T y: per-workgroup FP32 partial sums over a contiguous
// chunk of samples, then one workgroup adds the partials in FP64 and solves by Cholesky
// in FP64.
// ---------------------------------------------------------------------------------------
constexpr int kLlsThreads = 256;
constexpr size_t kLlsLdsBudget = 64 * 1024; // bytes of LDS per workgroup for one chunk of samples

// part[b][i*(L+1) + j] for j < L: sum a_i a_j + b_i b_j; j == L: sum a_i y_re + b_i y_im,
// over samples [b*chunk, (b+1)*chunk). The whole chunk sits in LDS; each thread owns
// entries e = tid + o*256 and accumulates one at a time over the chunk.
__global__ void __launch_bounds__( kLlsThreads ) lls_normal_partial( const float* __restrict__ rx, const float* __restrict__ tx,
                                                                   int N, int64_t T, int64_t chunk, float* __restrict__ part )
{
    const int L = 2 * N;
    const int ld = L + 1;
    const int E = L * ld;
    extern __shared__ float sA[]; // [chunk][L+1] rows a (+ y_re), then [chunk][L+1] rows b (+ y_im)
    const int64_t k0 = (int64_t)blockIdx.x * chunk;
    const int nt = (int)( ( k0 + chunk ) < T ? chunk : T - k0 );
    float* sBr = sA + (int64_t)chunk * ld;
    for ( int i = threadIdx.x; i < nt * N; i += kLlsThreads )
    {
        const int n = i / nt, t = i % nt; // consecutive threads read consecutive samples
        float2 v = reinterpret_cast<const float2*>( rx )[ (int64_t)n * T + k0 + t ];
        sA[ t * ld + n ] = v.x;
        sA[ t * ld + N + n ] = v.y;
        sBr[ t * ld + n ] = v.y;
        sBr[ t * ld + N + n ] = -v.x;
    }
    for ( int t = threadIdx.x; t < nt; t += kLlsThreads )
    {
        float2 y = reinterpret_cast<const float2*>( tx )[ k0 + t ];
        sA[ t * ld + L ] = y.x;
        sBr[ t * ld + L ] = y.y;
    }
    __syncthreads();
    for ( int e = threadIdx.x; e < E; e += kLlsThreads )
    {
        const int i = e / ld, j = e % ld;
        float s = 0.f;
        for ( int t = 0; t < nt; t++ )
            s += sA[ t * ld + i ] * sA[ t * ld + j ] + sBr[ t * ld + i ] * sBr[ t * ld + j ];
        part[ (int64_t)blockIdx.x * E + e ] = s;
    }
}

// G[e] = sum over partials in FP64.
__global__ void lls_reduce_partials( const float* __restrict__ part, int nPart, int E, double* __restrict__ G )
{
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if ( e >= E )
        return;
    double s = 0.0;
    for ( int b = 0; b < nPart; b++ )
        s += (double)part[ (int64_t)b * E + e ];
    G[ e ] = s;
}

// Cholesky-factor the normal matrix G = R R^T in FP64 (LDS-resident, L <= 128), solve
// R R^T w = r, write w (FP32) and the status (0 ok, -1 not positive definite).
constexpr int kSolveThreads = 256;
__global__ void __launch_bounds__( kSolveThreads ) lls_solve( const double* __restrict__ Gin, int L, float* __restrict__ w,
                                                            int* __restrict__ status )
{
    const int ld = L + 1;
    const int E = L * ld;
    extern __shared__ double Gm[]; // [L][L+1], 132 KB at L = 128
    for ( int e = threadIdx.x; e < E; e += kSolveThreads )
        Gm[ e ] = Gin[ e ];
    __shared__ int bad;
    if ( threadIdx.x == 0 )
        bad = 0;
    __syncthreads();
    for ( int j = 0; j < L; j++ )
    {
        if ( threadIdx.x == 0 )
        {
            double d = Gm[ j * ld + j ];
            if ( !( d > 0.0 ) )
                bad = 1;
            Gm[ j * ld + j ] = d > 0.0 ? sqrt( d ) : 1.0;
        }
        __syncthreads();
        const double djj = Gm[ j * ld + j ];
        for ( int i = j + 1 + threadIdx.x; i < L; i += kSolveThreads )
            Gm[ i * ld + j ] /= djj;
        __syncthreads();
        const int rem = L - j - 1;
        for ( int e = threadIdx.x; e < rem * rem; e += kSolveThreads )
        {
            const int i = j + 1 + e / rem, k = j + 1 + e % rem;
            if ( k <= i )
                Gm[ i * ld + k ] -= Gm[ i * ld + j ] * Gm[ k * ld + j ];
        }
        __syncthreads();
    }
    if ( threadIdx.x == 0 )
    {
        for ( int i = 0; i < L; i++ )
        {
            double s = Gm[ i * ld + L ];
            for ( int k = 0; k < i; k++ )
                s -= Gm[ i * ld + k ] * Gm[ k * ld + L ];
            Gm[ i * ld + L ] = s / Gm[ i * ld + i ];
        }
        for ( int i = L - 1; i >= 0; i-- )
        {
            double s = Gm[ i * ld + L ];
            for ( int k = i + 1; k < L; k++ )
                s -= Gm[ k * ld + i ] * Gm[ k * ld + L ];
            Gm[ i * ld + L ] = s / Gm[ i * ld + i ];
        }
        for ( int i = 0; i < L; i++ )
            w[ i ] = (float)Gm[ i * ld + L ];
        *status = bad ? -1 : 0;
    }
}

// est[m] = (a_m . w, b_m . w)  (lls_detector.cu:207-218, a GEMV over the embedding).
__global__ void lls_apply( const float* __restrict__ rx, int N, int64_t M, const float* __restrict__ w, float* __restrict__ est )
{
    const int64_t m = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if ( m >= M )
        return;
    float ea = 0.f, eb = 0.f;
    for ( int n = 0; n < N; n++ )
    {
        float2 v = reinterpret_cast<const float2*>( rx )[ (int64_t)n * M + m ];
        ea += v.x * w[ n ] + v.y * w[ N + n ];
        eb += v.y * w[ n ] - v.x * w[ N + n ];
    }
    reinterpret_cast<float2*>( est )[ m ] = make_float2( ea, eb );
}

// Low-level LLS over a general column-major real matrix A (m x n, leading dimension m) and
// target y (m), as trainLowLevel / detectLowLevel take them (lls_detector.cu:157-179,207-218).
// part[b][i*(n+1) + j]: sum over rows of chunk b of A_ri A_rj (j < n) and A_ri y_r (j == n).
__global__ void __launch_bounds__( kLlsThreads ) lls_normal_colmajor( const float* __restrict__ A, const float* __restrict__ y,
                                                                    int n, int64_t m, int64_t chunk, float* __restrict__ part )
{
    const int ld = n + 1;
    const int E = n * ld;
    extern __shared__ float sA[]; // [chunk][n+1]
    const int64_t r0 = (int64_t)blockIdx.x * chunk;
    const int nt = (int)( ( r0 + chunk ) < m ? chunk : m - r0 );
    for ( int i = threadIdx.x; i < nt * ld; i += kLlsThreads )
    {
        const int c = i / nt, t = i % nt; // consecutive threads read consecutive rows of one column
        sA[ t * ld + c ] = c < n ? A[ (int64_t)c * m + r0 + t ] : y[ r0 + t ];
    }
    __syncthreads();
    for ( int e = threadIdx.x; e < E; e += kLlsThreads )
    {
        const int i = e / ld, j = e % ld;
        float s = 0.f;
        for ( int t = 0; t < nt; t++ )
            s += sA[ t * ld + i ] * sA[ t * ld + j ];
        part[ (int64_t)blockIdx.x * E + e ] = s;
    }
}

// est[r] = alpha * sum_i A[i*m + r] w[i] + beta * est[r]  (cublasSgemm with one column).
__global__ void lls_gemv_colmajor( const float* __restrict__ A, int n, int64_t m, const float* __restrict__ w, float alpha,
                                   float beta, float* __restrict__ est )
{
    const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if ( r >= m )
        return;
    float s = 0.f;
    for ( int i = 0; i < n; i++ )
        s += A[ (int64_t)i * m + r ] * w[ i ];
    est[ r ] = beta == 0.f ? alpha * s : alpha * s + beta * est[ r ];
}

// out[b] = sum of squares over a grid-stride share of data (FP64 partials per workgroup).
__global__ void __launch_bounds__( 256 ) sum_squares_partial( const float* __restrict__ data, int64_t n, double* __restrict__ out )
{
    __shared__ double red[ 256 ];
    double s = 0.0;
    for ( int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256 )
    {
        const double v = data[ i ];
        s += v * v;
    }
    red[ threadIdx.x ] = s;
    __syncthreads();
    for ( int k = 128; k > 0; k >>= 1 )
    {
        if ( threadIdx.x < k )
            red[ threadIdx.x ] += red[ threadIdx.x + k ];
        __syncthreads();
    }
    if ( threadIdx.x == 0 )
        out[ blockIdx.x ] = red[ 0 ];
}

// theta_0 = w / wL for every user (apsm_noma_detector.cu:321-329).
__global__ void apsm_lls_init_theta( const float* __restrict__ w, float* __restrict__ thOld, int L, float wL )
{
    int i = threadIdx.x;
    if ( i < L )
        thOld[ blockIdx.x * L + i ] = w[ i ] / wL;
}

// Gaussian quadratic form c^T K c over G complex dictionary entries (2G real rows), the
// Gaussian norm of computeNormRatio (apsm_noma_detector.cu:220-319). Thread i owns complex
// entry i; the j entries stream through LDS in tiles of 64. For entries i, j with
// s = a_i . b_j the four real-row kernel values are K(a_i,a_j) = K(b_i,b_j) = E(n - 2p),
// K(a_i,b_j) = E(n - 2s) and K(b_i,a_j) = E(n + 2s).
constexpr int kQuadThreads = 256;
constexpr int kQuadTile = 64;
__global__ void __launch_bounds__( kQuadThreads ) apsm_quad_form_partial( const float* __restrict__ Z, const float* __restrict__ Zn,
                                                                        const float* __restrict__ C, int64_t G, int N, float g2,
                                                                        double* __restrict__ part )
{
    const int L = 2 * N;
    extern __shared__ float sZ[]; // [kQuadTile][L] then [kQuadTile] norms, [kQuadTile][2] coefficients
    float* sN = sZ + kQuadTile * L;
    float* sC = sN + kQuadTile;
    __shared__ double red[ kQuadThreads ];
    const int64_t i = (int64_t)blockIdx.x * kQuadThreads + threadIdx.x;
    const bool ok = i < G;
    const float* zi = Z + ( ok ? i : 0 ) * L;
    const float ni = ok ? Zn[ i ] : 0.f;
    const float cai = ok ? C[ 2 * i ] : 0.f, cbi = ok ? C[ 2 * i + 1 ] : 0.f;
    float acc = 0.f;
    for ( int64_t j0 = 0; j0 < G; j0 += kQuadTile )
    {
        const int nj = (int)( G - j0 < kQuadTile ? G - j0 : kQuadTile );
        __syncthreads();
        for ( int e = threadIdx.x; e < nj * L; e += kQuadThreads )
            sZ[ e ] = Z[ j0 * L + e ];
        for ( int e = threadIdx.x; e < nj; e += kQuadThreads )
        {
            sN[ e ] = Zn[ j0 + e ];
            sC[ 2 * e ] = C[ 2 * ( j0 + e ) ];
            sC[ 2 * e + 1 ] = C[ 2 * ( j0 + e ) + 1 ];
        }
        __syncthreads();
        if ( ok )
            for ( int j = 0; j < nj; j++ )
            {
                const float* zj = sZ + j * L;
                float p = 0.f, s = 0.f;
                for ( int n = 0; n < N; n++ )
                {
                    const float xr = zi[ n ], xi = zi[ N + n ], zr = zj[ n ], zim = zj[ N + n ];
                    p += xr * zr + xi * zim;
                    s += xr * zim - xi * zr;
                }
                const float nn = ni + sN[ j ];
                const float E0 = gaussFromDist( nn - 2.f * p, g2 );
                const float Eab = gaussFromDist( nn - 2.f * s, g2 );
                const float Eba = gaussFromDist( nn + 2.f * s, g2 );
                const float caj = sC[ 2 * j ], cbj = sC[ 2 * j + 1 ];
                acc += ( cai * caj + cbi * cbj ) * E0 + cai * cbj * Eab + cbi * caj * Eba;
            }
    }
    red[ threadIdx.x ] = acc;
    __syncthreads();
    for ( int k = kQuadThreads / 2; k > 0; k >>= 1 )
    {
        if ( threadIdx.x < k )
            red[ threadIdx.x ] += red[ threadIdx.x + k ];
        __syncthreads();
    }
    if ( threadIdx.x == 0 )
        part[ blockIdx.x ] = red[ 0 ];
}

// Per-user EVM / SER / BER of soft estimates against the true symbols, with the rules of
// cpp/lib/noma/noma_train_history.cu:105-236 (decision regions for SER, Gray-demapped PAM
// levels for BER). One workgroup per user: y_true [U][n][2] (row stride ldT complex),
// y_hat [U][n][2] (row stride ldH complex); out[u*outStride + 0..2] = evm, ser, ber.
constexpr int kMetricThreads = 1024;
__device__ __forceinline__ bool componentSymbolError( float t, float h, float scale, float maxVal )
{
    return fabsf( t - h ) > scale && !( h > maxVal && t > maxVal - scale ) && !( h < -maxVal && t < -maxVal + scale );
}
__device__ __forceinline__ unsigned grayLevel( float x, int levels )
{
    int xi = (int)roundf( ( x + (float)( levels - 1 ) ) * 0.5f );
    xi = min( levels - 1, max( 0, xi ) );
    return (unsigned)( xi ^ ( xi >> 1 ) );
}
__global__ void __launch_bounds__( kMetricThreads ) apsm_history_metrics( const float* __restrict__ yTrue, int64_t ldT,
                                                                        const float* __restrict__ yHat, int64_t ldH, int64_t n,
                                                                        int bps, float scale, float* __restrict__ out,
                                                                        int outStride )
{
    const int u = blockIdx.x;
    const float2* t2 = reinterpret_cast<const float2*>( yTrue ) + (int64_t)u * ldT;
    const float2* h2 = reinterpret_cast<const float2*>( yHat ) + (int64_t)u * ldH;
    const int levels = bps == 1 ? 2 : 1 << ( bps / 2 );
    const float maxVal = (float)( ( 1 << ( bps / 2 ) ) - 1 ) * scale;
    const float rev = 1.f / scale;
    double se = 0.0;
    unsigned sym = 0, bits = 0;
    for ( int64_t i = threadIdx.x; i < n; i += kMetricThreads )
    {
        const float2 t = t2[ i ], h = h2[ i ];
        const float dr = t.x - h.x, di = t.y - h.y;
        se += (double)( dr * dr + di * di );
        if ( bps == 1 )
            sym += signbit( t.x ) != signbit( h.x );
        else
            sym += componentSymbolError( t.x, h.x, scale, maxVal ) || componentSymbolError( t.y, h.y, scale, maxVal );
        bits += __popc( grayLevel( t.x * rev, levels ) ^ grayLevel( h.x * rev, levels ) );
        if ( bps > 1 )
            bits += __popc( grayLevel( t.y * rev, levels ) ^ grayLevel( h.y * rev, levels ) );
    }
    __shared__ double rs[ kMetricThreads ];
    __shared__ unsigned rsym[ kMetricThreads ], rbit[ kMetricThreads ];
    rs[ threadIdx.x ] = se;
    rsym[ threadIdx.x ] = sym;
    rbit[ threadIdx.x ] = bits;
    __syncthreads();
    for ( int k = kMetricThreads / 2; k > 0; k >>= 1 )
    {
        if ( threadIdx.x < k )
        {
            rs[ threadIdx.x ] += rs[ threadIdx.x + k ];
            rsym[ threadIdx.x ] += rsym[ threadIdx.x + k ];
            rbit[ threadIdx.x ] += rbit[ threadIdx.x + k ];
        }
        __syncthreads();
    }
    if ( threadIdx.x == 0 )
    {
        out[ (int64_t)u * outStride + 0 ] = (float)sqrt( rs[ 0 ] / (double)n );
        out[ (int64_t)u * outStride + 1 ] = (float)( (double)rsym[ 0 ] / (double)n );
        out[ (int64_t)u * outStride + 2 ] = (float)( (double)rbit[ 0 ] / ( (double)n * bps ) );
    }
}

// ---------------------------------------------------------------------------------------
// Gaussian-variance heuristic (cpp/lib/apsm/apsm_noma_detector.cu:194-218): the 70th
// percentile of d_ij = sqrt(sum_n |x_i[n] - x_j[n]|) over all pairs i < j of pilots. The
// percentile is found without storing the T(T-1)/2 values: non-negative floats order as
// their bit patterns, so three histogram passes over the pairs fix the top 12, the next 12
// and the last 8 bits of the k-th smallest value. Tiles of 64 x 64 pilots sit in LDS;
// per-workgroup histograms in LDS are added into a global 64-bit histogram.
// ---------------------------------------------------------------------------------------
constexpr int kVhTile = 64;
constexpr int kVhThreads = 256;
constexpr int kVhBins = 4096;

__device__ __forceinline__ uint32_t vhBin( uint32_t bits, int pass, uint32_t prefix, bool& take )
{
    if ( pass == 0 )
    {
        take = true;
        return bits >> 20;
    }
    if ( pass == 1 )
    {
        take = ( bits >> 20 ) == prefix;
        return ( bits >> 8 ) & 0xFFFu;
    }
    take = ( bits >> 8 ) == prefix;
    return bits & 0xFFu;
}

__global__ void __launch_bounds__( kVhThreads ) apsm_variance_histogram( const float* __restrict__ rx, int N, int64_t T, int pass,
                                                                        uint32_t prefix, unsigned long long* __restrict__ hist )
{
    const int bi = blockIdx.y, bj = blockIdx.x;
    if ( bj < bi )
        return;
    extern __shared__ float sv[]; // [2][N][kVhTile] complex as float2
    float2* sa = reinterpret_cast<float2*>( sv );
    float2* sb = sa + N * kVhTile;
    __shared__ uint32_t h[ kVhBins ];
    for ( int b = threadIdx.x; b < kVhBins; b += kVhThreads )
        h[ b ] = 0u;
    const int64_t i0 = (int64_t)bi * kVhTile, j0 = (int64_t)bj * kVhTile;
    for ( int e = threadIdx.x; e < N * kVhTile; e += kVhThreads )
    {
        const int n = e / kVhTile, k = e % kVhTile;
        const float2* r = reinterpret_cast<const float2*>( rx ) + (int64_t)n * T;
        sa[ e ] = i0 + k < T ? r[ i0 + k ] : make_float2( 0.f, 0.f );
        sb[ e ] = j0 + k < T ? r[ j0 + k ] : make_float2( 0.f, 0.f );
    }
    __syncthreads();
    const int il = threadIdx.x % kVhTile;
    const int jb = ( threadIdx.x / kVhTile ) * ( kVhTile * kVhTile / kVhThreads );
    const int64_t i = i0 + il;
    for ( int jj = 0; jj < kVhTile * kVhTile / kVhThreads; jj++ )
    {
        const int jl = jb + jj;
        const int64_t j = j0 + jl;
        if ( i >= T || j >= T || j <= i )
            continue;
        float dist = 0.f;
        for ( int n = 0; n < N; n++ )
        {
            const float2 a = sa[ n * kVhTile + il ], b = sb[ n * kVhTile + jl ];
            dist += hypotf( a.x - b.x, a.y - b.y );
        }
        bool take;
        const uint32_t bin = vhBin( __float_as_uint( sqrtf( dist ) ), pass, prefix, take );
        if ( take )
            atomicAdd( &h[ bin ], 1u );
    }
    __syncthreads();
    for ( int b = threadIdx.x; b < kVhBins; b += kVhThreads )
        if ( h[ b ] )
            atomicAdd( &hist[ b ], (unsigned long long)h[ b ] );
}

__global__ void fill_zero( float* __restrict__ p, int64_t n )
{
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if ( i < n )
        p[ i ] = 0.f;
}

} // namespace apsm
