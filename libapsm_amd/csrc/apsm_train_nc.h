// Train sequencer with the norm constraint (normConstraint > 0): after every projection step the
// Gaussian coefficients of the whole dictionary are scaled down when
// n = sqrt(wG * c^T K c) exceeds normConstraint (cpp/lib/apsm/apsm_train.cu:138-143,168-256).
//
// One workgroup per user runs every step of one train call (the per-call sequencer's schedule:
// passes, full-window start, ramp, FIFO dictionary). Differences to the sequencer without the
// constraint, all forced by the scaling of c:
//   * theta is explicit (theta += sum_t beta_t x_t every step, held in thOld) instead of being
//     derived from the coefficients, since scaling c must not change it;
//   * besides f, the Gaussian part g_u = sum_j c_j K(j, u) of every window row is tracked
//     (FG), with its own window Gram (GramG), so a scaling by alpha updates
//     f_u -= (1 - alpha) wG g_u and g_u *= alpha without touching the dictionary;
//   * Q = c^T K c is tracked exactly across steps in FP64: a step adds
//     2 beta^T g + beta^T K_W beta, leaving rows remove 2 c_E^T (K c)_E - c_E^T K_EE c_E,
//     a scaling multiplies it by alpha^2.
#pragma once

#include "apsm_device.h"

namespace apsm {

struct NcArgs
{
    float* GramG; // [U][w2][w2] Gaussian window Gram (kernel values only)
    float* FG; // [U][w2] Gaussian part of f by slot (without wG)
    double* Q; // [U] c^T K c over the dictionary
    float nc; // normConstraint
};

// Block-wide sum of one float per thread (result in every thread); red holds kSeqWaves floats.
__device__ __forceinline__ float blockSum( float v, float* red )
{
    const int lane = threadIdx.x & ( kWave - 1 ), wave = threadIdx.x / kWave;
    v = waveSum( v );
    __syncthreads();
    if ( lane == 0 )
        red[ wave ] = v;
    __syncthreads();
    float s = 0.f;
    for ( int w = 0; w < kSeqWaves; w++ )
        s += red[ w ];
    return s;
}

__global__ void __launch_bounds__( kSeqThreads ) apsm_train_seq_nc( const float* __restrict__ Z, const float* __restrict__ Zn,
                                                                  const float* __restrict__ Y, float* __restrict__ C,
                                                                  float* __restrict__ Fg, float* __restrict__ Gram,
                                                                  float* __restrict__ thOld, float* __restrict__ theta,
                                                                  const SeqArgs a, const NcArgs nca )
{
    const int u = blockIdx.x;
    const int tid = threadIdx.x;
    const int lane = tid & ( kWave - 1 );
    const int wave = tid / kWave;
    const int N = a.N, L = a.L, w2 = a.w2;

    extern __shared__ float sm[];
    float* sF = sm; // [w2] f by slot
    float* sG = sF + w2; // [w2] Gaussian part g by slot
    float* sB = sG + w2; // [w2] beta by slot
    float* sTh = sB + w2; // [L] explicit theta
    float* sX = sTh + L; // [kNewMax][L]
    float* sXn = sX + kNewMax * L; // [kNewMax]
    float* sRed = sXn + kNewMax; // [kSeqWaves][2*kNewMax]
    float* sPart = sRed + kSeqWaves * 2 * kNewMax; // [kSeqWaves][kWave]
    float* sPartG = sPart + kSeqWaves * kWave; // [kSeqWaves][kWave]
    float* sSum = sPartG + kSeqWaves * kWave; // [kSeqWaves]
    __shared__ double sQ;
    __shared__ float sAlpha;

    const float* Yu = Y + u * a.ldYC;
    float* Cu = C + u * a.ldYC;
    float* Gu = Gram + (int64_t)u * w2 * w2;
    float* GGu = nca.GramG + (int64_t)u * w2 * w2;
    float* Fu = Fg + (int64_t)u * w2;
    float* FGu = nca.FG + (int64_t)u * w2;
    const float wL = a.wL, wG = a.wG, g2 = a.g2;

    for ( int i = tid; i < w2; i += kSeqThreads )
    {
        sF[ i ] = Fu[ i ];
        sG[ i ] = FGu[ i ];
    }
    for ( int i = tid; i < L; i += kSeqThreads )
        sTh[ i ] = thOld[ u * L + i ];
    if ( tid == 0 )
        sQ = nca.Q[ u ];
    __syncthreads();

    const int uBlocks = ( w2 + kWave - 1 ) / kWave;
    const int tParts = kSeqWaves / uBlocks;
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
            const int64_t winLo = s - W;
            const int64_t lo = s > a.H ? s - a.H : 0;

            if ( moved )
            {
                const int64_t loPrev = sPrev > a.H ? sPrev - a.H : 0;
                // ---- 1. rows leaving the dictionary ----
                if ( lo > loPrev )
                {
                    const int64_t e0 = ( loPrev - a.rowBase ) / 2, e1 = ( lo - a.rowBase ) / 2;
                    const int ne = (int)( e1 - e0 ); // <= kNewMax
                    // (K c) at the leaving rows over the pre-step dictionary [loPrev, sPrev)
                    float ka[ kNewMax ], kb[ kNewMax ];
#pragma unroll
                    for ( int i = 0; i < kNewMax; i++ )
                        ka[ i ] = kb[ i ] = 0.f;
                    const int64_t jEnd = ( sPrev - a.rowBase ) / 2;
                    for ( int64_t j = e0 + tid; j < jEnd; j += kSeqThreads )
                    {
                        const float* zj = Z + j * L;
                        const float ca = Cu[ 2 * j ], cb = Cu[ 2 * j + 1 ];
#pragma unroll
                        for ( int i = 0; i < kNewMax; i++ )
                        {
                            if ( i < ne )
                            {
                                float p = 0.f, sp = 0.f;
                                const float* ze = Z + ( e0 + i ) * L;
                                for ( int n = 0; n < N; n++ )
                                {
                                    float xr = ze[ n ], xi = ze[ N + n ], zr = zj[ n ], zi = zj[ N + n ];
                                    p += xr * zr + xi * zi;
                                    sp += xr * zi - xi * zr;
                                }
                                const float nn = Zn[ e0 + i ] + Zn[ j ];
                                const float E0 = gaussFromDist( nn - 2.f * p, g2 );
                                const float E1 = gaussFromDist( nn - 2.f * sp, g2 );
                                const float E2 = gaussFromDist( nn + 2.f * sp, g2 );
                                ka[ i ] += ca * E0 + cb * E1;
                                kb[ i ] += ca * E2 + cb * E0;
                            }
                        }
                    }
                    float dq = 0.f;
#pragma unroll
                    for ( int i = 0; i < kNewMax; i++ )
                        if ( i < ne )
                            dq += Cu[ 2 * ( e0 + i ) ] * ka[ i ] + Cu[ 2 * ( e0 + i ) + 1 ] * kb[ i ];
                    dq = blockSum( dq, sSum );
                    if ( tid == 0 )
                    {
                        // c_E^T K_EE c_E among the leaving rows
                        double self = 0.0;
                        for ( int i = 0; i < ne; i++ )
                            for ( int k = 0; k < ne; k++ )
                            {
                                const float* x = Z + ( e0 + i ) * L;
                                const float* z = Z + ( e0 + k ) * L;
                                float p = 0.f, sp = 0.f;
                                for ( int n = 0; n < N; n++ )
                                {
                                    p += x[ n ] * z[ n ] + x[ N + n ] * z[ N + n ];
                                    sp += x[ n ] * z[ N + n ] - x[ N + n ] * z[ n ];
                                }
                                const float nn = Zn[ e0 + i ] + Zn[ e0 + k ];
                                const double ca = Cu[ 2 * ( e0 + i ) ], cb = Cu[ 2 * ( e0 + i ) + 1 ];
                                const double ca2 = Cu[ 2 * ( e0 + k ) ], cb2 = Cu[ 2 * ( e0 + k ) + 1 ];
                                self += ( ca * ca2 + cb * cb2 ) * gaussFromDist( nn - 2.f * p, g2 )
                                    + ca * cb2 * gaussFromDist( nn - 2.f * sp, g2 ) + cb * ca2 * gaussFromDist( nn + 2.f * sp, g2 );
                            }
                        sQ = sQ - 2.0 * (double)dq + self;
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
                        sG[ slotOf( rv, w2 ) ] -= dFa;
                        sG[ slotOf( rv + 1, w2 ) ] -= dFb;
                    }
                    __syncthreads();
                }
                // ---- 2. entering rows: Gaussian part against the dictionary, Gram entries ----
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
                    const int64_t jHi = tc + nNew;
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
                            pa[ t ] += ca * E0 + cb * E1;
                            pb[ t ] += ca * E2 + cb * E0;
                            if ( inWin )
                            {
                                const int64_t rt = a.rowBase + 2 * ( tc + t );
                                const int sta = slotOf( rt, w2 ), stb = slotOf( rt + 1, w2 );
                                const float kaa = wL * p + wG * E0, kba = wL * sp + wG * E1, kab = -wL * sp + wG * E2;
                                Gu[ sta * w2 + sja ] = kaa;
                                Gu[ sja * w2 + sta ] = kaa;
                                Gu[ stb * w2 + sjb ] = kaa;
                                Gu[ sjb * w2 + stb ] = kaa;
                                Gu[ sta * w2 + sjb ] = kba;
                                Gu[ sjb * w2 + sta ] = kba;
                                Gu[ stb * w2 + sja ] = kab;
                                Gu[ sja * w2 + stb ] = kab;
                                GGu[ sta * w2 + sja ] = E0;
                                GGu[ sja * w2 + sta ] = E0;
                                GGu[ stb * w2 + sjb ] = E0;
                                GGu[ sjb * w2 + stb ] = E0;
                                GGu[ sta * w2 + sjb ] = E1;
                                GGu[ sjb * w2 + sta ] = E1;
                                GGu[ stb * w2 + sja ] = E2;
                                GGu[ sja * w2 + stb ] = E2;
                            }
                        }
                    }
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
                        float g = 0.f;
                        for ( int w = 0; w < kSeqWaves; w++ )
                            g += sRed[ w * 2 * kNewMax + tid ];
                        const float* xt = sX + t * L;
                        float lin = 0.f;
                        for ( int n = 0; n < N; n++ )
                        {
                            float xr = xt[ n ], xi = xt[ N + n ];
                            lin += part == 0 ? ( sTh[ n ] * xr + sTh[ N + n ] * xi ) : ( sTh[ n ] * xi - sTh[ N + n ] * xr );
                        }
                        const int64_t r = a.rowBase + 2 * ( tc + t ) + part;
                        sF[ slotOf( r, w2 ) ] = wL * lin + wG * g;
                        sG[ slotOf( r, w2 ) ] = g;
                    }
                    __syncthreads();
                }
            }
            // ---- 3. projection weights; beta^T g before the update ----
            float bg = 0.f;
            for ( int i = tid; i < w2; i += kSeqThreads )
            {
                const int64_t r = s - w2 + i;
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
                    bg += b * sG[ sl_ ];
                }
                sB[ sl_ ] = beta;
            }
            bg = blockSum( bg, sSum ); // includes the barrier that publishes sB
            // theta += sum_t beta_t x_t over the window rows
            for ( int d = tid; d < L; d += kSeqThreads )
            {
                float acc = 0.f;
                for ( int64_t r = winLo; r < winLo + nb; r++ )
                {
                    const float bt = sB[ slotOf( r, w2 ) ];
                    const float* xr = Z + ( ( r - a.rowBase ) >> 1 ) * L;
                    // row 2k is a = [Re; Im], row 2k+1 is b = [Im; -Re]
                    const float xd = ( ( r - a.rowBase ) & 1 ) == 0 ? xr[ d ] : ( d < N ? xr[ N + d ] : -xr[ d - N ] );
                    acc += bt * xd;
                }
                sTh[ d ] += acc;
            }
            // ---- 4. f_u += sum_t kappa(t,u) beta_t and g_u += sum_t K(t,u) beta_t ----
            {
                float acc = 0.f, accG = 0.f;
                const int us = myUB * kWave + lane;
                if ( myTP < tParts && us < w2 )
                {
                    const int tb = myTP * tPer;
                    const int te = tb + tPer < w2 ? tb + tPer : w2;
                    for ( int t = tb; t < te; t++ )
                    {
                        float bt = sB[ t ];
                        if ( bt != 0.f )
                        {
                            acc += Gu[ t * w2 + us ] * bt;
                            accG += GGu[ t * w2 + us ] * bt;
                        }
                    }
                }
                sPart[ wave * kWave + lane ] = acc;
                sPartG[ wave * kWave + lane ] = accG;
                __syncthreads();
                float bdg = 0.f;
                for ( int us2 = tid; us2 < w2; us2 += kSeqThreads )
                {
                    const int ub = us2 / kWave, ln = us2 % kWave;
                    float sum = 0.f, sumG = 0.f;
                    for ( int tp = 0; tp < tParts; tp++ )
                    {
                        sum += sPart[ ( tp * uBlocks + ub ) * kWave + ln ];
                        sumG += sPartG[ ( tp * uBlocks + ub ) * kWave + ln ];
                    }
                    sF[ us2 ] += sum;
                    sG[ us2 ] += sumG;
                    bdg += sB[ us2 ] * sumG;
                }
                bdg = blockSum( bdg, sSum );
                // ---- 5. norm constraint over the dictionary [lo, s) ----
                if ( tid == 0 )
                {
                    double q = sQ + 2.0 * (double)bg + (double)bdg;
                    double n = (double)wG * q;
                    if ( n > 0.0 )
                        n = sqrt( n );
                    float alpha = 1.f;
                    if ( n > (double)nca.nc )
                    {
                        alpha = (float)( (double)nca.nc / n );
                        q *= (double)alpha * (double)alpha;
                    }
                    sQ = q;
                    sAlpha = alpha;
                }
                __syncthreads();
                const float alpha = sAlpha;
                if ( alpha != 1.f )
                {
                    for ( int64_t j = ( lo - a.rowBase ) + tid; j < s - a.rowBase; j += kSeqThreads )
                        Cu[ j ] *= alpha;
                    for ( int i = tid; i < w2; i += kSeqThreads )
                    {
                        sF[ i ] -= ( 1.f - alpha ) * wG * sG[ i ];
                        sG[ i ] *= alpha;
                    }
                }
                __syncthreads();
            }
            sPrev = s;
        }
    }
    for ( int i = tid; i < w2; i += kSeqThreads )
    {
        Fu[ i ] = sF[ i ];
        FGu[ i ] = sG[ i ];
    }
    for ( int i = tid; i < L; i += kSeqThreads )
    {
        thOld[ u * L + i ] = sTh[ i ];
        theta[ u * L + i ] = sTh[ i ];
    }
    if ( tid == 0 )
        nca.Q[ u ] = sQ;
}

inline size_t seqNcLdsBytes( int w2, int L )
{
    return sizeof( float ) * ( 3 * (size_t)w2 + L + kNewMax * L + kNewMax + kSeqWaves * 2 * kNewMax + 2 * kSeqWaves * kWave + kSeqWaves );
}

} // namespace apsm
