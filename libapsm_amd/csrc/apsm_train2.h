// Blocked APSM train: the pilot stream is processed in blocks of projection steps.
//
// For one block (steps s_first .. s_last, entering rows [s_prev, s_last)):
//   a0   = s_first - W(s_first): rows below a0 take no further projection in this block
//          (frozen); rows in [a0, s_last) are the block's active rows.
//   apsm_entry_frozen  (whole GPU) for every entering complex sample t and every user:
//          P_t = wL*thF.x_t + wG*sum_{j frozen, in the dictionary when t enters} c_j K(j,t),
//          with thF = theta_0 + sum_{j < a0} c_j x_j.
//   apsm_active_gram   kappa(t,j) = wL x_t.x_j + wG K(t,j) for every pair of active complex
//          samples j <= t, as 3 floats per complex pair:
//            A = kappa(a_t,a_j) = kappa(b_t,b_j), Bv = kappa(a_t,b_j), Cv = kappa(b_t,a_j).
//   apsm_evict_corr    wG*sum_e c_e K(e,u) for the rows e leaving the dictionary at each
//          step and the window rows u that were already in the window.
//   apsm_train_seq2    one workgroup per user runs the block's steps: f of an entering row
//          is P_t + sum over active rows of c_j kappa(j,t); the sliding window Gram matrix
//          lives in registers (complex-block form) and f, beta, c of the active rows in LDS,
//          so a step touches global memory only through a prefetch of the next step's inputs.
// Same pre-step-state semantics as cpp/lib/apsm/apsm_train.cu:333-428; the evaluation order
// differs from the reference (it sums per step over the whole dictionary).
#pragma once

#include "apsm_device.h"

namespace apsm {

constexpr int kActMax = 512; // active complex samples per block (host guarantees)
constexpr int kEntryThreads = 256;

__host__ __device__ __forceinline__ int64_t triIdx( int64_t i ) { return i * ( i + 1 ) / 2; }

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

struct EntryArgs
{
    int N, L, U, q;
    int64_t rowBase; // absolute real row of buffer row 0
    int64_t newC0; // local complex index of the first entering sample
    int64_t sFirst; // absolute s of the block's first step
    int64_t H;
    int64_t a0; // absolute frozen boundary (real row, even)
    int64_t ldYC;
    int nNewC;
    float wL, wG, g2;
};

// One workgroup per entering complex sample.
template <int UMAX>
__global__ void __launch_bounds__( kEntryThreads ) apsm_entry_frozen( const float* __restrict__ Z, const float* __restrict__ Zn,
                                                                     const float* __restrict__ C, const float* __restrict__ thF,
                                                                     float* __restrict__ P, const EntryArgs a )
{
    __shared__ float sx[ 128 ];
    __shared__ float sred[ kEntryThreads / kWave ][ 2 * UMAX ];
    const int i = blockIdx.x;
    const int tid = threadIdx.x;
    const int N = a.N, L = a.L;
    const int64_t tc = a.newC0 + i;
    for ( int d = tid; d < L; d += kEntryThreads )
        sx[ d ] = Z[ tc * L + d ];
    __syncthreads();
    const float xn = Zn[ tc ];
    // entry step of this sample and the dictionary lower bound at that step
    const int64_t rt = a.rowBase + 2 * tc;
    const int64_t se = rt < a.sFirst ? a.sFirst : a.sFirst + a.q * ( ( rt - a.sFirst ) / a.q + 1 );
    const int64_t lo = se > a.H ? se - a.H : 0;
    const int64_t j0 = ( lo - a.rowBase ) / 2, j1 = ( a.a0 - a.rowBase ) / 2;
    float pa[ UMAX ], pb[ UMAX ];
#pragma unroll
    for ( int u = 0; u < UMAX; u++ )
        pa[ u ] = pb[ u ] = 0.f;
    for ( int64_t j = j0 + tid; j < j1; j += kEntryThreads )
    {
        float p, sp;
        pairDots( sx, Z + j * L, N, p, sp );
        const float nn = xn + Zn[ j ];
        const float E0 = gaussFromDist( nn - 2.f * p, a.g2 );
        const float E1 = gaussFromDist( nn - 2.f * sp, a.g2 );
        const float E2 = gaussFromDist( nn + 2.f * sp, a.g2 );
#pragma unroll
        for ( int u = 0; u < UMAX; u++ )
        {
            if ( u < a.U )
            {
                const float ca = C[ u * a.ldYC + 2 * j ], cb = C[ u * a.ldYC + 2 * j + 1 ];
                pa[ u ] = fmaf( ca, E0, fmaf( cb, E1, pa[ u ] ) );
                pb[ u ] = fmaf( ca, E2, fmaf( cb, E0, pb[ u ] ) );
            }
        }
    }
    const int lane = tid & ( kWave - 1 ), wave = tid / kWave;
#pragma unroll
    for ( int u = 0; u < UMAX; u++ )
    {
        float va = waveSum( pa[ u ] ), vb = waveSum( pb[ u ] );
        if ( lane == 0 )
        {
            sred[ wave ][ 2 * u ] = va;
            sred[ wave ][ 2 * u + 1 ] = vb;
        }
    }
    __syncthreads();
    if ( tid < 2 * a.U && tid < 2 * UMAX )
    {
        const int u = tid >> 1, part = tid & 1;
        float g = 0.f;
        for ( int w = 0; w < kEntryThreads / kWave; w++ )
            g += sred[ w ][ tid ];
        const float* th = thF + u * L;
        float lin = 0.f;
        for ( int n = 0; n < N; n++ )
        {
            const float xr = sx[ n ], xi = sx[ N + n ];
            lin += part == 0 ? ( th[ n ] * xr + th[ N + n ] * xi ) : ( th[ n ] * xi - th[ N + n ] * xr );
        }
        P[ ( (int64_t)u * a.nNewC + i ) * 2 + part ] = a.wL * lin + a.wG * g;
    }
}

struct GramArgs
{
    int N, L;
    int64_t c0; // local complex index of the first active sample
    int64_t nA; // active complex samples
    float wL, wG, g2;
};

// Gact[tri(i) + j] (3 floats) for active samples i >= j (relative to c0).
__global__ void apsm_active_gram( const float* __restrict__ Z, const float* __restrict__ Zn, float* __restrict__ Gact, const GramArgs a )
{
    const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t total = triIdx( a.nA );
    if ( idx >= total )
        return;
    int64_t i = (int64_t)( ( sqrt( 8.0 * (double)idx + 1.0 ) - 1.0 ) * 0.5 );
    while ( triIdx( i + 1 ) <= idx )
        i++;
    while ( triIdx( i ) > idx )
        i--;
    const int64_t j = idx - triIdx( i );
    const int64_t t = a.c0 + i, jj = a.c0 + j;
    float p, sp;
    pairDots( Z + t * a.L, Z + jj * a.L, a.N, p, sp );
    const float nn = Zn[ t ] + Zn[ jj ];
    const float E0 = gaussFromDist( nn - 2.f * p, a.g2 );
    const float E1 = gaussFromDist( nn - 2.f * sp, a.g2 );
    const float E2 = gaussFromDist( nn + 2.f * sp, a.g2 );
    Gact[ idx * 3 + 0 ] = a.wL * p + a.wG * E0;
    Gact[ idx * 3 + 1 ] = a.wL * sp + a.wG * E1;
    Gact[ idx * 3 + 2 ] = -a.wL * sp + a.wG * E2;
}

struct EvArgs
{
    int N, L, U, q, w2, nSteps;
    int64_t rowBase, sPrev, sFirst, H, ldYC;
    float wG, g2;
};

// EvF[u][k][2*pc + part]: Gaussian share of the rows leaving the dictionary at step k, for
// the window row at position 2*pc + part (row s_k - W_k + 2*pc + part) if it entered before
// step k; zero otherwise. One thread per (step, complex window position).
__global__ void apsm_evict_corr( const float* __restrict__ Z, const float* __restrict__ Zn, const float* __restrict__ C,
                                 float* __restrict__ EvF, const EvArgs a )
{
    const int Wc = a.w2 / 2;
    const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if ( idx >= (int64_t)a.nSteps * Wc )
        return;
    const int kk = (int)( idx / Wc ), pc = (int)( idx % Wc );
    const int64_t s = a.sFirst + (int64_t)kk * a.q;
    const int64_t sp = kk == 0 ? a.sPrev : s - a.q;
    const int64_t W = s < a.w2 ? s : a.w2;
    const int64_t lo = s > a.H ? s - a.H : 0, loP = sp > a.H ? sp - a.H : 0;
    const int64_t r = s - W + 2 * pc; // absolute real row (a-row of the complex sample)
    const bool live = ( 2 * pc < W ) && ( r < sp ) && ( lo > loP );
    for ( int u = 0; u < a.U; u++ )
    {
        float dFa = 0.f, dFb = 0.f;
        if ( live )
        {
            const int64_t v = ( r - a.rowBase ) / 2;
            for ( int64_t e = ( loP - a.rowBase ) / 2; e < ( lo - a.rowBase ) / 2; e++ )
            {
                float p, sp2;
                pairDots( Z + v * a.L, Z + e * a.L, a.N, p, sp2 );
                const float nn = Zn[ v ] + Zn[ e ];
                const float E0 = gaussFromDist( nn - 2.f * p, a.g2 );
                const float E1 = gaussFromDist( nn - 2.f * sp2, a.g2 );
                const float E2 = gaussFromDist( nn + 2.f * sp2, a.g2 );
                const float ca = C[ u * a.ldYC + 2 * e ], cb = C[ u * a.ldYC + 2 * e + 1 ];
                dFa += ca * E0 + cb * E1;
                dFb += ca * E2 + cb * E0;
            }
        }
        float* o = EvF + ( (int64_t)u * a.nSteps + kk ) * a.w2;
        o[ 2 * pc ] = a.wG * dFa;
        o[ 2 * pc + 1 ] = a.wG * dFb;
    }
}

struct Seq2Args
{
    int N, L, U, w2, Wc, q, nSteps, hasEv;
    int nParts; // t-slot parts (threads per u-slot)
    int ramp; // 1 if a0 == 0 (the window may still be growing)
    int a0modW; // a0 mod w2 (real-row slot of the first active row)
    int sFirstRel; // s of the first step minus a0
    int nNewC, nActC; // entering / active complex samples
    int newRel0; // relative complex index of the first entering sample
    int nFr; // complex samples leaving the active set at the block end (thF update)
    int pkFloats; // floats in one step package
    int64_t ldYC; // user stride of Y and C
    float wL, wG, eB;
};

// Register-resident window Gram. Thread tid owns u-slot v = tid % Wc and t-slots
// [part*TP, part*TP + TP) with part = tid / Wc; per (t, v) it keeps
// A = kappa(a_t,a_v), B = kappa(a_t,b_v), C = kappa(b_t,a_v).
// All row indices are relative to the block's first active row a0 (even), so they fit in
// 32 bits; slots follow the absolute row (a0modW + rel) % w2, as the f array does.
// Pointers are offset by the host: Z, Zn to the first active complex sample; Y, C to the
// first active real row of user 0.
// Step package in LDS (double-buffered, filled one step ahead):
//   [0, NEWPER*3*nAct)       Gram rows of the entering samples
//   [+0, +2*NEWPER)          their frozen shares P
//   [+0, +w2)                eviction shares of the window rows
template <int TP, int NEWPER>
__global__ void __launch_bounds__( kSeqThreads ) apsm_train_seq2( const float* __restrict__ Z, const float* __restrict__ Zn,
                                                                const float* __restrict__ Y, float* __restrict__ C,
                                                                float* __restrict__ Fg, float* __restrict__ thF,
                                                                const float* __restrict__ P, const float* __restrict__ Gact,
                                                                const float* __restrict__ EvF, const Seq2Args a )
{
    constexpr int kRowRegs = ( NEWPER * 3 * kActMax + kSeqThreads - 1 ) / kSeqThreads;
    const int u = blockIdx.x;
    const int tid = threadIdx.x, lane = tid & ( kWave - 1 ), wave = tid / kWave;
    const int Wc = a.Wc, w2 = a.w2, L = a.L;
    const int v = tid % Wc, part = tid / Wc;
    const bool gThread = part < a.nParts;
    const int tBase = part * TP;
    const int nAct = a.nActC;
    const int cMod = a.a0modW >> 1; // complex slot of relative complex index 0
    const int rowFloats = 3 * nAct;
    const int pkRows = NEWPER * rowFloats;

    extern __shared__ float sm[];
    float* sF = sm; // [w2]
    float* sB = sF + w2; // [2 * kSeqThreads]
    float* sC = sB + 2 * kSeqThreads; // [2 * kActMax]
    float* sY = sC + 2 * kActMax; // [2 * kActMax]
    float* sZn = sY + 2 * kActMax; // [kActMax]
    float* sRedE = sZn + kActMax; // [kSeqWaves][2 * kNewMax]
    float* sRedF = sRedE + kSeqWaves * 2 * kNewMax; // [2 * kSeqThreads]
    float* sPk = sRedF + 2 * kSeqThreads; // [2][pkFloats]
    float* sCol = sPk + 2 * a.pkFloats; // [NEWPER][3][kSeqThreads]: Gram column of each entering sample by t-slot

    const float* Yu = Y + u * a.ldYC;
    float* Cu = C + u * a.ldYC;
    const float* Pu = P + (int64_t)u * a.nNewC * 2;
    const float* EvU = EvF + (int64_t)u * a.nSteps * w2;

    for ( int i = tid; i < w2; i += kSeqThreads )
        sF[ i ] = Fg[ (int64_t)u * w2 + i ];
    for ( int i = tid; i < 2 * kSeqThreads; i += kSeqThreads )
        sB[ i ] = 0.f;
    for ( int i = tid; i < 2 * nAct; i += kSeqThreads )
    {
        sC[ i ] = Cu[ i ];
        sY[ i ] = Yu[ i ];
    }
    for ( int i = tid; i < nAct; i += kSeqThreads )
        sZn[ i ] = Zn[ i ];

    // window Gram of the rows already in the window (active rows entered before the block);
    // the row in complex slot x, newest first from kLast, is kLast - ((pLast - x) mod Wc)
    float gA[ TP ], gB[ TP ], gC[ TP ];
    {
        const int kLast = a.newRel0 - 1;
        const int pLast = ( ( cMod + kLast ) % Wc + Wc ) % Wc;
        int dv = pLast - v;
        dv += dv < 0 ? Wc : 0;
        const int rv = kLast - dv;
        // branch-free: invalid pairs read entry 0 and are scaled by 0
#pragma unroll
        for ( int j = 0; j < TP; j++ )
        {
            const int tau = tBase + j;
            int dt = pLast - tau;
            dt += dt < 0 ? Wc : 0;
            const int rt = kLast - dt;
            const bool ok = gThread && tau < Wc && kLast >= 0 && rt >= 0 && rv >= 0;
            const bool ord = rt >= rv;
            const int hi = ok ? ( ord ? rt : rv ) : 0, lo = ok ? ( ord ? rv : rt ) : 0;
            const float* g = Gact + ( hi * ( hi + 1 ) / 2 + lo ) * 3;
            const float m = ok ? 1.f : 0.f;
            gA[ j ] = m * g[ 0 ];
            gB[ j ] = m * g[ ord ? 1 : 2 ];
            gC[ j ] = m * g[ ord ? 2 : 1 ];
            if ( ( j & 3 ) == 3 )
                __builtin_amdgcn_sched_barrier( 0 );
        }
    }

    float pfRow[ kRowRegs ];
    float pfP = 0.f, pfE = 0.f;
    auto pkLoad = [ & ]( int kk ) {
        const int c0 = a.newRel0 + kk * NEWPER;
        const float* rowBase = Gact + ( c0 * ( c0 + 1 ) / 2 ) * 3; // row of sample c0
#pragma unroll
        for ( int r = 0; r < kRowRegs; r++ )
        {
            const int e = tid + r * kSeqThreads;
            float val = 0.f;
            if ( e < pkRows )
            {
                // rows of consecutive samples c0, c0+1 are contiguous in the triangle, but the
                // package pads each to rowFloats
                const int i = NEWPER == 1 ? 0 : e / rowFloats;
                const int off = e - i * rowFloats;
                const int t = c0 + i;
                if ( off < 3 * ( t + 1 ) )
                    val = rowBase[ i * ( 3 * c0 + 3 ) + off ];
            }
            pfRow[ r ] = val;
        }
        pfP = tid < 2 * NEWPER ? Pu[ kk * NEWPER * 2 + tid ] : 0.f;
        pfE = ( a.hasEv && tid < w2 ) ? EvU[ kk * w2 + tid ] : 0.f;
    };
    auto pkStore = [ & ]( float* dst ) {
#pragma unroll
        for ( int r = 0; r < kRowRegs; r++ )
        {
            const int e = tid + r * kSeqThreads;
            if ( e < pkRows )
                dst[ e ] = pfRow[ r ];
        }
        if ( tid < 2 * NEWPER )
            dst[ pkRows + tid ] = pfP;
        if ( tid < w2 )
            dst[ pkRows + 2 * NEWPER + tid ] = pfE;
    };
    pkLoad( 0 );
    pkStore( sPk );
    __syncthreads();

    const int pkF = pkRows + 2 * NEWPER + w2;
    for ( int kk = 0; kk < a.nSteps; kk++ )
    {
        const float* pk = sPk + ( kk & 1 ) * pkF;
        if ( kk + 1 < a.nSteps )
            pkLoad( kk + 1 );
        const int sRel = a.sFirstRel + kk * a.q; // s - a0
        const int spRel = sRel - a.q;
        const int W = a.ramp ? ( sRel < w2 ? sRel : w2 ) : w2;
        const int c0 = a.newRel0 + kk * NEWPER; // first entering complex sample
        int sBase = ( a.a0modW + sRel ) % w2; // slot of row s - w2 (== slot of row s)

        // 1. rows leaving the dictionary: subtract their share from rows already in the window
        if ( a.hasEv && tid < W )
        {
            const int r = sRel - W + tid;
            if ( r < spRel )
            {
                int sl = sBase - W + tid;
                sl += sl < 0 ? w2 : 0;
                sF[ sl ] -= pk[ pkRows + 2 * NEWPER + tid ];
            }
        }
        // 2. entering samples: share of the active rows with non-zero c, and Gram registers
        {
            float ea[ NEWPER ], eb[ NEWPER ];
#pragma unroll
            for ( int i = 0; i < NEWPER; i++ )
                ea[ i ] = eb[ i ] = 0.f;
            if ( tid < c0 )
            {
                const float ca = sC[ 2 * tid ], cb = sC[ 2 * tid + 1 ];
#pragma unroll
                for ( int i = 0; i < NEWPER; i++ )
                {
                    const float* g = pk + i * rowFloats + 3 * tid;
                    ea[ i ] = ca * g[ 0 ] + cb * g[ 1 ];
                    eb[ i ] = ca * g[ 2 ] + cb * g[ 0 ];
                }
            }
#pragma unroll
            for ( int i = 0; i < NEWPER; i++ )
            {
                const float va = waveSum( ea[ i ] ), vb = waveSum( eb[ i ] );
                if ( lane == 0 )
                {
                    sRedE[ wave * 2 * kNewMax + 2 * i ] = va;
                    sRedE[ wave * 2 * kNewMax + 2 * i + 1 ] = vb;
                }
            }
        }
        // Gram entries of each entering sample k against the window as it stands after the
        // step (rows entering later in the same step included): the row in complex slot x is
        // kRef - ((pRef - x) mod Wc) with kRef the last entering sample. kappa(k, r) comes from
        // the package row of max(k, r) at the position of min(k, r).
        {
            const int kRef = c0 + NEWPER - 1;
            const int pRef = ( cMod + kRef ) % Wc;
#pragma unroll
            for ( int i = 0; i < NEWPER; i++ )
            {
                const int k = c0 + i;
                const int p = ( cMod + k ) % Wc;
                float* col = sCol + i * 3 * kSeqThreads;
                // column: sCol[i][.][tau] = (kappa(a_tau,a_k), kappa(a_tau,b_k), kappa(b_tau,a_k))
                if ( tid < a.nParts * TP )
                {
                    int dt = pRef - tid;
                    dt += dt < 0 ? Wc : 0;
                    const int rt = kRef - dt;
                    const bool ok = tid < Wc && rt >= 0;
                    const bool later = ok && rt > k; // entering after k in this step
                    const float* g = later ? pk + ( rt - c0 ) * rowFloats + 3 * k : pk + i * rowFloats + 3 * ( ok ? rt : 0 );
                    const float m = ok ? 1.f : 0.f;
                    col[ tid ] = m * g[ 0 ];
                    col[ kSeqThreads + tid ] = m * ( later ? g[ 1 ] : g[ 2 ] );
                    col[ 2 * kSeqThreads + tid ] = m * ( later ? g[ 2 ] : g[ 1 ] );
                }
                // row p: the threads whose t-slots contain p take kappa(k, v)
                const int jj = p - tBase;
                if ( gThread && jj >= 0 && jj < TP )
                {
                    int dv = pRef - v;
                    dv += dv < 0 ? Wc : 0;
                    const int rv = kRef - dv;
                    const bool ok = rv >= 0;
                    const bool later = rv > k;
                    const float* g = later ? pk + ( rv - c0 ) * rowFloats + 3 * k : pk + i * rowFloats + 3 * ( ok ? rv : 0 );
                    const float m = ok ? 1.f : 0.f;
                    const float A = m * g[ 0 ];
                    const float Bv = m * ( later ? g[ 2 ] : g[ 1 ] );
                    const float Cv = m * ( later ? g[ 1 ] : g[ 2 ] );
                    // per-lane select with static register indices; the empty asm keeps the
                    // compiler from turning this into a dynamically indexed vector insert
#pragma unroll
                    for ( int j = 0; j < TP; j++ )
                    {
                        float xa = j == jj ? A : gA[ j ];
                        float xb = j == jj ? Bv : gB[ j ];
                        float xc = j == jj ? Cv : gC[ j ];
                        asm volatile( "" : "+v"( xa ), "+v"( xb ), "+v"( xc ) );
                        gA[ j ] = xa;
                        gB[ j ] = xb;
                        gC[ j ] = xc;
                    }
                }
            }
        }
        __syncthreads();
        // column p of each entering sample: the thread holding u-slot p takes its t-slots
#pragma unroll
        for ( int i = 0; i < NEWPER; i++ )
        {
            const int p = ( cMod + c0 + i ) % Wc;
            if ( gThread && v == p )
            {
                const float* col = sCol + i * 3 * kSeqThreads + tBase;
#pragma unroll
                for ( int j = 0; j < TP; j++ )
                {
                    gA[ j ] = col[ j ];
                    gB[ j ] = col[ kSeqThreads + j ];
                    gC[ j ] = col[ 2 * kSeqThreads + j ];
                }
            }
        }
        // 3. projection weights over the window (entering rows complete their f first)
        if ( tid < w2 )
        {
            const int r = sRel - w2 + tid; // the last w2 rows cover every slot once
            int sl = sBase + tid;
            sl -= sl >= w2 ? w2 : 0;
            float beta = 0.f;
            if ( r >= sRel - W )
            {
                float f;
                if ( r >= spRel )
                {
                    const int e2 = r - spRel; // 2*entering index + part
                    float acc = pk[ pkRows + e2 ];
                    for ( int w = 0; w < kSeqWaves; w++ )
                        acc += sRedE[ w * 2 * kNewMax + e2 ];
                    f = acc;
                    sF[ sl ] = f;
                }
                else
                    f = sF[ sl ];
                const float res = sY[ r ] - f;
                float b = 0.f;
                if ( res > a.eB )
                    b = res - a.eB;
                else if ( res < -a.eB )
                    b = res + a.eB;
                b /= (float)W;
                b /= ( a.wL * sZn[ r >> 1 ] + a.wG );
                beta = b;
                sC[ r ] += b;
            }
            sB[ sl ] = beta; // sl = 2*complex slot + part
        }
        __syncthreads();
        // 4. f_v += sum_t kappa(t, v) beta_t from the register Gram
        if ( gThread )
        {
            float fa = 0.f, fb = 0.f;
            const float2* sB2 = reinterpret_cast<const float2*>( sB ) + tBase;
#pragma unroll
            for ( int j = 0; j < TP; j++ )
            {
                const float2 bt = sB2[ j ];
                fa = fmaf( bt.x, gA[ j ], fmaf( bt.y, gC[ j ], fa ) );
                fb = fmaf( bt.x, gB[ j ], fmaf( bt.y, gA[ j ], fb ) );
                if ( ( j & 3 ) == 3 )
                    __builtin_amdgcn_sched_barrier( 0 ); // bound the loads in flight (registers)
            }
            sRedF[ 2 * tid ] = fa;
            sRedF[ 2 * tid + 1 ] = fb;
        }
        __syncthreads();
        if ( tid < w2 )
        {
            const int cv = tid >> 1, pt = tid & 1;
            float acc = 0.f;
            for ( int t = 0; t < a.nParts; t++ )
                acc += sRedF[ 2 * ( t * Wc + cv ) + pt ];
            sF[ tid ] += acc;
        }
        if ( kk + 1 < a.nSteps )
            pkStore( sPk + ( ( kk + 1 ) & 1 ) * pkF );
        __syncthreads();
    }

    // write back f of the window and c of the active rows; thF += c_j x_j of rows leaving the
    // active set (they take no further projection)
    for ( int i = tid; i < w2; i += kSeqThreads )
        Fg[ (int64_t)u * w2 + i ] = sF[ i ];
    for ( int i = tid; i < 2 * nAct; i += kSeqThreads )
        Cu[ i ] = sC[ i ];
    const int chunks = kSeqThreads / L;
    {
        const int d = tid % L, ch = tid / L;
        if ( ch < chunks )
        {
            float acc = 0.f;
            for ( int j = ch; j < a.nFr; j += chunks )
            {
                const float* zj = Z + (int64_t)j * L;
                const float bd = d < a.N ? zj[ a.N + d ] : -zj[ d - a.N ];
                acc += sC[ 2 * j ] * zj[ d ] + sC[ 2 * j + 1 ] * bd;
            }
            sRedF[ ch * L + d ] = acc;
        }
    }
    __syncthreads();
    if ( tid < L )
    {
        float acc = thF[ u * L + tid ];
        for ( int ch = 0; ch < chunks; ch++ )
            acc += sRedF[ ch * L + tid ];
        thF[ u * L + tid ] = acc;
    }
}

} // namespace apsm
