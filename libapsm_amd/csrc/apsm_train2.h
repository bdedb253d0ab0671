// Blocked APSM train: the pilot stream is processed in blocks of projection steps.
//
// For one block (steps s_first .. s_last, entering rows [s_prev, s_last)):
//   a0   = s_first - W(s_first): rows below a0 take no further projection in this block
//          (frozen); rows in [a0, s_last) are the block's active rows.
// The first three run as roles of one launch (apsm_block_pre), the fourth after it:
//   entry share        (whole GPU) for every entering complex sample t and every user:
//          P_t = wL*thF.x_t + wG*sum_{j frozen, in the dictionary when t enters} c_j K(j,t),
//          with thF = theta_0 + sum_{j < a0} c_j x_j.
//   active Gram        kappa(t,j) = wL x_t.x_j + wG K(t,j) for every pair of active complex
//          samples j <= t, as 3 floats per complex pair:
//            A = kappa(a_t,a_j) = kappa(b_t,b_j), Bv = kappa(a_t,b_j), Cv = kappa(b_t,a_j).
//   eviction shares    wG*sum_e c_e K(e,u) for the rows e leaving the dictionary at each
//          step and the window rows u that were already in the window.
//   apsm_train_seq2    one workgroup per user runs the block's steps: f of an entering row
//          is P_t + sum over active rows of c_j kappa(j,t); the sliding window Gram matrix
//          lives in registers (complex-block form) and f, beta, c of the active rows in LDS,
//          so a step touches global memory only through a prefetch of the next step's inputs.
// Same pre-step-state semantics as cpp/lib/apsm/apsm_train.cu:333-428; the evaluation order
// differs from the reference (it sums per step over the whole dictionary).
#pragma once

#include "apsm_device.h"
#include "apsm_detect_launch.h"

namespace apsm {

constexpr int kActMax = 512; // active complex samples per block (host guarantees)
constexpr int kMaxParts = 32; // t-slot parts of the register Gram (1024 / Wc for Wc >= 32)
constexpr int kRedF = 2048; // per-part f partial slots (>= PMAX * Wc of every instantiation)
// Thread roles in the sequencer's entering phase: [0, w2) reduce f (w2 <= 320), [kColBase,
// kColBase + nParts*TP) stage the Gram column (<= 192 slots), the last wave takes the active
// rows' share of the entering f (c0 < kActMax).
constexpr int kColBase = 320;

// Parts (threads per u-slot) at most for a register-Gram run of TP t-slots: the launchers
// instantiate the sequencer with this PMAX, and a window of Wc complex slots fits when
// Wc <= TP * PMAX (the host checks nParts <= PMAX and nParts * TP >= Wc).
__host__ __device__ constexpr int seq2PmaxOfTp( int tp )
{
    return tp == 2 ? 24 : tp == 4 ? 16 : tp == 7 ? 12 : tp == 10 ? 10 : tp == 16 ? 8 : 6;
}

// Floats per step-package buffer: the Gram rows (<= NEWPER * 3 * kActMax, a multiple of 256)
// and P with the eviction shares (2 * NEWPER + w2 floats, w2 <= 2 * TP * PMAX, copied in
// 64-float chunks). Up to TP = 10 (at most 64 registers), rounded up to 256 floats, so that two
// workgroups fit in a CU's LDS. Beyond, where the registers allow one workgroup per CU anyway,
// the rows plus a whole kSeqThreads-float spare, in kSeqThreads units (3072 floats at TP = 27:
// measured 1 % faster there than the tight 2048).
__host__ __device__ constexpr int seq2PkCap( int newPer, int tp, int pmax )
{
    return tp <= 10 ? ( newPer * 3 * kActMax + ( ( 2 * newPer + 2 * tp * pmax + 63 ) & ~63 ) + 255 ) & ~255
                    : ( newPer * 3 * kActMax + 2 * newPer + 2 * kSeqThreads - 1 ) / kSeqThreads * kSeqThreads;
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
    int64_t ldYC; // user stride of Y and C
    float wL, wG, eB;
};

// LDS-only workgroup barrier: waits for this thread's LDS operations, not for outstanding
// global loads (the step packages stream in asynchronously across barriers).
__device__ __forceinline__ void ldsBarrier() { asm volatile( "s_waitcnt lgkmcnt(0)\n\ts_barrier" ::: "memory" ); }

// Asynchronous global -> LDS copy of one dword per lane (lane i writes ldsWaveBase + 4*i).
// M0 carries the wave's LDS destination (one wait state between the M0 write and the load);
// the sequencer kernel uses M0 for nothing else.
__device__ __forceinline__ void dmaDword( const float* src, uint32_t ldsWaveBase )
{
    asm volatile( "s_mov_b32 m0, %0\n\ts_nop 0\n\tglobal_load_lds_dword %1, off" ::"s"( ldsWaveBase ), "v"( src ) : "memory" );
}

// The same with 16 bytes per lane from a wave-uniform base plus a per-lane 32-bit byte offset
// (lane i writes ldsWaveBase + 16*i .. +16; base + byteOff 16-byte aligned): a quarter of the
// copy instructions for the same bytes (the issue cost is per instruction), and the load adds
// the offset, so no 64-bit address arithmetic runs on the vector ALU
__device__ __forceinline__ void dmaDwordx4( const float* base, uint32_t byteOff, uint32_t ldsWaveBase )
{
    asm volatile( "s_mov_b32 m0, %0\n\ts_nop 0\n\tglobal_load_lds_dwordx4 %1, %2" ::"s"( ldsWaveBase ), "v"( byteOff ), "s"( base ) : "memory" );
}

// threadIdx.x behind an empty asm: indices derived from it inside the step loop are recomputed
// each step (a few integer operations) instead of being hoisted and kept in registers across
// the loop, for which the register window Gram (3 TP of the 128 registers) leaves no room
__device__ __forceinline__ int tidInLoop()
{
    int t = threadIdx.x;
    asm volatile( "" : "+v"( t ) );
    return t;
}

template <int N>
__device__ __forceinline__ void waitVm()
{
    asm volatile( "s_waitcnt vmcnt(%0)" ::"n"( N ) : "memory" );
}

// s_waitcnt vmcnt(min(n, 15)) for a wave-uniform n (a smaller count only waits longer)
__device__ __forceinline__ void waitVmN( int n )
{
    switch ( n )
    {
    case 0: waitVm<0>(); break;
    case 1: waitVm<1>(); break;
    case 2: waitVm<2>(); break;
    case 3: waitVm<3>(); break;
    case 4: waitVm<4>(); break;
    case 5: waitVm<5>(); break;
    case 6: waitVm<6>(); break;
    case 7: waitVm<7>(); break;
    case 8: waitVm<8>(); break;
    case 9: waitVm<9>(); break;
    case 10: waitVm<10>(); break;
    case 11: waitVm<11>(); break;
    case 12: waitVm<12>(); break;
    case 13: waitVm<13>(); break;
    case 14: waitVm<14>(); break;
    default: waitVm<15>(); break;
    }
}

// Register-resident window Gram. Thread tid owns u-slot v = tid % Wc and t-slots
// [part*TP, part*TP + TP) with part = tid / Wc; per (t, v) it keeps
// A = kappa(a_t,a_v), B = kappa(a_t,b_v), C = kappa(b_t,a_v).
// All row indices are relative to the block's first active row a0 (even), so they fit in
// 32 bits; slots follow the absolute row (a0modW + rel) % w2, as the f array does.
// Pointers are offset by the host: Z, Zn to the first active complex sample; Y, C to the
// first active real row of user 0.
// Step package (four buffers in LDS, copied by global->LDS DMA during the product three steps
// ahead):
//   [0, pkRows)              Gram rows of the entering samples (row i at i*rowFloats, rowFloats =
//                            3*nAct rounded up to 4; pkRows = NEWPER*rowFloats rounded up to 256)
//   [+0, +2*NEWPER)          their frozen shares P
//   [+0, +w2)                eviction shares of the window rows
template <int TP, int NEWPER, int PMAX>
__device__ __forceinline__ void seq2Body( const float* __restrict__ Z, const float* __restrict__ Zn, const float* __restrict__ Y,
                                          float* __restrict__ C, float* __restrict__ Fg, float* __restrict__ thF,
                                          const float* __restrict__ P, const float* __restrict__ Gact, const float* __restrict__ EvF,
                                          const Seq2Args& a, const int u )
{
    static_assert( PMAX == seq2PmaxOfTp( TP ), "sequencer instantiated with another PMAX than its launchers assume" );
    constexpr int kPkCap = seq2PkCap( NEWPER, TP, PMAX ); // floats per package buffer
    constexpr int kPkBufs = 4; // package buffers: step k's package is issued at step k - 3
    constexpr int TPA = ( TP + 3 ) & ~3; // column segment per part, padded for float4 reads
    constexpr int kColStride = kMaxParts * TPA; // floats per column array (A, B or C)
    // the wave index as a scalar: everything derived from it stays in scalar registers (the
    // vector registers are all taken by the window Gram and the step's working set)
    const int tid = threadIdx.x, lane = tid & ( kWave - 1 ), wave = __builtin_amdgcn_readfirstlane( tid / kWave );
    const int Wc = a.Wc, w2 = a.w2, L = a.L;
    const int v = tid % Wc, part = tid / Wc;
    const bool gThread = part < a.nParts;
    const int tBase = part * TP;
    const int nAct = a.nActC;
    const int cMod = a.a0modW >> 1; // complex slot of relative complex index 0
    const int rowFloats = ( 3 * nAct + 3 ) & ~3; // package row stride: whole 16-byte units
    const int pkRows = ( NEWPER * rowFloats + 4 * kWave - 1 ) & ~( 4 * kWave - 1 ); // rows in whole 16-byte-lane copies

    extern __shared__ float sm[];
    float* sPk = sm; // [kPkBufs][kPkCap] (first, so the DMA destinations are 256-byte aligned)
    float* sF = sPk + kPkBufs * kPkCap; // [w2]
    float* sB = sF + w2; // [2 * kSeqThreads]
    float* sC = sB + 2 * kSeqThreads; // [2 * kActMax]
    float* sY = sC + 2 * kActMax; // [2 * kActMax]
    float* sInv = sY + 2 * kActMax; // [kActMax]: 1 / (wL |x|^2 + wG)
    float* sRedE = sInv + kActMax; // [2 * kNewMax] used: the entering samples' active-row shares
    float* sRedF = sRedE + kSeqWaves * 2 * kNewMax; // [2 * kRedF]: per-part f partials (zeros beyond nParts)
    float* sCol = sRedF + 2 * kRedF; // [NEWPER][3][kColStride]

    const float* Yu = Y + u * a.ldYC;
    float* Cu = C + u * a.ldYC;
    const float* Pu = P + (int64_t)u * a.nNewC * 2;
    const float* EvU = EvF + (int64_t)u * a.nSteps * w2;
    const int gEnd4Off = gramRow( nAct ) - 4; // last 16-byte unit of the block's Gram rows

    // issue the package of step kk into buffer kk % kPkBufs. Waves 0..3 (one per SIMD; they take
    // the product but do not stage) copy it after their product, while the staging waves
    // finish: wave w copies chunks w, w + kIssueWaves, ... (myChunks). Chunks [0, nX4) are the
    // Gram rows, 256 floats each (16 bytes per lane); the rest are P and the eviction shares,
    // 64 floats each.
    constexpr int kIssueWaves = 4;
    const int nX4 = pkRows / ( 4 * kWave );
    const int nChunks = nX4 + ( 2 * NEWPER + w2 + kWave - 1 ) / kWave;
    // wave-uniform (readfirstlane): the copy loop and the vmcnt waits below branch on scalars
    const int myChunks = __builtin_amdgcn_readfirstlane( wave < kIssueWaves && wave < nChunks ? ( nChunks - wave + kIssueWaves - 1 ) / kIssueWaves : 0 );
    // LDS byte address of the package buffers as a 32-bit scalar: the destinations below are
    // integer offsets from it (a generic pointer here costs a 64-bit register pair, which at
    // the kernel's 128 registers was spilled and reloaded, with a full vmcnt wait, every step)
    const uint32_t pkLds = __builtin_amdgcn_readfirstlane( (uint32_t)(uintptr_t)( __attribute__( ( address_space( 3 ) ) ) float* )( sPk ) );
    auto pkIssue = [ & ]( int kk ) {
        const int c0 = a.newRel0 + kk * NEWPER;
        const uint32_t bufLds = pkLds + (uint32_t)( ( kk % kPkBufs ) * kPkCap ) * 4u;
#pragma unroll 1
        for ( int d = 0; d < myChunks; d++ )
        {
            const int chunk = wave + d * kIssueWaves;
            if ( chunk < nX4 )
            {
                // rows start 16-byte aligned (gramRow) and rowFloats is a multiple of 4, so every
                // lane's 16 bytes are aligned and within one row (beyond the row's valid 3 (t + 1)
                // floats they are the next rows' values, which no reader uses)
                const int e = chunk * 4 * kWave + 4 * lane;
                int off = 0;
                if ( e < NEWPER * rowFloats )
                {
                    const int i = NEWPER == 1 ? 0 : e / rowFloats;
                    off = gramRow( c0 + i ) + ( e - i * rowFloats );
                    off = off > gEnd4Off ? gEnd4Off : off;
                }
                dmaDwordx4( Gact, 4u * (uint32_t)off, __builtin_amdgcn_readfirstlane( bufLds + (uint32_t)( chunk * 4 * kWave ) * 4u ) );
            }
            else
            {
                const int e = ( chunk - nX4 ) * kWave + lane;
                const float* src = Pu;
                if ( e < 2 * NEWPER )
                    src = Pu + kk * NEWPER * 2 + e;
                else if ( e < 2 * NEWPER + w2 )
                    src = EvU + kk * w2 + ( e - 2 * NEWPER );
                dmaDword( src, __builtin_amdgcn_readfirstlane( bufLds + (uint32_t)( pkRows + ( chunk - nX4 ) * kWave ) * 4u ) );
            }
        }
    };

    for ( int k = 0; k < kPkBufs - 1 && k < a.nSteps; k++ )
        pkIssue( k );

    for ( int i = tid; i < w2; i += kSeqThreads )
        sF[ i ] = Fg[ (int64_t)u * w2 + i ];
    for ( int i = tid; i < 2 * kSeqThreads; i += kSeqThreads )
        sB[ i ] = 0.f;
    for ( int i = tid; i < 2 * kRedF; i += kSeqThreads )
        sRedF[ i ] = 0.f;
    for ( int i = tid; i < 2 * nAct; i += kSeqThreads )
    {
        sC[ i ] = Cu[ i ];
        sY[ i ] = Yu[ i ];
    }
    for ( int i = tid; i < nAct; i += kSeqThreads )
        sInv[ i ] = 1.f / ( a.wL * Zn[ i ] + a.wG );

    // window Gram of the rows already in the window (active rows entered before the block);
    // the row in complex slot x, newest first from kLast, is kLast - ((pLast - x) mod Wc)
    // vector registers: the entering sample's entry is written at a wave-uniform index by VGPR
    // index mode (s_set_gpr_idx_on), the product reads them at constant indices
    typedef float GramRegs __attribute__( ( ext_vector_type( TP ) ) );
    GramRegs gA, gB, gC;
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
            const float* g = Gact + gramRow( hi ) + 3 * lo;
            const float m = ok ? 1.f : 0.f;
            gA[ j ] = m * g[ 0 ];
            gB[ j ] = m * g[ ord ? 1 : 2 ];
            gC[ j ] = m * g[ ord ? 2 : 1 ];
            if ( ( j & 3 ) == 3 )
                __builtin_amdgcn_sched_barrier( 0 );
        }
    }

    // Work of step ks that does not depend on the projection of step ks - 1 beyond c: waves
    // 8..15 take the active rows' share of each entering f (the c of rows that entered up to
    // step ks - 1 are final for it), waves 5..7 stage the Gram column of each entering sample.
    // It runs beside the Gram product of step ks - 1.
    auto stage = [ & ]( int ks, int pc0 ) { // pc0: complex slot of the step's first entering sample
        const float* pk = sPk + ( ks % kPkBufs ) * kPkCap;
        const int c0 = a.newRel0 + ks * NEWPER; // first entering complex sample
        if ( wave == kSeqWaves - 1 )
        {
            // the active rows' share of each entering f on the last wave, which holds no part
            // of the register Gram (the product waves' vector ALUs are the step's bottleneck):
            // kActMax / kWave active rows per lane, then one cross-lane sum per value
            float ea[ NEWPER ], eb[ NEWPER ];
#pragma unroll
            for ( int i = 0; i < NEWPER; i++ )
                ea[ i ] = eb[ i ] = 0.f;
            const int ln = tidInLoop() & ( kWave - 1 );
#pragma unroll
            for ( int r = 0; r < kActMax / kWave; r++ )
            {
                const int jd = r * kWave + ln;
                if ( jd < c0 )
                {
                    const float ca = sC[ 2 * jd ], cb = sC[ 2 * jd + 1 ];
#pragma unroll
                    for ( int i = 0; i < NEWPER; i++ )
                    {
                        const float* g = pk + i * rowFloats + 3 * jd;
                        ea[ i ] += ca * g[ 0 ] + cb * g[ 1 ];
                        eb[ i ] += ca * g[ 2 ] + cb * g[ 0 ];
                    }
                }
            }
#pragma unroll
            for ( int i = 0; i < NEWPER; i++ )
            {
                const float va = waveSum( ea[ i ] ), vb = waveSum( eb[ i ] );
                if ( lane == 0 )
                {
                    sRedE[ 2 * i ] = va;
                    sRedE[ 2 * i + 1 ] = vb;
                }
            }
        }
        else if ( tid >= kColBase )
        {
            // column: sCol[i][.][part*TPA + j] = (kappa(a_tau,a_k), kappa(a_tau,b_k), kappa(b_tau,a_k))
            // for tau = part*TP + j (segments padded to TPA floats for aligned vector reads),
            // against the window as it stands after the step (rows entering later in the same
            // step included): the row in complex slot x is kRef - ((pRef - x) mod Wc) with kRef
            // the last entering sample; kappa(k, r) comes from the package row of max(k, r) at
            // the position of min(k, r)
            const int kRef = c0 + NEWPER - 1;
            const int pRef = pc0 + NEWPER - 1 < Wc ? pc0 + NEWPER - 1 : pc0 + NEWPER - 1 - Wc;
            const int ct = tid - kColBase; // t-slot staged by this thread
            if ( ct < a.nParts * TP )
            {
                const int cpart = ct / TP, cj = ct - cpart * TP;
                const int cdst = cpart * TPA + cj;
                int dt = pRef - ct;
                dt += dt < 0 ? Wc : 0;
                const int rt = kRef - dt;
                const bool ok = ct < Wc && rt >= 0;
#pragma unroll
                for ( int i = 0; i < NEWPER; i++ )
                {
                    const int k = c0 + i;
                    float* col = sCol + i * 3 * kColStride;
                    const bool later = ok && rt > k; // entering after k in this step
                    const float* g = later ? pk + ( rt - c0 ) * rowFloats + 3 * k : pk + i * rowFloats + 3 * ( ok ? rt : 0 );
                    const float m = ok ? 1.f : 0.f;
                    col[ cdst ] = m * g[ 0 ];
                    col[ kColStride + cdst ] = m * ( later ? g[ 1 ] : g[ 2 ] );
                    col[ 2 * kColStride + cdst ] = m * ( later ? g[ 2 ] : g[ 1 ] );
                }
            }
        }
    };

    // slots advance by a fixed amount per step: tracked incrementally (no division in the loop)
    int sBase = ( a.a0modW + a.sFirstRel ) % w2; // slot of row s (== slot of row s - w2)
    int pC0 = ( cMod + a.newRel0 ) % Wc; // complex slot of the step's first entering sample
    waitVm<0>();
    __syncthreads();
    stage( 0, pC0 );
    ldsBarrier();

    for ( int kk = 0; kk < a.nSteps; kk++ )
    {
        const float* pk = sPk + ( kk % kPkBufs ) * kPkCap;
        const int sRel = a.sFirstRel + kk * a.q; // s - a0
        const int spRel = sRel - a.q;
        const int W = a.ramp ? ( sRel < w2 ? sRel : w2 ) : w2;
        const int c0 = a.newRel0 + kk * NEWPER; // first entering complex sample
        const int pNext = pC0 + NEWPER < Wc ? pC0 + NEWPER : pC0 + NEWPER - Wc;

        // 1. projection over the window. Thread tid < w2 takes row r = s - w2 + tid (the last
        //    w2 rows cover every slot once): its f gets the previous step's Gram product and,
        //    for a row already in the window, loses the share of rows leaving the dictionary;
        //    an entering row's f is its frozen share plus the active rows' share.
        if ( tid < w2 )
        {
            const int r = sRel - w2 + tid;
            int sl = sBase + tid;
            sl -= sl >= w2 ? w2 : 0;
            const int cv = sl >> 1, pt = sl & 1;
            const int wpos = r - ( sRel - W );
            const bool inWin = wpos >= 0, entering = r >= spRel; // entering implies inWin
            const int rc = r < 0 ? 0 : r, e2 = entering ? r - spRel : 0; // e2 = 2*entering index + part
            // every LDS read of the phase issued up front at clamped addresses (values a lane
            // does not need are discarded), so the phase is one round trip, not a chain of them
            float part[ PMAX ];
#pragma unroll
            for ( int t = 0; t < PMAX; t++ ) // parts beyond nParts hold zeros
                part[ t ] = sRedF[ 2 * ( t * Wc + cv ) + pt ];
            const float ent = sRedE[ e2 ];
            const float evv = pk[ pkRows + 2 * NEWPER + ( inWin ? wpos : 0 ) ];
            const float pv = pk[ pkRows + e2 ];
            const float fOld = sF[ sl ], yv = sY[ rc ], iv = sInv[ rc >> 1 ], cOld = sC[ rc ];
            const float invW = 1.f / (float)W;
            float acc = 0.f;
#pragma unroll
            for ( int t = 0; t < PMAX; t++ )
                acc += part[ t ];
            if ( a.hasEv && inWin && !entering )
                acc -= evv;
            float f = fOld + acc;
            float beta = 0.f;
            if ( inWin )
            {
                if ( entering )
                {
                    f = pv + ent;
                }
                const float res = yv - f;
                float bb = 0.f;
                if ( res > a.eB )
                    bb = res - a.eB;
                else if ( res < -a.eB )
                    bb = res + a.eB;
                bb = bb * invW * iv;
                beta = bb;
                sC[ r ] = cOld + bb;
            }
            sF[ sl ] = f;
            sB[ sl ] = beta; // sl = 2*complex slot + part
        }
        // 2. the entering samples' row and column of the register Gram: the threads whose
        //    t-slots contain p take kappa(k, v) (rows first), then the thread holding u-slot p
        //    takes the column staged beside the previous step's product
        {
            const int kRef = c0 + NEWPER - 1;
            const int pRef = pC0 + NEWPER - 1 < Wc ? pC0 + NEWPER - 1 : pC0 + NEWPER - 1 - Wc;
#pragma unroll
            for ( int i = 0; i < NEWPER; i++ )
            {
                const int k = c0 + i;
                const int p = pC0 + i < Wc ? pC0 + i : pC0 + i - Wc;
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
                    // the active lanes all belong to the part holding slot p, so jj is uniform
                    // among them: one indexed register write per array (VGPR index mode)
                    const int jjs = __builtin_amdgcn_readfirstlane( jj );
                    gA[ jjs ] = A;
                    gB[ jjs ] = Bv;
                    gC[ jjs ] = Cv;
                }
            }
#pragma unroll
            for ( int i = 0; i < NEWPER; i++ )
            {
                const int p = pC0 + i < Wc ? pC0 + i : pC0 + i - Wc;
                if ( gThread && v == p )
                {
                    const float4* col = reinterpret_cast<const float4*>( sCol + i * 3 * kColStride + part * TPA );
#pragma unroll
                    for ( int j4 = 0; j4 < TPA / 4; j4++ )
                    {
                        const float4 xa = col[ j4 ], xb = col[ kColStride / 4 + j4 ], xc = col[ 2 * kColStride / 4 + j4 ];
                        const float va[ 4 ] = { xa.x, xa.y, xa.z, xa.w }, vb[ 4 ] = { xb.x, xb.y, xb.z, xb.w },
                                    vc[ 4 ] = { xc.x, xc.y, xc.z, xc.w };
#pragma unroll
                        for ( int q4 = 0; q4 < 4; q4++ )
                        {
                            if ( 4 * j4 + q4 < TP )
                            {
                                gA[ 4 * j4 + q4 ] = va[ q4 ];
                                gB[ 4 * j4 + q4 ] = vb[ q4 ];
                                gC[ 4 * j4 + q4 ] = vc[ q4 ];
                            }
                        }
                    }
                }
            }
        }
        // the next step's package (issued two steps ago) must have landed: the staging of the
        // next step below reads it; the one issued in the previous step may still be in flight
        waitVmN( kk + 2 < a.nSteps ? myChunks : 0 );
        ldsBarrier(); // B
        if ( kk + 1 < a.nSteps )
            stage( kk + 1, pNext );
        // 3. partial f_v += sum_t kappa(t, v) beta_t from the register Gram (summed over the
        //    t-slot parts at the start of the next step)
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
        // the package kPkBufs - 1 steps ahead goes into the buffer read up to the previous step
        if ( kk + kPkBufs - 1 < a.nSteps )
            pkIssue( kk + kPkBufs - 1 );
        sBase += a.q;
        sBase -= sBase >= w2 ? w2 : 0;
        pC0 = pNext;
        ldsBarrier(); // C
    }
    if ( tid < w2 )
    {
        const int cv = tid >> 1, pt = tid & 1;
        float acc = 0.f;
#pragma unroll
        for ( int t = 0; t < PMAX; t++ )
            acc += sRedF[ 2 * ( t * Wc + cv ) + pt ];
        sF[ tid ] += acc;
    }
    __syncthreads();

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

template <int TP, int NEWPER, int PMAX>
__global__ void __launch_bounds__( kSeqThreads ) apsm_train_seq2( const float* __restrict__ Z, const float* __restrict__ Zn,
                                                                const float* __restrict__ Y, float* __restrict__ C,
                                                                float* __restrict__ Fg, float* __restrict__ thF,
                                                                const float* __restrict__ P, const float* __restrict__ Gact,
                                                                const float* __restrict__ EvF, const Seq2Args a )
{
    seq2Body<TP, NEWPER, PMAX>( Z, Zn, Y, C, Fg, thF, P, Gact, EvF, a, blockIdx.x );
}

// LDS bytes of apsm_train_seq2<TP, NEWPER> for a window of w2 real rows.
inline size_t seq2LdsBytes( int newPer, int w2, int tp )
{
    const int tpa = ( tp + 3 ) & ~3;
    const size_t floats = 4 * (size_t)seq2PkCap( newPer, tp, seq2PmaxOfTp( tp ) ) + w2 + 2 * kSeqThreads + 4 * kActMax + kActMax + kSeqWaves * 2 * kNewMax
        + 2 * kRedF + (size_t)newPer * 3 * kMaxParts * tpa;
    return floats * sizeof( float );
}

} // namespace apsm
