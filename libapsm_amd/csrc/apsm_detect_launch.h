// Host-side interface of the matrix-core detect kernels, which are compiled in translation
// units of their own (one per input precision, built in parallel): the batched detect
// descriptors and the launchers. The launchers return the launch status.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "apsm_common.h"

namespace apsm {

// Input precision of the matrix-core cross terms (accumulation is FP32 for every one).
constexpr int PREC_F32 = 0;
constexpr int PREC_BF16 = 1;
constexpr int PREC_F16 = 2;

constexpr int kMfmaThreads = 256; // 4 waves, 32 samples each
constexpr int kMfmaSamples = 128; // samples per workgroup tile

// Tile shapes of the persistent batched matrix-core detect, chosen by the batch autotuner:
// waves per workgroup (32 samples each) and dictionary entries per LDS tile (0: the single
// handle's choice, 64, or 32 at 64 antennas). Shape 0 is the single-handle shape; the others
// are built for one-user FP32 instances up to 32 antennas (mfmaShapeBuilt).
struct MfmaShape
{
    int waves, entries;
};
constexpr int kMfmaShapes = 5;
constexpr MfmaShape kMfmaShape[ kMfmaShapes ] = { { 4, 0 }, { 2, 64 }, { 4, 32 }, { 8, 64 }, { 4, 128 } };
inline bool mfmaShapeBuilt( int shape, int prec, int nmax, int umax )
{
    return shape == 0 || ( shape > 0 && shape < kMfmaShapes && prec == PREC_F32 && nmax <= 32 && umax == 1 );
}
inline int mfmaShapeSamples( int shape ) { return 32 * kMfmaShape[ shape ].waves; }
inline int mfmaShapeEntries( int shape, int nmax ) { return kMfmaShape[ shape ].entries ? kMfmaShape[ shape ].entries : ( nmax >= 64 ? 32 : 64 ); }

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

// Frozen-dictionary share of a block's entering pilots on the matrix cores: the frozen range
// is cut into chunks of kEntryChunk entries (one workgroup each), then combined.
constexpr int kEntryChunk = 64;

// Tiles of kMfmaSamples entering samples in the entry share (work items = tiles x chunks).
__host__ __device__ inline int entryTiles( int nNewC ) { return ( nNewC + kMfmaSamples - 1 ) / kMfmaSamples; }

// First frozen entry of the block's first entering sample and the number of chunks (chunk c
// covers [j1 - (c+1) kEntryChunk, j1 - c kEntryChunk) within [jBeg, j1)).
inline void entryChunks( const EntryArgs& a, int64_t& jBeg, int& nChunks )
{
    const int64_t rt = a.rowBase + 2 * a.newC0;
    const int64_t se = rt < a.sFirst ? a.sFirst : a.sFirst + a.q * ( ( rt - a.sFirst ) / a.q + 1 );
    const int64_t lo = se > a.H ? se - a.H : 0;
    jBeg = ( lo - a.rowBase ) / 2;
    const int64_t j1 = ( a.a0 - a.rowBase ) / 2;
    nChunks = j1 > jBeg ? (int)( ( j1 - jBeg + kEntryChunk - 1 ) / kEntryChunk ) : 0;
}

// One instance of a batched entry share: buffers, partial sums [U][nChunks][nNewC][2], output P.
struct EntryMfmaDesc
{
    const float *Z, *Zn, *C, *thF;
    float *P, *part;
    EntryArgs a;
    int64_t jBeg;
    int nChunks;
};

// Gram among a block's active samples: Gact[gramRow(i) + 3 j + (0..2)] for active samples j <= i (relative
// to c0), kappa = wL x.z + wG K in the three real forms A = kappa(a_i,a_j) = kappa(b_i,b_j),
// B = kappa(a_i,b_j), C = kappa(b_i,a_j); 32 x 32 sample blocks of the lower triangle on the
// matrix cores, four per workgroup.
struct GramArgs
{
    int N, L;
    int64_t c0; // local complex index of the first active sample
    int64_t nA; // active complex samples
    float wL, wG, g2;
};

struct GramDesc
{
    const float *Z, *Zn;
    float* Gact;
    GramArgs a;
};

// Row i of the active Gram (3 (i + 1) floats) starts at gramRow(i) floats: every row is padded
// to whole 16-byte units, so each starts 16-byte aligned (the sequencer copies rows to LDS in
// 16-byte lanes). The padding of rows 4m .. 4m+3 is 1, 2, 3, 0 floats; gramRow(nA) floats hold
// nA rows.
template <typename I>
__host__ __device__ inline I gramRow( I i )
{
    return 3 * ( i * ( i + 1 ) / 2 ) + 6 * ( i >> 2 ) + ( ( 0x6310 >> ( 4 * ( i & 3 ) ) ) & 15 );
}

__host__ __device__ inline int64_t gramBlocks( int64_t nA )
{
    const int64_t nb = ( nA + 31 ) / 32;
    return nb * ( nb + 1 ) / 2;
}

// One train block's precompute as one launch: the entry-share chunks (entryWg workgroups), the
// Gram blocks (gramWg workgroups of four 32 x 32 blocks) and the eviction shares (evWg
// workgroups of 256 threads), each workgroup taking one role; the chunk combine follows.
struct PreDesc
{
    EntryMfmaDesc e;
    GramDesc g;
    EvDesc v;
    int hasEv;
    int entryWg, gramWg, evWg;
};

inline void preShape( PreDesc& d )
{
    d.entryWg = d.e.nChunks * entryTiles( d.e.a.nNewC );
    d.gramWg = (int)( ( gramBlocks( d.g.a.nA ) + 3 ) / 4 );
    d.evWg = d.hasEv ? (int)( ( (int64_t)d.v.a.nSteps * ( d.v.a.w2 / 2 ) + 255 ) / 256 ) : 0;
}

// One instance of a batched detect: its data samples, dictionary, coefficients and weights.
struct DetDesc
{
    const float* rx;
    const float *Z, *Zn, *C, *theta;
    float* est;
    int64_t M, ldC, G;
    int U, N;
    float wL, wG, g2;
};

// One (instance, sample tile) work item of a persistent batched detect.
struct WorkItem
{
    int desc;
    int tile;
};

// Single-handle detect of M samples; nmax in {4, 8, 16, 32, 64}, umax in {1, 2, 4, 8, 12, 16}.
hipError_t launchDetectMfma( int prec, int nmax, int umax, hipStream_t s, const float* rx, int64_t M, const float* Z,
                             const float* Zn, const float* C, int64_t ldC, const float* theta, int64_t G, int U, float wL,
                             float wG, float g2, float* est, int N );

// Entry share of one block (work items [chunk0, chunk0 + count) of nChunks * entryTiles(nNewC)),
// then the combine of the chunks and the linear share into P [U][nNewC][2]; the batched combine
// runs over n descriptors (gridX: enough threads for the largest U * nNewC).
hipError_t launchEntryMfma( int nmax, int umax, hipStream_t s, const float* Z, const float* Zn, const float* C, float* part,
                            const EntryArgs& a, int64_t jBeg, int nChunks, int chunk0, int count );
hipError_t launchEntryCombine( hipStream_t s, const float* Z, const float* thF, const float* part, float* P, const EntryArgs& a,
                               int nChunks );
hipError_t launchBatchEntryCombine( hipStream_t s, const EntryMfmaDesc* d, int n, int gridX );

hipError_t launchBlockPre( int nmax, int umax, hipStream_t s, const PreDesc& d );
hipError_t launchBatchBlockPre( int nmax, int umax, hipStream_t s, const PreDesc* d, int n, int gridX );

// Persistent batched detect: `grid` workgroups of tile shape `shape` (kMfmaShape) walk nItems
// work items of tiles of mfmaShapeSamples(shape) samples.
hipError_t launchBatchDetectMfma( int prec, int nmax, int umax, hipStream_t s, const DetDesc* d, const WorkItem* items,
                                  int nItems, int grid, int shape = 0 );

} // namespace apsm
