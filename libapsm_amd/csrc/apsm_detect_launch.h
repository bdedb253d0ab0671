// Host-side interface of the matrix-core detect kernels, which are compiled in translation
// units of their own (one per input precision, built in parallel): the batched detect
// descriptors and the launchers. The launchers return the launch status.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace apsm {

// Input precision of the matrix-core cross terms (accumulation is FP32 for every one).
constexpr int PREC_F32 = 0;
constexpr int PREC_BF16 = 1;
constexpr int PREC_F16 = 2;

constexpr int kMfmaThreads = 256; // 4 waves, 32 samples each
constexpr int kMfmaSamples = 128; // samples per workgroup tile

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

// Single-handle detect of M samples; nmax in {4, 8, 16, 32, 64}, umax in {1, 2, 4, 8, 16}.
hipError_t launchDetectMfma( int prec, int nmax, int umax, hipStream_t s, const float* rx, int64_t M, const float* Z,
                             const float* Zn, const float* C, int64_t ldC, const float* theta, int64_t G, int U, float wL,
                             float wG, float g2, float* est, int N );

// Persistent batched detect: `grid` workgroups walk nItems work items of tiles of kMfmaSamples.
hipError_t launchBatchDetectMfma( int prec, int nmax, int umax, hipStream_t s, const DetDesc* d, const WorkItem* items,
                                  int nItems, int grid );

} // namespace apsm
