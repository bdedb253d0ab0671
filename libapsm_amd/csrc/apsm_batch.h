// Batched kernels: many independent detector instances per launch. Each launch reads one
// descriptor per instance (pointers into that instance's buffers plus the same argument block
// the single-handle kernel takes); blockIdx.y selects the descriptor, blockIdx.x plays the role
// it has in the single-handle launch, and workgroups beyond an instance's extent return before
// any barrier. Detect is persistent: a fixed grid walks a list of (instance, sample tile) work
// items. The device bodies are the single-handle ones, so a batched instance computes exactly
// what its own handle would.
#pragma once

#include "apsm_device.h"
#include "apsm_train2.h"
#include "apsm_detect_launch.h"

namespace apsm {

struct StageDesc
{
    const float *rx, *tx;
    float *Z, *Zn, *Y, *C;
    int64_t S, ldYC;
    int N, U;
};


struct Seq2Desc
{
    const float *Z, *Zn, *Y;
    float *C, *Fg, *thF;
    const float *P, *Gact, *EvF;
    Seq2Args a;
};

struct ThetaDesc
{
    const float *Z, *C, *thOld;
    float* theta;
    int64_t j0, j1, ldC;
    int N, U;
};

__global__ void batch_stage( const StageDesc* __restrict__ d )
{
    const StageDesc& s = d[ blockIdx.y ];
    stagePilotsBody( s.rx, s.tx, nullptr, s.N, s.U, s.S, s.Z, s.Zn, s.Y, s.C, s.ldYC, (int64_t)blockIdx.x * blockDim.x + threadIdx.x );
}

template <int TP, int NEWPER, int PMAX>
__global__ void __launch_bounds__( kSeqThreads ) batch_seq2( const Seq2Desc* __restrict__ d )
{
    const Seq2Desc& s = d[ blockIdx.y ];
    if ( (int)blockIdx.x >= s.a.U )
        return;
    seq2Body<TP, NEWPER, PMAX>( s.Z, s.Zn, s.Y, s.C, s.Fg, s.thF, s.P, s.Gact, s.EvF, s.a, blockIdx.x );
}

__global__ void batch_theta( const ThetaDesc* __restrict__ d )
{
    const ThetaDesc& t = d[ blockIdx.y ];
    if ( (int)blockIdx.x >= t.U )
        return;
    thetaFullBody( t.Z, t.C, t.thOld, t.theta, t.N, t.j0, t.j1, t.ldC, blockIdx.x );
}

// Persistent detect over a work list (all items of one launch share NMAX and UMAX).
template <int NMAX, int UMAX, int THREADS>
__global__ void __launch_bounds__( THREADS ) batch_detect( const DetDesc* __restrict__ d, const WorkItem* __restrict__ items,
                                                          int nItems )
{
    for ( int w = blockIdx.x; w < nItems; w += gridDim.x )
    {
        const WorkItem it = items[ w ];
        const DetDesc& x = d[ it.desc ];
        detectTile<NMAX, UMAX, THREADS>( x.rx, x.M, x.Z, x.Zn, x.C, x.ldC, x.theta, x.G, x.U, x.wL, x.wG, x.g2, x.est, x.N, it.tile );
    }
}

// Row-wise copies between device buffers: rows x n floats from src (row stride ldSrc) to dst
// (row stride ldDst); the buffers do not overlap.
struct CopyDesc
{
    const float* src;
    float* dst;
    int64_t n, ldSrc, ldDst;
    int rows;
};

__global__ void batch_copy( const CopyDesc* __restrict__ d )
{
    const CopyDesc& c = d[ blockIdx.y ];
    const int64_t total = c.n * c.rows;
    for ( int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < total; i += (int64_t)gridDim.x * blockDim.x )
    {
        const int64_t r = i / c.n, k = i - r * c.n;
        c.dst[ r * c.ldDst + k ] = c.src[ r * c.ldSrc + k ];
    }
}

// Clears n floats (used instead of memset nodes inside captured graphs).
__global__ void batch_fill_zero( float* const* __restrict__ ptrs, const int64_t* __restrict__ counts )
{
    float* p = ptrs[ blockIdx.y ];
    const int64_t n = counts[ blockIdx.y ];
    for ( int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x )
        p[ i ] = 0.f;
}

} // namespace apsm
