// Host side of the batched path (included by apsm_capi.hip after the handle definitions).
//
// A batch owns one detector handle per instance. A run is, for every instance at once: reset,
// train on its pilots, detect its data samples. Planning walks every instance's train schedule
// (the same blocks trainBlocked forms), groups the kernels of all instances by round (at most
// one block per instance, packed by the CUs the sequencer workgroups take) and by template
// instance, writes one descriptor per (launch, instance) into a single device buffer,
// picks the detect launch shape per bucket by timing candidates on the real batch, and records
// the launch sequence. A run replays it, captured once into a hipGraph, on the batch's stream.
#pragma once

#include <map>
#include <tuple>

constexpr int kBatchSideStreams = 3; // sequencer buckets of one round run side by side

struct apsm_batch
{
    std::vector<apsm_handle*> inst;
    hipStream_t stream = nullptr;
    hipStream_t side[ kBatchSideStreams ] = {};
    hipEvent_t fork = nullptr, join[ kBatchSideStreams ] = {};
    hipEvent_t ev0 = nullptr, evMid = nullptr, ev1 = nullptr;
    bool planned = false, useGraph = true, timed = false;
    DevBuf<char> desc; // every descriptor array of the plan
    std::vector<std::function<void()>> launches; // the run, in order (train launches, then detect)
    size_t trainLaunches = 0;
    hipGraph_t graph[ 2 ] = {};
    hipGraphExec_t exec[ 2 ] = {}; // train, detect
    std::vector<std::array<int, 8>> tuning; // (NMAX, UMAX, threads, blocks per CU, items, kernel, entries per tile, samples per workgroup)
    int cus = 256;
};

namespace {

// Detect launch shapes the autotuner chooses from: threads per workgroup (vector-ALU kernel) or
// tile shape (matrix-core kernel, apsm::kMfmaShape) x workgroups per CU.
constexpr int kTuneThreads[] = { 128, 256, 512 };
// workgroups per CU of the persistent detect; 0: one workgroup per work item (the end of the
// range: a larger grid than that changes nothing)
constexpr int kTuneBlocksPerCu[] = { 1, 2, 4, 8, 16, 32, 64, 128, 0 };
constexpr int kTuneCands = 5; // >= the thread counts and the matrix-core shapes

template <int NMAX, int UMAX, int THREADS>
void launchBatchDetect( hipStream_t s, const apsm::DetDesc* d, const apsm::WorkItem* items, int nItems, int grid )
{
    hipLaunchKernelGGL( ( apsm::batch_detect<NMAX, UMAX, THREADS> ), dim3( grid ), dim3( THREADS ), 0, s, d, items, nItems );
    HIP_OR_THROW( hipGetLastError() );
}

template <int NMAX, int UMAX>
void launchBatchDetectT( hipStream_t s, int threads, const apsm::DetDesc* d, const apsm::WorkItem* items, int nItems, int grid )
{
    switch ( threads )
    {
    case 128: launchBatchDetect<NMAX, UMAX, 128>( s, d, items, nItems, grid ); break;
    case 256: launchBatchDetect<NMAX, UMAX, 256>( s, d, items, nItems, grid ); break;
    default: launchBatchDetect<NMAX, UMAX, 512>( s, d, items, nItems, grid ); break;
    }
}

template <int NMAX>
void launchBatchDetectU( hipStream_t s, int umax, int threads, const apsm::DetDesc* d, const apsm::WorkItem* items, int nItems, int grid )
{
    switch ( umax )
    {
    case 1: launchBatchDetectT<NMAX, 1>( s, threads, d, items, nItems, grid ); break;
    case 2: launchBatchDetectT<NMAX, 2>( s, threads, d, items, nItems, grid ); break;
    case 4: launchBatchDetectT<NMAX, 4>( s, threads, d, items, nItems, grid ); break;
    case 8: launchBatchDetectT<NMAX, 8>( s, threads, d, items, nItems, grid ); break;
    case 12: launchBatchDetectT<NMAX, 12>( s, threads, d, items, nItems, grid ); break;
    default: launchBatchDetectT<NMAX, 16>( s, threads, d, items, nItems, grid ); break;
    }
}

void launchBatchDetectAny( hipStream_t s, int nmax, int umax, int threads, const apsm::DetDesc* d, const apsm::WorkItem* items,
                           int nItems, int grid )
{
    switch ( nmax )
    {
    case 4: launchBatchDetectU<4>( s, umax, threads, d, items, nItems, grid ); break;
    case 8: launchBatchDetectU<8>( s, umax, threads, d, items, nItems, grid ); break;
    case 16: launchBatchDetectU<16>( s, umax, threads, d, items, nItems, grid ); break;
    case 32: launchBatchDetectU<32>( s, umax, threads, d, items, nItems, grid ); break;
    default: launchBatchDetectU<64>( s, umax, threads, d, items, nItems, grid ); break;
    }
}

template <int TP, int NEWPER, int PMAX>
void launchBatchSeq2( hipStream_t s, const apsm::Seq2Desc* d, int n, int gridX, size_t lds )
{
    if ( lds > 64 * 1024 )
        HIP_OR_THROW( hipFuncSetAttribute( (const void*)apsm::batch_seq2<TP, NEWPER, PMAX>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                           (int)lds ) );
    hipLaunchKernelGGL( ( apsm::batch_seq2<TP, NEWPER, PMAX> ), dim3( gridX, n ), dim3( apsm::kSeqThreads ), lds, s, d );
    HIP_OR_THROW( hipGetLastError() );
}

template <int NEWPER>
void launchBatchSeq2Tp( hipStream_t s, int tp, const apsm::Seq2Desc* d, int n, int gridX, size_t lds )
{
    switch ( tp )
    {
    case 2: launchBatchSeq2<2, NEWPER, 24>( s, d, n, gridX, lds ); break;
    case 4: launchBatchSeq2<4, NEWPER, 16>( s, d, n, gridX, lds ); break;
    case 7: launchBatchSeq2<7, NEWPER, 12>( s, d, n, gridX, lds ); break;
    case 10: launchBatchSeq2<10, NEWPER, 10>( s, d, n, gridX, lds ); break;
    case 16: launchBatchSeq2<16, NEWPER, 8>( s, d, n, gridX, lds ); break;
    default: launchBatchSeq2<27, NEWPER, 6>( s, d, n, gridX, lds ); break;
    }
}

int pmaxOfTp( int tp ) { return apsm::seq2PmaxOfTp( tp ); }

// One block of one instance's train as batched-launch descriptors.
struct TrainBlk
{
    apsm::EntryMfmaDesc e;
    int nmax;
    apsm::GramDesc g;
    apsm::EvDesc v;
    apsm::Seq2Desc s;
    bool hasEv;
    int umax, tp, newPer;
    size_t lds;
};

// Host-side descriptor storage: typed arrays appended into one byte buffer; launches refer to
// offsets that become device pointers once the buffer is uploaded.
struct DescArena
{
    std::vector<char> bytes;
    template <typename T>
    size_t add( const std::vector<T>& v )
    {
        size_t off = ( bytes.size() + 255 ) & ~size_t( 255 );
        bytes.resize( off + v.size() * sizeof( T ) );
        std::memcpy( bytes.data() + off, v.data(), v.size() * sizeof( T ) );
        return off;
    }
};

// The block precompute of a round: one fused launch (entry chunks, Gram, eviction shares) per
// (NMAX, UMAX) bucket, then one combine over every instance of the round.
struct PreLaunch
{
    std::map<std::pair<int, int>, std::vector<apsm::PreDesc>> buckets;
    std::map<std::pair<int, int>, int> maxWg;
    int maxCombine = 0;
    void add( const TrainBlk& k )
    {
        apsm::PreDesc d {};
        d.e = k.e;
        d.g = k.g;
        d.v = k.v;
        d.hasEv = k.hasEv ? 1 : 0;
        apsm::preShape( d );
        const auto key = std::make_pair( k.nmax, k.umax );
        buckets[ key ].push_back( d );
        maxWg[ key ] = std::max( maxWg[ key ], d.entryWg + d.gramWg + d.evWg );
        maxCombine = std::max( maxCombine, k.e.a.U * k.e.a.nNewC );
    }
    // descriptors into the arena: per bucket, then every instance's entry descriptor for the combine
    size_t upload( DescArena& arena, std::vector<std::pair<std::pair<int, int>, size_t>>& offs, int& nComb ) const
    {
        std::vector<apsm::EntryMfmaDesc> all;
        for ( auto& kv : buckets )
        {
            offs.push_back( { kv.first, arena.add( kv.second ) } );
            for ( auto& d : kv.second )
                all.push_back( d.e );
        }
        nComb = (int)all.size();
        return arena.add( all );
    }
};

void launchRoundPre( hipStream_t s, const PreLaunch& pl, const char* base, const std::vector<std::pair<std::pair<int, int>, size_t>>& offs,
                     size_t combOff, int nComb )
{
    for ( auto& o : offs )
        HIP_OR_THROW( apsm::launchBatchBlockPre( o.first.first, o.first.second, s, (const apsm::PreDesc*)( base + o.second ),
                                                 (int)pl.buckets.at( o.first ).size(), pl.maxWg.at( o.first ) ) );
    if ( nComb > 0 )
        HIP_OR_THROW( apsm::launchBatchEntryCombine( s, (const apsm::EntryMfmaDesc*)( base + combOff ), nComb, ( pl.maxCombine + 255 ) / 256 ) );
}

// The same for a batch round: the (NMAX, UMAX) buckets are independent launches, run side by
// side on the batch's streams (largest NMAX on the main stream), then the combine after them.
void launchRoundPreForked( apsm_batch* B, const PreLaunch& pl, const char* base,
                           const std::vector<std::pair<std::pair<int, int>, size_t>>& offs, size_t combOff, int nComb )
{
    const int nside = std::min<int>( kBatchSideStreams, (int)offs.size() - 1 );
    if ( nside > 0 )
    {
        HIP_OR_THROW( hipEventRecord( B->fork, B->stream ) );
        for ( int k = 0; k < nside; k++ )
            HIP_OR_THROW( hipStreamWaitEvent( B->side[ k ], B->fork, 0 ) );
    }
    for ( size_t k = 0; k < offs.size(); k++ )
    {
        const auto& o = offs[ offs.size() - 1 - k ];
        const int lane = (int)( k % ( nside + 1 ) );
        hipStream_t st = lane == 0 ? B->stream : B->side[ lane - 1 ];
        HIP_OR_THROW( apsm::launchBatchBlockPre( o.first.first, o.first.second, st, (const apsm::PreDesc*)( base + o.second ),
                                                 (int)pl.buckets.at( o.first ).size(), pl.maxWg.at( o.first ) ) );
    }
    for ( int k = 0; k < nside; k++ )
    {
        HIP_OR_THROW( hipEventRecord( B->join[ k ], B->side[ k ] ) );
        HIP_OR_THROW( hipStreamWaitEvent( B->stream, B->join[ k ], 0 ) );
    }
    if ( nComb > 0 )
        HIP_OR_THROW( apsm::launchBatchEntryCombine( B->stream, (const apsm::EntryMfmaDesc*)( base + combOff ), nComb, ( pl.maxCombine + 255 ) / 256 ) );
}

void batchPlan( apsm_batch* B, const float* const* rxT, const float* const* txT, const uint32_t* T, const float* const* rxD, const uint32_t* M,
                float* const* est, bool useGraph )
{
    const int n = (int)B->inst.size();
    ARG_CHECK( n > 0, "empty batch" );
    for ( int i = 0; i < n; i++ )
    {
        apsm_handle* h = B->inst[ i ];
        const apsm_params& p = h->p;
        ARG_CHECK( rxT[ i ] && txT[ i ] && rxD[ i ] && est[ i ] && T[ i ] > 0 && M[ i ] > 0, "null or empty batch input" );
        if ( !h->blocked || p.gaussian_kernel_variance == 0.f || p.lls_initialization || p.training_shuffle || h->U > 16 )
            throw ApsmError { APSM_ERR_UNSUPPORTED,
                              "batched instances need the blocked train path, a configured kernel variance, no LLS initialisation, "
                              "no shuffle and at most 16 users" };
    }
    // a run enqueued before this call may still read the descriptors, graphs and instance
    // buffers replaced below (the descriptor upload is not ordered after the batch streams)
    HIP_OR_THROW( hipStreamSynchronize( B->stream ) );
    for ( int k = 0; k < kBatchSideStreams; k++ )
        HIP_OR_THROW( hipStreamSynchronize( B->side[ k ] ) );
    for ( int g = 0; g < 2; g++ )
    {
        if ( B->exec[ g ] )
            (void)hipGraphExecDestroy( B->exec[ g ] );
        if ( B->graph[ g ] )
            (void)hipGraphDestroy( B->graph[ g ] );
        B->exec[ g ] = nullptr;
        B->graph[ g ] = nullptr;
    }
    B->launches.clear();
    B->tuning.clear();

    DescArena arena;
    // per-run reset: f, theta share and theta of every instance (fill kernels, capturable)
    std::vector<float*> zp;
    std::vector<int64_t> zc;
    std::vector<apsm::StageDesc> stage;
    int64_t maxS = 0;
    std::vector<std::vector<TrainBlk>> blocks( n );
    std::vector<apsm::ThetaDesc> theta( n );
    std::vector<apsm::DetDesc> det( n );
    for ( int i = 0; i < n; i++ )
    {
        apsm_handle* h = B->inst[ i ];
        const apsm_params& p = h->p;
        const int N = h->N, L = h->L, U = h->U, w2 = h->w2, Wc = w2 / 2;
        const int64_t Ti = T[ i ];
        // fresh state: the stream buffer holds this call's pilots only
        h->varSet = true;
        h->var = p.gaussian_kernel_variance * (float)( N * N );
        h->anyTrain = true;
        StreamBuf& b = h->buf[ 0 ];
        b.ensure( Ti, N, U );
        const int64_t ld = 2 * b.capK;
        zp.push_back( h->F.ptr );
        zc.push_back( (int64_t)U * w2 );
        zp.push_back( h->thOld.ptr );
        zc.push_back( (int64_t)U * L );
        zp.push_back( h->theta.ptr );
        zc.push_back( (int64_t)U * L );
        stage.push_back( apsm::StageDesc { rxT[ i ], txT[ i ], b.Z.ptr, b.Zn.ptr, b.Y.ptr, b.C.ptr, Ti, ld, N, U } );
        maxS = std::max( maxS, Ti );

        const float wG = p.gaussian_kernel_weight, wL = p.linear_kernel_weight;
        const float g2 = 0.5f / h->var * kLog2e;
        const int64_t q = p.sample_step;
        const int newPer = (int)( q / 2 );
        const int64_t R2 = 2 * Ti;
        const int64_t nStepsTotal = R2 / q;
        const int64_t steps = blockedSteps( h );
        reserveBlockedScratch( h, steps );
        const int tp = seq2Tp( Wc );
        const int nParts = std::min( apsm::kSeqThreads / Wc, ( Wc + tp - 1 ) / tp );
        if ( tp < 0 || nParts > apsm::kMaxParts || nParts * tp < Wc || nParts > pmaxOfTp( tp ) )
            throw ApsmError { APSM_ERR_UNSUPPORTED, "window layout does not fit the blocked train path" };
        int64_t a0Next = 0;
        for ( int64_t k0 = 0; k0 < nStepsTotal; k0 += steps )
        {
            const int64_t nb = std::min( steps, nStepsTotal - k0 );
            const int64_t sPrev = k0 * q, sFirst = sPrev + q, sLast = sPrev + nb * q;
            const int64_t a0 = sFirst - std::min<int64_t>( w2, sFirst );
            const int64_t sNext = sLast + q;
            a0Next = sNext - std::min<int64_t>( w2, sNext );
            const int nNewC = (int)( nb * newPer );
            const int nActC = (int)( ( sLast - a0 ) / 2 );
            TrainBlk k {};
            k.umax = umaxOf( U );
            k.tp = tp;
            k.newPer = newPer;
            apsm::EntryArgs& ea = k.e.a;
            ea.N = N;
            ea.L = L;
            ea.U = U;
            ea.q = (int)q;
            ea.rowBase = 0;
            ea.newC0 = sPrev / 2;
            ea.sFirst = sFirst;
            ea.H = h->H;
            ea.a0 = a0;
            ea.ldYC = ld;
            ea.nNewC = nNewC;
            ea.wL = wL;
            ea.wG = wG;
            ea.g2 = g2;
            k.nmax = nmaxOf( N );
            k.e.Z = b.Z.ptr;
            k.e.Zn = b.Zn.ptr;
            k.e.C = b.C.ptr;
            k.e.thF = h->thOld.ptr;
            k.e.P = h->Pbuf.ptr;
            k.e.part = h->Ppart.ptr;
            apsm::entryChunks( ea, k.e.jBeg, k.e.nChunks );
            apsm::GramArgs& ga = k.g.a;
            ga.N = N;
            ga.L = L;
            ga.c0 = a0 / 2;
            ga.nA = nActC;
            ga.wL = wL;
            ga.wG = wG;
            ga.g2 = g2;
            k.g.Z = b.Z.ptr;
            k.g.Zn = b.Zn.ptr;
            k.g.Gact = h->Gact.ptr;
            const int64_t loLast = sLast > h->H ? sLast - h->H : 0, loPrev = sPrev > h->H ? sPrev - h->H : 0;
            k.hasEv = loLast > loPrev;
            apsm::EvArgs& ev = k.v.a;
            ev.N = N;
            ev.L = L;
            ev.U = U;
            ev.q = (int)q;
            ev.w2 = w2;
            ev.nSteps = (int)nb;
            ev.rowBase = 0;
            ev.sPrev = sPrev;
            ev.sFirst = sFirst;
            ev.H = h->H;
            ev.ldYC = ld;
            ev.wG = wG;
            ev.g2 = g2;
            k.v.Z = b.Z.ptr;
            k.v.Zn = b.Zn.ptr;
            k.v.C = b.C.ptr;
            k.v.EvF = h->EvF.ptr;
            apsm::Seq2Args& s2 = k.s.a;
            s2.N = N;
            s2.L = L;
            s2.U = U;
            s2.w2 = w2;
            s2.Wc = Wc;
            s2.q = (int)q;
            s2.nSteps = (int)nb;
            s2.hasEv = k.hasEv ? 1 : 0;
            s2.nParts = nParts;
            s2.ramp = a0 == 0 ? 1 : 0;
            s2.a0modW = (int)( a0 % w2 );
            s2.sFirstRel = (int)( sFirst - a0 );
            s2.nNewC = nNewC;
            s2.nActC = nActC;
            s2.newRel0 = (int)( ( sPrev - a0 ) / 2 );
            s2.nFr = (int)( ( a0Next - a0 ) / 2 );
            s2.ldYC = ld;
            s2.wL = wL;
            s2.wG = wG;
            s2.eB = p.eB;
            k.s.Z = b.Z.ptr + ( a0 / 2 ) * L;
            k.s.Zn = b.Zn.ptr + a0 / 2;
            k.s.Y = b.Y.ptr + a0;
            k.s.C = b.C.ptr + a0;
            k.s.Fg = h->F.ptr;
            k.s.thF = h->thOld.ptr;
            k.s.P = h->Pbuf.ptr;
            k.s.Gact = h->Gact.ptr;
            k.s.EvF = h->EvF.ptr;
            k.lds = apsm::seq2LdsBytes( newPer, w2, tp );
            blocks[ i ].push_back( k );
        }
        // stream bookkeeping after the train (same schedule as the kernels)
        const int64_t used = nStepsTotal * q;
        h->used = (uint64_t)used;
        h->a0Next = a0Next;
        h->cur = 0;
        h->bufBaseC = 0;
        h->bufK = Ti;
        h->trainTimed = false;
        const int64_t newUsedC = used / 2;
        const int64_t lo = std::min<int64_t>( a0Next / 2, newUsedC );
        theta[ i ] = apsm::ThetaDesc { b.Z.ptr, b.C.ptr, h->thOld.ptr, h->theta.ptr, lo, newUsedC, ld, N, U };
        const int64_t dlo = std::max<int64_t>( 0, newUsedC - h->Hc );
        const int64_t G = newUsedC - dlo;
        det[ i ] = apsm::DetDesc { rxD[ i ],
                                   G > 0 ? b.Z.ptr + dlo * L : nullptr,
                                   G > 0 ? b.Zn.ptr + dlo : nullptr,
                                   G > 0 ? b.C.ptr + 2 * dlo : nullptr,
                                   h->theta.ptr,
                                   est[ i ],
                                   (int64_t)M[ i ],
                                   ld,
                                   G,
                                   U,
                                   N,
                                   p.linear_kernel_weight,
                                   p.gaussian_kernel_weight,
                                   g2 };
    }

    // ---- launch list ----
    struct Pending
    {
        size_t off;
        std::function<void( const char* base )> fn;
    };
    std::vector<Pending> pend;
    const size_t zpOff = arena.add( zp ), zcOff = arena.add( zc );
    const int nz = (int)zp.size();
    pend.push_back( { 0, [ = ]( const char* base ) {
                         hipLaunchKernelGGL( apsm::batch_fill_zero, dim3( 4, nz ), dim3( 256 ), 0, B->stream, (float* const*)( base + zpOff ),
                                             (const int64_t*)( base + zcOff ) );
                         HIP_OR_THROW( hipGetLastError() );
                     } } );
    {
        const size_t off = arena.add( stage );
        const int ns = (int)stage.size(), gx = gridFor( maxS, 256 );
        pend.push_back( { 0, [ = ]( const char* base ) {
                             hipLaunchKernelGGL( apsm::batch_stage, dim3( gx, ns ), dim3( 256 ), 0, B->stream,
                                                 (const apsm::StageDesc*)( base + off ) );
                             HIP_OR_THROW( hipGetLastError() );
                         } } );
    }
    // Rounds: each instance's blocks in order, one per round at most, packed by the CUs their
    // sequencer workgroups hold (one user's workgroup takes a whole CU, or half of one where two
    // fit: at most 64 registers and 80 KB of LDS). A round fills up to the device's CUs, the
    // instances with the most blocks left first, so that the train takes as many rounds as its
    // longest chain of blocks. Instances are independent, so the order of their blocks across
    // rounds does not change any result.
    std::vector<std::vector<std::pair<int, int>>> rounds; // (instance, block) per round
    {
        std::vector<size_t> next( n, 0 );
        std::vector<double> weight( n, 0.0 );
        size_t left = 0;
        for ( int i = 0; i < n; i++ )
        {
            left += blocks[ i ].size();
            if ( !blocks[ i ].empty() )
            {
                const TrainBlk& k = blocks[ i ][ 0 ];
                const bool two = k.tp <= 10 && k.lds <= ( size_t( 80 ) << 10 );
                weight[ i ] = ( two ? 0.5 : 1.0 ) * k.s.a.U;
            }
        }
        std::vector<int> order( n );
        for ( int i = 0; i < n; i++ )
            order[ i ] = i;
        while ( left > 0 )
        {
            std::stable_sort( order.begin(), order.end(),
                              [ & ]( int x, int y ) { return blocks[ x ].size() - next[ x ] > blocks[ y ].size() - next[ y ]; } );
            std::vector<std::pair<int, int>> round;
            double cap = B->cus;
            for ( int i : order )
            {
                if ( next[ i ] >= blocks[ i ].size() )
                    break; // sorted: no instance after this one has blocks left
                if ( weight[ i ] > cap && !round.empty() )
                    continue;
                round.push_back( { i, (int)next[ i ]++ } );
                cap -= weight[ i ];
                left--;
            }
            rounds.push_back( std::move( round ) );
        }
    }
    for ( const auto& round : rounds )
    {
        PreLaunch entry;
        std::map<std::pair<int, int>, std::vector<apsm::Seq2Desc>> seq;
        std::map<std::pair<int, int>, std::pair<int, size_t>> seqDim; // (max users, max lds)
        for ( const auto& ib : round )
        {
            const TrainBlk& k = blocks[ ib.first ][ ib.second ];
            entry.add( k );
            const auto key = std::make_pair( k.tp, k.newPer );
            seq[ key ].push_back( k.s );
            auto& dim = seqDim[ key ];
            dim.first = std::max( dim.first, k.s.a.U );
            dim.second = std::max( dim.second, k.lds );
        }
        {
            std::vector<std::pair<std::pair<int, int>, size_t>> offs;
            int nComb = 0;
            const size_t combOff = entry.upload( arena, offs, nComb );
            pend.push_back( { 0, [ = ]( const char* base ) { launchRoundPreForked( B, entry, base, offs, combOff, nComb ); } } );
        }
        // the round's sequencer buckets are independent: fork them onto side streams
        struct SeqLaunch
        {
            size_t off;
            int cnt, tp, newPer, gx;
            size_t lds;
        };
        std::vector<SeqLaunch> sl;
        for ( auto& kv : seq )
            sl.push_back( SeqLaunch { arena.add( kv.second ), (int)kv.second.size(), kv.first.first, kv.first.second,
                                      seqDim[ kv.first ].first, seqDim[ kv.first ].second } );
        pend.push_back( { 0, [ = ]( const char* base ) {
                             const int nside = std::min<int>( kBatchSideStreams, (int)sl.size() - 1 );
                             if ( nside > 0 )
                                 HIP_OR_THROW( hipEventRecord( B->fork, B->stream ) );
                             for ( int k = 0; k < nside; k++ )
                                 HIP_OR_THROW( hipStreamWaitEvent( B->side[ k ], B->fork, 0 ) );
                             for ( size_t k = 0; k < sl.size(); k++ )
                             {
                                 const int lane = (int)( k % ( nside + 1 ) );
                                 hipStream_t st = lane == 0 ? B->stream : B->side[ lane - 1 ];
                                 const apsm::Seq2Desc* d = (const apsm::Seq2Desc*)( base + sl[ k ].off );
                                 if ( sl[ k ].newPer == 1 )
                                     launchBatchSeq2Tp<1>( st, sl[ k ].tp, d, sl[ k ].cnt, sl[ k ].gx, sl[ k ].lds );
                                 else
                                     launchBatchSeq2Tp<2>( st, sl[ k ].tp, d, sl[ k ].cnt, sl[ k ].gx, sl[ k ].lds );
                             }
                             for ( int k = 0; k < nside; k++ )
                             {
                                 HIP_OR_THROW( hipEventRecord( B->join[ k ], B->side[ k ] ) );
                                 HIP_OR_THROW( hipStreamWaitEvent( B->stream, B->join[ k ], 0 ) );
                             }
                         } } );
    }
    {
        int maxU = 1;
        for ( auto& t : theta )
            maxU = std::max( maxU, t.U );
        const size_t off = arena.add( theta );
        pend.push_back( { 0, [ = ]( const char* base ) {
                             hipLaunchKernelGGL( apsm::batch_theta, dim3( maxU, n ), dim3( 256 ), 0, B->stream,
                                                 (const apsm::ThetaDesc*)( base + off ) );
                             HIP_OR_THROW( hipGetLastError() );
                         } } );
    }
    const size_t trainLaunches = pend.size();

    // detect buckets by (kernel, NMAX, UMAX); work items for every candidate tile size. The
    // kernel (vector ALU or a matrix-core precision) is the one the instance's own handle runs,
    // so only launch shapes of the same arithmetic compete.
    std::map<std::array<int, 3>, std::vector<int>> dBuckets;
    for ( int i = 0; i < n; i++ )
        dBuckets[ { detectKernelOf( B->inst[ i ]->p.detect_precision, (uint32_t)det[ i ].N ), nmaxOf( det[ i ].N ),
                    umaxOf( det[ i ].U ) } ]
            .push_back( i );
    const size_t detOff = arena.add( det );
    struct DetBucket
    {
        int kernel, nmax, umax, nCand;
        int threads[ kTuneCands ], shape[ kTuneCands ];
        size_t itemOff[ kTuneCands ];
        int nItems[ kTuneCands ];
    };
    std::vector<DetBucket> dbs;
    for ( auto& kv : dBuckets )
    {
        DetBucket db {};
        db.kernel = kv.first[ 0 ];
        db.nmax = kv.first[ 1 ];
        db.umax = kv.first[ 2 ];
        db.nCand = 0;
        if ( db.kernel < 0 )
            for ( int t : kTuneThreads )
            {
                db.threads[ db.nCand ] = t;
                db.shape[ db.nCand++ ] = 0;
            }
        else
            for ( int sh = 0; sh < apsm::kMfmaShapes; sh++ )
                if ( apsm::mfmaShapeBuilt( sh, db.kernel, db.nmax, db.umax ) )
                {
                    db.threads[ db.nCand ] = 64 * apsm::kMfmaShape[ sh ].waves;
                    db.shape[ db.nCand++ ] = sh;
                }
        for ( int c = 0; c < db.nCand; c++ )
        {
            const int tileSamples = db.kernel < 0 ? db.threads[ c ] : apsm::mfmaShapeSamples( db.shape[ c ] );
            // instances with the largest dictionaries first (a tile's cost is proportional to
            // G): the workgroups of the persistent launch end on cheap tiles, not on a tail of
            // expensive ones
            std::vector<int> order( kv.second );
            std::stable_sort( order.begin(), order.end(), [ & ]( int x, int y ) { return det[ x ].G > det[ y ].G; } );
            std::vector<apsm::WorkItem> items;
            for ( int i : order )
                for ( int64_t t = 0; t < ( det[ i ].M + tileSamples - 1 ) / tileSamples; t++ )
                    items.push_back( apsm::WorkItem { i, (int)t } );
            db.itemOff[ c ] = arena.add( items );
            db.nItems[ c ] = (int)items.size();
        }
        dbs.push_back( db );
    }

    // upload the descriptors
    B->desc.ensure( arena.bytes.size() );
    HIP_OR_THROW( hipMemcpy( B->desc.ptr, arena.bytes.data(), arena.bytes.size(), hipMemcpyHostToDevice ) );
    const char* base = B->desc.ptr;
    for ( auto& pd : pend )
    {
        auto fn = pd.fn;
        B->launches.push_back( [ fn, base ]() { fn( base ); } );
    }

    // run the train once eagerly (the detect candidates are timed on the trained state)
    for ( size_t l = 0; l < trainLaunches; l++ )
        B->launches[ l ]();
    HIP_OR_THROW( hipStreamSynchronize( B->stream ) );

    // autotune: every (threads or tile shape, workgroups per CU) candidate per bucket, best of two
    // runs (launch shapes of the same arithmetic: the results do not depend on the choice)
    std::vector<std::function<void( hipStream_t )>> detLaunches;
    hipEvent_t t0, t1;
    HIP_OR_THROW( hipEventCreate( &t0 ) );
    HIP_OR_THROW( hipEventCreate( &t1 ) );
    for ( auto& db : dbs )
    {
        const apsm::DetDesc* dd = (const apsm::DetDesc*)( base + detOff );
        const int kernel = db.kernel, nmax = db.nmax, umax = db.umax;
        auto launch = [ = ]( hipStream_t st, int threads, int shape, const apsm::WorkItem* it, int nItems, int grid ) {
            if ( kernel < 0 )
                launchBatchDetectAny( st, nmax, umax, threads, dd, it, nItems, grid );
            else
                HIP_OR_THROW( apsm::launchBatchDetectMfma( kernel, nmax, umax, st, dd, it, nItems, grid, shape ) );
        };
        float best = 1e30f;
        int bc = 0, bb = 4;
        for ( int c = 0; c < db.nCand; c++ )
        {
            int lastGrid = -1;
            for ( int bpc : kTuneBlocksPerCu )
            {
                const int grid = std::max( 1, bpc == 0 ? db.nItems[ c ] : std::min( db.nItems[ c ], B->cus * bpc ) );
                if ( grid == lastGrid )
                    continue; // the same launch as the previous candidate
                lastGrid = grid;
                float ms = 1e30f;
                for ( int rep = 0; rep < 2; rep++ )
                {
                    HIP_OR_THROW( hipEventRecord( t0, B->stream ) );
                    launch( B->stream, db.threads[ c ], db.shape[ c ], (const apsm::WorkItem*)( base + db.itemOff[ c ] ), db.nItems[ c ],
                            grid );
                    HIP_OR_THROW( hipEventRecord( t1, B->stream ) );
                    HIP_OR_THROW( hipEventSynchronize( t1 ) );
                    float t = 0.f;
                    HIP_OR_THROW( hipEventElapsedTime( &t, t0, t1 ) );
                    ms = std::min( ms, t );
                }
                if ( ms < best )
                {
                    best = ms;
                    bc = c;
                    bb = bpc == 0 ? ( db.nItems[ c ] + B->cus - 1 ) / B->cus : bpc;
                }
            }
        }
        const int threads = db.threads[ bc ], shape = db.shape[ bc ], nItems = db.nItems[ bc ];
        const int grid = std::max( 1, std::min( nItems, B->cus * bb ) );
        const apsm::WorkItem* it = (const apsm::WorkItem*)( base + db.itemOff[ bc ] );
        detLaunches.push_back( [ = ]( hipStream_t st ) { launch( st, threads, shape, it, nItems, grid ); } );
        B->tuning.push_back( { nmax, umax, threads, bb, nItems, kernel, kernel < 0 ? apsm::kDetTile : apsm::mfmaShapeEntries( shape, nmax ),
                               kernel < 0 ? threads : apsm::mfmaShapeSamples( shape ) } );
    }
    // the buckets run side by side (their tails overlap instead of draining the GPU in turn)
    B->launches.push_back( [ B, detLaunches ]() {
        const int nb = (int)detLaunches.size();
        const int nside = std::min<int>( kBatchSideStreams, nb - 1 );
        if ( nside > 0 )
        {
            HIP_OR_THROW( hipEventRecord( B->fork, B->stream ) );
            for ( int k = 0; k < nside; k++ )
                HIP_OR_THROW( hipStreamWaitEvent( B->side[ k ], B->fork, 0 ) );
        }
        for ( int b = 0; b < nb; b++ )
        {
            const int lane = b % ( nside + 1 );
            detLaunches[ b ]( lane == 0 ? B->stream : B->side[ lane - 1 ] );
        }
        for ( int k = 0; k < nside; k++ )
        {
            HIP_OR_THROW( hipEventRecord( B->join[ k ], B->side[ k ] ) );
            HIP_OR_THROW( hipStreamWaitEvent( B->stream, B->join[ k ], 0 ) );
        }
    } );
    (void)hipEventDestroy( t0 );
    (void)hipEventDestroy( t1 );

    B->useGraph = useGraph;
    B->trainLaunches = trainLaunches;
    if ( useGraph )
    {
        // two graphs (train, detect) so a run can time its two parts
        for ( int g = 0; g < 2; g++ )
        {
            const size_t l0 = g == 0 ? 0 : trainLaunches, l1 = g == 0 ? trainLaunches : B->launches.size();
            HIP_OR_THROW( hipStreamBeginCapture( B->stream, hipStreamCaptureModeRelaxed ) );
            try
            {
                for ( size_t l = l0; l < l1; l++ )
                    B->launches[ l ]();
            }
            catch ( ... )
            {
                hipGraph_t gg = nullptr;
                (void)hipStreamEndCapture( B->stream, &gg );
                if ( gg )
                    (void)hipGraphDestroy( gg );
                throw;
            }
            HIP_OR_THROW( hipStreamEndCapture( B->stream, &B->graph[ g ] ) );
            HIP_OR_THROW( hipGraphInstantiate( &B->exec[ g ], B->graph[ g ], nullptr, nullptr, 0 ) );
        }
    }
    B->planned = true;
}

void batchRun( apsm_batch* B )
{
    ARG_CHECK( B->planned, "batch has no plan (apsm_batch_plan)" );
    HIP_OR_THROW( hipEventRecord( B->ev0, B->stream ) );
    if ( B->useGraph )
    {
        HIP_OR_THROW( hipGraphLaunch( B->exec[ 0 ], B->stream ) );
        HIP_OR_THROW( hipEventRecord( B->evMid, B->stream ) );
        HIP_OR_THROW( hipGraphLaunch( B->exec[ 1 ], B->stream ) );
    }
    else
    {
        for ( size_t l = 0; l < B->launches.size(); l++ )
        {
            if ( l == B->trainLaunches )
                HIP_OR_THROW( hipEventRecord( B->evMid, B->stream ) );
            B->launches[ l ]();
        }
    }
    HIP_OR_THROW( hipEventRecord( B->ev1, B->stream ) );
    B->timed = true;
}

} // namespace
