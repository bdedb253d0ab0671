// Host side of the serving engine (included by apsm_capi.hip after the batched path).
//
// A server holds links: detector handles whose state (stream buffers, dictionary, weights)
// stays on the device between frames, each admitted against an HBM budget with its buffers
// reserved up front, so that serving never allocates. A frame of a link is: detect its data
// samples with the link's current state, then train the link incrementally on the frame's
// pilots. Frames are served by continuous batching on two streams:
//   detect stream: every frame admitted since the previous detect batch is uploaded and
//     detected in one batch (a persistent batched detect per kernel bucket), and its estimates
//     are copied out; up to kServeDetectInFlight batches are queued;
//   train stream: rounds; a round starts the trains of frames whose detect was enqueued
//     (after an event wait on their detect batch), and advances every link train in progress
//     by one block (the blocks trainBlocked forms, as batched launches); up to
//     kServeRoundsInFlight rounds are queued.
// A frame arriving while others are on the GPU is detected in the next detect batch, without
// waiting for the trains, and its train joins the next round. A link takes its next frame once
// the round with its previous frame's last block has completed. Frame buffers come from a
// device arena of fixed size: a frame waits until the arena has room (first fit, in arrival
// order).
#pragma once

#include <chrono>
#include <deque>

constexpr int kServeRoundsInFlight = 2;
constexpr int kServeDetectInFlightMax = 8; // detect batch slots (apsm_server_config.detects_in_flight <= this)
constexpr int kServeDetectInFlight = 4; // default
// One stream each for detect batches, uploads, train rounds and one side stream for a round's
// further sequencer buckets: four streams, created in that order, so that each gets a hardware
// queue of its own (a process has four; streams beyond share them round-robin, and a detect
// batch behind a train round's sequencer on a shared queue would wait for it).
constexpr int kServeSideStreams = 1;
constexpr size_t kServeDescBytes = size_t( 1 ) << 20; // descriptor staging per batch / round
// A detect batch closes at this many frames or bytes, so that the upload of the next batch,
// this batch's detect and the trains overlap instead of running as one serial burst.
constexpr size_t kServeBatchFrames = 32;
constexpr size_t kServeBatchBytes = size_t( 64 ) << 20;

// Descriptors of one incremental train call of one handle on the blocked path.
struct CallPlan
{
    std::vector<apsm::CopyDesc> copies; // the live dictionary into the other stream buffer
    apsm::StageDesc stage {};
    std::vector<TrainBlk> blocks;
    apsm::ThetaDesc theta {};
    int64_t copyMax = 0;
};

namespace {

// Serving and batched links: blocked path, configured kernel variance, no LLS initialisation,
// no shuffle, at most 16 users.
void checkServable( const apsm_handle* h )
{
    const apsm_params& p = h->p;
    if ( !h->blocked || p.gaussian_kernel_variance == 0.f || p.lls_initialization || p.training_shuffle || h->U > 16 )
        throw ApsmError { APSM_ERR_UNSUPPORTED,
                          "served links need the blocked train path, a configured kernel variance, no LLS initialisation, "
                          "no shuffle and at most 16 users" };
}

// The detect of M samples with the handle's current state (as detectDev does for <= 16 users).
apsm::DetDesc detectDescOf( apsm_handle* h, const float* d_rx, int64_t M, float* d_est )
{
    if ( !h->varSet )
    {
        h->var = h->p.gaussian_kernel_variance * (float)( h->N * h->N );
        h->varSet = true;
    }
    const int64_t usedC = (int64_t)( h->used / 2 );
    const int64_t lo = std::max<int64_t>( 0, usedC - h->Hc );
    const int64_t G = usedC - lo;
    const StreamBuf& b = h->buf[ h->cur ];
    const int64_t off = lo - h->bufBaseC;
    return apsm::DetDesc { d_rx,
                           G > 0 ? b.Z.ptr + off * h->L : nullptr,
                           G > 0 ? b.Zn.ptr + off : nullptr,
                           G > 0 ? b.C.ptr + 2 * off : nullptr,
                           h->theta.ptr,
                           d_est,
                           M,
                           2 * b.capK,
                           G,
                           h->U,
                           h->N,
                           h->p.linear_kernel_weight,
                           h->p.gaussian_kernel_weight,
                           0.5f / h->var * kLog2e };
}

// One train call of T pilots (device pointers, host layouts) as descriptors: the same
// compaction, staging, blocks and theta as trainDev + trainBlocked, with the handle's
// bookkeeping advanced as they advance it. The stream buffers must already hold the call
// (reserved for the link's largest frame).
void planTrainCall( apsm_handle* h, const float* d_rx, const float* d_tx, int64_t T, CallPlan& cp )
{
    const apsm_params& p = h->p;
    const int N = h->N, L = h->L, U = h->U, w2 = h->w2, Wc = w2 / 2;
    ARG_CHECK( T > 0, "train needs at least one pilot" );
    if ( !h->varSet )
    {
        h->var = p.gaussian_kernel_variance * (float)( N * N );
        h->varSet = true;
    }
    h->anyTrain = true;
    const float wG = p.gaussian_kernel_weight, wL = p.linear_kernel_weight;
    const float g2 = 0.5f / h->var * kLog2e;

    const int64_t usedC = (int64_t)( h->used / 2 );
    const int64_t nOld = std::min<int64_t>( usedC, h->Hc );
    const int nxt = h->cur ^ 1;
    StreamBuf& nb = h->buf[ nxt ];
    const int64_t K = nOld + T;
    ARG_CHECK( K <= nb.capK, "frame has more pilots than the link reserved" );
    const int64_t ld = 2 * nb.capK;
    cp.copies.clear();
    cp.blocks.clear();
    if ( nOld > 0 )
    {
        const StreamBuf& ob = h->buf[ h->cur ];
        const int64_t off = usedC - nOld - h->bufBaseC;
        const int64_t ldOld = 2 * ob.capK;
        cp.copies.push_back( apsm::CopyDesc { ob.Z.ptr + off * L, nb.Z.ptr, nOld * L, 0, 0, 1 } );
        cp.copies.push_back( apsm::CopyDesc { ob.Zn.ptr + off, nb.Zn.ptr, nOld, 0, 0, 1 } );
        cp.copies.push_back( apsm::CopyDesc { ob.Y.ptr + 2 * off, nb.Y.ptr, 2 * nOld, ldOld, ld, U } );
        cp.copies.push_back( apsm::CopyDesc { ob.C.ptr + 2 * off, nb.C.ptr, 2 * nOld, ldOld, ld, U } );
        cp.copyMax = std::max<int64_t>( nOld * L, 2 * nOld * U );
    }
    cp.stage = apsm::StageDesc { d_rx, d_tx, nb.Z.ptr + nOld * L, nb.Zn.ptr + nOld, nb.Y.ptr + 2 * nOld, nb.C.ptr + 2 * nOld, T, ld, N, U };

    const int64_t q = p.sample_step;
    const int newPer = (int)( q / 2 );
    const int64_t usedOld = (int64_t)h->used, R2 = 2 * T, rowBase = 2 * ( usedC - nOld );
    const int64_t nStepsTotal = R2 / q;
    const int64_t steps = blockedSteps( h );
    const int tp = seq2Tp( Wc );
    const int nParts = std::min( apsm::kSeqThreads / Wc, ( Wc + tp - 1 ) / tp );
    if ( tp < 0 || nParts > apsm::kMaxParts || nParts * tp < Wc || nParts > pmaxOfTp( tp ) )
        throw ApsmError { APSM_ERR_UNSUPPORTED, "window layout does not fit the blocked train path" };
    int64_t a0Next = h->a0Next;
    for ( int64_t k0 = 0; k0 < nStepsTotal; k0 += steps )
    {
        const int64_t n = std::min( steps, nStepsTotal - k0 );
        const int64_t sPrev = usedOld + k0 * q, sFirst = sPrev + q, sLast = sPrev + n * q;
        const int64_t a0 = sFirst - std::min<int64_t>( w2, sFirst );
        const int64_t sNext = sLast + q;
        a0Next = sNext - std::min<int64_t>( w2, sNext );
        const int nNewC = (int)( n * newPer );
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
        ea.rowBase = rowBase;
        ea.newC0 = ( sPrev - rowBase ) / 2;
        ea.sFirst = sFirst;
        ea.H = h->H;
        ea.a0 = a0;
        ea.ldYC = ld;
        ea.nNewC = nNewC;
        ea.wL = wL;
        ea.wG = wG;
        ea.g2 = g2;
        k.nmax = nmaxOf( N );
        k.e.Z = nb.Z.ptr;
        k.e.Zn = nb.Zn.ptr;
        k.e.C = nb.C.ptr;
        k.e.thF = h->thOld.ptr;
        k.e.P = h->Pbuf.ptr;
        k.e.part = h->Ppart.ptr;
        apsm::entryChunks( ea, k.e.jBeg, k.e.nChunks );
        apsm::GramArgs& ga = k.g.a;
        ga.N = N;
        ga.L = L;
        ga.c0 = ( a0 - rowBase ) / 2;
        ga.nA = nActC;
        ga.wL = wL;
        ga.wG = wG;
        ga.g2 = g2;
        k.g.Z = nb.Z.ptr;
        k.g.Zn = nb.Zn.ptr;
        k.g.Gact = h->Gact.ptr;
        const int64_t loLast = sLast > h->H ? sLast - h->H : 0, loPrev = sPrev > h->H ? sPrev - h->H : 0;
        k.hasEv = loLast > loPrev;
        apsm::EvArgs& ev = k.v.a;
        ev.N = N;
        ev.L = L;
        ev.U = U;
        ev.q = (int)q;
        ev.w2 = w2;
        ev.nSteps = (int)n;
        ev.rowBase = rowBase;
        ev.sPrev = sPrev;
        ev.sFirst = sFirst;
        ev.H = h->H;
        ev.ldYC = ld;
        ev.wG = wG;
        ev.g2 = g2;
        k.v.Z = nb.Z.ptr;
        k.v.Zn = nb.Zn.ptr;
        k.v.C = nb.C.ptr;
        k.v.EvF = h->EvF.ptr;
        apsm::Seq2Args& s2 = k.s.a;
        s2.N = N;
        s2.L = L;
        s2.U = U;
        s2.w2 = w2;
        s2.Wc = Wc;
        s2.q = (int)q;
        s2.nSteps = (int)n;
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
        const int64_t aR = a0 - rowBase;
        k.s.Z = nb.Z.ptr + ( aR / 2 ) * L;
        k.s.Zn = nb.Zn.ptr + aR / 2;
        k.s.Y = nb.Y.ptr + aR;
        k.s.C = nb.C.ptr + aR;
        k.s.Fg = h->F.ptr;
        k.s.thF = h->thOld.ptr;
        k.s.P = h->Pbuf.ptr;
        k.s.Gact = h->Gact.ptr;
        k.s.EvF = h->EvF.ptr;
        k.lds = apsm::seq2LdsBytes( newPer, w2, tp );
        cp.blocks.push_back( k );
    }
    // bookkeeping, as trainDev
    const int64_t first = ( p.start_with_full_window && h->used == 0 ) ? w2 : q;
    if ( R2 >= first )
        h->used += (uint64_t)( first + ( ( R2 - first ) / q ) * q );
    h->a0Next = a0Next;
    h->cur = nxt;
    h->bufBaseC = usedC - nOld;
    h->bufK = K;
    const int64_t newUsedC = (int64_t)( h->used / 2 );
    const int64_t lo = std::min<int64_t>( h->a0Next / 2, newUsedC );
    cp.theta = apsm::ThetaDesc { nb.Z.ptr, nb.C.ptr, h->thOld.ptr, h->theta.ptr, lo - h->bufBaseC, newUsedC - h->bufBaseC, ld, N, U };
}

// First-fit allocator over [0, size) in 256-byte units, free ranges kept sorted and merged.
struct RangeAlloc
{
    std::map<size_t, size_t> freeR; // offset -> length
    void reset( size_t size )
    {
        freeR.clear();
        if ( size > 0 )
            freeR[ 0 ] = size;
    }
    bool alloc( size_t n, size_t& off )
    {
        n = ( n + 255 ) & ~size_t( 255 );
        for ( auto it = freeR.begin(); it != freeR.end(); ++it )
            if ( it->second >= n )
            {
                off = it->first;
                const size_t rest = it->second - n;
                freeR.erase( it );
                if ( rest > 0 )
                    freeR[ off + n ] = rest;
                return true;
            }
        return false;
    }
    void release( size_t off, size_t n )
    {
        n = ( n + 255 ) & ~size_t( 255 );
        auto it = freeR.emplace( off, n ).first;
        auto nx = std::next( it );
        if ( nx != freeR.end() && it->first + it->second == nx->first )
        {
            it->second += nx->second;
            freeR.erase( nx );
        }
        if ( it != freeR.begin() )
        {
            auto pv = std::prev( it );
            if ( pv->first + pv->second == it->first )
            {
                pv->second += it->second;
                freeR.erase( it );
            }
        }
    }
};

} // namespace

struct ServeLink
{
    apsm_handle* h = nullptr;
    size_t bytes = 0;
    uint32_t maxPilots = 0;
    bool busy = false; // a frame of this link is admitted and its train has not completed
};

struct ServeFrame
{
    apsm_frame f {};
    uint64_t ticket = 0;
    double arrival = 0.0; // server clock (s)
    double admit = -1.0, output = -1.0, trained = -1.0;
    size_t arenaOff = 0, arenaBytes = 0;
    float *dRxD = nullptr, *dRxP = nullptr, *dTxP = nullptr, *dEst = nullptr;
    CallPlan plan;
    size_t nextBlock = 0;
    int64_t detectBatch = -1, lastRound = -1;
    int32_t status = APSM_OK;
};

// Pinned host staging and device copy of one launch's descriptors.
struct DescSlot
{
    DevBuf<char> dev;
    char* host = nullptr;
    size_t hostCap = 0;
    // Allocates the staging at server creation: growing it later frees memory, and the
    // runtime's frees wait for every stream of the device (a stall of the whole pipeline).
    void reserve( size_t bytes )
    {
        HIP_OR_THROW( hipHostMalloc( (void**)&host, bytes, hipHostMallocDefault ) );
        hostCap = bytes;
        dev.ensure( bytes );
    }
    const char* upload( const DescArena& a, hipStream_t st )
    {
        const size_t bytes = std::max<size_t>( a.bytes.size(), 256 );
        if ( hostCap < bytes )
        {
            if ( host )
                HIP_OR_THROW( hipHostFree( host ) );
            host = nullptr;
            hostCap = 0;
            HIP_OR_THROW( hipHostMalloc( (void**)&host, bytes * 2, hipHostMallocDefault ) );
            hostCap = bytes * 2;
        }
        std::memcpy( host, a.bytes.data(), a.bytes.size() );
        dev.ensure( hostCap );
        HIP_OR_THROW( hipMemcpyAsync( dev.ptr, host, a.bytes.size(), hipMemcpyHostToDevice, st ) );
        return dev.ptr;
    }
    void release()
    {
        if ( host )
            (void)hipHostFree( host );
        host = nullptr;
        hostCap = 0;
        dev.release();
    }
};

struct ServeDetect
{
    int64_t id = -1;
    hipEvent_t uploaded = nullptr, detected = nullptr;
    std::vector<ServeFrame*> frames; // frames without pilots finish here
    DescSlot desc;
};

struct ServeRound
{
    int64_t id = -1;
    hipEvent_t done = nullptr;
    std::vector<ServeFrame*> finishing;
    DescSlot desc;
};

struct apsm_server
{
    apsm_server_config cfg {};
    std::vector<ServeLink> links;
    size_t linkBytes = 0;
    DevBuf<char> arena;
    RangeAlloc arenaAlloc;
    hipStream_t stream = nullptr; // train rounds
    hipStream_t dstream = nullptr; // detect batches
    hipStream_t ustream = nullptr; // frame uploads (overlap the previous batch's detect)
    hipStream_t side[ kServeSideStreams ] = {};
    hipEvent_t fork = nullptr, join[ kServeSideStreams ] = {};
    ServeRound rounds[ kServeRoundsInFlight ];
    ServeDetect detects[ kServeDetectInFlightMax ];
    int detectSlots = kServeDetectInFlight; // batches in flight at most
    size_t batchFrames = kServeBatchFrames, batchBytes = kServeBatchBytes; // a detect batch closes here
    std::deque<int64_t> roundsInFlight, detectsInFlight; // ids, oldest first
    int64_t nextRound = 0, nextDetect = 0;
    std::deque<ServeFrame*> waiting; // arrival order
    std::vector<ServeFrame*> pendingTrain; // detected (enqueued), train not started
    std::vector<ServeFrame*> training; // trains in progress (blocks left)
    std::vector<ServeFrame*> finished; // completed, not yet reported by poll
    uint64_t nextTicket = 0;
    std::chrono::steady_clock::time_point epoch;
    int cus = 256;
    uint64_t roundsRun = 0, detectsRun = 0, maxBatch = 0;
};

namespace {

double serverNow( const apsm_server* S )
{
    return std::chrono::duration<double>( std::chrono::steady_clock::now() - S->epoch ).count();
}

size_t handleDeviceBytes( const apsm_handle* h )
{
    size_t b = 0;
    for ( const StreamBuf& s : h->buf )
        b += ( s.Z.cap + s.Zn.cap + s.Y.cap + s.C.cap ) * sizeof( float );
    b += ( h->F.cap + h->Gram.cap + h->thOld.cap + h->theta.cap + h->Pbuf.cap + h->Gact.cap + h->EvF.cap + h->Ppart.cap + h->GramG.cap +
           h->FG.cap ) *
        sizeof( float );
    b += h->Qn.cap * sizeof( double );
    return b;
}

size_t alignUp256( size_t n ) { return ( n + 255 ) & ~size_t( 255 ); }

size_t frameBytes( const ServeLink& L, const apsm_frame& f )
{
    const size_t N = (size_t)L.h->N, U = (size_t)L.h->U;
    return alignUp256( N * f.n_data * 8 ) + alignUp256( N * f.n_pilots * 8 ) + alignUp256( U * f.n_pilots * 8 ) +
        alignUp256( U * f.n_data * 8 );
}

bool serverIdle( const apsm_server* S )
{
    return S->waiting.empty() && S->pendingTrain.empty() && S->training.empty() && S->roundsInFlight.empty() && S->detectsInFlight.empty();
}

void finishFrame( apsm_server* S, ServeFrame* f, double t )
{
    f->trained = t;
    S->arenaAlloc.release( f->arenaOff, f->arenaBytes );
    S->links[ f->f.link ].busy = false;
    S->finished.push_back( f );
}

// Retires completed detect batches and train rounds (each in order).
void serverRetire( apsm_server* S )
{
    while ( !S->detectsInFlight.empty() )
    {
        ServeDetect& D = S->detects[ S->detectsInFlight.front() % S->detectSlots ];
        const hipError_t e = hipEventQuery( D.detected );
        if ( e == hipErrorNotReady )
            break;
        HIP_OR_THROW( e );
        const double t = serverNow( S );
        for ( ServeFrame* f : D.frames )
        {
            f->output = t;
            if ( f->f.n_pilots == 0 )
                finishFrame( S, f, t );
        }
        D.frames.clear();
        S->detectsInFlight.pop_front();
    }
    while ( !S->roundsInFlight.empty() )
    {
        ServeRound& R = S->rounds[ S->roundsInFlight.front() % kServeRoundsInFlight ];
        const hipError_t e = hipEventQuery( R.done );
        if ( e == hipErrorNotReady )
            break;
        HIP_OR_THROW( e );
        const double t = serverNow( S );
        for ( ServeFrame* f : R.finishing )
            finishFrame( S, f, t );
        R.finishing.clear();
        S->roundsInFlight.pop_front();
    }
}

// Admits waiting frames and launches their detect batch on the detect stream; returns whether
// it launched one.
bool serverLaunchDetect( apsm_server* S )
{
    if ( S->waiting.empty() || (int)S->detectsInFlight.size() >= S->detectSlots )
        return false;
    const double now = serverNow( S );
    // arrival order; a frame whose link is busy waits, a frame without arena room stops admission
    std::vector<ServeFrame*> admitted;
    size_t batchBytes = 0;
    for ( auto it = S->waiting.begin(); it != S->waiting.end(); )
    {
        if ( admitted.size() >= S->batchFrames || batchBytes >= S->batchBytes )
            break;
        ServeFrame* f = *it;
        ServeLink& L = S->links[ f->f.link ];
        if ( L.busy )
        {
            ++it;
            continue;
        }
        size_t off = 0;
        if ( !S->arenaAlloc.alloc( f->arenaBytes, off ) )
            break;
        f->arenaOff = off;
        char* base = S->arena.ptr + off;
        const size_t N = (size_t)L.h->N, U = (size_t)L.h->U;
        f->dRxD = (float*)base;
        base += alignUp256( N * f->f.n_data * 8 );
        f->dRxP = (float*)base;
        base += alignUp256( N * f->f.n_pilots * 8 );
        f->dTxP = (float*)base;
        base += alignUp256( U * f->f.n_pilots * 8 );
        f->dEst = (float*)base;
        f->admit = now;
        L.busy = true;
        batchBytes += f->arenaBytes;
        admitted.push_back( f );
        it = S->waiting.erase( it );
    }
    if ( admitted.empty() )
        return false;

    const int64_t id = S->nextDetect++;
    ServeDetect& D = S->detects[ id % S->detectSlots ];
    D.id = id;
    D.frames = admitted;
    DescArena arena;
    std::vector<apsm::DetDesc> det;
    std::map<std::array<int, 3>, std::vector<int>> dBuckets;
    for ( ServeFrame* f : admitted )
    {
        apsm_handle* h = S->links[ f->f.link ].h;
        det.push_back( detectDescOf( h, f->dRxD, f->f.n_data, f->dEst ) );
        dBuckets[ { detectKernelOf( h->p.detect_precision, (uint32_t)h->N ), nmaxOf( h->N ), umaxOf( h->U ) } ].push_back(
            (int)det.size() - 1 );
        f->detectBatch = id;
    }
    const size_t detOff = arena.add( det );
    struct DetL
    {
        int kernel, nmax, umax, nItems;
        size_t itemOff;
    };
    std::vector<DetL> detL;
    for ( auto& kv : dBuckets )
    {
        const int kernel = kv.first[ 0 ];
        const int tile = kernel < 0 ? apsm::kDetThreads : apsm::mfmaShapeSamples( 0 );
        // frames of the largest dictionaries first (see the batched detect)
        std::vector<int> order( kv.second );
        std::stable_sort( order.begin(), order.end(), [ & ]( int x, int y ) { return det[ x ].G > det[ y ].G; } );
        std::vector<apsm::WorkItem> items;
        for ( int i : order )
            for ( int64_t t = 0; t < ( det[ i ].M + tile - 1 ) / tile; t++ )
                items.push_back( apsm::WorkItem { i, (int)t } );
        detL.push_back( DetL { kernel, kv.first[ 1 ], kv.first[ 2 ], (int)items.size(), arena.add( items ) } );
    }
    // the large copies (data and pilot samples) go on the upload stream; the small ones
    // (descriptors, transmitted pilots) on the detect stream, where the runtime's copy kernel
    // for them does not interleave with the upload stream's engine copies
    hipStream_t st = S->dstream, up = S->ustream;
    const char* base = D.desc.upload( arena, st );
    for ( const ServeFrame* f : admitted )
    {
        const apsm_handle* h = S->links[ f->f.link ].h;
        const size_t N = (size_t)h->N, U = (size_t)h->U;
        HIP_OR_THROW( hipMemcpyAsync( f->dRxD, f->f.rx_data, N * f->f.n_data * 8, hipMemcpyHostToDevice, up ) );
        if ( f->f.n_pilots > 0 )
        {
            HIP_OR_THROW( hipMemcpyAsync( f->dRxP, f->f.rx_pilots, N * f->f.n_pilots * 8, hipMemcpyHostToDevice, up ) );
            HIP_OR_THROW( hipMemcpyAsync( f->dTxP, f->f.tx_pilots, U * f->f.n_pilots * 8, hipMemcpyHostToDevice, st ) );
        }
    }
    HIP_OR_THROW( hipEventRecord( D.uploaded, up ) );
    HIP_OR_THROW( hipStreamWaitEvent( st, D.uploaded, 0 ) );
    const apsm::DetDesc* dd = (const apsm::DetDesc*)( base + detOff );
    for ( const DetL& d : detL )
    {
        const int grid = std::max( 1, std::min( d.nItems, S->cus * 4 ) );
        const apsm::WorkItem* it = (const apsm::WorkItem*)( base + d.itemOff );
        if ( d.kernel < 0 )
            launchBatchDetectAny( st, d.nmax, d.umax, apsm::kDetThreads, dd, it, d.nItems, grid );
        else
            HIP_OR_THROW( apsm::launchBatchDetectMfma( d.kernel, d.nmax, d.umax, st, dd, it, d.nItems, grid ) );
    }
    for ( ServeFrame* f : admitted )
    {
        const size_t U = (size_t)S->links[ f->f.link ].h->U;
        HIP_OR_THROW( hipMemcpyAsync( f->f.est, f->dEst, U * f->f.n_data * 8, hipMemcpyDeviceToHost, st ) );
    }
    HIP_OR_THROW( hipEventRecord( D.detected, st ) );
    for ( ServeFrame* f : admitted )
        if ( f->f.n_pilots > 0 )
            S->pendingTrain.push_back( f );
    S->detectsInFlight.push_back( id );
    S->detectsRun++;
    S->maxBatch = std::max<uint64_t>( S->maxBatch, admitted.size() );
    return true;
}

// Builds and launches one train round on the train stream; returns whether it launched one.
bool serverLaunchTrainRound( apsm_server* S )
{
    if ( (int)S->roundsInFlight.size() >= kServeRoundsInFlight || ( S->pendingTrain.empty() && S->training.empty() ) )
        return false;
    const int64_t rid = S->nextRound++;
    ServeRound& R = S->rounds[ rid % kServeRoundsInFlight ];
    R.id = rid;
    R.finishing.clear();
    hipStream_t st = S->stream;

    // trains starting now: after their detect batch (it reads the state they change)
    std::vector<int64_t> waitOn;
    for ( ServeFrame* f : S->pendingTrain )
    {
        planTrainCall( S->links[ f->f.link ].h, f->dRxP, f->dTxP, f->f.n_pilots, f->plan );
        f->nextBlock = 0;
        if ( std::find( waitOn.begin(), waitOn.end(), f->detectBatch ) == waitOn.end() )
            waitOn.push_back( f->detectBatch );
    }
    for ( int64_t b : waitOn )
    {
        // a batch still in flight: wait for its event (a retired one has completed)
        if ( std::find( S->detectsInFlight.begin(), S->detectsInFlight.end(), b ) != S->detectsInFlight.end() )
            HIP_OR_THROW( hipStreamWaitEvent( st, S->detects[ b % S->detectSlots ].detected, 0 ) );
    }
    std::vector<apsm::CopyDesc> copies;
    std::vector<apsm::StageDesc> stage;
    int64_t copyMax = 0, maxS = 0;
    for ( ServeFrame* f : S->pendingTrain )
    {
        copies.insert( copies.end(), f->plan.copies.begin(), f->plan.copies.end() );
        copyMax = std::max( copyMax, f->plan.copyMax );
        stage.push_back( f->plan.stage );
        maxS = std::max<int64_t>( maxS, f->f.n_pilots );
        S->training.push_back( f );
    }
    S->pendingTrain.clear();

    DescArena arena;
    PreLaunch entry;
    std::map<std::pair<int, int>, std::vector<apsm::Seq2Desc>> seq;
    std::map<std::pair<int, int>, std::pair<int, size_t>> seqDim;
    std::vector<apsm::ThetaDesc> theta;
    int maxU = 1;
    for ( auto it = S->training.begin(); it != S->training.end(); )
    {
        ServeFrame* f = *it;
        const TrainBlk& k = f->plan.blocks[ f->nextBlock++ ];
        entry.add( k );
        const auto key = std::make_pair( k.tp, k.newPer );
        seq[ key ].push_back( k.s );
        auto& dim = seqDim[ key ];
        dim.first = std::max( dim.first, k.s.a.U );
        dim.second = std::max( dim.second, k.lds );
        if ( f->nextBlock == f->plan.blocks.size() )
        {
            theta.push_back( f->plan.theta );
            maxU = std::max( maxU, f->plan.theta.U );
            f->lastRound = rid;
            R.finishing.push_back( f );
            it = S->training.erase( it );
        }
        else
            ++it;
    }
    const size_t copyOff = arena.add( copies ), stageOff = arena.add( stage ), thetaOff = arena.add( theta );
    std::vector<std::pair<std::pair<int, int>, size_t>> entryOffs;
    int nComb = 0;
    const size_t combOff = entry.upload( arena, entryOffs, nComb );
    struct SeqL
    {
        size_t off;
        int cnt, tp, newPer, gx;
        size_t lds;
    };
    std::vector<SeqL> sl;
    for ( auto& kv : seq )
        sl.push_back( SeqL { arena.add( kv.second ), (int)kv.second.size(), kv.first.first, kv.first.second, seqDim[ kv.first ].first,
                             seqDim[ kv.first ].second } );
    const char* base = R.desc.upload( arena, st );

    if ( !copies.empty() )
    {
        hipLaunchKernelGGL( apsm::batch_copy, dim3( std::min( 64, gridFor( copyMax, 256 ) ), (unsigned)copies.size() ), dim3( 256 ), 0, st,
                            (const apsm::CopyDesc*)( base + copyOff ) );
        HIP_OR_THROW( hipGetLastError() );
    }
    if ( !stage.empty() )
    {
        hipLaunchKernelGGL( apsm::batch_stage, dim3( gridFor( maxS, 256 ), (unsigned)stage.size() ), dim3( 256 ), 0, st,
                            (const apsm::StageDesc*)( base + stageOff ) );
        HIP_OR_THROW( hipGetLastError() );
    }
    launchRoundPre( st, entry, base, entryOffs, combOff, nComb );
    if ( !sl.empty() )
    {
        const int nside = std::min<int>( kServeSideStreams, (int)sl.size() - 1 );
        if ( nside > 0 )
            HIP_OR_THROW( hipEventRecord( S->fork, st ) );
        for ( int k = 0; k < nside; k++ )
            HIP_OR_THROW( hipStreamWaitEvent( S->side[ k ], S->fork, 0 ) );
        for ( size_t k = 0; k < sl.size(); k++ )
        {
            const int lane = (int)( k % ( nside + 1 ) );
            hipStream_t ss = lane == 0 ? st : S->side[ lane - 1 ];
            const apsm::Seq2Desc* d = (const apsm::Seq2Desc*)( base + sl[ k ].off );
            if ( sl[ k ].newPer == 1 )
                launchBatchSeq2Tp<1>( ss, sl[ k ].tp, d, sl[ k ].cnt, sl[ k ].gx, sl[ k ].lds );
            else
                launchBatchSeq2Tp<2>( ss, sl[ k ].tp, d, sl[ k ].cnt, sl[ k ].gx, sl[ k ].lds );
        }
        for ( int k = 0; k < nside; k++ )
        {
            HIP_OR_THROW( hipEventRecord( S->join[ k ], S->side[ k ] ) );
            HIP_OR_THROW( hipStreamWaitEvent( st, S->join[ k ], 0 ) );
        }
    }
    if ( !theta.empty() )
    {
        hipLaunchKernelGGL( apsm::batch_theta, dim3( maxU, (unsigned)theta.size() ), dim3( 256 ), 0, st,
                            (const apsm::ThetaDesc*)( base + thetaOff ) );
        HIP_OR_THROW( hipGetLastError() );
    }
    HIP_OR_THROW( hipEventRecord( R.done, st ) );
    S->roundsInFlight.push_back( rid );
    S->roundsRun++;
    return true;
}

// One step of the serving loop: retire, then launch what can be launched.
void serverProgress( apsm_server* S )
{
    serverRetire( S );
    serverLaunchDetect( S );
    serverLaunchTrainRound( S );
}

// Throws APSM_ERR_ARG for a frame the server cannot take; enqueues nothing.
void serverCheckFrame( const apsm_server* S, const apsm_frame& f )
{
    ARG_CHECK( f.link < S->links.size() && S->links[ f.link ].h, "unknown link" );
    const ServeLink& L = S->links[ f.link ];
    ARG_CHECK( f.rx_data && f.est && f.n_data > 0, "frame without data samples" );
    ARG_CHECK( f.n_pilots == 0 || ( f.rx_pilots && f.tx_pilots ), "frame pilots missing" );
    ARG_CHECK( f.n_pilots <= L.maxPilots, "frame has more pilots than the link reserved" );
    ARG_CHECK( frameBytes( L, f ) <= S->cfg.frame_arena, "frame larger than the frame arena" );
}

void serverEnqueue( apsm_server* S, const apsm_frame& f, double arrival, uint64_t ticket )
{
    serverCheckFrame( S, f );
    auto* sf = new ServeFrame;
    sf->f = f;
    sf->ticket = ticket;
    sf->arrival = arrival;
    sf->arenaBytes = frameBytes( S->links[ f.link ], f );
    S->waiting.push_back( sf );
}

void fillResult( const ServeFrame* f, double t0, apsm_frame_result* r )
{
    r->status = f->status;
    r->ticket = f->ticket;
    r->arrival_s = f->arrival - t0;
    r->admit_s = f->admit - t0;
    r->output_s = f->output - t0;
    r->trained_s = f->trained - t0;
    r->detect_round = f->detectBatch;
    r->last_round = f->lastRound;
}

// Deletes every frame the server still holds (each is in exactly one list).
void serverDropFrames( apsm_server* S )
{
    std::vector<ServeFrame*> all( S->waiting.begin(), S->waiting.end() );
    all.insert( all.end(), S->pendingTrain.begin(), S->pendingTrain.end() );
    all.insert( all.end(), S->training.begin(), S->training.end() );
    all.insert( all.end(), S->finished.begin(), S->finished.end() );
    for ( ServeRound& R : S->rounds )
    {
        all.insert( all.end(), R.finishing.begin(), R.finishing.end() );
        R.finishing.clear();
    }
    for ( ServeDetect& D : S->detects )
    {
        // frames of a detect batch are also in pendingTrain, training, a round or finished,
        // except frames without pilots, which finish at the batch
        for ( ServeFrame* f : D.frames )
            if ( f->f.n_pilots == 0 )
                all.push_back( f );
        D.frames.clear();
    }
    std::sort( all.begin(), all.end() );
    all.erase( std::unique( all.begin(), all.end() ), all.end() );
    for ( ServeFrame* f : all )
        delete f;
    S->waiting.clear();
    S->pendingTrain.clear();
    S->training.clear();
    S->finished.clear();
}

// After a failure inside the serving loop: waits for everything enqueued on the server's
// streams, then drops every frame it holds (their host buffers are no longer touched) and
// returns the server to idle (arena free, no link busy, nothing in flight).
void serverAbort( apsm_server* S )
{
    for ( hipStream_t st : { S->stream, S->dstream, S->ustream } )
        if ( st )
            (void)hipStreamSynchronize( st );
    for ( hipStream_t st : S->side )
        if ( st )
            (void)hipStreamSynchronize( st );
    serverDropFrames( S );
    S->roundsInFlight.clear();
    S->detectsInFlight.clear();
    S->arenaAlloc.reset( (size_t)S->cfg.frame_arena );
    for ( ServeLink& L : S->links )
        L.busy = false;
}

} // namespace
