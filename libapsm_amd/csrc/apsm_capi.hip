// Host side of libapsm_amd: the C ABI declared in include/apsm_capi.h.
//
// One apsm_handle owns its device state (stream buffer, window state, linear weights), one
// HIP stream and two events. Host entry points upload into handle-owned device buffers and
// call the *_dev entry points; those enqueue kernels on the handle's stream.

#include "apsm_capi.h"
#include "apsm_device.h"
#include "apsm_train2.h"
#include "apsm_train_nc.h"
#include "apsm_batch.h"
#include "apsm_detect_launch.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <functional>
#include <cmath>
#include <complex>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <random>
#include <vector>
#include <stdexcept>
#include <string>
#include <vector>

namespace {

thread_local std::string g_lastError;

struct ApsmError
{
    int code;
    std::string msg;
};

#define HIP_OR_THROW( expr )                                                                                                  \
    do                                                                                                                        \
    {                                                                                                                         \
        hipError_t e_ = ( expr );                                                                                             \
        if ( e_ != hipSuccess )                                                                                               \
            throw ApsmError { APSM_ERR_DEVICE, std::string( #expr ) + ": " + hipGetErrorString( e_ ) };                      \
    } while ( 0 )

#define ARG_CHECK( cond, msg )                                                                                                \
    do                                                                                                                        \
    {                                                                                                                         \
        if ( !( cond ) )                                                                                                      \
            throw ApsmError { APSM_ERR_ARG, msg };                                                                            \
    } while ( 0 )

template <typename F>
int guarded( F&& f )
{
    try
    {
        f();
        return APSM_OK;
    }
    catch ( const ApsmError& e )
    {
        g_lastError = e.msg;
        return e.code;
    }
    catch ( const std::exception& e )
    {
        g_lastError = e.what();
        return APSM_ERR_ARG;
    }
}

// Device buffer that grows on demand (contents are not preserved on growth).
template <typename T>
struct DevBuf
{
    T* ptr = nullptr;
    size_t cap = 0;
    DevBuf() = default;
    DevBuf( const DevBuf& ) = delete;
    DevBuf& operator=( const DevBuf& ) = delete;
    ~DevBuf() { release(); }
    void ensure( size_t n )
    {
        if ( n <= cap )
            return;
        if ( ptr )
            HIP_OR_THROW( hipFree( ptr ) );
        ptr = nullptr;
        cap = 0;
        if ( n == 0 )
            return;
        HIP_OR_THROW( hipMalloc( &ptr, n * sizeof( T ) ) );
        cap = n;
    }
    void release()
    {
        if ( ptr )
            (void)hipFree( ptr );
        ptr = nullptr;
        cap = 0;
    }
};

struct StreamBuf
{
    DevBuf<float> Z, Zn, Y, C;
    int64_t capK = 0;
    void ensure( int64_t K, int N, int U )
    {
        if ( K <= capK )
            return;
        int64_t k2 = std::max<int64_t>( K, capK + capK / 2 );
        Z.ensure( (size_t)k2 * 2 * N );
        Zn.ensure( (size_t)k2 );
        Y.ensure( (size_t)k2 * 2 * U );
        C.ensure( (size_t)k2 * 2 * U );
        capK = k2;
    }
    void release()
    {
        Z.release();
        Zn.release();
        Y.release();
        C.release();
        capK = 0;
    }
};

constexpr float kLog2e = 1.4426950408889634f;

// Template buckets of the antenna count and of a user group.
int nmaxOf( int N ) { return N <= 4 ? 4 : N <= 8 ? 8 : N <= 16 ? 16 : N <= 32 ? 32 : 64; }
int umaxOf( int U ) { return U <= 1 ? 1 : U <= 2 ? 2 : U <= 4 ? 4 : U <= 8 ? 8 : U <= 12 ? 12 : 16; }

// APSM_PRECISION_FP32 runs the cross terms on the matrix cores from this many antennas on
// (below, the vector-ALU kernel is at least as fast).
constexpr uint32_t kMatrixFp32MinAntennas = 1;

// Permutation of the reference shuffle (cpp/lib/noma/noma_detector.cu:59-74).
std::vector<int32_t> shufflePermutation( int64_t T )
{
    std::vector<int32_t> perm( T );
    for ( int64_t i = 0; i < T; i++ )
        perm[ i ] = (int32_t)i;
    std::seed_seq seed { 1, 2, 3, 4, 5, 6, 7, 8 };
    std::mt19937 rng( seed );
    std::shuffle( perm.begin(), perm.end(), rng );
    return perm;
}

int gridFor( int64_t n, int threads ) { return (int)( ( n + threads - 1 ) / threads ); }

} // namespace

// ---------------------------------------------------------------------------------------
// LLS
// ---------------------------------------------------------------------------------------
constexpr int kLlsMaxDims = 128; // unknowns of one least-squares solve (the FP64 Cholesky is LDS-resident)

struct lls_handle
{
    int N = 0, L = 0;
    hipStream_t stream = nullptr;
    bool ownStream = true; // false when the caller lent its stream (lls_create_on_stream)
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    DevBuf<float> w, part, rx, tx, est;
    DevBuf<double> gram, sums;
    DevBuf<int> status;
    int nW = 0; // unknowns of the last solve (2N for the embedding, num_input_dims for the low-level call)
    bool trained = false;
};

namespace {

// FP64 sum of the nPart partial normal equations of n unknowns, then the FP64 Cholesky solve
// into h->w (first n entries) and h->status.
void llsReduceSolve( lls_handle* h, int nPart, int n )
{
    const int E = n * ( n + 1 );
    h->nW = n;
    h->gram.ensure( (size_t)E );
    hipLaunchKernelGGL( apsm::lls_reduce_partials, dim3( gridFor( E, 256 ) ), dim3( 256 ), 0, h->stream, h->part.ptr, nPart, E,
                        h->gram.ptr );
    HIP_OR_THROW( hipGetLastError() );
    const size_t ldsSolve = (size_t)E * sizeof( double );
    if ( ldsSolve > 64 * 1024 )
        HIP_OR_THROW( hipFuncSetAttribute( (const void*)apsm::lls_solve, hipFuncAttributeMaxDynamicSharedMemorySize, (int)ldsSolve ) );
    hipLaunchKernelGGL( apsm::lls_solve, dim3( 1 ), dim3( apsm::kSolveThreads ), ldsSolve, h->stream, h->gram.ptr, n, h->w.ptr,
                        h->status.ptr );
    HIP_OR_THROW( hipGetLastError() );
}

void llsTrainDev( lls_handle* h, const float* d_rx, const float* d_tx, int64_t T )
{
    ARG_CHECK( T > 0, "LLS train needs at least one pilot" );
    const int L = h->L;
    const int E = L * ( L + 1 );
    const int64_t chunk = std::max<int64_t>( 1, (int64_t)( apsm::kLlsLdsBudget / ( 2 * ( L + 1 ) * sizeof( float ) ) ) );
    const int nPart = (int)( ( T + chunk - 1 ) / chunk );
    h->part.ensure( (size_t)nPart * E );
    const size_t lds = 2 * (size_t)chunk * ( L + 1 ) * sizeof( float );
    hipLaunchKernelGGL( apsm::lls_normal_partial, dim3( nPart ), dim3( apsm::kLlsThreads ), lds, h->stream, d_rx, d_tx, h->N,
                        T, chunk, h->part.ptr );
    HIP_OR_THROW( hipGetLastError() );
    llsReduceSolve( h, nPart, L );
}

// General column-major A (m x n, ld m), target y (m): trainLowLevel's operands.
void llsTrainColMajor( lls_handle* h, const float* d_A, const float* d_y, int n, int64_t m )
{
    ARG_CHECK( m > 0 && n > 0, "LLS train needs at least one row and one column" );
    ARG_CHECK( n <= kLlsMaxDims, "LLS supports at most 128 input dimensions" );
    const int E = n * ( n + 1 );
    const int64_t chunk = std::max<int64_t>( 1, (int64_t)( apsm::kLlsLdsBudget / ( ( n + 1 ) * sizeof( float ) ) ) );
    const int64_t nPart64 = ( m + chunk - 1 ) / chunk;
    ARG_CHECK( nPart64 < ( 1 << 30 ), "too many rows" );
    const int nPart = (int)nPart64;
    h->part.ensure( (size_t)nPart * E );
    const size_t lds = (size_t)chunk * ( n + 1 ) * sizeof( float );
    hipLaunchKernelGGL( apsm::lls_normal_colmajor, dim3( nPart ), dim3( apsm::kLlsThreads ), lds, h->stream, d_A, d_y, n, m, chunk,
                        h->part.ptr );
    HIP_OR_THROW( hipGetLastError() );
    llsReduceSolve( h, nPart, n );
}

// ---- least squares without allocation (lls_train_lowlevel_noalloc) ----
// Device scratch of the allocation-free least-squares entry point: the per-chunk FP32 partial
// normal equations, then their FP64 sum. One copy per device, held until the process ends;
// every call on a device uses the same copy (calls that overlap on one device race). Host
// code never passes the array's name to a kernel or copy: its device address comes from
// hipGetSymbolAddress, once per device (llsScratchOnCurrentDevice).
constexpr size_t kLlsScratchBytes = size_t( 32 ) << 20;
static __device__ __align__( 256 ) unsigned char scratch[ kLlsScratchBytes ];

constexpr int kMaxScratchDevices = 64;
std::mutex g_llsScratchMutex;
unsigned char* g_llsScratch[ kMaxScratchDevices ] = {};

unsigned char* llsScratchOnCurrentDevice()
{
    int dev = 0;
    HIP_OR_THROW( hipGetDevice( &dev ) );
    ARG_CHECK( dev >= 0 && dev < kMaxScratchDevices, "device index beyond the scratch table" );
    std::lock_guard<std::mutex> lock( g_llsScratchMutex );
    if ( !g_llsScratch[ dev ] )
    {
        void* ptr = nullptr;
        HIP_OR_THROW( hipGetSymbolAddress( &ptr, HIP_SYMBOL( scratch ) ) );
        ARG_CHECK( ptr != nullptr, "no device address for the least-squares scratch" );
        g_llsScratch[ dev ] = static_cast<unsigned char*>( ptr );
    }
    return g_llsScratch[ dev ];
}

// Scratch layout of one no-allocation solve of n unknowns over m rows (the chunking of
// llsTrainColMajor): nPart * E FP32 partials, then E FP64 sums at a 256-byte boundary.
struct LlsScratchPlan
{
    int E = 0, nPart = 0;
    int64_t chunk = 0;
    size_t gramOff = 0, bytes = 0;
};

LlsScratchPlan llsScratchPlan( int n, int64_t m )
{
    LlsScratchPlan sp;
    sp.E = n * ( n + 1 );
    sp.chunk = std::max<int64_t>( 1, (int64_t)( apsm::kLlsLdsBudget / ( ( n + 1 ) * sizeof( float ) ) ) );
    const int64_t nPart = ( m + sp.chunk - 1 ) / sp.chunk;
    const size_t partBytes = (size_t)nPart * sp.E * sizeof( float );
    sp.gramOff = ( partBytes + 255 ) & ~size_t( 255 );
    sp.bytes = sp.gramOff + (size_t)sp.E * sizeof( double );
    sp.nPart = nPart < ( 1 << 30 ) ? (int)nPart : ( 1 << 30 );
    return sp;
}

void llsCheckStatus( lls_handle* h )
{
    int st = 0;
    HIP_OR_THROW( hipMemcpyAsync( &st, h->status.ptr, sizeof( int ), hipMemcpyDeviceToHost, h->stream ) );
    HIP_OR_THROW( hipStreamSynchronize( h->stream ) );
    if ( st != 0 )
        throw ApsmError { APSM_ERR_NUMERIC, "least-squares normal matrix is not positive definite" };
}

} // namespace

// ---------------------------------------------------------------------------------------
// APSM handle
// ---------------------------------------------------------------------------------------
struct apsm_handle
{
    apsm_params p {};
    int N = 0, L = 0, U = 0, w2 = 0;
    int64_t H = 0, Hc = 0;
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr; // train sequencer
    hipEvent_t ev2 = nullptr, ev3 = nullptr; // detect
    bool trainTimed = false, detectTimed = false;
    float var = 0.f;
    bool varSet = false;
    bool anyTrain = false;
    uint64_t used = 0; // absolute real rows entered into the pilot stream
    StreamBuf buf[ 2 ];
    int cur = 0;
    int64_t bufBaseC = 0; // absolute complex index of entry 0 of buf[cur]
    int64_t bufK = 0; // complex entries staged in buf[cur]
    DevBuf<float> F, Gram, thOld, theta; // thOld: evicted-row linear share (per-call path) or thF (blocked path)
    bool blocked = false; // blocked train path (register-resident window Gram)
    int64_t a0Next = 0; // blocked path: frozen boundary after the last step (absolute real row)
    DevBuf<float> Pbuf, Gact, EvF, Ppart; // blocked path work buffers (Ppart: entry-share chunk partials)
    DevBuf<float> GramG, FG; // norm-constraint path: Gaussian window Gram and Gaussian part of f
    DevBuf<double> Qn; // norm-constraint path: c^T K c per user
    DevBuf<float> rx, tx, est; // staging for the host entry points
    DevBuf<int32_t> perm;
    lls_handle* lls = nullptr;
    // per-kernel profile (apsm_profile_*): event pairs bracketing launches of a kind
    bool prof = false;
    std::vector<hipEvent_t> profEv; // pool, reused after every read
    size_t profUsed = 0;
    std::vector<std::array<int, 3>> profMarks; // (kind, start event, end event)
};

namespace {

// Next event of the handle's profile pool, recorded on the handle's stream (index returned).
int profRecord( apsm_handle* h )
{
    if ( h->profUsed == h->profEv.size() )
    {
        hipEvent_t e = nullptr;
        HIP_OR_THROW( hipEventCreate( &e ) );
        h->profEv.push_back( e );
    }
    const int i = (int)h->profUsed++;
    HIP_OR_THROW( hipEventRecord( h->profEv[ i ], h->stream ) );
    return i;
}

void profMark( apsm_handle* h, int kind, int start, int end )
{
    if ( h->prof )
        h->profMarks.push_back( { kind, start, end } );
}

void resetDeviceState( apsm_handle* h )
{
    const size_t nF = (size_t)h->U * h->w2, nG = (size_t)h->U * h->w2 * h->w2, nT = (size_t)h->U * h->L;
    HIP_OR_THROW( hipMemsetAsync( h->F.ptr, 0, nF * sizeof( float ), h->stream ) );
    if ( h->Gram.ptr )
        HIP_OR_THROW( hipMemsetAsync( h->Gram.ptr, 0, nG * sizeof( float ), h->stream ) );
    HIP_OR_THROW( hipMemsetAsync( h->thOld.ptr, 0, nT * sizeof( float ), h->stream ) );
    HIP_OR_THROW( hipMemsetAsync( h->theta.ptr, 0, nT * sizeof( float ), h->stream ) );
    if ( h->GramG.ptr )
        HIP_OR_THROW( hipMemsetAsync( h->GramG.ptr, 0, nG * sizeof( float ), h->stream ) );
    if ( h->FG.ptr )
        HIP_OR_THROW( hipMemsetAsync( h->FG.ptr, 0, nF * sizeof( float ), h->stream ) );
    if ( h->Qn.ptr )
        HIP_OR_THROW( hipMemsetAsync( h->Qn.ptr, 0, h->U * sizeof( double ), h->stream ) );
    h->used = 0;
    h->a0Next = 0;
    h->bufBaseC = 0;
    h->bufK = 0;
    h->anyTrain = false;
}

void validateParams( const apsm_params* p, uint32_t N, uint32_t U )
{
    ARG_CHECK( p != nullptr, "null parameter set" );
    ARG_CHECK( N >= 1 && N <= 64, "numAntennas must be in [1, 64]" );
    ARG_CHECK( U >= 1 && U <= 1024, "numUsers must be in [1, 1024]" );
    ARG_CHECK( p->dictionary_size >= 1, "dictionarySize must be positive" );
    ARG_CHECK( p->window_size >= 1 && p->window_size <= p->dictionary_size, "windowSize must be in [1, dictionarySize]" );
    ARG_CHECK( 2 * p->window_size <= 1024, "windowSize must be at most 512" );
    ARG_CHECK( p->sample_step >= 2 && p->sample_step % 2 == 0, "sampleStep must be a positive even number" );
    ARG_CHECK( p->train_passes >= 1, "trainPasses must be at least 1" );
    ARG_CHECK( p->gaussian_kernel_weight >= 0.f && p->gaussian_kernel_weight <= 1.f, "gaussianKernelWeight must be in [0, 1]" );
    ARG_CHECK( p->linear_kernel_weight >= 0.f, "linearKernelWeight must be non-negative" );
    ARG_CHECK( p->linear_kernel_weight > 0.f || p->gaussian_kernel_weight > 0.f, "kernel weights must not both be zero" );
    ARG_CHECK( !p->lls_initialization || p->linear_kernel_weight > 0.f, "llsInitialization needs a positive linear kernel weight" );
    ARG_CHECK( p->gaussian_kernel_variance >= 0.f, "gaussianKernelVariance must be non-negative" );
    ARG_CHECK( p->eB >= 0.f, "eB must be non-negative" );
    ARG_CHECK( p->norm_constraint >= 0.f, "normConstraint must be non-negative" );
    if ( p->norm_constraint > 0.f && p->sample_step > 2 * apsm::kNewMax )
        throw ApsmError { APSM_ERR_UNSUPPORTED, "normConstraint > 0 supports sampleStep <= 8" };
    switch ( p->detect_version )
    {
    case APSM_VERSION_DETECT_ORIGINAL:
    case APSM_VERSION_DETECT_OLDFAST:
    case APSM_VERSION_DETECT_SHMEM:
    case APSM_VERSION_DETECT_BALANCED:
        break;
    default:
        throw ApsmError { APSM_ERR_ARG, "unknown detect version" };
    }
    ARG_CHECK( p->detect_precision >= APSM_PRECISION_FP32 && p->detect_precision <= APSM_PRECISION_FP16, "unknown detect precision" );
}

// Blocked-path eligibility: single pass, no full-window start, window Gram fits the register
// layouts (windowSize <= 160), at most kNewMax complex samples entering per step.
bool blockedEligible( const apsm_params& p )
{
    return p.train_passes == 1 && !p.start_with_full_window && p.window_size <= 160 && ( p.sample_step == 2 || p.sample_step == 4 )
        && p.norm_constraint == 0.f;
}

constexpr int64_t kBlockRowsMax = 1024; // real rows entering per block (the active-sample limit binds first)

template <int TP, int NEWPER, int PMAX>
void launchSeq2( apsm_handle* h, const float* Z, const float* Zn, const float* Y, float* C, const float* Gact, const float* EvF,
                 const apsm::Seq2Args& s2, size_t lds )
{
    if ( s2.nParts > PMAX || PMAX * s2.Wc > apsm::kRedF )
        throw ApsmError { APSM_ERR_UNSUPPORTED, "window layout does not fit the blocked train path" };
    if ( lds > 64 * 1024 )
        HIP_OR_THROW( hipFuncSetAttribute( (const void*)apsm::apsm_train_seq2<TP, NEWPER, PMAX>,
                                           hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds ) );
    hipLaunchKernelGGL( ( apsm::apsm_train_seq2<TP, NEWPER, PMAX> ), dim3( h->U ), dim3( apsm::kSeqThreads ), lds, h->stream, Z, Zn,
                        Y, C, h->F.ptr, h->thOld.ptr, h->Pbuf.ptr, Gact, EvF, s2 );
    HIP_OR_THROW( hipGetLastError() );
}

// Instantiations: (register triples per thread, upper bound of the t-slot part count).
template <int NEWPER>
void launchSeq2Tp( apsm_handle* h, int tp, const float* Z, const float* Zn, const float* Y, float* C, const float* Gact, const float* EvF,
                   const apsm::Seq2Args& s2, size_t lds )
{
    switch ( tp )
    {
    case 2: launchSeq2<2, NEWPER, 24>( h, Z, Zn, Y, C, Gact, EvF, s2, lds ); break;
    case 4: launchSeq2<4, NEWPER, 16>( h, Z, Zn, Y, C, Gact, EvF, s2, lds ); break;
    case 7: launchSeq2<7, NEWPER, 12>( h, Z, Zn, Y, C, Gact, EvF, s2, lds ); break;
    case 10: launchSeq2<10, NEWPER, 10>( h, Z, Zn, Y, C, Gact, EvF, s2, lds ); break;
    case 16: launchSeq2<16, NEWPER, 8>( h, Z, Zn, Y, C, Gact, EvF, s2, lds ); break;
    default: launchSeq2<27, NEWPER, 6>( h, Z, Zn, Y, C, Gact, EvF, s2, lds ); break;
    }
}

// Gram register triples per thread for a window of Wc complex slots (see apsm_train_seq2).
int seq2Tp( int Wc )
{
    const int parts = apsm::kSeqThreads / Wc;
    const int need = ( Wc + parts - 1 ) / parts;
    for ( int tp : { 2, 4, 7, 10, 16, 27 } )
        if ( tp >= need )
            return tp;
    return -1;
}

// Projection steps per block of the blocked path: entering rows <= H - W + q (no active row
// leaves the dictionary inside a block), active samples <= kActMax.
int64_t blockedSteps( const apsm_handle* h )
{
    const int64_t q = h->p.sample_step;
    const int w2 = h->w2;
    int64_t steps = std::max<int64_t>( 1, std::min<int64_t>( kBlockRowsMax / q, ( h->H - w2 ) / q + 1 ) );
    while ( steps > 1 && ( w2 + steps * q ) / 2 > apsm::kActMax )
        steps--;
    if ( ( w2 + steps * q ) / 2 > apsm::kActMax )
        throw ApsmError { APSM_ERR_UNSUPPORTED, "window too large for the blocked train path" };
    return steps;
}

// Work buffers of the blocked path for blocks of `steps` steps.
void reserveBlockedScratch( apsm_handle* h, int64_t steps )
{
    const int64_t q = h->p.sample_step;
    const int newPer = (int)( q / 2 );
    h->Pbuf.ensure( (size_t)h->U * steps * newPer * 2 );
    h->Ppart.ensure( (size_t)h->U * ( h->Hc / apsm::kEntryChunk + 2 ) * steps * newPer * 2 );
    h->Gact.ensure( (size_t)apsm::gramRow<int64_t>( ( h->w2 + steps * q ) / 2 ) );
    h->EvF.ensure( (size_t)h->U * steps * h->w2 );
}

// Runs the projection steps of one call in blocks (see apsm_train2.h).
void trainBlocked( apsm_handle* h, const StreamBuf& b, int64_t ld, int64_t usedOld, int64_t R2, int64_t rowBase, float wL, float wG,
                   float g2 )
{
    const int64_t q = h->p.sample_step;
    const int w2 = h->w2, Wc = w2 / 2, U = h->U;
    const int newPer = (int)( q / 2 );
    const int64_t nStepsTotal = R2 / q;
    const int64_t steps = blockedSteps( h );
    reserveBlockedScratch( h, steps );

    const int tp = seq2Tp( Wc );
    if ( tp < 0 )
        throw ApsmError { APSM_ERR_UNSUPPORTED, "window too large for the blocked train path" };
    const int nParts = std::min( apsm::kSeqThreads / Wc, ( Wc + tp - 1 ) / tp );
    if ( nParts > apsm::kMaxParts || nParts * tp < Wc )
        throw ApsmError { APSM_ERR_UNSUPPORTED, "window layout does not fit the blocked train path" };

    const int nmax = nmaxOf( h->N );
    for ( int64_t k0 = 0; k0 < nStepsTotal; k0 += steps )
    {
        const int64_t n = std::min( steps, nStepsTotal - k0 );
        const int64_t sPrev = usedOld + k0 * q, sFirst = sPrev + q, sLast = sPrev + n * q;
        const int64_t a0 = sFirst - std::min<int64_t>( w2, sFirst );
        const int64_t sNext = sLast + q;
        const int64_t a0Next = sNext - std::min<int64_t>( w2, sNext );
        const int nNewC = (int)( n * newPer );
        const int nActC = (int)( ( sLast - a0 ) / 2 );

        apsm::EntryArgs ea {};
        ea.N = h->N;
        ea.L = h->L;
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
        int64_t jBeg = 0;
        int nChunks = 0;
        apsm::entryChunks( ea, jBeg, nChunks );

        apsm::GramArgs ga {};
        ga.N = h->N;
        ga.L = h->L;
        ga.c0 = ( a0 - rowBase ) / 2;
        ga.nA = nActC;
        ga.wL = wL;
        ga.wG = wG;
        ga.g2 = g2;
        const int64_t loLast = sLast > h->H ? sLast - h->H : 0, loPrev = sPrev > h->H ? sPrev - h->H : 0;
        const int hasEv = loLast > loPrev ? 1 : 0;
        apsm::EvArgs ev {};
        ev.N = h->N;
        ev.L = h->L;
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

        // one launch for the entry share of the first 16 users, the Gram and the eviction
        // shares; further user groups get entry launches of their own; then the combines
        const int evPre = h->prof ? profRecord( h ) : -1;
        for ( int u0 = 0; u0 < U; u0 += 16 )
        {
            apsm::EntryArgs eu = ea;
            eu.U = std::min( 16, U - u0 );
            const float* Cg = b.C.ptr + (size_t)u0 * ld;
            float* partG = h->Ppart.ptr + (size_t)u0 * nChunks * nNewC * 2;
            if ( u0 == 0 )
            {
                apsm::PreDesc pd {};
                pd.e = apsm::EntryMfmaDesc { b.Z.ptr, b.Zn.ptr, Cg, h->thOld.ptr, h->Pbuf.ptr, partG, eu, jBeg, nChunks };
                pd.g = apsm::GramDesc { b.Z.ptr, b.Zn.ptr, h->Gact.ptr, ga };
                pd.v = apsm::EvDesc { b.Z.ptr, b.Zn.ptr, b.C.ptr, h->EvF.ptr, ev };
                pd.hasEv = hasEv;
                apsm::preShape( pd );
                HIP_OR_THROW( apsm::launchBlockPre( nmax, umaxOf( eu.U ), h->stream, pd ) );
            }
            else
                HIP_OR_THROW( apsm::launchEntryMfma( nmax, umaxOf( eu.U ), h->stream, b.Z.ptr, b.Zn.ptr, Cg, partG, eu, jBeg, nChunks, 0,
                                                     nChunks * apsm::entryTiles( nNewC ) ) );
        }
        for ( int u0 = 0; u0 < U; u0 += 16 )
        {
            apsm::EntryArgs eu = ea;
            eu.U = std::min( 16, U - u0 );
            const float* thg = h->thOld.ptr + (size_t)u0 * h->L;
            float* Pg = h->Pbuf.ptr + (size_t)u0 * nNewC * 2;
            float* partG = h->Ppart.ptr + (size_t)u0 * nChunks * nNewC * 2;
            HIP_OR_THROW( apsm::launchEntryCombine( h->stream, b.Z.ptr, thg, partG, Pg, eu, nChunks ) );
        }

        apsm::Seq2Args s2 {};
        s2.N = h->N;
        s2.L = h->L;
        s2.U = U;
        s2.w2 = w2;
        s2.Wc = Wc;
        s2.q = (int)q;
        s2.nSteps = (int)n;
        s2.hasEv = hasEv;
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
        s2.eB = h->p.eB;
        const size_t lds = apsm::seq2LdsBytes( newPer, w2, tp );
        const int evSeq = h->prof ? profRecord( h ) : -1;
        profMark( h, APSM_PROFILE_BLOCK_PRE, evPre, evSeq );
        const int64_t aR = a0 - rowBase; // local real row of the first active row
        const float* Za = b.Z.ptr + ( aR / 2 ) * h->L;
        const float* Zna = b.Zn.ptr + aR / 2;
        const float* Ya = b.Y.ptr + aR;
        float* Ca = b.C.ptr + aR;
        if ( newPer == 1 )
            launchSeq2Tp<1>( h, tp, Za, Zna, Ya, Ca, h->Gact.ptr, h->EvF.ptr, s2, lds );
        else
            launchSeq2Tp<2>( h, tp, Za, Zna, Ya, Ca, h->Gact.ptr, h->EvF.ptr, s2, lds );
        if ( h->prof )
            profMark( h, APSM_PROFILE_SEQUENCER, evSeq, profRecord( h ) );
        h->a0Next = a0Next;
    }
}

// Pairwise-distance heuristic for the Gaussian variance (cpp/lib/apsm/apsm_noma_detector.cu:194-218)
// on the device: the k-th smallest pair distance by a three-pass radix select, k as the
// reference indexes its sorted vector (size * 0.7f in float). The set of pair distances does
// not depend on the pilot order, so the training shuffle does not enter.
float varianceHeuristicDevice( hipStream_t stream, const float* d_rx, int N, int64_t T )
{
    ARG_CHECK( T >= 2, "variance heuristic needs at least two pilots" );
    const uint64_t P = (uint64_t)T * (uint64_t)( T - 1 ) / 2;
    uint64_t k = (uint64_t)( (float)P * 0.7f );
    if ( k >= P )
        k = P - 1;
    DevBuf<unsigned long long> hist;
    hist.ensure( apsm::kVhBins );
    std::vector<unsigned long long> hh( apsm::kVhBins );
    const int tiles = gridFor( T, apsm::kVhTile );
    ARG_CHECK( tiles <= 65535, "too many pilots for the variance heuristic" );
    const size_t lds = 2 * (size_t)N * apsm::kVhTile * sizeof( float2 );
    uint32_t prefix = 0;
    for ( int pass = 0; pass < 3; pass++ )
    {
        HIP_OR_THROW( hipMemsetAsync( hist.ptr, 0, apsm::kVhBins * sizeof( unsigned long long ), stream ) );
        hipLaunchKernelGGL( apsm::apsm_variance_histogram, dim3( tiles, tiles ), dim3( apsm::kVhThreads ), lds, stream, d_rx, N, T, pass,
                            prefix, hist.ptr );
        HIP_OR_THROW( hipGetLastError() );
        HIP_OR_THROW( hipMemcpyAsync( hh.data(), hist.ptr, apsm::kVhBins * sizeof( unsigned long long ), hipMemcpyDeviceToHost, stream ) );
        HIP_OR_THROW( hipStreamSynchronize( stream ) );
        const int nb = pass == 2 ? 256 : apsm::kVhBins;
        uint64_t cum = 0;
        int b = 0;
        for ( ; b < nb; b++ )
        {
            if ( cum + hh[ b ] > k )
                break;
            cum += hh[ b ];
        }
        if ( b == nb )
            throw ApsmError { APSM_ERR_NUMERIC, "variance heuristic: percentile not found" };
        k -= cum;
        prefix = pass == 0 ? (uint32_t)b : ( prefix << ( pass == 1 ? 12 : 8 ) ) | (uint32_t)b;
    }
    float value;
    std::memcpy( &value, &prefix, sizeof( value ) );
    // float element times a double power, rounded to float, then squared in float
    const float v = (float)( (double)value * std::pow( (double)T, -1. / 5. ) );
    return v * v;
}

// Kernel variance (configured value scaled by N^2, or the heuristic on the call's pilots) and
// the LLS initialisation of theta, both from the whole call (apsm_noma_detector.cu:96-105).
void trainPrepare( apsm_handle* h, const float* d_rx, const float* d_tx, int64_t T )
{
    const apsm_params& p = h->p;
    const int N = h->N, L = h->L, U = h->U;
    if ( !h->varSet )
    {
        if ( p.gaussian_kernel_variance != 0.f )
            h->var = p.gaussian_kernel_variance * (float)( N * N );
        else
            h->var = varianceHeuristicDevice( h->stream, d_rx, N, T );
        h->varSet = true;
    }
    const float wL = p.linear_kernel_weight;

    if ( p.lls_initialization && !h->anyTrain )
    {
        // theta_0 = w_LLS / wL per user (apsm_noma_detector.cu:321-346)
        if ( !h->lls )
        {
            lls_handle* l = nullptr;
            if ( lls_create( N, &l ) != APSM_OK )
                throw ApsmError { APSM_ERR_DEVICE, g_lastError };
            h->lls = l;
        }
        for ( int u = 0; u < U; u++ )
        {
            HIP_OR_THROW( hipStreamSynchronize( h->stream ) );
            llsTrainDev( h->lls, d_rx, d_tx + (size_t)u * T * 2, T );
            llsCheckStatus( h->lls );
            hipLaunchKernelGGL( apsm::apsm_lls_init_theta, dim3( 1 ), dim3( 128 ), 0, h->lls->stream, h->lls->w.ptr,
                                h->thOld.ptr + (size_t)u * L, L, wL );
            HIP_OR_THROW( hipGetLastError() );
            HIP_OR_THROW( hipStreamSynchronize( h->lls->stream ) );
        }
    }
    h->anyTrain = true;
}

// One train call. prepared: the caller already ran trainPrepare over the whole pilot set and
// the rows arrive in final order (no shuffle here).
void trainDev( apsm_handle* h, const float* d_rx, const float* d_tx, int64_t T, bool prepared = false )
{
    const apsm_params& p = h->p;
    ARG_CHECK( T > 0, "train needs at least one pilot" );
    ARG_CHECK( p.train_passes == 1 || h->used == 0, "trainPasses > 1 requires a fresh state (resetState)" );
    const int N = h->N, L = h->L, U = h->U;

    std::vector<int32_t> perm;
    const bool shuffle = p.training_shuffle && !prepared;
    if ( shuffle )
        perm = shufflePermutation( T );
    if ( !prepared )
        trainPrepare( h, d_rx, d_tx, T );
    const float wG = p.gaussian_kernel_weight, wL = p.linear_kernel_weight;

    // New stream buffer: the dictionary rows still live (at most Dc complex entries) + this call.
    const int64_t usedC = (int64_t)( h->used / 2 );
    const int64_t nOld = std::min<int64_t>( usedC, h->Hc );
    const int nxt = h->cur ^ 1;
    StreamBuf& nb = h->buf[ nxt ];
    const int64_t K = nOld + T;
    nb.ensure( K, N, U );
    const int64_t ld = 2 * nb.capK;
    if ( nOld > 0 )
    {
        StreamBuf& ob = h->buf[ h->cur ];
        const int64_t off = usedC - nOld - h->bufBaseC;
        const int64_t ldOld = 2 * ob.capK;
        HIP_OR_THROW( hipMemcpyAsync( nb.Z.ptr, ob.Z.ptr + off * L, (size_t)nOld * L * sizeof( float ), hipMemcpyDeviceToDevice,
                                      h->stream ) );
        HIP_OR_THROW( hipMemcpyAsync( nb.Zn.ptr, ob.Zn.ptr + off, (size_t)nOld * sizeof( float ), hipMemcpyDeviceToDevice, h->stream ) );
        HIP_OR_THROW( hipMemcpy2DAsync( nb.Y.ptr, ld * sizeof( float ), ob.Y.ptr + 2 * off, ldOld * sizeof( float ),
                                        2 * nOld * sizeof( float ), U, hipMemcpyDeviceToDevice, h->stream ) );
        HIP_OR_THROW( hipMemcpy2DAsync( nb.C.ptr, ld * sizeof( float ), ob.C.ptr + 2 * off, ldOld * sizeof( float ),
                                        2 * nOld * sizeof( float ), U, hipMemcpyDeviceToDevice, h->stream ) );
    }
    const int32_t* dPerm = nullptr;
    if ( shuffle )
    {
        h->perm.ensure( (size_t)T );
        HIP_OR_THROW( hipMemcpyAsync( h->perm.ptr, perm.data(), T * sizeof( int32_t ), hipMemcpyHostToDevice, h->stream ) );
        dPerm = h->perm.ptr;
    }
    hipLaunchKernelGGL( apsm::stage_pilots, dim3( gridFor( T, 256 ) ), dim3( 256 ), 0, h->stream, d_rx, d_tx, dPerm, N, U, T,
                        nb.Z.ptr + nOld * L, nb.Zn.ptr + nOld, nb.Y.ptr + 2 * nOld, nb.C.ptr + 2 * nOld, ld );
    HIP_OR_THROW( hipGetLastError() );

    apsm::SeqArgs a {};
    a.N = N;
    a.L = L;
    a.U = U;
    a.w2 = h->w2;
    a.q = (int)p.sample_step;
    a.passes = (int)p.train_passes;
    a.fullStart = p.start_with_full_window ? 1 : 0;
    a.H = h->H;
    a.usedOld = (int64_t)h->used;
    a.R2 = 2 * T;
    a.rowBase = 2 * ( usedC - nOld );
    a.ldYC = ld;
    a.wL = wL;
    a.wG = wG;
    a.g2 = 0.5f / h->var * kLog2e;
    a.eB = p.eB;

    const size_t lds = sizeof( float )
        * ( 2 * (size_t)h->w2 + L + apsm::kNewMax * L + apsm::kNewMax + apsm::kSeqWaves * 2 * apsm::kNewMax
            + apsm::kSeqWaves * apsm::kWave );
    HIP_OR_THROW( hipEventRecord( h->ev0, h->stream ) );
    if ( h->blocked )
        trainBlocked( h, nb, ld, (int64_t)h->used, 2 * T, a.rowBase, wL, wG, a.g2 );
    else if ( p.norm_constraint > 0.f )
    {
        apsm::NcArgs nca { h->GramG.ptr, h->FG.ptr, h->Qn.ptr, p.norm_constraint };
        hipLaunchKernelGGL( apsm::apsm_train_seq_nc, dim3( U ), dim3( apsm::kSeqThreads ), apsm::seqNcLdsBytes( h->w2, L ), h->stream,
                            nb.Z.ptr, nb.Zn.ptr, nb.Y.ptr, nb.C.ptr, h->F.ptr, h->Gram.ptr, h->thOld.ptr, h->theta.ptr, a, nca );
        HIP_OR_THROW( hipGetLastError() );
    }
    else
    {
        hipLaunchKernelGGL( apsm::apsm_train_seq, dim3( U ), dim3( apsm::kSeqThreads ), lds, h->stream, nb.Z.ptr, nb.Zn.ptr, nb.Y.ptr,
                            nb.C.ptr, h->F.ptr, h->Gram.ptr, h->thOld.ptr, a );
        HIP_OR_THROW( hipGetLastError() );
    }
    HIP_OR_THROW( hipEventRecord( h->ev1, h->stream ) );
    h->trainTimed = true;

    // stream bookkeeping, same schedule as the kernel
    const int64_t first = ( p.start_with_full_window && h->used == 0 ) ? h->w2 : p.sample_step;
    const int64_t R2 = 2 * T;
    if ( R2 >= first )
        h->used += (uint64_t)( first + ( ( R2 - first ) / p.sample_step ) * p.sample_step );
    h->cur = nxt;
    h->bufBaseC = usedC - nOld;
    h->bufK = K;

    if ( p.norm_constraint > 0.f )
        return; // the norm-constraint sequencer keeps theta explicitly and wrote it
    const int64_t newUsedC = (int64_t)( h->used / 2 );
    const int64_t lo = h->blocked ? std::min<int64_t>( h->a0Next / 2, newUsedC ) : std::max<int64_t>( 0, newUsedC - h->Hc );
    hipLaunchKernelGGL( apsm::apsm_theta_full, dim3( U ), dim3( 256 ), 0, h->stream, nb.Z.ptr, nb.C.ptr, h->thOld.ptr,
                        h->theta.ptr, N, lo - h->bufBaseC, newUsedC - h->bufBaseC, ld );
    HIP_OR_THROW( hipGetLastError() );
}

template <int NMAX, int UMAX>
void launchDetect( apsm_handle* h, const float* d_rx, int64_t M, float* d_est, const float* Z, const float* Zn, const float* C,
                   int64_t ldC, const float* theta, int64_t G, int U )
{
    const float wG = h->p.gaussian_kernel_weight, wL = h->p.linear_kernel_weight;
    const float g2 = 0.5f / h->var * kLog2e;
    hipLaunchKernelGGL( ( apsm::apsm_detect_kernel<NMAX, UMAX> ), dim3( gridFor( M, apsm::kDetThreads ) ), dim3( apsm::kDetThreads ),
                        0, h->stream, d_rx, M, Z, Zn, C, ldC, theta, G, U, wL, wG, g2, d_est, h->N );
    HIP_OR_THROW( hipGetLastError() );
}

template <int NMAX>
void launchDetectU( apsm_handle* h, const float* d_rx, int64_t M, float* d_est, const float* Z, const float* Zn, const float* C,
                    int64_t ldC, const float* theta, int64_t G, int U )
{
    if ( U <= 1 )
        launchDetect<NMAX, 1>( h, d_rx, M, d_est, Z, Zn, C, ldC, theta, G, U );
    else if ( U <= 2 )
        launchDetect<NMAX, 2>( h, d_rx, M, d_est, Z, Zn, C, ldC, theta, G, U );
    else if ( U <= 4 )
        launchDetect<NMAX, 4>( h, d_rx, M, d_est, Z, Zn, C, ldC, theta, G, U );
    else if ( U <= 8 )
        launchDetect<NMAX, 8>( h, d_rx, M, d_est, Z, Zn, C, ldC, theta, G, U );
    else if ( U <= 12 )
        launchDetect<NMAX, 12>( h, d_rx, M, d_est, Z, Zn, C, ldC, theta, G, U );
    else
        launchDetect<NMAX, 16>( h, d_rx, M, d_est, Z, Zn, C, ldC, theta, G, U );
}

// The detect kernel of a precision mode for N antennas: -1 for the vector-ALU kernel, else the
// matrix-core input precision (apsm::PREC_*).
int detectKernelOf( int32_t precision, uint32_t N )
{
    switch ( precision )
    {
    case APSM_PRECISION_FP32_VECTOR: return -1;
    case APSM_PRECISION_FP32_MATRIX: return apsm::PREC_F32;
    case APSM_PRECISION_BF16: return apsm::PREC_BF16;
    case APSM_PRECISION_FP16: return apsm::PREC_F16;
    default: return N >= kMatrixFp32MinAntennas ? apsm::PREC_F32 : -1;
    }
}


void detectDev( apsm_handle* h, const float* d_rx, int64_t M, float* d_est )
{
    ARG_CHECK( M > 0, "detect needs at least one sample" );
    if ( !h->varSet )
        h->var = h->p.gaussian_kernel_variance != 0.f ? h->p.gaussian_kernel_variance * (float)( h->N * h->N ) : 1.f;
    const int64_t usedC = (int64_t)( h->used / 2 );
    const int64_t lo = std::max<int64_t>( 0, usedC - h->Hc );
    const int64_t G = usedC - lo;
    StreamBuf& b = h->buf[ h->cur ];
    const int64_t off = lo - h->bufBaseC;
    const int64_t ld = 2 * b.capK;
    const float* Z = G > 0 ? b.Z.ptr + off * h->L : nullptr;
    const float* Zn = G > 0 ? b.Zn.ptr + off : nullptr;
    const float* C = G > 0 ? b.C.ptr + 2 * off : nullptr;
    HIP_OR_THROW( hipEventRecord( h->ev2, h->stream ) );
    const int evDet = h->prof ? profRecord( h ) : -1;
    for ( int u0 = 0; u0 < h->U; u0 += 16 )
    {
        const int Ug = std::min( 16, h->U - u0 );
        const float* Cg = C ? C + (int64_t)u0 * ld : nullptr;
        const float* th = h->theta.ptr + (size_t)u0 * h->L;
        float* est = d_est + (size_t)u0 * M * 2;
        const int kernel = detectKernelOf( h->p.detect_precision, h->N );
        if ( kernel >= 0 )
        {
            const float g2 = 0.5f / h->var * kLog2e;
            HIP_OR_THROW( apsm::launchDetectMfma( kernel, nmaxOf( (int)h->N ), umaxOf( Ug ), h->stream, d_rx, M, Z, Zn, Cg, ld,
                                                  th, G, Ug, h->p.linear_kernel_weight, h->p.gaussian_kernel_weight, g2, est,
                                                  (int)h->N ) );
        }
        else if ( h->N <= 4 )
            launchDetectU<4>( h, d_rx, M, est, Z, Zn, Cg, ld, th, G, Ug );
        else if ( h->N <= 8 )
            launchDetectU<8>( h, d_rx, M, est, Z, Zn, Cg, ld, th, G, Ug );
        else if ( h->N <= 16 )
            launchDetectU<16>( h, d_rx, M, est, Z, Zn, Cg, ld, th, G, Ug );
        else if ( h->N <= 32 )
            launchDetectU<32>( h, d_rx, M, est, Z, Zn, Cg, ld, th, G, Ug );
        else
            launchDetectU<64>( h, d_rx, M, est, Z, Zn, Cg, ld, th, G, Ug );
    }
    if ( h->prof )
        profMark( h, APSM_PROFILE_DETECT, evDet, profRecord( h ) );
    HIP_OR_THROW( hipEventRecord( h->ev3, h->stream ) );
    h->detectTimed = true;
}

float elapsedMs( hipEvent_t a, hipEvent_t b )
{
    float ms = 0.f;
    HIP_OR_THROW( hipEventSynchronize( b ) );
    HIP_OR_THROW( hipEventElapsedTime( &ms, a, b ) );
    return ms;
}

} // namespace

#include "apsm_batch_host.h"
#include "apsm_serve_host.h"

extern "C" {

void apsm_params_default( apsm_params* p )
{
    if ( !p )
        return;
    std::memset( p, 0, sizeof( *p ) );
    p->gaussian_kernel_weight = 0.5f;
    p->gaussian_kernel_variance = 0.05f;
    p->eB = 0.001f;
    p->norm_constraint = 0.f;
    p->window_size = 20;
    p->sample_step = 2;
    p->train_passes = 1;
    p->dictionary_size = 685;
    p->start_with_full_window = 0;
    p->lls_initialization = 0;
    p->training_shuffle = 0;
    p->detect_version = APSM_VERSION_DETECT_BALANCED;
    p->linear_kernel_weight = 0.5f;
    p->detect_precision = APSM_PRECISION_FP32;
}

const char* apsm_last_error( void ) { return g_lastError.c_str(); }

int apsm_device_count( void )
{
    int n = 0;
    if ( hipGetDeviceCount( &n ) != hipSuccess )
        return 0;
    return n;
}

int apsm_device_malloc( void** ptr, uint64_t bytes )
{
    return guarded( [ & ] {
        ARG_CHECK( ptr != nullptr, "null output pointer" );
        *ptr = nullptr;
        if ( bytes )
            HIP_OR_THROW( hipMalloc( ptr, bytes ) );
    } );
}

int apsm_device_free( void* ptr )
{
    return guarded( [ & ] {
        if ( ptr )
            HIP_OR_THROW( hipFree( ptr ) );
    } );
}

int apsm_memcpy_h2d( void* dst, const void* src, uint64_t bytes )
{
    return guarded( [ & ] {
        ARG_CHECK( ( dst && src ) || bytes == 0, "null argument" );
        if ( bytes )
        {
            HIP_OR_THROW( hipMemcpy( dst, src, bytes, hipMemcpyHostToDevice ) );
            HIP_OR_THROW( hipDeviceSynchronize() );
        }
    } );
}

int apsm_memcpy_d2h( void* dst, const void* src, uint64_t bytes )
{
    return guarded( [ & ] {
        ARG_CHECK( ( dst && src ) || bytes == 0, "null argument" );
        if ( bytes )
        {
            HIP_OR_THROW( hipDeviceSynchronize() ); // work on every stream, non-blocking ones included
            HIP_OR_THROW( hipMemcpy( dst, src, bytes, hipMemcpyDeviceToHost ) );
        }
    } );
}

int apsm_device_synchronize( void )
{
    return guarded( [ & ] { HIP_OR_THROW( hipDeviceSynchronize() ); } );
}

int apsm_create( const apsm_params* p, uint32_t num_antennas, uint32_t num_users, apsm_handle** out )
{
    return guarded( [ & ] {
        ARG_CHECK( out != nullptr, "null output handle" );
        *out = nullptr;
        validateParams( p, num_antennas, num_users );
        apsm_handle* h = new apsm_handle();
        try
        {
            h->p = *p;
            h->N = (int)num_antennas;
            h->L = 2 * h->N;
            h->U = (int)num_users;
            h->w2 = 2 * (int)p->window_size;
            h->Hc = p->dictionary_size;
            h->H = 2 * h->Hc;
            h->blocked = blockedEligible( *p );
            HIP_OR_THROW( hipStreamCreateWithFlags( &h->stream, hipStreamNonBlocking ) );
            HIP_OR_THROW( hipEventCreate( &h->ev0 ) );
            HIP_OR_THROW( hipEventCreate( &h->ev1 ) );
            HIP_OR_THROW( hipEventCreate( &h->ev2 ) );
            HIP_OR_THROW( hipEventCreate( &h->ev3 ) );
            h->F.ensure( (size_t)h->U * h->w2 );
            if ( !h->blocked )
                h->Gram.ensure( (size_t)h->U * h->w2 * h->w2 );
            if ( p->norm_constraint > 0.f )
            {
                h->GramG.ensure( (size_t)h->U * h->w2 * h->w2 );
                h->FG.ensure( (size_t)h->U * h->w2 );
                h->Qn.ensure( (size_t)h->U );
            }
            h->thOld.ensure( (size_t)h->U * h->L );
            h->theta.ensure( (size_t)h->U * h->L );
            resetDeviceState( h );
            HIP_OR_THROW( hipStreamSynchronize( h->stream ) );
        }
        catch ( ... )
        {
            apsm_destroy( h );
            throw;
        }
        *out = h;
    } );
}

void apsm_destroy( apsm_handle* h )
{
    if ( !h )
        return;
    if ( h->stream )
        (void)hipStreamSynchronize( h->stream );
    for ( auto& b : h->buf )
        b.release();
    h->F.release();
    h->Gram.release();
    h->GramG.release();
    h->FG.release();
    h->Qn.release();
    h->Pbuf.release();
    h->Ppart.release();
    h->Gact.release();
    h->EvF.release();

    h->thOld.release();
    h->theta.release();
    h->rx.release();
    h->tx.release();
    h->est.release();
    h->perm.release();
    if ( h->lls )
        lls_destroy( h->lls );
    for ( hipEvent_t e : h->profEv )
        (void)hipEventDestroy( e );
    if ( h->ev0 )
        (void)hipEventDestroy( h->ev0 );
    if ( h->ev1 )
        (void)hipEventDestroy( h->ev1 );
    if ( h->ev2 )
        (void)hipEventDestroy( h->ev2 );
    if ( h->ev3 )
        (void)hipEventDestroy( h->ev3 );
    if ( h->stream )
        (void)hipStreamDestroy( h->stream );
    delete h;
}

int apsm_kernel_times( apsm_handle* h, float* train_ms, float* detect_ms )
{
    return guarded( [ & ] {
        ARG_CHECK( h != nullptr, "null handle" );
        if ( train_ms )
            *train_ms = h->trainTimed ? elapsedMs( h->ev0, h->ev1 ) : 0.f;
        if ( detect_ms )
            *detect_ms = h->detectTimed ? elapsedMs( h->ev2, h->ev3 ) : 0.f;
    } );
}

int apsm_profile_enable( apsm_handle* h, int enable )
{
    return guarded( [ & ] {
        ARG_CHECK( h != nullptr, "null handle" );
        h->prof = enable != 0;
    } );
}

int apsm_profile_read( apsm_handle* h, double* ms, uint64_t* launches )
{
    return guarded( [ & ] {
        ARG_CHECK( h && ms && launches, "null argument" );
        for ( int k = 0; k < APSM_PROFILE_KINDS; k++ )
        {
            ms[ k ] = 0.0;
            launches[ k ] = 0;
        }
        if ( h->profUsed > 0 )
            HIP_OR_THROW( hipEventSynchronize( h->profEv[ h->profUsed - 1 ] ) );
        for ( const auto& m : h->profMarks )
        {
            float t = 0.f;
            HIP_OR_THROW( hipEventElapsedTime( &t, h->profEv[ m[ 1 ] ], h->profEv[ m[ 2 ] ] ) );
            ms[ m[ 0 ] ] += t;
            launches[ m[ 0 ] ]++;
        }
        h->profMarks.clear();
        h->profUsed = 0;
    } );
}

int apsm_reset( apsm_handle* h )
{
    return guarded( [ & ] {
        ARG_CHECK( h != nullptr, "null handle" );
        resetDeviceState( h );
        HIP_OR_THROW( hipStreamSynchronize( h->stream ) );
    } );
}

int apsm_train( apsm_handle* h, const float* rx, const float* tx, uint32_t T, float* ms )
{
    return guarded( [ & ] {
        ARG_CHECK( h && rx && tx, "null argument" );
        const size_t nrx = (size_t)h->N * T * 2, ntx = (size_t)h->U * T * 2;
        h->rx.ensure( nrx );
        h->tx.ensure( ntx );
        HIP_OR_THROW( hipMemcpyAsync( h->rx.ptr, rx, nrx * sizeof( float ), hipMemcpyHostToDevice, h->stream ) );
        HIP_OR_THROW( hipMemcpyAsync( h->tx.ptr, tx, ntx * sizeof( float ), hipMemcpyHostToDevice, h->stream ) );
        trainDev( h, h->rx.ptr, h->tx.ptr, T );
        float t = elapsedMs( h->ev0, h->ev1 );
        HIP_OR_THROW( hipStreamSynchronize( h->stream ) );
        if ( ms )
            *ms = t;
    } );
}

int apsm_train_dev( apsm_handle* h, const float* d_rx, const float* d_tx, uint32_t T )
{
    return guarded( [ & ] {
        ARG_CHECK( h && d_rx && d_tx, "null argument" );
        trainDev( h, d_rx, d_tx, T );
    } );
}

int apsm_detect( apsm_handle* h, const float* rx, uint32_t M, float* est, float* ms )
{
    return guarded( [ & ] {
        ARG_CHECK( h && rx && est, "null argument" );
        const size_t nrx = (size_t)h->N * M * 2, nest = (size_t)h->U * M * 2;
        h->rx.ensure( nrx );
        h->est.ensure( nest );
        HIP_OR_THROW( hipMemcpyAsync( h->rx.ptr, rx, nrx * sizeof( float ), hipMemcpyHostToDevice, h->stream ) );
        detectDev( h, h->rx.ptr, M, h->est.ptr );
        HIP_OR_THROW( hipMemcpyAsync( est, h->est.ptr, nest * sizeof( float ), hipMemcpyDeviceToHost, h->stream ) );
        float t = elapsedMs( h->ev2, h->ev3 );
        HIP_OR_THROW( hipStreamSynchronize( h->stream ) );
        if ( ms )
            *ms = t;
    } );
}

int apsm_detect_dev( apsm_handle* h, const float* d_rx, uint32_t M, float* d_est )
{
    return guarded( [ & ] {
        ARG_CHECK( h && d_rx && d_est, "null argument" );
        detectDev( h, d_rx, M, d_est );
    } );
}

int apsm_synchronize( apsm_handle* h )
{
    return guarded( [ & ] {
        ARG_CHECK( h != nullptr, "null handle" );
        HIP_OR_THROW( hipStreamSynchronize( h->stream ) );
    } );
}

void* apsm_stream( apsm_handle* h ) { return h ? (void*)h->stream : nullptr; }

int apsm_get_state( apsm_handle* h, uint64_t* used_rows, uint32_t* dict_entries, float* variance, float* theta, float* coeffs,
                    uint32_t coeffs_cap )
{
    return guarded( [ & ] {
        ARG_CHECK( h != nullptr, "null handle" );
        HIP_OR_THROW( hipStreamSynchronize( h->stream ) );
        const int64_t usedC = (int64_t)( h->used / 2 );
        const int64_t lo = std::max<int64_t>( 0, usedC - h->Hc );
        const int64_t G = usedC - lo;
        if ( used_rows )
            *used_rows = h->used;
        if ( dict_entries )
            *dict_entries = (uint32_t)G;
        if ( variance )
            *variance = h->var;
        if ( theta )
            HIP_OR_THROW( hipMemcpy( theta, h->theta.ptr, (size_t)h->U * h->L * sizeof( float ), hipMemcpyDeviceToHost ) );
        if ( coeffs && G > 0 )
        {
            ARG_CHECK( coeffs_cap >= 2 * G, "coefficient buffer too small" );
            StreamBuf& b = h->buf[ h->cur ];
            const int64_t off = lo - h->bufBaseC;
            HIP_OR_THROW( hipMemcpy2D( coeffs, 2 * G * sizeof( float ), b.C.ptr + 2 * off, 2 * b.capK * sizeof( float ),
                                       2 * G * sizeof( float ), h->U, hipMemcpyDeviceToHost ) );
        }
    } );
}

int apsm_get_dictionary( apsm_handle* h, float* dict, uint32_t cap_entries, uint32_t* dict_entries )
{
    return guarded( [ & ] {
        ARG_CHECK( h != nullptr, "null handle" );
        HIP_OR_THROW( hipStreamSynchronize( h->stream ) );
        const int64_t usedC = (int64_t)( h->used / 2 );
        const int64_t lo = std::max<int64_t>( 0, usedC - h->Hc );
        const int64_t G = usedC - lo;
        if ( dict_entries )
            *dict_entries = (uint32_t)G;
        if ( !dict || G == 0 )
            return;
        ARG_CHECK( cap_entries >= G, "dictionary buffer too small" );
        const StreamBuf& b = h->buf[ h->cur ];
        std::vector<float> z( (size_t)G * h->L );
        HIP_OR_THROW( hipMemcpy( z.data(), b.Z.ptr + ( lo - h->bufBaseC ) * h->L, z.size() * sizeof( float ), hipMemcpyDeviceToHost ) );
        for ( int64_t k = 0; k < G; k++ )
            for ( int n = 0; n < h->N; n++ )
            {
                dict[ ( (size_t)n * G + k ) * 2 ] = z[ k * h->L + n ];
                dict[ ( (size_t)n * G + k ) * 2 + 1 ] = z[ k * h->L + h->N + n ];
            }
    } );
}

int apsm_set_detect_precision( apsm_handle* h, int32_t precision )
{
    return guarded( [ & ] {
        ARG_CHECK( h != nullptr, "null handle" );
        ARG_CHECK( precision >= APSM_PRECISION_FP32 && precision <= APSM_PRECISION_FP16, "unknown detect precision" );
        h->p.detect_precision = precision;
    } );
}

int apsm_norm_ratio( apsm_handle* h, uint32_t user, float* ratio )
{
    return guarded( [ & ] {
        ARG_CHECK( h && ratio, "null argument" );
        ARG_CHECK( user < (uint32_t)h->U, "user index out of range" );
        const int64_t usedC = (int64_t)( h->used / 2 );
        const int64_t lo = std::max<int64_t>( 0, usedC - h->Hc );
        const int64_t G = usedC - lo;
        const float wG = h->p.gaussian_kernel_weight, wL = h->p.linear_kernel_weight;
        double gauss = 0.0;
        if ( G > 0 )
        {
            const StreamBuf& b = h->buf[ h->cur ];
            const int64_t off = lo - h->bufBaseC;
            const int nb = gridFor( G, apsm::kQuadThreads );
            DevBuf<double> part;
            part.ensure( (size_t)nb );
            const size_t lds = sizeof( float ) * ( (size_t)apsm::kQuadTile * h->L + 3 * apsm::kQuadTile );
            hipLaunchKernelGGL( apsm::apsm_quad_form_partial, dim3( nb ), dim3( apsm::kQuadThreads ), lds, h->stream,
                                b.Z.ptr + off * h->L, b.Zn.ptr + off, b.C.ptr + (size_t)user * 2 * b.capK + 2 * off, G, h->N,
                                0.5f / h->var * kLog2e, part.ptr );
            HIP_OR_THROW( hipGetLastError() );
            std::vector<double> hp( nb );
            HIP_OR_THROW( hipMemcpyAsync( hp.data(), part.ptr, nb * sizeof( double ), hipMemcpyDeviceToHost, h->stream ) );
            HIP_OR_THROW( hipStreamSynchronize( h->stream ) );
            part.release();
            for ( double v : hp )
                gauss += v;
        }
        std::vector<float> th( h->L );
        HIP_OR_THROW( hipStreamSynchronize( h->stream ) );
        HIP_OR_THROW( hipMemcpy( th.data(), h->theta.ptr + (size_t)user * h->L, h->L * sizeof( float ), hipMemcpyDeviceToHost ) );
        double lin = 0.0;
        for ( float v : th )
            lin += (double)v * v;
        // linear norm / Gaussian norm (apsm_noma_detector.cu:300-318)
        *ratio = (float)( ( wL * lin ) / ( wG * gauss ) );
    } );
}

int apsm_train_history( apsm_handle* h, const float* rx, const float* tx, uint32_t T, const float* rx_val, const float* tx_val,
                        uint32_t T_val, uint32_t bits_per_symbol, float modulation_scale, float* history, uint32_t history_cap,
                        uint32_t* num_steps, float* ms )
{
    return guarded( [ & ] {
        ARG_CHECK( h && rx && tx && rx_val && tx_val && history && num_steps, "null argument" );
        ARG_CHECK( T > 0 && T_val > 0, "training and validation sets must not be empty" );
        ARG_CHECK( bits_per_symbol >= 1 && bits_per_symbol <= 16 && ( bits_per_symbol == 1 || bits_per_symbol % 2 == 0 ),
                   "bitsPerSymbol must be 1 or even" );
        ARG_CHECK( modulation_scale > 0.f, "modulation scale must be positive" );
        const apsm_params& p = h->p;
        if ( p.train_passes != 1 )
            throw ApsmError { APSM_ERR_UNSUPPORTED, "a training history needs trainPasses == 1" };
        const int N = h->N, U = h->U;
        const int64_t q2 = p.sample_step / 2; // complex samples per step
        const int64_t firstC = ( p.start_with_full_window && h->used == 0 ) ? h->w2 / 2 : q2;
        const int64_t steps = T >= firstC ? 1 + ( T - firstC ) / q2 : 0;
        ARG_CHECK( (uint64_t)steps <= history_cap, "history buffer too small" );

        // pilots in training order (the reference shuffles before anything else, apsm_noma_detector.cu:102-106)
        std::vector<float> hrx( rx, rx + (size_t)N * T * 2 ), htx( tx, tx + (size_t)U * T * 2 );
        if ( p.training_shuffle )
        {
            const std::vector<int32_t> perm = shufflePermutation( T );
            for ( int n = 0; n < N; n++ )
                for ( int64_t k = 0; k < T; k++ )
                {
                    hrx[ ( n * T + k ) * 2 ] = rx[ ( n * T + perm[ k ] ) * 2 ];
                    hrx[ ( n * T + k ) * 2 + 1 ] = rx[ ( n * T + perm[ k ] ) * 2 + 1 ];
                }
            for ( int u = 0; u < U; u++ )
                for ( int64_t k = 0; k < T; k++ )
                {
                    htx[ ( u * T + k ) * 2 ] = tx[ ( u * T + perm[ k ] ) * 2 ];
                    htx[ ( u * T + k ) * 2 + 1 ] = tx[ ( u * T + perm[ k ] ) * 2 + 1 ];
                }
        }
        DevBuf<float> dRx, dTx, dRv, dTv, dEst, dChunkRx, dChunkTx, dHist;
        const int64_t Tmax = std::max<int64_t>( T, T_val );
        dRx.ensure( hrx.size() );
        dTx.ensure( htx.size() );
        dRv.ensure( (size_t)N * T_val * 2 );
        dTv.ensure( (size_t)U * T_val * 2 );
        dEst.ensure( (size_t)U * Tmax * 2 );
        dChunkRx.ensure( (size_t)N * std::max( firstC, q2 ) * 2 );
        dChunkTx.ensure( (size_t)U * std::max( firstC, q2 ) * 2 );
        dHist.ensure( (size_t)std::max<int64_t>( steps, 1 ) * U * 6 );
        HIP_OR_THROW( hipMemcpyAsync( dRx.ptr, hrx.data(), hrx.size() * sizeof( float ), hipMemcpyHostToDevice, h->stream ) );
        HIP_OR_THROW( hipMemcpyAsync( dTx.ptr, htx.data(), htx.size() * sizeof( float ), hipMemcpyHostToDevice, h->stream ) );
        HIP_OR_THROW( hipMemcpyAsync( dRv.ptr, rx_val, (size_t)N * T_val * 2 * sizeof( float ), hipMemcpyHostToDevice, h->stream ) );
        HIP_OR_THROW( hipMemcpyAsync( dTv.ptr, tx_val, (size_t)U * T_val * 2 * sizeof( float ), hipMemcpyHostToDevice, h->stream ) );
        trainPrepare( h, dRx.ptr, dTx.ptr, T );

        hipEvent_t e0 = nullptr, e1 = nullptr;
        HIP_OR_THROW( hipEventCreate( &e0 ) );
        HIP_OR_THROW( hipEventCreate( &e1 ) );
        HIP_OR_THROW( hipEventRecord( e0, h->stream ) );
        int64_t c0 = 0;
        for ( int64_t st = 0; st < steps; st++ )
        {
            const int64_t nC = st == 0 ? firstC : q2;
            // one step's pilots as a contiguous [N][nC][2] / [U][nC][2] call
            HIP_OR_THROW( hipMemcpy2DAsync( dChunkRx.ptr, nC * 2 * sizeof( float ), dRx.ptr + c0 * 2, T * 2 * sizeof( float ),
                                            nC * 2 * sizeof( float ), N, hipMemcpyDeviceToDevice, h->stream ) );
            HIP_OR_THROW( hipMemcpy2DAsync( dChunkTx.ptr, nC * 2 * sizeof( float ), dTx.ptr + c0 * 2, T * 2 * sizeof( float ),
                                            nC * 2 * sizeof( float ), U, hipMemcpyDeviceToDevice, h->stream ) );
            trainDev( h, dChunkRx.ptr, dChunkTx.ptr, nC, true );
            c0 += nC;
            // detect the whole training set and the validation set with the state after this step
            // (apsm_train.cu:147-158)
            detectDev( h, dRx.ptr, T, dEst.ptr );
            hipLaunchKernelGGL( apsm::apsm_history_metrics, dim3( U ), dim3( apsm::kMetricThreads ), 0, h->stream, dTx.ptr,
                                (int64_t)T, dEst.ptr, (int64_t)T, (int64_t)T, (int)bits_per_symbol, modulation_scale,
                                dHist.ptr + (size_t)st * U * 6, 6 );
            HIP_OR_THROW( hipGetLastError() );
            detectDev( h, dRv.ptr, T_val, dEst.ptr );
            hipLaunchKernelGGL( apsm::apsm_history_metrics, dim3( U ), dim3( apsm::kMetricThreads ), 0, h->stream, dTv.ptr,
                                (int64_t)T_val, dEst.ptr, (int64_t)T_val, (int64_t)T_val, (int)bits_per_symbol,
                                modulation_scale, dHist.ptr + (size_t)st * U * 6 + 3, 6 );
            HIP_OR_THROW( hipGetLastError() );
        }
        HIP_OR_THROW( hipEventRecord( e1, h->stream ) );
        const float t = elapsedMs( e0, e1 );
        (void)hipEventDestroy( e0 );
        (void)hipEventDestroy( e1 );
        if ( steps > 0 )
            HIP_OR_THROW( hipMemcpyAsync( history, dHist.ptr, (size_t)steps * U * 6 * sizeof( float ), hipMemcpyDeviceToHost,
                                          h->stream ) );
        HIP_OR_THROW( hipStreamSynchronize( h->stream ) );
        *num_steps = (uint32_t)steps;
        if ( ms )
            *ms = t;
        dRx.release();
        dTx.release();
        dRv.release();
        dTv.release();
        dEst.release();
        dChunkRx.release();
        dChunkTx.release();
        dHist.release();
    } );
}

// ---- LLS ----
static int llsCreate( uint32_t num_antennas, void* stream, lls_handle** out )
{
    return guarded( [ & ] {
        ARG_CHECK( out != nullptr, "null output handle" );
        ARG_CHECK( num_antennas >= 1 && num_antennas <= 64, "numAntennas must be in [1, 64]" );
        *out = nullptr;
        lls_handle* h = new lls_handle();
        try
        {
            h->N = (int)num_antennas;
            h->L = 2 * h->N;
            if ( stream )
            {
                h->stream = (hipStream_t)stream;
                h->ownStream = false;
            }
            else
                HIP_OR_THROW( hipStreamCreateWithFlags( &h->stream, hipStreamNonBlocking ) );
            HIP_OR_THROW( hipEventCreate( &h->ev0 ) );
            HIP_OR_THROW( hipEventCreate( &h->ev1 ) );
            h->w.ensure( (size_t)kLlsMaxDims );
            h->status.ensure( 1 );
            HIP_OR_THROW( hipMemsetAsync( h->w.ptr, 0, kLlsMaxDims * sizeof( float ), h->stream ) );
            HIP_OR_THROW( hipStreamSynchronize( h->stream ) );
        }
        catch ( ... )
        {
            lls_destroy( h );
            throw;
        }
        *out = h;
    } );
}

int lls_create( uint32_t num_antennas, lls_handle** out ) { return llsCreate( num_antennas, nullptr, out ); }

int lls_create_on_stream( uint32_t num_antennas, void* stream, lls_handle** out )
{
    if ( !stream )
        return guarded( [] { throw ApsmError { APSM_ERR_ARG, "null stream" }; } );
    return llsCreate( num_antennas, stream, out );
}

void lls_destroy( lls_handle* h )
{
    if ( !h )
        return;
    if ( h->stream )
        (void)hipStreamSynchronize( h->stream );
    h->w.release();
    h->part.release();
    h->gram.release();
    h->rx.release();
    h->tx.release();
    h->est.release();
    h->status.release();
    h->sums.release();
    if ( h->ev0 )
        (void)hipEventDestroy( h->ev0 );
    if ( h->ev1 )
        (void)hipEventDestroy( h->ev1 );
    if ( h->stream && h->ownStream )
        (void)hipStreamDestroy( h->stream );
    delete h;
}

int lls_train_dev( lls_handle* h, const float* d_rx, const float* d_tx, uint32_t T )
{
    return guarded( [ & ] {
        ARG_CHECK( h && d_rx && d_tx, "null argument" );
        llsTrainDev( h, d_rx, d_tx, T );
        llsCheckStatus( h );
        h->trained = true;
    } );
}

int lls_train( lls_handle* h, const float* rx, const float* tx, uint32_t T, float* ms )
{
    return guarded( [ & ] {
        ARG_CHECK( h && rx && tx, "null argument" );
        const size_t nrx = (size_t)h->N * T * 2, ntx = (size_t)T * 2;
        h->rx.ensure( nrx );
        h->tx.ensure( ntx );
        HIP_OR_THROW( hipMemcpyAsync( h->rx.ptr, rx, nrx * sizeof( float ), hipMemcpyHostToDevice, h->stream ) );
        HIP_OR_THROW( hipMemcpyAsync( h->tx.ptr, tx, ntx * sizeof( float ), hipMemcpyHostToDevice, h->stream ) );
        HIP_OR_THROW( hipEventRecord( h->ev0, h->stream ) );
        llsTrainDev( h, h->rx.ptr, h->tx.ptr, T );
        HIP_OR_THROW( hipEventRecord( h->ev1, h->stream ) );
        float t = elapsedMs( h->ev0, h->ev1 );
        llsCheckStatus( h );
        h->trained = true;
        if ( ms )
            *ms = t;
    } );
}

int lls_detect_dev( lls_handle* h, const float* d_rx, uint32_t M, float* d_est )
{
    return guarded( [ & ] {
        ARG_CHECK( h && d_rx && d_est, "null argument" );
        ARG_CHECK( M > 0, "detect needs at least one sample" );
        ARG_CHECK( h->nW == 0 || h->nW == h->L, "weights come from a low-level solve of another size" );
        hipLaunchKernelGGL( apsm::lls_apply, dim3( gridFor( M, 256 ) ), dim3( 256 ), 0, h->stream, d_rx, h->N, (int64_t)M, h->w.ptr,
                            d_est );
        HIP_OR_THROW( hipGetLastError() );
    } );
}

int lls_detect( lls_handle* h, const float* rx, uint32_t M, float* est, float* ms )
{
    return guarded( [ & ] {
        ARG_CHECK( h && rx && est, "null argument" );
        ARG_CHECK( M > 0, "detect needs at least one sample" );
        ARG_CHECK( h->nW == 0 || h->nW == h->L, "weights come from a low-level solve of another size" );
        const size_t nrx = (size_t)h->N * M * 2, nest = (size_t)M * 2;
        h->rx.ensure( nrx );
        h->est.ensure( nest );
        HIP_OR_THROW( hipMemcpyAsync( h->rx.ptr, rx, nrx * sizeof( float ), hipMemcpyHostToDevice, h->stream ) );
        HIP_OR_THROW( hipEventRecord( h->ev0, h->stream ) );
        hipLaunchKernelGGL( apsm::lls_apply, dim3( gridFor( M, 256 ) ), dim3( 256 ), 0, h->stream, h->rx.ptr, h->N, (int64_t)M,
                            h->w.ptr, h->est.ptr );
        HIP_OR_THROW( hipGetLastError() );
        HIP_OR_THROW( hipEventRecord( h->ev1, h->stream ) );
        HIP_OR_THROW( hipMemcpyAsync( est, h->est.ptr, nest * sizeof( float ), hipMemcpyDeviceToHost, h->stream ) );
        float t = elapsedMs( h->ev0, h->ev1 );
        HIP_OR_THROW( hipStreamSynchronize( h->stream ) );
        if ( ms )
            *ms = t;
    } );
}

int lls_train_lowlevel( lls_handle* h, const float* d_A, const float* d_y, int32_t num_input_dims, int32_t num_output_dims,
                        int32_t num_samples )
{
    return guarded( [ & ] {
        ARG_CHECK( h && d_A && d_y, "null argument" );
        ARG_CHECK( num_output_dims == 1, "LLS solves for one output dimension" );
        llsTrainColMajor( h, d_A, d_y, num_input_dims, num_samples );
        llsCheckStatus( h );
        h->trained = true;
    } );
}

int lls_train_lowlevel_noalloc( const float* d_A, const float* d_y, int32_t num_input_dims, int32_t num_output_dims,
                                int32_t num_samples, float* d_w, int32_t* d_status, void* stream )
{
    return guarded( [ & ] {
        ARG_CHECK( d_A && d_y && d_w && d_status, "null argument" );
        ARG_CHECK( num_output_dims == 1, "LLS solves for one output dimension" );
        ARG_CHECK( num_samples > 0 && num_input_dims > 0, "LLS train needs at least one row and one column" );
        ARG_CHECK( num_input_dims <= kLlsMaxDims, "LLS supports at most 128 input dimensions" );
        const int n = num_input_dims;
        const int64_t m = num_samples;
        const LlsScratchPlan sp = llsScratchPlan( n, m );
        if ( sp.bytes > kLlsScratchBytes )
            throw ApsmError { APSM_ERR_ARG, "least-squares problem needs " + std::to_string( sp.bytes ) +
                                                " bytes of scratch; the allocation-free entry point holds " +
                                                std::to_string( kLlsScratchBytes ) };
        unsigned char* base = llsScratchOnCurrentDevice();
        float* part = reinterpret_cast<float*>( base );
        double* gram = reinterpret_cast<double*>( base + sp.gramOff );
        hipStream_t st = static_cast<hipStream_t>( stream );
        // every scratch entry a kernel reads was written before it in this call (partials by
        // the normal-equation kernel, sums by the reduction): nothing needs clearing
        hipLaunchKernelGGL( apsm::lls_normal_colmajor, dim3( sp.nPart ), dim3( apsm::kLlsThreads ),
                            (size_t)sp.chunk * ( n + 1 ) * sizeof( float ), st, d_A, d_y, n, m, sp.chunk, part );
        HIP_OR_THROW( hipGetLastError() );
        hipLaunchKernelGGL( apsm::lls_reduce_partials, dim3( gridFor( sp.E, 256 ) ), dim3( 256 ), 0, st, part, sp.nPart, sp.E, gram );
        HIP_OR_THROW( hipGetLastError() );
        const size_t ldsSolve = (size_t)sp.E * sizeof( double );
        if ( ldsSolve > 64 * 1024 )
            HIP_OR_THROW(
                hipFuncSetAttribute( (const void*)apsm::lls_solve, hipFuncAttributeMaxDynamicSharedMemorySize, (int)ldsSolve ) );
        hipLaunchKernelGGL( apsm::lls_solve, dim3( 1 ), dim3( apsm::kSolveThreads ), ldsSolve, st, gram, n, d_w, d_status );
        HIP_OR_THROW( hipGetLastError() );
    } );
}

int lls_detect_lowlevel( lls_handle* h, const float* d_A, float* d_est, int32_t num_input_dims, int32_t num_output_dims,
                         int32_t num_samples, float alpha, float beta )
{
    return guarded( [ & ] {
        ARG_CHECK( h && d_A && d_est, "null argument" );
        ARG_CHECK( num_output_dims == 1, "LLS solves for one output dimension" );
        ARG_CHECK( num_input_dims > 0 && num_input_dims <= kLlsMaxDims && num_samples > 0, "invalid matrix shape" );
        hipLaunchKernelGGL( apsm::lls_gemv_colmajor, dim3( gridFor( num_samples, 256 ) ), dim3( 256 ), 0, h->stream, d_A,
                            num_input_dims, (int64_t)num_samples, h->w.ptr, alpha, beta, d_est );
        HIP_OR_THROW( hipGetLastError() );
    } );
}

int lls_signal_power( lls_handle* h, const float* d_data, uint32_t num_values, float* out )
{
    return guarded( [ & ] {
        ARG_CHECK( h && d_data && out, "null argument" );
        ARG_CHECK( num_values > 0, "no values" );
        const int nb = std::min( 1024, gridFor( num_values, 256 ) );
        h->sums.ensure( (size_t)nb );
        hipLaunchKernelGGL( apsm::sum_squares_partial, dim3( nb ), dim3( 256 ), 0, h->stream, d_data, (int64_t)num_values,
                            h->sums.ptr );
        HIP_OR_THROW( hipGetLastError() );
        std::vector<double> part( nb );
        HIP_OR_THROW( hipMemcpyAsync( part.data(), h->sums.ptr, nb * sizeof( double ), hipMemcpyDeviceToHost, h->stream ) );
        HIP_OR_THROW( hipStreamSynchronize( h->stream ) );
        double s = 0.0;
        for ( double v : part )
            s += v;
        *out = (float)std::sqrt( s / num_values );
    } );
}

float* lls_weights_dev( lls_handle* h ) { return h ? h->w.ptr : nullptr; }

void* lls_stream( lls_handle* h ) { return h ? (void*)h->stream : nullptr; }

int apsm_stream_create( void** out )
{
    return guarded( [ & ] {
        ARG_CHECK( out != nullptr, "null output" );
        hipStream_t s = nullptr;
        HIP_OR_THROW( hipStreamCreateWithFlags( &s, hipStreamNonBlocking ) );
        *out = (void*)s;
    } );
}

int apsm_stream_destroy( void* stream )
{
    return guarded( [ & ] {
        if ( stream )
            HIP_OR_THROW( hipStreamDestroy( (hipStream_t)stream ) );
    } );
}

int apsm_stream_synchronize( void* stream )
{
    return guarded( [ & ] {
        ARG_CHECK( stream != nullptr, "null stream" );
        HIP_OR_THROW( hipStreamSynchronize( (hipStream_t)stream ) );
    } );
}

int lls_get_weights( lls_handle* h, float* w )
{
    return guarded( [ & ] {
        ARG_CHECK( h && w, "null argument" );
        HIP_OR_THROW( hipStreamSynchronize( h->stream ) );
        HIP_OR_THROW( hipMemcpy( w, h->w.ptr, ( h->nW ? h->nW : h->L ) * sizeof( float ), hipMemcpyDeviceToHost ) );
    } );
}

int lls_synchronize( lls_handle* h )
{
    return guarded( [ & ] {
        ARG_CHECK( h != nullptr, "null handle" );
        HIP_OR_THROW( hipStreamSynchronize( h->stream ) );
    } );
}

// ---- batched instances ----
int apsm_batch_create( uint32_t n, const apsm_params* params, const uint32_t* num_antennas, const uint32_t* num_users, apsm_batch** out )
{
    return guarded( [ & ] {
        ARG_CHECK( out && params && num_antennas && num_users, "null argument" );
        ARG_CHECK( n >= 1 && n <= 65535, "batch size must be in [1, 65535]" );
        *out = nullptr;
        apsm_batch* B = new apsm_batch();
        try
        {
            HIP_OR_THROW( hipStreamCreateWithFlags( &B->stream, hipStreamNonBlocking ) );
            for ( int k = 0; k < kBatchSideStreams; k++ )
            {
                HIP_OR_THROW( hipStreamCreateWithFlags( &B->side[ k ], hipStreamNonBlocking ) );
                HIP_OR_THROW( hipEventCreateWithFlags( &B->join[ k ], hipEventDisableTiming ) );
            }
            HIP_OR_THROW( hipEventCreateWithFlags( &B->fork, hipEventDisableTiming ) );
            HIP_OR_THROW( hipEventCreate( &B->ev0 ) );
            HIP_OR_THROW( hipEventCreate( &B->evMid ) );
            HIP_OR_THROW( hipEventCreate( &B->ev1 ) );
            int dev = 0;
            HIP_OR_THROW( hipGetDevice( &dev ) );
            HIP_OR_THROW( hipDeviceGetAttribute( &B->cus, hipDeviceAttributeMultiprocessorCount, dev ) );
            for ( uint32_t i = 0; i < n; i++ )
            {
                apsm_handle* h = nullptr;
                const int rc = apsm_create( &params[ i ], num_antennas[ i ], num_users[ i ], &h );
                if ( rc != APSM_OK )
                    throw ApsmError { rc, std::string( "instance " ) + std::to_string( i ) + ": " + g_lastError };
                B->inst.push_back( h );
            }
        }
        catch ( ... )
        {
            apsm_batch_destroy( B );
            throw;
        }
        *out = B;
    } );
}

void apsm_batch_destroy( apsm_batch* B )
{
    if ( !B )
        return;
    if ( B->stream )
        (void)hipStreamSynchronize( B->stream );
    for ( int k = 0; k < kBatchSideStreams; k++ )
        if ( B->side[ k ] )
            (void)hipStreamSynchronize( B->side[ k ] );
    for ( int g = 0; g < 2; g++ )
    {
        if ( B->exec[ g ] )
            (void)hipGraphExecDestroy( B->exec[ g ] );
        if ( B->graph[ g ] )
            (void)hipGraphDestroy( B->graph[ g ] );
    }
    for ( apsm_handle* h : B->inst )
        apsm_destroy( h );
    B->desc.release();
    for ( hipEvent_t e : { B->ev0, B->evMid, B->ev1, B->fork } )
        if ( e )
            (void)hipEventDestroy( e );
    for ( int k = 0; k < kBatchSideStreams; k++ )
    {
        if ( B->join[ k ] )
            (void)hipEventDestroy( B->join[ k ] );
        if ( B->side[ k ] )
            (void)hipStreamDestroy( B->side[ k ] );
    }
    if ( B->stream )
        (void)hipStreamDestroy( B->stream );
    delete B;
}

int apsm_batch_plan( apsm_batch* B, const float* const* d_rx_train, const float* const* d_tx_train, const uint32_t* T,
                     const float* const* d_rx_data, const uint32_t* M, float* const* d_est, int use_graph )
{
    return guarded( [ & ] {
        ARG_CHECK( B && d_rx_train && d_tx_train && T && d_rx_data && M && d_est, "null argument" );
        B->planned = false;
        batchPlan( B, d_rx_train, d_tx_train, T, d_rx_data, M, d_est, use_graph != 0 );
    } );
}

int apsm_batch_run( apsm_batch* B )
{
    return guarded( [ & ] {
        ARG_CHECK( B != nullptr, "null batch" );
        batchRun( B );
    } );
}

int apsm_batch_synchronize( apsm_batch* B )
{
    return guarded( [ & ] {
        ARG_CHECK( B != nullptr, "null batch" );
        HIP_OR_THROW( hipStreamSynchronize( B->stream ) );
    } );
}

int apsm_batch_last_ms( apsm_batch* B, float* ms )
{
    return guarded( [ & ] {
        ARG_CHECK( B && ms, "null argument" );
        ARG_CHECK( B->timed, "no run recorded" );
        *ms = elapsedMs( B->ev0, B->ev1 );
    } );
}

int apsm_batch_last_split( apsm_batch* B, float* train_ms, float* detect_ms )
{
    return guarded( [ & ] {
        ARG_CHECK( B && train_ms && detect_ms, "null argument" );
        ARG_CHECK( B->timed, "no run recorded" );
        *train_ms = elapsedMs( B->ev0, B->evMid );
        *detect_ms = elapsedMs( B->evMid, B->ev1 );
    } );
}

apsm_handle* apsm_batch_instance( apsm_batch* B, uint32_t i ) { return ( B && i < B->inst.size() ) ? B->inst[ i ] : nullptr; }

int apsm_batch_tuning( apsm_batch* B, int32_t* table, uint32_t cap_rows, uint32_t* rows )
{
    return guarded( [ & ] {
        ARG_CHECK( B && rows, "null argument" );
        *rows = (uint32_t)B->tuning.size();
        if ( table )
        {
            ARG_CHECK( cap_rows >= B->tuning.size(), "tuning table too small" );
            for ( size_t r = 0; r < B->tuning.size(); r++ )
                for ( int c = 0; c < 8; c++ )
                    table[ r * 8 + c ] = B->tuning[ r ][ c ];
        }
    } );
}

// ---- serving ----
int apsm_server_create( const apsm_server_config* cfg, apsm_server** out )
{
    return guarded( [ & ] {
        ARG_CHECK( cfg && out, "null argument" );
        ARG_CHECK( cfg->frame_arena > 0 && cfg->frame_arena <= cfg->hbm_budget, "frame arena must be positive and within the budget" );
        *out = nullptr;
        apsm_server* S = new apsm_server();
        try
        {
            S->cfg = *cfg;
            ARG_CHECK( cfg->detects_in_flight <= (uint32_t)kServeDetectInFlightMax, "detects_in_flight must be at most 8" );
            S->detectSlots = cfg->detects_in_flight ? (int)cfg->detects_in_flight : kServeDetectInFlight;
            S->batchFrames = cfg->batch_max_frames ? cfg->batch_max_frames : kServeBatchFrames;
            S->batchBytes = cfg->batch_max_bytes ? (size_t)cfg->batch_max_bytes : kServeBatchBytes;
            S->epoch = std::chrono::steady_clock::now();
            // creation order: detect, upload, train, side (one hardware queue each while the
            // process has no other streams)
            HIP_OR_THROW( hipStreamCreateWithFlags( &S->dstream, hipStreamNonBlocking ) );
            HIP_OR_THROW( hipStreamCreateWithFlags( &S->ustream, hipStreamNonBlocking ) );
            HIP_OR_THROW( hipStreamCreateWithFlags( &S->stream, hipStreamNonBlocking ) );
            for ( int k = 0; k < kServeSideStreams; k++ )
            {
                HIP_OR_THROW( hipStreamCreateWithFlags( &S->side[ k ], hipStreamNonBlocking ) );
                HIP_OR_THROW( hipEventCreateWithFlags( &S->join[ k ], hipEventDisableTiming ) );
            }
            HIP_OR_THROW( hipEventCreateWithFlags( &S->fork, hipEventDisableTiming ) );
            for ( ServeRound& R : S->rounds )
            {
                HIP_OR_THROW( hipEventCreateWithFlags( &R.done, hipEventDisableTiming ) );
                R.desc.reserve( kServeDescBytes );
            }
            for ( ServeDetect& D : S->detects )
            {
                HIP_OR_THROW( hipEventCreateWithFlags( &D.uploaded, hipEventDisableTiming ) );
                D.desc.reserve( kServeDescBytes );
                HIP_OR_THROW( hipEventCreateWithFlags( &D.detected, hipEventDisableTiming ) );
            }
            int dev = 0;
            HIP_OR_THROW( hipGetDevice( &dev ) );
            HIP_OR_THROW( hipDeviceGetAttribute( &S->cus, hipDeviceAttributeMultiprocessorCount, dev ) );
            S->arena.ensure( (size_t)cfg->frame_arena );
            S->arenaAlloc.reset( (size_t)cfg->frame_arena );
        }
        catch ( ... )
        {
            apsm_server_destroy( S );
            throw;
        }
        *out = S;
    } );
}

void apsm_server_destroy( apsm_server* S )
{
    if ( !S )
        return;
    if ( S->stream )
        (void)hipStreamSynchronize( S->stream );
    if ( S->dstream )
        (void)hipStreamSynchronize( S->dstream );
    if ( S->ustream )
        (void)hipStreamSynchronize( S->ustream );
    serverDropFrames( S );
    for ( ServeRound& R : S->rounds )
    {
        R.desc.release();
        if ( R.done )
            (void)hipEventDestroy( R.done );
    }
    for ( ServeDetect& D : S->detects )
    {
        D.desc.release();
        if ( D.uploaded )
            (void)hipEventDestroy( D.uploaded );
        if ( D.detected )
            (void)hipEventDestroy( D.detected );
    }
    for ( ServeLink& L : S->links )
        apsm_destroy( L.h );
    for ( int k = 0; k < kServeSideStreams; k++ )
    {
        if ( S->side[ k ] )
            (void)hipStreamDestroy( S->side[ k ] );
        if ( S->join[ k ] )
            (void)hipEventDestroy( S->join[ k ] );
    }
    if ( S->fork )
        (void)hipEventDestroy( S->fork );
    if ( S->stream )
        (void)hipStreamDestroy( S->stream );
    if ( S->dstream )
        (void)hipStreamDestroy( S->dstream );
    if ( S->ustream )
        (void)hipStreamDestroy( S->ustream );
    delete S;
}

int apsm_server_add_link( apsm_server* S, const apsm_params* p, uint32_t num_antennas, uint32_t num_users, uint32_t max_pilots,
                          uint32_t* link )
{
    return guarded( [ & ] {
        ARG_CHECK( S && p && link, "null argument" );
        ARG_CHECK( max_pilots >= 1, "max_pilots must be positive" );
        apsm_handle* h = nullptr;
        const int rc = apsm_create( p, num_antennas, num_users, &h );
        if ( rc != APSM_OK )
            throw ApsmError { rc, g_lastError };
        try
        {
            checkServable( h );
            // reserve every buffer a frame can need (serving never allocates)
            const int64_t K = h->Hc + (int64_t)max_pilots;
            h->buf[ 0 ].ensure( K, h->N, h->U );
            h->buf[ 1 ].ensure( K, h->N, h->U );
            reserveBlockedScratch( h, blockedSteps( h ) );
            const size_t bytes = handleDeviceBytes( h );
            if ( S->linkBytes + bytes + S->cfg.frame_arena > S->cfg.hbm_budget )
                throw ApsmError { APSM_ERR_BUSY, "HBM budget exhausted: link not admitted" };
            ServeLink L;
            L.h = h;
            L.bytes = bytes;
            L.maxPilots = max_pilots;
            S->links.push_back( L );
            S->linkBytes += bytes;
            *link = (uint32_t)( S->links.size() - 1 );
        }
        catch ( ... )
        {
            apsm_destroy( h );
            throw;
        }
    } );
}

apsm_handle* apsm_server_link( apsm_server* S, uint32_t link ) { return ( S && link < S->links.size() ) ? S->links[ link ].h : nullptr; }

int apsm_server_submit( apsm_server* S, const apsm_frame* f, uint64_t* ticket )
{
    return guarded( [ & ] {
        ARG_CHECK( S && f, "null argument" );
        const uint64_t t = S->nextTicket++;
        serverEnqueue( S, *f, serverNow( S ), t );
        if ( ticket )
            *ticket = t;
    } );
}

int apsm_server_poll( apsm_server* S, int wait, apsm_frame_result* out, uint32_t cap, uint32_t* n )
{
    return guarded( [ & ] {
        ARG_CHECK( S && n, "null argument" );
        try
        {
            do
                serverProgress( S );
            while ( wait && !serverIdle( S ) );
        }
        catch ( ... )
        {
            serverAbort( S ); // the frames in flight are dropped; the server is idle again
            throw;
        }
        // host-side ends of the estimate copies into pageable buffers: the detect event of a
        // batch completes before the runtime has finished such a copy on the host
        if ( wait || ( out && cap > 0 && !S->finished.empty() ) )
            HIP_OR_THROW( hipStreamSynchronize( S->dstream ) );
        uint32_t k = 0;
        while ( k < cap && !S->finished.empty() && out )
        {
            ServeFrame* f = S->finished.front();
            fillResult( f, 0.0, &out[ k++ ] );
            S->finished.erase( S->finished.begin() );
            delete f;
        }
        *n = k;
    } );
}

int apsm_server_replay( apsm_server* S, const apsm_frame* frames, uint32_t n, apsm_frame_result* results )
{
    return guarded( [ & ] {
        ARG_CHECK( S && ( n == 0 || ( frames && results ) ), "null argument" );
        ARG_CHECK( serverIdle( S ) && S->finished.empty(), "replay needs an idle server" );
        for ( uint32_t i = 1; i < n; i++ )
            ARG_CHECK( frames[ i ].arrival_s >= frames[ i - 1 ].arrival_s, "replay frames must be in arrival order" );
        for ( uint32_t i = 0; i < n; i++ ) // a bad frame fails the call before any work is enqueued
            serverCheckFrame( S, frames[ i ] );
        const uint64_t ticket0 = S->nextTicket;
        S->nextTicket += n;
        const double t0 = serverNow( S );
        uint32_t next = 0, done = 0;
        try
        {
            while ( done < n )
            {
                const double now = serverNow( S ) - t0;
                while ( next < n && frames[ next ].arrival_s <= now )
                {
                    serverEnqueue( S, frames[ next ], t0 + frames[ next ].arrival_s, ticket0 + next );
                    next++;
                }
                serverProgress( S );
                for ( ServeFrame* f : S->finished )
                {
                    fillResult( f, t0, &results[ f->ticket - ticket0 ] );
                    delete f;
                    done++;
                }
                S->finished.clear();
            }
            HIP_OR_THROW( hipStreamSynchronize( S->dstream ) ); // host-side ends of pageable copies
        }
        catch ( ... )
        {
            // every frame of this replay is dropped, so no later call touches the caller's
            // buffers, and the server is idle again
            serverAbort( S );
            throw;
        }
    } );
}

int apsm_server_stats( apsm_server* S, uint64_t* rounds, uint64_t* max_frames_per_round, uint64_t* link_bytes )
{
    return guarded( [ & ] {
        ARG_CHECK( S, "null server" );
        if ( rounds )
            *rounds = S->roundsRun + S->detectsRun;
        if ( max_frames_per_round )
            *max_frames_per_round = S->maxBatch;
        if ( link_bytes )
            *link_bytes = S->linkBytes;
    } );
}

int apsm_host_malloc( void** ptr, uint64_t bytes )
{
    return guarded( [ & ] {
        ARG_CHECK( ptr && bytes > 0, "bad allocation request" );
        HIP_OR_THROW( hipHostMalloc( ptr, (size_t)bytes, hipHostMallocDefault ) );
    } );
}

int apsm_host_free( void* ptr )
{
    return guarded( [ & ] {
        if ( ptr )
            HIP_OR_THROW( hipHostFree( ptr ) );
    } );
}

} // extern "C"
