// APSM oracle: a CPU restatement of the reference APSM / LLS detector algorithm.
//
// TEST INFRASTRUCTURE ONLY. Nothing in the product (libapsm_amd/, include/) links,
// loads or calls this file. Only tests/, __graft_entry__.smoke() and the cpu_baseline
// leg of bench.py use it, as the checker or as the timed CPU baseline.
//
// Every function is templated on the real type: `double` is the FP64 oracle used for
// parity checks, `float` is the reference algorithm in FP32 on host cores (the CPU
// baseline). Paths below are relative to the reference repository root.
//
// What each part restates:
//   embedding          cpp/lib/cuda/cuda_matrix_device.cuh:180-188 (dedup matrix) and
//                      :288-296 (dedup ring buffer): row 2k = [Re x_k; Im x_k],
//                      row 2k+1 = [Im x_k; -Re x_k]; targets Re s_k / Im s_k
//                      (cuda_matrix_host.cu:131-149).
//   kernel / model     cpp/lib/apsm/apsm_detect.cu:103-152 (linear and Gaussian kernels),
//                      :170-226 (f(u) = wL*theta.u + wG*sum_j c_j exp(-0.5/var*|u-d_j|^2)).
//   variance scaling   cpp/lib/apsm/apsm_noma_detector.cu:79-80 (var <- v*N^2 if v != 0).
//   train schedule     cpp/lib/apsm/apsm_train.cu:101-160 (host loop, moveWindow, passes)
//                      and :333-428 (window projection, soft threshold, normalisation,
//                      coefficient and linear-weight update from the pre-step state).
//   norm constraint    cpp/lib/apsm/apsm_train.cu:192-256.
//   norm ratio         cpp/lib/apsm/apsm_noma_detector.cu:220-319.
//   train history      cpp/lib/noma/noma_train_history.cu:105-236 (per-step EVM/SER/BER).
//   variance heuristic cpp/lib/apsm/apsm_noma_detector.cu:194-218.
//   shuffle            cpp/lib/noma/noma_detector.cu:59-74.
//   LLS                cpp/lib/noma/lls_detector.cu:136-218 (least squares over the real
//                      embedding, estimate = A_d w).
//   LLS init           cpp/lib/apsm/apsm_noma_detector.cu:321-346 (theta = w / wL).
//   metrics            cpp/cli/common/util.cpp:62-192, cpp/cli/common/modulation.cpp:68-151.
//
// Decisions where the reference is undefined (each is also what the HIP path implements):
//   * Dictionary wrap (2T > 2*dictionarySize): FIFO. The dictionary holds the newest
//     H = 2*dictionarySize real rows of the pilot stream; an evicted row's Gaussian
//     coefficient is dropped, its share of theta stays. Window rows are the newest
//     W rows of the stream. For 2T <= H this is exactly the reference.
//   * A second train() without resetState() continues the same stream:
//     train(A); train(B) == train(A||B).
//   * Rows of a call that do not fill a whole sampleStep never enter (the reference's
//     loop bound `sample_idx <= width`, apsm_train.cu:109-110).
//   * trainPasses > 1 is allowed only on the first call after a reset (the reference
//     asserts 2*dictionarySize == 2T there, apsm_train.cu:88); later passes re-project
//     the final window exactly as the reference does (apsm_train.cu:112-118,333-346).
//
// Two train forms:
//   naive       per step, f_t for every window row against the whole dictionary, exactly
//               the reference's arithmetic order of operations (O(W*G*L) per step).
//   incremental f of a row is computed once against the dictionary when the row enters
//               the window; afterwards it is updated with the window Gram matrix times the
//               projection weights, and evictions subtract their Gaussian share
//               (O(G*L + W^2) per step). Algebraically identical to naive.

#include <algorithm>
#include <cmath>
#include <complex>
#include <cstdint>
#include <cstring>
#include <random>
#include <stdexcept>
#include <string>
#include <vector>

#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

struct OrcParams
{
    double gaussian_kernel_weight; // wG
    double gaussian_kernel_variance; // JSON value v; 0 selects the heuristic
    double eB;
    double norm_constraint;
    int32_t window_size; // complex samples; real window 2*window_size
    int32_t sample_step; // real rows per step
    int32_t train_passes;
    int32_t dictionary_size; // complex samples; real capacity 2*dictionary_size
    int32_t start_with_full_window;
    int32_t training_shuffle;
    int32_t lls_initialization;
    int32_t pad;
    double linear_kernel_weight; // wL (the reference factory sets 1 - wG)
};

thread_local std::string g_err;

template <typename R>
struct Model
{
    OrcParams p;
    int N = 0, L = 0;
    int64_t H = 0;
    R wL = 0, wG = 0, var = 0;
    bool varSet = false;
    std::vector<R> X; // every row ever entered, row-major [row][L]
    std::vector<R> y;
    std::vector<R> c;
    std::vector<R> theta;
    int64_t used = 0; // rows entered into the stream
    bool anyCall = false;
    // incremental-form state
    std::vector<R> F; // f of every row while it is in the window (indexed by absolute row)

    int64_t dictLo( int64_t s ) const { return std::max<int64_t>( 0, s - H ); }
};

template <typename R>
inline R dot( const R* a, const R* b, int L )
{
    R acc = 0;
    for ( int d = 0; d < L; d++ )
        acc += a[ d ] * b[ d ];
    return acc;
}

// Gaussian kernel exp(-0.5/var*|a-b|^2) (apsm_detect.cu:131-152).
template <typename R>
inline R gauss( const R* a, const R* b, int L, R var )
{
    R acc = 0;
    for ( int d = 0; d < L; d++ )
    {
        R e = a[ d ] - b[ d ];
        acc += e * e;
    }
    return std::exp( R( -0.5 ) / var * acc );
}

// Embedding of complex sample k of rx laid out [ant][sample][2] into the two real rows
// (cuda_matrix_device.cuh:180-188).
template <typename R>
inline void embed( const float* rx, int N, int64_t S, int64_t k, R* a, R* b )
{
    for ( int n = 0; n < N; n++ )
    {
        R re = rx[ ( (int64_t)n * S + k ) * 2 + 0 ];
        R im = rx[ ( (int64_t)n * S + k ) * 2 + 1 ];
        a[ n ] = re;
        a[ N + n ] = im;
        b[ n ] = im;
        b[ N + n ] = -re;
    }
}

// f(u) over dictionary rows [lo, hi) (apsm_detect.cu:170-226).
template <typename R>
R evalModel( const Model<R>& m, const R* u, int64_t lo, int64_t hi )
{
    R lin = m.wL * dot( m.theta.data(), u, m.L );
    R gs = 0;
    for ( int64_t j = lo; j < hi; j++ )
    {
        R cj = m.c[ j ];
        if ( cj == R( 0 ) )
            continue; // exact: contributes 0
        gs += m.wG * gauss( &m.X[ j * m.L ], u, m.L, m.var ) * cj;
    }
    return lin + gs;
}

// Soft threshold onto the hyperslab |r| <= eB (apsm_train.cu:385-392).
template <typename R>
inline R softThreshold( R r, R eB )
{
    if ( r > eB )
        return r - eB;
    if ( r < -eB )
        return r + eB;
    return 0;
}

// Norm constraint (apsm_train.cu:192-256): n = sqrt(wG * c^T K c) over the dictionary;
// if n > normConstraint, scale c by normConstraint / n.
template <typename R>
void adaptCoeffs( Model<R>& m, int64_t s )
{
    int64_t lo = m.dictLo( s );
    R acc = 0;
    for ( int64_t i = lo; i < s; i++ )
        for ( int64_t j = lo; j < s; j++ )
            acc += m.wG * gauss( &m.X[ i * m.L ], &m.X[ j * m.L ], m.L, m.var ) * m.c[ i ] * m.c[ j ];
    R n = acc;
    if ( n > 0 )
        n = std::sqrt( n );
    if ( n > R( m.p.norm_constraint ) )
    {
        R f = R( m.p.norm_constraint ) / n;
        for ( int64_t i = lo; i < s; i++ )
            m.c[ i ] *= f;
    }
}

// One projection step of the naive form over window rows [s-W, s-W+nb), dictionary
// [lo, s), normalised by W (apsm_train.cu:333-428).
template <typename R>
void naiveStep( Model<R>& m, int64_t s, int64_t W, int64_t nb )
{
    const int L = m.L;
    const int64_t lo = m.dictLo( s );
    const int64_t t0 = s - W;
    std::vector<R> beta( nb );
#pragma omp parallel for schedule( static )
    for ( int64_t b = 0; b < nb; b++ )
    {
        const int64_t t = t0 + b;
        const R* xt = &m.X[ t * L ];
        R f = evalModel( m, xt, lo, s );
        R r = m.y[ t ] - f;
        R w = softThreshold( r, R( m.p.eB ) );
        R linearNorm = m.wL * dot( xt, xt, L );
        R gaussianNorm = m.wG;
        w /= R( W );
        w /= ( linearNorm + gaussianNorm );
        beta[ b ] = w;
    }
    for ( int64_t b = 0; b < nb; b++ )
    {
        const int64_t t = t0 + b;
        m.c[ t ] += beta[ b ];
        for ( int d = 0; d < L; d++ )
            m.theta[ d ] += beta[ b ] * m.X[ t * L + d ];
    }
}

template <typename R>
R kappa( const Model<R>& m, int64_t i, int64_t j )
{
    const R* a = &m.X[ i * m.L ];
    const R* b = &m.X[ j * m.L ];
    return m.wL * dot( a, b, m.L ) + m.wG * gauss( a, b, m.L, m.var );
}

// One step of the incremental form. New rows [sPrev, s) enter; rows [loPrev, lo) leave
// the dictionary. F holds f of every row of the current window under the pre-step state.
template <typename R>
void incrementalStep( Model<R>& m, int64_t sPrev, int64_t s, int64_t W, int64_t nb, bool moved )
{
    const int L = m.L;
    const int64_t t0 = s - W;
    if ( moved )
    {
        const int64_t loPrev = m.dictLo( sPrev );
        const int64_t lo = m.dictLo( s );
        // evictions: rows still in the window lose the evicted rows' Gaussian share
        for ( int64_t e = loPrev; e < lo; e++ )
        {
            if ( m.c[ e ] == R( 0 ) )
                continue;
            for ( int64_t u = t0; u < sPrev; u++ )
                m.F[ u ] -= m.wG * m.c[ e ] * gauss( &m.X[ e * L ], &m.X[ u * L ], L, m.var );
        }
        // entering rows: f against the current dictionary (rows in [sPrev, s) have c = 0)
        if ( (int64_t)m.F.size() < s )
            m.F.resize( s, 0 );
#pragma omp parallel for schedule( static )
        for ( int64_t t = sPrev; t < s; t++ )
            m.F[ t ] = evalModel( m, &m.X[ t * L ], lo, sPrev );
    }
    std::vector<R> beta( nb );
    for ( int64_t b = 0; b < nb; b++ )
    {
        const int64_t t = t0 + b;
        const R* xt = &m.X[ t * L ];
        R r = m.y[ t ] - m.F[ t ];
        R w = softThreshold( r, R( m.p.eB ) );
        w /= R( W );
        w /= ( m.wL * dot( xt, xt, L ) + m.wG );
        beta[ b ] = w;
    }
    for ( int64_t b = 0; b < nb; b++ )
    {
        const int64_t t = t0 + b;
        m.c[ t ] += beta[ b ];
        for ( int d = 0; d < L; d++ )
            m.theta[ d ] += beta[ b ] * m.X[ t * L + d ];
    }
    // window rows absorb the update through the window Gram matrix
#pragma omp parallel for schedule( static )
    for ( int64_t u = t0; u < s; u++ )
    {
        R acc = 0;
        for ( int64_t b = 0; b < nb; b++ )
            if ( beta[ b ] != R( 0 ) )
                acc += beta[ b ] * kappa( m, t0 + b, u );
        m.F[ u ] += acc;
    }
}

template <typename R>
void recomputeWindowF( Model<R>& m, int64_t s, int64_t W )
{
    const int64_t lo = m.dictLo( s );
    for ( int64_t u = s - W; u < s; u++ )
        m.F[ u ] = evalModel( m, &m.X[ u * m.L ], lo, s );
}

// Gaussian variance heuristic (apsm_noma_detector.cu:194-218), arithmetic in FP32 as in
// the reference so the percentile index is the same.
float varianceHeuristic( const float* rx, int N, int64_t T, const std::vector<int64_t>& perm )
{
    std::vector<float> nd;
    nd.reserve( T * ( T - 1 ) / 2 );
    for ( int64_t i = 0; i < T; i++ )
        for ( int64_t j = i + 1; j < T; j++ )
        {
            float dist = 0;
            for ( int n = 0; n < N; n++ )
            {
                int64_t pi = perm[ i ], pj = perm[ j ];
                std::complex<float> a( rx[ ( n * T + pi ) * 2 ], rx[ ( n * T + pi ) * 2 + 1 ] );
                std::complex<float> b( rx[ ( n * T + pj ) * 2 ], rx[ ( n * T + pj ) * 2 + 1 ] );
                dist += std::abs( a - b );
            }
            nd.push_back( std::sqrt( dist ) );
        }
    std::sort( nd.begin(), nd.end() );
    const float p = 0.7f, q = 5.f;
    // float element times a double power, rounded to float, then squared in float
    float v = (float)( (double)nd[ (size_t)( nd.size() * p ) ] * std::pow( (double)T, -1. / q ) );
    return v * v;
}

// Permutation applied by the reference shuffle (noma_detector.cu:59-74): every vector is
// shuffled with its own copy of the same engine, so all see the same permutation.
std::vector<int64_t> shufflePermutation( int64_t T, bool shuffle )
{
    std::vector<int64_t> perm( T );
    for ( int64_t i = 0; i < T; i++ )
        perm[ i ] = i;
    if ( shuffle )
    {
        std::seed_seq seed { 1, 2, 3, 4, 5, 6, 7, 8 };
        std::mt19937 rng( seed );
        std::shuffle( perm.begin(), perm.end(), rng );
    }
    return perm;
}

// ---------------------------------------------------------------------------------------
// Least squares (lls_detector.cu:136-218): min_w |A w - y|_2 over the real embedding.
// ---------------------------------------------------------------------------------------

// Normal equations A^T A w = A^T y solved by Cholesky, all FP64.
int llsCholesky( const float* rx, const float* tx, int N, int64_t T, double* w )
{
    const int L = 2 * N;
    std::vector<double> G( L * L, 0.0 ), r( L, 0.0 ), a( L ), b( L );
    for ( int64_t k = 0; k < T; k++ )
    {
        embed<double>( rx, N, T, k, a.data(), b.data() );
        double yr = tx[ 2 * k ], yi = tx[ 2 * k + 1 ];
        for ( int i = 0; i < L; i++ )
        {
            for ( int j = 0; j < L; j++ )
                G[ i * L + j ] += a[ i ] * a[ j ] + b[ i ] * b[ j ];
            r[ i ] += a[ i ] * yr + b[ i ] * yi;
        }
    }
    // G = C C^T
    std::vector<double> C( L * L, 0.0 );
    for ( int j = 0; j < L; j++ )
    {
        double d = G[ j * L + j ];
        for ( int k = 0; k < j; k++ )
            d -= C[ j * L + k ] * C[ j * L + k ];
        if ( !( d > 0 ) )
            return -1;
        C[ j * L + j ] = std::sqrt( d );
        for ( int i = j + 1; i < L; i++ )
        {
            double s = G[ i * L + j ];
            for ( int k = 0; k < j; k++ )
                s -= C[ i * L + k ] * C[ j * L + k ];
            C[ i * L + j ] = s / C[ j * L + j ];
        }
    }
    std::vector<double> z( L );
    for ( int i = 0; i < L; i++ )
    {
        double s = r[ i ];
        for ( int k = 0; k < i; k++ )
            s -= C[ i * L + k ] * z[ k ];
        z[ i ] = s / C[ i * L + i ];
    }
    for ( int i = L - 1; i >= 0; i-- )
    {
        double s = z[ i ];
        for ( int k = i + 1; k < L; k++ )
            s -= C[ k * L + i ] * w[ k ];
        w[ i ] = s / C[ i * L + i ];
    }
    return 0;
}

// Householder QR of the 2T x 2N real matrix, then R w = Q^T y. FP64.
int llsQr( const float* rx, const float* tx, int N, int64_t T, double* w )
{
    const int L = 2 * N;
    const int64_t m = 2 * T;
    if ( m < L )
        return -1;
    std::vector<double> A( m * L ), y( m ), a( L ), b( L );
    for ( int64_t k = 0; k < T; k++ )
    {
        embed<double>( rx, N, T, k, a.data(), b.data() );
        for ( int d = 0; d < L; d++ )
        {
            A[ ( 2 * k ) * L + d ] = a[ d ];
            A[ ( 2 * k + 1 ) * L + d ] = b[ d ];
        }
        y[ 2 * k ] = tx[ 2 * k ];
        y[ 2 * k + 1 ] = tx[ 2 * k + 1 ];
    }
    std::vector<double> v( m );
    for ( int j = 0; j < L; j++ )
    {
        double nrm = 0;
        for ( int64_t i = j; i < m; i++ )
            nrm += A[ i * L + j ] * A[ i * L + j ];
        nrm = std::sqrt( nrm );
        if ( nrm == 0 )
            return -1;
        double alpha = A[ j * L + j ] > 0 ? -nrm : nrm;
        for ( int64_t i = j; i < m; i++ )
            v[ i ] = A[ i * L + j ];
        v[ j ] -= alpha;
        double vn = 0;
        for ( int64_t i = j; i < m; i++ )
            vn += v[ i ] * v[ i ];
        if ( vn == 0 )
            continue;
        for ( int k = j; k < L; k++ )
        {
            double s = 0;
            for ( int64_t i = j; i < m; i++ )
                s += v[ i ] * A[ i * L + k ];
            s = 2 * s / vn;
            for ( int64_t i = j; i < m; i++ )
                A[ i * L + k ] -= s * v[ i ];
        }
        double s = 0;
        for ( int64_t i = j; i < m; i++ )
            s += v[ i ] * y[ i ];
        s = 2 * s / vn;
        for ( int64_t i = j; i < m; i++ )
            y[ i ] -= s * v[ i ];
    }
    for ( int i = L - 1; i >= 0; i-- )
    {
        double s = y[ i ];
        for ( int k = i + 1; k < L; k++ )
            s -= A[ i * L + k ] * w[ k ];
        w[ i ] = s / A[ i * L + i ];
    }
    return 0;
}

template <typename R>
int trainImpl( Model<R>& m, const float* rx, const float* tx, int64_t T, int incremental )
{
    const OrcParams& p = m.p;
    const int N = m.N, L = m.L;
    const int64_t q = p.sample_step;
    if ( q <= 0 || ( q % 2 ) != 0 )
    {
        g_err = "sampleStep must be a positive even number";
        return -1;
    }
    if ( p.window_size <= 0 || p.window_size > p.dictionary_size )
    {
        g_err = "windowSize must be in [1, dictionarySize]";
        return -1;
    }
    const int64_t base = m.used;
    if ( p.train_passes > 1 && base != 0 )
    {
        g_err = "trainPasses > 1 requires a fresh state";
        return -1;
    }
    std::vector<int64_t> perm = shufflePermutation( T, p.training_shuffle != 0 );

    if ( !m.varSet )
    {
        if ( p.gaussian_kernel_variance != 0 )
            m.var = R( R( p.gaussian_kernel_variance ) * R( N * N ) );
        else
            m.var = R( varianceHeuristic( rx, N, T, perm ) );
        m.varSet = true;
    }
    if ( p.lls_initialization && !m.anyCall )
    {
        std::vector<double> w( L );
        if ( llsCholesky( rx, tx, N, T, w.data() ) != 0 )
        {
            g_err = "LLS initialisation failed";
            return -1;
        }
        for ( int d = 0; d < L; d++ )
            m.theta[ d ] = R( w[ d ] ) / m.wL;
    }
    m.anyCall = true;

    // append rows of this call
    const int64_t R2 = 2 * T;
    m.X.resize( ( base + R2 ) * L );
    m.y.resize( base + R2 );
    m.c.resize( base + R2, 0 );
    for ( int64_t k = 0; k < T; k++ )
    {
        int64_t pk = perm[ k ];
        embed<R>( rx, N, T, pk, &m.X[ ( base + 2 * k ) * L ], &m.X[ ( base + 2 * k + 1 ) * L ] );
        m.y[ base + 2 * k ] = tx[ 2 * pk ];
        m.y[ base + 2 * k + 1 ] = tx[ 2 * pk + 1 ];
    }
    const int64_t w2 = 2 * (int64_t)p.window_size;
    const bool fullStart = p.start_with_full_window && base == 0;
    const int64_t first = fullStart ? w2 : q;
    int64_t sPrev = base;
    for ( int pass = 0; pass < std::max( 1, p.train_passes ); pass++ )
    {
        for ( int64_t sl = first; sl <= R2; sl += q )
        {
            int64_t s, W, nb;
            bool moved = ( pass == 0 );
            if ( moved )
            {
                s = base + sl;
                m.used = s;
                W = std::min( w2, s );
                nb = W;
            }
            else
            {
                s = m.used;
                W = std::min( w2, s );
                nb = std::min( W, std::min( w2, sl ) );
            }
            if ( incremental )
                incrementalStep( m, sPrev, s, W, nb, moved );
            else
                naiveStep( m, s, W, nb );
            if ( p.norm_constraint > 0 )
            {
                adaptCoeffs( m, s );
                if ( incremental )
                    recomputeWindowF( m, s, W );
            }
            sPrev = s;
        }
    }
    return 0;
}

template <typename R>
void detectImpl( const Model<R>& m, const float* rx, int64_t M, double* est )
{
    const int N = m.N, L = m.L;
    const int64_t lo = m.dictLo( m.used ), hi = m.used;
#pragma omp parallel
    {
        std::vector<R> a( L ), b( L );
#pragma omp for schedule( static )
        for ( int64_t k = 0; k < M; k++ )
        {
            embed<R>( rx, N, M, k, a.data(), b.data() );
            est[ 2 * k ] = (double)evalModel( m, a.data(), lo, hi );
            est[ 2 * k + 1 ] = (double)evalModel( m, b.data(), lo, hi );
        }
    }
}

template <typename R>
Model<R>* createImpl( const OrcParams* p, int N )
{
    if ( N <= 0 || p->dictionary_size <= 0 )
    {
        g_err = "numAntennas and dictionarySize must be positive";
        return nullptr;
    }
    Model<R>* m = new Model<R>();
    m->p = *p;
    m->N = N;
    m->L = 2 * N;
    m->H = 2 * (int64_t)p->dictionary_size;
    m->wG = R( p->gaussian_kernel_weight );
    m->wL = R( p->linear_kernel_weight );
    m->theta.assign( m->L, 0 );
    return m;
}

// Modulation (modulation.cpp:68-151, modulation.hpp:157-222).
double modScale( int bps )
{
    if ( bps == 1 || bps == 2 )
        return 1.0 / std::sqrt( (double)bps );
    uint32_t v = ( ( 2u << bps ) - 2 ) / 3;
    return 1.0 / std::sqrt( (double)v );
}

uint32_t demap1d( double x, uint32_t ns )
{
    int32_t xi = (int32_t)std::round( ( x + ns - 1 ) / 2 );
    xi = std::min( (int32_t)ns - 1, std::max( 0, xi ) );
    return xi ^ ( xi >> 1 );
}

// ---------------------------------------------------------------------------------------
// Streamed multi-user form (FP64): the incremental form above for U detectors that share
// one received pilot stream (one target row per user), with the window Gram matrix kept
// between steps. Per step only the Gram entries of the entering rows are new (O(W*L)
// kernel evaluations instead of O(W^2*L)), and the kernel values of an entering row
// against the dictionary are computed once for all users. Every value is computed with
// the same expressions, in the same order, as incrementalStep / evalModel, so each user's
// result equals the single-user incremental form (and, through it, the naive reference
// schedule of apsm_train.cu:101-160,333-428). It exists so that workloads of 10^5 steps
// (200k pilots, window 160, dictionary 4096, 10 users) finish in minutes on host cores.
// Scope: trainPasses 1, normConstraint 0, no LLS initialisation, no shuffle.
// ---------------------------------------------------------------------------------------
struct Multi
{
    OrcParams p;
    int N = 0, L = 0, U = 0;
    int64_t H = 0, w2 = 0;
    double wL = 0, wG = 0, var = 0;
    bool varSet = false;
    std::vector<double> X; // every row ever entered [row][L]
    std::vector<double> y; // [row][U]
    std::vector<double> c; // [row][U]
    std::vector<double> theta; // [U][L]
    std::vector<double> F; // [slot][U], slot = row mod w2 (rows of the window)
    std::vector<double> Kw; // [slot][slot] window Gram (wL x.x + wG g)
    int64_t used = 0;

    int64_t dictLo( int64_t s ) const { return std::max<int64_t>( 0, s - H ); }
    const double* row( int64_t r ) const { return &X[ r * L ]; }
    int64_t slot( int64_t r ) const { return r % w2; }
};

double multiKappa( const Multi& m, int64_t i, int64_t j )
{
    const double* a = m.row( i );
    const double* b = m.row( j );
    return m.wL * dot( a, b, m.L ) + m.wG * gauss( a, b, m.L, m.var );
}

void multiStep( Multi& m, int64_t sPrev, int64_t s, int64_t W )
{
    const int L = m.L, U = m.U;
    const int64_t t0 = s - W;
    const int64_t loPrev = m.dictLo( sPrev ), lo = m.dictLo( s );
    // evictions: window rows that stay lose the evicted rows' Gaussian share
    for ( int64_t e = loPrev; e < lo; e++ )
    {
#pragma omp parallel for schedule( static )
        for ( int64_t u = t0; u < sPrev; u++ )
        {
            double g = -1.0;
            for ( int k = 0; k < U; k++ )
            {
                const double ce = m.c[ e * U + k ];
                if ( ce == 0.0 )
                    continue;
                if ( g < 0.0 )
                    g = gauss( m.row( e ), m.row( u ), L, m.var );
                m.F[ m.slot( u ) * U + k ] -= m.wG * ce * g;
            }
        }
    }
    // entering rows: f against the dictionary [lo, sPrev) under the current state
    const int64_t nd = sPrev - lo;
    std::vector<double> g( std::max<int64_t>( nd, 1 ) );
    for ( int64_t t = std::max( sPrev, t0 ); t < s; t++ ) // rows below t0 never enter the window
    {
        const double* xt = m.row( t );
#pragma omp parallel for schedule( static )
        for ( int64_t j = 0; j < nd; j++ )
            g[ j ] = gauss( m.row( lo + j ), xt, L, m.var );
#pragma omp parallel for schedule( static )
        for ( int k = 0; k < U; k++ )
        {
            double lin = m.wL * dot( &m.theta[ (size_t)k * L ], xt, L );
            double gs = 0;
            for ( int64_t j = 0; j < nd; j++ )
            {
                const double cj = m.c[ ( lo + j ) * U + k ];
                if ( cj == 0.0 )
                    continue;
                gs += m.wG * g[ j ] * cj;
            }
            m.F[ m.slot( t ) * U + k ] = lin + gs;
        }
    }
    // Gram entries of the entering rows against the window
#pragma omp parallel for schedule( static )
    for ( int64_t v = t0; v < s; v++ )
        for ( int64_t t = std::max( sPrev, t0 ); t < s; t++ )
        {
            const double kv = multiKappa( m, t, v );
            m.Kw[ m.slot( t ) * m.w2 + m.slot( v ) ] = kv;
            m.Kw[ m.slot( v ) * m.w2 + m.slot( t ) ] = kv;
        }
    // projections of every window row under the pre-step state, per user
    std::vector<double> beta( (size_t)W * U );
#pragma omp parallel for schedule( static )
    for ( int k = 0; k < U; k++ )
    {
        for ( int64_t b = 0; b < W; b++ )
        {
            const int64_t t = t0 + b;
            const double* xt = m.row( t );
            double r = m.y[ t * U + k ] - m.F[ m.slot( t ) * U + k ];
            double w = softThreshold( r, m.p.eB );
            w /= double( W );
            w /= ( m.wL * dot( xt, xt, L ) + m.wG );
            beta[ (size_t)k * W + b ] = w;
        }
        for ( int64_t b = 0; b < W; b++ )
        {
            const int64_t t = t0 + b;
            const double bb = beta[ (size_t)k * W + b ];
            m.c[ t * U + k ] += bb;
            for ( int d = 0; d < L; d++ )
                m.theta[ (size_t)k * L + d ] += bb * m.X[ t * L + d ];
        }
    }
    // window rows absorb the update through the window Gram matrix
#pragma omp parallel for collapse( 2 ) schedule( static )
    for ( int k = 0; k < U; k++ )
        for ( int64_t u = t0; u < s; u++ )
        {
            const double* kr = &m.Kw[ m.slot( u ) * m.w2 ];
            const double* bk = &beta[ (size_t)k * W ];
            double acc = 0;
            for ( int64_t b = 0; b < W; b++ )
                if ( bk[ b ] != 0.0 )
                    acc += bk[ b ] * kr[ m.slot( t0 + b ) ];
            m.F[ m.slot( u ) * U + k ] += acc;
        }
}

int multiTrain( Multi& m, const float* rx, const float* tx, int64_t T )
{
    const OrcParams& p = m.p;
    const int N = m.N, L = m.L, U = m.U;
    const int64_t q = p.sample_step;
    if ( q <= 0 || ( q % 2 ) != 0 || p.window_size <= 0 || p.window_size > p.dictionary_size )
    {
        g_err = "sampleStep must be a positive even number and windowSize in [1, dictionarySize]";
        return -1;
    }
    if ( p.train_passes > 1 || p.norm_constraint > 0 || p.lls_initialization || p.training_shuffle )
    {
        g_err = "the streamed form needs trainPasses 1, no normConstraint, no LLS initialisation, no shuffle";
        return -1;
    }
    const int64_t base = m.used;
    if ( !m.varSet )
    {
        if ( p.gaussian_kernel_variance != 0 )
            m.var = double( p.gaussian_kernel_variance ) * double( N * N );
        else
            m.var = double( varianceHeuristic( rx, N, T, shufflePermutation( T, false ) ) );
        m.varSet = true;
    }
    const int64_t R2 = 2 * T;
    m.X.resize( ( base + R2 ) * L );
    m.y.resize( ( base + R2 ) * U );
    m.c.resize( ( base + R2 ) * U, 0.0 );
    std::vector<double> a( L ), b( L );
    for ( int64_t k = 0; k < T; k++ )
    {
        embed<double>( rx, N, T, k, &m.X[ ( base + 2 * k ) * L ], &m.X[ ( base + 2 * k + 1 ) * L ] );
        for ( int u = 0; u < U; u++ )
        {
            m.y[ ( base + 2 * k ) * U + u ] = tx[ ( (int64_t)u * T + k ) * 2 ];
            m.y[ ( base + 2 * k + 1 ) * U + u ] = tx[ ( (int64_t)u * T + k ) * 2 + 1 ];
        }
    }
    const bool fullStart = p.start_with_full_window && base == 0;
    const int64_t first = fullStart ? m.w2 : q;
    int64_t sPrev = base;
    for ( int64_t sl = first; sl <= R2; sl += q )
    {
        const int64_t s = base + sl;
        m.used = s;
        multiStep( m, sPrev, s, std::min( m.w2, s ) );
        sPrev = s;
    }
    return 0;
}

// est [U][M][2]
void multiDetect( const Multi& m, const float* rx, int64_t M, double* est )
{
    const int N = m.N, L = m.L, U = m.U;
    const int64_t lo = m.dictLo( m.used ), hi = m.used;
#pragma omp parallel
    {
        std::vector<double> a( L ), b( L ), ga( U ), gb( U );
#pragma omp for schedule( dynamic, 64 )
        for ( int64_t k = 0; k < M; k++ )
        {
            embed<double>( rx, N, M, k, a.data(), b.data() );
            std::fill( ga.begin(), ga.end(), 0.0 );
            std::fill( gb.begin(), gb.end(), 0.0 );
            for ( int64_t j = lo; j < hi; j++ )
            {
                const double* xj = m.row( j );
                double ea = -1.0, eb = -1.0;
                for ( int u = 0; u < U; u++ )
                {
                    const double cj = m.c[ j * U + u ];
                    if ( cj == 0.0 )
                        continue;
                    if ( ea < 0.0 )
                    {
                        ea = gauss( xj, a.data(), L, m.var );
                        eb = gauss( xj, b.data(), L, m.var );
                    }
                    ga[ u ] += m.wG * ea * cj;
                    gb[ u ] += m.wG * eb * cj;
                }
            }
            for ( int u = 0; u < U; u++ )
            {
                const double* th = &m.theta[ (size_t)u * L ];
                est[ ( (int64_t)u * M + k ) * 2 ] = m.wL * dot( th, a.data(), L ) + ga[ u ];
                est[ ( (int64_t)u * M + k ) * 2 + 1 ] = m.wL * dot( th, b.data(), L ) + gb[ u ];
            }
        }
    }
}

} // namespace

extern "C" {

const char* orc_last_error( void ) { return g_err.c_str(); }

int orc_omp_threads( void )
{
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

void orc_set_threads( int n )
{
#ifdef _OPENMP
    if ( n > 0 )
        omp_set_num_threads( n );
#else
    (void)n;
#endif
}

// ---- FP64 oracle ----
void* orc_apsm_create( const OrcParams* p, int N ) { return createImpl<double>( p, N ); }
void orc_apsm_destroy( void* h ) { delete static_cast<Model<double>*>( h ); }

void orc_apsm_reset( void* h )
{
    Model<double>* m = static_cast<Model<double>*>( h );
    OrcParams p = m->p;
    int N = m->N;
    bool varSet = m->varSet;
    double var = m->var;
    Model<double>* fresh = createImpl<double>( &p, N );
    *m = *fresh;
    delete fresh;
    // the scaled variance survives a reset (it lives in the parameter set)
    m->varSet = varSet;
    m->var = var;
}

int orc_apsm_train( void* h, const float* rx, const float* tx, int64_t T, int incremental )
{
    try
    {
        return trainImpl( *static_cast<Model<double>*>( h ), rx, tx, T, incremental );
    }
    catch ( const std::exception& e )
    {
        g_err = e.what();
        return -1;
    }
}

void orc_apsm_detect( void* h, const float* rx, int64_t M, double* est )
{
    detectImpl( *static_cast<Model<double>*>( h ), rx, M, est );
}

// State dump: stream length, dictionary range [lo, used), scaled variance, theta (L),
// coefficients of dictionary rows in stream order (up to `cap` values).
int64_t orc_apsm_state( void* h, double* theta, double* coeffs, int64_t cap, int64_t* lo, double* var )
{
    const Model<double>* m = static_cast<Model<double>*>( h );
    for ( int d = 0; d < m->L; d++ )
        theta[ d ] = m->theta[ d ];
    int64_t l = m->dictLo( m->used );
    *lo = l;
    *var = m->var;
    int64_t n = std::min<int64_t>( cap, m->used - l );
    for ( int64_t i = 0; i < n; i++ )
        coeffs[ i ] = m->c[ l + i ];
    return m->used;
}

// computeNormRatio (apsm_noma_detector.cu:220-319): (wL * |theta|^2) / (wG * sum_ij c_i c_j
// K(d_i, d_j)) over the dictionary rows [lo, used).
double orc_apsm_norm_ratio( void* h )
{
    const Model<double>* m = static_cast<Model<double>*>( h );
    const int64_t lo = m->dictLo( m->used );
    double g = 0;
    for ( int64_t i = lo; i < m->used; i++ )
        for ( int64_t j = lo; j < m->used; j++ )
            g += gauss( &m->X[ i * m->L ], &m->X[ j * m->L ], m->L, m->var ) * m->c[ i ] * m->c[ j ];
    const double lin = dot( m->theta.data(), m->theta.data(), m->L );
    return ( m->wL * lin ) / ( m->wG * g );
}

// f(u) for arbitrary real rows u [n][L] under the current state (for kernel-level checks).
void orc_apsm_eval_rows( void* h, const double* u, int64_t n, double* out )
{
    const Model<double>* m = static_cast<Model<double>*>( h );
    const int64_t lo = m->dictLo( m->used );
    for ( int64_t i = 0; i < n; i++ )
        out[ i ] = evalModel( *m, u + i * m->L, lo, m->used );
}

// Load an explicit state: dictionary rows [G][L] (stream order), coefficients [G], theta [L].
int orc_apsm_set_state( void* h, const double* rows, const double* coeffs, const double* theta, int64_t G, double var )
{
    Model<double>* m = static_cast<Model<double>*>( h );
    if ( G > m->H )
    {
        g_err = "more rows than the dictionary holds";
        return -1;
    }
    m->X.assign( rows, rows + G * m->L );
    m->c.assign( coeffs, coeffs + G );
    m->y.assign( G, 0 );
    m->theta.assign( theta, theta + m->L );
    m->used = G;
    m->var = var;
    m->varSet = true;
    m->anyCall = true;
    return 0;
}

// ---- FP64 oracle, streamed multi-user form ----
void* orc_multi_create( const OrcParams* p, int N, int U )
{
    if ( N <= 0 || U <= 0 || p->dictionary_size <= 0 )
    {
        g_err = "numAntennas, users and dictionarySize must be positive";
        return nullptr;
    }
    Multi* m = new Multi();
    m->p = *p;
    m->N = N;
    m->L = 2 * N;
    m->U = U;
    m->H = 2 * (int64_t)p->dictionary_size;
    m->w2 = 2 * (int64_t)p->window_size;
    m->wG = p->gaussian_kernel_weight;
    m->wL = p->linear_kernel_weight;
    m->theta.assign( (size_t)U * m->L, 0.0 );
    m->F.assign( (size_t)std::max<int64_t>( m->w2, 1 ) * U, 0.0 );
    m->Kw.assign( (size_t)std::max<int64_t>( m->w2, 1 ) * std::max<int64_t>( m->w2, 1 ), 0.0 );
    return m;
}

void orc_multi_destroy( void* h ) { delete static_cast<Multi*>( h ); }

// rx [N][T][2], tx [U][T][2]
int orc_multi_train( void* h, const float* rx, const float* tx, int64_t T )
{
    try
    {
        return multiTrain( *static_cast<Multi*>( h ), rx, tx, T );
    }
    catch ( const std::exception& e )
    {
        g_err = e.what();
        return -1;
    }
}

// est [U][M][2]
void orc_multi_detect( void* h, const float* rx, int64_t M, double* est ) { multiDetect( *static_cast<Multi*>( h ), rx, M, est ); }

// theta [U][L]; coeffs [U][cap] of the dictionary rows [lo, used) in stream order.
int64_t orc_multi_state( void* h, double* theta, double* coeffs, int64_t cap, int64_t* lo, double* var )
{
    const Multi* m = static_cast<Multi*>( h );
    std::copy( m->theta.begin(), m->theta.end(), theta );
    const int64_t l = m->dictLo( m->used );
    *lo = l;
    *var = m->var;
    const int64_t n = std::min<int64_t>( cap, m->used - l );
    for ( int u = 0; u < m->U; u++ )
        for ( int64_t i = 0; i < n; i++ )
            coeffs[ (int64_t)u * cap + i ] = m->c[ ( l + i ) * m->U + u ];
    return m->used;
}

// ---- FP32 restatement (CPU baseline, naive reference schedule) ----
void* cpu_apsm_create( const OrcParams* p, int N ) { return createImpl<float>( p, N ); }
void cpu_apsm_destroy( void* h ) { delete static_cast<Model<float>*>( h ); }
int cpu_apsm_train( void* h, const float* rx, const float* tx, int64_t T )
{
    try
    {
        return trainImpl( *static_cast<Model<float>*>( h ), rx, tx, T, 0 );
    }
    catch ( const std::exception& e )
    {
        g_err = e.what();
        return -1;
    }
}
void cpu_apsm_detect( void* h, const float* rx, int64_t M, double* est )
{
    detectImpl( *static_cast<Model<float>*>( h ), rx, M, est );
}

// ---- LLS ----
int orc_lls_train( const float* rx, const float* tx, int N, int64_t T, int method, double* w )
{
    return method == 0 ? llsCholesky( rx, tx, N, T, w ) : llsQr( rx, tx, N, T, w );
}

void orc_lls_detect( const float* rx, int N, int64_t M, const double* w, double* est )
{
    const int L = 2 * N;
    std::vector<double> a( L ), b( L );
    for ( int64_t k = 0; k < M; k++ )
    {
        embed<double>( rx, N, M, k, a.data(), b.data() );
        est[ 2 * k ] = dot( a.data(), w, L );
        est[ 2 * k + 1 ] = dot( b.data(), w, L );
    }
}

// ---- shuffle permutation and variance heuristic (exposed for tests) ----
void orc_shuffle_perm( int64_t T, int64_t* perm )
{
    std::vector<int64_t> p = shufflePermutation( T, true );
    std::copy( p.begin(), p.end(), perm );
}

float orc_variance_heuristic( const float* rx, int N, int64_t T )
{
    return varianceHeuristic( rx, N, T, shufflePermutation( T, false ) );
}

// ---- modulation and metrics ----
double orc_mod_scale( int bps ) { return modScale( bps ); }

// Gray map of one symbol's bits to a unit-energy point (modulation.cpp:112-151).
void orc_mod_map( int bps, uint32_t bits, double* re, double* im )
{
    auto map1 = []( uint32_t b, uint32_t ns ) {
        uint32_t g = b; // inverse Gray code
        while ( b )
        {
            b >>= 1;
            g ^= b;
        }
        return 2.0 * (int32_t)g - (int32_t)ns + 1;
    };
    double s = modScale( bps );
    if ( bps == 1 )
    {
        *re = map1( bits, 2 ) * s;
        *im = 0;
        return;
    }
    uint32_t ns = 1u << ( bps / 2 );
    *re = map1( bits >> ( bps / 2 ), ns ) * s;
    *im = map1( bits & ( ( 1u << ( bps / 2 ) ) - 1 ), ns ) * s;
}

uint32_t orc_mod_demap( int bps, double re, double im )
{
    double rs = 1.0 / modScale( bps );
    if ( bps == 1 )
        return demap1d( re * rs, 2 );
    uint32_t ns = 1u << ( bps / 2 );
    return ( demap1d( re * rs, ns ) << ( bps / 2 ) ) | demap1d( im * rs, ns );
}

// Symbol errors by decision regions (util.cpp:62-110 with the real-part rule applied to
// both dimensions). cliCompat = 1 reproduces the reference CLI's imaginary-branch clamp.
int64_t orc_symbol_errors( const float* est, const float* ref, int64_t M, int bps, int cliCompat )
{
    const double s = modScale( bps );
    const double mx = ( bps == 1 ) ? 1.0 : ( ( 1 << ( bps / 2 ) ) - 1 ) * s;
    int64_t errs = 0;
    for ( int64_t k = 0; k < M; k++ )
    {
        double yr = est[ 2 * k ], yi = est[ 2 * k + 1 ];
        double tr = ref[ 2 * k ], ti = ref[ 2 * k + 1 ];
        if ( std::isnan( yr ) || std::isnan( yi ) )
        {
            errs++;
            continue;
        }
        bool er = std::abs( yr - tr ) > s && !( yr > mx && tr > mx - s ) && !( yr < -mx && tr < -mx + s );
        bool ei;
        if ( cliCompat )
            ei = std::abs( yi - ti ) > s && !( yi > mx && ti < mx - s ) && !( yi < -mx && ti > -mx + s );
        else
            ei = std::abs( yi - ti ) > s && !( yi > mx && ti > mx - s ) && !( yi < -mx && ti < -mx + s );
        if ( er || ( ei && bps > 1 ) )
            errs++;
    }
    return errs;
}

int64_t orc_bit_errors( const float* est, const float* ref, int64_t M, int bps )
{
    int64_t errs = 0;
    for ( int64_t k = 0; k < M; k++ )
    {
        uint32_t a = orc_mod_demap( bps, est[ 2 * k ], est[ 2 * k + 1 ] );
        uint32_t b = orc_mod_demap( bps, ref[ 2 * k ], ref[ 2 * k + 1 ] );
        errs += __builtin_popcount( a ^ b );
    }
    return errs;
}

double orc_evm( const float* est, const float* ref, int64_t M )
{
    double acc = 0;
    for ( int64_t k = 0; k < M; k++ )
    {
        double dr = (double)est[ 2 * k ] - ref[ 2 * k ], di = (double)est[ 2 * k + 1 ] - ref[ 2 * k + 1 ];
        acc += dr * dr + di * di;
    }
    return std::sqrt( acc / M );
}

} // extern "C"
