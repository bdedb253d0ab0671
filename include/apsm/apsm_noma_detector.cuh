// ApsmNomaDetector: the APSM multi-user detector (cpp/lib/apsm/apsm_noma_detector.cuh:83-115)
// over the C ABI. Train and detect run as hand-written gfx950 kernels inside libapsm_amd.so.
//
// Beyond the reference, one detector may carry several users (numUsers): they train on the
// same received pilots, so they share the Gaussian dictionary and every kernel evaluation;
// use the overloads that take [user][sample] pilots and return [user][sample] estimates.
#pragma once

#include <memory>
#include <stdexcept>
#include <string>
#include <utility>

#include "apsm/apsm_parameters.cuh"
#include "apsm/apsm_versions.cuh"
#include "apsm_capi.h"
#include "cuda/cuda_stream.cuh"
#include "libapsm/capi_check.hpp"
#include "libapsm/json_select.hpp"
#include "noma/noma_detector.cuh"

// Snapshot of the device training state (the reference keeps HostTrainingState on the device,
// cpp/lib/cuda/cuda_matrix_host.cuh:476-529); refreshed by ApsmNomaDetector::getTrainingState().
class HostTrainingState
{
public:
    uint64_t usedHeight = 0; ///< real rows that entered the pilot stream
    uint32_t dictionaryEntries = 0; ///< complex basis vectors in the dictionary
    RealSample gaussianKernelVariance = 0; ///< variance after scaling by N^2 (or the heuristic)
    RealSampleMatrix linearCoeffs; ///< theta, [user][2N]
    RealSampleMatrix gaussianCoeffs; ///< real-row coefficients, [user][2 * dictionaryEntries]
    ComplexSampleMatrix basis; ///< dictionary vectors, [antenna][dictionaryEntries], oldest first
};

// Everything of the detector that does not depend on the detectVersion template argument;
// ApsmNomaDetector<version_id> adds only the version.
class ApsmDetectorBase : public NomaDetector
{
public:
    ApsmDetectorBase( const apsm_parameters& par, uint32_t numAntennas, uint32_t numUsers, int32_t version )
        : par( par )
        , numAntennas( numAntennas )
        , numUsers( numUsers )
    {
        const apsm_params p = par.toC( version );
        libapsm::check( apsm_create( &p, numAntennas, numUsers, &handle ), "ApsmNomaDetector" );
    }
    ~ApsmDetectorBase() { apsm_destroy( handle ); }
    ApsmDetectorBase( const ApsmDetectorBase& ) = delete;
    ApsmDetectorBase& operator=( const ApsmDetectorBase& ) = delete;

    float train( const ComplexSampleMatrix& rxSigTraining, const ComplexSampleVector& txSigTraining, NomaValSet valSet = std::nullopt ) override
    {
        return train( rxSigTraining, ComplexSampleMatrix { txSigTraining }, std::move( valSet ), nullptr );
    }

    // Several users at once: txSigTraining[user][sample]; a validation set records user 0.
    float train( const ComplexSampleMatrix& rxSigTraining, const ComplexSampleMatrix& txSigTraining, NomaValSet valSet = std::nullopt,
                 const ComplexSampleMatrix* txValPerUser = nullptr )
    {
        size_t T = 0;
        std::vector<float> rx = libapsm::detail::flatten( rxSigTraining, T );
        checkAntennas( rxSigTraining );
        std::vector<float> tx = flattenUsers( txSigTraining, T );
        float ms = 0.f;
        if ( !valSet.has_value() )
        {
            libapsm::check( apsm_train( handle, rx.data(), tx.data(), (uint32_t)T, &ms ), "ApsmNomaDetector::train" );
            return ms;
        }
        auto& [ history, rxVal, txVal ] = *valSet;
        size_t Tv = 0;
        std::vector<float> rv = libapsm::detail::flatten( rxVal, Tv );
        checkAntennas( rxVal );
        std::vector<float> tv;
        if ( txValPerUser )
            tv = flattenUsers( *txValPerUser, Tv );
        else
        {
            if ( numUsers != 1 )
                throw std::invalid_argument( "a validation set for several users needs their validation symbols" );
            tv = flattenUsers( ComplexSampleMatrix { txVal }, Tv );
        }
        std::vector<float> hist( ( T + 1 ) * numUsers * 6 );
        uint32_t steps = 0;
        libapsm::check( apsm_train_history( handle, rx.data(), tx.data(), (uint32_t)T, rv.data(), tv.data(), (uint32_t)Tv,
                                            history.getBitPerSymbol(), history.getModulationScale(), hist.data(),
                                            (uint32_t)( T + 1 ), &steps, &ms ),
                        "ApsmNomaDetector::train" );
        for ( uint32_t s = 0; s < steps; s++ )
        {
            const float* r = &hist[ (size_t)s * numUsers * 6 ]; // user 0
            history.evmTrain.push_back( r[ 0 ] );
            history.serTrain.push_back( r[ 1 ] );
            history.berTrain.push_back( r[ 2 ] );
            history.evmVal.push_back( r[ 3 ] );
            history.serVal.push_back( r[ 4 ] );
            history.berVal.push_back( r[ 5 ] );
        }
        return ms;
    }

    float detect( const ComplexSampleMatrix& rxSigData, ComplexSampleVector& estData ) override
    {
        if ( numUsers != 1 )
            throw std::invalid_argument( "this detector carries several users: use detect(rx, ComplexSampleMatrix&)" );
        size_t M = 0;
        std::vector<float> rx = libapsm::detail::flatten( rxSigData, M );
        checkAntennas( rxSigData );
        estData.resize( M );
        float ms = 0.f;
        libapsm::check( apsm_detect( handle, rx.data(), (uint32_t)M, reinterpret_cast<float*>( estData.data() ), &ms ),
                        "ApsmNomaDetector::detect" );
        return ms;
    }

    // estData[user][sample]
    float detect( const ComplexSampleMatrix& rxSigData, ComplexSampleMatrix& estData )
    {
        size_t M = 0;
        std::vector<float> rx = libapsm::detail::flatten( rxSigData, M );
        checkAntennas( rxSigData );
        std::vector<float> est( (size_t)numUsers * M * 2 );
        float ms = 0.f;
        libapsm::check( apsm_detect( handle, rx.data(), (uint32_t)M, est.data(), &ms ), "ApsmNomaDetector::detect" );
        estData.assign( numUsers, ComplexSampleVector( M ) );
        for ( uint32_t u = 0; u < numUsers; u++ )
            std::copy( est.begin() + (size_t)u * M * 2, est.begin() + (size_t)( u + 1 ) * M * 2,
                       reinterpret_cast<float*>( estData[ u ].data() ) );
        return ms;
    }

    void resetState() override { libapsm::check( apsm_reset( handle ), "ApsmNomaDetector::resetState" ); }

    // (w_L |theta|^2) / (w_G c^T K c) of user 0 (apsm_noma_detector.cu:220-319)
    // Detect arithmetic for the following detects (APSM_PRECISION_*); keeps the trained state.
    void setDetectPrecision( int32_t precision )
    {
        libapsm::check( apsm_set_detect_precision( handle, precision ), "ApsmNomaDetector::setDetectPrecision" );
        par.detectPrecision = precision;
    }

    RealSample computeNormRatio( uint32_t user = 0 )
    {
        float r = 0.f;
        libapsm::check( apsm_norm_ratio( handle, user, &r ), "ApsmNomaDetector::computeNormRatio" );
        return r;
    }

    HostTrainingState* getTrainingState()
    {
        HostTrainingState& s = state;
        uint64_t used = 0;
        uint32_t entries = 0;
        float var = 0.f;
        libapsm::check( apsm_get_state( handle, &used, &entries, &var, nullptr, nullptr, 0 ), "getTrainingState" );
        std::vector<float> theta( (size_t)numUsers * 2 * numAntennas ), coeffs( (size_t)numUsers * 2 * std::max( entries, 1u ) );
        libapsm::check( apsm_get_state( handle, nullptr, nullptr, nullptr, theta.data(), coeffs.data(), 2 * std::max( entries, 1u ) ),
                        "getTrainingState" );
        std::vector<float> dict( (size_t)numAntennas * 2 * std::max( entries, 1u ) );
        libapsm::check( apsm_get_dictionary( handle, dict.data(), std::max( entries, 1u ), &entries ), "getTrainingState" );
        s.usedHeight = used;
        s.dictionaryEntries = entries;
        s.gaussianKernelVariance = var;
        s.linearCoeffs.assign( numUsers, RealSampleVector() );
        s.gaussianCoeffs.assign( numUsers, RealSampleVector() );
        for ( uint32_t u = 0; u < numUsers; u++ )
        {
            s.linearCoeffs[ u ].assign( theta.begin() + (size_t)u * 2 * numAntennas, theta.begin() + (size_t)( u + 1 ) * 2 * numAntennas );
            s.gaussianCoeffs[ u ].assign( coeffs.begin() + (size_t)u * 2 * entries, coeffs.begin() + (size_t)u * 2 * entries + 2 * entries );
        }
        s.basis.assign( numAntennas, ComplexSampleVector( entries ) );
        for ( uint32_t a = 0; a < numAntennas; a++ )
            for ( uint32_t k = 0; k < entries; k++ )
                s.basis[ a ][ k ] = ComplexSample( dict[ ( (size_t)a * entries + k ) * 2 ], dict[ ( (size_t)a * entries + k ) * 2 + 1 ] );
        return &s;
    }

    const apsm_parameters& parameters() const { return par; }
    uint32_t getNumUsers() const { return numUsers; }
    apsm_handle* nativeHandle() const { return handle; }

private:
    void checkAntennas( const ComplexSampleMatrix& m ) const
    {
        if ( m.size() != numAntennas )
            throw std::invalid_argument( "expected " + std::to_string( numAntennas ) + " antennas, got " + std::to_string( m.size() ) );
    }
    std::vector<float> flattenUsers( const ComplexSampleMatrix& tx, size_t T ) const
    {
        if ( tx.size() != numUsers )
            throw std::invalid_argument( "expected pilots of " + std::to_string( numUsers ) + " users, got " + std::to_string( tx.size() ) );
        std::vector<float> out( (size_t)numUsers * T * 2 );
        for ( uint32_t u = 0; u < numUsers; u++ )
        {
            if ( tx[ u ].size() != T )
                throw std::invalid_argument( "pilot count differs from the received training samples" );
            const float* src = reinterpret_cast<const float*>( tx[ u ].data() );
            std::copy( src, src + 2 * T, out.begin() + (size_t)u * T * 2 );
        }
        return out;
    }

    apsm_parameters par;
    uint32_t numAntennas;
    uint32_t numUsers;
    apsm_handle* handle = nullptr;
    HostTrainingState state;
};

template <int32_t version_id>
class ApsmNomaDetector : public ApsmDetectorBase
{
public:
    ApsmNomaDetector( const apsm_parameters& par, uint32_t numAntennas, uint32_t numUsers = 1 )
        : ApsmDetectorBase( par, numAntennas, numUsers, version_id )
    {
    }
};

class ApsmNomaDetectorFactory
{
public:
    // JSON -> apsm_parameters with the reference's key whitelist and defaults
    // (apsm_noma_detector.cu:348-383).
    static apsm_parameters parameters( const nlohmann::json& config )
    {
        static const char* const validKeys[] = { "otype", "llsInitialization", "trainingShuffle", "gaussianKernelWeight",
                                                 "gaussianKernelVariance", "windowSize", "sampleStep", "trainPasses",
                                                 "startWithFullWindow", "dictionarySize", "normConstraint", "eB", "detectVersion",
                                                 "detectPrecision" };
        for ( auto& item : config.items() )
        {
            bool ok = false;
            for ( const char* k : validKeys )
                ok = ok || item.key() == k;
            if ( !ok )
                throw std::invalid_argument( "Invalid key: " + item.key() );
        }
        apsm_parameters par;
        par.llsInitialization = config.value( "llsInitialization", false );
        par.trainingShuffle = config.value( "trainingShuffle", false );
        par.gaussianKernelWeight = config.value( "gaussianKernelWeight", 0.5 );
        par.linearKernelWeight = 1 - par.gaussianKernelWeight;
        par.gaussianKernelVariance = config.value( "gaussianKernelVariance", 0.05 );
        par.windowSize = config.value( "windowSize", 20 );
        par.sampleStep = config.value( "sampleStep", 2 );
        par.trainPasses = config.value( "trainPasses", 1 );
        par.startWithFullWindow = config.value( "startWithFullWindow", false );
        par.dictionarySize = config.value( "dictionarySize", 685 );
        par.normConstraint = config.value( "normConstraint", 0. );
        par.eB = config.value( "eB", 0.001 );
        par.detectPrecision = precision( config );
        return par;
    }

    // "detectPrecision": arithmetic of the detect cross terms (no counterpart in the reference)
    static int32_t precision( const nlohmann::json& config )
    {
        const std::string v = config.value( "detectPrecision", "fp32" );
        if ( v == "fp32" )
            return APSM_PRECISION_FP32;
        if ( v == "fp32-vector" )
            return APSM_PRECISION_FP32_VECTOR;
        if ( v == "fp32-matrix" )
            return APSM_PRECISION_FP32_MATRIX;
        if ( v == "bf16" )
            return APSM_PRECISION_BF16;
        if ( v == "fp16" )
            return APSM_PRECISION_FP16;
        throw std::invalid_argument( "Unknown detect precision: " + v );
    }

    static int32_t version( const nlohmann::json& config )
    {
        const std::string v = config.value( "detectVersion", "balanced" );
        if ( v == "original" )
            return apsm_versions::APSM_DETECT_ORIGINAL;
        if ( v == "oldfast" )
            return apsm_versions::APSM_DETECT_OLDFAST;
        if ( v == "shmem" )
            return apsm_versions::APSM_DETECT_SHMEM;
        if ( v == "balanced" )
            return apsm_versions::APSM_DETECT_BALANCED;
        throw std::invalid_argument( "Unknown detect version: " + v );
    }

    static std::unique_ptr<NomaDetector> build( const nlohmann::json& config, uint32_t numAntennas, uint32_t numUsers = 1 )
    {
        const apsm_parameters par = parameters( config );
        switch ( version( config ) )
        {
        case apsm_versions::APSM_DETECT_ORIGINAL: return std::make_unique<ApsmNomaDetector<apsm_versions::APSM_DETECT_ORIGINAL>>( par, numAntennas, numUsers );
        case apsm_versions::APSM_DETECT_OLDFAST: return std::make_unique<ApsmNomaDetector<apsm_versions::APSM_DETECT_OLDFAST>>( par, numAntennas, numUsers );
        case apsm_versions::APSM_DETECT_SHMEM: return std::make_unique<ApsmNomaDetector<apsm_versions::APSM_DETECT_SHMEM>>( par, numAntennas, numUsers );
        default: return std::make_unique<ApsmNomaDetector<apsm_versions::APSM_DETECT_BALANCED>>( par, numAntennas, numUsers );
        }
    }
};
