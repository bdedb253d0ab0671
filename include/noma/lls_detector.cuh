// LlsNomaDetector: least-squares linear detector (cpp/lib/noma/lls_detector.cuh:75-105) over the
// C ABI. The normal equations are accumulated on the device in FP32 and solved by an FP64
// Cholesky on the device; detect is est = A w.
#pragma once

#include <algorithm>
#include <initializer_list>
#include <memory>
#include <stdexcept>
#include <string>

#include "apsm_capi.h"
#include "cuda/cuda_stream.cuh"
#include "libapsm/capi_check.hpp"
#include "libapsm/json_select.hpp"
#include "noma/noma_detector.cuh"

class LlsNomaDetector : public NomaDetector
{
public:
    LlsNomaDetector( uint32_t numAntennas, std::shared_ptr<CudaStream> extStream = {} )
        : numAntennas( numAntennas )
        , stream( extStream )
    {
        if ( stream )
            libapsm::check( lls_create_on_stream( numAntennas, stream->get(), &handle ), "LlsNomaDetector" );
        else
            libapsm::check( lls_create( numAntennas, &handle ), "LlsNomaDetector" );
    }
    ~LlsNomaDetector() { lls_destroy( handle ); }
    LlsNomaDetector( const LlsNomaDetector& ) = delete;
    LlsNomaDetector& operator=( const LlsNomaDetector& ) = delete;

    float train( const ComplexSampleMatrix& rxSigTraining, const ComplexSampleVector& txSigTraining, NomaValSet = std::nullopt ) override
    {
        size_t T = 0;
        std::vector<float> rx = libapsm::detail::flatten( rxSigTraining, T );
        checkAntennas( rxSigTraining );
        if ( txSigTraining.size() != T )
            throw std::invalid_argument( "pilot count differs from the received training samples" );
        std::vector<float> tx = libapsm::detail::flatten( txSigTraining );
        float ms = 0.f;
        libapsm::check( lls_train( handle, rx.data(), tx.data(), (uint32_t)T, &ms ), "LlsNomaDetector::train" );
        return ms;
    }

    // Least squares over a device column-major A (numSamples x numInputDims) and device target
    // y; A and y are left unchanged (lls_detector.cu:157-179).
    void trainLowLevel( const RealSample* rxSigMat, const RealSample* txSigMat, const int32_t numInputDims, const int32_t numOutputDims,
                        const int32_t numSamples )
    {
        libapsm::check( lls_train_lowlevel( handle, rxSigMat, txSigMat, numInputDims, numOutputDims, numSamples ),
                        "LlsNomaDetector::trainLowLevel" );
    }

    float detect( const ComplexSampleMatrix& rxSigData, ComplexSampleVector& estData ) override
    {
        size_t M = 0;
        std::vector<float> rx = libapsm::detail::flatten( rxSigData, M );
        checkAntennas( rxSigData );
        estData.resize( M );
        float ms = 0.f;
        libapsm::check( lls_detect( handle, rx.data(), (uint32_t)M, reinterpret_cast<float*>( estData.data() ), &ms ),
                        "LlsNomaDetector::detect" );
        return ms;
    }

    // est = alpha * A w + beta * est on device memory (lls_detector.cu:207-218)
    void detectLowLevel( const RealSample* rxSigMat, RealSample* estMat, const int32_t numInputDims, const int32_t numOutputDims,
                         const int32_t numSamples, const float alpha = 1, const float beta = 0 )
    {
        libapsm::check( lls_detect_lowlevel( handle, rxSigMat, estMat, numInputDims, numOutputDims, numSamples, alpha, beta ),
                        "LlsNomaDetector::detectLowLevel" );
    }

    void resetState() override {} // the next train replaces the whole state (lls_detector.cu:220-223)

    float* getWeightMat() { return lls_weights_dev( handle ); }

    float computeSignalPower( const RealSample* data, const uint32_t numValues )
    {
        float p = 0.f;
        libapsm::check( lls_signal_power( handle, data, numValues, &p ), "LlsNomaDetector::computeSignalPower" );
        return p;
    }

    // host copy of the weights (2 * numAntennas values after train)
    RealSampleVector getWeights()
    {
        RealSampleVector w( 128 );
        libapsm::check( lls_get_weights( handle, w.data() ), "LlsNomaDetector::getWeights" );
        w.resize( 2 * numAntennas );
        return w;
    }

private:
    void checkAntennas( const ComplexSampleMatrix& m ) const
    {
        if ( m.size() != numAntennas )
            throw std::invalid_argument( "expected " + std::to_string( numAntennas ) + " antennas, got " + std::to_string( m.size() ) );
    }

    uint32_t numAntennas;
    std::shared_ptr<CudaStream> stream;
    lls_handle* handle = nullptr;
};

class LlsNomaDetectorFactory
{
public:
    // Only "otype" is accepted (lls_detector.cu:244-254).
    static std::unique_ptr<NomaDetector> build( const nlohmann::json& config, uint32_t numAntennas )
    {
        for ( auto& item : config.items() )
            if ( item.key() != "otype" )
                throw std::invalid_argument( "Invalid key: " + item.key() );
        return std::make_unique<LlsNomaDetector>( numAntennas );
    }
};
