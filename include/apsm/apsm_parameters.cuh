// apsm_parameters: the detector's parameter set with the fields of
// cpp/lib/apsm/apsm_parameters.cuh:68-86 and its printer (cpp/lib/apsm/apsm_parameters.cu:66-87).
#pragma once

#include <iomanip>
#include <iostream>

#include "apsm_capi.h"
#include "noma/types.cuh"

class apsm_parameters
{
public:
    // APSM kernel parameters
    RealSample linearKernelWeight; ///< Linear Kernel Weight
    RealSample gaussianKernelWeight; ///< Gaussian Kernel Weight
    RealSample gaussianKernelVariance; ///< Gaussian Kernel Variance (scaled by N^2 by the detector)

    // training parameters
    uint32_t windowSize; ///< number of past (complex) samples to reuse
    uint32_t sampleStep; ///< real rows entering per train step
    uint32_t trainPasses = 1; ///< number of training passes over the training data
    bool startWithFullWindow = false; ///< if true skip the smaller windows at the beginning of training
    bool llsInitialization = false; ///< initialise the linear weights with the LLS solution
    bool trainingShuffle = false; ///< shuffle the training data
    uint32_t dictionarySize; ///< maximum number of (complex) basis vectors kept in the dictionary
    RealSample normConstraint; ///< norm constraint for dictionary sparsification (0 disables)
    RealSample eB; ///< hyperslab width

    // detect arithmetic (extension: APSM_PRECISION_* of the C ABI, key "detectPrecision")
    int32_t detectPrecision = APSM_PRECISION_FP32;

    // The C ABI parameter block for a given detectVersion.
    apsm_params toC( int32_t detectVersion ) const
    {
        apsm_params p;
        apsm_params_default( &p );
        p.gaussian_kernel_weight = gaussianKernelWeight;
        p.linear_kernel_weight = linearKernelWeight;
        p.gaussian_kernel_variance = gaussianKernelVariance;
        p.eB = eB;
        p.norm_constraint = normConstraint;
        p.window_size = windowSize;
        p.sample_step = sampleStep;
        p.train_passes = trainPasses;
        p.dictionary_size = dictionarySize;
        p.start_with_full_window = startWithFullWindow ? 1 : 0;
        p.lls_initialization = llsInitialization ? 1 : 0;
        p.training_shuffle = trainingShuffle ? 1 : 0;
        p.detect_version = detectVersion;
        p.detect_precision = detectPrecision;
        return p;
    }
};

inline std::ostream& operator<<( std::ostream& os, const apsm_parameters& par )
{
    os << "{" << std::endl << std::endl;
    os << "    APSM kernel parameters" << std::endl;
    os << "    ----------------------" << std::endl;
    os << "    par.linearKernelWeight = " << par.linearKernelWeight << std::endl;
    os << "    par.gaussianKernelWeight = " << par.gaussianKernelWeight << std::endl;
    os << "    par.gaussianKernelVariance = " << par.gaussianKernelVariance << std::endl << std::endl;
    os << "    training parameters" << std::endl;
    os << "    -------------------" << std::endl;
    os << "    par.windowSize = " << par.windowSize << std::endl;
    os << "    par.sampleStep = " << par.sampleStep << std::endl;
    os << "    par.trainPasses = " << par.trainPasses << std::endl;
    os << "    par.dictionarySize = " << par.dictionarySize << std::endl;
    os << "    par.normConstraint = " << par.normConstraint << std::endl;
    os << "    par.eB = " << std::setprecision( 4 ) << par.eB << std::endl;
    os << "}";
    return os;
}
