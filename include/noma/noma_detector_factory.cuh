// NomaDetectorFactory: "otype" selects the APSM or LLS factory
// (cpp/lib/noma/noma_detector_factory.cuh:70-74, noma_detector_factory.cu:57-65).
#pragma once

#include <memory>
#include <stdexcept>
#include <string>

#include "apsm/apsm_noma_detector.cuh"
#include "libapsm/json_select.hpp"
#include "noma/lls_detector.cuh"
#include "noma/noma_detector.cuh"

class NomaDetectorFactory
{
public:
    static std::unique_ptr<NomaDetector> build( const nlohmann::json& config, uint32_t numAntennas )
    {
        if ( config[ "otype" ] == "APSM" )
            return ApsmNomaDetectorFactory::build( config, numAntennas );
        else if ( config[ "otype" ] == "LLS" )
            return LlsNomaDetectorFactory::build( config, numAntennas );
        else
            throw std::invalid_argument( "Unknown algorithm: " + config[ "otype" ].get<std::string>() );
    }
};
