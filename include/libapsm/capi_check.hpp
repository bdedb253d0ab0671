// Status-code to exception mapping for the C++ classes over the C ABI.
// Configuration errors throw std::invalid_argument as the reference factories do
// (cpp/lib/apsm/apsm_noma_detector.cu:348-383); device and numeric failures throw
// std::runtime_error (the reference prints and exits, cpp/lib/cuda/cuda_errorhandling.cuh:98-108;
// the NOMA_detect tool restores that by catching and exiting with status 1).
#pragma once

#include <stdexcept>
#include <string>

#include "apsm_capi.h"

namespace libapsm {

inline void check( int rc, const char* what )
{
    if ( rc == APSM_OK )
        return;
    std::string msg = std::string( what ) + ": " + apsm_last_error();
    if ( rc == APSM_ERR_ARG || rc == APSM_ERR_UNSUPPORTED )
        throw std::invalid_argument( msg );
    throw std::runtime_error( msg );
}

} // namespace libapsm
