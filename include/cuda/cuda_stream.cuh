// CudaStream: RAII owner of one device stream, the class cpp/lib/cuda/cuda_stream.cuh:67-96
// declares, backed by a HIP stream of default priority created through the C ABI.
#pragma once

#include "apsm_capi.h"
#include "libapsm/capi_check.hpp"

// Same declaration as in <hip/hip_runtime_api.h>; repeating an identical typedef is valid C++.
typedef struct ihipStream_t* hipStream_t;

class CudaStream
{
public:
    CudaStream() { libapsm::check( apsm_stream_create( &stream ), "CudaStream" ); }
    ~CudaStream() { (void)apsm_stream_destroy( stream ); }

    // object can't be copied
    CudaStream( const CudaStream& ) = delete;
    CudaStream& operator=( const CudaStream& ) = delete;

    operator hipStream_t() const { return static_cast<hipStream_t>( stream ); }
    void* get() const { return stream; }

    void synchronize() const { libapsm::check( apsm_stream_synchronize( stream ), "CudaStream::synchronize" ); }

private:
    void* stream = nullptr;
};
