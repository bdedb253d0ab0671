// Sample types of the detector API, as cpp/lib/noma/types.cuh:69-100 declares them:
// RealSample is float; complex matrices are indexed [antenna][sample].
#pragma once

#include <complex>
#include <cstdint>
#include <iostream>
#include <type_traits>
#include <vector>

typedef float RealSample;
typedef std::vector<RealSample> RealSampleVector;
typedef std::vector<RealSampleVector> RealSampleMatrix;

typedef std::complex<RealSample> ComplexSample;
typedef std::vector<ComplexSample> ComplexSampleVector;
typedef std::vector<ComplexSampleVector> ComplexSampleMatrix;

static_assert( std::is_same<RealSample, float>::value, "Only float data is supported" );

namespace libapsm {
namespace detail {
template <typename V>
inline void printRow( std::ostream& os, const V& v )
{
    os << "  [";
    for ( size_t i = 0; i < v.size(); i++ )
        os << v[ i ] << ( i + 1 < v.size() ? ", " : "" );
    os << "]" << std::endl;
}
template <typename M>
inline void printMatrix( std::ostream& os, const M& m )
{
    os << "[" << std::endl;
    for ( const auto& row : m )
        printRow( os, row );
    os << "]" << std::endl;
}
} // namespace detail
} // namespace libapsm

// Printers with the reference's text layout (cpp/lib/noma/types.cu:69-150).
inline std::ostream& operator<<( std::ostream& os, const RealSampleVector& m )
{
    libapsm::detail::printRow( os, m );
    return os;
}
inline std::ostream& operator<<( std::ostream& os, const RealSampleMatrix& m )
{
    libapsm::detail::printMatrix( os, m );
    return os;
}
inline std::ostream& operator<<( std::ostream& os, const ComplexSampleVector& m )
{
    libapsm::detail::printRow( os, m );
    return os;
}
inline std::ostream& operator<<( std::ostream& os, const ComplexSampleMatrix& m )
{
    libapsm::detail::printMatrix( os, m );
    return os;
}
