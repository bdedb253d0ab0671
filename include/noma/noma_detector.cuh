// NomaDetector: the abstract multi-user detector of cpp/lib/noma/noma_detector.cuh:65-75.
// train / detect return the device time in milliseconds, as the reference's event timer does.
#pragma once

#include <algorithm>
#include <random>
#include <stdexcept>
#include <vector>

#include "noma/noma_train_history.cuh"
#include "noma/types.cuh"

class NomaDetector
{
public:
    virtual ~NomaDetector() {}

    virtual float train( const ComplexSampleMatrix& rxSigTraining, const ComplexSampleVector& txSigTraining, NomaValSet valSet = std::nullopt ) = 0;
    virtual float detect( const ComplexSampleMatrix& rxSigData, ComplexSampleVector& estData ) = 0;
    virtual void resetState() = 0;

protected:
    // The same permutation for every antenna stream and the pilots: std::shuffle driven by a
    // std::mt19937 seeded with seed_seq {1..8}, one fresh copy of the engine per vector
    // (cpp/lib/noma/noma_detector.cu:59-74).
    void shuffleData( ComplexSampleMatrix& rxSigTraining, ComplexSampleVector& txSigTraining )
    {
        std::seed_seq seed { 1, 2, 3, 4, 5, 6, 7, 8 };
        const std::mt19937 engine( seed );
        for ( auto& row : rxSigTraining )
        {
            std::mt19937 e = engine;
            std::shuffle( row.begin(), row.end(), e );
        }
        std::mt19937 e = engine;
        std::shuffle( txSigTraining.begin(), txSigTraining.end(), e );
    }
};

namespace libapsm {
namespace detail {

// [antenna][sample] complex vectors -> contiguous float [antenna][sample][2]
inline std::vector<float> flatten( const ComplexSampleMatrix& m, size_t& samples )
{
    if ( m.empty() )
        throw std::invalid_argument( "received signal has no antennas" );
    samples = m[ 0 ].size();
    std::vector<float> out( m.size() * samples * 2 );
    for ( size_t a = 0; a < m.size(); a++ )
    {
        if ( m[ a ].size() != samples )
            throw std::invalid_argument( "all antennas must carry the same number of samples" );
        const float* src = reinterpret_cast<const float*>( m[ a ].data() );
        std::copy( src, src + 2 * samples, out.begin() + a * samples * 2 );
    }
    return out;
}

inline std::vector<float> flatten( const ComplexSampleVector& v )
{
    const float* src = reinterpret_cast<const float*>( v.data() );
    return std::vector<float>( src, src + 2 * v.size() );
}

} // namespace detail
} // namespace libapsm
