// NomaTrainHistory: per-train-step EVM / SER / BER on the training and validation sets,
// the class of cpp/lib/noma/noma_train_history.cuh:64-99 with the metric rules of
// cpp/lib/noma/noma_train_history.cu:105-236. Here the detector computes the per-step values
// on the device (apsm_train_history) and appends them; addTrainStep / addValStep take host
// arrays of interleaved (Re, Im) values and compute the same metrics on the host.
#pragma once

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <fstream>
#include <optional>
#include <string>
#include <tuple>

#include "noma/types.cuh"

class NomaTrainHistory
{
public:
    NomaTrainHistory( uint32_t bitPerSymbol, float modulationScale )
        : bitPerSymbol( bitPerSymbol )
        , modulationScale( modulationScale )
    {
    }

    NomaTrainHistory() = delete;
    NomaTrainHistory( const NomaTrainHistory& ) = delete;
    NomaTrainHistory& operator=( const NomaTrainHistory& ) = delete;

    void addTrainStep( const RealSample* y_true, const RealSample* y_hat, uint32_t numElements )
    {
        evmTrain.push_back( computeEvm( y_true, y_hat, numElements ) );
        serTrain.push_back( computeSer( y_true, y_hat, numElements ) );
        berTrain.push_back( computeBer( y_true, y_hat, numElements ) );
    }
    void addValStep( const RealSample* y_true, const RealSample* y_hat, uint32_t numElements )
    {
        evmVal.push_back( computeEvm( y_true, y_hat, numElements ) );
        serVal.push_back( computeSer( y_true, y_hat, numElements ) );
        berVal.push_back( computeBer( y_true, y_hat, numElements ) );
    }

    void toCsv( std::string filename )
    {
        std::ofstream file( filename, std::ios::out | std::ios::trunc );
        file << "step,train_evm,train_ser,train_ber,val_evm,val_ser,val_ber" << std::endl;
        for ( size_t i = 0; i < evmTrain.size(); i++ )
        {
            file << i << "," << evmTrain[ i ] << "," << serTrain[ i ] << "," << berTrain[ i ] << ",";
            file << ( i < evmVal.size() ? evmVal[ i ] : 0.f ) << "," << ( i < serVal.size() ? serVal[ i ] : 0.f ) << ","
                 << ( i < berVal.size() ? berVal[ i ] : 0.f ) << std::endl;
        }
    }

    uint32_t getBitPerSymbol() const { return bitPerSymbol; }
    float getModulationScale() const { return modulationScale; }

    RealSampleVector evmTrain;
    RealSampleVector evmVal;
    RealSampleVector serTrain;
    RealSampleVector serVal;
    RealSampleVector berTrain;
    RealSampleVector berVal;

private:
    // root mean squared complex error over numElements / 2 samples
    RealSample computeEvm( const RealSample* y_true, const RealSample* y_hat, uint32_t numElements ) const
    {
        double s = 0.0;
        for ( uint32_t i = 0; i < numElements / 2; i++ )
        {
            const double dr = y_true[ 2 * i ] - y_hat[ 2 * i ], di = y_true[ 2 * i + 1 ] - y_hat[ 2 * i + 1 ];
            s += dr * dr + di * di;
        }
        return (RealSample)std::sqrt( s / ( numElements / 2 ) );
    }
    // decision-region symbol error per component: wrong region unless both lie beyond the
    // outermost constellation level on the same side
    static bool componentError( RealSample t, RealSample h, RealSample scale, RealSample maxVal )
    {
        return std::fabs( t - h ) > scale && !( h > maxVal && t > maxVal - scale ) && !( h < -maxVal && t < -maxVal + scale );
    }
    RealSample computeSer( const RealSample* y_true, const RealSample* y_hat, uint32_t numElements ) const
    {
        uint64_t errors = 0;
        const RealSample maxVal = ( ( 1u << ( bitPerSymbol / 2 ) ) - 1 ) * modulationScale;
        for ( uint32_t i = 0; i < numElements / 2; i++ )
        {
            if ( bitPerSymbol == 1 )
                errors += std::signbit( y_true[ 2 * i ] ) != std::signbit( y_hat[ 2 * i ] );
            else
                errors += componentError( y_true[ 2 * i ], y_hat[ 2 * i ], modulationScale, maxVal )
                    || componentError( y_true[ 2 * i + 1 ], y_hat[ 2 * i + 1 ], modulationScale, maxVal );
        }
        return (RealSample)errors / ( numElements / 2 );
    }
    // Gray index of the nearest PAM level of x (levels ... -3 -1 1 3 ... in units of the scale)
    static uint32_t grayLevel( RealSample x, uint32_t levels )
    {
        int32_t xi = (int32_t)std::round( ( x + levels - 1 ) / 2 );
        xi = std::min<int32_t>( (int32_t)levels - 1, std::max<int32_t>( 0, xi ) );
        return (uint32_t)( xi ^ ( xi >> 1 ) );
    }
    RealSample computeBer( const RealSample* y_true, const RealSample* y_hat, uint32_t numElements ) const
    {
        const uint32_t levels = bitPerSymbol == 1 ? 2 : 1u << ( bitPerSymbol / 2 );
        const RealSample rev = 1.f / modulationScale;
        uint64_t errors = 0;
        for ( uint32_t i = 0; i < numElements / 2; i++ )
        {
            errors += __builtin_popcount( grayLevel( y_true[ 2 * i ] * rev, levels ) ^ grayLevel( y_hat[ 2 * i ] * rev, levels ) );
            if ( bitPerSymbol > 1 )
                errors += __builtin_popcount( grayLevel( y_true[ 2 * i + 1 ] * rev, levels )
                                              ^ grayLevel( y_hat[ 2 * i + 1 ] * rev, levels ) );
        }
        return (RealSample)errors / ( (RealSample)numElements * bitPerSymbol / 2 );
    }

    const uint32_t bitPerSymbol;
    const float modulationScale;
};

typedef std::optional<std::tuple<NomaTrainHistory&, const ComplexSampleMatrix&, const ComplexSampleVector&>> NomaValSet;
