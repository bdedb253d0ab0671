// Host-only checks of the C++ drop-in headers and the NOMA_detect helpers (no device calls).
// Usage: api_check <command> [args]; prints results as JSON on stdout.
//   factory <json>                       apsm_parameters / version / errors of the factories
//   history <bps> <scale> <true.f32> <hat.f32>   NomaTrainHistory metrics of raw float32 files
//   shuffle <T>                          permutation NomaDetector::shuffleData applies
//   refdata <seed> <bps> <users> <ntrain> <ndata>  reference symbols of the -rs generator
//   modulation <bps>                     scale, and encode -> hardDecode round trip of all symbols
#include <cstdio>
#include <fstream>
#include <iostream>
#include <iterator>
#include <string>
#include <vector>

#include "../../libapsm_amd/cli/cli_common.hpp"
#include "noma/noma_detector_factory.cuh"

namespace {

std::vector<float> readF32( const char* path )
{
    std::ifstream f( path, std::ios::binary );
    std::vector<char> b( ( std::istreambuf_iterator<char>( f ) ), std::istreambuf_iterator<char>() );
    std::vector<float> v( b.size() / 4 );
    std::memcpy( v.data(), b.data(), v.size() * 4 );
    return v;
}

class ShuffleProbe : public NomaDetector
{
public:
    float train( const ComplexSampleMatrix&, const ComplexSampleVector&, NomaValSet ) override { return 0; }
    float detect( const ComplexSampleMatrix&, ComplexSampleVector& ) override { return 0; }
    void resetState() override {}
    void run( ComplexSampleMatrix& rx, ComplexSampleVector& tx ) { shuffleData( rx, tx ); }
};

} // namespace

int main( int argc, char** argv )
{
    const std::string cmd = argc > 1 ? argv[ 1 ] : "";
    std::cout.precision( 9 );
    if ( cmd == "factory" )
    {
        try
        {
            const auto cfg = nlohmann::json::parse( std::string( argv[ 2 ] ) );
            const apsm_parameters p = ApsmNomaDetectorFactory::parameters( cfg );
            const int32_t v = ApsmNomaDetectorFactory::version( cfg );
            std::cout << "{\"linearKernelWeight\": " << p.linearKernelWeight << ", \"gaussianKernelWeight\": " << p.gaussianKernelWeight
                      << ", \"gaussianKernelVariance\": " << p.gaussianKernelVariance << ", \"windowSize\": " << p.windowSize
                      << ", \"sampleStep\": " << p.sampleStep << ", \"trainPasses\": " << p.trainPasses
                      << ", \"startWithFullWindow\": " << p.startWithFullWindow << ", \"llsInitialization\": " << p.llsInitialization
                      << ", \"trainingShuffle\": " << p.trainingShuffle << ", \"dictionarySize\": " << p.dictionarySize
                      << ", \"normConstraint\": " << p.normConstraint << ", \"eB\": " << p.eB << ", \"version\": " << v
                      << ", \"detectPrecision\": " << p.detectPrecision
                      << ", \"versionString\": \"" << apsm_get_version_string( v ) << "\"}" << std::endl;
        }
        catch ( const std::invalid_argument& e )
        {
            std::cout << "{\"invalid_argument\": \"" << e.what() << "\"}" << std::endl;
        }
        return 0;
    }
    if ( cmd == "history" )
    {
        NomaTrainHistory h( (uint32_t)std::stoul( argv[ 2 ] ), std::stof( argv[ 3 ] ) );
        const std::vector<float> t = readF32( argv[ 4 ] ), y = readF32( argv[ 5 ] );
        h.addTrainStep( t.data(), y.data(), (uint32_t)t.size() );
        h.addValStep( t.data(), y.data(), (uint32_t)t.size() );
        std::cout << "{\"evm\": " << h.evmTrain[ 0 ] << ", \"ser\": " << h.serTrain[ 0 ] << ", \"ber\": " << h.berTrain[ 0 ]
                  << ", \"same\": " << ( h.evmVal[ 0 ] == h.evmTrain[ 0 ] && h.serVal[ 0 ] == h.serTrain[ 0 ] ) << "}" << std::endl;
        return 0;
    }
    if ( cmd == "shuffle" )
    {
        const size_t T = std::stoul( argv[ 2 ] );
        ComplexSampleMatrix rx( 3, ComplexSampleVector( T ) );
        ComplexSampleVector tx( T );
        for ( size_t k = 0; k < T; k++ )
        {
            for ( auto& row : rx )
                row[ k ] = ComplexSample( (float)k, 0.f );
            tx[ k ] = ComplexSample( (float)k, 1.f );
        }
        ShuffleProbe().run( rx, tx );
        bool same = true;
        for ( size_t k = 0; k < T; k++ )
            for ( auto& row : rx )
                same = same && row[ k ].real() == tx[ k ].real();
        std::cout << "{\"same\": " << same << ", \"perm\": [";
        for ( size_t k = 0; k < T; k++ )
            std::cout << ( k ? ", " : "" ) << (long)tx[ k ].real();
        std::cout << "]}" << std::endl;
        return 0;
    }
    if ( cmd == "refdata" )
    {
        const cli::Modulation mod( (uint32_t)std::stoul( argv[ 3 ] ) );
        ComplexSampleMatrix tr, da;
        cli::generateReferenceData( tr, da, (uint32_t)std::stoul( argv[ 2 ] ), mod, (uint32_t)std::stoul( argv[ 4 ] ),
                                    (uint32_t)std::stoul( argv[ 5 ] ), (uint32_t)std::stoul( argv[ 6 ] ) );
        std::cout << "{\"train\": [";
        for ( size_t u = 0; u < tr.size(); u++ )
            for ( size_t k = 0; k < tr[ u ].size(); k++ )
                std::cout << ( u || k ? ", " : "" ) << "[" << tr[ u ][ k ].real() << ", " << tr[ u ][ k ].imag() << "]";
        std::cout << "], \"data\": [";
        for ( size_t u = 0; u < da.size(); u++ )
            for ( size_t k = 0; k < da[ u ].size(); k++ )
                std::cout << ( u || k ? ", " : "" ) << "[" << da[ u ][ k ].real() << ", " << da[ u ][ k ].imag() << "]";
        std::cout << "]}" << std::endl;
        return 0;
    }
    if ( cmd == "modulation" )
    {
        const uint32_t bps = (uint32_t)std::stoul( argv[ 2 ] );
        const cli::Modulation mod( bps );
        std::vector<bool> bits;
        for ( uint32_t s = 0; s < ( 1u << bps ); s++ )
            for ( uint32_t b = 0; b < bps; b++ )
                bits.push_back( ( s >> ( bps - 1 - b ) ) & 1u );
        const ComplexSampleVector sym = mod.encode( bits );
        const std::vector<bool> back = mod.hardDecode( sym );
        double energy = 0;
        for ( auto& z : sym )
            energy += std::norm( z );
        std::cout << "{\"scale\": " << mod.getScale() << ", \"roundtrip\": " << ( back == bits ) << ", \"energy\": " << energy / sym.size()
                  << ", \"points\": [";
        for ( size_t k = 0; k < sym.size(); k++ )
            std::cout << ( k ? ", " : "" ) << "[" << sym[ k ].real() << ", " << sym[ k ].imag() << "]";
        std::cout << "]}" << std::endl;
        return 0;
    }
    std::cerr << "unknown command" << std::endl;
    return 2;
}
