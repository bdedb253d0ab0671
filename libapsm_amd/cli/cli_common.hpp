// Host-side helpers of the NOMA_detect tool: the binary container format, the Gray-mapped
// QAM constellations, the error printers, antenna selection and a small argument parser.
// Behaviour follows the reference CLI (cpp/cli/common/*.cpp, cpp/TOOL_overview.md:76-181);
// file:line citations are given per function.
#pragma once

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <iomanip>
#include <iostream>
#include <map>
#include <random>
#include <stdexcept>
#include <string>
#include <vector>

#include "libapsm/json_select.hpp"
#include "noma/types.cuh"

namespace cli {

// ---------------------------------------------------------------------------------------
// Binary container (cpp/external/binary/include/binary/binary_file.hpp:159-228): i32 key
// count; per key an i32 length and its characters; then per key, in the same order, i32
// rank, i32 dimensions and float32 row-major data. Keys are written in sorted order.
// ---------------------------------------------------------------------------------------
struct Array
{
    std::vector<int32_t> dims;
    std::vector<float> data;
};

inline std::map<std::string, Array> readBinary( const std::string& path )
{
    std::ifstream f( path, std::ios::binary );
    if ( !f )
        throw std::runtime_error( "Error loading the file " + path + ": File not found." );
    auto i32 = [ & ]() {
        int32_t v = 0;
        if ( !f.read( reinterpret_cast<char*>( &v ), 4 ) )
            throw std::runtime_error( "Error loading the file " + path + ": File truncated." );
        return v;
    };
    const int32_t n = i32();
    if ( n < 0 )
        throw std::runtime_error( "Error loading the file " + path + ": bad key count" );
    std::vector<std::string> keys;
    for ( int32_t k = 0; k < n; k++ )
    {
        const int32_t len = i32();
        if ( len < 0 || len > ( 1 << 20 ) )
            throw std::runtime_error( "Error loading the file " + path + ": bad key length" );
        std::string key( (size_t)len, '\0' );
        if ( !f.read( &key[ 0 ], len ) )
            throw std::runtime_error( "Error loading the file " + path + ": File truncated." );
        keys.push_back( key );
    }
    std::map<std::string, Array> out;
    for ( const auto& key : keys )
    {
        Array a;
        const int32_t rank = i32();
        size_t count = 1;
        for ( int32_t d = 0; d < rank; d++ )
        {
            const int32_t dim = i32();
            if ( dim < 0 )
                throw std::runtime_error( "Error loading the file " + path + ": negative dimension" );
            a.dims.push_back( dim );
            count *= (size_t)dim;
        }
        a.data.resize( count );
        if ( count && !f.read( reinterpret_cast<char*>( a.data.data() ), count * 4 ) )
            throw std::runtime_error( "Error loading the file " + path + ": File truncated." );
        out[ key ] = std::move( a );
    }
    return out;
}

inline void writeBinary( const std::string& path, const std::map<std::string, Array>& arrays )
{
    std::ofstream f( path, std::ios::binary | std::ios::trunc );
    if ( !f )
        throw std::runtime_error( "Error writing the file " + path );
    auto i32 = [ & ]( int32_t v ) { f.write( reinterpret_cast<const char*>( &v ), 4 ); };
    i32( (int32_t)arrays.size() );
    for ( const auto& kv : arrays )
    {
        i32( (int32_t)kv.first.size() );
        f.write( kv.first.data(), kv.first.size() );
    }
    for ( const auto& kv : arrays )
    {
        i32( (int32_t)kv.second.dims.size() );
        for ( int32_t d : kv.second.dims )
            i32( d );
        f.write( reinterpret_cast<const char*>( kv.second.data.data() ), kv.second.data.size() * 4 );
    }
}

// [sample][antenna][2] floats -> [antenna][sample] complex times scale
// (cpp/cli/common/binary_load.cpp:55-79)
inline ComplexSampleMatrix loadComplex( const std::string& path, const std::string& key, float scale = 1.f )
{
    const auto arrays = readBinary( path );
    auto it = arrays.find( key );
    if ( it == arrays.end() )
        throw std::runtime_error( "Error loading the file " + path + ": no variable '" + key + "'" );
    const Array& a = it->second;
    if ( a.dims.size() != 3 || a.dims[ 2 ] != 2 )
        throw std::runtime_error( "variable '" + key + "' is not a [samples][antennas][2] array" );
    const size_t S = a.dims[ 0 ], A = a.dims[ 1 ];
    ComplexSampleMatrix m( A, ComplexSampleVector( S ) );
    for ( size_t s = 0; s < S; s++ )
        for ( size_t n = 0; n < A; n++ )
            m[ n ][ s ] = scale * ComplexSample( a.data[ ( s * A + n ) * 2 ], a.data[ ( s * A + n ) * 2 + 1 ] );
    return m;
}

inline Array toArray( const ComplexSampleMatrix& m )
{
    Array a;
    const size_t A = m.size(), S = A ? m[ 0 ].size() : 0;
    a.dims = { (int32_t)S, (int32_t)A, 2 };
    a.data.resize( S * A * 2 );
    for ( size_t n = 0; n < A; n++ )
        for ( size_t s = 0; s < S; s++ )
        {
            a.data[ ( s * A + n ) * 2 ] = m[ n ][ s ].real();
            a.data[ ( s * A + n ) * 2 + 1 ] = m[ n ][ s ].imag();
        }
    return a;
}

// ---------------------------------------------------------------------------------------
// Modulation (cpp/cli/common/modulation.hpp:65-222, modulation.cpp:68-151): square Gray QAM
// with PAM levels ... -3 -1 1 3 ... per dimension, scaled to unit average energy:
// 1/sqrt(bps) for BPSK / QPSK, 1/sqrt((2^(bps+1) - 2) / 3) for QAM.
// ---------------------------------------------------------------------------------------
class Modulation
{
public:
    explicit Modulation( uint32_t bps = 4 )
        : bps( bps )
    {
        if ( bps > 14 || ( bps > 2 && bps % 2 ) )
            throw std::invalid_argument( "Unsupported modulation" );
    }
    static Modulation fromString( std::string name )
    {
        std::transform( name.begin(), name.end(), name.begin(), ::toupper );
        static const std::pair<const char*, uint32_t> table[] = { { "OFF", 0 }, { "BPSK", 1 }, { "QPSK", 2 }, { "QAM16", 4 }, { "QAM64", 6 },
                                                                  { "QAM256", 8 }, { "QAM1024", 10 }, { "QAM4096", 12 }, { "QAM16384", 14 } };
        for ( const auto& e : table )
            if ( name == e.first )
                return Modulation( e.second );
        throw std::invalid_argument( "Unsupported modulation: " + name );
    }
    uint32_t getBitPerSymbol() const { return bps; }
    RealSample getScale() const
    {
        if ( bps == 0 )
            return 0.f;
        if ( bps <= 2 )
            return (RealSample)( 1.0 / std::sqrt( (double)bps ) );
        return (RealSample)( 1.0 / std::sqrt( (double)( ( 2u << bps ) - 2 ) / 3.0 ) );
    }
    RealSample getRevScale() const
    {
        if ( bps == 0 )
            return 0.f;
        if ( bps <= 2 )
            return (RealSample)std::sqrt( (double)bps );
        return (RealSample)std::sqrt( (double)( ( 2u << bps ) - 2 ) / 3.0 );
    }

    // Gray index of the nearest of `levels` PAM levels, clamped at the outermost ones
    static uint32_t demap1d( RealSample x, uint32_t levels )
    {
        int32_t xi = (int32_t)std::round( ( x + levels - 1 ) / 2 );
        xi = std::min( (int32_t)levels - 1, std::max( 0, xi ) );
        return (uint32_t)( xi ^ ( xi >> 1 ) );
    }
    // PAM level (odd integer) of a Gray-coded index
    static RealSample map1d( uint32_t gray, uint32_t levels )
    {
        uint32_t bin = gray;
        for ( uint32_t g = gray >> 1; g; g >>= 1 )
            bin ^= g;
        return (RealSample)( 2 * (int32_t)bin - (int32_t)levels + 1 );
    }

    std::vector<bool> hardDecode( const ComplexSampleVector& in ) const
    {
        std::vector<bool> out;
        out.reserve( in.size() * bps );
        const RealSample rev = getRevScale();
        for ( const auto& s : in )
        {
            const ComplexSample y = s * rev;
            uint32_t bits;
            if ( bps == 1 )
                bits = demap1d( y.real(), 2 );
            else
            {
                const uint32_t levels = 1u << ( bps / 2 );
                bits = ( demap1d( y.real(), levels ) << ( bps / 2 ) ) | demap1d( y.imag(), levels );
            }
            for ( uint32_t b = 0; b < bps; b++ )
                out.push_back( ( bits >> ( bps - 1 - b ) ) & 1u );
        }
        return out;
    }

    ComplexSampleVector encode( const std::vector<bool>& in ) const
    {
        ComplexSampleVector out;
        out.reserve( in.size() / std::max( bps, 1u ) );
        const RealSample s = getScale();
        uint32_t bits = 0;
        for ( size_t i = 0; i < in.size(); i++ )
        {
            bits = ( bits << 1 ) | ( in[ i ] ? 1u : 0u );
            if ( i % bps == bps - 1 )
            {
                if ( bps == 1 )
                    out.push_back( ComplexSample( map1d( bits, 2 ), 0 ) * s );
                else
                {
                    const uint32_t levels = 1u << ( bps / 2 );
                    out.push_back( ComplexSample( map1d( bits >> ( bps / 2 ), levels ), map1d( bits & ( levels - 1 ), levels ) ) * s );
                }
                bits = 0;
            }
        }
        return out;
    }

private:
    uint32_t bps;
};

// ---------------------------------------------------------------------------------------
// Error printers (cpp/cli/common/util.cpp:62-192): one character per symbol (80 per line),
// then a summary; each returns the value the scripting line prints.
// ---------------------------------------------------------------------------------------
enum OutputMode
{
    BER,
    SER,
    EVM,
    Scripting
};

inline float printSymbolErrors( const ComplexSampleVector& est, const ComplexSampleVector& ref, const Modulation& mod )
{
    const float s = mod.getScale();
    const float mx = mod.getBitPerSymbol() == 1 ? 1.f : ( ( 1u << ( mod.getBitPerSymbol() / 2 ) ) - 1 ) * s;
    uint32_t errors = 0;
    for ( size_t k = 0; k < est.size(); k++ )
    {
        const ComplexSample y = est[ k ], t = ref[ k ];
        if ( k % 80 == 0 )
            std::cout << std::endl;
        // real part: outside the decision region unless both lie beyond the outer level
        const bool re = std::abs( y.real() - t.real() ) > s && !( y.real() > +mx && t.real() > +mx - s )
            && !( y.real() < -mx && t.real() < -mx + s );
        // imaginary part: the reference tool's clamp comparison is kept as it is (util.cpp:89-97)
        const bool im = std::abs( y.imag() - t.imag() ) > s && !( y.imag() > +mx && t.imag() < +mx - s )
            && !( y.imag() < -mx && t.imag() > -mx + s ) && mod.getBitPerSymbol() > 1;
        if ( std::isnan( std::abs( y ) ) )
        {
            errors++;
            std::cout << "n";
        }
        else if ( re )
        {
            errors++;
            std::cout << "r";
        }
        else if ( im )
        {
            errors++;
            std::cout << "i";
        }
        else
            std::cout << ".";
    }
    const float pct = 100.f * errors / est.size();
    std::cout << " symbol errors: " << errors << " of " << est.size() << " - " << std::fixed << std::setprecision( 2 ) << pct << " percent"
              << std::endl;
    return pct;
}

inline float printBitErrors( const ComplexSampleVector& est, const ComplexSampleVector& ref, const Modulation& mod )
{
    const std::vector<bool> a = mod.hardDecode( est ), b = mod.hardDecode( ref );
    const uint32_t bps = mod.getBitPerSymbol();
    uint32_t errors = 0;
    for ( size_t i = 0; i < a.size(); i++ )
    {
        if ( i % ( bps * 80 ) == 0 )
            std::cout << std::endl;
        if ( i % bps == 0 )
            std::cout << " ";
        if ( a[ i ] != b[ i ] )
        {
            errors++;
            std::cout << "x";
        }
        else
            std::cout << ".";
    }
    const float pct = 100.f * errors / a.size();
    std::cout << " bit errors: " << errors << " of " << a.size() << " - " << std::fixed << std::setprecision( 2 ) << pct << " percent" << std::endl;
    return pct;
}

inline float printEvm( const ComplexSampleVector& est, const ComplexSampleVector& ref )
{
    static const char* const bars[] = { " ", "▁", "▂", "▃", "▄", "▅", "▆", "▇", "█" };
    float sq = 0.f;
    for ( size_t k = 0; k < est.size(); k++ )
    {
        if ( k % 80 == 0 )
            std::cout << std::endl;
        const float e = std::abs( est[ k ] - ref[ k ] );
        sq += e * e;
        const uint32_t level = (uint32_t)( e / 0.5f * 8 ); // a full bar from an error of 0.5
        std::cout << bars[ std::min<uint32_t>( level, 8 ) ];
    }
    const float evm = std::sqrt( sq / est.size() );
    std::cout << " EVM: " << std::fixed << std::setprecision( 4 ) << evm << std::endl;
    return evm;
}

inline void printErrors( OutputMode mode, const ComplexSampleVector& est, const ComplexSampleVector& ref, const Modulation& mod )
{
    switch ( mode )
    {
    case BER: printBitErrors( est, ref, mod ); break;
    case SER: printSymbolErrors( est, ref, mod ); break;
    case EVM: printEvm( est, ref ); break;
    case Scripting:
    {
        const float ber = printBitErrors( est, ref, mod );
        const float ser = printSymbolErrors( est, ref, mod );
        const float evm = printEvm( est, ref );
        std::cout.clear();
        std::cout << std::setprecision( -1 ) << ber << "," << ser << "," << evm << std::endl;
        std::cout.setstate( std::ios_base::failbit );
        break;
    }
    }
}

// ---------------------------------------------------------------------------------------
// Antenna selection (cpp/cli/common/util.cpp:194-260). The reference fixes 16 antenna slots;
// here the slot count is max(16, antennas in the file), so files with 32 or 64 antennas work
// and 16-antenna files select exactly as before.
// ---------------------------------------------------------------------------------------
typedef std::vector<bool> AntennaPattern;

inline AntennaPattern antennaPattern( const nlohmann::json& config, uint32_t slots )
{
    for ( auto& item : config.items() )
        if ( item.key() != "otype" && item.key() != "pattern" && item.key() != "number" )
            throw std::invalid_argument( "Invalid key: " + item.key() );
    AntennaPattern p( slots, false );
    auto count = [ & ]() {
        const uint32_t n = config.value( "number", 16 );
        if ( n > slots )
            throw std::invalid_argument( "more antennas requested than available: " + std::to_string( n ) );
        return n;
    };
    const std::string otype = config[ "otype" ].get<std::string>();
    if ( otype == "all" )
        std::fill( p.begin(), p.end(), true );
    else if ( otype == "pattern" )
    {
        // std::bitset string constructor semantics: the last character is antenna 0
        const std::string s = config.value( "pattern", "11111111111111111111" );
        const size_t n = std::min<size_t>( s.size(), slots );
        for ( size_t i = 0; i < n; i++ )
        {
            const char c = s[ n - 1 - i ];
            if ( c != '0' && c != '1' )
                throw std::invalid_argument( "antenna pattern must consist of 0 and 1" );
            p[ i ] = c == '1';
        }
    }
    else if ( otype == "equidistant" )
    {
        const uint32_t n = count();
        for ( uint32_t i = 0; i < n; i++ )
            p[ (size_t)i * slots / n ] = true;
    }
    else if ( otype == "random" )
    {
        const uint32_t n = count();
        std::vector<uint32_t> idx( slots );
        for ( uint32_t i = 0; i < slots; i++ )
            idx[ i ] = i;
        std::random_device rd;
        std::mt19937 g( rd() );
        std::shuffle( idx.begin(), idx.end(), g );
        for ( uint32_t i = 0; i < n; i++ )
            p[ idx[ i ] ] = true;
    }
    else if ( otype == "first" )
    {
        const uint32_t n = count();
        for ( uint32_t i = 0; i < n; i++ )
            p[ i ] = true;
    }
    else
        throw std::invalid_argument( "Invalid antenna scheme: " + otype );
    return p;
}

inline ComplexSampleMatrix selectAntennas( const ComplexSampleMatrix& in, const AntennaPattern& p )
{
    ComplexSampleMatrix out;
    for ( size_t i = 0; i < std::min( in.size(), p.size() ); i++ )
        if ( p[ i ] )
            out.push_back( in[ i ] );
    return out;
}

// ---------------------------------------------------------------------------------------
// Reference symbols from a seed (cpp/cli/common/util.cpp:274-325): fair bits from
// std::mt19937(seed) through std::bernoulli_distribution(0.5), Gray-mapped; user u takes the
// u-th run of (training + data) symbols. Defaults: 6 users, 685 training and 3840 data
// symbols (cpp/cli/common/util.hpp:67-113).
// ---------------------------------------------------------------------------------------
constexpr uint32_t kDefaultUsers = 6;
constexpr uint32_t kDefaultTrainingSamples = 685;
constexpr uint32_t kDefaultDataSamples = 3840;

inline void generateReferenceData( ComplexSampleMatrix& training, ComplexSampleMatrix& data, uint32_t seed, const Modulation& mod,
                                   uint32_t users = kDefaultUsers, uint32_t nTrain = kDefaultTrainingSamples,
                                   uint32_t nData = kDefaultDataSamples )
{
    std::mt19937 gen( seed );
    std::bernoulli_distribution coin( 0.5 );
    std::vector<bool> bits( (size_t)( nTrain + nData ) * users * mod.getBitPerSymbol() );
    for ( size_t i = 0; i < bits.size(); i++ )
        bits[ i ] = coin( gen );
    const ComplexSampleVector sym = mod.encode( bits );
    training.clear();
    data.clear();
    for ( uint32_t u = 0; u < users; u++ )
    {
        auto it = sym.begin() + (size_t)u * ( nTrain + nData );
        training.emplace_back( it, it + nTrain );
        data.emplace_back( it + nTrain, it + nTrain + nData );
    }
}

// ---------------------------------------------------------------------------------------
// Argument parser for the flags of cpp/TOOL_overview.md:80-96.
// ---------------------------------------------------------------------------------------
struct Option
{
    std::string shortName, longName, help, value;
    bool isFlag = false, present = false, required = false;
    std::string defaultValue;
};

class ArgParser
{
public:
    ArgParser( std::string program, std::string version, std::string description )
        : program( std::move( program ) )
        , version( std::move( version ) )
        , description( std::move( description ) )
    {
    }
    void add( const std::string& s, const std::string& l, const std::string& help, const std::string& def = "", bool flag = false,
              bool required = false )
    {
        Option o;
        o.shortName = s;
        o.longName = l;
        o.help = help;
        o.defaultValue = def;
        o.isFlag = flag;
        o.required = required;
        options.push_back( o );
    }
    // returns false when the program should exit successfully (help / version shown)
    bool parse( int argc, char** argv )
    {
        for ( int i = 1; i < argc; i++ )
        {
            const std::string a = argv[ i ];
            if ( a == "-h" || a == "--help" )
            {
                std::cout << usage();
                return false;
            }
            if ( a == "-v" || a == "--version" )
            {
                std::cout << version << std::endl;
                return false;
            }
            Option* o = find( a );
            if ( !o )
                throw std::runtime_error( "Unknown argument: " + a );
            o->present = true;
            if ( o->isFlag )
                o->value = "1";
            else
            {
                if ( i + 1 >= argc )
                    throw std::runtime_error( a + ": expected 1 argument(s). 0 provided." );
                o->value = argv[ ++i ];
            }
        }
        for ( const auto& o : options )
            if ( o.required && !o.present )
                throw std::runtime_error( o.longName + ": required." );
        return true;
    }
    bool present( const std::string& name ) const
    {
        const Option* o = findConst( name );
        return o && o->present;
    }
    std::string get( const std::string& name ) const
    {
        const Option* o = findConst( name );
        if ( !o )
            throw std::logic_error( "no option " + name );
        return o->present ? o->value : o->defaultValue;
    }
    bool flag( const std::string& name ) const { return present( name ); }
    std::string usage() const
    {
        std::string u = "Usage: " + program + " [options]\n\n" + description + "\n\nOptional arguments:\n";
        u += "-h  --help                       shows help message and exits\n";
        u += "-v  --version                    prints version information and exits\n";
        for ( const auto& o : options )
        {
            std::string n = o.shortName + std::string( 4 - std::min<size_t>( 4, o.shortName.size() ), ' ' ) + o.longName;
            n.resize( std::max<size_t>( n.size() + 1, 33 ), ' ' );
            u += n + o.help;
            if ( o.required )
                u += " [required]";
            else if ( !o.defaultValue.empty() )
                u += " [default: " + o.defaultValue + "]";
            u += "\n";
        }
        return u;
    }

private:
    Option* find( const std::string& a )
    {
        for ( auto& o : options )
            if ( a == o.shortName || a == o.longName )
                return &o;
        return nullptr;
    }
    const Option* findConst( const std::string& a ) const
    {
        for ( const auto& o : options )
            if ( a == o.shortName || a == o.longName )
                return &o;
        return nullptr;
    }
    std::string program, version, description;
    std::vector<Option> options;
};

} // namespace cli
