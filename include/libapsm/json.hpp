// Minimal JSON value and parser: the subset of nlohmann::json that the detector factories
// and the NOMA_detect tool use (parse, value(key, default), operator[], items(), contains,
// get<T>(), comparison with strings and numbers). Used when <nlohmann/json.hpp> is not
// available; the reference pulls nlohmann/json in as a git submodule (.gitmodules:1-10).
#pragma once

#include <cmath>
#include <cstdint>
#include <istream>
#include <iterator>
#include <map>
#include <memory>
#include <sstream>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <utility>
#include <vector>

namespace libapsm {

class Json
{
public:
    enum class Kind
    {
        Null,
        Bool,
        Number,
        String,
        Array,
        Object
    };

    Json() = default;
    Json( std::nullptr_t ) {}
    Json( bool b )
        : kind_( Kind::Bool )
        , b_( b )
    {
    }
    template <typename T, typename std::enable_if<std::is_arithmetic<T>::value && !std::is_same<T, bool>::value, int>::type = 0>
    Json( T v )
        : kind_( Kind::Number )
        , n_( (double)v )
        , integral_( std::is_integral<T>::value )
    {
    }
    Json( const char* s )
        : kind_( Kind::String )
        , s_( s )
    {
    }
    Json( std::string s )
        : kind_( Kind::String )
        , s_( std::move( s ) )
    {
    }
    // {{"key", value}, ...} builds an object, as in nlohmann::json
    Json( std::initializer_list<std::pair<const std::string, Json>> init )
        : kind_( Kind::Object )
    {
        for ( auto& kv : init )
            o_.emplace_back( kv.first, kv.second );
    }

    static Json array() { Json j; j.kind_ = Kind::Array; return j; }
    static Json object() { Json j; j.kind_ = Kind::Object; return j; }

    static Json parse( const std::string& text )
    {
        size_t pos = 0;
        Json j = parseValue( text, pos );
        skipWs( text, pos );
        if ( pos != text.size() )
            throw std::invalid_argument( "json: trailing characters at offset " + std::to_string( pos ) );
        return j;
    }
    static Json parse( std::istream& in )
    {
        std::string text( ( std::istreambuf_iterator<char>( in ) ), std::istreambuf_iterator<char>() );
        return parse( text );
    }

    Kind kind() const { return kind_; }
    bool is_null() const { return kind_ == Kind::Null; }
    bool is_boolean() const { return kind_ == Kind::Bool; }
    bool is_number() const { return kind_ == Kind::Number; }
    bool is_string() const { return kind_ == Kind::String; }
    bool is_array() const { return kind_ == Kind::Array; }
    bool is_object() const { return kind_ == Kind::Object; }

    bool contains( const std::string& key ) const { return find( key ) != nullptr; }

    const Json& operator[]( const std::string& key ) const
    {
        const Json* v = find( key );
        if ( !v )
            throw std::out_of_range( "json: key '" + key + "' not found" );
        return *v;
    }
    Json& operator[]( const std::string& key )
    {
        if ( kind_ == Kind::Null )
            kind_ = Kind::Object;
        if ( kind_ != Kind::Object )
            throw std::domain_error( "json: not an object" );
        for ( auto& kv : o_ )
            if ( kv.first == key )
                return kv.second;
        o_.emplace_back( key, Json() );
        return o_.back().second;
    }
    const Json& operator[]( size_t i ) const
    {
        if ( kind_ != Kind::Array || i >= a_.size() )
            throw std::out_of_range( "json: index out of range" );
        return a_[ i ];
    }
    void push_back( Json v )
    {
        if ( kind_ == Kind::Null )
            kind_ = Kind::Array;
        a_.push_back( std::move( v ) );
    }
    size_t size() const { return kind_ == Kind::Array ? a_.size() : ( kind_ == Kind::Object ? o_.size() : 0 ); }

    template <typename T>
    T get() const
    {
        return convert<T>();
    }

    // value(key, default): the key's value converted to the default's type, or the default
    template <typename T>
    typename std::conditional<std::is_same<typename std::decay<T>::type, const char*>::value || std::is_array<typename std::remove_reference<T>::type>::value,
                              std::string, typename std::decay<T>::type>::type
    value( const std::string& key, T&& def ) const
    {
        using R = typename std::conditional<std::is_same<typename std::decay<T>::type, const char*>::value
                                                || std::is_array<typename std::remove_reference<T>::type>::value,
                                            std::string, typename std::decay<T>::type>::type;
        const Json* v = find( key );
        if ( !v || v->is_null() )
            return R( def );
        return v->convert<R>();
    }

    struct Item
    {
        const std::string& k;
        const Json& v;
        const std::string& key() const { return k; }
        const Json& value() const { return v; }
    };
    class ItemRange
    {
    public:
        explicit ItemRange( const Json& j )
        {
            if ( j.kind_ == Kind::Object )
                for ( auto& kv : j.o_ )
                    items_.push_back( Item { kv.first, kv.second } );
        }
        typename std::vector<Item>::const_iterator begin() const { return items_.begin(); }
        typename std::vector<Item>::const_iterator end() const { return items_.end(); }

    private:
        std::vector<Item> items_;
    };
    ItemRange items() const { return ItemRange( *this ); }

    bool operator==( const std::string& s ) const { return kind_ == Kind::String && s_ == s; }
    bool operator==( const char* s ) const { return kind_ == Kind::String && s_ == s; }
    bool operator!=( const std::string& s ) const { return !( *this == s ); }
    bool operator!=( const char* s ) const { return !( *this == s ); }

    std::string dump() const
    {
        std::ostringstream os;
        write( os );
        return os.str();
    }

private:
    Kind kind_ = Kind::Null;
    bool b_ = false;
    double n_ = 0.0;
    bool integral_ = false;
    std::string s_;
    std::vector<Json> a_;
    std::vector<std::pair<std::string, Json>> o_;

    const Json* find( const std::string& key ) const
    {
        if ( kind_ != Kind::Object )
            return nullptr;
        for ( auto& kv : o_ )
            if ( kv.first == key )
                return &kv.second;
        return nullptr;
    }

    template <typename T>
    T convert() const
    {
        if constexpr ( std::is_same<T, std::string>::value )
        {
            if ( kind_ != Kind::String )
                throw std::domain_error( "json: type must be string" );
            return s_;
        }
        else if constexpr ( std::is_same<T, bool>::value )
        {
            if ( kind_ != Kind::Bool )
                throw std::domain_error( "json: type must be boolean" );
            return b_;
        }
        else if constexpr ( std::is_arithmetic<T>::value )
        {
            if ( kind_ != Kind::Number )
                throw std::domain_error( "json: type must be number" );
            return (T)n_;
        }
        else if constexpr ( std::is_same<T, Json>::value )
            return *this;
        else
            static_assert( std::is_same<T, void>::value, "unsupported json conversion" );
    }

    static void skipWs( const std::string& t, size_t& p )
    {
        while ( p < t.size() && ( t[ p ] == ' ' || t[ p ] == '\n' || t[ p ] == '\r' || t[ p ] == '\t' ) )
            p++;
    }
    [[noreturn]] static void fail( const std::string& what, size_t p )
    {
        throw std::invalid_argument( "json: " + what + " at offset " + std::to_string( p ) );
    }
    static std::string parseString( const std::string& t, size_t& p )
    {
        if ( t[ p ] != '"' )
            fail( "expected string", p );
        p++;
        std::string out;
        while ( p < t.size() && t[ p ] != '"' )
        {
            char c = t[ p++ ];
            if ( c == '\\' )
            {
                if ( p >= t.size() )
                    fail( "bad escape", p );
                char e = t[ p++ ];
                switch ( e )
                {
                case '"': out += '"'; break;
                case '\\': out += '\\'; break;
                case '/': out += '/'; break;
                case 'b': out += '\b'; break;
                case 'f': out += '\f'; break;
                case 'n': out += '\n'; break;
                case 'r': out += '\r'; break;
                case 't': out += '\t'; break;
                case 'u':
                {
                    if ( p + 4 > t.size() )
                        fail( "bad unicode escape", p );
                    unsigned cp = std::stoul( t.substr( p, 4 ), nullptr, 16 );
                    p += 4;
                    if ( cp < 0x80 )
                        out += (char)cp;
                    else if ( cp < 0x800 )
                    {
                        out += (char)( 0xC0 | ( cp >> 6 ) );
                        out += (char)( 0x80 | ( cp & 0x3F ) );
                    }
                    else
                    {
                        out += (char)( 0xE0 | ( cp >> 12 ) );
                        out += (char)( 0x80 | ( ( cp >> 6 ) & 0x3F ) );
                        out += (char)( 0x80 | ( cp & 0x3F ) );
                    }
                    break;
                }
                default: fail( "bad escape", p );
                }
            }
            else
                out += c;
        }
        if ( p >= t.size() )
            fail( "unterminated string", p );
        p++;
        return out;
    }
    static Json parseValue( const std::string& t, size_t& p )
    {
        skipWs( t, p );
        if ( p >= t.size() )
            fail( "unexpected end", p );
        char c = t[ p ];
        if ( c == '{' )
        {
            Json j = object();
            p++;
            skipWs( t, p );
            if ( p < t.size() && t[ p ] == '}' )
            {
                p++;
                return j;
            }
            while ( true )
            {
                skipWs( t, p );
                std::string k = parseString( t, p );
                skipWs( t, p );
                if ( p >= t.size() || t[ p ] != ':' )
                    fail( "expected ':'", p );
                p++;
                Json v = parseValue( t, p );
                j.o_.emplace_back( k, std::move( v ) );
                skipWs( t, p );
                if ( p < t.size() && t[ p ] == ',' )
                {
                    p++;
                    continue;
                }
                if ( p < t.size() && t[ p ] == '}' )
                {
                    p++;
                    return j;
                }
                fail( "expected ',' or '}'", p );
            }
        }
        if ( c == '[' )
        {
            Json j = array();
            p++;
            skipWs( t, p );
            if ( p < t.size() && t[ p ] == ']' )
            {
                p++;
                return j;
            }
            while ( true )
            {
                j.a_.push_back( parseValue( t, p ) );
                skipWs( t, p );
                if ( p < t.size() && t[ p ] == ',' )
                {
                    p++;
                    continue;
                }
                if ( p < t.size() && t[ p ] == ']' )
                {
                    p++;
                    return j;
                }
                fail( "expected ',' or ']'", p );
            }
        }
        if ( c == '"' )
            return Json( parseString( t, p ) );
        if ( t.compare( p, 4, "true" ) == 0 )
        {
            p += 4;
            return Json( true );
        }
        if ( t.compare( p, 5, "false" ) == 0 )
        {
            p += 5;
            return Json( false );
        }
        if ( t.compare( p, 4, "null" ) == 0 )
        {
            p += 4;
            return Json();
        }
        size_t q = p;
        if ( q < t.size() && ( t[ q ] == '-' || t[ q ] == '+' ) )
            q++;
        bool integral = true;
        while ( q < t.size() && ( isdigit( (unsigned char)t[ q ] ) || t[ q ] == '.' || t[ q ] == 'e' || t[ q ] == 'E' || t[ q ] == '-' || t[ q ] == '+' ) )
        {
            if ( t[ q ] == '.' || t[ q ] == 'e' || t[ q ] == 'E' )
                integral = false;
            q++;
        }
        if ( q == p )
            fail( "unexpected character", p );
        Json j( std::stod( t.substr( p, q - p ) ) );
        j.integral_ = integral;
        p = q;
        return j;
    }
    void write( std::ostream& os ) const
    {
        switch ( kind_ )
        {
        case Kind::Null: os << "null"; break;
        case Kind::Bool: os << ( b_ ? "true" : "false" ); break;
        case Kind::Number:
            if ( integral_ && std::fabs( n_ ) < 9e15 )
                os << (long long)n_;
            else
                os << n_;
            break;
        case Kind::String: os << '"' << s_ << '"'; break;
        case Kind::Array:
            os << '[';
            for ( size_t i = 0; i < a_.size(); i++ )
            {
                if ( i )
                    os << ',';
                a_[ i ].write( os );
            }
            os << ']';
            break;
        case Kind::Object:
            os << '{';
            for ( size_t i = 0; i < o_.size(); i++ )
            {
                if ( i )
                    os << ',';
                os << '"' << o_[ i ].first << "\":";
                o_[ i ].second.write( os );
            }
            os << '}';
            break;
        }
    }
};

} // namespace libapsm
