// detectVersion identifiers, as cpp/lib/apsm/apsm_versions.cuh:58-105 defines them. Every
// name selects the same dictionary-streaming detect kernel in this library.
#pragma once

#include <cstdint>
#include <string>

enum apsm_function_id
{
    APSM_UNKNOWN_ID = 0,
    APSM_DETECT_ID = 1
};

constexpr uint32_t version_id( const enum apsm_function_id function, const int32_t version )
{
    return ( static_cast<int32_t>( function ) << 16 ) + version;
}

enum apsm_versions
{
    UNKNOWN = 0,
    APSM_DETECT_ORIGINAL = version_id( APSM_DETECT_ID, 1 ),
    APSM_DETECT_OLDFAST = version_id( APSM_DETECT_ID, 2 ),
    APSM_DETECT_SHMEM = version_id( APSM_DETECT_ID, 6 ),
    APSM_DETECT_BALANCED = version_id( APSM_DETECT_ID, 7 )
};

inline const std::string apsm_get_version_string( const int32_t id )
{
    switch ( id )
    {
    case APSM_DETECT_ORIGINAL: return "APSM Detect original";
    case APSM_DETECT_OLDFAST: return "APSM Detect old fast code";
    case APSM_DETECT_SHMEM: return "APSM Detect Store vectors in shared memory";
    case APSM_DETECT_BALANCED: return "APSM Detect Balance memory and computation";
    default: return "Unknown";
    }
}

#ifndef APSM_DETECT_VERSION
#define APSM_DETECT_VERSION apsm_versions::APSM_DETECT_BALANCED
#endif
