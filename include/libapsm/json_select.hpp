// nlohmann::json when the real library is on the include path (the reference's submodule,
// .gitmodules:1-10), otherwise the bundled subset in libapsm/json.hpp under the same name.
#pragma once

#if __has_include( <nlohmann/json.hpp> )
#include <nlohmann/json.hpp>
#else
#include "libapsm/json.hpp"
namespace nlohmann {
using json = libapsm::Json;
}
#endif
