// Opaque handle types behind the C ABI (include/tnsgpu.h).
#pragma once

#include <string>

#include "core.hpp"

struct tns_shape {
    tns::Shape s;
};

namespace tns {
// sets the thread-local message returned by tns_last_error
int set_error(const char* what);
}  // namespace tns

#define TNS_TRY(...)                                    \
    try {                                               \
        __VA_ARGS__;                                    \
        return 0;                                       \
    } catch (const std::exception& e) {                 \
        return tns::set_error(e.what());                \
    } catch (...) {                                     \
        return tns::set_error("unknown C++ exception"); \
    }
