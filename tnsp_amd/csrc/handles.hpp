// Opaque handle types behind the C ABI (include/tnsgpu.h).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>

#include "core.hpp"

struct tns_shape {
    tns::Shape s;
};

struct tns_plan;

namespace tns {
// sets the thread-local message returned by tns_last_error
int set_error(const char* what);

class Staging;
// body of tns_plan_run; staging = nullptr uses the calling thread's own
// (growable) task staging
void run_plan(tns_plan* p, int dtype, const void* const* inputs, void* const* outputs, void* workspace,
              hipStream_t stream, Staging* staging);
// device workspace bytes tns_plan_run needs for the plan
int64_t plan_workspace_bytes(const tns_plan* p, int dtype);
bool plan_is_svd(const tns_plan* p);
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
