// tns_plan_run_noalloc: tns_plan_run for callers that cannot allocate device
// memory (code under HIP graph capture, tight serving loops, harnesses that
// forbid hipMalloc).  The plan's workspace and its task records live in one
// library-owned static device array instead of a caller workspace and the
// growable per-thread staging buffers.
//
// The array's name is a host-side stand-in: host code never passes it to a
// kernel or a copy.  Its device address is looked up once per device with
// hipGetSymbolAddress (each GPU holds its own copy of the array) and that
// address is what the kernels and the task upload receive.
#include <hip/hip_runtime.h>

#include <mutex>
#include <string>

#include "../../include/tnsgpu.h"
#include "core.hpp"
#include "handles.hpp"
#include "runtime.hpp"

namespace {

// workspace part (contract operands and product, QR factors and the global
// scratch of large sectors) and task-record part of the array
constexpr int64_t NOALLOC_WORKSPACE_BYTES = int64_t(64) << 20;
constexpr int64_t NOALLOC_TASK_BYTES = int64_t(4) << 20;

__device__ __align__(256) unsigned char scratch[NOALLOC_WORKSPACE_BYTES + NOALLOC_TASK_BYTES];

constexpr int MAX_DEVICES = 64;
std::mutex scratch_mu;
unsigned char* scratch_ptr[MAX_DEVICES];  // device address of `scratch` per device

unsigned char* scratch_on_current_device() {
    int dev = -1;
    tns::check_hip(hipGetDevice(&dev), "noalloc: current device");
    if (dev < 0 || dev >= MAX_DEVICES) throw tns::Error("noalloc: device index out of range");
    std::lock_guard<std::mutex> lk(scratch_mu);
    if (!scratch_ptr[dev]) {
        void* p = nullptr;
        tns::check_hip(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)), "noalloc: scratch address");
        if (!p) throw tns::Error("noalloc: scratch address is null");
        scratch_ptr[dev] = static_cast<unsigned char*>(p);
    }
    return scratch_ptr[dev];
}

}  // namespace

extern "C" {

int tns_plan_run_noalloc(tns_plan* p, int dtype, const void* const* inputs, void* const* outputs, void* stream) {
    TNS_TRY({
        if (!p) throw tns::Error("noalloc: null plan");
        if (dtype != 0 && dtype != 1) throw tns::Error("unsupported dtype");
        if (tns::plan_is_svd(p))
            throw tns::Error("noalloc: svd plans are not supported (tns_svd_factor writes into the caller's workspace)");
        // size checks come before any HIP call, so an oversized plan fails the
        // same way with or without a device
        const int64_t need = tns::plan_workspace_bytes(p, dtype);
        if (need > NOALLOC_WORKSPACE_BYTES)
            throw tns::Error("noalloc: the plan needs " + std::to_string(need) +
                             " bytes of workspace, more than the static maximum of " +
                             std::to_string(NOALLOC_WORKSPACE_BYTES) + " bytes; use tns_plan_run");
        unsigned char* base = scratch_on_current_device();
        tns::Staging staging(reinterpret_cast<char*>(base + NOALLOC_WORKSPACE_BYTES), size_t(NOALLOC_TASK_BYTES));
        tns::run_plan(p, dtype, inputs, outputs, base, static_cast<hipStream_t>(stream), &staging);
    })
}

}  // extern "C"
