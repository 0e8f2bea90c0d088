// noalloc.hip -- the device scratch of the allocation-free per-rule entry
// point (gpu_evaluate_rule_noalloc): a file-scope device array that every
// GPU holds from module load to process end, so calls through that entry
// point allocate nothing.
//
// In host code `scratch` is a host-side stand-in, not a device address: the
// array's address on each device comes from hipGetSymbolAddress, taken once
// per device with that device current (at context set-up) and kept here; all
// host code uses that pointer, never the array's name. Each GPU has its own
// copy at its own address. Calls on one GPU share it and it keeps the last
// call's contents, so the per-rule path fills its counters before every
// launch; calls that overlap on one GPU race.
#include <hip/hip_runtime.h>

#include <cstdio>
#include <mutex>

#include "noalloc.hpp"

namespace gr {

static __device__ __align__(256) unsigned char scratch[kNoAllocBytes];

namespace {
constexpr int kMaxDevices = 64;
void* g_scratch[kMaxDevices] = {};
std::mutex g_mu;
}  // namespace

bool noalloc_init(int device) {
    if (device < 0 || device >= kMaxDevices) {
        std::fprintf(stderr, "[gpu_rules] device %d beyond the %d the allocation-free scratch table holds\n", device,
                     kMaxDevices);
        return false;
    }
    std::lock_guard<std::mutex> g(g_mu);
    if (g_scratch[device]) return true;
    int cur = -1;
    if (hipGetDevice(&cur) != hipSuccess || cur != device) {
        std::fprintf(stderr, "[gpu_rules] allocation-free scratch: device %d is not current\n", device);
        return false;
    }
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) {
        std::fprintf(stderr, "[gpu_rules] hipGetSymbolAddress(scratch) on device %d: %s\n", device,
                     hipGetErrorString(e));
        return false;
    }
    g_scratch[device] = p;
    return true;
}

void* noalloc_scratch(int device) {
    if (device < 0 || device >= kMaxDevices) return nullptr;
    std::lock_guard<std::mutex> g(g_mu);
    return g_scratch[device];
}

}  // namespace gr
