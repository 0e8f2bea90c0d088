// kernels.hip -- utility kernels (copy yardstick, fills) and the launchers of
// the counting kernels: each launch goes to the slot class (4, 8 or 16
// variable slots) that holds every plan it runs; the counting kernels of each
// class are compiled in search_s4.hip, search_s8.hip and search_s16.hip.
#include <hip/hip_runtime.h>

#include <cstdint>

#include "kernels.hpp"
#include "search.hpp"

namespace gr {

// Plain streaming copy: the yardstick that scales hardware byte counters.
__global__ void k_copy_u32(const uint4* __restrict__ src, uint4* __restrict__ dst, int64_t n16) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n16; i += (int64_t)gridDim.x * blockDim.x)
        dst[i] = src[i];
}

__global__ void k_fill_u64(unsigned long long* p, int64_t n, unsigned long long v) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        p[i] = v;
}
__global__ void k_fill_u32(uint32_t* p, int64_t n, uint32_t v) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        p[i] = v;
}
// One range per grid row.
__global__ void k_zero_ranges(ZeroRanges z) {
    const int r = blockIdx.y;
    if (r >= z.count) return;
    uint32_t* p = z.p[r];
    const int64_t n = z.n[r];
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        p[i] = 0u;
}

void launch_thread_batch(int kind, int slots, const ThreadArgs& a, const DevGraph& G, int blocks, hipStream_t s) {
    switch (slot_class(slots)) {
        case 4: SlotKernels<4>::thread_batch(kind, a, G, blocks, s); break;
        case 8: SlotKernels<8>::thread_batch(kind, a, G, blocks, s); break;
        default: SlotKernels<16>::thread_batch(kind, a, G, blocks, s); break;
    }
}

void launch_thread_one(int kind, int slots, const Plan& P, const OneArgs& a, const DevGraph& G, int blocks,
                       hipStream_t s) {
    switch (slot_class(slots)) {
        case 4: SlotKernels<4>::thread_one(kind, P, a, G, blocks, s); break;
        case 8: SlotKernels<8>::thread_one(kind, P, a, G, blocks, s); break;
        default: SlotKernels<16>::thread_one(kind, P, a, G, blocks, s); break;
    }
}

void launch_wave_batch(int kind, int slots, int dedup, const WaveArgs& a, const DevGraph& G, int blocks,
                       hipStream_t s) {
    switch (slot_class(slots)) {
        case 4: SlotKernels<4>::wave_batch(kind, dedup, a, G, blocks, s); break;
        case 8: SlotKernels<8>::wave_batch(kind, dedup, a, G, blocks, s); break;
        default: SlotKernels<16>::wave_batch(kind, dedup, a, G, blocks, s); break;
    }
}

void launch_wave_one(int kind, int slots, int dedup, const Plan& P, const WaveArgs& a, const DevGraph& G, int blocks,
                     hipStream_t s) {
    switch (slot_class(slots)) {
        case 4: SlotKernels<4>::wave_one(kind, dedup, P, a, G, blocks, s); break;
        case 8: SlotKernels<8>::wave_one(kind, dedup, P, a, G, blocks, s); break;
        default: SlotKernels<16>::wave_one(kind, dedup, P, a, G, blocks, s); break;
    }
}

void launch_giants(int kind, int slots, int dedup, const WaveArgs& a, const DevGraph& G, int blocks, hipStream_t s) {
    switch (slot_class(slots)) {
        case 4: SlotKernels<4>::giants(kind, dedup, a, G, blocks, s); break;
        case 8: SlotKernels<8>::giants(kind, dedup, a, G, blocks, s); break;
        default: SlotKernels<16>::giants(kind, dedup, a, G, blocks, s); break;
    }
}

size_t wave_static_lds(int slots, int engine) {
    switch (slot_class(slots)) {
        case 4: return SlotKernels<4>::lds_static(engine);
        case 8: return SlotKernels<8>::lds_static(engine);
        default: return SlotKernels<16>::lds_static(engine);
    }
}

int wave_lds_blocks_per_cu(int slots, size_t dynLds, int engine) {
    switch (slot_class(slots)) {
        case 4: return SlotKernels<4>::lds_blocks_per_cu(dynLds, engine);
        case 8: return SlotKernels<8>::lds_blocks_per_cu(dynLds, engine);
        default: return SlotKernels<16>::lds_blocks_per_cu(dynLds, engine);
    }
}

void prepare_wave_kernels(size_t dynLds) {
    SlotKernels<4>::prepare(dynLds);
    SlotKernels<8>::prepare(dynLds);
    SlotKernels<16>::prepare(dynLds);
}

void launch_copy(const void* src, void* dst, int64_t bytes, hipStream_t s) {
    const int64_t n16 = bytes / 16;
    hipLaunchKernelGGL(k_copy_u32, dim3(2048), dim3(256), 0, s, static_cast<const uint4*>(src),
                       static_cast<uint4*>(dst), n16);
}

static int fill_blocks(int64_t n) {
    const int64_t b = (n + 255) / 256;
    return (int)(b < 1 ? 1 : (b > 1024 ? 1024 : b));
}
void launch_fill_u64(unsigned long long* p, int64_t n, unsigned long long v, hipStream_t s) {
    hipLaunchKernelGGL(k_fill_u64, dim3(fill_blocks(n)), dim3(256), 0, s, p, n, v);
}
void launch_fill_u32(uint32_t* p, int64_t n, uint32_t v, hipStream_t s) {
    hipLaunchKernelGGL(k_fill_u32, dim3(fill_blocks(n)), dim3(256), 0, s, p, n, v);
}
void launch_zero(const ZeroRanges& z, hipStream_t s) {
    if (z.count <= 0) return;
    int64_t most = 1;
    for (int r = 0; r < z.count; ++r) most = z.n[r] > most ? z.n[r] : most;
    hipLaunchKernelGGL(k_zero_ranges, dim3(fill_blocks(most), z.count), dim3(256), 0, s, z);
}

}  // namespace gr
