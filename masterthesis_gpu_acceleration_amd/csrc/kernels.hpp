// kernels.hpp -- argument blocks and host launchers of the counting kernels.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

#include "plan.hpp"

namespace gr {

// Persistent thread-mode pass over a chunk list (batched evaluation).
struct ThreadArgs {
    const Plan* plans;
    const ConflictRefs* refs;     // same indexing as plans
    const Chunk* chunks;
    uint32_t nChunks;
    uint32_t heavyCap;
    uint32_t* cursor;
    unsigned long long* counts;   // [2*rule + kind]
    HeavyItem* heavy;
    uint32_t* heavyCount;
};

// Thread-mode pass over all seeds of one plan given by value.
struct OneArgs {
    const ConflictRefs* refs;
    uint32_t nItems;
    uint32_t heavyCap;
    unsigned long long* count;
    HeavyItem* heavy;
    uint32_t* heavyCount;
};

// Wave-mode pass over the heavy queue.
struct WaveArgs {
    const Plan* plans;            // wave-order plans (unused by the by-value kernel)
    const ConflictRefs* refs;
    const HeavyItem* heavy;
    const uint32_t* heavyCount;
    uint32_t heavyCap;
    uint32_t pad;
    uint32_t* cursor;
    unsigned long long* counts;   // batch: [2*rule + kind]; by-value kernel: counts[0]
    uint32_t* rows;               // per-wave dedup rows, all zero between seeds
    int64_t rowWords;
};

// `kind` is the plan kind (kSupport / kPairs) of every plan the launch sees.
void launch_thread_batch(int kind, const ThreadArgs& a, const DevGraph& G, int blocks, hipStream_t s);
void launch_thread_one(int kind, const Plan& P, const OneArgs& a, const DevGraph& G, int blocks, hipStream_t s);
void launch_wave_batch(int kind, const WaveArgs& a, const DevGraph& G, int blocks, hipStream_t s);
void launch_wave_one(int kind, const Plan& P, const WaveArgs& a, const DevGraph& G, int blocks, hipStream_t s);
void launch_copy(const void* src, void* dst, int64_t bytes, hipStream_t s);
void launch_fill_u64(unsigned long long* p, int64_t n, unsigned long long v, hipStream_t s);
void launch_fill_u32(uint32_t* p, int64_t n, uint32_t v, hipStream_t s);

}  // namespace gr
