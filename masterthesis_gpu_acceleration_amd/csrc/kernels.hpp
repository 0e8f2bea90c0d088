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

// Dedup strategy of the body-pair wave-mode kernels.
enum DedupMode : int32_t {
    kDedupRows = 0,   // per-wave bitmap row of nE bits in HBM (atomicOr)
    kDedupSort = 1,   // keys (seed << hoBits | head object) emitted to HBM, then radix sort + unique
    kDedupLds = 2,    // per-wave bitmap of nE bits in LDS (dynamic shared memory)
};

// Emission target of the sort-unique path: a wave appends the head objects it
// finds for heavy seed h as keys (h << hoBits) | object. `count` keeps
// counting past `cap` (nothing is written there), so the host can tell an
// overflow and rerun with a larger buffer.
struct SortSink {
    unsigned long long* keys;
    uint32_t* count;
    uint32_t cap;
    int32_t hoBits;
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
    uint32_t* rows;               // per-wave dedup rows, all zero between seeds (kDedupRows)
    int64_t rowWords;
    SortSink sink;                // kDedupSort
};

// `kind` is the plan kind (kSupport / kPairs) of every plan the launch sees.
void launch_thread_batch(int kind, const ThreadArgs& a, const DevGraph& G, int blocks, hipStream_t s);
void launch_thread_one(int kind, const Plan& P, const OneArgs& a, const DevGraph& G, int blocks, hipStream_t s);
// `dedup` (DedupMode) matters for kPairs only.
void launch_wave_batch(int kind, int dedup, const WaveArgs& a, const DevGraph& G, int blocks, hipStream_t s);
void launch_wave_one(int kind, int dedup, const Plan& P, const WaveArgs& a, const DevGraph& G, int blocks,
                     hipStream_t s);
// Static LDS of a wave-mode block (bytes), and the dynamic-LDS limit to
// request for the kDedupLds kernels.
size_t wave_static_lds();
void prepare_wave_kernels(size_t dynLds);
void launch_copy(const void* src, void* dst, int64_t bytes, hipStream_t s);
void launch_fill_u64(unsigned long long* p, int64_t n, unsigned long long v, hipStream_t s);
void launch_fill_u32(uint32_t* p, int64_t n, uint32_t v, hipStream_t s);

}  // namespace gr
