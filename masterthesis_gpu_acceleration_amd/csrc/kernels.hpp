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
    int32_t fuse;                 // PAIRS: plans with supPred != -1 also count support (into [2*rule + kSupport])
    const int32_t* supTab;        // member lists of plans with supPred <= -2 (Plan::supPred)
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
    kDedupRows = 0,   // per-wave bitmap row in HBM over head-object ranks (atomicOr)
    kDedupSort = 1,   // keys (seed << hoBits | head object) emitted to HBM, then radix sort + unique
    kDedupLds = 2,    // the same bitmap per wave in LDS (dynamic shared memory)
};

// Emission target of the sort-unique path: a wave appends the head objects it
// finds for heavy seed h as keys (h << hoBits) | object. `count` (64-bit, so
// it cannot wrap) keeps counting past `cap` (nothing is written there), so
// the host can tell an overflow and rerun with a larger buffer or fail.
struct SortSink {
    unsigned long long* keys;
    unsigned long long* count;
    uint32_t cap;
    int32_t hoBits;
};

// Giant seeds. A wave-mode seed still running after `after` steps is
// abandoned (its partial count dropped) and listed; an expansion kernel
// cuts each listed seed into sub-items -- ranges of its level-0 row, or of
// its level-1 rows when the level-0 row is short, each with the bindings
// above -- and a sub-item kernel runs them on all waves. Sub-items of one
// giant count against one HBM bitmap row (PAIRS: marked directly, the bits an
// atomicOr sets first counting each head object once, or, with LDS bitmaps,
// ORed in from the sub-item's own bitmap when it ends and counted whole by
// the giant's last sub-item) or share one found
// flag (SUPPORT); the sub-item that finishes last clears the row. Nothing
// waits on another wave.
struct alignas(16) SubItem {
    uint32_t giant;      // index in the giant list
    uint32_t plan;       // plan index (wave-order table)
    int32_t level;       // entries [begin, end) of this level
    int32_t begin;
    int32_t end;
    uint32_t heavy;      // heavy-queue index of the seed (sort-path keys)
    int32_t pad[2];
    int32_t v[kMaxSlots];   // bindings above `level`
};

constexpr int kSubsPerGiant = 512;   // sub-items one giant is cut into, at most
constexpr int kSubMinRow = 64;       // a level-0 row at least this long is cut at level 0

struct GiantArgs {
    int32_t after;           // wave-mode steps before a seed is listed as a giant (0: never)
    uint32_t cap;            // giants listed at most (more are run to the end as usual)
    HeavyItem* list;         // [cap] giant seeds (plan, item)
    uint32_t* heavyIdx;      // [cap] their heavy-queue index
    uint32_t* count;         // giants listed (may pass cap)
    SubItem* subs;           // [cap * kSubsPerGiant]
    uint32_t* subCount;
    uint32_t* subCursor;
    uint32_t* remaining;     // [cap] sub-items of each giant not yet finished
    uint32_t* found;         // [cap] SUPPORT: a sub-item found a grounding
    uint32_t* rows;          // [cap] HBM bitmap rows (rowWords each), all zero between uses
    unsigned long long* stats;   // kWaveStats counters of the launch (wave and sub-item kernels), or null
};

// Wave-mode counters of one launch: seeds and sub-items run, loop steps,
// donations, pool draws at the last level, pool draws above it, lanes busy
// with their own search at the end of a step and lanes that took pool
// entries in a step (both summed over steps); then kProfSlots cycle sums of
// the profiling build (kernels compiled with GR_STEP_PROFILE; zero
// otherwise): taking a seed and its plan, setting the seed up, pool draws at
// the last level, other pool draws, own-search steps, the end of a step
// (donation), the end of a seed (fused support, bitmap clear), sub-items.
constexpr int kProfBase = 7;
constexpr int kProfSlots = 8;
constexpr int kWaveStats = kProfBase + kProfSlots;

// Wave-mode pass over the heavy queue.
struct WaveArgs {
    const Plan* plans;            // wave-order plans (unused by the by-value kernel)
    const ConflictRefs* refs;
    const HeavyItem* heavy;
    const uint32_t* heavyCount;
    uint32_t heavyCap;
    uint32_t fuse;                // batch PAIRS launches: plans with supPred != -1 also count support (bitmap dedup only)
    uint32_t* cursor;
    unsigned long long* counts;   // batch: [2*rule + kind]; by-value kernel: counts[0]
    uint32_t* rows;               // per-wave dedup rows, all zero between seeds (kDedupRows)
    int64_t rowWords;
    SortSink sink;                // kDedupSort
    GiantArgs giant;              // batch launches; after == 0 for per-rule launches
    const int32_t* supTab;        // member lists of plans with supPred <= -2 (batch launches)
    Plan* planOut;                // per-rule launches with giants: the by-value kernel stores its plan here
                                  // (plans[0] of the giant kernels)
    int32_t engine;               // wave-mode search: 0 pool of entries, 1 level-synchronous (search.hpp)
};

// `kind` is the plan kind (kSupport / kPairs) of every plan the launch sees;
// `slots` is at least plan_slots() of each of them (the kernels of the
// smallest slot class holding it run).
void launch_thread_batch(int kind, int slots, const ThreadArgs& a, const DevGraph& G, int blocks, hipStream_t s);
void launch_thread_one(int kind, int slots, const Plan& P, const OneArgs& a, const DevGraph& G, int blocks,
                       hipStream_t s);
// `dedup` (DedupMode) matters for kPairs only.
void launch_wave_batch(int kind, int slots, int dedup, const WaveArgs& a, const DevGraph& G, int blocks,
                       hipStream_t s);
void launch_wave_one(int kind, int slots, int dedup, const Plan& P, const WaveArgs& a, const DevGraph& G, int blocks,
                     hipStream_t s);
// Static LDS of a kDedupLds wave-mode block of the slot class for `slots`
// (bytes, from the runtime; 16: the largest class), and the dynamic-LDS
// limit to request for the kDedupLds kernels.
size_t wave_static_lds(int slots = 16, int engine = 0);
int wave_lds_blocks_per_cu(int slots, size_t dynLds, int engine = 0);
void prepare_wave_kernels(size_t dynLds);
// The giant-seed pass that follows a wave-mode batch launch: expansion of
// the listed giants, then their sub-items (same `a` as the wave launch).
void launch_giants(int kind, int slots, int dedup, const WaveArgs& a, const DevGraph& G, int blocks, hipStream_t s);

// Launchers of one slot class (kernels with NS variable slots per lane),
// defined in search.hpp and instantiated once per class.
template <int NS>
struct SlotKernels {
    static void thread_batch(int kind, const ThreadArgs& a, const DevGraph& G, int blocks, hipStream_t s);
    static void thread_one(int kind, const Plan& P, const OneArgs& a, const DevGraph& G, int blocks, hipStream_t s);
    static void wave_batch(int kind, int dedup, const WaveArgs& a, const DevGraph& G, int blocks, hipStream_t s);
    static void wave_one(int kind, int dedup, const Plan& P, const WaveArgs& a, const DevGraph& G, int blocks,
                         hipStream_t s);
    static void giants(int kind, int dedup, const WaveArgs& a, const DevGraph& G, int blocks, hipStream_t s);
    static void prepare(size_t dynLds);
    static size_t lds_static(int engine);   // static LDS of the kDedupLds wave kernel (bytes)
    // resident one-wave blocks per CU of the kDedupLds pairs wave kernel with
    // `dynLds` bytes of dynamic LDS (the runtime's occupancy query), 0 on error
    static int lds_blocks_per_cu(size_t dynLds, int engine);
};
extern template struct SlotKernels<4>;
extern template struct SlotKernels<8>;
extern template struct SlotKernels<16>;

// Word ranges zeroed by one launch (a pass's counters and flags in one
// kernel instead of one fill each).
constexpr int kZeroRanges = 12;
struct ZeroRanges {
    uint32_t* p[kZeroRanges];
    int64_t n[kZeroRanges];   // 32-bit words
    int32_t count;
};
// Appends a range; false when the list is full.
inline bool zero_add(ZeroRanges& z, void* p, int64_t words) {
    if (z.count >= kZeroRanges) return false;
    z.p[z.count] = static_cast<uint32_t*>(p);
    z.n[z.count] = words;
    ++z.count;
    return true;
}
void launch_zero(const ZeroRanges& z, hipStream_t s);

void launch_copy(const void* src, void* dst, int64_t bytes, hipStream_t s);
void launch_fill_u64(unsigned long long* p, int64_t n, unsigned long long v, hipStream_t s);
void launch_fill_u32(uint32_t* p, int64_t n, uint32_t v, hipStream_t s);

}  // namespace gr
