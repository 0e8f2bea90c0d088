// engine_internal.hpp -- library-internal types shared by the C ABI
// (engine.cpp) and the serving loop (serve.cpp).
#pragma once
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <cstdio>
#include <memory>
#include <vector>

#include "../../include/gpu_rules_api.h"
#include "graph_host.hpp"
#include "kernels.hpp"
#include "planner.hpp"
#include "rules_host.hpp"
#include "sort_unique.hpp"

namespace gr {


constexpr uint32_t kMagicHost = 0x47524831;  // "GRH1"
constexpr uint32_t kMagicCtx = 0x47524331;   // "GRC1"
constexpr uint32_t kMagicBatch = 0x47524231; // "GRB1"

// Per-rule launch header (the context's own; the allocation-free entry
// point's is kNoAllocHeader bytes and stops before the plan).
constexpr size_t kOnePlanOff = 512;
constexpr size_t kOneBytes = 1024;
static_assert(kOnePlanOff + sizeof(Plan) <= kOneBytes, "plan beyond the per-rule header");

inline void log_err(const char* what, hipError_t e) {
    std::fprintf(stderr, "[gpu_rules] %s: %s\n", what, hipGetErrorString(e));
}

#define GR_HIP(call)                                   \
    do {                                               \
        hipError_t e_ = (call);                        \
        if (e_ != hipSuccess) {                        \
            log_err(#call, e_);                        \
            return false;                              \
        }                                              \
    } while (0)

struct Handle {
    uint32_t magic;
};

struct HostHandle : Handle {
    HostGraph g;
    std::unique_ptr<TermResolver> tr;
};

// Device buffer that only grows.
//   ensure: hipMalloc / hipFree (hipFree waits for the whole device);
//   ensure_on: stream-ordered (hipMallocAsync / hipFreeAsync on stream st),
//     at least doubling, so a buffer of one stream -- a serving slot's batch --
//     grows without waiting for work on other streams. Such a buffer is
//     freed on the same stream, which must outlive it (release first).
struct DevBuf {
    void* p = nullptr;
    size_t bytes = 0;
    hipStream_t ordered = nullptr;   // stream of a stream-ordered allocation
    bool ensure(size_t need) {
        if (need <= bytes) return true;
        release();
        size_t b = std::max<size_t>(need, 256);
        hipError_t e = hipMalloc(&p, b);
        if (e != hipSuccess) {
            log_err("hipMalloc", e);
            p = nullptr;
            return false;
        }
        bytes = b;
        return true;
    }
    bool ensure_on(size_t need, hipStream_t st) {
        if (need <= bytes && (ordered == st || !p)) return true;
        const size_t b = std::max<size_t>({need, bytes * 2, 256});
        release();
        hipError_t e = hipMallocAsync(&p, b, st);
        if (e != hipSuccess) {
            log_err("hipMallocAsync", e);
            p = nullptr;
            return false;
        }
        bytes = b;
        ordered = st;
        return true;
    }
    void release() {
        if (p) {
            if (ordered) (void)hipFreeAsync(p, ordered);
            else (void)hipFree(p);
        }
        p = nullptr;
        bytes = 0;
        ordered = nullptr;
    }
    template <class T> T* as() const { return static_cast<T*>(p); }
};

// Buffers of the sort-unique dedup path: two key arrays (LSD ping-pong) of
// `cap` keys each and the radix histogram. The capacity is a power of two;
// it doubles (at least) whenever a pass emitted more keys than it holds, up
// to kSortCapMax keys (32 GiB of key arrays -- sized for a 288 GB device;
// key positions stay below 2^31, so the sort kernels index in 32 bits).
constexpr uint32_t kSortCapMin = 1u << 22;
constexpr uint32_t kSortCapMax = 1u << 31;
struct SortBufs {
    DevBuf keys[2], hist;
    uint32_t cap = 0;
    // Capacity for at least `want` keys (clamped to capMax); the first
    // allocation holds at least `first` keys.
    bool ensure(uint64_t want, uint64_t first = kSortCapMin, uint64_t capMax = kSortCapMax) {
        uint64_t c = cap ? cap : kSortTile;
        while ((c < want || (!cap && c < first)) && c < capMax) c <<= 1;
        if (c == cap) return true;
        if (!keys[0].ensure(c * 8) || !keys[1].ensure(c * 8) || !hist.ensure(sort_hist_words((uint32_t)c) * 4)) {
            release();
            return false;
        }
        cap = (uint32_t)c;
        return true;
    }
    void release() {
        keys[0].release();
        keys[1].release();
        hist.release();
        cap = 0;
    }
    size_t bytes() const { return keys[0].bytes + keys[1].bytes + hist.bytes; }
};

// Buffers of the giant-seed pass for one execution target (the context or
// a serving slot): the giant list, sub-items, per-giant counters and found
// flags, and per-giant HBM bitmap rows (all zero between uses: the last
// sub-item of a giant clears its row). Counters, remaining counts and flags
// are zeroed by fill kernels before each wave-mode phase.
// Giants listed per wave-mode phase: up to kGiantCapMax (sub-items 96 B x
// 512 per giant, 201 MB at the maximum), fewer when their bitmap rows would
// pass kGiantRowBudget, never fewer than kGiantCapMin.
// A seed past the step cap when the list is full runs to its end as usual.
constexpr uint32_t kGiantCapMin = 512;
constexpr uint32_t kGiantCapMax = 4096;
constexpr size_t kGiantRowBudget = (size_t)256 << 20;
struct GiantBufs {
    DevBuf list, heavyIdx, subs, ctl, remaining, found, rows, stats;
    int64_t rowWords = 0;
    uint32_t cap = 0;
    bool ensure(int64_t words) {
        words = std::max<int64_t>(words, 1);
        if (rows.p && words == rowWords) return true;
        release();
        const size_t perRow = (size_t)words * 4;
        const size_t fit = kGiantRowBudget / perRow;
        const uint32_t c = (uint32_t)std::min<size_t>(kGiantCapMax, std::max<size_t>(kGiantCapMin, fit));
        const size_t rowBytes = (size_t)c * perRow;
        if (!list.ensure(c * sizeof(HeavyItem)) || !heavyIdx.ensure(c * 4) ||
            !subs.ensure((size_t)c * kSubsPerGiant * sizeof(SubItem)) || !ctl.ensure(16) ||
            !remaining.ensure(c * 4) || !found.ensure(c * 4) || !rows.ensure(rowBytes) ||
            !stats.ensure(2 * kWaveStats * 8))
            return false;
        cap = c;
        if (hipMemset(rows.p, 0, rowBytes) != hipSuccess) return false;
        rowWords = words;
        return true;
    }
    void release() {
        list.release(); heavyIdx.release(); subs.release(); ctl.release(); remaining.release(); found.release();
        rows.release(); stats.release();
        rowWords = 0;
        cap = 0;
    }
    size_t bytes() const {
        return list.bytes + heavyIdx.bytes + subs.bytes + ctl.bytes + remaining.bytes + found.bytes + rows.bytes;
    }
};

struct Context;

// What the host does with a pass's device counters (decided apart from the
// device code, so the decisions can be checked with forced values):
//   sort_count_action: keys emitted vs the sort buffers' capacity and the
//     largest capacity allowed -- counted, regrow and run again, or fail;
//   heavy_count_action: seeds escalated vs the heavy queue's capacity (the
//     queue holds every seed of the phase, so a larger count means a broken
//     pass) -- counted or fail.
enum CountAction : int32_t { kCountsOk = 0, kCountsRegrow = 1, kCountsFail = 2 };
inline int32_t sort_count_action(uint64_t emitted, uint64_t cap, uint64_t capMax) {
    if (emitted <= cap) return kCountsOk;
    return cap < capMax ? kCountsRegrow : kCountsFail;
}
inline int32_t heavy_count_action(uint64_t escalated, uint64_t cap) {
    return escalated <= cap ? kCountsOk : kCountsFail;
}
// Seeds one phase of a batch may hold: the heavy queue's count, cursor and
// capacity are 32-bit (a larger batch is refused by batch_build).
constexpr uint64_t kMaxPhaseSeeds = 0xFFFFFFF0ull;

// First error of a sequence of asynchronous HIP calls.
struct HipFirst {
    hipError_t e = hipSuccess;
    const char* what = nullptr;
    void operator()(hipError_t x, const char* w) {
        if (e == hipSuccess && x != hipSuccess) {
            e = x;
            what = w;
        }
    }
    bool ok() const { return e == hipSuccess; }
};

// Support thresholds of a gated batch: body pairs are searched only for rules
// whose support reaches minSupport and whose head coverage (support /
// headSize, the IEEE quotient the results report) reaches minHeadCoverage --
// except where the body-pair pass itself counts the support (fused rules),
// whose body pairs are then reported only if the rule meets them. The
// defaults (1, 0) are the gate of every gated batch: support > 0.
struct Thresholds {
    int64_t minSupport = 1;
    double minHeadCoverage = 0.0;
};

// The smallest support that meets t for a rule of headSize head triples, at
// least 1, clamped to kGateMax (support never exceeds the triple count).
constexpr uint32_t kGateMax = 0x7fffffffu;
uint32_t support_threshold(int64_t headSize, const Thresholds& t);

// A compiled, device-resident batch of rules.
struct Batch : Handle {
    Context* ctx = nullptr;
    int32_t n = 0;
    int32_t flags = 0;
    Thresholds thr;                 // gated batches: the gate of the pairs phase
    bool seedsOverflow = false;     // batch_build refused: a phase's seeds exceed the context's cap
    std::vector<uint32_t> ruleThr;  // per rule: support_threshold (gated batches)
    std::vector<CompiledRule> comp;
    std::vector<int32_t> status;
    // chunk lists and heavy-queue segments per phase (0 support, 1 pairs)
    uint32_t nChunks[2] = {};
    uint32_t chunkOff[2] = {};
    uint32_t heavyCap[2] = {};
    uint64_t heavyOff[2] = {};
    // support chunks of rules whose support a fused pairs launch counts
    // (CompiledRule::supPred); they end the support chunk list, so such
    // a launch runs only the first nChunks[0] - nFusable
    uint32_t nFusable = 0;
    int64_t nFused = 0;   // rules with a fusable support (pairs plans with supPred >= 0)
    // per rule: the rule whose body-pair search it shares (same body up to
    // renaming), or -1; nShared rules have one
    std::vector<int32_t> pairsFrom;
    int64_t nShared = 0;
    int64_t nSearches[2] = {0, 0};   // plans launched per phase (support searches, body-pair searches)
    uint32_t capRules = 0;
    int slots[2] = {1, 1};   // variable slots of the plans of each phase (picks the kernels' slot class)
    double kernelMs[4] = {};
    DevBuf plans, wplans, refs, wrefs, chunks, counts, heavy, ctl;
    DevBuf supTab;   // member lists of shared searches that count several rules' support (Plan::supPred)
    unsigned long long* hCounts = nullptr;   // pinned
    size_t hCountsCap = 0;                   // words
    uint32_t* hCtl = nullptr;                // pinned copy of ctl (kCtlWords words)
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    const void* capturedDedup = nullptr;   // context scratch the graph was captured with
    const void* capturedSort = nullptr;
    uint32_t capturedSortCap = 0;
    int capturedMode = -1;
    const void* capturedGiants = nullptr;
    int capturedGiantAfter = -1;
    int capturedEngine = -1;
    int capturedOverlap = -1;
    int capturedFuse = -1;
    int capturedBlocks[2] = {0, 0};
    double lastMs = 0, lastSupMs = 0, lastPairMs = 0, lastSortMs = 0;
    int64_t lastHeavy = 0;
    int64_t lastEmitted = 0;   // sort path: keys the last pass emitted
    uint32_t* hGiant = nullptr;   // pinned: giants listed in each phase of the last pass
    unsigned long long* hWaveStats = nullptr;   // pinned: kWaveStats counters per phase of the last pass
    ~Batch();
    // Frees the device buffers (stream-ordered ones on the stream they were
    // built on, which must still exist).
    void free_device();
};

struct Context : Handle {
    HostGraph g;
    std::unique_ptr<TermResolver> tr;
    Tuning tune;
    int device = 0;
    int numCUs = 256;
    int threadWavesPerCU = 16;
    int waveWavesPerCU = 8;
    bool useGraph = true;
    int dedupMode = 0;        // body-pair dedup: 0 auto, 1 HBM bitmap rows, 2 sort-unique, 3 LDS bitmap
    bool rowsFit = true;      // HBM bitmap rows fit kRowBudget at one wave per CU
    bool ldsFits = false;     // one head-object bitmap (rowWords words) fits a wave-mode block's LDS
    bool ldsPreferred = false;  // ... and still leaves room for kLdsMinWaves wave-mode waves per CU
    size_t ldsPerBlock = 65536, ldsPerCU = 65536;
    hipStream_t st = nullptr;
    hipEvent_t ev[8] = {};
    DevBuf blocks, rowStart, col, coo, keys, crank;
    DevGraph dg{};
    // per-rule launch scratch
    DevBuf one;       // kOneBytes: the per-rule launch header (counts, control words, ConflictRefs, plan)
    DevBuf heavy1;
    uint32_t heavy1Cap = 0;
    unsigned long long* hOne = nullptr;   // pinned
    // wave-mode dedup rows
    DevBuf dedup;
    int64_t rowWords = 0;
    int waveSlots = 0;
    SortBufs sortb;   // sort-unique dedup buffers (allocated when that path is used)
    uint32_t sortCapMax = kSortCapMax;   // largest key capacity they may grow to ("sort_cap_max")
    bool noallocReady = false;           // the allocation-free entry point's device scratch is mapped
    uint64_t noallocMaxSeeds = ~0ull;    // its seeds-per-launch limit ("noalloc_max_seeds"; capped by the scratch)
    GiantBufs giants; // giant-seed pass
    // Ungated batches replayed as graphs run the support and body-pair phases
    // side by side: the pairs phase on st2 (forked from st by forkEv, joined
    // back by joinEv) with its own giant buffers.
    bool overlap = true;
    hipStream_t st2 = nullptr;
    hipEvent_t forkEv = nullptr, joinEv = nullptr;
    GiantBufs giants2;
    int giantAfter = 2048;   // wave-mode steps on one seed before it is listed as a giant (0: never)
    int waveEngine = 1;      // wave-mode search: 0 pool of entries, 1 level-synchronous (WaveArgs::engine)
    // Serving: requests whose estimated search work (estimate_work) passes
    // serveHeavyWork run in batches of their own, one such batch in flight at
    // a time, on serveHeavyPercent of the grid, so that light requests are
    // neither batched with them nor kept off the device by them (0: off).
    double serveHeavyWork = 1e11;
    int serveHeavyPercent = 50;
    RowStats rowStats;       // of g, for the estimates
    // Per-rule launches cut giant seeds into sub-items only for rules of at
    // least this estimated work (0: every rule).
    double perRuleGiantWork = 1e7;
    uint64_t phaseSeedCap = kMaxPhaseSeeds;   // seeds per batch phase (tuning key phase_seed_cap)
    // Count the support of qualifying rules in the body-pair pass instead of
    // a support pass of their own; needs bitmap dedup (rows or LDS).
    bool fuseSupport = true;
    // Thread-mode work units per thread-mode wave a batch aims for (smaller
    // units when the seeds are few): more units even out the waves' ends.
    int chunksPerWave = 2;
    // Rules with the same body (up to renaming, head variables fixed) share
    // one body-pair search.
    bool shareBodies = true;
    bool fuseActive(int mode) const { return fuseSupport && (mode == kDedupRows || mode == kDedupLds); }
    int64_t lastGiants = 0;
    int64_t lastFused = 0;   // rules whose support the last batch counted in its pairs pass
    int64_t lastShared = 0;  // rules of the last batch that took another rule's body-pair count
    int64_t lastSearches[2] = {0, 0};   // support / body-pair plans the last batch launched
    // per rule of the last GPU_RULES_PER_RULE evaluation: device time of its
    // support launches and of its body-size launches (ms; 0 where none ran)
    std::vector<double> lastRuleSupMs, lastRuleConfMs;
    unsigned long long lastWaveStats[2 * kWaveStats] = {};
    Batch* lastBatch = nullptr;
    double lastMs = 0, lastSupMs = 0, lastPairMs = 0, lastSortMs = 0;
    double lastKernelMs[4] = {};
    int64_t lastHeavy = 0;
    int64_t lastEmitted = 0;

    int waveBlocks() const { return std::max(1, waveSlots); }   // one wave per wave-mode block
    // Wave-mode blocks of a launch whose plans need `slots` (plan_slots):
    // with LDS bitmaps, as many as fit a CU's LDS for that slot class
    // (smaller classes keep fewer frames), else waveBlocks().
    int waveBlocksFor(int slots) const;
    int threadBlocks() const { return std::max(1, numCUs * threadWavesPerCU / 4); }
    int sortBlocks() const { return std::max(1, numCUs * 4); }
    bool sortActive() const { return dedupMode == 2 || (dedupMode == 0 && !rowsFit && !ldsFits); }
    // DedupMode of the body-pair wave-mode kernels
    int pairsDedup() const {
        if (sortActive()) return kDedupSort;
        if ((dedupMode == 0 && ldsPreferred) || (dedupMode == 3 && ldsFits)) return kDedupLds;
        return kDedupRows;
    }

    bool upload();
    bool ensure_dedup();
    ~Context();
};

inline void Batch::free_device() {
    plans.release(); wplans.release(); refs.release(); wrefs.release(); chunks.release(); counts.release();
    heavy.release(); ctl.release(); supTab.release();
}

inline Batch::~Batch() {
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
    free_device();
    if (hCounts) (void)hipHostFree(hCounts);
    if (hCtl) (void)hipHostFree(hCtl);
    if (hGiant) (void)hipHostFree(hGiant);
    if (hWaveStats) (void)hipHostFree(hWaveStats);
    if (ctx && ctx->lastBatch == this) ctx->lastBatch = nullptr;
}

inline Context::~Context() {
    if (st) (void)hipStreamSynchronize(st);
    if (st2) (void)hipStreamSynchronize(st2);
    blocks.release(); rowStart.release(); col.release(); coo.release(); keys.release(); crank.release();
    one.release(); heavy1.release(); dedup.release(); sortb.release(); giants.release(); giants2.release();
    if (hOne) (void)hipHostFree(hOne);
    for (auto& e : ev) if (e) (void)hipEventDestroy(e);
    if (forkEv) (void)hipEventDestroy(forkEv);
    if (joinEv) (void)hipEventDestroy(joinEv);
    if (st2) (void)hipStreamDestroy(st2);
    if (st) (void)hipStreamDestroy(st);
}

template <class T>
inline bool put(DevBuf& b, const std::vector<T>& v, hipStream_t st) {
    if (!b.ensure(std::max<size_t>(v.size(), 1) * sizeof(T))) return false;
    if (!v.empty()) GR_HIP(hipMemcpyAsync(b.p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice, st));
    return true;
}

// The same into a stream-ordered buffer of st.
template <class T>
inline bool put_on(DevBuf& b, const std::vector<T>& v, hipStream_t st) {
    if (!b.ensure_on(std::max<size_t>(v.size(), 1) * sizeof(T), st)) return false;
    if (!v.empty()) GR_HIP(hipMemcpyAsync(b.p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice, st));
    return true;
}

// Where a batch's device pass runs: stream, wave-mode dedup rows or (when
// `sort` is set) the sort-unique buffers, and (optionally) timing events
// ev[0..4] around the four kernels, ev[6] before the sort-unique pass.
struct ExecTarget {
    hipStream_t st;
    uint32_t* dedup;
    int64_t rowWords;
    int waveBlocks;
    hipEvent_t* ev;   // nullptr: untimed
    SortBufs* sort;   // the sort-unique buffers (pairsDedup == kDedupSort)
    int sortBlocks;
    int pairsDedup;   // DedupMode of the body-pair wave-mode kernel
    GiantBufs* giants;
    int giantAfter;
    int gridPercent = 100;   // share of the thread-mode and wave-mode grids the pass launches
    // optional: run the pairs phase of an ungated batch on st2 beside the
    // support phase (untimed passes only), with giants2 as its giant buffers
    hipStream_t st2 = nullptr;
    hipEvent_t forkEv = nullptr, joinEv = nullptr;
    GiantBufs* giants2 = nullptr;
};

// `percent` of a grid of `blocks` blocks, at least one.
inline int scale_grid(int blocks, int percent) {
    return percent >= 100 ? blocks : std::max(1, (int)((int64_t)blocks * percent / 100));
}

// Zeroes the giant pass's counters and flags for one wave-mode phase and
// fills a GiantArgs for it (after == 0: no seed is listed).
void giant_prepare(GiantBufs& gb, int after, GiantArgs& ga, hipStream_t s, int phase);
// The same without the launch: fills ga and appends the words to zero to z.
void giant_args(GiantBufs& gb, int after, GiantArgs& ga, int phase, ZeroRanges& z);

// Bytes of bitmap rows allowed before body-pair dedup switches to the
// sort-unique path (auto mode).
constexpr int64_t kRowBudget = int64_t(8) << 30;
// Auto mode keeps head-object bitmaps in LDS only while that leaves at
// least this many wave-mode waves per CU: wave mode is latency-bound, and on
// the 500k-term graph LDS bitmaps (2 waves per CU) ran 2.4x slower than HBM
// rows (16 waves per CU).
constexpr int kLdsMinWaves = 8;

// Control words of a batch pass (device `ctl`, host copy `hCtl`): per phase
// ph, words 4*ph + {0 thread cursor, 1 heavy count, 2 wave cursor, 3 spare},
// then the phases' 64-bit emitted-key counters (sort path) at words 8..11.
constexpr int kCtlWords = 12;
inline unsigned long long* ctl_emitted(uint32_t* ctl, int ph) {
    return reinterpret_cast<unsigned long long*>(ctl + 8) + ph;
}
// Keys emitted by B's last pass (sort path), from its host control copy.
int64_t batch_emitted(const Batch& B);
// Escalated seeds of B's last pass within each phase's heavy queue?
bool batch_heavy_ok(const Batch& B);

// Compiles rules into B and uploads plans and chunk lists on stream st.
bool batch_build(Context& C, Batch& B, const std::vector<RuleIR>& rules, const std::vector<int32_t>& status,
                 int32_t flags, hipStream_t st);
// For each rule, the earlier rule with the same body (up to a renaming that
// fixes the head variables) whose body-pair count it can take, or -1.
std::vector<int32_t> shared_pairs(const std::vector<RuleIR>& rules, const std::vector<CompiledRule>& comp,
                                  const std::vector<int32_t>& status);
// Enqueues B's device pass (counter fills, kernels, result copies) on t.st.
// Returns the first error of the enqueued calls (logged), or hipSuccess.
hipError_t batch_enqueue(Context& C, Batch& B, const ExecTarget& t);
// Fills out[B.n] from B's host-side counters after its pass completed.
void batch_collect(Batch& B, GpuRuleResult64* out);
// Builds-if-needed, runs (graph replay or eager) and collects on the context's stream.
bool batch_run(Context& C, Batch& B, GpuRuleResult64* out);

}  // namespace gr
