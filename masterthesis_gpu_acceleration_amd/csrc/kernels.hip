// kernels.hip -- rule support / body-pair counting on gfx950.
//
// Work model
//   * A SEED is one head triple (SUPPORT plans) or one head-subject key
//     (PAIRS plans). Seeds are independent, so a seed can be abandoned and
//     redone elsewhere without changing any count.
//   * Every lane runs an iterative depth-first search over the plan's levels:
//     one uniform loop in which each lane advances its own frame (entry
//     cursor, end) at its own depth. Frames and the plan live in LDS, the
//     variable slots in registers, so the control flow stays flat and the
//     register count does not grow with the body length.
//   * Thread mode: one lane per seed, with a work budget; a wave works through
//     a chunk of up to kChunkSeeds seeds of one plan and a lane takes the next
//     seed as soon as its own is done. A seed that uses up its budget, or
//     (PAIRS) whose distinct head objects outgrow the lane's register set, is
//     appended to the heavy queue and its partial work discarded.
//   * Wave mode: one wave64 per heavy seed. Lanes take level-0 entries one at
//     a time from a shared LDS cursor and search below them on their own.
//     SUPPORT seeds stop as soon as __any shows a complete grounding; PAIRS
//     seeds mark head objects in a per-wave bitmap row; the words that were
//     set are listed in LDS, so counting (popcount) and clearing the row
//     touch only those words.
//
// Semantics, as the reference CPU counters define them
// (Rewritten_CPP/SupportCounting.cpp:299-479, CPU_01_nonparallel/
// SupportCounting.cpp:83-247): a new variable binding must differ from every
// variable bound before it; a ground triple may not repeat an earlier body
// triple or (SUPPORT) the head triple -- levels whose atoms could coincide
// carry a conflict mask computed by the planner.
#include <hip/hip_runtime.h>

#include <cstdint>

#include "kernels.hpp"
#include "plan.hpp"

namespace gr {

constexpr int kBlock = 256;
constexpr int kWavesPerBlock = kBlock / kWave;
constexpr int kPlanWords = (int)(sizeof(Plan) / 4);

constexpr int kTouchCap = 256;      // dedup-row words listed per wave-mode seed
constexpr int kOneChunkSeeds = 64;  // seeds per wave in per-rule launches (one rule fills the GPU)

struct Frame {
    int i;     // next entry
    int end;   // one past the last entry
};

// Per-wave LDS: the plan being evaluated, one frame per (level, lane), and
// the wave-mode seed state (level-0 cursor, touched dedup-row words).
struct WaveLds {
    Plan plan;
    Frame frame[kMaxBody][kWave];
    int next0;
    int end0;
    uint32_t nTouched;
    uint32_t pad;
    uint32_t touched[kTouchCap];
};

static_assert(sizeof(Plan) % 4 == 0, "plan must be word-sized");

__device__ __forceinline__ int lane_id() { return __lane_id(); }
__device__ __forceinline__ int wave_in_block() { return threadIdx.x / kWave; }

// Value of a term reference: slot r (r >= 0) or constant -(r+1).
__device__ __forceinline__ int term_value(const int (&v)[kMaxSlots], int r) {
    int x = -(r + 1);
#pragma unroll
    for (int j = 0; j < kMaxSlots; ++j) x = (r == j) ? v[j] : x;
    return x;
}

__device__ __forceinline__ void set_slot(int (&v)[kMaxSlots], int slot, int x) {
#pragma unroll
    for (int j = 0; j < kMaxSlots; ++j) v[j] = (slot == j) ? x : v[j];
}

// Is x already bound to one of the first nbound slots?
__device__ __forceinline__ bool taken(const int (&v)[kMaxSlots], int nbound, int x) {
    bool d = false;
#pragma unroll
    for (int j = 0; j < kMaxSlots; ++j) d |= (j < nbound) & (v[j] == x);
    return d;
}

// Index in col[lo,hi) holding `target` (rows are sorted), or -1.
__device__ __forceinline__ int find_in_row(const int32_t* __restrict__ col, int lo, int hi, int target) {
    int l = lo, h = hi;
    while (l < h) {
        const int m = (l + h) >> 1;
        if (col[m] < target) l = m + 1;
        else h = m;
    }
    return (l < hi && col[l] == target) ? l : -1;
}

// Would ground triple (s, o) of a level repeat an earlier one? Only rules
// with constants reach this; it reads the conflict table from memory.
__device__ __forceinline__ bool repeats(const ConflictRefs* __restrict__ R, const int (&v)[kMaxSlots], int mask, int s,
                                        int o) {
    for (int j = 0; j < kMaxBody; ++j)
        if ((mask & (1 << j)) && term_value(v, R->sref[j]) == s && term_value(v, R->oref[j]) == o) return true;
    return (mask & 256) && term_value(v, R->hs) == s && term_value(v, R->ho) == o;
}

// Distinct head objects of one seed, in registers (thread mode).
struct RegSet {
    static constexpr int kCap = 8;
    int e[kCap];
    int n;
    __device__ __forceinline__ bool seen(int x) const {
        bool f = false;
#pragma unroll
        for (int j = 0; j < kCap; ++j) f |= (j < n) & (e[j] == x);
        return f;
    }
    __device__ __forceinline__ bool insert(int x) {
        if (n == kCap) return false;
#pragma unroll
        for (int j = kCap - 1; j > 0; --j) e[j] = e[j - 1];
        e[0] = x;
        ++n;
        return true;
    }
};

// Distinct head objects of one seed, one bit per term id (wave mode). The
// load bypasses the vector L1 so it sees the other lanes' atomics. The lane
// whose atomic finds a word still zero lists that word in LDS.
struct BitRow {
    uint32_t* row;
    WaveLds* W;
    __device__ __forceinline__ bool seen(int x) const {
        const uint32_t w = __hip_atomic_load(row + (x >> 5), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        return (w >> (x & 31)) & 1u;
    }
    __device__ __forceinline__ bool insert(int x) {
        const uint32_t old = atomicOr(row + (x >> 5), 1u << (x & 31));
        if (old == 0) {
            const uint32_t k = atomicAdd(&W->nTouched, 1u);
            if (k < (uint32_t)kTouchCap) W->touched[k] = (uint32_t)(x >> 5);
        }
        return true;
    }
};

// Entry range of a level under the current bindings.
__device__ __forceinline__ Frame enter_level(const Level& lv, const DevGraph& G, const int (&v)[kMaxSlots]) {
    const int kind = level_kind(lv.flags);
    if (kind == kScan) return Frame{lv.lo, lv.hi};
    const int keyv = term_value(v, lv.key);
    const int64_t base = (int64_t)(uint32_t)lv.lo | ((int64_t)lv.hi << 32);
    const KeyBlock kb = G.blocks[base + (keyv >> 6)];
    const uint64_t bit = 1ull << (keyv & 63);
    if (!(kb.bits & bit)) return Frame{0, 0};
    const uint32_t r = kb.rank + (uint32_t)__popcll(kb.bits & (bit - 1));
    const int begin = G.rowStart[r];
    const int end = G.rowStart[r + 1];
    if (kind == kExt) return Frame{begin, end};
    const int at = find_in_row(G.col, begin, end, term_value(v, lv.val));
    return at < 0 ? Frame{0, 0} : Frame{at, at + 1};
}

// Binds entry i of a level. False if the candidate violates injectivity or
// repeats an earlier ground triple.
__device__ __forceinline__ bool bind_entry(const Level& lv, const DevGraph& G, const ConflictRefs* R,
                                           int (&v)[kMaxSlots], int i) {
    const int f = lv.flags;
    const int kind = level_kind(f);
    const int nbound = level_nbound(f);
    const int x = G.col[i];
    int k = 0;
    if (kind == kExt) {
        if (taken(v, nbound, x)) return false;
        set_slot(v, lv.val, x);
    } else if (kind == kScan) {
        k = G.coo[i];
        if (lv.val == lv.val2) {
            if (k != x || taken(v, nbound, k)) return false;
            set_slot(v, lv.val, k);
        } else {
            if (k == x || taken(v, nbound, k) || taken(v, nbound, x)) return false;
            set_slot(v, lv.val, k);
            set_slot(v, lv.val2, x);
        }
    }
    const int conf = level_conflicts(f);
    if (conf) {
        if (kind != kScan) k = term_value(v, lv.key);
        const bool spo = level_dir(f) == kSPO;
        if (repeats(R, v, conf, spo ? k : x, spo ? x : k)) return false;
    }
    return true;
}

// Binds the head of a SUPPORT seed; false if the head triple cannot match.
__device__ __forceinline__ bool bind_support_seed(const Plan& P, const DevGraph& G, int item, int (&v)[kMaxSlots]) {
    const int e = P.seedBegin + item;
    const int key = G.coo[e];
    const int val = G.col[e];
    const int s = P.seedDir == kSPO ? key : val;
    const int o = P.seedDir == kSPO ? val : key;
    if (P.hsRef < 0 && -(P.hsRef + 1) != s) return false;
    if (P.hoRef < 0 && -(P.hoRef + 1) != o) return false;
    if (P.hsRef >= 0 && P.hoRef >= 0) {
        if (P.hsRef == P.hoRef) {
            if (s != o) return false;
        } else if (s == o) {
            return false;
        }
    }
    if (P.hsRef >= 0) set_slot(v, P.hsRef, s);
    if (P.hoRef >= 0) set_slot(v, P.hoRef, o);
    return true;
}

// Binds a seed's head slots. False if the seed is not a match.
__device__ __forceinline__ bool bind_seed(const Plan& P, const DevGraph& G, int item, int (&v)[kMaxSlots]) {
#pragma unroll
    for (int j = 0; j < kMaxSlots; ++j) v[j] = -1;
    if (P.kind == kSupport) return bind_support_seed(P, G, item, v);
    v[0] = G.keys[P.seedBegin + item];
    return true;
}

__device__ __forceinline__ void copy_plan(WaveLds& W, const Plan* src) {
    const int32_t* s = reinterpret_cast<const int32_t*>(src);
    int32_t* d = reinterpret_cast<int32_t*>(&W.plan);
    for (int w = lane_id(); w < kPlanWords; w += kWave) d[w] = s[w];
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// ------------------------------------------------------------ helpers

__device__ __forceinline__ uint64_t wave_sum(uint64_t x) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) x += __shfl_xor(x, off, kWave);
    return x;
}

__device__ __forceinline__ uint32_t wave_take(uint32_t* cursor) {
    uint32_t c = 0;
    if (lane_id() == 0) c = atomicAdd(cursor, 1u);
    c = __shfl(c, 0, kWave);
    return __builtin_amdgcn_readfirstlane(c);
}

// Wave-aggregated append of this wave's escalated seeds.
__device__ __forceinline__ void push_heavy(bool heavy, uint32_t plan, uint32_t item, HeavyItem* q, uint32_t* qCount,
                                           uint32_t qCap) {
    const uint64_t mask = __ballot(heavy);
    if (mask == 0) return;
    uint32_t base = 0;
    if (lane_id() == 0) base = atomicAdd(qCount, (uint32_t)__popcll(mask));
    base = __shfl(base, 0, kWave);
    if (heavy) {
        const uint32_t pos = base + (uint32_t)__popcll(mask & ((1ull << lane_id()) - 1ull));
        if (pos < qCap) q[pos] = HeavyItem{plan, item};
    }
}

// Thread mode over seeds [begin, begin + count) of the plan in W. Every lane
// runs its own depth-first search; a lane whose seed is finished takes the
// next one (ballot prefix count), so no lane waits for the slowest seed.
// Returns this lane's count; escalated seeds go to the heavy queue. K is the
// plan kind of every plan the launch sees (kSupport or kPairs).
template <int K>
__device__ __forceinline__ uint64_t thread_chunk(WaveLds& W, const DevGraph& G, const ConflictRefs* R,
                                                 uint32_t planIdx, uint32_t begin, uint32_t count, HeavyItem* hq,
                                                 uint32_t* hCount, uint32_t hCap) {
    const int lane = lane_id();
    const Plan& P = W.plan;
    const int nlev = P.nlev;
    const bool pairs = K == kPairs && P.pairLevel >= 0;
    const int pl = P.pairLevel;
    const uint64_t below = (1ull << lane) - 1ull;
    int v[kMaxSlots];
    RegSet set;
    set.n = 0;
    int d = 0, budget = 0;
    bool active = false;
    uint32_t item = 0, next = 0;
    uint64_t mine = 0;
    for (;;) {
        const bool idle = !active;
        const uint64_t want = __ballot(idle);
        if (want && next < count) {
            const uint32_t off = next + (uint32_t)__popcll(want & below);
            if (idle && off < count) {
                item = begin + off;
                active = bind_seed(P, G, (int)item, v);
                if (active) {
                    W.frame[0][lane] = enter_level(P.lev[0], G, v);
                    d = 0;
                    budget = P.budget;
                    set.n = 0;
                }
            }
            next += (uint32_t)__popcll(want);
        }
        if (!__any(active)) {
            if (next >= count) break;
            continue;
        }
        bool heavyNow = false;
        if (active) {
            const Frame fr = W.frame[d][lane];
            if (fr.i >= fr.end) {
                if (d == 0) {
                    active = false;
                    if (pairs) mine += (uint64_t)set.n;
                } else {
                    --d;
                }
            } else {
                W.frame[d][lane].i = fr.i + 1;
                if (--budget < 0) {
                    heavyNow = true;
                    active = false;
                } else {
                    bool ok = bind_entry(P.lev[d], G, R, v, fr.i);
                    if (ok && pairs && d == pl) ok = !set.seen(term_value(v, P.hoRef));
                    if (ok) {
                        if (d + 1 == nlev) {
                            if (pairs) {
                                if (!set.insert(term_value(v, P.hoRef))) {
                                    heavyNow = true;
                                    active = false;
                                } else {
                                    d = pl;   // head object settled: next candidate at the pair level
                                }
                            } else {
                                mine += 1;
                                active = false;
                            }
                        } else {
                            ++d;
                            W.frame[d][lane] = enter_level(P.lev[d], G, v);
                        }
                    }
                }
            }
        }
        push_heavy(heavyNow, planIdx, item, hq, hCount, hCap);
    }
    return mine;
}

// Wave mode: the whole wave on one seed. Level-0 entries are handed out one
// at a time from an LDS cursor.
template <int K>
__device__ __forceinline__ uint64_t wave_seed(WaveLds& W, const DevGraph& G, const ConflictRefs* R, uint32_t item,
                                              uint32_t* row, int64_t rowWords) {
    const int lane = lane_id();
    const Plan& P = W.plan;
    const int nlev = P.nlev;
    const bool pairs = K == kPairs && P.pairLevel >= 0;
    const int pl = P.pairLevel;
    int v[kMaxSlots];
    if (!bind_seed(P, G, (int)item, v)) return 0;
    const Frame f0 = enter_level(P.lev[0], G, v);
    if (f0.i >= f0.end) return 0;
    if (lane == 0) {
        W.next0 = f0.i;
        W.end0 = f0.end;
        W.nTouched = 0;
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    const int vs[kMaxSlots] = {v[0], v[1], v[2], v[3], v[4], v[5], v[6], v[7],
                               v[8], v[9], v[10], v[11], v[12], v[13], v[14], v[15]};
    BitRow br{row, &W};
    int d = 0;
    bool active = true, found = false;
    // take the first level-0 entry
    {
        const int i0 = atomicAdd(&W.next0, 1);
        if (i0 < f0.end) W.frame[0][lane] = Frame{i0, i0 + 1};
        else active = false;
    }
    while (__any(active)) {
        if (active) {
            const Frame fr = W.frame[d][lane];
            if (fr.i >= fr.end) {
                if (d == 0) {
                    const int i0 = atomicAdd(&W.next0, 1);
                    if (i0 < f0.end) {
#pragma unroll
                        for (int j = 0; j < kMaxSlots; ++j) v[j] = vs[j];
                        W.frame[0][lane] = Frame{i0, i0 + 1};
                    } else {
                        active = false;
                    }
                } else {
                    --d;
                }
            } else {
                W.frame[d][lane].i = fr.i + 1;
                bool ok = bind_entry(P.lev[d], G, R, v, fr.i);
                if (ok && pairs && d == pl) ok = !br.seen(term_value(v, P.hoRef));
                if (ok) {
                    if (d + 1 == nlev) {
                        if (pairs) {
                            br.insert(term_value(v, P.hoRef));
                            d = pl;
                        } else {
                            found = true;
                            active = false;
                        }
                    } else {
                        ++d;
                        W.frame[d][lane] = enter_level(P.lev[d], G, v);
                    }
                }
            }
        }
        if (!pairs && __any(found)) return 1;
    }
    if (!pairs) return 0;
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
    const uint32_t nt = W.nTouched;
    uint64_t c = 0;
    if (nt <= (uint32_t)kTouchCap) {
        for (uint32_t k = lane; k < nt; k += kWave) c += (uint64_t)__popc(atomicAnd(row + W.touched[k], 0u));
    } else {
        for (int64_t w = lane; w < rowWords; w += kWave) c += (uint64_t)__popc(atomicAnd(row + w, 0u));
    }
    return wave_sum(c);
}

// ------------------------------------------------------------- kernels

// Persistent thread mode over a chunk list: each wave repeatedly takes <= 64
// seeds of one plan, one per lane.
template <int K>
__global__ void __launch_bounds__(kBlock) k_thread_batch(ThreadArgs a, DevGraph G) {
    __shared__ WaveLds lds[kWavesPerBlock];
    WaveLds& W = lds[wave_in_block()];
    const int lane = lane_id();
    uint32_t loaded = 0xffffffffu;
    for (;;) {
        const uint32_t c = wave_take(a.cursor);
        if (c >= a.nChunks) break;
        const Chunk ch = a.chunks[c];
        if (ch.gate && a.counts[2 * a.plans[ch.plan].rule + kSupport] == 0) continue;
        if (ch.plan != loaded) {
            copy_plan(W, a.plans + ch.plan);
            loaded = ch.plan;
        }
        const uint64_t mine = thread_chunk<K>(W, G, a.refs + ch.plan, ch.plan, ch.begin, ch.count, a.heavy,
                                           a.heavyCount, a.heavyCap);
        const uint64_t tot = wave_sum(mine);
        if (lane == 0 && tot) atomicAdd(a.counts + 2 * W.plan.rule + K, (unsigned long long)tot);
    }
}

// One plan passed by value in the kernel arguments (per-rule launches);
// waves stride over its seeds.
template <int K>
__global__ void __launch_bounds__(kBlock) k_thread_one(Plan P, OneArgs a, DevGraph G) {
    __shared__ WaveLds lds[kWavesPerBlock];
    WaveLds& W = lds[wave_in_block()];
    const int lane = lane_id();
    copy_plan(W, &P);
    const uint32_t waves = (gridDim.x * blockDim.x) / kWave;
    for (uint32_t base = ((blockIdx.x * blockDim.x + threadIdx.x) / kWave) * kOneChunkSeeds; base < a.nItems;
         base += waves * kOneChunkSeeds) {
        const uint32_t cnt = a.nItems - base < (uint32_t)kOneChunkSeeds ? a.nItems - base : (uint32_t)kOneChunkSeeds;
        const uint64_t mine = thread_chunk<K>(W, G, a.refs, 0u, base, cnt, a.heavy, a.heavyCount, a.heavyCap);
        const uint64_t tot = wave_sum(mine);
        if (lane == 0 && tot) atomicAdd(a.count, (unsigned long long)tot);
    }
}

// Persistent wave mode over the heavy queue.
template <int K>
__global__ void __launch_bounds__(kBlock) k_wave_batch(WaveArgs a, DevGraph G) {
    __shared__ WaveLds lds[kWavesPerBlock];
    WaveLds& W = lds[wave_in_block()];
    const int lane = lane_id();
    const uint32_t wave = (blockIdx.x * blockDim.x + threadIdx.x) / kWave;
    uint32_t* row = a.rows + (int64_t)wave * a.rowWords;
    uint32_t n = *a.heavyCount;
    if (n > a.heavyCap) n = a.heavyCap;
    uint32_t loaded = 0xffffffffu;
    for (;;) {
        const uint32_t h = wave_take(a.cursor);
        if (h >= n) break;
        const HeavyItem it = a.heavy[h];
        if (it.plan != loaded) {
            copy_plan(W, a.plans + it.plan);
            loaded = it.plan;
        }
        const uint64_t c = wave_seed<K>(W, G, a.refs + it.plan, it.item, row, a.rowWords);
        if (lane == 0 && c) atomicAdd(a.counts + 2 * W.plan.rule + K, (unsigned long long)c);
    }
}

template <int K>
__global__ void __launch_bounds__(kBlock) k_wave_one(Plan P, WaveArgs a, DevGraph G) {
    __shared__ WaveLds lds[kWavesPerBlock];
    WaveLds& W = lds[wave_in_block()];
    const int lane = lane_id();
    copy_plan(W, &P);
    const uint32_t wave = (blockIdx.x * blockDim.x + threadIdx.x) / kWave;
    uint32_t* row = a.rows + (int64_t)wave * a.rowWords;
    uint32_t n = *a.heavyCount;
    if (n > a.heavyCap) n = a.heavyCap;
    for (;;) {
        const uint32_t h = wave_take(a.cursor);
        if (h >= n) break;
        const uint64_t c = wave_seed<K>(W, G, a.refs, a.heavy[h].item, row, a.rowWords);
        if (lane == 0 && c) atomicAdd(a.counts, (unsigned long long)c);
    }
}

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

// ---------------------------------------------------------------- launchers

// One instantiation per plan kind, so a profile names the phase and the
// SUPPORT kernels carry no distinct-set code.
void launch_thread_batch(int kind, const ThreadArgs& a, const DevGraph& G, int blocks, hipStream_t s) {
    if (kind == kSupport) hipLaunchKernelGGL(k_thread_batch<kSupport>, dim3(blocks), dim3(kBlock), 0, s, a, G);
    else hipLaunchKernelGGL(k_thread_batch<kPairs>, dim3(blocks), dim3(kBlock), 0, s, a, G);
}
void launch_thread_one(int kind, const Plan& P, const OneArgs& a, const DevGraph& G, int blocks, hipStream_t s) {
    if (kind == kSupport) hipLaunchKernelGGL(k_thread_one<kSupport>, dim3(blocks), dim3(kBlock), 0, s, P, a, G);
    else hipLaunchKernelGGL(k_thread_one<kPairs>, dim3(blocks), dim3(kBlock), 0, s, P, a, G);
}
void launch_wave_batch(int kind, const WaveArgs& a, const DevGraph& G, int blocks, hipStream_t s) {
    if (kind == kSupport) hipLaunchKernelGGL(k_wave_batch<kSupport>, dim3(blocks), dim3(kBlock), 0, s, a, G);
    else hipLaunchKernelGGL(k_wave_batch<kPairs>, dim3(blocks), dim3(kBlock), 0, s, a, G);
}
void launch_wave_one(int kind, const Plan& P, const WaveArgs& a, const DevGraph& G, int blocks, hipStream_t s) {
    if (kind == kSupport) hipLaunchKernelGGL(k_wave_one<kSupport>, dim3(blocks), dim3(kBlock), 0, s, P, a, G);
    else hipLaunchKernelGGL(k_wave_one<kPairs>, dim3(blocks), dim3(kBlock), 0, s, P, a, G);
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

}  // namespace gr
