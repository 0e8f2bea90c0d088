// search.hpp -- the counting kernels' device code, templated on the number of
// variable slots a launch's plans use (NS = 4, 8 or 16). Each slot class is
// compiled in its own translation unit (search_s4.hip, search_s8.hip,
// search_s16.hip); kernels.hip picks the class per launch.
//
// Work model
//   * A SEED is one head triple (SUPPORT plans) or one head-subject key
//     (PAIRS plans). Seeds are independent, so a seed can be abandoned and
//     redone elsewhere without changing any count.
//   * Every lane runs an iterative depth-first search over the plan's levels:
//     one uniform loop in which each lane advances its own frame (entry
//     cursor, end) at its own depth. Frames and the plan live in LDS, the
//     variable slots in registers (NS of them: binding, lookup and the
//     injectivity test are select chains over the slots, so a launch whose
//     rules have few variables runs short chains), so the control flow stays
//     flat and the register count does not grow with the body length.
//   * Thread mode: one lane per seed, with a work budget; a wave works through
//     a chunk of seeds of one plan and a lane takes the next seed as soon as
//     its own is done. A seed that uses up its budget, or (PAIRS) whose
//     distinct head objects outgrow the lane's register set, is appended to
//     the heavy queue and its partial work discarded.
//   * Wave mode: one wave64 per heavy seed, with a shared LDS pool of entries
//     that idle lanes draw from and busy lanes donate to (wave_run). PAIRS
//     seeds dedup their head objects in a bitmap indexed by the object's rank
//     among the keys of one (direction, predicate) -- an HBM row per wave or
//     dynamic LDS -- whose set words are listed so that counting and clearing
//     touch only those words, or emit (seed, object) keys that the radix sort
//     + unique pass of sort_unique.hip counts.
//
// Semantics, as the reference CPU counters define them
// (Rewritten_CPP/SupportCounting.cpp:299-479, CPU_01_nonparallel/
// SupportCounting.cpp:83-247): a new variable binding must differ from every
// variable bound before it; a ground triple may not repeat an earlier body
// triple or (SUPPORT) the head triple -- levels whose atoms could coincide
// carry a conflict mask computed by the planner.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

#include "kernels.hpp"
#include "plan.hpp"

namespace gr {

// Cycle attribution (profiling build only): s_memtime sums per section of
// the wave-mode search, into the wave counters from kProfBase on.
#ifdef GR_STEP_PROFILE
#define GR_PROF_NOW() ((unsigned long long)__builtin_readcyclecounter())
#define GR_PROF_ADD(slot, t0) (stats[kProfBase + (slot)] += GR_PROF_NOW() - (t0))
#else
#define GR_PROF_NOW() 0ull
#define GR_PROF_ADD(slot, t0) ((void)(t0))
#endif

constexpr int kBlock = 256;
constexpr int kWavesPerBlock = kBlock / kWave;
constexpr int kPlanWords = (int)(sizeof(Plan) / 4);

constexpr int kTouchCap = 256;      // dedup-row words listed per wave-mode seed
constexpr int kCacheSlots = 512;    // sort path: LDS cache of emitted head objects per wave (direct mapped)
constexpr int kOneChunkSeeds = 64;  // seeds per wave in per-rule launches (one rule fills the GPU)
constexpr int kLeafBatch = 16;      // last-level candidates a lane tries per step of its own search (wave mode)
constexpr int kLeafBatchT = 4;      // the same in thread mode (keeps it within 128 VGPRs: 4 waves per SIMD)

struct Frame {
    int i;     // next entry
    int end;   // one past the last entry
};

// Per-wave LDS of thread mode: the plan being evaluated and one frame per
// (level, lane).
// Levels the kernels of slot class NS keep frames for: as many as slots, at
// most kMaxBody (the host sends a plan to a class that covers both its
// variables and its levels).
constexpr int levels_of(int ns) { return ns < kMaxBody ? ns : kMaxBody; }

template <int NL>
struct ThreadLdsT {
    Plan plan;
    Frame frame[NL][kWave];
    // ISECT level L: cursor into row 2 and its end, at cur[L >> 1] (an ISECT
    // level is always followed by its CHECK, so no two share an index)
    Frame cur[NL / 2][kWave];
};

// Wave mode adds the seed's shared pool -- a frame at level poolLevel whose
// entries [poolNext, poolEnd) any idle lane may take, with the bindings
// poolV that hold above that level -- and the number of dedup-row words
// touched so far (HBM rows; the list itself is the kernel's `cache` array).
template <int NL>
struct WaveLdsT : ThreadLdsT<NL> {
    static constexpr bool kLevelSync = false;
    int poolV[kMaxSlots];
    int poolLevel;
    int poolNext;
    int poolEnd;
    uint32_t poolGen;
    uint32_t nTouched;
    uint32_t pad[3];
};

// Per-wave LDS array beside WaveLdsT: the sort path's cache of emitted head
// objects, the HBM-row path's list of touched row words; the LDS-bitmap
// path needs neither (it clears its whole bitmap per seed).
template <int D>
constexpr int wave_cache_words() { return D == kDedupSort ? kCacheSlots : (D == kDedupRows ? kTouchCap : 1); }

static_assert(sizeof(Plan) % 4 == 0, "plan must be word-sized");

__device__ __forceinline__ int lane_id() { return __lane_id(); }
__device__ __forceinline__ int wave_in_block() { return threadIdx.x / kWave; }
// A value every lane holds alike, as a scalar.
__device__ __forceinline__ int uniform(int x) { return __builtin_amdgcn_readfirstlane(x); }

// Value of a term reference: slot r (r >= 0) or constant -(r+1).
template <int NS>
__device__ __forceinline__ int term_value(const int (&v)[NS], int r) {
    int x = -(r + 1);
#pragma unroll
    for (int j = 0; j < NS; ++j) x = (r == j) ? v[j] : x;
    return x;
}

template <int NS>
__device__ __forceinline__ void set_slot(int (&v)[NS], int slot, int x) {
#pragma unroll
    for (int j = 0; j < NS; ++j) v[j] = (slot == j) ? x : v[j];
}

// Is x already bound to one of the first nbound slots?
template <int NS>
__device__ __forceinline__ bool taken(const int (&v)[NS], int nbound, int x) {
    bool d = false;
#pragma unroll
    for (int j = 0; j < NS; ++j) d |= (j < nbound) & (v[j] == x);
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

// Index in col[lo,hi) holding `target`, or -1, for a row whose first and
// last values vlo = col[lo], vhi = col[hi-1] the caller holds: a target
// outside [vlo, vhi] costs no load; inside, up to three interpolation probes
// narrow the bracket (term ids of a row are spread fairly evenly), then a
// binary search finishes it. The probes only pick where to look, so the
// answer is find_in_row's.
__device__ __forceinline__ int find_in_row_interp(const int32_t* __restrict__ col, int lo, int hi, int target, int vlo,
                                                  int vhi) {
    if (lo >= hi || target < vlo || target > vhi) return -1;
    if (target == vlo) return lo;
    if (target == vhi) return hi - 1;
    int l = lo, h = hi - 1, vl = vlo, vh = vhi;   // vl < target < vh, vl = col[l], vh = col[h]
#pragma unroll
    for (int it = 0; it < 3; ++it) {
        if (h - l <= 8) break;
        const float f = (float)(target - vl) / (float)(vh - vl);
        int m = l + (int)(f * (float)(h - l));
        m = m <= l ? l + 1 : (m >= h ? h - 1 : m);
        const int y = col[m];
        if (y == target) return m;
        if (y < target) {
            l = m;
            vl = y;
        } else {
            h = m;
            vh = y;
        }
    }
    return find_in_row(col, l + 1, h, target);
}

// Would ground triple (s, o) of a level repeat an earlier one? Only rules
// with constants reach this; it reads the conflict table from memory.
template <int NS>
__device__ __forceinline__ bool repeats(const ConflictRefs* __restrict__ R, const int (&v)[NS], int mask, int s,
                                        int o) {
    for (int j = 0; j < kMaxBody; ++j)
        if ((mask & (1 << j)) && term_value(v, R->sref[j]) == s && term_value(v, R->oref[j]) == o) return true;
    return (mask & 256) && term_value(v, R->hs) == s && term_value(v, R->ho) == o;
}

// Distinct head objects of one seed, in registers (thread mode).
struct RegSet {
    // 12 entries: on the DBpedia-shaped batch 8 sent seeds with 9-12 head
    // objects to wave mode (wave pass 0.23 -> 0.14 ms with 12); 16 cost the
    // thread pass more registers than it saved
    static constexpr int kCap = 12;
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

// Entry range of term `keyv` in the (direction, predicate) block run `base`.
__device__ __forceinline__ Frame key_row(const DevGraph& G, int64_t base, int keyv) {
    const KeyBlock kb = G.blocks[base + (keyv >> 6)];
    const uint64_t bit = 1ull << (keyv & 63);
    if (!(kb.bits & bit)) return Frame{0, 0};
    const uint32_t r = kb.rank + (uint32_t)__popcll(kb.bits & (bit - 1));
    return Frame{G.rowStart[r], G.rowStart[r + 1]};
}

// Fused support: the key-index block run of a plan's head-predicate SPO rows
// (SPO runs come first, one per predicate).
__device__ __forceinline__ int64_t sup_base(const Plan& P, const DevGraph& G) {
    return (int64_t)P.supPred * G.nBlocks;
}

// Thread mode, fused support: how many of seed s's distinct head objects
// (set) are objects of s in the head predicate -- the supported head triples
// of s. One binary search in s's head row per object.
__device__ __forceinline__ uint32_t head_hits(const DevGraph& G, int64_t base, int s, const RegSet& set) {
    const Frame row = key_row(G, base, s);
    uint32_t n = 0;
    if (row.i < row.end) {
#pragma unroll
        for (int j = 0; j < RegSet::kCap; ++j)
            if (j < set.n && find_in_row(G.col, row.i, row.end, set.e[j]) >= 0) ++n;
    }
    return n;
}

// Entry range of a level under the current bindings (ISECT: its row 1).
template <int NS>
__device__ __forceinline__ Frame enter_level(const Level& lv, const DevGraph& G, const int (&v)[NS]) {
    const int kind = level_kind(lv.flags);
    if (kind == kScan) return Frame{lv.lo, lv.hi};
    const int64_t base = (int64_t)(uint32_t)lv.lo | ((int64_t)lv.hi << 32);
    const Frame row = key_row(G, base, term_value(v, lv.key));
    if (kind != kCheck) return row;
    const int at = find_in_row(G.col, row.i, row.end, term_value(v, lv.val));
    return at < 0 ? Frame{0, 0} : Frame{at, at + 1};
}

// Row 2 of an ISECT level under the current bindings.
template <int NS>
__device__ __forceinline__ Frame isect_row(const Level& lv, const DevGraph& G, const int (&v)[NS]) {
    const int64_t base = (int64_t)(uint32_t)lv.lo2 | ((int64_t)lv.hi2 << 32);
    return key_row(G, base, term_value(v, lv.val2));
}

// Entering level L in a lane's own search: an ISECT level also places the
// lane's row-2 cursor (no entry can match if row 2 is empty); the CHECK
// after an ISECT level is the entry the intersection matched.
template <int NS>
__device__ __forceinline__ Frame enter(ThreadLdsT<levels_of(NS)>& W, const Plan& P, int L, const DevGraph& G,
                                       const int (&v)[NS]) {
    const Level& lv = P.lev[L];
    const int kind = level_kind(lv.flags);
    if (kind == kIsect) {
        const Frame r2 = isect_row(lv, G, v);
        if (r2.i >= r2.end) return Frame{0, 0};
        W.cur[L >> 1][lane_id()] = r2;
    } else if (kind == kCheck && L > 0 && level_kind(P.lev[L - 1].flags) == kIsect) {
        const int j = W.cur[(L - 1) >> 1][lane_id()].i;
        return Frame{j, j + 1};
    }
    return enter_level(lv, G, v);
}

// Is x in row 2 of ISECT level L? Row 1 is enumerated in increasing x, so
// the lane's cursor only moves forward: gallop from it, then binary search.
// Leaves the cursor at the first entry >= x (the match, if any).
template <int NL>
__device__ __forceinline__ bool isect_next(ThreadLdsT<NL>& W, int L, const DevGraph& G, int x) {
    const int lane = lane_id();
    int j = W.cur[L >> 1][lane].i;
    const int e = W.cur[L >> 1][lane].end;
    bool hit = false;
    if (j < e) {
        const int y = G.col[j];
        if (y < x) {
            int lo = j, hi = e, step = 1;
            for (;;) {
                const int t = lo + step;
                if (t >= e) break;
                if (G.col[t] >= x) {
                    hi = t;
                    break;
                }
                lo = t;
                step <<= 1;
            }
            int l = lo + 1, h = hi;
            while (l < h) {
                const int m = (l + h) >> 1;
                if (G.col[m] < x) l = m + 1;
                else h = m;
            }
            j = l;
            hit = j < e && G.col[j] == x;
        } else {
            hit = y == x;
        }
    }
    W.cur[L >> 1][lane].i = j;
    return hit;
}

// The same test for one entry drawn from a shared pool (no cursor to reuse):
// binary search of row 2; on a match the lane's cursor points at it.
template <int NS>
__device__ __forceinline__ bool isect_find(ThreadLdsT<levels_of(NS)>& W, int L, const Level& lv, const DevGraph& G,
                                           const int (&v)[NS], int x) {
    const Frame r2 = isect_row(lv, G, v);
    const int at = find_in_row(G.col, r2.i, r2.end, x);
    if (at < 0) return false;
    W.cur[L >> 1][lane_id()] = Frame{at, r2.end};
    return true;
}

// Binds entry i of a level. False if the candidate violates injectivity or
// repeats an earlier ground triple.
// x is col[i], loaded by the caller.
template <int NS>
__device__ __forceinline__ bool bind_entry_x(const Level& lv, const DevGraph& G, const ConflictRefs* R,
                                             int (&v)[NS], int i, int x) {
    const int f = lv.flags;
    const int kind = level_kind(f);
    const int nbound = level_nbound(f);
    int k = 0;
    if (kind == kExt || kind == kIsect) {
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

template <int NS>
__device__ __forceinline__ bool bind_entry(const Level& lv, const DevGraph& G, const ConflictRefs* R,
                                           int (&v)[NS], int i) {
    return bind_entry_x(lv, G, R, v, i, G.col[i]);
}

// Binds the head of a SUPPORT seed; false if the head triple cannot match.
template <int NS>
__device__ __forceinline__ bool bind_support_seed(const Plan& P, const DevGraph& G, int item, int (&v)[NS]) {
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
template <int NS>
__device__ __forceinline__ bool bind_seed(const Plan& P, const DevGraph& G, int item, int (&v)[NS]) {
#pragma unroll
    for (int j = 0; j < NS; ++j) v[j] = -1;
    if (P.kind == kSupport) return bind_support_seed(P, G, item, v);
    v[0] = G.keys[P.seedBegin + item];
    return true;
}

template <class Lds>
__device__ __forceinline__ void copy_plan(Lds& W, const Plan* src) {
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

// Candidates i[k] (valid where in[k]) of a last level that is a plain EXT --
// no triple conflicts -- whose bound variable is, for PAIRS, the head object
// with rank crank[i]: a candidate grounds iff its value is not bound yet.
// Returns the grounding mask; xs[k] gets each candidate's value, rs[k] the
// rank of each grounding (when `pairs`; else 0), -1 where none. Straight-line
// code: every load is issued (lanes without a candidate read entry `safe`, a
// valid index, and drop the value), and the injectivity test is a minimum of
// XORs against the bound slots (unbound slots hold -1, which no entry does),
// so no candidate gets a branch of its own or a wait between its loads.
template <int U, int NS>
__device__ __forceinline__ uint32_t leaf_fast(const DevGraph& G, const int (&v)[NS], int nbound, bool pairs,
                                              const int (&i)[U], const bool (&in)[U], int safe, int (&xs)[U],
                                              int (&rs)[U]) {
    int at[U];
#pragma unroll
    for (int k = 0; k < U; ++k) {
        at[k] = in[k] ? i[k] : safe;
        xs[k] = G.col[at[k]];
    }
    if (pairs) {
#pragma unroll
        for (int k = 0; k < U; ++k) rs[k] = G.crank[at[k]];
    } else {
#pragma unroll
        for (int k = 0; k < U; ++k) rs[k] = 0;
    }
    int vb[NS];
#pragma unroll
    for (int j = 0; j < NS; ++j) vb[j] = j < nbound ? v[j] : -1;
    uint32_t m = 0;
#pragma unroll
    for (int k = 0; k < U; ++k) {
        uint32_t d = 0xFFFFFFFFu;
#pragma unroll
        for (int j = 0; j < NS; ++j) d = min(d, (uint32_t)(vb[j] ^ xs[k]));
        const bool ok = in[k] & (d != 0u);
        m |= (uint32_t)ok << k;
        rs[k] = ok ? rs[k] : -1;
    }
    return m;
}

// Whether level `last` is a plain EXT that, for PAIRS, binds the head object.
__device__ __forceinline__ bool leaf_plain(const Plan& P, int last, bool pairs) {
    const int f = P.lev[last].flags;
    if (level_kind(f) != kExt || level_conflicts(f) != 0) return false;
    return !pairs || P.pairLevel == last;
}

// Single-frame count (thread mode). A PAIRS seed whose search stands at a
// plain-EXT pair level that is the last level, with no head object counted
// yet and no untried entry left in any frame above, has one row left to
// give head objects: each entry of it grounds unless a bound variable holds
// its value (rows hold each value once), so its distinct head objects number
// the entries minus the bound values found in the row -- one binary search
// per bound variable instead of enumerating the row. The fused support of
// such a seed is the head-row entries (at most kShortHead) in the row that no
// bound variable holds. Frames shorter than kShortMin are enumerated.
constexpr int kShortMin = 16;
constexpr int kShortHead = 16;

template <int NL>
__device__ __forceinline__ bool frames_above_done(const ThreadLdsT<NL>& W, int d) {
    const int lane = lane_id();
    bool done = true;
#pragma unroll
    for (int l = 0; l < NL; ++l)
        if (l < d) done = done && W.frame[l][lane].i >= W.frame[l][lane].end;
    return done;
}

// Entries of row fr (at the last level) that no bound slot j < nb holds.
template <int NS>
__device__ __forceinline__ int frame_count(const DevGraph& G, Frame fr, const int (&v)[NS], int nb) {
    int n = fr.end - fr.i;
#pragma unroll
    for (int j = 0; j < NS; ++j)
        if (j < nb && find_in_row(G.col, fr.i, fr.end, v[j]) >= 0) --n;
    return n;
}

// Head-row entries of seed s (head predicate run `base`) that row fr holds
// and no bound slot takes: the seed's supported head triples when fr is its
// only frame. Returns -1 if the head row is longer than kShortHead.
template <int NS>
__device__ __forceinline__ int frame_head_hits(const DevGraph& G, int64_t base, int s, Frame fr, const int (&v)[NS],
                                               int nb) {
    const Frame hr = key_row(G, base, s);
    if (hr.end - hr.i > kShortHead) return -1;
    int n = 0;
    for (int x = hr.i; x < hr.end; ++x) {
        const int h = G.col[x];
        if (!taken(v, nb, h) && find_in_row(G.col, fr.i, fr.end, h) >= 0) ++n;
    }
    return n;
}

// Whether level `last` takes the leaf_fast path of wave mode under dedup D
// (bitmap marks by crank rank).
template <int D>
__device__ __forceinline__ bool leaf_fast_ok(const Plan& P, int last, bool pairs) {
    return leaf_plain(P, last, pairs) && (!pairs || (D != kDedupSort && P.hoByEntry));
}

// Thread mode over seeds [begin, begin + count) of the plan in W. Every lane
// runs its own depth-first search; a lane whose seed is finished takes the
// next one (ballot prefix count), so no lane waits for the slowest seed.
// Returns this lane's count; escalated seeds go to the heavy queue. K is the
// plan kind of every plan the launch sees (kSupport or kPairs). With `fuse`,
// a PAIRS plan with supPred >= 0 adds each finished seed's supported head triples
// to `sup`, one with a member list (supPred <= -2) adds each member's to its
// counter in `counts` (seeds moved to the heavy queue count there instead).
template <int K, int NS>
__device__ __forceinline__ uint64_t thread_chunk(ThreadLdsT<levels_of(NS)>& W, const DevGraph& G, const ConflictRefs* R,
                                                 uint32_t planIdx, uint32_t begin, uint32_t count, HeavyItem* hq,
                                                 uint32_t* hCount, uint32_t hCap, bool fuse, uint64_t& sup,
                                                 unsigned long long* counts, const int32_t* supTab) {
    const int lane = lane_id();
    const Plan& P = W.plan;
    // plan-wide fields as wave-uniform values (scalar registers)
    const int nlev = uniform(P.nlev);
    const int pl = uniform(P.pairLevel);
    const int hoRef = uniform(P.hoRef);
    const int budget0 = uniform(P.budget);
    const bool pairs = K == kPairs && pl >= 0;
    const bool seedRow0 = K == kPairs && uniform(P.seedRow0) != 0;
    // also count the seeds' supported head triples: of this plan's rule, or of
    // each rule in its member list (added to their counters directly)
    const int supPred = uniform(P.supPred);
    const bool fs = pairs && fuse && supPred != -1;
    const int32_t* supList = (fs && supPred <= -2) ? supTab + (-supPred - 2) : nullptr;
    // single-frame counts apply to this plan (a plain-EXT pair level that is the last level)
    const bool shortOk = pairs && pl == nlev - 1 && leaf_plain(P, nlev - 1, pairs);
    const uint64_t below = (1ull << lane) - 1ull;
    int v[NS];
    RegSet set;
    set.n = 0;
    int d = 0, budget = 0;
    bool active = false;
    bool noCut = false;   // this seed's head rows are too long for a single-frame count
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
                    const int r0 = P.seedBegin + (int)item;
                    const Frame f0 = seedRow0 ? Frame{G.rowStart[r0], G.rowStart[r0 + 1]} : enter(W, P, 0, G, v);
                    active = f0.i < f0.end;   // an empty level-0 row: nothing to count
                    W.frame[0][lane] = f0;
                    d = 0;
                    // a distinct-pair search visits every level-0 entry at one
                    // budget unit or more each: a longer level-0 row than the
                    // budget is sure to escalate, so it does at its first step
                    budget = (pairs && f0.end - f0.i > budget0) ? 0 : budget0;
                    set.n = 0;
                    noCut = false;
                }
            }
            next += (uint32_t)__popcll(want);
        }
        if (!__any(active)) {
            if (next >= count) break;
            continue;
        }
        bool heavyNow = false;
        if (active && shortOk && !noCut && d + 1 == nlev && set.n == 0) {
            const Frame fr = W.frame[d][lane];
            if (fr.end - fr.i >= kShortMin && frames_above_done(W, d)) {
                // the seed's only row left, nothing counted: count it without enumerating
                const int nb = level_nbound(P.lev[d].flags);
                bool cut = true;
                if (fs) {   // the head rows must be short enough to search them
                    const int nm = supList ? supList[0] : 1;
                    for (int m = 0; m < nm && cut; ++m) {
                        const int pred = supList ? supList[2 + 2 * m] : supPred;
                        const Frame hr = key_row(G, (int64_t)pred * G.nBlocks, v[0]);
                        cut = hr.end - hr.i <= kShortHead;
                    }
                    noCut = !cut;
                }
                if (cut) {
                    mine += (uint64_t)frame_count(G, fr, v, nb);
                    if (fs) {
                        const int nm = supList ? supList[0] : 1;
                        for (int m = 0; m < nm; ++m) {
                            const int pred = supList ? supList[2 + 2 * m] : supPred;
                            const int h = frame_head_hits(G, (int64_t)pred * G.nBlocks, v[0], fr, v, nb);
                            if (h <= 0) continue;
                            if (!supList) sup += (uint64_t)h;
                            else atomicAdd(counts + 2 * supList[1 + 2 * m] + kSupport, (unsigned long long)h);
                        }
                    }
                    active = false;
                }
            }
        }
        if (active) {
            const Frame fr = W.frame[d][lane];
            bool spent = false;      // level d has nothing left to try after this step
            bool finished = false;   // the seed's search is over
            if (fr.i >= fr.end) {
                if (d == 0) finished = true;
                else --d;
            } else if (d + 1 == nlev && fr.end - fr.i > 1 && leaf_plain(P, d, pairs)) {
                // last level, plain EXT: up to kLeafBatchT candidates per step, straight-line
                const int take = fr.end - fr.i < kLeafBatchT ? fr.end - fr.i : kLeafBatchT;
                W.frame[d][lane].i = fr.i + take;
                spent = fr.i + take >= fr.end;
                budget -= take;
                if (budget < 0) {
                    heavyNow = true;
                    active = false;
                } else {
                    int is[kLeafBatchT], xs[kLeafBatchT], rs[kLeafBatchT];
                    bool in[kLeafBatchT];
#pragma unroll
                    for (int k = 0; k < kLeafBatchT; ++k) {
                        is[k] = fr.i + k;
                        in[k] = k < take;
                    }
                    const uint32_t m = leaf_fast(G, v, level_nbound(P.lev[d].flags), false, is, in, fr.i, xs, rs);
                    if (!pairs) {
                        if (m) {
                            mine += 1;
                            active = false;
                        }
                    } else {
                        // a row holds each value once: the new objects are distinct
                        bool over = false;
#pragma unroll
                        for (int k = 0; k < kLeafBatchT; ++k)
                            if (((m >> k) & 1u) && !set.seen(xs[k])) over = over || !set.insert(xs[k]);
                        if (over) {
                            heavyNow = true;
                            active = false;
                        }
                    }
                }
            } else if (d + 1 == nlev && fr.end - fr.i > 1) {
                // last level: up to kLeafBatchT candidates per step, loaded together
                const int take = fr.end - fr.i < kLeafBatchT ? fr.end - fr.i : kLeafBatchT;
                W.frame[d][lane].i = fr.i + take;
                spent = fr.i + take >= fr.end;
                budget -= take;
                if (budget < 0) {
                    heavyNow = true;
                    active = false;
                } else {
                    int xs[kLeafBatchT];
#pragma unroll
                    for (int k = 0; k < kLeafBatchT; ++k) xs[k] = k < take ? G.col[fr.i + k] : 0;
                    bool stop = false;
#pragma unroll
                    for (int k = 0; k < kLeafBatchT; ++k) {
                        if (stop || k >= take || !bind_entry_x(P.lev[d], G, R, v, fr.i + k, xs[k])) continue;
                        if (!pairs) {
                            mine += 1;
                            active = false;
                            stop = true;
                            continue;
                        }
                        const int ho = term_value(v, hoRef);
                        if (d == pl && set.seen(ho)) continue;
                        if (!set.insert(ho)) {
                            heavyNow = true;
                            active = false;
                            stop = true;
                        } else if (d != pl) {
                            d = pl;   // head object settled: next candidate at the pair level
                            stop = true;
                            spent = false;
                        }
                    }
                }
            } else {
                W.frame[d][lane].i = fr.i + 1;
                spent = fr.i + 1 >= fr.end;   // unless the candidate descends
                if (--budget < 0) {
                    heavyNow = true;
                    active = false;
                } else {
                    const int x = G.col[fr.i];
                    bool ok = bind_entry_x(P.lev[d], G, R, v, fr.i, x);
                    if (ok && level_kind(P.lev[d].flags) == kIsect) ok = isect_next(W, d, G, x);
                    if (ok && pairs && d == pl) ok = !set.seen(term_value(v, hoRef));
                    if (ok) {
                        if (d + 1 == nlev) {
                            if (pairs) {
                                if (!set.insert(term_value(v, hoRef))) {
                                    heavyNow = true;
                                    active = false;
                                } else {
                                    d = pl;   // head object settled: next candidate at the pair level
                                    spent = false;
                                }
                            } else {
                                mine += 1;
                                active = false;
                            }
                        } else {
                            const Frame nf = enter(W, P, d + 1, G, v);
                            if (nf.i < nf.end) {   // an empty row is not entered
                                ++d;
                                W.frame[d][lane] = nf;
                                spent = false;
                            }
                        }
                    }
                }
            }
            // leave spent levels now rather than one per step; a seed whose
            // level-0 row is spent ends in this step
            if (active && spent) {
                finished = d == 0;
                while (d > 0) {
                    --d;
                    const Frame up = W.frame[d][lane];
                    finished = up.i >= up.end;
                    if (!finished) break;
                }
            }
            if (active && finished) {
                active = false;
                if (pairs) mine += (uint64_t)set.n;
                if (fs && set.n > 0) {
                    const int nm = supList ? supList[0] : 1;
                    for (int m = 0; m < nm; ++m) {
                        const int pred = supList ? supList[2 + 2 * m] : supPred;
                        const uint32_t h = head_hits(G, (int64_t)pred * G.nBlocks, v[0], set);
                        if (!supList) sup += h;
                        else if (h) atomicAdd(counts + 2 * supList[1 + 2 * m] + kSupport, (unsigned long long)h);
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
__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

__device__ __forceinline__ uint32_t bcast(uint32_t x) { return (uint32_t)__shfl((int)x, 0, kWave); }
__device__ __forceinline__ uint32_t ld_rlx(const uint32_t* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Head-object dedup of one PAIRS seed in wave mode.
//   kDedupRows: the wave's bitmap row in HBM (atomicOr);
//   kDedupLds:  the same bitmap in the wave's LDS (cleared whole per seed);
//   kDedupSort: nothing is marked; the caller emits keys to HBM, and a small
//               direct-mapped LDS cache of emitted objects prunes repeats.
// For HBM rows, the lane whose atomicOr finds a word still zero lists it in
// LDS so that clearing touches only those words; insert returns 1 to the one
// lane whose call set the bit, so the lanes' sums are the distinct count.
// The bitmaps index a head object x by its rank among the keys of the
// plan's head-object space (Plan::hoLo/hoHi/hoRow0/hoKeys) -- crank[e] when
// the pair level binds x as col[e], else one key-index block load and a
// popcount -- so a bitmap needs only as many bits as the largest key set of
// one (direction, predicate), not one per term id.
template <int D, int NS>
struct Dedup {
    uint32_t* bits;     // the wave's own bitmap: its HBM row (kDedupRows) or LDS (kDedupLds)
    uint32_t* shared;   // a giant's HBM row, shared by its sub-items (then used instead of bits), or null
    uint32_t* merge;    // LDS bitmaps: a giant's HBM row the wave's bitmap is ORed into when the sub-item ends
    const Plan* plan;   // the wave's plan (LDS)
    uint32_t* nTouched; // HBM rows: words listed in `cache` so far (LDS)
    int* cache;         // wave_cache_words<D>() words: emitted-object cache (sort) or touched-word list (rows)
    const KeyBlock* blocks;
    const int32_t* crank;
    // Rank of head object x (sort path: x itself). x is a key of the space by
    // construction; the clamp only keeps a wrong plan from writing out of bounds.
    __device__ __forceinline__ int rank(int x) const {
        if (D == kDedupSort) return x;
        const Plan& P = *plan;
        const int64_t base = (int64_t)(uint32_t)P.hoLo | ((int64_t)P.hoHi << 32);
        const KeyBlock kb = blocks[base + (x >> 6)];
        const uint64_t bit = 1ull << (x & 63);
        const int r = (int)(kb.rank + (uint32_t)__popcll(kb.bits & (bit - 1))) - P.hoRow0;
        return r < 0 ? 0 : (r >= P.hoKeys ? P.hoKeys - 1 : r);
    }
    // Rank of head object x bound by entry i at the pair level: crank[i] when
    // the pair level binds it as col[i] (the plan's rank space by
    // construction), else from the key index (sort path: x itself).
    __device__ __forceinline__ int rank_at(int i, int x) const {
        if (D == kDedupSort) return x;
        return plan->hoByEntry ? crank[i] : rank(x);
    }
    // Rank of x if it is a key of the head-object space, else -1.
    __device__ __forceinline__ int rank_key(int x) const {
        const Plan& P = *plan;
        const int64_t base = (int64_t)(uint32_t)P.hoLo | ((int64_t)P.hoHi << 32);
        const KeyBlock kb = blocks[base + (x >> 6)];
        const uint64_t bit = 1ull << (x & 63);
        if (!(kb.bits & bit)) return -1;
        return (int)(kb.rank + (uint32_t)__popcll(kb.bits & (bit - 1))) - P.hoRow0;
    }
    __device__ __forceinline__ bool seen(int x) const { return seen_r(rank(x)); }
    __device__ __forceinline__ uint32_t insert(int x) { return insert_r(rank(x)); }
    __device__ __forceinline__ bool seen_r(int x) const {
        if (D == kDedupSort) return cache[x & (kCacheSlots - 1)] == x;
        uint32_t w;
        if (shared) w = __hip_atomic_load(shared + (x >> 5), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        else if (D == kDedupRows) w = __hip_atomic_load(bits + (x >> 5), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        else w = bits[x >> 5];
        return (w >> (x & 31)) & 1u;
    }
    __device__ __forceinline__ uint32_t insert_r(int x) {
        if (D == kDedupSort) {
            cache[x & (kCacheSlots - 1)] = x;
            return 0;
        }
        const uint32_t bit = 1u << (x & 31);
        if (shared) return (atomicOr(shared + (x >> 5), bit) & bit) ? 0u : 1u;
        const uint32_t old = atomicOr(bits + (x >> 5), bit);
        if (D == kDedupRows && old == 0) {   // list the word for the clear
            const uint32_t k = atomicAdd(nTouched, 1u);
            if (k < (uint32_t)kTouchCap) cache[k] = x >> 5;
        }
        return (old & bit) ? 0u : 1u;
    }
    __device__ __forceinline__ void clear_word(uint32_t w) {
        if (D == kDedupRows) atomicAnd(bits + w, 0u);
        else bits[w] = 0u;
    }
    // The bitmap word holding rank x (bitmap modes), for tests batched ahead of the marks.
    __device__ __forceinline__ uint32_t word_r(int x) const {
        if (shared) return __hip_atomic_load(shared + (x >> 5), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (D == kDedupRows) return __hip_atomic_load(bits + (x >> 5), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        return bits[x >> 5];
    }
};

// Wave mode, fused support: how many head objects marked in a PAIRS seed's
// bitmap are objects of the seed in the head predicate (supported head
// triples); hr is the seed's row there. `row`: a giant's HBM row; null: the
// wave's own bitmap. Called by the whole wave after its marks; returns the
// wave's total.
template <int D, int NS>
__device__ __forceinline__ uint64_t head_hits_wave(const Dedup<D, NS>& dd, const DevGraph& G, Frame hr,
                                                   const uint32_t* row) {
    uint64_t n = 0;
    for (int i = hr.i + lane_id(); i < hr.end; i += kWave) {
        const int r = dd.rank_key(G.col[i]);
        if (r < 0) continue;   // not a value the pair level can bind: no body pair
        const uint32_t w = row ? __hip_atomic_load(row + (r >> 5), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)
                               : dd.word_r(r);
        n += (w >> (r & 31)) & 1u;
    }
    return wave_sum(n);
}

// LDS bitmap (the wave's own): marks ranks r[k] (>= 0; -1: none). The bit
// tests are read together, unconditionally (a slot without a rank reads
// word 0: lanes reading one address share a broadcast, and its bit is 0);
// then an atomic goes out for each bit found clear, all before any returned
// word is used, so a batch costs two LDS round trips however many marks it
// makes (no atomic for slots without a rank or for bits already set). A
// mark whose bit the returned word still lacked was set first by this lane.
template <int D, int U, int NS>
__device__ __forceinline__ uint32_t mark_lds(Dedup<D, NS>& dd, const int (&r)[U]) {
    uint32_t bit[U], w[U], old[U];
    int idx[U];
#pragma unroll
    for (int k = 0; k < U; ++k) {
        const bool in = r[k] >= 0;
        bit[k] = in ? 1u << (r[k] & 31) : 0u;
        idx[k] = in ? r[k] >> 5 : 0;
        w[k] = dd.bits[idx[k]];
    }
#pragma unroll
    for (int k = 0; k < U; ++k) {
        old[k] = ~0u;
        if (bit[k] & ~w[k]) old[k] = atomicOr(dd.bits + idx[k], bit[k]);
    }
    uint32_t n = 0;
#pragma unroll
    for (int k = 0; k < U; ++k) n += (bit[k] & ~old[k]) ? 1u : 0u;
    return n;
}

// Marks the head objects of a batch of last-level groundings (ho[k] >= 0;
// -1: none), with their ranks hr[k] (bitmap paths). The word tests are
// issued together, then the marks together, so a batch costs two memory
// round trips instead of two per candidate. Returns the objects this lane
// marked first; on the sort path emit[k] gets the objects to emit.
template <int D, int U, int NS>
__device__ __forceinline__ uint32_t mark_batch(Dedup<D, NS>& dd, const int (&ho)[U], const int (&hr)[U], int (&emit)[U]) {
    uint32_t n = 0;
    if (D == kDedupSort) {
#pragma unroll
        for (int k = 0; k < U; ++k) {
            emit[k] = -1;
            if (ho[k] >= 0 && !dd.seen(ho[k])) {
                dd.insert(ho[k]);
                emit[k] = ho[k];
            }
        }
        return 0;
    }
    int r[U];
    uint32_t w[U];
#pragma unroll
    for (int k = 0; k < U; ++k) r[k] = ho[k] >= 0 ? hr[k] : -1;
    if (D == kDedupLds) {
#pragma unroll
        for (int k = 0; k < U; ++k) emit[k] = -1;
        return mark_lds<D, U, NS>(dd, r);
    }
#pragma unroll
    for (int k = 0; k < U; ++k) w[k] = r[k] >= 0 ? dd.word_r(r[k]) : 0xffffffffu;
#pragma unroll
    for (int k = 0; k < U; ++k) {
        emit[k] = -1;
        if (r[k] >= 0 && !((w[k] >> (r[k] & 31)) & 1u)) n += dd.insert_r(r[k]);
    }
    return n;
}

// Sort path: wave-aggregated append of this step's emitted head objects as
// keys keyHi | object. Called by the whole wave.
__device__ __forceinline__ void emit_keys(const SortSink& sink, unsigned long long keyHi, int emit) {
    const uint64_t em = __ballot(emit >= 0);
    if (em == 0) return;
    unsigned long long base = 0;
    if (lane_id() == 0) base = atomicAdd(sink.count, (unsigned long long)__popcll(em));
    base = __shfl(base, 0, kWave);
    if (emit >= 0) {
        const unsigned long long pos = base + (unsigned long long)__popcll(em & ((1ull << lane_id()) - 1ull));
        if (pos < (unsigned long long)sink.cap) sink.keys[pos] = keyHi | (unsigned long long)(uint32_t)emit;
    }
}

// Bitmap paths: marks ranks hr[k] (>= 0; -1: none); HBM rows test first, as mark_batch.
template <int D, int U, int NS>
__device__ __forceinline__ uint32_t mark_ranks(Dedup<D, NS>& dd, const int (&hr)[U]) {
    if (D == kDedupLds) return mark_lds<D, U, NS>(dd, hr);
    uint32_t w[U];
#pragma unroll
    for (int k = 0; k < U; ++k) w[k] = hr[k] >= 0 ? dd.word_r(hr[k]) : 0xffffffffu;
    uint32_t n = 0;
#pragma unroll
    for (int k = 0; k < U; ++k)
        if (!((w[k] >> (hr[k] & 31)) & 1u)) n += dd.insert_r(hr[k]);
    return n;
}

// The end of a wave-mode seed or sub-item, shared by both wave engines.
// SUPPORT (and existence-only PAIRS): 0/1, a giant's sub-item counting only
// if it is the one to set the giant's found flag. PAIRS with bitmap dedup:
// the fused support of the seed from its bitmap (supOn), then the bitmap is
// cleared (a giant's LDS sub-item ORs its bitmap into the giant's row first:
// the bits that OR sets first are the objects it adds); returns the distinct
// head objects this wave marked first (0 if aborted, and on the sort path,
// which counts later).
template <int K, int D, int NS>
__device__ __forceinline__ uint64_t seed_finish(Dedup<D, NS>& dd, const DevGraph& G, int64_t rowWords, bool pairs,
                                                bool aborted, bool stopped, bool found, uint32_t* foundFlag,
                                                uint64_t mine, bool supOn, Frame supRow, const int32_t* supList,
                                                int supSubject, unsigned long long* counts, uint64_t& supOut,
                                                unsigned long long (&stats)[kWaveStats], unsigned long long tEnd) {
    const int lane = lane_id();
    if (!pairs) {
        if (aborted || stopped || !__any(found)) return 0;
        if (!foundFlag) return 1;
        uint32_t win = 0;
        if (lane == 0) win = atomicCAS(foundFlag, 0u, 1u) == 0u ? 1u : 0u;
        return bcast(win);
    }
    if (D == kDedupSort) return 0;
    // clear the wave's own bitmap (a sub-item marked only its giant's row: nothing listed)
    wave_sync();
    if (supOn && !aborted && wave_sum(mine) != 0) {
        const int nm = supList ? supList[0] : 1;
        for (int m = 0; m < nm; ++m) {
            const Frame hr = supList ? key_row(G, (int64_t)supList[2 + 2 * m] * G.nBlocks, supSubject) : supRow;
            const uint64_t h = head_hits_wave(dd, G, hr, nullptr);
            if (!supList) supOut += h;
            else if (lane == 0 && h) atomicAdd(counts + 2 * supList[1 + 2 * m] + kSupport, (unsigned long long)h);
        }
    }
    if (D == kDedupLds) {
        // a giant's sub-item ORs its bitmap into the giant's row (atomics with no
        // return value, so none is waited for); the giant's last sub-item counts
        // the row whole, so a sub-item counts nothing itself
        for (int64_t w = lane; w < rowWords; w += kWave) {
            if (dd.merge) {
                const uint32_t x = dd.bits[w];
                if (x) (void)__hip_atomic_fetch_or(dd.merge + w, x, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            }
            dd.clear_word((uint32_t)w);
        }
        wave_sync();
        if (dd.merge) mine = 0;
    } else {
        const uint32_t nt = *dd.nTouched;
        if (nt <= (uint32_t)kTouchCap) {
            for (uint32_t k = lane; k < nt; k += kWave) dd.clear_word((uint32_t)dd.cache[k]);
        } else {
            for (int64_t w = lane; w < rowWords; w += kWave) dd.clear_word((uint32_t)w);
        }
    }
    const uint64_t total = wave_sum(mine);
    GR_PROF_ADD(6, tEnd);
    return aborted ? 0 : total;
}

constexpr int kLeafUnroll = 16;  // last-level pool entries an idle lane takes per draw
// The same for last levels off the straight-line path (conflict checks,
// SCAN levels, a head object bound above): each candidate inlines the full
// binder, so a short batch keeps the kernel's code small.
constexpr int kLeafGeneric = 4;

// Wave mode: the whole wave on one seed, with a shared pool of work.
//   * The pool starts as the seed's level-0 row. Each step, the idle lanes
//     draw consecutive pool entries (coalesced): from a last-level pool each
//     takes kLeafUnroll entries and finishes them on the spot; from any other
//     level each takes one entry, binds it on top of the pool's bindings and
//     searches below it on its own (frames in LDS), as in thread mode.
//   * When the pool is empty and some lanes are idle, the busy lane with the
//     most untried entries in its shallowest open frame donates the rest of
//     that frame (with its bindings) as the new pool. Long rows at any depth
//     -- hub rows -- are thus spread over the wave instead of scanned by one
//     lane.
//   * SUPPORT seeds end at the first complete grounding (__any). PAIRS seeds
//     count each head object once (Dedup); a lane whose head object is
//     already counted drops the rest of that subtree.
//   * A seed (not a sub-item) still running after ga.after steps is listed
//     as a giant and abandoned (`aborted`; its partial count is dropped).
//     A SUPPORT sub-item shares its giant's `found` flag: it polls it and
//     counts only if it is the one to set it.
//   * supOn (fused support, bitmap dedup): before the bitmap is cleared, the
//     supported head triples of the seed, whose head row is supRow, are
//     added to supOut -- or, with a member list, each member's (head rows of
//     seed supSubject) to its counter in `counts` -- nothing if the seed was
//     abandoned or found no pair.
// The pool (poolLevel/Next/End/V) is set up by the caller, all lanes idle.
// Returns the count this wave adds (SUPPORT: 0/1; PAIRS: distinct head
// objects it marked first, or 0 on the sort path, which counts later).
template <int K, int D, int NS>
__device__ __forceinline__ uint64_t wave_run(WaveLdsT<levels_of(NS)>& W, const DevGraph& G, const ConflictRefs* R,
                                             Dedup<D, NS>& dd,
                                             int64_t rowWords, const SortSink& sink, unsigned long long keyHi,
                                             const GiantArgs& ga, bool mayGiant, HeavyItem self, uint32_t heavyIdx,
                                             uint32_t* foundFlag, bool& aborted, bool supOn, Frame supRow,
                                             const int32_t* supList, int supSubject, unsigned long long* counts,
                                             uint64_t& supOut,
                                             unsigned long long (&stats)[kWaveStats]) {
    const int lane = lane_id();
    const Plan& P = W.plan;
    const int nlev = uniform(P.nlev);
    const int last = nlev - 1;
    const int pl = uniform(P.pairLevel);
    const int hoRef = uniform(P.hoRef);
    const bool hoByEntry = uniform(P.hoByEntry) != 0;
    const bool pairs = K == kPairs && pl >= 0;
    const uint64_t below = (1ull << lane) - 1ull;
    int v[NS];
#pragma unroll
    for (int j = 0; j < NS; ++j) v[j] = -1;
    bool busy = false, found = false, stopped = false;
    aborted = false;
    int d = 0, base = 0;            // own search: current level, first level it owns
    int hr = -1;                    // rank of the head object once v binds it (pairs)
    uint32_t gen = 0xffffffffu;     // pool whose bindings v holds (idle lanes)
    uint64_t mine = 0;
    int steps = 0;
    uint32_t nDonate = 0, nLeafDraw = 0, nDraw = 0;
    uint64_t busyLanes = 0, drawLanes = 0;
    for (;;) {
        // ---- A: idle lanes draw from the pool
        unsigned long long tp = GR_PROF_NOW();
        [[maybe_unused]] int profA = 3;
        const uint64_t idleM = __ballot(!busy);
        const int pn = uniform(W.poolNext), pe = uniform(W.poolEnd), plv = uniform(W.poolLevel);
        const uint32_t pg = (uint32_t)uniform((int)W.poolGen);
        if (idleM != 0 && pn < pe) {
            const int n = __popcll(idleM);
            const int r = __popcll(idleM & below);
            drawLanes += (uint64_t)(pe - pn < n ? pe - pn : n);
            int next = pe;
            bool drop = false;
            int hru = -1;              // rank of a head object bound above the pool
            if (pairs && pl < plv) {   // every entry shares a head object bound above the pool
                const int ref = hoRef;
                const int hov = ref >= 0 ? W.poolV[ref] : -(ref + 1);
                hru = uniform(dd.rank(uniform(hov)));
                drop = dd.seen_r(hru);
            }
            if (!drop) {
                if (!busy && gen != pg) {
#pragma unroll
                    for (int j = 0; j < NS; ++j) v[j] = W.poolV[j];
                    gen = pg;
                }
                if (plv == last && leaf_fast_ok<D>(P, last, pairs)) {
                    ++nLeafDraw;
                    profA = 2;
                    next = pn + n * kLeafUnroll < pe ? pn + n * kLeafUnroll : pe;
                    int is[kLeafUnroll], xs[kLeafUnroll], rs[kLeafUnroll];
                    bool in[kLeafUnroll];
#pragma unroll
                    for (int k = 0; k < kLeafUnroll; ++k) {
                        is[k] = pn + r + n * k;
                        in[k] = !busy && is[k] < pe;
                    }
                    const uint32_t m = leaf_fast(G, v, level_nbound(P.lev[last].flags), pairs, is, in, pn, xs, rs);
                    if (!pairs) found = found || m != 0u;
                    else mine += mark_ranks<D, kLeafUnroll>(dd, rs);
                } else if (plv == last) {
                    ++nLeafDraw;
                    profA = 2;
                    next = pn + n * kLeafGeneric < pe ? pn + n * kLeafGeneric : pe;
                    const bool byEntry = pairs && pl == last && D != kDedupSort && hoByEntry;
                    int xs[kLeafGeneric], rs[kLeafGeneric];
#pragma unroll
                    for (int k = 0; k < kLeafGeneric; ++k) {   // loads issued together (entry pn is valid)
                        const int i = pn + r + n * k;
                        const int at = (!busy && i < pe) ? i : pn;
                        xs[k] = G.col[at];
                        rs[k] = byEntry ? G.crank[at] : hru;
                    }
                    bool hit = false;
                    int hos[kLeafGeneric];
#pragma unroll
                    for (int k = 0; k < kLeafGeneric; ++k) {
                        const int i = pn + r + n * k;
                        hos[k] = -1;
                        if (!busy && i < pe && bind_entry_x(P.lev[last], G, R, v, i, xs[k])) {
                            if (!pairs) {
                                found = true;
                            } else {
                                hos[k] = term_value(v, hoRef);
                                if (pl == last && !byEntry) rs[k] = dd.rank_at(i, hos[k]);
                            }
                        }
                    }
                    if (pairs) {
                        if (pl < last) {   // one head object for the whole pool: mark it once
#pragma unroll
                            for (int k = 1; k < kLeafGeneric; ++k)
                                if (hos[0] < 0) hos[0] = hos[k];
#pragma unroll
                            for (int k = 1; k < kLeafGeneric; ++k) hos[k] = -1;
                        }
                        int emits[kLeafGeneric];
                        mine += mark_batch<D, kLeafGeneric>(dd, hos, rs, emits);
                        hit = hos[0] >= 0;
                        if (D == kDedupSort) {
#pragma unroll
                            for (int k = 0; k < kLeafGeneric; ++k) emit_keys(sink, keyHi, emits[k]);
                        }
                    }
                    // head object bound above the pool and now counted: the rest adds nothing
                    if (pairs && pl < last && __any(hit)) next = pe;
                } else {
                    ++nDraw;
                    next = pn + n < pe ? pn + n : pe;
                    const int i = pn + r;
                    if (!busy && i < pe) {
                        const int x = G.col[i];
                        bool ok = bind_entry_x(P.lev[plv], G, R, v, i, x);
                        if (ok && level_kind(P.lev[plv].flags) == kIsect) ok = isect_find(W, plv, P.lev[plv], G, v, x);
                        if (ok && pairs && plv == pl) {
                            hr = dd.rank_at(i, term_value(v, hoRef));
                            ok = !dd.seen_r(hr);
                        } else if (pairs && pl < plv) {
                            hr = hru;
                        }
                        if (ok) {
                            d = base = plv + 1;
                            W.frame[d][lane] = enter(W, P, d, G, v);
                            busy = true;
                        }
                    }
                }
            }
            wave_sync();
            if (lane == 0) W.poolNext = next;
        }
        GR_PROF_ADD(profA, tp);
        tp = GR_PROF_NOW();
        if (!pairs && __any(found)) break;
        // ---- B: busy lanes advance their own search by one step
        int emit = -1;
        int emits[kLeafGeneric];
#pragma unroll
        for (int k = 0; k < kLeafGeneric; ++k) emits[k] = -1;
        if (busy) {
            if (pairs && pl < base && dd.seen_r(hr)) {
                busy = false;   // head object counted by another lane
            } else {
                const Frame fr = W.frame[d][lane];
                if (fr.i >= fr.end) {
                    if (d == base) busy = false;
                    else --d;
                } else if (d == last && fr.end - fr.i > 1 && leaf_fast_ok<D>(P, last, pairs)) {
                    // last level, plain EXT: up to kLeafBatch candidates per step, straight-line
                    const int take = fr.end - fr.i < kLeafBatch ? fr.end - fr.i : kLeafBatch;
                    W.frame[d][lane].i = fr.i + take;
                    int is[kLeafBatch], xs[kLeafBatch], rs[kLeafBatch];
                    bool in[kLeafBatch];
#pragma unroll
                    for (int k = 0; k < kLeafBatch; ++k) {
                        is[k] = fr.i + k;
                        in[k] = k < take;
                    }
                    const uint32_t m = leaf_fast(G, v, level_nbound(P.lev[d].flags), pairs, is, in, fr.i, xs, rs);
                    if (!pairs) {
                        if (m) {
                            found = true;
                            busy = false;
                        }
                    } else {
                        mine += mark_ranks<D, kLeafBatch>(dd, rs);
                    }
                } else if (d == last && fr.end - fr.i > 1) {
                    // last level: up to kLeafGeneric candidates per step, loaded together
                    const int take = fr.end - fr.i < kLeafGeneric ? fr.end - fr.i : kLeafGeneric;
                    W.frame[d][lane].i = fr.i + take;
                    const bool byEntry = pairs && d == pl && D != kDedupSort && hoByEntry;
                    int xs[kLeafGeneric], rs[kLeafGeneric];
#pragma unroll
                    for (int k = 0; k < kLeafGeneric; ++k) {
                        xs[k] = k < take ? G.col[fr.i + k] : 0;
                        rs[k] = (k < take && byEntry) ? G.crank[fr.i + k] : hr;
                    }
                    int hos[kLeafGeneric];
                    bool any = false;
#pragma unroll
                    for (int k = 0; k < kLeafGeneric; ++k) {
                        hos[k] = -1;
                        if (k < take && bind_entry_x(P.lev[d], G, R, v, fr.i + k, xs[k])) {
                            hos[k] = pairs ? term_value(v, hoRef) : 0;
                            if (pairs && d == pl && !byEntry) rs[k] = dd.rank_at(fr.i + k, hos[k]);
                            any = true;
                        }
                    }
                    if (!pairs) {
                        if (any) {
                            found = true;
                            busy = false;
                        }
                    } else {
                        if (d != pl) {   // head object bound above: one grounding settles it
#pragma unroll
                            for (int k = 1; k < kLeafGeneric; ++k)
                                if (hos[0] < 0) hos[0] = hos[k];
#pragma unroll
                            for (int k = 1; k < kLeafGeneric; ++k) hos[k] = -1;
                            if (any) {
                                if (pl >= base) d = pl;   // head object settled: next candidate at the pair level
                                else busy = false;
                            }
                        }
                        mine += mark_batch<D, kLeafGeneric>(dd, hos, rs, emits);
                    }
                } else {
                    W.frame[d][lane].i = fr.i + 1;
                    const int x = G.col[fr.i];
                    bool ok = bind_entry_x(P.lev[d], G, R, v, fr.i, x);
                    if (ok && level_kind(P.lev[d].flags) == kIsect) ok = isect_next(W, d, G, x);
                    if (ok && pairs && d == pl) {
                        hr = dd.rank_at(fr.i, term_value(v, hoRef));
                        ok = !dd.seen_r(hr);
                    }
                    if (ok) {
                        if (d == last) {
                            if (pairs) {
                                mine += dd.insert_r(hr);
                                emit = term_value(v, hoRef);
                                if (pl >= base) d = pl;   // head object settled: next candidate at the pair level
                                else busy = false;
                            } else {
                                found = true;
                                busy = false;
                            }
                        } else {
                            ++d;
                            W.frame[d][lane] = enter(W, P, d, G, v);
                        }
                    }
                }
            }
        }
        if (D == kDedupSort && pairs) {
            emit_keys(sink, keyHi, emit);
#pragma unroll
            for (int k = 0; k < kLeafGeneric; ++k) emit_keys(sink, keyHi, emits[k]);
        }
        GR_PROF_ADD(4, tp);
        tp = GR_PROF_NOW();
        if (!pairs && __any(found)) break;
        // ---- C: done, hand the seed over as a giant, or donate work to the idle lanes
        wave_sync();
        const uint64_t busyM = __ballot(busy);
        const bool poolEmpty = uniform(W.poolNext) >= uniform(W.poolEnd);
        if (busyM == 0 && poolEmpty) break;
        ++steps;
        busyLanes += (uint64_t)__popcll(busyM);
        if (mayGiant && steps >= ga.after) {
            mayGiant = false;   // one try per seed
            uint32_t slot = 0xffffffffu;
            if (lane == 0) {
                const uint32_t gi = atomicAdd(ga.count, 1u);
                if (gi < ga.cap) {
                    ga.list[gi] = self;
                    ga.heavyIdx[gi] = heavyIdx;
                    slot = gi;
                }
            }
            if (bcast(slot) != 0xffffffffu) {
                aborted = true;
                break;
            }
        }
        if (foundFlag && (steps & 31) == 0) {   // another sub-item of this giant found a grounding
            uint32_t f = lane == 0 ? ld_rlx(foundFlag) : 0u;
            if (bcast(f)) {
                stopped = true;
                break;
            }
        }
        if (busyM != ~0ull && poolEmpty) {
            int rem = 0, lvl = 0;
            if (busy) {
                for (int l = base; l <= d; ++l) {
                    const Frame f = W.frame[l][lane];
                    if (f.end > f.i) {
                        rem = f.end - f.i;
                        lvl = l;
                        break;
                    }
                }
            }
            uint64_t key = ((uint64_t)(uint32_t)rem << 8) | (uint64_t)lane;
#pragma unroll
            for (int off = 32; off > 0; off >>= 1) {
                const uint64_t o = __shfl_xor(key, off, kWave);
                key = o > key ? o : key;
            }
            const int donor = (int)(key & 255u);
            if ((key >> 8) != 0) {
                ++nDonate;
                if (lane == donor) {
                    const Frame f = W.frame[lvl][lane];
                    W.poolLevel = lvl;
                    W.poolNext = f.i;
                    W.poolEnd = f.end;
                    W.poolGen = W.poolGen + 1u;
#pragma unroll
                    for (int j = 0; j < NS; ++j) W.poolV[j] = v[j];
                    W.frame[lvl][lane].end = f.i;
                }
                wave_sync();
            }
        }
        GR_PROF_ADD(5, tp);
    }
    const unsigned long long tEnd = GR_PROF_NOW();
    stats[0] += 1;
    stats[1] += (unsigned long long)steps;
    stats[2] += nDonate;
    stats[3] += nLeafDraw;
    stats[4] += nDraw;
    stats[5] += busyLanes;
    stats[6] += drawLanes;
    return seed_finish<K, D, NS>(dd, G, rowWords, pairs, aborted, stopped, found, foundFlag, mine, supOn, supRow,
                                 supList, supSubject, counts, supOut, stats, tEnd);
}

// Sets up the pool of heavy seed `item` (its level-0 row) and runs it.
template <int K, int D, int NS>
__device__ __forceinline__ uint64_t wave_seed(WaveLdsT<levels_of(NS)>& W, const DevGraph& G, const ConflictRefs* R,
                                              uint32_t item, Dedup<D, NS>& dd, int64_t rowWords, const SortSink& sink,
                                              unsigned long long keyHi, const GiantArgs& ga, uint32_t planIdx,
                                              uint32_t heavyIdx, bool fuse, uint64_t& sup,
                                              unsigned long long* counts, const int32_t* supTab,
                                              unsigned long long (&stats)[kWaveStats]) {
    const int lane = lane_id();
    const Plan& P = W.plan;
    const unsigned long long tp = GR_PROF_NOW();
    int v[NS];
    if (!bind_seed(P, G, (int)item, v)) return 0;
    const Frame f0 = enter(W, P, 0, G, v);
    if (f0.i >= f0.end) return 0;
    if (lane < NS) W.poolV[lane] = term_value(v, lane);
    if (lane == 0) {
        W.poolLevel = 0;
        W.poolNext = f0.i;
        W.poolEnd = f0.end;
        W.poolGen = W.poolGen + 1u;
        W.nTouched = 0;
    }
    if (D == kDedupSort && K == kPairs)
        for (int k = lane; k < kCacheSlots; k += kWave) dd.cache[k] = -1;
    wave_sync();
    bool aborted = false;
    // fused support: the seed's head row, looked up now so its loads overlap the search
    const bool supOn = K == kPairs && D != kDedupSort && fuse && P.supPred != -1;
    const int32_t* supList = (supOn && P.supPred <= -2) ? supTab + (-P.supPred - 2) : nullptr;
    const Frame supRow = (supOn && !supList) ? key_row(G, sup_base(P, G), v[0]) : Frame{0, 0};
    GR_PROF_ADD(1, tp);
    return wave_run<K, D, NS>(W, G, R, dd, rowWords, sink, keyHi, ga, ga.after > 0, HeavyItem{planIdx, item}, heavyIdx,
                          nullptr, aborted, supOn, supRow, supList, uniform(v[0]), counts, sup, stats);
}

// ------------------------------------------------- level-synchronous wave mode
//
// The second wave-mode engine (WaveArgs::engine = 1). The pool engine above
// gives every lane a subtree of its own, so the lanes of a wave stand at
// different levels: a step runs the code of every level some lane is at, and
// each lane's step is a chain of dependent loads. Here the whole wave expands
// one level per step:
//   * every level keeps a ring of open ITEMS: an entry range of the level
//     (an EXT/ISECT/SCAN row, a CHECK's match), the bindings above it and,
//     below the pair level, the rank of the head object already bound;
//   * a step takes the deepest level with items -- depth first, so a ring
//     never holds more than the wave's children of one step -- or the level
//     above it while the deepest has less work than kLsFill candidates and
//     room for more, and deals that level's next entries to the lanes, one
//     per lane, balanced over its items (an inclusive scan of the items'
//     remaining lengths and a binary search of it per lane): every lane runs
//     the same level's code on a candidate of its own;
//   * candidates that bind become items of the next level, appended in lane
//     order, or, at the last level, groundings: marks in the seed's dedup set
//     (PAIRS) or the end of the search (existence). A CHECK level without
//     conflict checks binds nothing and its one entry always passes, so a
//     child at such a level goes on to the level below it at once;
//   * a last level that is a plain EXT binding the head object (or, for
//     existence, any plain EXT) is tested with the straight-line leaf test:
//     an item at least kLsStream entries long is streamed by the whole wave,
//     up to kLsStreamU entries per lane per step; shorter items are cut into
//     chunks of kLsLeafChunk entries, one chunk per lane;
//   * an ISECT level's two rows both list candidates of its new variable:
//     an item enumerates the shorter and searches the other;
//   * below the pair level, an item whose head object is counted already is
//     dropped, as the pool engine drops such subtrees.
// The dedup set, giants, found flags and the end of a seed are shared with
// the pool engine.

// Open items per level: at most one per lane, and a power of two.
#ifndef GR_LS_CAP4
#define GR_LS_CAP4 64
#endif
#ifndef GR_LS_STREAM_U
#define GR_LS_STREAM_U 16
#endif
constexpr int ls_cap(int ns) { return ns <= 4 ? GR_LS_CAP4 : (ns <= 8 ? 32 : 16); }
constexpr int kLsStream = 512;
constexpr int kLsStreamU = GR_LS_STREAM_U;   // entries per lane per step of a streamed item
constexpr int kLsFill = 32;
#ifndef GR_LS_CHUNK
#define GR_LS_CHUNK 8
#endif
constexpr int kLsLeafChunk = GR_LS_CHUNK;   // last-level entries a lane takes from one item per step (chunked leaf steps)
#ifndef GR_LS_ISECT_U
#define GR_LS_ISECT_U 2
#endif
// ISECT candidates a lane takes per step where every candidate found ends as a grounding
constexpr int kLsIsectU = GR_LS_ISECT_U;

template <int NS>
struct LsLdsT {
    static constexpr bool kLevelSync = true;
    static constexpr int NL = levels_of(NS);
    static constexpr int C = ls_cap(NS);
    Plan plan;
    int i0, e0, h0;           // level 0's one item (a seed's level-0 row, or a sub-item's range there)
    uint32_t nTouched;        // HBM rows: dedup-row words listed so far
    int oi0, oe0;
    int v0[NS];
    int itI[NL - 1][C];       // levels 1..NL-1, per ring slot: next entry,
    int itEnd[NL - 1][C];     // end of the range,
    int itHr[NL - 1][C];      // rank of the head object bound above the level (PAIRS), else -1,
    int itOI[NL - 1][C];      // ISECT: the row the item searches (the one it does not enumerate):
    int itOE[NL - 1][C];      //   begin, and end (~end when the item enumerates row 2),
    int itV[NL - 1][NS][C];   // bindings above the level
    int own[kWave];           // a step's deal: the item whose candidates start at each position
};

// One level's item arrays (pointers chosen once per wave-uniform level):
// entry k of slot s's bindings is v[k * vs + s]. Level 0 holds one item, at
// slot 0.
template <int NS>
struct LsLevel {
    int* i;
    int* e;
    int* hr;
    int* oi;
    int* oe;
    int* v;
    int vs;
};
template <int NS>
__device__ __forceinline__ LsLevel<NS> ls_level(LsLdsT<NS>& W, int L) {
    if (L == 0) return LsLevel<NS>{&W.i0, &W.e0, &W.h0, &W.oi0, &W.oe0, W.v0, 1};
    return LsLevel<NS>{W.itI[L - 1], W.itEnd[L - 1], W.itHr[L - 1], W.itOI[L - 1], W.itOE[L - 1], &W.itV[L - 1][0][0],
                       LsLdsT<NS>::C};
}

template <int NS>
__device__ __forceinline__ int* ls_i(LsLdsT<NS>& W, int L, int s) { return L == 0 ? &W.i0 : &W.itI[L - 1][s]; }
template <int NS>
__device__ __forceinline__ int* ls_e(LsLdsT<NS>& W, int L, int s) { return L == 0 ? &W.e0 : &W.itEnd[L - 1][s]; }
template <int NS>
__device__ __forceinline__ int* ls_hr(LsLdsT<NS>& W, int L, int s) { return L == 0 ? &W.h0 : &W.itHr[L - 1][s]; }
template <int NS>
__device__ __forceinline__ int* ls_oi(LsLdsT<NS>& W, int L, int s) { return L == 0 ? &W.oi0 : &W.itOI[L - 1][s]; }
template <int NS>
__device__ __forceinline__ int* ls_oe(LsLdsT<NS>& W, int L, int s) { return L == 0 ? &W.oe0 : &W.itOE[L - 1][s]; }
template <int NS>
__device__ __forceinline__ int* ls_v(LsLdsT<NS>& W, int L, int k, int s) {
    return L == 0 ? &W.v0[k] : &W.itV[L - 1][k][s];
}

// Wave-uniform ring state, one value per level (select chains, no scratch).
template <int NL>
__device__ __forceinline__ int rget(const int (&a)[NL], int L) {
    int x = 0;
#pragma unroll
    for (int l = 0; l < NL; ++l) x = l == L ? a[l] : x;
    return x;
}
template <int NL>
__device__ __forceinline__ void radd(int (&a)[NL], int L, int d) {
#pragma unroll
    for (int l = 0; l < NL; ++l) a[l] += l == L ? d : 0;
}

// Inclusive sum over lanes 0..lane, with DPP moves (no LDS round trip):
// shifts by 1, 2, 4, 8 within each 16-lane row, then the last lane of row 0
// (of row 2) added to row 1 (row 3), then the last lane of row 1 added to
// rows 2 and 3. A lane whose DPP source lies outside its row, and the rows a
// step masks off, add 0.
__device__ __forceinline__ int wave_incl_scan(int x) {
    x += __builtin_amdgcn_update_dpp(0, x, 0x111, 0xf, 0xf, false);   // row_shr:1
    x += __builtin_amdgcn_update_dpp(0, x, 0x112, 0xf, 0xf, false);   // row_shr:2
    x += __builtin_amdgcn_update_dpp(0, x, 0x114, 0xf, 0xf, false);   // row_shr:4
    x += __builtin_amdgcn_update_dpp(0, x, 0x118, 0xf, 0xf, false);   // row_shr:8
    x += __builtin_amdgcn_update_dpp(0, x, 0x142, 0xa, 0xf, false);   // row_bcast:15 -> rows 1, 3
    x += __builtin_amdgcn_update_dpp(0, x, 0x143, 0xc, 0xf, false);   // row_bcast:31 -> rows 2, 3
    return x;
}

// Inclusive maximum over lanes 0..lane, with the same DPP moves.
__device__ __forceinline__ int wave_incl_max(int x) {
    constexpr int lo = -2147483647 - 1;
    x = max(x, __builtin_amdgcn_update_dpp(lo, x, 0x111, 0xf, 0xf, false));
    x = max(x, __builtin_amdgcn_update_dpp(lo, x, 0x112, 0xf, 0xf, false));
    x = max(x, __builtin_amdgcn_update_dpp(lo, x, 0x114, 0xf, 0xf, false));
    x = max(x, __builtin_amdgcn_update_dpp(lo, x, 0x118, 0xf, 0xf, false));
    x = max(x, __builtin_amdgcn_update_dpp(lo, x, 0x142, 0xa, 0xf, false));
    x = max(x, __builtin_amdgcn_update_dpp(lo, x, 0x143, 0xc, 0xf, false));
    return x;
}

// The items of level L (ring head h, n of them): lane j < n loads item j
// and returns its remaining length. With dropBound, an item whose head
// object is counted already has nothing left to add: it is marked consumed
// and reported empty.
template <int D, int NS>
__device__ __forceinline__ int ls_items(LsLdsT<NS>& W, const Dedup<D, NS>& dd, int L, int h, int n, bool dropBound,
                                        int& ii, int& ie, int& ihr, int& sj) {
    constexpr int C = LsLdsT<NS>::C;
    const int lane = lane_id();
    const LsLevel<NS> lv = ls_level(W, L);
    sj = (h + lane) & (C - 1);
    ii = 0;
    ie = 0;
    ihr = -1;
    if (lane < n) {
        ii = lv.i[sj];
        ie = lv.e[sj];
        ihr = lv.hr[sj];
    }
    int rem = ie - ii;
    if (dropBound && rem > 0 && dd.seen_r(ihr)) {
        rem = 0;
        lv.i[sj] = ie;
    }
    return rem;
}

// The deal of a step: positions 0..T-1 go to the items in order, item j
// (j < n, the lane holding it) taking cnt_j positions from excl_j on. Returns
// the item of position `lane` (for lane < T): each item with positions
// writes its index at its first position, and an inclusive maximum over the
// positions carries it forward.
template <int NS>
__device__ __forceinline__ int ls_deal(LsLdsT<NS>& W, int n, int cnt, int excl, int T) {
    const int lane = lane_id();
    W.own[lane] = -1;
    wave_sync();
    if (lane < n && cnt > 0 && excl < T) W.own[excl] = lane;
    wave_sync();
    const int j = wave_incl_max(W.own[lane]);
    return j < 0 ? 0 : j;
}

// Entry range of an item of level lv under bindings v. An ISECT level's two
// rows both list candidates of its new variable, sorted: the item enumerates
// the shorter one and searches the other, returned in `other` with its end
// as ~end when the item enumerates row 2 (other levels: {0, 0}).
template <int NS>
__device__ __forceinline__ Frame ls_enter(const Level& lv, const DevGraph& G, const int (&v)[NS], Frame& other) {
    other = Frame{0, 0};
    const Frame r1 = enter_level(lv, G, v);
    if (level_kind(lv.flags) != kIsect || r1.i >= r1.end) return r1;
    const Frame r2 = isect_row(lv, G, v);
    if (r2.i >= r2.end) return Frame{0, 0};
    if (r2.end - r2.i < r1.end - r1.i) {
        other = Frame{r1.i, ~r1.end};
        return r2;
    }
    other = r2;
    return r1;
}

// Runs the search from the one item the caller placed at level `start`
// (ring slot 0), with the same contract as wave_run.
template <int K, int D, int NS>
__device__ __forceinline__ uint64_t ls_run(LsLdsT<NS>& W, const DevGraph& G, const ConflictRefs* R, Dedup<D, NS>& dd,
                                           int64_t rowWords, const SortSink& sink, unsigned long long keyHi,
                                           const GiantArgs& ga, bool mayGiant, HeavyItem self, uint32_t heavyIdx,
                                           uint32_t* foundFlag, bool& aborted, bool supOn, Frame supRow,
                                           const int32_t* supList, int supSubject, unsigned long long* counts,
                                           uint64_t& supOut, unsigned long long (&stats)[kWaveStats], int start) {
    constexpr int NL = LsLdsT<NS>::NL;
    constexpr int C = LsLdsT<NS>::C;
    const int lane = lane_id();
    const Plan& P = W.plan;
    const int nlev = uniform(P.nlev);
    const int last = nlev - 1;
    const int pl = uniform(P.pairLevel);
    const int hoRef = uniform(P.hoRef);
    const bool pairs = K == kPairs && pl >= 0;
    const bool fastLeaf = leaf_fast_ok<D>(P, last, pairs);
    const bool byEntry = D != kDedupSort && uniform(P.hoByEntry) != 0;   // head-object ranks from crank
    // an ISECT level right above its conflict-free CHECK as the last level: every
    // candidate it finds is a grounding, so its steps keep no children (bitmap
    // dedup or existence: the sort path emits keys wave by wave)
    const bool isectLeaf = nlev >= 2 && level_kind(P.lev[nlev >= 2 ? last - 1 : 0].flags) == kIsect &&
                           level_conflicts(P.lev[last].flags) == 0 && (!pairs || D != kDedupSort);
    const uint64_t below = (1ull << lane) - 1ull;
    int hd[NL], ct[NL];
#pragma unroll
    for (int l = 0; l < NL; ++l) {
        hd[l] = 0;
        ct[l] = l == start ? 1 : 0;
    }
    int deep = start;
    bool found = false, stopped = false;
    aborted = false;
    uint64_t mine = 0;
    int steps = 0;
    uint32_t nFill = 0, nStream = 0, nBal = 0;
    uint64_t nCand = 0, nPass = 0;
    for (;;) {
        unsigned long long tp = GR_PROF_NOW();
        while (deep >= start && rget(ct, deep) == 0) --deep;
        if (deep < start) break;
        int X = deep;
        bool streamed = false;
        if (X == last && fastLeaf) {
            // a long last-level item: the whole wave streams it
            const int s = rget(hd, X) & (C - 1);
            const int i0 = uniform(*ls_i(W, X, s));
            const int e0 = uniform(*ls_e(W, X, s));
            if (e0 - i0 >= kLsStream) {
                streamed = true;
                int v[NS];
#pragma unroll
                for (int k = 0; k < NS; ++k) v[k] = uniform(*ls_v(W, X, k, s));
                const int lim = e0 - i0 > kWave * kLsStreamU ? i0 + kWave * kLsStreamU : e0;
                int is[kLsStreamU], xs[kLsStreamU], rs[kLsStreamU];
                bool in[kLsStreamU];
#pragma unroll
                for (int k = 0; k < kLsStreamU; ++k) {
                    is[k] = i0 + lane + kWave * k;
                    in[k] = is[k] < lim;
                }
                const uint32_t m = leaf_fast(G, v, level_nbound(P.lev[last].flags), pairs, is, in, i0, xs, rs);
                if (!pairs) found = found || m != 0u;
                else mine += mark_ranks<D, kLsStreamU>(dd, rs);
                nCand += (uint64_t)(lim - i0);
                if (lane == 0) *ls_i(W, X, s) = lim;
                if (lim == e0) {
                    radd(hd, X, 1);
                    radd(ct, X, -1);
                }
                ++nStream;
                GR_PROF_ADD(2, tp);
            }
        }
        if (!streamed && isectLeaf && X == last - 1) {
            // An ISECT level whose candidates end as groundings: with no children to
            // keep, a step deals up to kLsIsectU candidates per lane (positions lane,
            // lane + 64, ...), and their row searches run in lockstep, so a lane has
            // that many independent loads in flight.
            streamed = true;
            constexpr int U = kLsIsectU;
            int ii, ie, ihr, sj;
            const int n = rget(ct, X);
            const int rem = ls_items(W, dd, X, rget(hd, X), n, pairs && pl < X, ii, ie, ihr, sj);
            const int incl = wave_incl_scan(rem);
            const int total = uniform(__shfl(incl, kWave - 1, kWave));
            const int T = total < kWave * U ? total : kWave * U;
            const int excl = incl - rem;
            int took = T - excl;
            took = took < 0 ? 0 : (took > rem ? rem : took);
            const LsLevel<NS> lx = ls_level(W, X);
            if (lane < n && took > 0) lx.i[sj] = ii + took;
            const uint64_t doneM = __ballot(lane < n && rem == took);
            const int pre = doneM == ~0ull ? kWave : (int)__builtin_ctzll(~doneM);
            const Level& lv = P.lev[X];
            // the item of each position, window by window (an item starting before a
            // window carries into it)
            int jk[U];
            int carry = -1;
#pragma unroll
            for (int k = 0; k < U; ++k) {
                W.own[lane] = -1;
                wave_sync();
                const int rel = excl - kWave * k;
                if (lane < n && rem > 0 && excl < T && rel >= 0 && rel < kWave) W.own[rel] = lane;
                wave_sync();
                int m = wave_incl_max(W.own[lane]);
                m = m > carry ? m : carry;
                jk[k] = m < 0 ? 0 : m;
                carry = uniform(__shfl(m, kWave - 1, kWave));
                wave_sync();
            }
            int x[U], e[U], hrk[U], lo[U], hi[U], vl[U], vh[U], res[U], swpk[U];
            bool ok[U], go[U];
#pragma unroll
            for (int k = 0; k < U; ++k) {
                const int j = jk[k];
                const int jx = __shfl(excl, j, kWave);
                const int ji = __shfl(ii, j, kWave);
                const int js = __shfl(sj, j, kWave);
                hrk[k] = __shfl(ihr, j, kWave);
                const int q = lane + kWave * k;
                const bool act = q < T;
                e[k] = act ? ji + (q - jx) : 0;
                x[k] = G.col[e[k]];
                int v[NS];
#pragma unroll
                for (int t = 0; t < NS; ++t) v[t] = act ? lx.v[t * lx.vs + js] : -1;
                ok[k] = act && bind_entry_x(lv, G, R, v, e[k], x[k]);
                int r2i = 0, r2e = 1;
                swpk[k] = 0;
                if (act) {
                    r2i = lx.oi[js];
                    r2e = lx.oe[js];
                    swpk[k] = r2e < 0 ? 1 : 0;
                    r2e = swpk[k] ? ~r2e : r2e;
                }
                lo[k] = r2i;
                hi[k] = r2e - 1;
            }
            // the searched rows' end values, then up to three interpolation probes, all in lockstep
#pragma unroll
            for (int k = 0; k < U; ++k) {
                vl[k] = G.col[lo[k]];
                vh[k] = G.col[hi[k]];
            }
#pragma unroll
            for (int k = 0; k < U; ++k) {
                res[k] = -1;
                go[k] = ok[k] && x[k] >= vl[k] && x[k] <= vh[k];
                if (go[k] && x[k] == vl[k]) {
                    res[k] = lo[k];
                    go[k] = false;
                }
                if (go[k] && x[k] == vh[k]) {
                    res[k] = hi[k];
                    go[k] = false;
                }
            }
#pragma unroll
            for (int it = 0; it < 3; ++it) {
                int m[U], y[U];
#pragma unroll
                for (int k = 0; k < U; ++k) {
                    const bool p = go[k] && hi[k] - lo[k] > 8;
                    int mm = lo[k];
                    if (p) {
                        const float f = (float)(x[k] - vl[k]) / (float)(vh[k] - vl[k]);
                        mm = lo[k] + (int)(f * (float)(hi[k] - lo[k]));
                        mm = mm <= lo[k] ? lo[k] + 1 : (mm >= hi[k] ? hi[k] - 1 : mm);
                    }
                    m[k] = mm;
                }
#pragma unroll
                for (int k = 0; k < U; ++k) y[k] = G.col[m[k]];
#pragma unroll
                for (int k = 0; k < U; ++k) {
                    if (!(go[k] && hi[k] - lo[k] > 8)) continue;
                    if (y[k] == x[k]) {
                        res[k] = m[k];
                        go[k] = false;
                    } else if (y[k] < x[k]) {
                        lo[k] = m[k];
                        vl[k] = y[k];
                    } else {
                        hi[k] = m[k];
                        vh[k] = y[k];
                    }
                }
            }
            // binary search of (lo, hi), in lockstep
            int a[U], b[U];
#pragma unroll
            for (int k = 0; k < U; ++k) {
                a[k] = lo[k] + 1;
                b[k] = hi[k];
            }
            for (;;) {
                bool any = false;
#pragma unroll
                for (int k = 0; k < U; ++k) any = any || (go[k] && a[k] < b[k]);
                if (!__any(any)) break;
                int m[U], y[U];
#pragma unroll
                for (int k = 0; k < U; ++k) m[k] = (go[k] && a[k] < b[k]) ? (a[k] + b[k]) >> 1 : lo[k];
#pragma unroll
                for (int k = 0; k < U; ++k) y[k] = G.col[m[k]];
#pragma unroll
                for (int k = 0; k < U; ++k) {
                    if (!(go[k] && a[k] < b[k])) continue;
                    if (y[k] < x[k]) a[k] = m[k] + 1;
                    else b[k] = m[k];
                }
            }
            {
                int y[U];
#pragma unroll
                for (int k = 0; k < U; ++k) y[k] = G.col[go[k] && a[k] < hi[k] ? a[k] : lo[k]];
#pragma unroll
                for (int k = 0; k < U; ++k)
                    if (go[k] && a[k] < hi[k] && y[k] == x[k]) res[k] = a[k];
            }
            // groundings
#pragma unroll
            for (int k = 0; k < U; ++k) {
                if (res[k] < 0) continue;
                if (!pairs) {
                    found = true;
                } else if (X == pl) {   // the ISECT binds the head object: each match is one
                    const int e1 = swpk[k] ? res[k] : e[k];
                    const int h1 = byEntry ? G.crank[e1] : dd.rank(x[k]);
                    if (!dd.seen_r(h1)) mine += dd.insert_r(h1);
                } else {                // head object bound above: the grounding settles it
                    mine += dd.insert_r(hrk[k]);
                }
            }
            nCand += (uint64_t)T;
            radd(hd, X, pre);
            radd(ct, X, -pre);
            ++nBal;
            GR_PROF_ADD(4, tp);
        }
        if (!streamed && X == last && fastLeaf) {
            // Last level, plain EXT: the items are cut into chunks of kLsLeafChunk
            // entries and each lane takes one chunk (the wave takes the first 64
            // chunks in item order), so a lane's candidates share one item's
            // bindings and up to 64 * kLsLeafChunk candidates are tested per step.
            streamed = true;
            constexpr int U = kLsLeafChunk;
            int ii, ie, ihr, sj;
            const int n = rget(ct, X);
            const int rem = ls_items(W, dd, X, rget(hd, X), n, false, ii, ie, ihr, sj);
            const int cnk = (rem + U - 1) / U;
            const int incl = wave_incl_scan(cnk);
            const int totalC = uniform(__shfl(incl, kWave - 1, kWave));
            const int TC = totalC < kWave ? totalC : kWave;
            const int excl = incl - cnk;
            int tookC = TC - excl;
            tookC = tookC < 0 ? 0 : (tookC > cnk ? cnk : tookC);
            const int took = tookC * U < rem ? tookC * U : rem;
            const LsLevel<NS> lx = ls_level(W, X);
            if (lane < n && took > 0) lx.i[sj] = ii + took;
            const uint64_t doneM = __ballot(lane < n && rem == took);
            const int pre = doneM == ~0ull ? kWave : (int)__builtin_ctzll(~doneM);
            const int j = ls_deal(W, n, cnk, excl, TC);
            const int jx = __shfl(excl, j, kWave);
            const int ji = __shfl(ii, j, kWave);
            const int je = __shfl(ie, j, kWave);
            const int js = __shfl(sj, j, kWave);
            const bool act = lane < TC;
            const int start = act ? ji + (lane - jx) * U : 0;
            const int end = act ? (start + U < je ? start + U : je) : 0;
            int vb[NS];
#pragma unroll
            for (int k = 0; k < NS; ++k) vb[k] = act ? lx.v[k * lx.vs + js] : -1;
            int is[U], xs[U], rs[U];
            bool in[U];
#pragma unroll
            for (int k = 0; k < U; ++k) {
                is[k] = start + k;
                in[k] = is[k] < end;
            }
            const uint32_t m = leaf_fast(G, vb, level_nbound(P.lev[last].flags), pairs, is, in, start, xs, rs);
            if (!pairs) found = found || m != 0u;
            else mine += mark_ranks<D, U>(dd, rs);
            nCand += (uint64_t)(TC * U);   // candidate slots dealt (the last chunk of an item may be short)
            radd(hd, X, pre);
            radd(ct, X, -pre);
            ++nBal;
            GR_PROF_ADD(3, tp);
        }
        if (!streamed) {
            int ii, ie, ihr, sj;
            int rem = ls_items(W, dd, X, rget(hd, X), rget(ct, X), pairs && pl < X, ii, ie, ihr, sj);
            int incl = wave_incl_scan(rem);
            int total = uniform(__shfl(incl, kWave - 1, kWave));
            if (total < kLsFill && X > start && rget(ct, X - 1) > 0 && C - rget(ct, X) >= kLsFill) {
                // little work at the deepest level and room for more: expand the level above first
                --X;
                rem = ls_items(W, dd, X, rget(hd, X), rget(ct, X), pairs && pl < X, ii, ie, ihr, sj);
                incl = wave_incl_scan(rem);
                total = uniform(__shfl(incl, kWave - 1, kWave));
                ++nFill;
            }
            const int n = rget(ct, X);
            const int room = X < last ? (C - rget(ct, X + 1) < kWave ? C - rget(ct, X + 1) : kWave) : kWave;
            const int T = total < room ? total : room;
            const int excl = incl - rem;
            int took = T - excl;
            took = took < 0 ? 0 : (took > rem ? rem : took);
            const LsLevel<NS> lx = ls_level(W, X);
            if (lane < n && took > 0) lx.i[sj] = ii + took;
            const uint64_t doneM = __ballot(lane < n && rem == took);
            const int pre = doneM == ~0ull ? kWave : (int)__builtin_ctzll(~doneM);
            // the item of this lane's candidate
            const int j = ls_deal(W, n, rem, excl, T);
            const int jx = __shfl(excl, j, kWave);
            const int ji = __shfl(ii, j, kWave);
            const int jhr = __shfl(ihr, j, kWave);
            const int js = __shfl(sj, j, kWave);
            const bool act = lane < T;
            const int e = act ? ji + (lane - jx) : 0;
            const Level& lv = P.lev[X];
            const int kind = uniform(level_kind(lv.flags));
            const int x = G.col[e];
            // ISECT: the row the candidate's item searches, with its first and last
            // values (loads shared by the item's lanes)
            int r2i = 0, r2e = 0, swp = 0, r2lo = 0, r2hi = -1;
            if (kind == kIsect && act) {
                r2i = lx.oi[js];
                r2e = lx.oe[js];
                swp = r2e < 0 ? 1 : 0;
                r2e = swp ? ~r2e : r2e;
                r2lo = G.col[r2i];
                r2hi = G.col[r2e - 1];
            }
            int v[NS];
#pragma unroll
            for (int k = 0; k < NS; ++k) v[k] = act ? lx.v[k * lx.vs + js] : -1;
            bool ok = act && bind_entry_x(lv, G, R, v, e, x);
            int at = -1;    // ISECT: the match's entry in row 2 (the CHECK level's row)
            int e1 = e;     // the candidate's entry in the level's own row (row 1)
            if (kind == kIsect && ok) {
                const int m = find_in_row_interp(G.col, r2i, r2e, x, r2lo, r2hi);
                ok = m >= 0;
                at = swp ? e : m;
                e1 = swp ? m : e;
            }
            int hr = jhr;
            if (pairs && X == pl && ok) {
                hr = byEntry ? G.crank[e1] : dd.rank(term_value(v, hoRef));
                ok = !dd.seen_r(hr);
            }
            nCand += (uint64_t)T;
            nPass += (uint64_t)__popcll(__ballot(ok));
            if (X == last) {
                if (!pairs) {
                    found = found || ok;
                } else if (D == kDedupSort) {
                    int emit = -1;
                    if (ok) {
                        const int ho = term_value(v, hoRef);
                        if (!dd.seen(ho)) {
                            dd.insert(ho);
                            emit = ho;
                        }
                    }
                    emit_keys(sink, keyHi, emit);
                } else if (ok) {
                    mine += dd.insert_r(hr);
                }
            } else {
                int c = X + 1;
                Frame f{0, 0}, fo{0, 0};
                if (ok) {
                    const Level& lc = P.lev[c];
                    f = (level_kind(lc.flags) == kCheck && kind == kIsect) ? Frame{at, at + 1} : ls_enter(lc, G, v, fo);
                }
                bool has = ok && f.i < f.end;
                // A CHECK level without conflict checks binds nothing, and its one entry
                // (the match enter_level found) always passes: go on to the level below
                // it at once, or, below the last level, count the grounding.
                bool ground = false;
                while (has && level_kind(P.lev[c].flags) == kCheck && level_conflicts(P.lev[c].flags) == 0) {
                    if (c == last) {
                        ground = true;
                        has = false;
                        break;
                    }
                    ++c;
                    f = ls_enter(P.lev[c], G, v, fo);
                    has = f.i < f.end;
                }
                // the head object of such a grounding is bound at or above X (no CHECK binds it)
                if (!pairs) {
                    found = found || ground;
                } else if (D == kDedupSort) {
                    int emit = -1;
                    if (ground) {
                        const int ho = term_value(v, hoRef);
                        if (!dd.seen(ho)) {
                            dd.insert(ho);
                            emit = ho;
                        }
                    }
                    emit_keys(sink, keyHi, emit);
                } else if (ground) {
                    mine += dd.insert_r(hr);
                }
                // children, appended level by level in lane order (below X + 1 only after a CHECK passed)
                const int cEnd = __any(has && c > X + 1) ? last : X + 1;
                for (int cc = X + 1; cc <= cEnd; ++cc) {
                    const uint64_t cm = __ballot(has && c == cc);
                    if (cm == 0) continue;
                    if (has && c == cc) {
                        const LsLevel<NS> lc = ls_level(W, cc);
                        const int s = (rget(hd, cc) + rget(ct, cc) + (int)__popcll(cm & below)) & (C - 1);
                        lc.i[s] = f.i;
                        lc.e[s] = f.end;
                        lc.hr[s] = hr;
                        lc.oi[s] = fo.i;
                        lc.oe[s] = fo.end;
#pragma unroll
                        for (int k = 0; k < NS; ++k) lc.v[k * lc.vs + s] = v[k];
                    }
                    radd(ct, cc, (int)__popcll(cm));
                    if (cc > deep) deep = cc;
                }
            }
            radd(hd, X, pre);
            radd(ct, X, -pre);
            ++nBal;
            GR_PROF_ADD(X == last ? 3 : 4, tp);
        }
        tp = GR_PROF_NOW();
        wave_sync();
        ++steps;
        if (!pairs && __any(found)) break;
        if (mayGiant && steps >= ga.after) {
            mayGiant = false;   // one try per seed
            uint32_t slot = 0xffffffffu;
            if (lane == 0) {
                const uint32_t gi = atomicAdd(ga.count, 1u);
                if (gi < ga.cap) {
                    ga.list[gi] = self;
                    ga.heavyIdx[gi] = heavyIdx;
                    slot = gi;
                }
            }
            if (bcast(slot) != 0xffffffffu) {
                aborted = true;
                break;
            }
        }
        if (foundFlag && (steps & 31) == 0) {   // another sub-item of this giant found a grounding
            uint32_t f = lane == 0 ? ld_rlx(foundFlag) : 0u;
            if (bcast(f)) {
                stopped = true;
                break;
            }
        }
        GR_PROF_ADD(5, tp);
    }
    const unsigned long long tEnd = GR_PROF_NOW();
    stats[0] += 1;
    stats[1] += (unsigned long long)steps;
    stats[2] += nFill;
    stats[3] += nStream;
    stats[4] += nBal;
    stats[5] += nCand;
    stats[6] += nPass;
    return seed_finish<K, D, NS>(dd, G, rowWords, pairs, aborted, stopped, found, foundFlag, mine, supOn, supRow,
                                 supList, supSubject, counts, supOut, stats, tEnd);
}

// Places heavy seed `item` (its level-0 row) as level 0's item and runs it.
template <int K, int D, int NS>
__device__ __forceinline__ uint64_t ls_seed(LsLdsT<NS>& W, const DevGraph& G, const ConflictRefs* R, uint32_t item,
                                            Dedup<D, NS>& dd, int64_t rowWords, const SortSink& sink,
                                            unsigned long long keyHi, const GiantArgs& ga, uint32_t planIdx,
                                            uint32_t heavyIdx, bool fuse, uint64_t& sup, unsigned long long* counts,
                                            const int32_t* supTab, unsigned long long (&stats)[kWaveStats]) {
    const int lane = lane_id();
    const Plan& P = W.plan;
    const unsigned long long tp = GR_PROF_NOW();
    int v[NS];
    if (!bind_seed(P, G, (int)item, v)) return 0;
    Frame o0{0, 0};
    const Frame f0 = ls_enter(P.lev[0], G, v, o0);
    if (f0.i >= f0.end) return 0;
    if (lane < NS) W.v0[lane] = term_value(v, lane);
    if (lane == 0) {
        W.i0 = f0.i;
        W.e0 = f0.end;
        W.h0 = -1;
        W.oi0 = o0.i;
        W.oe0 = o0.end;
        W.nTouched = 0;
    }
    if (D == kDedupSort && K == kPairs)
        for (int k = lane; k < kCacheSlots; k += kWave) dd.cache[k] = -1;
    wave_sync();
    bool aborted = false;
    const bool supOn = K == kPairs && D != kDedupSort && fuse && P.supPred != -1;
    const int32_t* supList = (supOn && P.supPred <= -2) ? supTab + (-P.supPred - 2) : nullptr;
    const Frame supRow = (supOn && !supList) ? key_row(G, sup_base(P, G), v[0]) : Frame{0, 0};
    GR_PROF_ADD(1, tp);
    return ls_run<K, D, NS>(W, G, R, dd, rowWords, sink, keyHi, ga, ga.after > 0, HeavyItem{planIdx, item}, heavyIdx,
                            nullptr, aborted, supOn, supRow, supList, uniform(v[0]), counts, sup, stats, 0);
}

// A giant's sub-item as the one item of its level (ring slot 0): entries
// [begin, end) of `level` with the bindings above it; returns the level.
template <int K, int D, int NS>
__device__ __forceinline__ int ls_place_sub(LsLdsT<NS>& W, const Dedup<D, NS>& dd, const SubItem& S,
                                            const DevGraph& G) {
    const int lane = lane_id();
    const Plan& P = W.plan;
    const int lvl = uniform(S.level);
    // a range of an ISECT level's row 1 (as the giant expansion cuts it) searches row 2
    Frame other{0, 0};
    const bool isect = level_kind(P.lev[lvl].flags) == kIsect;
    if (isect) {
        int v[NS];
#pragma unroll
        for (int k = 0; k < NS; ++k) v[k] = uniform(S.v[k]);
        other = isect_row(P.lev[lvl], G, v);
    }
    const int pl = uniform(P.pairLevel);
    int hr = -1;
    if (K == kPairs && pl >= 0 && pl < lvl) {   // the head object is bound above the sub-item's level
        const int ref = uniform(P.hoRef);
        const int hov = ref >= 0 ? uniform(S.v[ref]) : -(ref + 1);
        hr = uniform(dd.rank(hov));
    }
    if (lane < NS) *ls_v(W, lvl, lane, 0) = S.v[lane];
    if (lane == 0) {
        *ls_i(W, lvl, 0) = S.begin;
        *ls_e(W, lvl, 0) = (isect && other.i >= other.end) ? S.begin : S.end;   // row 2 empty: nothing matches
        *ls_hr(W, lvl, 0) = hr;
        *ls_oi(W, lvl, 0) = other.i;
        *ls_oe(W, lvl, 0) = other.end;
        W.nTouched = 0;
    }
    return lvl;
}

// ------------------------------------------------------------- kernels

// Persistent thread mode over a chunk list: each wave repeatedly takes a
// chunk of seeds of one plan; its lanes take the chunk's seeds one at a time.
// For the 4-slot class the second launch bound asks for registers that let
// 6 waves share a SIMD: its pairs kernel then keeps 78 registers and spills
// 12 bytes per lane. With 8 (64 registers) it spilled 72 bytes and ran 6%
// slower on wide, shallow graphs; with 7, 1% slower. The larger slot
// classes need more registers than that.
constexpr int kThreadMinWaves = 6;
template <int K, int NS>
__global__ void __launch_bounds__(kBlock, NS == 4 ? kThreadMinWaves : 1) k_thread_batch(ThreadArgs a, DevGraph G) {
    __shared__ ThreadLdsT<levels_of(NS)> lds[kWavesPerBlock];
    ThreadLdsT<levels_of(NS)>& W = lds[wave_in_block()];
    const int lane = lane_id();
    uint32_t loaded = 0xffffffffu;
    for (;;) {
        const uint32_t c = wave_take(a.cursor);
        if (c >= a.nChunks) break;
        const Chunk ch = a.chunks[c];
        const uint32_t thr = ch.gate & ~kGateFused;
        if (thr && (!(ch.gate & kGateFused) || !a.fuse) && a.counts[2 * a.plans[ch.plan].rule + kSupport] < thr)
            continue;
        if (ch.plan != loaded) {
            copy_plan(W, a.plans + ch.plan);
            loaded = ch.plan;
        }
        uint64_t sup = 0;
        const uint64_t mine = thread_chunk<K, NS>(W, G, a.refs + ch.plan, ch.plan, ch.begin, ch.count, a.heavy,
                                           a.heavyCount, a.heavyCap, a.fuse != 0, sup, a.counts, a.supTab);
        const uint64_t tot = wave_sum(mine);
        if (lane == 0 && tot) atomicAdd(a.counts + 2 * W.plan.rule + K, (unsigned long long)tot);
        if (K == kPairs && a.fuse && W.plan.supPred >= 0) {
            const uint64_t ts = wave_sum(sup);
            if (lane == 0 && ts) atomicAdd(a.counts + 2 * W.plan.rule + kSupport, (unsigned long long)ts);
        }
    }
}

// One plan passed by value in the kernel arguments (per-rule launches);
// waves stride over its seeds.
template <int K, int NS>
__global__ void __launch_bounds__(kBlock) k_thread_one(Plan P, OneArgs a, DevGraph G) {
    __shared__ ThreadLdsT<levels_of(NS)> lds[kWavesPerBlock];
    ThreadLdsT<levels_of(NS)>& W = lds[wave_in_block()];
    const int lane = lane_id();
    copy_plan(W, &P);
    const uint32_t waves = (gridDim.x * blockDim.x) / kWave;
    for (uint32_t base = ((blockIdx.x * blockDim.x + threadIdx.x) / kWave) * kOneChunkSeeds; base < a.nItems;
         base += waves * kOneChunkSeeds) {
        const uint32_t cnt = a.nItems - base < (uint32_t)kOneChunkSeeds ? a.nItems - base : (uint32_t)kOneChunkSeeds;
        uint64_t sup = 0;
        const uint64_t mine =
            thread_chunk<K, NS>(W, G, a.refs, 0u, base, cnt, a.heavy, a.heavyCount, a.heavyCap, false, sup, nullptr,
                                nullptr);
        const uint64_t tot = wave_sum(mine);
        if (lane == 0 && tot) atomicAdd(a.count, (unsigned long long)tot);
    }
}

// Dedup state of one wave-mode wave: D selects the bitmap in HBM (this
// wave's row), the bitmap in LDS (dynamic shared memory, cleared once here)
// or the sort path's LDS cache.
template <int D, int NS, class Lds>
__device__ __forceinline__ Dedup<D, NS> wave_dedup(Lds& W, int* cache, uint32_t* dyn, const WaveArgs& a,
                                                   const DevGraph& G) {
    Dedup<D, NS> dd;
    dd.plan = &W.plan;
    dd.nTouched = &W.nTouched;
    dd.cache = cache;
    dd.bits = nullptr;
    dd.shared = nullptr;
    dd.merge = nullptr;
    dd.blocks = G.blocks;
    dd.crank = G.crank;
    if (D == kDedupRows) dd.bits = a.rows + (int64_t)blockIdx.x * a.rowWords;
    if (D == kDedupLds) {
        dd.bits = dyn;
        for (int64_t w = lane_id(); w < a.rowWords; w += kWave) dyn[w] = 0u;
    }
    if constexpr (!Lds::kLevelSync)
        if (lane_id() == 0) W.poolGen = 0;
    wave_sync();
    return dd;
}

// Adds a wave's counters to the launch's (once per wave).
__device__ __forceinline__ void flush_stats(const GiantArgs& ga, const unsigned long long (&stats)[kWaveStats]) {
    if (ga.stats && lane_id() == 0)
        for (int q = 0; q < kWaveStats; ++q)
            if (stats[q]) atomicAdd(ga.stats + q, stats[q]);
}

// LDS of a wave-mode block under engine E (0: pool, 1: level-synchronous).
template <int E, int NS>
struct WaveLdsOf {
    using T = WaveLdsT<levels_of(NS)>;
};
template <int NS>
struct WaveLdsOf<1, NS> {
    using T = LsLdsT<NS>;
};

// One heavy seed under engine E.
template <int K, int D, int NS, class Lds>
__device__ __forceinline__ uint64_t run_seed(Lds& W, const DevGraph& G, const ConflictRefs* R, uint32_t item,
                                             Dedup<D, NS>& dd, int64_t rowWords, const SortSink& sink,
                                             unsigned long long keyHi, const GiantArgs& ga, uint32_t planIdx,
                                             uint32_t heavyIdx, bool fuse, uint64_t& sup, unsigned long long* counts,
                                             const int32_t* supTab, unsigned long long (&stats)[kWaveStats]) {
    if constexpr (Lds::kLevelSync)
        return ls_seed<K, D, NS>(W, G, R, item, dd, rowWords, sink, keyHi, ga, planIdx, heavyIdx, fuse, sup, counts,
                                 supTab, stats);
    else
        return wave_seed<K, D, NS>(W, G, R, item, dd, rowWords, sink, keyHi, ga, planIdx, heavyIdx, fuse, sup, counts,
                                   supTab, stats);
}

// Persistent wave mode over the heavy queue: one wave per block, one heavy
// seed at a time.
template <int K, int D, int NS, int E>
__global__ void __launch_bounds__(kWave) k_wave_batch(WaveArgs a, DevGraph G) {
    __shared__ typename WaveLdsOf<E, NS>::T W;
    __shared__ int cache[wave_cache_words<D>()];
    extern __shared__ uint32_t dyn[];
    const int lane = lane_id();
    uint32_t n = *a.heavyCount;
    if (n > a.heavyCap) n = a.heavyCap;
    if (n == 0) return;   // nothing escalated: no cursor traffic
    Dedup<D, NS> dd = wave_dedup<D, NS>(W, cache, dyn, a, G);
    uint32_t loaded = 0xffffffffu;
    unsigned long long stats[kWaveStats] = {};
    for (;;) {
        const unsigned long long tp = GR_PROF_NOW();
        const uint32_t h = wave_take(a.cursor);
        if (h >= n) break;
        const HeavyItem it = a.heavy[h];
        if (it.plan != loaded) {
            copy_plan(W, a.plans + it.plan);
            loaded = it.plan;
        }
        GR_PROF_ADD(0, tp);
        uint64_t sup = 0;
        const uint64_t c = run_seed<K, D, NS>(W, G, a.refs + it.plan, it.item, dd, a.rowWords, a.sink,
                                              (unsigned long long)h << a.sink.hoBits, a.giant, it.plan, h,
                                              a.fuse != 0, sup, a.counts, a.supTab, stats);
        if (lane == 0 && c) atomicAdd(a.counts + 2 * W.plan.rule + K, (unsigned long long)c);
        if (lane == 0 && sup) atomicAdd(a.counts + 2 * W.plan.rule + kSupport, (unsigned long long)sup);
    }
    flush_stats(a.giant, stats);
}

template <int K, int D, int NS, int E>
__global__ void __launch_bounds__(kWave) k_wave_one(Plan P, WaveArgs a, DevGraph G) {
    __shared__ typename WaveLdsOf<E, NS>::T W;
    __shared__ int cache[wave_cache_words<D>()];
    extern __shared__ uint32_t dyn[];
    const int lane = lane_id();
    uint32_t n = *a.heavyCount;
    if (n > a.heavyCap) n = a.heavyCap;
    if (n == 0) return;
    if (a.planOut && blockIdx.x == 0) {   // the giant kernels read the plan from memory
        const int32_t* src = reinterpret_cast<const int32_t*>(&P);
        int32_t* dst = reinterpret_cast<int32_t*>(a.planOut);
        for (int k = lane; k < kPlanWords; k += kWave) dst[k] = src[k];
    }
    copy_plan(W, &P);
    Dedup<D, NS> dd = wave_dedup<D, NS>(W, cache, dyn, a, G);
    unsigned long long stats[kWaveStats] = {};
    for (;;) {
        const uint32_t h = wave_take(a.cursor);
        if (h >= n) break;
        uint64_t sup = 0;
        const uint64_t c = run_seed<K, D, NS>(W, G, a.refs, a.heavy[h].item, dd, a.rowWords, a.sink,
                                              (unsigned long long)h << a.sink.hoBits, a.giant, 0u, h, false, sup,
                                              nullptr, nullptr, stats);
        if (lane == 0 && c) atomicAdd(a.counts, (unsigned long long)c);
    }
    flush_stats(a.giant, stats);
}

// ------------------------------------------------------------ giant seeds

// One wave per listed giant: binds the seed again and writes its sub-items
// -- kSubsPerGiant ranges of its level-0 row, or, when that row is shorter
// than kSubMinRow, up to 4 ranges of the level-1 row below each level-0
// entry (one lane per entry), with that entry's bindings -- and records
// how many it wrote. Capacity is never short: cap * kSubsPerGiant slots.
template <int NS>
__global__ void __launch_bounds__(kWave) k_giant_expand(GiantArgs ga, const Plan* plans, const ConflictRefs* refs,
                                                        DevGraph G) {
    __shared__ ThreadLdsT<levels_of(NS)> W;
    const int lane = lane_id();
    uint32_t n = *ga.count;
    if (n > ga.cap) n = ga.cap;
    for (uint32_t g = blockIdx.x; g < n; g += gridDim.x) {
        const HeavyItem it = ga.list[g];
        const uint32_t hv = ga.heavyIdx[g];
        copy_plan(W, plans + it.plan);
        const Plan& P = W.plan;
        int v[NS];
        uint32_t total = 0;
        if (bind_seed(P, G, (int)it.item, v)) {
            const Frame f0 = enter(W, P, 0, G, v);
            const int len0 = f0.end - f0.i;
            if (len0 >= kSubMinRow || (len0 > 0 && P.nlev == 1)) {
                const int parts = len0 < kSubsPerGiant ? len0 : kSubsPerGiant;
                const int chunk = (len0 + parts - 1) / parts;
                const int items = (len0 + chunk - 1) / chunk;
                uint32_t base = 0;
                if (lane == 0) base = atomicAdd(ga.subCount, (uint32_t)items);
                base = bcast(base);
                for (int c = lane; c < items; c += kWave) {
                    SubItem& S = ga.subs[base + c];
                    S.giant = g;
                    S.plan = it.plan;
                    S.level = 0;
                    S.begin = f0.i + c * chunk;
                    S.end = f0.i + (c + 1) * chunk < f0.end ? f0.i + (c + 1) * chunk : f0.end;
                    S.heavy = hv;
#pragma unroll
                    for (int j = 0; j < kMaxSlots; ++j) S.v[j] = -1;
#pragma unroll
                    for (int j = 0; j < NS; ++j) S.v[j] = v[j];
                }
                total = (uint32_t)items;
            } else if (len0 > 0) {
                // short level-0 row (< kSubMinRow <= 64 entries): one lane per entry, cut its level-1 row
                int mine = 0, chunk = 1;
                Frame f1{0, 0};
                if (lane < len0) {
                    const int i = f0.i + lane;
                    const int x = G.col[i];
                    bool ok = bind_entry_x(P.lev[0], G, refs + it.plan, v, i, x);
                    if (ok && level_kind(P.lev[0].flags) == kIsect) ok = isect_find(W, 0, P.lev[0], G, v, x);
                    if (ok) {
                        f1 = enter(W, P, 1, G, v);
                        const int len1 = f1.end - f1.i;
                        if (len1 > 0) {
                            const int parts = len1 < 4 ? len1 : 4;
                            chunk = (len1 + parts - 1) / parts;
                            mine = (len1 + chunk - 1) / chunk;
                        }
                    }
                }
                // exclusive prefix of the lanes' item counts
                int incl = mine;
#pragma unroll
                for (int off = 1; off < kWave; off <<= 1) {
                    const int y = __shfl_up(incl, off, kWave);
                    if (lane >= off) incl += y;
                }
                const int all = __shfl(incl, kWave - 1, kWave);
                uint32_t base = 0;
                if (lane == 0 && all > 0) base = atomicAdd(ga.subCount, (uint32_t)all);
                base = bcast(base) + (uint32_t)(incl - mine);
                for (int c = 0; c < mine; ++c) {
                    SubItem& S = ga.subs[base + c];
                    S.giant = g;
                    S.plan = it.plan;
                    S.level = 1;
                    S.begin = f1.i + c * chunk;
                    S.end = f1.i + (c + 1) * chunk < f1.end ? f1.i + (c + 1) * chunk : f1.end;
                    S.heavy = hv;
#pragma unroll
                    for (int j = 0; j < kMaxSlots; ++j) S.v[j] = -1;
#pragma unroll
                    for (int j = 0; j < NS; ++j) S.v[j] = v[j];
                }
                total = (uint32_t)all;
            }
        }
        if (lane == 0) ga.remaining[g] = total;
        __builtin_amdgcn_wave_barrier();
    }
}

// The sub-items of all giants, over all waves: each one is a pool (its range
// and bindings) run by wave_run against the giant's shared row or found
// flag. The sub-item that finishes a giant's last clears its row, after
// counting the seed's supported head triples from it (fused support).
template <int K, int D, int NS, int E>
__global__ void __launch_bounds__(kWave) k_wave_sub(WaveArgs a, DevGraph G) {
    __shared__ typename WaveLdsOf<E, NS>::T W;
    __shared__ int cache[wave_cache_words<D>()];
    extern __shared__ uint32_t dyn[];
    const int lane = lane_id();
    const GiantArgs& ga = a.giant;
    uint32_t n = *ga.subCount;
    const uint32_t capSubs = ga.cap * (uint32_t)kSubsPerGiant;
    if (n > capSubs) n = capSubs;
    if (n == 0) return;   // no giants listed
    Dedup<D, NS> dd = wave_dedup<D, NS>(W, cache, dyn, a, G);
    uint32_t loaded = 0xffffffffu;
    unsigned long long stats[kWaveStats] = {};
    for (;;) {
        const unsigned long long tsub = GR_PROF_NOW();
        const uint32_t si = wave_take(ga.subCursor);
        if (si >= n) break;
        const SubItem& S = ga.subs[si];
        const uint32_t g = (uint32_t)uniform((int)S.giant);
        const uint32_t plan = (uint32_t)uniform((int)S.plan);
        const uint32_t hv = (uint32_t)uniform((int)S.heavy);
        if (plan != loaded) {
            copy_plan(W, a.plans + plan);
            loaded = plan;
        }
        // existence searches (SUPPORT, and PAIRS plans whose head object is bound by the seed) count a giant once
        uint32_t* foundFlag = (K == kSupport || W.plan.pairLevel < 0) ? ga.found + g : nullptr;
        uint32_t skip = 0;
        if (foundFlag && lane == 0) skip = ld_rlx(foundFlag);
        uint64_t c = 0;
        if (!bcast(skip)) {
            int start = 0;
            if constexpr (E == 1) {
                start = ls_place_sub<K, D, NS>(W, dd, S, G);
            } else {
                if (lane < kMaxSlots) W.poolV[lane] = S.v[lane];
                if (lane == 0) {
                    W.poolLevel = S.level;
                    W.poolNext = S.begin;
                    W.poolEnd = S.end;
                    W.poolGen = W.poolGen + 1u;
                    W.nTouched = 0;
                }
            }
            if (D == kDedupSort && K == kPairs)
                for (int k = lane; k < kCacheSlots; k += kWave) cache[k] = -1;
            wave_sync();
            uint32_t* const row = ga.rows + (int64_t)g * a.rowWords;
            dd.shared = (K == kPairs && D == kDedupRows) ? row : nullptr;   // marks go straight to the row
            dd.merge = (K == kPairs && D == kDedupLds) ? row : nullptr;     // marks go to LDS, merged at the end
            bool aborted = false;
            uint64_t none = 0;
            if constexpr (E == 1)
                c = ls_run<K, D, NS>(W, G, a.refs + plan, dd, a.rowWords, a.sink,
                                     (unsigned long long)hv << a.sink.hoBits, ga, false, HeavyItem{plan, 0u}, hv,
                                     foundFlag, aborted, false, Frame{0, 0}, nullptr, 0, nullptr, none, stats, start);
            else
                c = wave_run<K, D, NS>(W, G, a.refs + plan, dd, a.rowWords, a.sink,
                                       (unsigned long long)hv << a.sink.hoBits, ga, false, HeavyItem{plan, 0u}, hv,
                                       foundFlag, aborted, false, Frame{0, 0}, nullptr, 0, nullptr, none, stats);
            dd.shared = nullptr;
            dd.merge = nullptr;
        }
        if (lane == 0 && c) atomicAdd(a.counts + 2 * W.plan.rule + K, (unsigned long long)c);
        // The giant's row is complete once its last sub-item is done: that
        // sub-item counts the seed's supported head triples from it (fused
        // support) and, with LDS bitmaps, the seed's distinct head objects (the
        // set bits; the sub-items merged their bitmaps without counting).
        const bool fs = K == kPairs && D != kDedupSort && a.fuse && W.plan.supPred != -1;
        const bool rowCount = K == kPairs && D == kDedupLds;
        if (fs || rowCount) __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
        uint32_t last = 0;
        if (lane == 0) last = atomicSub(ga.remaining + g, 1u) == 1u ? 1u : 0u;
        if (bcast(last) && K == kPairs && D != kDedupSort) {
            uint32_t* row = ga.rows + (int64_t)g * a.rowWords;
            if (fs || rowCount) __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
            if (rowCount) {
                uint64_t nb = 0;
                for (int64_t w = lane; w < a.rowWords; w += kWave)
                    nb += (uint64_t)__popc(__hip_atomic_load(row + w, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));
                nb = wave_sum(nb);
                if (lane == 0 && nb) atomicAdd(a.counts + 2 * W.plan.rule + K, (unsigned long long)nb);
            }
            if (fs) {
                const int supPred = W.plan.supPred;
                const int32_t* supList = supPred <= -2 ? a.supTab + (-supPred - 2) : nullptr;
                const int s0 = uniform(S.v[0]);
                const int nm = supList ? supList[0] : 1;
                for (int m = 0; m < nm; ++m) {
                    const int pred = supList ? supList[2 + 2 * m] : supPred;
                    const int rule = supList ? supList[1 + 2 * m] : W.plan.rule;
                    const Frame hr = key_row(G, (int64_t)pred * G.nBlocks, s0);
                    const uint64_t sup = head_hits_wave(dd, G, hr, row);
                    if (lane == 0 && sup) atomicAdd(a.counts + 2 * rule + kSupport, (unsigned long long)sup);
                }
            }
            for (int64_t w = lane; w < a.rowWords; w += kWave) row[w] = 0u;   // read again only by later launches
        }
        GR_PROF_ADD(7, tsub);
    }
    flush_stats(ga, stats);
}

// ---------------------------------------------------------------- launchers

// One instantiation per plan kind, so a profile names the phase and the
// SUPPORT kernels carry no distinct-set code. Wave-mode launches run one wave
// per block; kDedupLds blocks carry their bitmap (rowWords words) in dynamic
// LDS.
template <int NS>
void SlotKernels<NS>::thread_batch(int kind, const ThreadArgs& a, const DevGraph& G, int blocks, hipStream_t s) {
    if (kind == kSupport) hipLaunchKernelGGL((k_thread_batch<kSupport, NS>), dim3(blocks), dim3(kBlock), 0, s, a, G);
    else hipLaunchKernelGGL((k_thread_batch<kPairs, NS>), dim3(blocks), dim3(kBlock), 0, s, a, G);
}

template <int NS>
void SlotKernels<NS>::thread_one(int kind, const Plan& P, const OneArgs& a, const DevGraph& G, int blocks,
                                 hipStream_t s) {
    if (kind == kSupport) hipLaunchKernelGGL((k_thread_one<kSupport, NS>), dim3(blocks), dim3(kBlock), 0, s, P, a, G);
    else hipLaunchKernelGGL((k_thread_one<kPairs, NS>), dim3(blocks), dim3(kBlock), 0, s, P, a, G);
}

template <int NS, int E>
static void wave_batch_e(int kind, int dedup, const WaveArgs& a, const DevGraph& G, int blocks, hipStream_t s) {
    if (kind == kSupport)
        hipLaunchKernelGGL((k_wave_batch<kSupport, kDedupRows, NS, E>), dim3(blocks), dim3(kWave), 0, s, a, G);
    else if (dedup == kDedupSort)
        hipLaunchKernelGGL((k_wave_batch<kPairs, kDedupSort, NS, E>), dim3(blocks), dim3(kWave), 0, s, a, G);
    else if (dedup == kDedupLds)
        hipLaunchKernelGGL((k_wave_batch<kPairs, kDedupLds, NS, E>), dim3(blocks), dim3(kWave), (size_t)a.rowWords * 4,
                           s, a, G);
    else
        hipLaunchKernelGGL((k_wave_batch<kPairs, kDedupRows, NS, E>), dim3(blocks), dim3(kWave), 0, s, a, G);
}

template <int NS>
void SlotKernels<NS>::wave_batch(int kind, int dedup, const WaveArgs& a, const DevGraph& G, int blocks,
                                 hipStream_t s) {
    if (a.engine == 1) wave_batch_e<NS, 1>(kind, dedup, a, G, blocks, s);
    else wave_batch_e<NS, 0>(kind, dedup, a, G, blocks, s);
}

template <int NS, int E>
static void wave_one_e(int kind, int dedup, const Plan& P, const WaveArgs& a, const DevGraph& G, int blocks,
                       hipStream_t s) {
    if (kind == kSupport)
        hipLaunchKernelGGL((k_wave_one<kSupport, kDedupRows, NS, E>), dim3(blocks), dim3(kWave), 0, s, P, a, G);
    else if (dedup == kDedupSort)
        hipLaunchKernelGGL((k_wave_one<kPairs, kDedupSort, NS, E>), dim3(blocks), dim3(kWave), 0, s, P, a, G);
    else if (dedup == kDedupLds)
        hipLaunchKernelGGL((k_wave_one<kPairs, kDedupLds, NS, E>), dim3(blocks), dim3(kWave), (size_t)a.rowWords * 4, s,
                           P, a, G);
    else
        hipLaunchKernelGGL((k_wave_one<kPairs, kDedupRows, NS, E>), dim3(blocks), dim3(kWave), 0, s, P, a, G);
}

template <int NS>
void SlotKernels<NS>::wave_one(int kind, int dedup, const Plan& P, const WaveArgs& a, const DevGraph& G, int blocks,
                               hipStream_t s) {
    if (a.engine == 1) wave_one_e<NS, 1>(kind, dedup, P, a, G, blocks, s);
    else wave_one_e<NS, 0>(kind, dedup, P, a, G, blocks, s);
}

template <int NS, int E>
static void wave_sub_e(int kind, int dedup, const WaveArgs& a, const DevGraph& G, int blocks, hipStream_t s) {
    if (kind == kSupport)
        hipLaunchKernelGGL((k_wave_sub<kSupport, kDedupRows, NS, E>), dim3(blocks), dim3(kWave), 0, s, a, G);
    else if (dedup == kDedupSort)
        hipLaunchKernelGGL((k_wave_sub<kPairs, kDedupSort, NS, E>), dim3(blocks), dim3(kWave), 0, s, a, G);
    else if (dedup == kDedupLds)
        hipLaunchKernelGGL((k_wave_sub<kPairs, kDedupLds, NS, E>), dim3(blocks), dim3(kWave), (size_t)a.rowWords * 4, s,
                           a, G);
    else
        hipLaunchKernelGGL((k_wave_sub<kPairs, kDedupRows, NS, E>), dim3(blocks), dim3(kWave), 0, s, a, G);
}

template <int NS>
void SlotKernels<NS>::giants(int kind, int dedup, const WaveArgs& a, const DevGraph& G, int blocks, hipStream_t s) {
    const int eb = blocks < 1024 ? blocks : 1024;
    hipLaunchKernelGGL(k_giant_expand<NS>, dim3(eb), dim3(kWave), 0, s, a.giant, a.plans, a.refs, G);
    if (a.engine == 1) wave_sub_e<NS, 1>(kind, dedup, a, G, blocks, s);
    else wave_sub_e<NS, 0>(kind, dedup, a, G, blocks, s);
}

template <int NS>
size_t SlotKernels<NS>::lds_static(int engine) {
    hipFuncAttributes fa{};
    const void* k = engine == 1 ? reinterpret_cast<const void*>(&k_wave_batch<kPairs, kDedupLds, NS, 1>)
                                : reinterpret_cast<const void*>(&k_wave_batch<kPairs, kDedupLds, NS, 0>);
    if (hipFuncGetAttributes(&fa, k) != hipSuccess)
        return (engine == 1 ? sizeof(LsLdsT<NS>) : sizeof(WaveLdsT<levels_of(NS)>)) + kCacheSlots * sizeof(int);
    return fa.sharedSizeBytes;
}

template <int NS>
int SlotKernels<NS>::lds_blocks_per_cu(size_t dynLds, int engine) {
    int n = 0;
    const void* k = engine == 1 ? reinterpret_cast<const void*>(&k_wave_batch<kPairs, kDedupLds, NS, 1>)
                                : reinterpret_cast<const void*>(&k_wave_batch<kPairs, kDedupLds, NS, 0>);
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, k, kWave, dynLds) != hipSuccess) return 0;
    return n;
}

template <int NS, int E>
static void prepare_e(size_t dynLds) {
    (void)hipFuncSetAttribute(reinterpret_cast<const void*>(&k_wave_batch<kPairs, kDedupLds, NS, E>),
                              hipFuncAttributeMaxDynamicSharedMemorySize, (int)dynLds);
    (void)hipFuncSetAttribute(reinterpret_cast<const void*>(&k_wave_one<kPairs, kDedupLds, NS, E>),
                              hipFuncAttributeMaxDynamicSharedMemorySize, (int)dynLds);
    (void)hipFuncSetAttribute(reinterpret_cast<const void*>(&k_wave_sub<kPairs, kDedupLds, NS, E>),
                              hipFuncAttributeMaxDynamicSharedMemorySize, (int)dynLds);
}

template <int NS>
void SlotKernels<NS>::prepare(size_t dynLds) {
    prepare_e<NS, 0>(dynLds);
    prepare_e<NS, 1>(dynLds);
}

}  // namespace gr
