// plan.hpp -- plain-old-data shared by the host planner and the HIP kernels.
//
// A rule is compiled into two Plans: a SUPPORT plan (head variables bound by
// each seed head triple, existence search over the body) and a PAIRS plan
// (head subject bound by each seed key, all head-object bindings counted
// once). Each plan is a list of Levels, one per body atom, in execution
// order. Variables live in "slots" numbered in binding order, so a level
// reads only slots bound before it and binds at most two new ones.
//
// Term references: r >= 0 is slot r; r < 0 is the constant term id -(r+1).
#pragma once
#include <cstdint>

#ifndef __HIPCC__
#ifndef __host__
#define __host__
#define __device__
#endif
#endif

namespace gr {

constexpr int kMaxBody  = 8;    // body atoms per rule (GPU_03/gpu_common.cuh:25-27)
constexpr int kMaxSlots = 16;   // distinct variables per rule (same source)
constexpr int kWave     = 64;   // CDNA wavefront width
constexpr int kChunkSeeds = 1024;   // seeds per thread-mode work unit, at most
constexpr int kMinChunkSeeds = 64;  // ... and at least (small batches use smaller units)

enum LevelKind : int32_t {
    kCheck = 0,   // both terms bound: binary search of one row
    kExt   = 1,   // key bound, other term a new variable: enumerate one row
    kScan  = 2,   // neither bound: enumerate all triples of the predicate
    kIsect = 3,   // kExt whose new variable must also lie in a second row (the CHECK level after it):
                  // enumerate row 1, walk row 2 with a cursor; the CHECK's frame is the match
};

enum Dir : int32_t { kSPO = 0, kPOS = 1 };

enum PlanKind : int32_t { kSupport = 0, kPairs = 1 };

// Level descriptor: 32 bytes, so a plan's levels stay in scalar registers.
// flags: bits 0-1 kind, bit 2 dir, bits 3-7 nbound (variable slots bound
// before this level), bits 8-16 conflict mask (bit 8+j: the ground triple
// may repeat level j's; bit 16: may repeat the head triple).
// CHECK/EXT/ISECT: key-index block of term id k = blockBase + (k >> 6), with
// blockBase = lo | hi << 32 (the (direction, predicate) block run).
// SCAN: entries [lo, hi) of the merged col/coo arrays.
// ISECT: the second row is that of term val2 in the block run lo2 | hi2 << 32.
struct alignas(16) Level {
    int32_t flags;
    int32_t key;   // term ref of the key position (CHECK, EXT, ISECT)
    int32_t val;   // CHECK: term ref searched in the row; EXT/ISECT: slot bound; SCAN: slot of the key position
    int32_t val2;  // SCAN: slot of the other position (== val for (?x p ?x)); ISECT: key term ref of row 2
    int32_t lo;
    int32_t hi;
    int32_t lo2, hi2;   // ISECT: block run of row 2
};

__host__ __device__ constexpr int32_t level_kind(int32_t f) { return f & 3; }
__host__ __device__ constexpr int32_t level_dir(int32_t f) { return (f >> 2) & 1; }
__host__ __device__ constexpr int32_t level_nbound(int32_t f) { return (f >> 3) & 31; }
__host__ __device__ constexpr int32_t level_conflicts(int32_t f) { return (f >> 8) & 511; }
__host__ __device__ constexpr int32_t make_level_flags(int32_t kind, int32_t dir, int32_t nbound, int32_t conf) {
    return kind | (dir << 2) | (nbound << 3) | (conf << 8);
}

struct alignas(16) Plan {
    int32_t nlev;        // number of levels (= body atoms)
    int32_t kind;        // PlanKind
    int32_t rule;        // rule index within its batch (result slot)
    int32_t seedDir;     // SUPPORT: orientation of the seed entries (kSPO: key = subject)
    int32_t seedBegin;   // SUPPORT: entries [seedBegin, seedEnd) of the merged col/coo arrays;
    int32_t seedEnd;     // PAIRS: entries of the merged keys array
    int32_t hsRef;       // SUPPORT: head subject term ref; PAIRS: slot of the head subject (0)
    int32_t hoRef;       // SUPPORT: head object term ref; PAIRS: slot of the head object
    int32_t pairLevel;   // PAIRS: level binding the head object; -1 if bound by the seed
    int32_t budget;      // thread-mode candidates per seed before escalation
    // PAIRS with pairLevel >= 0: the head objects are keys of one (direction,
    // predicate) -- the values the pair level enumerates -- so wave-mode
    // bitmaps index them by rank among those keys: key-index block run
    // hoLo | hoHi << 32, global row number hoRow0 of its first key, hoKeys keys.
    int32_t hoLo, hoHi, hoRow0, hoKeys;
    int32_t hoByEntry;   // 1: the pair level binds the head object as col[e], whose rank is crank[e]
    // PAIRS, 1: level 0 is an EXT keyed by the seed over the seeds' own
    // (direction, predicate), so seed i's level-0 row is row seedBegin + i
    // (rows are numbered in the order of the keys array), read from rowStart
    // without the key index
    int32_t seedRow0;
    // PAIRS, supPred >= 0: the plan also counts the rule's support (when the
    // launch's `fuse` is set): a head triple (s, p_head, o) is supported iff
    // (s, o) is a distinct body pair, so each seed s adds the head objects it
    // found that are objects of s in the SPO rows of predicate supPred (the
    // head predicate; its key-index block run starts at supPred * nBlocks).
    // Only rules where no body atom can ground to the head triple itself
    // qualify (the support search excludes that triple, the pair search does
    // not). -1: no support counting here. <= -2: the plan's search is shared
    // by rules with equal bodies, and it counts the support of several of
    // them: supTab[-supPred - 2] = n, then n (rule, head predicate) pairs.
    int32_t supPred;
    Level lev[kMaxBody];
};

// Subject/object term refs of each level and of the head, read only when a
// level's conflict mask is non-zero (rules with constants).
struct alignas(16) ConflictRefs {
    int32_t sref[kMaxBody];
    int32_t oref[kMaxBody];
    int32_t hs, ho, pad0, pad1;
};

// The slot class a plan needs: its variable slots (one past the largest
// slot any term ref of the plan or its conflict refs names; constants are
// negative refs) and its levels, since a class keeps frames for as many
// levels as it has slots (at most kMaxBody).
inline int plan_slots(const Plan& P, const ConflictRefs& R) {
    int n = P.nlev > 1 ? P.nlev : 1;
    auto see = [&](int32_t r) { if (r >= 0 && r + 1 > n) n = r + 1; };
    see(P.hsRef);
    see(P.hoRef);
    for (int L = 0; L < P.nlev && L < kMaxBody; ++L) {
        see(P.lev[L].key);
        see(P.lev[L].val);
        see(P.lev[L].val2);
    }
    for (int j = 0; j < kMaxBody; ++j) {
        see(R.sref[j]);
        see(R.oref[j]);
    }
    see(R.hs);
    see(R.ho);
    return n;
}

// The slot class whose kernels run a launch of plans needing `slots` (plan_slots).
constexpr int slot_class(int slots) { return slots <= 4 ? 4 : (slots <= 8 ? 8 : 16); }

// Chunk::gate of a PAIRS chunk: bits 0-30 a support threshold (0: none), the
// chunk is skipped when the rule's support counter is below it; with
// kGateFused set, only when the launch does not fuse support (the pairs pass
// itself then counts that support).
constexpr uint32_t kGateFused = 0x80000000u;

// One unit of thread-mode work: up to kChunkSeeds consecutive seeds of one plan.
struct alignas(16) Chunk {
    uint32_t plan;       // index into the plan table
    uint32_t begin;      // first seed (relative to the plan's seed range)
    uint32_t count;      // seeds in this chunk (<= kChunkSeeds)
    uint32_t gate;       // PAIRS chunks: support threshold | kGateFused, or 0
};

// A seed escalated from thread mode to wave mode.
struct alignas(8) HeavyItem {
    uint32_t plan;
    uint32_t item;
};

// Key index of one (direction, predicate): one 16-byte block per 64
// consecutive term ids -- `bits` marks the ids that have a row, `rank` is
// the global row number of the block's first marked id. Row r spans entries
// [rowStart[r], rowStart[r+1]) of the merged col/coo arrays. Rows are
// numbered in (direction, predicate, key) order, the order of `keys`, so
// rowStart also gives every row's key list position.
struct alignas(16) KeyBlock {
    uint32_t rank;
    uint32_t pad;
    uint64_t bits;
};

// Device-resident graph. Direction d's entries of predicate p occupy
// [d*T + predOff[p], d*T + predOff[p+1]) of col/coo; SPO entries are sorted
// by (s, o) within a predicate (coo = s, col = o), POS by (o, s) (coo = o,
// col = s). keys holds the distinct subjects then the distinct objects of
// every predicate (row r's key is keys[r]).
struct DevGraph {
    const KeyBlock* blocks;   // [2][P][ceil(nE/64)]
    const int32_t* rowStart;  // [rows + 1]
    const int32_t* col;       // [2T]
    const int32_t* coo;       // [2T]
    const int32_t* keys;      // [rows]
    // [2T] for entry e of (direction d, predicate p): the rank of col[e] among
    // the keys of (1-d, p) -- the bitmap index of a head object bound by e
    const int32_t* crank;
    int64_t nBlocks;          // key-index blocks per (direction, predicate): ceil(nE/64)
};

}  // namespace gr
