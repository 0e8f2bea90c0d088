// sort_unique.hpp -- radix sort + unique count of the body-pair keys that the
// sort-dedup wave kernels emit (the counterpart of the reference's overflow
// path, Thrust sort + unique over (headS, headO) pairs,
// GPU_03/gpu_rules_lib.cu:42-51,508-511).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

#include "plan.hpp"

namespace gr {

constexpr int kSortBlock = 256;                      // threads per sort block
constexpr int kSortItems = 16;                       // keys per thread per tile
constexpr int kSortTile = kSortBlock * kSortItems;   // keys per tile
constexpr int kRadixBits = 8;
constexpr int kRadix = 1 << kRadixBits;

// Histogram words a buffer of `cap` keys needs: kRadix counters per tile,
// then kRadix digit totals.
inline uint32_t sort_max_tiles(uint32_t cap) { return (cap + kSortTile - 1) / kSortTile; }
inline uint64_t sort_hist_words(uint32_t cap) {
    return (uint64_t)kRadix * (sort_max_tiles(cap) + 1);
}

// One sort-unique job: the keys are in keys[0][0, min(*count, cap)) (the
// count is 64-bit and may pass cap; cap <= 2^31, so positions fit 32 bits), each
// key (h << hoBits) | object for heavy seed h. `bits` significant key bits
// decide the number of 8-bit LSD passes; the sorted keys end in
// keys[passes % 2]. The unique pass adds the number of distinct keys of each
// seed to counts[2 * rule + 1] with rule = plans[heavy[h].plan].rule, or,
// when heavy is null, all distinct keys to counts[0].
struct SortJob {
    unsigned long long* keys[2];
    const unsigned long long* count;
    uint32_t cap;
    uint32_t maxTiles;
    uint32_t* hist;
    int32_t bits;
    int32_t hoBits;
    const HeavyItem* heavy;
    const Plan* plans;
    unsigned long long* counts;
    int blocks;   // persistent grid of the tile kernels
};

inline int sort_passes(int bits) { return (bits + kRadixBits - 1) / kRadixBits; }

// Number of bits that hold values 0..n-1 (at least 1).
inline int bits_for(uint64_t n) {
    int b = 1;
    while (b < 63 && (1ull << b) < n) ++b;
    return b;
}

// Enqueues all LSD passes and the unique count on stream s. Every kernel
// reads the key count from device memory, so the sequence can be captured
// in a HIP graph before the count is known.
void launch_sort_unique(const SortJob& j, hipStream_t s);

// Just the LSD passes: sorts the first min(*count, cap) keys of keys[0] by
// their low `bits` bits (stable), ping-ponging through keys[1]; hist holds
// sort_hist_words(cap) words. Returns the index of the buffer holding the
// sorted keys (passes % 2).
int launch_radix_sort(unsigned long long* const keys[2], const unsigned long long* count, uint32_t cap, uint32_t* hist,
                      int bits, int blocks, hipStream_t s);

}  // namespace gr
