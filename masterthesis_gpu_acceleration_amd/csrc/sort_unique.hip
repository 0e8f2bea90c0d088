// sort_unique.hip -- LSD radix sort and unique count of body-pair keys on gfx950.
//
// The keys come from the sort-dedup wave kernels: key = (h << hoBits) | o
// for head object o of heavy seed h. Sorting them groups each seed's
// objects; a key that differs from its predecessor is one distinct
// (head subject, head object) pair of that seed's rule.
//
// One LSD pass over 8 key bits is three kernels:
//   k_radix_hist    per tile of kSortTile keys, a 256-bin digit histogram
//                   built with LDS atomics, stored digit-major
//                   (hist[d * maxTiles + t]) so that
//   k_radix_scan    one block per digit can turn its row into the exclusive
//                   prefix over tiles and write the digit's total, and
//   k_radix_scatter rank each key stably inside its tile -- rounds of 256
//                   keys in index order; within a wave, the lanes holding the
//                   same digit are found with 8 ballots (one per digit bit),
//                   across the block's four waves by per-wave digit counts in
//                   LDS -- and write it to digitStart[d] + tileOffset[d][t] +
//                   its rank.
// Stability of every pass makes the whole sort correct. All kernels read the
// key count from device memory and walk tiles with a grid stride, so the
// pass sequence has a fixed launch shape and can be captured in a HIP graph.
#include <hip/hip_runtime.h>

#include <cstdint>

#include "sort_unique.hpp"

namespace gr {

namespace {

constexpr int kSortWaves = kSortBlock / kWave;

struct RadixPass {
    const unsigned long long* in;
    unsigned long long* out;
    const unsigned long long* count;
    uint32_t cap;
    uint32_t maxTiles;
    uint32_t* hist;
    int32_t shift;
};

struct UniqueArgs {
    const unsigned long long* keys;
    const unsigned long long* count;
    uint32_t cap;
    int32_t hoBits;
    const HeavyItem* heavy;
    const Plan* plans;
    unsigned long long* counts;
};

__device__ __forceinline__ uint32_t key_count(const unsigned long long* count, uint32_t cap) {
    const unsigned long long n = *count;
    return n < (unsigned long long)cap ? (uint32_t)n : cap;
}

__device__ __forceinline__ uint32_t wave_incl_scan(uint32_t x) {
    const int lane = __lane_id();
#pragma unroll
    for (int off = 1; off < kWave; off <<= 1) {
        const uint32_t y = __shfl_up(x, off, kWave);
        if (lane >= off) x += y;
    }
    return x;
}

__global__ void __launch_bounds__(kSortBlock) k_radix_hist(RadixPass a) {
    __shared__ uint32_t h[kRadix];
    const uint32_t n = key_count(a.count, a.cap);
    const uint32_t tiles = (n + kSortTile - 1) / kSortTile;
    const int tid = threadIdx.x;
    for (uint32_t t = blockIdx.x; t < tiles; t += gridDim.x) {
        h[tid] = 0;
        __syncthreads();
        const uint32_t base = t * (uint32_t)kSortTile;
#pragma unroll 4
        for (int r = 0; r < kSortItems; ++r) {
            const uint32_t i = base + (uint32_t)(r * kSortBlock + tid);
            if (i < n) atomicAdd(&h[(uint32_t)(a.in[i] >> a.shift) & (kRadix - 1)], 1u);
        }
        __syncthreads();
        a.hist[(size_t)tid * a.maxTiles + t] = h[tid];
        __syncthreads();
    }
}

// One block per digit: exclusive prefix over the digit's tile counts, and
// the digit total after the last tile.
__global__ void __launch_bounds__(kSortBlock) k_radix_scan(RadixPass a) {
    __shared__ uint32_t ws[kSortWaves];
    const uint32_t n = key_count(a.count, a.cap);
    const uint32_t tiles = (n + kSortTile - 1) / kSortTile;
    const int tid = threadIdx.x, lane = __lane_id(), wv = tid / kWave;
    uint32_t* row = a.hist + (size_t)blockIdx.x * a.maxTiles;
    uint32_t carry = 0;
    for (uint32_t b = 0; b < tiles; b += kSortBlock) {
        const uint32_t i = b + (uint32_t)tid;
        const uint32_t x = i < tiles ? row[i] : 0u;
        const uint32_t incl = wave_incl_scan(x);
        if (lane == kWave - 1) ws[wv] = incl;
        __syncthreads();
        uint32_t off = 0, total = 0;
#pragma unroll
        for (int w = 0; w < kSortWaves; ++w) {
            off += w < wv ? ws[w] : 0u;
            total += ws[w];
        }
        if (i < tiles) row[i] = carry + off + incl - x;
        carry += total;
        __syncthreads();
    }
    if (tid == 0) a.hist[(size_t)kRadix * a.maxTiles + blockIdx.x] = carry;
}

__global__ void __launch_bounds__(kSortBlock) k_radix_scatter(RadixPass a) {
    __shared__ uint32_t base[kRadix];               // next output position of each digit in this tile
    __shared__ uint32_t cnt[kSortWaves][kRadix];    // keys of each digit per wave in the current round
    __shared__ uint32_t ws[kSortWaves];
    const uint32_t n = key_count(a.count, a.cap);
    const uint32_t tiles = (n + kSortTile - 1) / kSortTile;
    const int tid = threadIdx.x, lane = __lane_id(), wv = tid / kWave;
    const uint64_t below = (1ull << lane) - 1ull;
    // start of each digit in the output: exclusive prefix of the digit totals
    const uint32_t tot = a.hist[(size_t)kRadix * a.maxTiles + tid];
    const uint32_t incl = wave_incl_scan(tot);
    if (lane == kWave - 1) ws[wv] = incl;
#pragma unroll
    for (int w = 0; w < kSortWaves; ++w) cnt[w][tid] = 0;
    __syncthreads();
    uint32_t digitStart = incl - tot;
#pragma unroll
    for (int w = 0; w < kSortWaves; ++w) digitStart += w < wv ? ws[w] : 0u;
    for (uint32_t t = blockIdx.x; t < tiles; t += gridDim.x) {
        base[tid] = digitStart + a.hist[(size_t)tid * a.maxTiles + t];
        __syncthreads();
        for (int r = 0; r < kSortItems; ++r) {
            const uint32_t i = t * (uint32_t)kSortTile + (uint32_t)(r * kSortBlock + tid);
            const bool valid = i < n;
            const unsigned long long k = valid ? a.in[i] : 0ull;
            const uint32_t d = valid ? (uint32_t)(k >> a.shift) & (kRadix - 1) : 0u;
            uint64_t peers = __ballot(valid);
#pragma unroll
            for (int b = 0; b < kRadixBits; ++b) {
                const bool bit = (d >> b) & 1u;
                const uint64_t m = __ballot(bit);
                peers &= bit ? m : ~m;
            }
            if (valid && (peers & below) == 0) cnt[wv][d] = (uint32_t)__popcll(peers);
            __syncthreads();
            if (valid) {
                uint32_t pos = base[d] + (uint32_t)__popcll(peers & below);
#pragma unroll
                for (int w = 0; w < kSortWaves; ++w) pos += w < wv ? cnt[w][d] : 0u;
                a.out[pos] = k;
            }
            __syncthreads();
            uint32_t add = 0;
#pragma unroll
            for (int w = 0; w < kSortWaves; ++w) {
                add += cnt[w][tid];
                cnt[w][tid] = 0;
            }
            base[tid] += add;
            __syncthreads();
        }
    }
}

// Distinct keys per seed's rule: a key unequal to its predecessor starts a
// new (seed, object) pair. Counter updates are aggregated per wave (the keys
// are sorted, so a wave's new pairs mostly share one rule).
__global__ void __launch_bounds__(kSortBlock) k_unique_count(UniqueArgs a) {
    const uint32_t n = key_count(a.count, a.cap);
    const int lane = __lane_id();
    const uint32_t wave = (blockIdx.x * blockDim.x + threadIdx.x) / kWave;
    const uint32_t waves = gridDim.x * blockDim.x / kWave;
    for (uint64_t b = (uint64_t)wave * kWave; b < n; b += (uint64_t)waves * kWave) {
        const uint64_t i = b + (uint64_t)lane;
        int rule = -1;
        if (i < n) {
            const unsigned long long k = a.keys[i];
            if (i == 0 || a.keys[i - 1] != k) {
                if (a.heavy) {
                    const uint32_t h = (uint32_t)(k >> a.hoBits);
                    rule = a.plans[a.heavy[h].plan].rule;
                } else {
                    rule = 0;
                }
            }
        }
        for (;;) {
            const uint64_t m = __ballot(rule >= 0);
            if (m == 0) break;
            const int leader = (int)__ffsll((long long)m) - 1;
            const int rr = __shfl(rule, leader, kWave);
            const uint64_t same = __ballot(rule == rr);
            if (lane == leader)
                atomicAdd(a.counts + (a.heavy ? 2 * (size_t)rr + kPairs : 0), (unsigned long long)__popcll(same));
            if (rule == rr) rule = -1;
        }
    }
}

}  // namespace

int launch_radix_sort(unsigned long long* const keys[2], const unsigned long long* count, uint32_t cap, uint32_t* hist,
                      int bits, int blocks, hipStream_t s) {
    const int passes = sort_passes(bits);
    if (blocks < 1) blocks = 1;
    int cur = 0;
    for (int p = 0; p < passes; ++p) {
        RadixPass a{keys[cur], keys[cur ^ 1], count, cap, sort_max_tiles(cap), hist, p * kRadixBits};
        hipLaunchKernelGGL(k_radix_hist, dim3(blocks), dim3(kSortBlock), 0, s, a);
        hipLaunchKernelGGL(k_radix_scan, dim3(kRadix), dim3(kSortBlock), 0, s, a);
        hipLaunchKernelGGL(k_radix_scatter, dim3(blocks), dim3(kSortBlock), 0, s, a);
        cur ^= 1;
    }
    return cur;
}

void launch_sort_unique(const SortJob& j, hipStream_t s) {
    const int cur = launch_radix_sort(j.keys, j.count, j.cap, j.hist, j.bits, j.blocks, s);
    UniqueArgs u{j.keys[cur], j.count, j.cap, j.hoBits, j.heavy, j.plans, j.counts};
    hipLaunchKernelGGL(k_unique_count, dim3(j.blocks > 0 ? j.blocks : 1), dim3(kSortBlock), 0, s, u);
}

}  // namespace gr
