// noalloc.hpp -- device scratch of the allocation-free per-rule entry point.
#pragma once
#include <cstddef>
#include <cstdint>

namespace gr {

// Bytes of the per-device scratch array: 512 bytes of counters, control
// words and conflict tables, then the heavy-seed queue (8 bytes per seed).
constexpr size_t kNoAllocBytes = (size_t)16 << 20;
constexpr size_t kNoAllocHeader = 512;
// Seeds one launch of the allocation-free path may escalate at most: the
// queue holds every seed of the launch, so this bounds a rule's support
// seeds (head triples) and its body-pair seeds (head-subject keys).
constexpr uint64_t kNoAllocMaxSeeds = (kNoAllocBytes - kNoAllocHeader) / 8;

// Takes the scratch array's address on `device` (which must be current)
// once and keeps it; false (logged) if that fails.
bool noalloc_init(int device);
// The scratch array's device address on `device`, or null before a
// successful noalloc_init.
void* noalloc_scratch(int device);

}  // namespace gr
