// refine.hpp -- one refinement level of a rule miner (refine.cpp).
#pragma once
#include <cstdint>

namespace gr {

// Canonical keys of every refinement of n parent key records (2 + 3 *
// kMaxBody int32 each) over nPreds body predicates, at most maxAtoms atoms,
// sorted and without repeats. Writes up to cap records to out and returns
// how many there are (a larger count than cap: call again with room), or
// -1 on a bad argument.
int64_t refine_rules(const int32_t* parents, int32_t n, const int32_t* preds, int32_t nPreds, int32_t maxAtoms,
                     int32_t* out, int64_t cap);

}  // namespace gr
