// search_s16.hip -- the counting kernels for plans with at most 16 variable slots.
#include "search.hpp"

namespace gr {
template struct SlotKernels<16>;
}  // namespace gr
