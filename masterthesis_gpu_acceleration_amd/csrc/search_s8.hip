// search_s8.hip -- the counting kernels for plans with at most 8 variable slots.
#include "search.hpp"

namespace gr {
template struct SlotKernels<8>;
}  // namespace gr
