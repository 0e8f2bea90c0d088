// search_s4.hip -- the counting kernels for plans with at most 4 variable slots.
#include "search.hpp"

namespace gr {
template struct SlotKernels<4>;
}  // namespace gr
