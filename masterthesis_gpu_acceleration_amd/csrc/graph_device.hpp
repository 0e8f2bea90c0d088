// graph_device.hpp -- building the grouped adjacency arrays on the GPU.
#pragma once
#include <cstdint>
#include <string>

#include "graph_host.hpp"

namespace gr {

// Fills g (row space nTerms, no term strings) from n term-id triples, with
// the triple sorts on the current GPU; the arrays are identical to
// build_graph's. False (with err) if the keys do not fit 64 bits, device
// memory is short or a HIP call fails -- the caller then builds on the host.
bool build_graph_device(HostGraph& g, const int32_t* s, const int32_t* p, const int32_t* o, int64_t n,
                        int64_t nTerms, std::string& err);

// Whether a load of n triples uses the device build: GPU_RULES_DEVICE_BUILD
// = 0 never, 1 always, unset from 2^20 triples on.
bool device_build_wanted(int64_t n);

}  // namespace gr
