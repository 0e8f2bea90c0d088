// traffic.hpp -- host model of the bytes a plan's search reads from HBM.
#pragma once
#include <cstdint>

#include "graph_host.hpp"
#include "plan.hpp"

namespace gr {

struct TrafficEstimate {
    int64_t bytes;   // algorithmic bytes over all seeds
    int64_t nodes;   // search levels entered
};

TrafficEstimate plan_traffic(const Plan& P, const ConflictRefs& R, const HostGraph& g, int64_t items);

}  // namespace gr
