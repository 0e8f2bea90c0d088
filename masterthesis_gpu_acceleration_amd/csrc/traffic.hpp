// traffic.hpp -- host model of the bytes a plan's search reads from HBM.
#pragma once
#include <cstdint>
#include <vector>

#include "graph_host.hpp"
#include "plan.hpp"

namespace gr {

struct TrafficEstimate {
    int64_t bytes;   // algorithmic bytes over all seeds
    int64_t nodes;   // search levels entered
};

TrafficEstimate plan_traffic(const Plan& P, const ConflictRefs& R, const HostGraph& g, int64_t items);

// Many plans at once on `threads` host threads (seeds split into tasks).
struct TrafficJob {
    const Plan* plan;
    const ConflictRefs* refs;
    int64_t items;
    // PAIRS: head predicates whose support the search also counts (a shared
    // search's member list); null: the plan's own supPred, if any
    const std::vector<int32_t>* heads = nullptr;
};
void plan_traffic_many(const std::vector<TrafficJob>& jobs, const HostGraph& g, std::vector<TrafficEstimate>& out,
                       int threads);

}  // namespace gr
