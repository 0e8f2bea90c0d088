// traffic.hpp -- host model of the bytes a plan's search reads from HBM.
#pragma once
#include <cstdint>
#include <vector>

#include "graph_host.hpp"
#include "plan.hpp"
#include "rules_host.hpp"

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

// Injective groundings of r's body: the pairs the reference's body-size
// kernel emits before they are deduplicated (GPU_03/gpu_kernels.cuh:
// 218-317, which emits one (head subject, head object) pair per grounding).
// -1 for a body with constants, more than kMaxBody atoms, or without both
// head variables (no such pairs).
int64_t body_groundings(const RuleIR& r, const HostGraph& g);

}  // namespace gr
