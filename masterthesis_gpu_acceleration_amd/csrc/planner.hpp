// planner.hpp -- compiles a rule into device plans.
#pragma once
#include <cstdint>
#include <vector>

#include "graph_host.hpp"
#include "plan.hpp"
#include "rules_host.hpp"

namespace gr {

enum RuleStatus : int32_t {
    kRuleOk = 0,        // counted on the device (or statically zero)
    kRuleSkipped = 1,   // unsupported shape: body not in [1,8] atoms or > 16 variables
    kRuleError = 2,     // parse error
};

struct Tuning {
    int32_t budget = 4096;       // thread-mode candidates per seed before escalation
};

struct CompiledRule {
    int32_t status = kRuleOk;
    int64_t headSize = 0;
    int64_t headSupport = 0;
    bool supportLaunch = false;  // false: support is statically 0
    bool pairsLaunch = false;    // false: body pairs statically 0
    Plan sup{}, supWave{};       // thread order / wave order (level 0 is the split level)
    Plan pairs{}, pairsWave{};
    ConflictRefs supRefs{}, supWaveRefs{}, pairsRefs{}, pairsWaveRefs{};
    bool supWaveOk = false, pairsWaveOk = false;
    // The pair plans can count the support too (Plan::supPred): distinct head
    // variables, the head object bound at a pair level in both plans, and no
    // body atom that could ground to the head triple itself. The head
    // predicate, or -1.
    int32_t supPred = -1;
    int64_t supItems = 0, pairItems = 0;
};

CompiledRule compile_rule(const RuleIR& r, const HostGraph& g, const Tuning& t);

// Row statistics per (direction, predicate), index dir * P + p: keys, entries
// and the sum of squared row lengths (for size-biased fan-outs).
struct RowStats {
    int32_t P = 0;
    std::vector<double> keys, nnz, sq;
};
RowStats row_stats(const HostGraph& g);

// Estimated candidates the searches of a compiled rule enumerate (support
// plan + body-pair plan), from row statistics alone: the seeds' level-0 rows
// at their mean length (at most the predicate's entries), each later level
// that extends from a variable at the larger of the mean and the size-biased
// mean row length (sum of squares / entries: the expected row of a key
// reached over an edge, which hubs dominate), ISECT at the smaller of its
// two rows, SCAN at the predicate's entries, CHECK at 1. An order-of-magnitude
// figure for scheduling, not a bound.
double estimate_work(const CompiledRule& c, const HostGraph& g, const RowStats& st);

}  // namespace gr
