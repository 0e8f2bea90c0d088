// oracle.cpp -- CPU restatement of the rule support / body-size counting
// semantics. TEST INFRASTRUCTURE ONLY: tests/, __graft_entry__.smoke() and
// bench.py's cpu_baseline leg use it to check results. The product library
// (libgpu_rules.so) never links, loads or calls it.
//
// Semantics restated from the reference's two CPU counters (the algorithm is
// re-expressed here, not copied):
//   * head matching: constants must equal, two distinct head variables must
//     bind distinct entities, a reflexive head (?a p ?a) only matches
//     self-loops -- Rewritten_CPP/SupportCounting.cpp:242-284,
//     CPU_01_nonparallel/SupportCounting.cpp:9-34.
//   * headSize = |P_head| (deduplicated); headSupport is constant-aware --
//     Rewritten_CPP/SupportCounting.cpp:142-201,
//     CPU_01_nonparallel/SupportCounting.cpp:474-507.
//   * a body grounding is injective: distinct variables bind distinct
//     entities (constants take no part), and no ground triple is used twice,
//     the head triple included for support -- Rewritten_CPP/
//     SupportCounting.cpp:299-479 (tryRecurse / containsAtom),
//     CPU_01_nonparallel/SupportCounting.cpp:83-247.
//   * support = number of matching head triples whose body has >= 1 such
//     grounding. This is CPU_01's full count; Rewritten's minSupport=1 scan
//     (Rewritten_CPP/SupportCounting.cpp:208-236) can stop one head binding
//     early and report 0 where the true count is 1 -- that divergence is
//     recorded in the golden fixtures, not reproduced here.
//   * bodyPairs = number of distinct (headS, headO) bindings over all body
//     groundings (no head triple in the used set) --
//     Rewritten_CPP/SupportCounting.cpp:557-733,
//     CPU_01_nonparallel/SupportCounting.cpp:563-631. 0 if a head term is a
//     constant or a head variable does not occur in the body (Rewritten's
//     "both bound" condition, :585-593).
//   * the ABI's bodySize is bodyPairs gated by support > 0
//     (Rewritten_CPP/SupportCounting.cpp:557-561, GPU_03/gpu_rules_lib.cu:441).
//
// Rule encoding (shared with the library's structured batch entry points):
// a term >= 0 is a constant term id, a term < 0 is variable -(term+1).
#include <algorithm>
#include <array>
#include <cstdint>
#include <cstring>
#include <unordered_map>
#include <unordered_set>
#include <vector>

namespace {

struct PredIdx {
    std::vector<std::pair<int, int>> spo;                  // sorted, unique
    std::unordered_map<int, std::vector<int>> out, in;     // s -> objs, o -> subjs (sorted)
};

struct Graph {
    std::unordered_map<int, PredIdx> preds;

    const PredIdx* get(int p) const {
        auto it = preds.find(p);
        return it == preds.end() ? nullptr : &it->second;
    }
    static const std::vector<int>* row(const std::unordered_map<int, std::vector<int>>& m, int k) {
        auto it = m.find(k);
        return it == m.end() ? nullptr : &it->second;
    }
    bool has(int p, int s, int o) const {
        const PredIdx* pi = get(p);
        if (!pi) return false;
        const std::vector<int>* r = row(pi->out, s);
        return r && std::binary_search(r->begin(), r->end(), o);
    }
};

struct Term { bool var; int v; };
struct Atom { Term s; int p; Term o; };

struct Rule {
    Atom head;
    std::vector<Atom> body;
    int nvars = 0;
};

Term decode(int32_t t, std::unordered_map<int, int>& vmap, int& nv) {
    if (t >= 0) return {false, t};
    int key = -(t + 1);
    auto it = vmap.find(key);
    if (it != vmap.end()) return {true, it->second};
    vmap[key] = nv;
    return {true, nv++};
}

struct StopAtFirst { bool operator()() const { return true; } };
struct NeverStop { bool operator()() const { return false; } };

// Backtracking search over the body. `val[v]` < 0 means unbound.
struct Search {
    const Graph& g;
    const Rule& r;
    std::vector<int> val;
    std::vector<char> done;
    std::vector<std::array<int, 3>> used;
    int64_t nodes = 0;

    Search(const Graph& g_, const Rule& r_) : g(g_), r(r_), val(r_.nvars, -1), done(r_.body.size(), 0) {}

    int tv(const Term& t) const { return t.var ? val[t.v] : t.v; }
    bool bound(const Term& t) const { return !t.var || val[t.v] >= 0; }

    bool taken(int x) const {
        for (int b : val) if (b == x) return true;
        return false;
    }
    bool usedTriple(int s, int p, int o) const {
        for (auto& u : used) if (u[0] == s && u[1] == p && u[2] == o) return true;
        return false;
    }

    // Pick the next atom: most bound terms first, then the smallest
    // candidate list (any order gives the same set of groundings).
    int pick() const {
        int best = -1;
        long bestScore = 0;
        for (size_t i = 0; i < r.body.size(); ++i) {
            if (done[i]) continue;
            const Atom& a = r.body[i];
            const PredIdx* pi = g.get(a.p);
            long cand;
            if (!pi) cand = 0;
            else if (bound(a.s) && bound(a.o)) cand = 1;
            else if (bound(a.s)) { auto* rr = Graph::row(pi->out, tv(a.s)); cand = rr ? (long)rr->size() : 0; }
            else if (bound(a.o)) { auto* rr = Graph::row(pi->in, tv(a.o)); cand = rr ? (long)rr->size() : 0; }
            else cand = (long)pi->spo.size() + 1;
            if (best < 0 || cand < bestScore) { best = (int)i; bestScore = cand; }
        }
        return best;
    }

    // Try to bind term t to x. Returns 0 = mismatch, 1 = ok (nothing bound),
    // 2 = ok (t's variable newly bound; caller must unbind).
    int bindTerm(const Term& t, int x) {
        if (!t.var) return t.v == x ? 1 : 0;
        int cur = val[t.v];
        if (cur >= 0) return cur == x ? 1 : 0;
        if (taken(x)) return 0;
        val[t.v] = x;
        return 2;
    }

    // Optional pair pruning for the body-pair count: once both head
    // variables are bound, the rest of the subtree can only yield that one
    // pair, so it is skipped if already recorded and otherwise decided by an
    // existence search.
    std::unordered_set<uint64_t>* pairSet = nullptr;
    int hs = -1, ho = -1;

    // visit(): called at a full grounding; returns true to stop the search.
    template <class Visit>
    bool run(int depth, Visit& visit) {
        ++nodes;
        if (pairSet && val[hs] >= 0 && val[ho] >= 0) {
            uint64_t k = ((uint64_t)(uint32_t)val[hs] << 32) | (uint32_t)val[ho];
            if (pairSet->count(k)) return false;
            auto* keep = pairSet;
            pairSet = nullptr;
            StopAtFirst found;
            bool ok = run(depth, found);
            pairSet = keep;
            if (ok) pairSet->insert(k);
            return false;
        }
        if (depth == (int)r.body.size()) return visit();
        int ai = pick();
        const Atom& a = r.body[ai];
        const PredIdx* pi = g.get(a.p);
        if (!pi) return false;
        done[ai] = 1;
        bool stop = false;
        auto tryTriple = [&](int s, int o) -> bool {
            int bs = bindTerm(a.s, s);
            if (!bs) return false;
            int bo = bindTerm(a.o, o);
            if (!bo) { if (bs == 2) val[a.s.v] = -1; return false; }
            bool st = false;
            if (!usedTriple(s, a.p, o)) {
                used.push_back({s, a.p, o});
                st = run(depth + 1, visit);
                used.pop_back();
            }
            if (bo == 2) val[a.o.v] = -1;
            if (bs == 2) val[a.s.v] = -1;
            return st;
        };
        if (bound(a.s) && bound(a.o)) {
            int s = tv(a.s), o = tv(a.o);
            if (g.has(a.p, s, o)) stop = tryTriple(s, o);
        } else if (bound(a.s)) {
            int s = tv(a.s);
            if (auto* rr = Graph::row(pi->out, s))
                for (int o : *rr) if ((stop = tryTriple(s, o))) break;
        } else if (bound(a.o)) {
            int o = tv(a.o);
            if (auto* rr = Graph::row(pi->in, o))
                for (int s : *rr) if ((stop = tryTriple(s, o))) break;
        } else {
            for (auto& t : pi->spo) if ((stop = tryTriple(t.first, t.second))) break;
        }
        done[ai] = 0;
        return stop;
    }
};

struct Result {
    int64_t support, headSize, headSupport, bodyPairs, bodySize, skipped, nodes;
};

Result evalRule(const Graph& g, const Rule& r, bool withPairs = true) {
    Result res{};
    const PredIdx* hp = g.get(r.head.p);
    res.headSize = hp ? (int64_t)hp->spo.size() : 0;
    // constant-aware head support
    if (hp) {
        bool sc = !r.head.s.var, oc = !r.head.o.var;
        if (!sc && !oc) res.headSupport = res.headSize;
        else if (sc && !oc) { auto* rr = Graph::row(hp->out, r.head.s.v); res.headSupport = rr ? rr->size() : 0; }
        else if (!sc && oc) { auto* rr = Graph::row(hp->in, r.head.o.v); res.headSupport = rr ? rr->size() : 0; }
        else res.headSupport = g.has(r.head.p, r.head.s.v, r.head.o.v) ? 1 : 0;
    }
    int nb = (int)r.body.size();
    if (nb < 1 || nb > 8 || r.nvars > 16) { res.skipped = 1; return res; }
    if (!hp) return res;

    // ---- support
    {
        Search S(g, r);
        for (auto& t : hp->spo) {
            int s = t.first, o = t.second;
            std::fill(S.val.begin(), S.val.end(), -1);
            S.used.clear();
            if (S.bindTerm(r.head.s, s) == 0) continue;
            if (S.bindTerm(r.head.o, o) == 0) continue;
            S.used.push_back({s, r.head.p, o});
            StopAtFirst visit;
            if (S.run(0, visit)) ++res.support;
        }
        res.nodes += S.nodes;
    }

    // ---- distinct body pairs
    bool closed = withPairs && r.head.s.var && r.head.o.var;
    if (closed) {
        bool sIn = false, oIn = false;
        for (auto& a : r.body) {
            for (const Term* t : {&a.s, &a.o}) {
                if (t->var && t->v == r.head.s.v) sIn = true;
                if (t->var && t->v == r.head.o.v) oIn = true;
            }
        }
        closed = sIn && oIn;
    }
    if (closed) {
        Search S(g, r);
        std::unordered_set<uint64_t> pairs;
        S.pairSet = &pairs;
        S.hs = r.head.s.v;
        S.ho = r.head.o.v;
        NeverStop visit;  // leaves are reached only through the pruned existence search
        S.run(0, visit);
        res.bodyPairs = (int64_t)pairs.size();
        res.nodes += S.nodes;
    }
    res.bodySize = res.support > 0 ? res.bodyPairs : 0;
    return res;
}

}  // namespace

struct OracleGraph { Graph g; };

extern "C" {

// Builds the oracle's index from n deduplicated-or-not triples.
void* oracle_graph_new(const int32_t* s, const int32_t* p, const int32_t* o, int64_t n) {
    auto* G = new OracleGraph();
    for (int64_t i = 0; i < n; ++i) G->g.preds[p[i]].spo.emplace_back(s[i], o[i]);
    for (auto& kv : G->g.preds) {
        auto& pi = kv.second;
        std::sort(pi.spo.begin(), pi.spo.end());
        pi.spo.erase(std::unique(pi.spo.begin(), pi.spo.end()), pi.spo.end());
        for (auto& t : pi.spo) { pi.out[t.first].push_back(t.second); pi.in[t.second].push_back(t.first); }
        for (auto& r : pi.out) std::sort(r.second.begin(), r.second.end());
        for (auto& r : pi.in) std::sort(r.second.begin(), r.second.end());
    }
    return G;
}

void oracle_graph_free(void* g) { delete static_cast<OracleGraph*>(g); }

// rules: n_rules records of 28 int32 words: [nBody, head(s,p,o), body[8](s,p,o)].
// out: n_rules records of 7 int64: support, headSize, headSupport, bodyPairs,
// bodySize (gated), skipped, nodes.
static int eval_records(void* gp, const int32_t* rules, int64_t n_rules, int64_t* out, bool withPairs) {
    const Graph& g = static_cast<OracleGraph*>(gp)->g;
    for (int64_t i = 0; i < n_rules; ++i) {
        const int32_t* w = rules + i * 28;
        int nb = w[0];
        Rule r;
        std::unordered_map<int, int> vmap;
        int nv = 0;
        int nbody = std::max(0, std::min(nb, 8));
        for (int j = 0; j < nbody; ++j) {
            const int32_t* a = w + 4 + 3 * j;
            Atom at;
            at.s = decode(a[0], vmap, nv);
            at.o = decode(a[2], vmap, nv);
            at.p = a[1];
            r.body.push_back(at);
        }
        r.head.s = decode(w[1], vmap, nv);
        r.head.o = decode(w[3], vmap, nv);
        r.head.p = w[2];
        r.nvars = nv;
        if (nb > 8) r.body.resize(9);  // marks "too long" for the skip rule
        Result res = evalRule(g, r, withPairs);
        int64_t* o = out + i * 7;
        o[0] = res.support; o[1] = res.headSize; o[2] = res.headSupport;
        o[3] = res.bodyPairs; o[4] = res.bodySize; o[5] = res.skipped; o[6] = res.nodes;
    }
    return 0;
}

int oracle_eval(void* gp, const int32_t* rules, int64_t n_rules, int64_t* out) {
    return eval_records(gp, rules, n_rules, out, true);
}

// The same records without the body-pair search: bodyPairs and bodySize 0
// (support and the head figures only).
int oracle_eval_support(void* gp, const int32_t* rules, int64_t n_rules, int64_t* out) {
    return eval_records(gp, rules, n_rules, out, false);
}

}  // extern "C"
