// traffic.cpp -- algorithmic HBM bytes of a plan.
//
// Walks the device plan seed by seed on the host, in the order a thread-mode
// lane would (same levels, same early exits, same pair-level settling), and
// counts the bytes the search must read from the graph arrays:
//   seed              8 B (SUPPORT: coo + col of the head entry) or 4 B (PAIRS key)
//   row lookup        16 B key-index block per EXT/CHECK level entered,
//                     plus 8 B of row bounds when the key has a row (a PAIRS
//                     seed's level-0 row, when Plan::seedRow0, 8 B alone)
//   CHECK probe       4 B per binary-search step (plus the final compare)
//   EXT candidate     4 B (col)
//   SCAN candidate    8 B (coo + col)
//   ISECT             an EXT candidate, plus a second row lookup per level
//                     entered and 4 B per row-2 probe of the forward cursor
//                     (gallop + binary search); the CHECK after it costs nothing
//   single frame      a PAIRS seed whose search reaches a plain-EXT last pair
//                     level with nothing counted and no untried entry above,
//                     on a row of at least 16 entries, counts that row
//                     without enumerating it (as the thread-mode kernel
//                     does): one binary search per bound variable; with
//                     fused support the head row lookups first, then the
//                     head-row entries (4 B each) and a binary search each
//   fused support     PAIRS plans with supPred >= 0, per seed with a body pair:
//                     the seed's head row lookup (as a row lookup), then, as the device
//                     checks it: up to 12 distinct head objects (a thread-mode
//                     seed) a binary search of the row (as a CHECK probe) per
//                     object; more (only wave mode holds them) 20 B per row
//                     entry (col + the key-index block of its rank)
// It returns no counts to callers; the counting path is the device's.
#include <algorithm>
#include <atomic>
#include <cstdint>
#include <thread>
#include <vector>

#include "graph_host.hpp"
#include "plan.hpp"
#include "traffic.hpp"

namespace gr {

namespace {

struct Walker {
    const Plan& P;
    const ConflictRefs& R;
    const HostGraph& g;
    int64_t T;
    int64_t bytes = 0;
    int64_t nodes = 0;
    int v[kMaxSlots];
    // distinct head objects of the current PAIRS seed: one bit per term id,
    // with the words set listed so a seed clears only those
    std::vector<uint64_t>& bits;
    std::vector<int64_t>& touched;
    const std::vector<int32_t>* heads = nullptr;   // fused support of several head predicates
    bool shortOk = false;    // the plan's last level is a plain-EXT pair level (single-frame counts apply)
    bool counted = false;    // the current seed was counted by a single-frame count (its support too)
    static constexpr int64_t kShortMin = 16, kShortHead = 16;   // as the thread-mode kernel

    bool seen(int x) const { return (bits[x >> 6] >> (x & 63)) & 1u; }
    void insert(int x) {
        uint64_t& w = bits[x >> 6];
        if (!w) touched.push_back(x >> 6);
        w |= 1ull << (x & 63);
    }
    void clear_set() {
        for (int64_t w : touched) bits[w] = 0;
        touched.clear();
    }

    int64_t cur[kMaxBody] = {};                          // ISECT: entry of the match in row 2
    int64_t oLo[kMaxBody] = {}, oHi[kMaxBody] = {};      // ISECT: the row searched
    bool swp[kMaxBody] = {};                             // ISECT: enumerating row 2

    int col(int64_t e) const { return e < T ? g.spoVal[e] : g.posVal[e - T]; }
    int coo(int64_t e) const { return e < T ? g.spoKey[e] : g.posKey[e - T]; }
    int tv(int r) const { return r >= 0 ? v[r] : -(r + 1); }

    bool repeats(int L, int mask, int s, int o) const {
        for (int j = 0; j < kMaxBody; ++j)
            if ((mask & (1 << j)) && tv(R.sref[j]) == s && tv(R.oref[j]) == o) return true;
        return (mask & 256) && tv(R.hs) == s && tv(R.ho) == o;
        (void)L;
    }

    // Entry range of the row of term `key` in block run `base`.
    void key_row(int64_t base, int key, int64_t& lo, int64_t& hi) {
        const int64_t per = (int64_t)g.P() * g.nBlocks();
        const int dir = (int)(base / per);
        const int pred = (int)((base % per) / g.nBlocks());
        auto r = g.row(dir, pred, key);
        bytes += 16;
        if (r.second > r.first) bytes += 8;
        lo = dir * T + r.first;
        hi = dir * T + r.second;
    }

    // Membership test of ISECT candidate entry i in the row the level
    // searches (a binary search, as the device's find_in_row reads it); on a
    // match cur[L] is the match's entry in row 2.
    bool isect_find(int L, int64_t i) {
        const int x = col(i);
        const int64_t m = find_interp(oLo[L], oHi[L], x);
        if (m < 0) return false;
        cur[L] = swp[L] ? i : m;
        return true;
    }

    // The device's find_in_row_interp: the row's first and last values are
    // read once per level entered (counted there); a target outside them
    // reads nothing; up to three interpolation probes, then a binary search.
    int64_t find_interp(int64_t lo, int64_t hi, int x) {
        if (lo >= hi) return -1;
        const int vlo = col(lo), vhi = col(hi - 1);
        if (x < vlo || x > vhi) return -1;
        if (x == vlo) return lo;
        if (x == vhi) return hi - 1;
        int64_t l = lo, h = hi - 1;
        int vl = vlo, vh = vhi;
        for (int it = 0; it < 3; ++it) {
            if (h - l <= 8) break;
            const float f = (float)(x - vl) / (float)(vh - vl);
            int64_t m = l + (int64_t)(int)(f * (float)(int)(h - l));
            m = m <= l ? l + 1 : (m >= h ? h - 1 : m);
            bytes += 4;
            const int y = col(m);
            if (y == x) return m;
            if (y < x) { l = m; vl = y; } else { h = m; vh = y; }
        }
        int64_t a = l + 1, b = h;
        while (a < b) {
            const int64_t m = (a + b) >> 1;
            bytes += 4;
            if (col(m) < x) a = m + 1; else b = m;
        }
        if (a < h) bytes += 4;
        return (a < h && col(a) == x) ? a : -1;
    }

    void range(int L, int64_t& lo, int64_t& hi) {
        const Level& lv = P.lev[L];
        const int kind = level_kind(lv.flags);
        if (kind == kScan) { lo = lv.lo; hi = lv.hi; return; }
        if (kind == kCheck && L > 0 && level_kind(P.lev[L - 1].flags) == kIsect) {   // the intersection's match
            lo = cur[L - 1];
            hi = lo + 1;
            return;
        }
        const int64_t base = (int64_t)(uint32_t)lv.lo | ((int64_t)lv.hi << 32);
        if (kind == kIsect) {
            // both rows list candidates of the new variable: enumerate the
            // shorter, search the other (as the level-synchronous kernels do)
            const int64_t base2 = (int64_t)(uint32_t)lv.lo2 | ((int64_t)lv.hi2 << 32);
            int64_t a2, b2;
            key_row(base2, tv(lv.val2), a2, b2);
            if (a2 >= b2) { lo = hi = 0; return; }
            key_row(base, tv(lv.key), lo, hi);
            if (lo >= hi) return;
            swp[L] = b2 - a2 < hi - lo;
            if (swp[L]) {
                oLo[L] = lo; oHi[L] = hi;
                lo = a2; hi = b2;
            } else {
                oLo[L] = a2; oHi[L] = b2;
            }
            bytes += 8;   // the searched row's first and last values
            return;
        }
        key_row(base, tv(lv.key), lo, hi);
        if (L == 0 && P.kind == kPairs && P.seedRow0) bytes -= 16;   // the seed's row from rowStart alone
        if (kind == kCheck) {
            const int target = tv(lv.val);
            int64_t l = lo, h = hi;
            while (l < h) {
                const int64_t m = (l + h) >> 1;
                bytes += 4;
                if (col(m) < target) l = m + 1; else h = m;
            }
            if (l < hi) bytes += 4;
            if (l < hi && col(l) == target) { lo = l; hi = l + 1; }
            else { lo = hi = 0; }
        }
    }

    bool bind(int L, int64_t i) {
        const Level& lv = P.lev[L];
        const int f = lv.flags, kind = level_kind(f), nb = level_nbound(f);
        const int x = col(i);
        int k = 0;
        auto taken = [&](int y) { for (int j = 0; j < nb; ++j) if (v[j] == y) return true; return false; };
        if (kind == kExt || kind == kIsect) {
            bytes += 4;
            if (taken(x)) return false;
            v[lv.val] = x;
        } else if (kind == kScan) {
            bytes += 8;
            k = coo(i);
            if (lv.val == lv.val2) {
                if (k != x || taken(k)) return false;
                v[lv.val] = k;
            } else {
                if (k == x || taken(k) || taken(x)) return false;
                v[lv.val] = k;
                v[lv.val2] = x;
            }
        }
        const int conf = level_conflicts(f);
        if (conf) {
            if (kind != kScan) k = tv(lv.key);
            const bool spo = level_dir(f) == kSPO;
            if (repeats(L, conf, spo ? k : x, spo ? x : k)) return false;
        }
        return true;
    }

    // Binary search of x in [lo, hi) as the device's find_in_row reads it.
    bool search(int64_t lo, int64_t hi, int x) {
        int64_t l = lo, h = hi;
        while (l < h) {
            const int64_t m = (l + h) >> 1;
            bytes += 4;
            if (col(m) < x) l = m + 1; else h = m;
        }
        if (l < hi) bytes += 4;
        return l < hi && col(l) == x;
    }

    // Single-frame count of the seed's last row [lo, hi) (see the header);
    // false if a head row is too long, when the row is enumerated instead.
    bool single_frame(int L, int64_t lo, int64_t hi) {
        const int nb = level_nbound(P.lev[L].flags);
        std::vector<int32_t> preds;
        if (heads) preds = *heads;
        else if (P.supPred >= 0) preds.push_back(P.supPred);
        for (int32_t q : preds) {
            int64_t a, b;
            key_row(((int64_t)kSPO * g.P() + q) * g.nBlocks(), v[0], a, b);
            if (b - a > kShortHead) return false;
        }
        for (int j = 0; j < nb; ++j) search(lo, hi, v[j]);
        for (int32_t q : preds) {
            int64_t a, b;
            key_row(((int64_t)kSPO * g.P() + q) * g.nBlocks(), v[0], a, b);
            for (int64_t x = a; x < b; ++x) {
                bytes += 4;
                const int h = col(x);
                bool t = false;
                for (int j = 0; j < nb; ++j) t = t || v[j] == h;
                if (!t) search(lo, hi, h);
            }
        }
        counted = true;
        return true;
    }

    // Returns true when a complete grounding was found (existence mode).
    // `only`: no untried entry is left in the frames above this level.
    bool dfs(int L, bool pairs, bool only) {
        ++nodes;
        int64_t lo, hi;
        range(L, lo, hi);
        if (pairs && shortOk && L + 1 == P.nlev && only && touched.empty() && hi - lo >= kShortMin &&
            single_frame(L, lo, hi))
            return false;
        const bool isect = level_kind(P.lev[L].flags) == kIsect;
        for (int64_t i = lo; i < hi; ++i) {
            if (!bind(L, i)) continue;
            if (isect && !isect_find(L, i)) continue;
            if (pairs && L == P.pairLevel) {
                const int ho = tv(P.hoRef);
                if (seen(ho)) continue;
                if (L + 1 == P.nlev || dfs(L + 1, false, false)) insert(ho);
                continue;
            }
            if (L + 1 == P.nlev) {
                if (!pairs) return true;
                continue;
            }
            if (dfs(L + 1, pairs, only && i + 1 == hi) && !pairs) return true;
            if (counted) return false;
        }
        return false;
    }

    void seed(int64_t item) {
        for (int& x : v) x = -1;
        clear_set();
        if (P.kind == kSupport) {
            bytes += 8;
            const int64_t e = P.seedBegin + item;
            const int key = coo(e), val = col(e);
            const int s = P.seedDir == kSPO ? key : val, o = P.seedDir == kSPO ? val : key;
            if (P.hsRef < 0 && -(P.hsRef + 1) != s) return;
            if (P.hoRef < 0 && -(P.hoRef + 1) != o) return;
            if (P.hsRef >= 0 && P.hoRef >= 0 && ((P.hsRef == P.hoRef) != (s == o))) return;
            if (P.hsRef >= 0) v[P.hsRef] = s;
            if (P.hoRef >= 0) v[P.hoRef] = o;
            dfs(0, false, false);
            return;
        }
        bytes += 4;
        const int64_t ks = P.seedBegin + item;
        v[0] = ks < (int64_t)g.keysS.size() ? g.keysS[ks] : g.keysO[ks - g.keysS.size()];
        const int last = P.nlev - 1;
        shortOk = P.pairLevel >= 0 && P.pairLevel == last && level_kind(P.lev[last].flags) == kExt &&
                  level_conflicts(P.lev[last].flags) == 0;
        counted = false;
        dfs(0, P.pairLevel >= 0, true);
        if (P.pairLevel >= 0 && !counted) {
            if (heads)
                for (int32_t q : *heads) head_hits(q);
            else if (P.supPred >= 0)
                head_hits(P.supPred);
        }
    }

    // Fused support: the seed's head row, then one binary search of it per
    // distinct head object found (thread mode, at most kThreadSet objects),
    // or one pass over it (wave mode).
    static constexpr int64_t kThreadSet = 12;
    void head_hits(int32_t pred) {
        int64_t n = 0;
        for (int64_t w : touched) n += __builtin_popcountll(bits[w]);
        if (n == 0) return;   // no body pair: nothing to look up
        const int64_t base = ((int64_t)kSPO * g.P() + pred) * g.nBlocks();
        int64_t lo, hi;
        key_row(base, v[0], lo, hi);
        if (lo >= hi) return;
        if (n > kThreadSet) {
            bytes += 20 * (hi - lo);
            return;
        }
        for (int64_t w : touched)
            for (uint64_t m = bits[w]; m; m &= m - 1) {
                const int x = (int)(w * 64 + __builtin_ctzll(m));
                int64_t l = lo, h = hi;
                while (l < h) {
                    const int64_t mid = (l + h) >> 1;
                    bytes += 4;
                    if (col(mid) < x) l = mid + 1; else h = mid;
                }
                if (l < hi) bytes += 4;
            }
    }
};

}  // namespace

TrafficEstimate plan_traffic(const Plan& P, const ConflictRefs& R, const HostGraph& g, int64_t items) {
    std::vector<uint64_t> bits((size_t)g.nBlocks(), 0);
    std::vector<int64_t> touched;
    Walker w{P, R, g, g.T(), 0, 0, {}, bits, touched};
    for (int64_t i = 0; i < items; ++i) w.seed(i);
    return TrafficEstimate{w.bytes, w.nodes};
}

void plan_traffic_many(const std::vector<TrafficJob>& jobs, const HostGraph& g, std::vector<TrafficEstimate>& out,
                       int threads) {
    constexpr int64_t kSeedsPerTask = 512;
    struct Task {
        size_t job;
        int64_t begin, end;
    };
    std::vector<Task> tasks;
    for (size_t j = 0; j < jobs.size(); ++j)
        for (int64_t b = 0; b < jobs[j].items; b += kSeedsPerTask)
            tasks.push_back(Task{j, b, std::min(jobs[j].items, b + kSeedsPerTask)});
    std::vector<std::atomic<int64_t>> bytes(jobs.size()), nodes(jobs.size());
    for (size_t j = 0; j < jobs.size(); ++j) {
        bytes[j] = 0;
        nodes[j] = 0;
    }
    std::atomic<size_t> next{0};
    auto work = [&]() {
        std::vector<uint64_t> bits((size_t)g.nBlocks(), 0);
        std::vector<int64_t> touched;
        for (;;) {
            const size_t t = next.fetch_add(1);
            if (t >= tasks.size()) break;
            const TrafficJob& J = jobs[tasks[t].job];
            Walker w{*J.plan, *J.refs, g, g.T(), 0, 0, {}, bits, touched};
            w.heads = J.heads;
            for (int64_t i = tasks[t].begin; i < tasks[t].end; ++i) w.seed(i);
            w.clear_set();
            bytes[tasks[t].job] += w.bytes;
            nodes[tasks[t].job] += w.nodes;
        }
    };
    threads = std::max(1, std::min<int>(threads, (int)std::max<size_t>(tasks.size(), 1)));
    std::vector<std::thread> pool;
    for (int k = 1; k < threads; ++k) pool.emplace_back(work);
    work();
    for (auto& th : pool) th.join();
    out.assign(jobs.size(), TrafficEstimate{0, 0});
    for (size_t j = 0; j < jobs.size(); ++j) out[j] = TrafficEstimate{bytes[j].load(), nodes[j].load()};
}

}  // namespace gr

namespace gr {

namespace {

// Injective groundings of a rule body (variables only) by depth-first search
// over its atoms, each atom taken when it has a bound term (the one with the
// fewest triples first); a last atom that binds one fresh variable is
// counted without enumerating its row: the row's entries minus the bound
// values it holds (a row holds each value once).
struct GroundingCounter {
    const HostGraph& g;
    const std::vector<AtomIR>& body;
    std::vector<int32_t> pred;    // dense predicate per atom
    std::vector<int> order;       // atoms in search order
    std::vector<int32_t> val;     // per variable: bound value or -1
    std::vector<int32_t> boundVals;
    int64_t count = 0;

    int32_t vcol(int dir, int64_t e) const { return dir == 0 ? g.spoVal[(size_t)e] : g.posVal[(size_t)e]; }
    bool taken(int32_t x) const {
        for (int32_t y : val)
            if (y == x) return true;
        return false;
    }
    bool in_row(int dir, std::pair<int32_t, int32_t> r, int32_t x) const {
        int32_t lo = r.first, hi = r.second;
        while (lo < hi) {
            const int32_t m = (lo + hi) >> 1;
            if (vcol(dir, m) < x) lo = m + 1; else hi = m;
        }
        return lo < r.second && vcol(dir, lo) == x;
    }
    void dfs(size_t k) {
        if (k == order.size()) {
            ++count;
            return;
        }
        const AtomIR& a = body[(size_t)order[k]];
        const int32_t p = pred[(size_t)order[k]];
        const int sv = a.s.v, ov = a.o.v;
        const bool sb = val[(size_t)sv] >= 0, ob = val[(size_t)ov] >= 0;
        if (sb && ob) {
            if (g.has(p, val[(size_t)sv], val[(size_t)ov])) dfs(k + 1);
            return;
        }
        if (sb || ob) {
            const int dir = sb ? 0 : 1;   // key = the bound term
            const int32_t key = sb ? val[(size_t)sv] : val[(size_t)ov];
            const int fv = sb ? ov : sv;
            if (fv == (sb ? sv : ov)) {   // p(x, x) with x bound
                if (g.has(p, key, key)) dfs(k + 1);
                return;
            }
            const auto r = g.row(dir, p, key);
            if (k + 1 == order.size()) {   // last atom: count the row
                int64_t n = r.second - r.first;
                for (int32_t y : val)
                    if (y >= 0 && in_row(dir, r, y)) --n;
                count += n;
                return;
            }
            for (int32_t e = r.first; e < r.second; ++e) {
                const int32_t x = vcol(dir, e);
                if (taken(x)) continue;
                val[(size_t)fv] = x;
                dfs(k + 1);
                val[(size_t)fv] = -1;
            }
            return;
        }
        // neither bound: every triple of the predicate
        for (int32_t e = g.predOff[(size_t)p]; e < g.predOff[(size_t)p + 1]; ++e) {
            const int32_t s = g.spoKey[(size_t)e], o = g.spoVal[(size_t)e];
            if (sv == ov) {
                if (s != o || taken(s)) continue;
                val[(size_t)sv] = s;
                dfs(k + 1);
                val[(size_t)sv] = -1;
                continue;
            }
            if (s == o || taken(s) || taken(o)) continue;
            val[(size_t)sv] = s;
            val[(size_t)ov] = o;
            dfs(k + 1);
            val[(size_t)sv] = val[(size_t)ov] = -1;
        }
    }
};

}  // namespace

int64_t body_groundings(const RuleIR& r, const HostGraph& g) {
    if (r.body.empty() || r.body.size() > (size_t)kMaxBody) return -1;
    if (!r.head.s.var || !r.head.o.var) return -1;
    int32_t nv = 0;
    bool hs = false, ho = false;
    for (const AtomIR& a : r.body) {
        if (!a.s.var || !a.o.var) return -1;
        nv = std::max({nv, a.s.v + 1, a.o.v + 1});
        hs = hs || a.s.v == r.head.s.v || a.o.v == r.head.s.v;
        ho = ho || a.s.v == r.head.o.v || a.o.v == r.head.o.v;
    }
    if (!hs || !ho) return -1;
    GroundingCounter c{g, r.body, {}, {}, std::vector<int32_t>((size_t)nv, -1), {}, 0};
    for (const AtomIR& a : r.body) {
        const int32_t p = g.dense_pred(a.pterm);
        if (p < 0) return 0;
        c.pred.push_back(p);
    }
    // order: the atom with the fewest triples, then atoms with a bound term
    // (both bound first, then the fewest triples)
    std::vector<char> used(r.body.size(), 0), bound((size_t)nv, 0);
    for (size_t k = 0; k < r.body.size(); ++k) {
        int best = -1;
        int64_t bestKey = 0;
        for (size_t j = 0; j < r.body.size(); ++j) {
            if (used[j]) continue;
            const AtomIR& a = r.body[j];
            const int nb = (bound[(size_t)a.s.v] ? 1 : 0) + (bound[(size_t)a.o.v] ? 1 : 0);
            if (k > 0 && nb == 0) continue;
            const int64_t key = (int64_t)(2 - nb) * (int64_t(1) << 40) + g.nnz(c.pred[j]);
            if (best < 0 || key < bestKey) {
                best = (int)j;
                bestKey = key;
            }
        }
        if (best < 0) {   // a part of the body not joined to the rest: take it as a new start
            for (size_t j = 0; j < r.body.size(); ++j)
                if (!used[j] && (best < 0 || g.nnz(c.pred[j]) < g.nnz(c.pred[(size_t)best]))) best = (int)j;
        }
        used[(size_t)best] = 1;
        bound[(size_t)r.body[(size_t)best].s.v] = bound[(size_t)r.body[(size_t)best].o.v] = 1;
        c.order.push_back(best);
    }
    c.dfs(0);
    return c.count;
}

}  // namespace gr
