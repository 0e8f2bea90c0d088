// planner.cpp -- rule -> device plans.
//
// Both plans evaluate the same injective-grounding semantics as the
// reference's CPU counters (Rewritten_CPP/SupportCounting.cpp,
// CPU_01_nonparallel/SupportCounting.cpp); the atom order only changes the
// work, never the result. The order is greedy: the atom with the most bound
// terms first, then the smallest expected candidate list -- the same idea as
// the reference's chooseBestAtom / order_body (GPU_03/gpu_rules_lib.cu:93-176),
// fixed per rule so the device needs no scoring.
#include "planner.hpp"

#include <algorithm>
#include <cstring>
#include <limits>

namespace gr {

namespace {

enum Eq { kNever, kMaybe, kAlways };

Eq term_eq(const TermIR& a, const TermIR& b) {
    if (a.var && b.var) return a.v == b.v ? kAlways : kNever;   // distinct variables never share a value
    if (!a.var && !b.var) return a.v == b.v ? kAlways : kNever;
    return kMaybe;                                               // a variable may equal a constant
}

// Can the ground triples of two atoms coincide? (same predicate assumed)
Eq atom_eq(const AtomIR& a, const AtomIR& b) {
    Eq s = term_eq(a.s, b.s), o = term_eq(a.o, b.o);
    if (s == kNever || o == kNever) return kNever;
    if (s == kAlways && o == kAlways) return kAlways;
    return kMaybe;
}

struct Builder {
    const RuleIR& r;
    const HostGraph& g;
    int32_t slotOf[64];
    int32_t nslots = 0;

    Builder(const RuleIR& r_, const HostGraph& g_) : r(r_), g(g_) {
        for (auto& s : slotOf) s = -1;
    }
    bool bound(const TermIR& t) const { return !t.var || slotOf[t.v] >= 0; }
    int32_t ref(const TermIR& t) const { return t.var ? slotOf[t.v] : -(t.v + 1); }
    int32_t bind(int32_t var) {
        if (slotOf[var] < 0) slotOf[var] = nslots++;
        return slotOf[var];
    }

    double cost(const AtomIR& a, int nb2, bool sB) const {
        int32_t p = g.dense_pred(a.pterm);
        if (p < 0) return 0;
        if (nb2 == 2) return 0;
        if (nb2 == 0) return (double)g.nnz(p) + 1.0;
        int dir = sB ? 0 : 1;
        const TermIR& k = sB ? a.s : a.o;
        if (!k.var) {
            if (k.v >= g.nE) return 0;
            auto rr = g.row(dir, p, k.v);
            return rr.second - rr.first;
        }
        int32_t nk = g.nkeys(dir, p);
        return nk ? (double)g.nnz(p) / nk : 0;
    }

    // Greedy order from the current bound set (does not bind anything).
    std::vector<int> order() const {
        int nb = (int)r.body.size();
        std::vector<int> out;
        std::vector<char> used(nb, 0);
        bool vb[64];
        for (int v = 0; v < 64; ++v) vb[v] = slotOf[v] >= 0;
        auto isb = [&](const TermIR& t) { return !t.var || vb[t.v]; };
        for (int step = 0; step < nb; ++step) {
            int best = -1, bestNb = -1;
            double bestCost = 0;
            for (int i = 0; i < nb; ++i) {
                if (used[i]) continue;
                const AtomIR& a = r.body[i];
                bool sB = isb(a.s), oB = isb(a.o);
                int nb2 = (int)sB + (int)oB;
                double c = cost(a, nb2, sB);
                if (best < 0 || nb2 > bestNb || (nb2 == bestNb && c < bestCost)) {
                    best = i; bestNb = nb2; bestCost = c;
                }
            }
            used[best] = 1;
            out.push_back(best);
            const AtomIR& a = r.body[best];
            if (a.s.var) vb[a.s.v] = true;
            if (a.o.var) vb[a.o.v] = true;
        }
        return out;
    }

    // Builds levels for `ord`, binding slots as it goes. `withHead` adds
    // head-triple comparisons (support plans).
    void levels(const std::vector<int>& ord, Plan& P, ConflictRefs& R, bool withHead) {
        P.nlev = (int32_t)ord.size();
        P.seedRow0 = 0;
        const int64_t T = g.T();
        for (size_t L = 0; L < ord.size(); ++L) {
            const AtomIR& a = r.body[ord[L]];
            Level lv{};
            const int32_t p = g.dense_pred(a.pterm);
            const int32_t nbound = nslots;
            bool sB = bound(a.s), oB = bound(a.o);
            int32_t kind, dir;
            if (sB && oB) {
                kind = kCheck; dir = kSPO; lv.key = ref(a.s); lv.val = ref(a.o);
            } else if (sB) {
                kind = kExt; dir = kSPO; lv.key = ref(a.s); lv.val = bind(a.o.v);
            } else if (oB) {
                kind = kExt; dir = kPOS; lv.key = ref(a.o); lv.val = bind(a.s.v);
            } else {
                kind = kScan; dir = kSPO; lv.key = 0;
                lv.val = bind(a.s.v);
                lv.val2 = bind(a.o.v);   // same slot when the atom is (?x p ?x)
            }
            if (kind == kScan) {
                lv.lo = g.predOff[p];
                lv.hi = g.predOff[p + 1];
            } else {
                const int64_t base = ((int64_t)dir * g.P() + p) * g.nBlocks();
                lv.lo = (int32_t)(uint32_t)(base & 0xffffffffLL);
                lv.hi = (int32_t)(base >> 32);
            }
            (void)T;
            R.sref[L] = ref(a.s);
            R.oref[L] = ref(a.o);
            int32_t m = 0;
            for (size_t j = 0; j < L; ++j) {
                const AtomIR& b = r.body[ord[j]];
                if (b.pterm == a.pterm && atom_eq(a, b) != kNever) m |= 1 << j;
            }
            if (withHead && r.head.pterm == a.pterm && atom_eq(a, r.head) != kNever) m |= 1 << 8;
            lv.flags = make_level_flags(kind, dir, nbound, m);
            P.lev[L] = lv;
        }
        fuse(P, ord);
    }

    // An EXT level binding x followed by a CHECK of (x, y) or (y, x) with y
    // bound before the EXT becomes an intersection: x runs over the EXT row
    // and must also be in y's row of the CHECK predicate -- (y p x): y's
    // objects (SPO), (x p y): y's subjects (POS). The CHECK stays (same
    // conflict mask), keyed by y; its frame is the intersection's match, so
    // it costs no lookup or search of its own.
    void fuse(Plan& P, const std::vector<int>& ord) const {
        for (int L = 0; L + 1 < P.nlev; ++L) {
            Level& a = P.lev[L];
            Level& m = P.lev[L + 1];
            if (level_kind(a.flags) != kExt || level_kind(m.flags) != kCheck) continue;
            const AtomIR& am = r.body[ord[L + 1]];
            const int32_t x = a.val;
            const int32_t nb = level_nbound(a.flags);
            const int32_t rs = ref(am.s), ro = ref(am.o);
            auto before = [&](int32_t t) { return t < 0 || t < nb; };
            int32_t y, dir2;
            if (ro == x && rs != x && before(rs)) {
                y = rs;
                dir2 = kSPO;
            } else if (rs == x && ro != x && before(ro)) {
                y = ro;
                dir2 = kPOS;
            } else {
                continue;
            }
            const int32_t p2 = g.dense_pred(am.pterm);
            const int64_t base2 = ((int64_t)dir2 * g.P() + p2) * g.nBlocks();
            a.flags = make_level_flags(kIsect, level_dir(a.flags), nb, level_conflicts(a.flags));
            a.val2 = y;
            a.lo2 = (int32_t)(uint32_t)(base2 & 0xffffffffLL);
            a.hi2 = (int32_t)(base2 >> 32);
            m.flags = make_level_flags(kCheck, dir2, level_nbound(m.flags), level_conflicts(m.flags));
            m.key = y;
            m.val = x;
            ++L;   // the CHECK is taken
        }
    }
};

bool const_ok(const TermIR& t, const HostGraph& g) { return t.var || (t.v >= 0 && t.v < g.nE); }

// Moves the first non-CHECK atom of `ord` to the front (CHECK atoms bind
// nothing, so slots and kinds are unchanged). Returns false if every level
// is a CHECK.
bool wave_order(const Plan& thread, std::vector<int>& ord) {
    int f = -1;
    for (int L = 0; L < thread.nlev; ++L)
        if (level_kind(thread.lev[L].flags) != kCheck) { f = L; break; }
    if (f < 0) return false;
    int a = ord[f];
    ord.erase(ord.begin() + f);
    ord.insert(ord.begin(), a);
    return true;
}

// Rank space of a PAIRS plan's head objects: the values its pair level
// enumerates are the keys of one (direction, predicate) -- objects (POS keys)
// for an SPO row, subjects (SPO keys) for a POS row; for a SCAN, the side the
// head object takes. Rows are numbered direction-major, then by predicate,
// then by key (as the device key index is built), so the first key's row
// number is the number of keys before it.
void head_object_ranks(Plan& Q, const HostGraph& g) {
    const Level& lv = Q.lev[Q.pairLevel];
    const int32_t kd = level_kind(lv.flags);
    const int64_t nB = g.nBlocks();
    int32_t p, dir2;
    if (kd == kScan) {
        p = (int32_t)(std::upper_bound(g.predOff.begin(), g.predOff.end(), lv.lo) - g.predOff.begin()) - 1;
        dir2 = lv.val == Q.hoRef ? kSPO : kPOS;   // the scan binds subject (val) and object (val2)
    } else {
        const int64_t base = (int64_t)(uint32_t)lv.lo | ((int64_t)lv.hi << 32);
        const int32_t dir = (int32_t)(base / ((int64_t)g.P() * nB));
        p = (int32_t)((base / nB) % g.P());
        dir2 = dir == kSPO ? kPOS : kSPO;
    }
    const int64_t base2 = ((int64_t)dir2 * g.P() + p) * nB;
    Q.hoLo = (int32_t)(uint32_t)(base2 & 0xffffffffLL);
    Q.hoHi = (int32_t)(base2 >> 32);
    Q.hoRow0 = dir2 == kSPO ? g.keysSOff[p] : (int32_t)g.keysS.size() + g.keysOOff[p];
    Q.hoKeys = g.nkeys(dir2, p);
    Q.hoByEntry = kd == kScan ? 0 : 1;
}

}  // namespace

CompiledRule compile_rule(const RuleIR& r, const HostGraph& g, const Tuning& t) {
    CompiledRule c;
    for (Plan* Q : {&c.sup, &c.supWave, &c.pairs, &c.pairsWave}) Q->supPred = -1;
    const int nb = (int)r.body.size();
    const int32_t hp = g.dense_pred(r.head.pterm);

    // head size and constant-aware head support
    // (Rewritten_CPP/SupportCounting.cpp:142-201, CPU_01_nonparallel/SupportCounting.cpp:474-507)
    if (hp >= 0) {
        c.headSize = g.nnz(hp);
        const TermIR &hs = r.head.s, &ho = r.head.o;
        if (hs.var && ho.var) {
            c.headSupport = c.headSize;
        } else if (!hs.var && ho.var) {
            if (const_ok(hs, g)) { auto rr = g.row(0, hp, hs.v); c.headSupport = rr.second - rr.first; }
        } else if (hs.var && !ho.var) {
            if (const_ok(ho, g)) { auto rr = g.row(1, hp, ho.v); c.headSupport = rr.second - rr.first; }
        } else {
            c.headSupport = const_ok(hs, g) && const_ok(ho, g) && g.has(hp, hs.v, ho.v) ? 1 : 0;
        }
    }
    // unsupported shapes: reported with head figures only (GPU_03/gpu_rules_lib.cu:352-357)
    if (nb < 1 || nb > kMaxBody || r.nvars > kMaxSlots) {
        c.status = kRuleSkipped;
        return c;
    }

    bool bodyMatchable = true;
    bool repeatedBody = false;
    for (int i = 0; i < nb; ++i) {
        const AtomIR& a = r.body[i];
        if (g.dense_pred(a.pterm) < 0 || !const_ok(a.s, g) || !const_ok(a.o, g)) bodyMatchable = false;
        for (int j = 0; j < i; ++j)
            if (r.body[j].pterm == a.pterm && atom_eq(a, r.body[j]) == kAlways) repeatedBody = true;
    }
    bool bodyIsHead = false;
    for (const AtomIR& a : r.body)
        if (a.pterm == r.head.pterm && atom_eq(a, r.head) == kAlways) bodyIsHead = true;

    // ---------------------------------------------------------- support
    c.supportLaunch = bodyMatchable && !repeatedBody && !bodyIsHead && hp >= 0 && c.headSupport > 0;
    if (c.supportLaunch) {
        Builder B(r, g);
        Plan& P = c.sup;
        P.kind = kSupport;
        P.budget = t.budget;
        const TermIR &hs = r.head.s, &ho = r.head.o;
        if (hs.var) B.bind(hs.v);
        if (ho.var) B.bind(ho.v);
        P.hsRef = B.ref(hs);
        P.hoRef = B.ref(ho);
        if (!hs.var) {
            auto rr = g.row(0, hp, hs.v);
            P.seedDir = kSPO; P.seedBegin = rr.first; P.seedEnd = rr.second;
        } else if (!ho.var) {
            auto rr = g.row(1, hp, ho.v);
            P.seedDir = kPOS; P.seedBegin = (int32_t)(g.T() + rr.first); P.seedEnd = (int32_t)(g.T() + rr.second);
        } else {
            P.seedDir = kSPO; P.seedBegin = g.predOff[hp]; P.seedEnd = g.predOff[hp + 1];
        }
        P.pairLevel = -1;
        Builder B0 = B;
        std::vector<int> ord = B.order();
        B.levels(ord, P, c.supRefs, true);
        c.supRefs.hs = P.hsRef;
        c.supRefs.ho = P.hoRef;
        c.supItems = P.seedEnd - P.seedBegin;
        c.supWave = P;
        c.supWaveRefs = c.supRefs;
        if (wave_order(P, ord)) {
            B0.levels(ord, c.supWave, c.supWaveRefs, true);
            c.supWaveOk = true;
        }
    }

    // ------------------------------------------------------- body pairs
    bool headVars = r.head.s.var && r.head.o.var;
    bool sIn = false, oIn = false;
    for (const AtomIR& a : r.body)
        for (const TermIR* x : {&a.s, &a.o}) {
            if (x->var && headVars && x->v == r.head.s.v) sIn = true;
            if (x->var && headVars && x->v == r.head.o.v) oIn = true;
        }
    c.pairsLaunch = bodyMatchable && !repeatedBody && headVars && sIn && oIn;
    if (c.pairsLaunch) {
        // seed keys: the head-subject side of the body atom with the fewest
        // distinct keys there (CPU_01_nonparallel/SupportCounting.cpp:577-605
        // seeds from the smallest such atom)
        int32_t bestDir = -1, bestP = -1, bestN = std::numeric_limits<int32_t>::max();
        for (const AtomIR& a : r.body) {
            int32_t p = g.dense_pred(a.pterm);
            if (a.s.var && a.s.v == r.head.s.v && g.nkeys(0, p) < bestN) { bestDir = 0; bestP = p; bestN = g.nkeys(0, p); }
            if (a.o.var && a.o.v == r.head.s.v && g.nkeys(1, p) < bestN) { bestDir = 1; bestP = p; bestN = g.nkeys(1, p); }
        }
        Builder B(r, g);
        Plan& P = c.pairs;
        P.kind = kPairs;
        P.budget = t.budget;
        B.bind(r.head.s.v);
        P.hsRef = 0;
        P.seedDir = bestDir;
        const int32_t kbase = bestDir == 0 ? 0 : (int32_t)g.keysS.size();
        const std::vector<int32_t>& off = bestDir == 0 ? g.keysSOff : g.keysOOff;
        P.seedBegin = kbase + off[bestP];
        P.seedEnd = kbase + off[bestP + 1];
        Builder B0 = B;
        std::vector<int> ord = B.order();
        B.levels(ord, P, c.pairsRefs, false);
        auto finish = [&](Plan& Q, const Builder& b) {
            Q.hoRef = b.slotOf[r.head.o.v];
            Q.pairLevel = -1;
            if (r.head.o.v != r.head.s.v)
                for (int L = 0; L < Q.nlev; ++L) {
                    const Level& lv = Q.lev[L];
                    const int32_t kd = level_kind(lv.flags);
                    if (((kd == kExt || kd == kIsect) && lv.val == Q.hoRef) ||
                        (kd == kScan && (lv.val == Q.hoRef || lv.val2 == Q.hoRef))) {
                        Q.pairLevel = L;
                        break;
                    }
                }
            if (Q.pairLevel >= 0) head_object_ranks(Q, g);
            const Level& l0 = Q.lev[0];
            const int64_t b0 = (int64_t)(uint32_t)l0.lo | ((int64_t)l0.hi << 32);
            Q.seedRow0 = Q.nlev > 0 && level_kind(l0.flags) == kExt && l0.key == 0 &&
                         level_dir(l0.flags) == bestDir && b0 == ((int64_t)bestDir * g.P() + bestP) * g.nBlocks();
        };
        finish(P, B);
        c.pairItems = P.seedEnd - P.seedBegin;
        c.pairsWave = P;
        c.pairsWaveRefs = c.pairsRefs;
        if (wave_order(P, ord)) {
            B0.levels(ord, c.pairsWave, c.pairsWaveRefs, false);
            finish(c.pairsWave, B0);
            c.pairsWaveOk = true;
        }
    }

    // -------------------------------------- support from the body pairs
    // A head triple (s, o) is supported iff the body has an injective
    // grounding with those values that does not use the triple itself; a
    // distinct body pair (s, o) has one, and it cannot use that triple when
    // no body atom can ground to it (atom_eq == kNever against the head, the
    // test that sets the support plan's head-conflict bit). So the support
    // is the number of distinct body pairs (s, o) with (s, p_head, o) a
    // triple, which the pair plans can count at the end of each seed.
    if (c.supportLaunch && c.pairsLaunch && r.head.s.v != r.head.o.v && c.pairs.pairLevel >= 0 &&
        c.pairsWave.pairLevel >= 0) {
        bool headFree = true;
        for (const AtomIR& a : r.body)
            if (a.pterm == r.head.pterm && atom_eq(a, r.head) != kNever) headFree = false;
        if (headFree) c.supPred = hp;   // the batch sets Plan::supPred when it fuses
    }
    return c;
}

RowStats row_stats(const HostGraph& g) {
    RowStats st;
    st.P = g.P();
    st.keys.assign(2 * (size_t)st.P, 0.0);
    st.nnz.assign(2 * (size_t)st.P, 0.0);
    st.sq.assign(2 * (size_t)st.P, 0.0);
    for (int dir = 0; dir < 2; ++dir) {
        const std::vector<int32_t>& key = dir == 0 ? g.spoKey : g.posKey;
        for (int32_t p = 0; p < st.P; ++p) {
            const size_t at = (size_t)dir * st.P + p;
            int32_t i = g.predOff[p];
            const int32_t end = g.predOff[p + 1];
            while (i < end) {
                int32_t j = i + 1;
                while (j < end && key[j] == key[i]) ++j;
                const double len = j - i;
                st.keys[at] += 1;
                st.nnz[at] += len;
                st.sq[at] += len * len;
                i = j;
            }
        }
    }
    return st;
}

namespace {

// (direction * P + predicate) of a level's key-index block run.
int64_t level_row_set(int32_t lo, int32_t hi, const HostGraph& g) {
    const int64_t base = (int64_t)(uint32_t)lo | ((int64_t)hi << 32);
    return base / std::max<int64_t>(1, g.nBlocks());
}

double plan_work(const Plan& P, int64_t items, const HostGraph& g, const RowStats& st) {
    if (items <= 0) return 0;
    const int64_t nsets = 2 * (int64_t)st.P;
    auto mean = [&](int64_t set, bool reached) {
        if (set < 0 || set >= nsets || st.keys[set] <= 0) return 0.0;
        const double m = st.nnz[set] / st.keys[set];
        return reached ? std::max(m, st.sq[set] / st.nnz[set]) : m;
    };
    double nodes = (double)items, work = 0;
    for (int L = 0; L < P.nlev; ++L) {
        const Level& lv = P.lev[L];
        const int kind = level_kind(lv.flags);
        double f = 1;
        if (kind == kScan) {
            f = (double)(lv.hi - lv.lo);
        } else if (kind == kExt || kind == kIsect) {
            const int64_t set = level_row_set(lv.lo, lv.hi, g);
            if (lv.key < 0) {   // constant key: its row
                const int32_t p = (int32_t)(set % std::max(1, st.P));
                const auto rr = g.row((int)(set / std::max(1, st.P)), p, -(lv.key + 1));
                f = rr.second - rr.first;
            } else {
                f = mean(set, L > 0);
            }
            if (kind == kIsect) f = std::min(f, mean(level_row_set(lv.lo2, lv.hi2, g), true));
            if (L == 0 && set >= 0 && set < nsets) {   // the seeds' rows hold at most the predicate's entries
                nodes = std::min(nodes * f, st.nnz[set]);
                work += nodes;
                continue;
            }
        }
        nodes *= f;
        work += nodes;
        if (work > 1e30) return 1e30;
    }
    return work;
}

}  // namespace

double estimate_work(const CompiledRule& c, const HostGraph& g, const RowStats& st) {
    if (c.status != kRuleOk) return 0;
    double w = 0;
    if (c.supportLaunch) w += plan_work(c.sup, c.supItems, g, st);
    if (c.pairsLaunch) w += plan_work(c.pairs, c.pairItems, g, st);
    return w;
}

}  // namespace gr
