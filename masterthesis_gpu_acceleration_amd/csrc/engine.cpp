// engine.cpp -- contexts, device residency and the C ABI of libgpu_rules.so.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <atomic>
#include <chrono>
#include <cmath>
#include <climits>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <exception>
#include <functional>
#include <memory>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include "engine_internal.hpp"
#include "graph_device.hpp"
#include "noalloc.hpp"
#include "refine.hpp"
#include "snapshot.hpp"
#include "traffic.hpp"

namespace gr {

bool Context::upload() {
    GR_HIP(hipGetDevice(&device));
    hipDeviceProp_t prop;
    GR_HIP(hipGetDeviceProperties(&prop, device));
    numCUs = prop.multiProcessorCount > 0 ? prop.multiProcessorCount : 256;
    if (prop.sharedMemPerBlock > 0) ldsPerBlock = prop.sharedMemPerBlock;
    ldsPerCU = prop.maxSharedMemoryPerMultiProcessor > 0 ? prop.maxSharedMemoryPerMultiProcessor : ldsPerBlock;
    GR_HIP(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
    GR_HIP(hipStreamCreateWithFlags(&st2, hipStreamNonBlocking));
    GR_HIP(hipEventCreateWithFlags(&forkEv, hipEventDisableTiming));
    GR_HIP(hipEventCreateWithFlags(&joinEv, hipEventDisableTiming));
    for (auto& e : ev) GR_HIP(hipEventCreate(&e));
    GR_HIP(hipHostMalloc((void**)&hOne, 512, hipHostMallocDefault));
    // batch buffers grow stream-ordered from the device's default pool: let the
    // pool keep what they free, so regrowing reuses it instead of mapping anew
    hipMemPool_t pool = nullptr;
    if (hipDeviceGetDefaultMemPool(&pool, device) == hipSuccess && pool) {
        uint64_t keep = UINT64_MAX;
        (void)hipMemPoolSetAttribute(pool, hipMemPoolAttrReleaseThreshold, &keep);
    }
    // the allocation-free entry point's scratch on this device (reported there if unavailable)
    noallocReady = noalloc_init(device);

    const int32_t P = g.P();
    const int64_t T = g.T();
    // key index: per (dir, pred) one block per 64 term ids; rows numbered in
    // (dir, pred, key) order, rowStart[r] = first merged entry of row r
    const int64_t nB = g.nBlocks();
    std::vector<KeyBlock> hb((size_t)2 * P * nB, KeyBlock{0, 0, 0});
    std::vector<int32_t> rs;
    rs.reserve(g.keysS.size() + g.keysO.size() + 1);
    uint32_t row = 0;
    for (int d = 0; d < 2; ++d) {
        const std::vector<int32_t>& key = d == 0 ? g.spoKey : g.posKey;
        for (int32_t p = 0; p < P; ++p) {
            KeyBlock* base = hb.data() + ((size_t)d * P + p) * nB;
            int64_t nextBlock = 0;   // blocks before this one already have their rank
            for (int32_t e = g.predOff[p]; e < g.predOff[p + 1]; ++e) {
                if (e != g.predOff[p] && key[e] == key[e - 1]) continue;
                const int64_t blk = key[e] >> 6;
                for (; nextBlock <= blk; ++nextBlock) base[nextBlock].rank = row;
                base[blk].bits |= 1ull << (key[e] & 63);
                rs.push_back((int32_t)(d * T + e));
                ++row;
            }
            for (; nextBlock < nB; ++nextBlock) base[nextBlock].rank = row;
        }
    }
    rs.push_back((int32_t)(2 * T));
    if (!put(blocks, hb, st) || !put(rowStart, rs, st)) return false;
    GR_HIP(hipStreamSynchronize(st));
    std::vector<KeyBlock>().swap(hb);
    std::vector<int32_t> m;
    m.reserve((size_t)2 * T);
    m.insert(m.end(), g.spoVal.begin(), g.spoVal.end());
    m.insert(m.end(), g.posVal.begin(), g.posVal.end());
    if (!put(col, m, st)) return false;
    GR_HIP(hipStreamSynchronize(st));
    m.clear();
    m.insert(m.end(), g.spoKey.begin(), g.spoKey.end());
    m.insert(m.end(), g.posKey.begin(), g.posKey.end());
    if (!put(coo, m, st)) return false;
    GR_HIP(hipStreamSynchronize(st));
    m.clear();
    m.insert(m.end(), g.keysS.begin(), g.keysS.end());
    m.insert(m.end(), g.keysO.begin(), g.keysO.end());
    if (!put(keys, m, st)) return false;
    GR_HIP(hipStreamSynchronize(st));
    // crank: each entry's value ranked among the keys of the other direction
    // of its predicate (every value of (d, p) is a key of (1-d, p))
    m.assign((size_t)2 * T, 0);
    {
        std::vector<int32_t> pos((size_t)std::max<int64_t>(g.nE, 1), 0);
        for (int d = 0; d < 2; ++d) {
            const std::vector<int32_t>& val = d == 0 ? g.spoVal : g.posVal;
            const std::vector<int32_t>& ok = d == 0 ? g.keysO : g.keysS;
            const std::vector<int32_t>& oo = d == 0 ? g.keysOOff : g.keysSOff;
            for (int32_t p = 0; p < P; ++p) {
                for (int32_t r = oo[p]; r < oo[p + 1]; ++r) pos[ok[r]] = r - oo[p];
                for (int32_t e = g.predOff[p]; e < g.predOff[p + 1]; ++e) m[(size_t)d * T + e] = pos[val[e]];
            }
        }
    }
    if (!put(crank, m, st)) return false;
    GR_HIP(hipStreamSynchronize(st));
    dg.blocks = blocks.as<KeyBlock>();
    dg.rowStart = rowStart.as<int32_t>();
    dg.col = col.as<int32_t>();
    dg.coo = coo.as<int32_t>();
    dg.keys = keys.as<int32_t>();
    dg.crank = crank.as<int32_t>();
    dg.nBlocks = g.nBlocks();
    if (!one.ensure(kOneBytes)) return false;
    GR_HIP(hipMemset(one.p, 0, kOneBytes));
    return ensure_dedup();
}

int Context::waveBlocksFor(int slots) const {
    if (pairsDedup() != kDedupLds) return waveBlocks();
    const size_t blockLds = (wave_static_lds(slots, waveEngine) + (size_t)rowWords * 4 + 511) / 512 * 512;
    const int perCU = std::max<int>(1, (int)(ldsPerCU / blockLds));
    return std::max(1, numCUs * std::min(waveWavesPerCU, perCU));
}

bool Context::ensure_dedup() {
    // wave-mode bitmaps index head objects by rank among the keys of one
    // (direction, predicate): as many bits as the largest such key set
    int64_t maxKeys = 1;
    for (int32_t p = 0; p < g.P(); ++p) maxKeys = std::max<int64_t>(maxKeys, std::max(g.nkeys(0, p), g.nkeys(1, p)));
    rowWords = (maxKeys + 31) / 32;
    const int64_t rowBytes = rowWords * 4;
    rowsFit = rowBytes * numCUs <= kRowBudget;
    const size_t blockLds = (wave_static_lds(16, waveEngine) + (size_t)rowBytes + 511) / 512 * 512;   // LDS granules
    ldsFits = blockLds <= ldsPerBlock;
    ldsPreferred = ldsFits && ldsPerCU / blockLds >= (size_t)kLdsMinWaves;
    waveSlots = numCUs * waveWavesPerCU;
    const int mode = pairsDedup();
    if (mode == kDedupLds) {
        prepare_wave_kernels((size_t)rowBytes);
        const int perCU = std::max<int>(1, (int)(ldsPerCU / blockLds));   // resident bitmaps per CU
        waveSlots = numCUs * std::min(waveWavesPerCU, perCU);
    }
    // LDS bitmaps and the sort path never touch the HBM rows; one keeps the kernels' pointer valid
    size_t bytes = (size_t)rowBytes;
    if (mode == kDedupRows) {
        // keep the dedup rows within kRowBudget
        const int64_t maxSlots = std::max<int64_t>(4, kRowBudget / rowBytes);
        if (waveSlots > maxSlots) waveSlots = (int)(maxSlots / 4 * 4);
        bytes = (size_t)waveSlots * rowBytes;
    }
    if (dedup.bytes < bytes) {
        if (!dedup.ensure(bytes)) return false;
        GR_HIP(hipMemset(dedup.p, 0, dedup.bytes));
    }
    return giants.ensure(rowWords) && giants2.ensure(rowWords);
}

void giant_prepare(GiantBufs& gb, int after, GiantArgs& ga, hipStream_t s, int phase) {
    ZeroRanges z{};
    giant_args(gb, after, ga, phase, z);
    launch_zero(z, s);
}

void giant_args(GiantBufs& gb, int after, GiantArgs& ga, int phase, ZeroRanges& z) {
    uint32_t* ctl = gb.ctl.as<uint32_t>();
    ga.stats = gb.stats.as<unsigned long long>() + phase * kWaveStats;
    zero_add(z, ga.stats, 2 * kWaveStats);
    zero_add(z, ctl, 4);
    zero_add(z, gb.remaining.as<uint32_t>(), gb.cap);
    zero_add(z, gb.found.as<uint32_t>(), gb.cap);
    ga.after = after;
    ga.cap = gb.cap;
    ga.list = gb.list.as<HeavyItem>();
    ga.heavyIdx = gb.heavyIdx.as<uint32_t>();
    ga.count = ctl + 0;
    ga.subs = gb.subs.as<SubItem>();
    ga.subCount = ctl + 1;
    ga.subCursor = ctl + 2;
    ga.remaining = gb.remaining.as<uint32_t>();
    ga.found = gb.found.as<uint32_t>();
    ga.rows = gb.rows.as<uint32_t>();
}

// --------------------------------------------------------- per-rule path

bool has_conflicts(const Plan& P) {
    for (int L = 0; L < P.nlev; ++L)
        if (level_conflicts(P.lev[L].flags)) return true;
    return false;
}

// Device scratch of the per-rule launches: a header -- [0] count (u64);
// [8] the giant sub-items' count (u64); +64: u32 heavy count, wave cursor;
// +96: u64 keys emitted (sort path); +128: ConflictRefs of both plans (to
// +288); +kOnePlanOff: the wave-mode plan for the giant kernels -- and the
// heavy-seed queue. The context's own (C.one, kOneBytes, and C.heavy1,
// grown as rules need) or, for the allocation-free entry point, the fixed
// device scratch array (a kNoAllocHeader-byte header), where nothing may
// grow (mayAlloc false) and giant seeds are not cut.
struct OneScratch {
    char* hdr;
    HeavyItem* heavy;
    uint64_t heavyCap;
    bool mayAlloc;
};

OneScratch context_scratch(Context& C) { return OneScratch{C.one.as<char>(), nullptr, 0, true}; }

// One plan: thread-mode launch over all seeds, wave-mode launch over the
// escalated ones, sync, copy the count back.
// With the sort-unique dedup path, a body-pair plan's wave-mode launch
// emits keys and the sort-unique pass counts them; a pass that emitted more
// keys than the buffer holds is redone with a larger one (when S may grow
// buffers; else it fails).
bool run_one(Context& C, OneScratch S, const Plan& P, const ConflictRefs& R, const Plan& PW, const ConflictRefs& RW,
             int64_t items, bool cutGiants, uint64_t& count, double& ms, int64_t& heavySeeds) {
    count = 0;
    ms = 0;
    if (items <= 0) return true;
    if ((uint64_t)items > kMaxPhaseSeeds) {
        std::fprintf(stderr, "[gpu_rules] a rule with %lld seeds exceeds the heavy queue's 32-bit range\n",
                     (long long)items);
        return false;
    }
    if (S.mayAlloc) {
        if ((uint64_t)items > C.heavy1Cap) {
            if (!C.heavy1.ensure((size_t)items * sizeof(HeavyItem))) return false;
            C.heavy1Cap = (uint32_t)items;
        }
        S.heavy = C.heavy1.as<HeavyItem>();
        S.heavyCap = C.heavy1Cap;
    } else if ((uint64_t)items > S.heavyCap) {
        std::fprintf(stderr, "[gpu_rules] allocation-free evaluation: a plan with %lld seeds exceeds the scratch "
                             "queue's %llu\n", (long long)items, (unsigned long long)S.heavyCap);
        return false;
    }
    const uint32_t heavyCap = (uint32_t)S.heavyCap;
    unsigned long long* d_count = reinterpret_cast<unsigned long long*>(S.hdr);
    uint32_t* ctl = reinterpret_cast<uint32_t*>(S.hdr + 64);
    unsigned long long* d_emitted = reinterpret_cast<unsigned long long*>(S.hdr + 96);
    ConflictRefs* d_refs = reinterpret_cast<ConflictRefs*>(S.hdr + 128);
    const bool conf = has_conflicts(P) || has_conflicts(PW);
    if (conf) {
        ConflictRefs both[2] = {R, RW};
        GR_HIP(hipMemcpyAsync(d_refs, both, sizeof(both), hipMemcpyHostToDevice, C.st));
    }
    const int mode = PW.kind == kPairs ? C.pairsDedup() : kDedupRows;
    const bool sort = mode == kDedupSort;
    const int slots = std::max(plan_slots(P, R), plan_slots(PW, RW));
    if (sort && !S.mayAlloc && C.sortb.cap == 0) {
        std::fprintf(stderr, "[gpu_rules] allocation-free evaluation: sort-unique dedup needs key buffers, which "
                             "this context has not allocated (set \"sort_cap\" first)\n");
        return false;
    }
    if (sort && S.mayAlloc && !C.sortb.ensure(1, kSortCapMin, C.sortCapMax)) return false;
    const int hoBits = bits_for((uint64_t)C.g.nE);
    const int hBits = bits_for((uint64_t)items);
    // Seeds still running after giantAfter wave-mode steps are cut into
    // sub-items that all waves share, as in batch passes (the wave kernel
    // stores its plan, as rule 0, where the giant kernels read it), for
    // rules the caller marks (cutGiants). The allocation-free path runs every
    // seed to its end on one wave.
    const bool giants = cutGiants && S.mayAlloc && C.giantAfter > 0;
    Plan* d_plan = reinterpret_cast<Plan*>(S.hdr + kOnePlanOff);
    Plan pw = PW;
    pw.rule = 0;
    for (;;) {
        GiantArgs ga{};
        if (giants) {
            ZeroRanges z{};
            zero_add(z, d_count, 4);   // the count and the sub-items' count
            zero_add(z, ctl, 16);      // control words and the emitted-key counter
            giant_args(C.giants, C.giantAfter, ga, 0, z);
            launch_zero(z, C.st);
        } else {
            // both count words: the count is read as their sum, and a giant pass of an
            // earlier rule leaves its sub-items' count in the second
            launch_fill_u64(d_count, 2, 0, C.st);
            launch_fill_u32(ctl, 16, 0, C.st);   // control words and the emitted-key counter
        }
        GR_HIP(hipEventRecord(C.ev[0], C.st));
        const int64_t waves = (items + kWave - 1) / kWave;
        const int blocks = (int)std::min<int64_t>((waves + 3) / 4, C.threadBlocks());
        OneArgs oa{conf ? d_refs : nullptr, (uint32_t)items, heavyCap, d_count, S.heavy, ctl + 0};
        launch_thread_one(P.kind, slots, P, oa, C.dg, blocks, C.st);
        SortSink sink{};
        if (sort) sink = SortSink{C.sortb.keys[0].as<unsigned long long>(), d_emitted, C.sortb.cap, hoBits};
        WaveArgs wa{giants ? d_plan : nullptr, conf ? d_refs + 1 : nullptr, S.heavy, ctl + 0, heavyCap, 0, ctl + 1,
                    d_count, C.dedup.as<uint32_t>(), C.rowWords, sink};
        wa.giant = ga;
        wa.planOut = giants ? d_plan : nullptr;
        wa.engine = C.waveEngine;
        launch_wave_one(PW.kind, slots, mode, pw, wa, C.dg, C.waveBlocksFor(slots), C.st);
        if (giants) launch_giants(PW.kind, slots, mode, wa, C.dg, C.waveBlocksFor(slots), C.st);
        if (sort) {
            SortJob j{{C.sortb.keys[0].as<unsigned long long>(), C.sortb.keys[1].as<unsigned long long>()},
                      d_emitted, C.sortb.cap, sort_max_tiles(C.sortb.cap), C.sortb.hist.as<uint32_t>(),
                      hoBits + hBits, hoBits, nullptr, nullptr, d_count, C.sortBlocks()};
            launch_sort_unique(j, C.st);
        }
        GR_HIP(hipGetLastError());
        GR_HIP(hipEventRecord(C.ev[1], C.st));
        GR_HIP(hipMemcpyAsync(C.hOne, S.hdr, 128, hipMemcpyDeviceToHost, C.st));
        GR_HIP(hipStreamSynchronize(C.st));
        const uint32_t* hctl = reinterpret_cast<uint32_t*>(reinterpret_cast<char*>(C.hOne) + 64);
        const uint64_t emitted = *reinterpret_cast<const unsigned long long*>(reinterpret_cast<char*>(C.hOne) + 96);
        if (heavy_count_action(hctl[0], heavyCap) != kCountsOk) {
            std::fprintf(stderr, "[gpu_rules] %u seeds escalated, beyond the heavy queue's %u\n", hctl[0], heavyCap);
            return false;
        }
        if (sort) {
            C.lastEmitted = (int64_t)emitted;
            const int32_t act = sort_count_action(emitted, C.sortb.cap, C.sortCapMax);
            if (act == kCountsFail) {
                std::fprintf(stderr, "[gpu_rules] body-pair keys of one rule (%llu) exceed the sort capacity %u\n",
                             (unsigned long long)emitted, C.sortb.cap);
                return false;
            }
            if (act == kCountsRegrow) {
                if (!S.mayAlloc) {
                    std::fprintf(stderr, "[gpu_rules] allocation-free evaluation: body-pair keys of one rule "
                                         "(%llu) exceed the sort buffers' %u\n", (unsigned long long)emitted,
                                 C.sortb.cap);
                    return false;
                }
                if (!C.sortb.ensure(emitted, kSortCapMin, C.sortCapMax)) return false;
                continue;
            }
        }
        float f = 0;
        GR_HIP(hipEventElapsedTime(&f, C.ev[0], C.ev[1]));
        ms = f;
        count = C.hOne[0] + C.hOne[1];
        heavySeeds += hctl[0];
        return true;
    }
}

uint32_t support_threshold(int64_t headSize, const Thresholds& t) {
    int64_t thr = std::max<int64_t>(1, t.minSupport);
    if (t.minHeadCoverage > 0.0) {
        if (headSize <= 0) return kGateMax;   // coverage 0: no support meets a positive minimum
        // the smallest s with s / headSize >= minHeadCoverage in the same
        // double division as the reported headCoverage
        const double h = (double)headSize;
        int64_t s = (int64_t)std::ceil(t.minHeadCoverage * h);
        if (s > headSize + 1) s = headSize + 1;   // more than every head triple: never met
        while (s > 0 && (double)(s - 1) / h >= t.minHeadCoverage) --s;
        while (s <= headSize && (double)s / h < t.minHeadCoverage) ++s;
        thr = std::max(thr, s);
    }
    return (uint32_t)std::min<int64_t>(thr, kGateMax);
}

void finish_measures(GpuRuleResult64& r) {
    r.headCoverage = r.headSize > 0 ? (double)r.support / (double)r.headSize : 0.0;
    r.confidence = r.bodySize > 0 ? (double)r.support / (double)r.bodySize : 0.0;
}

bool eval_rule_plain(Context& C, OneScratch S, const RuleIR& r, bool gated, const Thresholds& thr,
                     GpuRuleResult64& res, double& supMs, double& confMs) {
    std::memset(&res, 0, sizeof(res));
    res.bodyPairs = -1;
    supMs = confMs = 0;
    CompiledRule cr = compile_rule(r, C.g, C.tune);
    res.headSize = cr.headSize;
    res.headSupport = cr.headSupport;
    res.status = cr.status;
    if (cr.status != kRuleOk) return true;
    int64_t heavy = 0;
    uint64_t sup = 0;
    // The giant pass costs two launches per phase; a rule whose estimated
    // work is below perRuleGiantWork has no seed long enough to pay for them.
    const bool cut = C.perRuleGiantWork <= 0 || estimate_work(cr, C.g, C.rowStats) >= C.perRuleGiantWork;
    if (cr.supportLaunch &&
        !run_one(C, S, cr.sup, cr.supRefs, cr.supWave, cr.supWaveRefs, cr.supItems, cut, sup, supMs, heavy))
        return false;
    res.support = (int64_t)sup;
    if (!gated || sup >= support_threshold(cr.headSize, thr)) {
        uint64_t pairs = 0;
        if (cr.pairsLaunch &&
            !run_one(C, S, cr.pairs, cr.pairsRefs, cr.pairsWave, cr.pairsWaveRefs, cr.pairItems, cut, pairs, confMs,
                     heavy))
            return false;
        res.bodyPairs = (int64_t)pairs;
    }
    res.bodySize = res.support > 0 && res.bodyPairs > 0 ? res.bodyPairs : 0;
    finish_measures(res);
    C.lastHeavy = heavy;
    return true;
}

// ------------------------------------------------------------ batch path

// Body pairs depend only on a rule's body and on which of its variables are
// the head subject and object, not on the head predicate. Rules are keyed by
// their body with the head subject named 0, the head object 1 and the other
// variables numbered by first appearance, atoms sorted (equal keys: the
// bodies are equal up to a renaming that fixes the head variables, so their
// injective groundings -- and distinct pairs -- correspond one to one). The
// first rule of a key runs the search; from[i] is that rule for the others,
// -1 for rules with their own search (or none).
std::vector<int32_t> shared_pairs(const std::vector<RuleIR>& rules, const std::vector<CompiledRule>& comp,
                                  const std::vector<int32_t>& status) {
    struct KeyHash {
        size_t operator()(const std::vector<int64_t>& k) const {
            uint64_t h = 1469598103934665603ull;
            for (int64_t x : k) h = (h ^ (uint64_t)x) * 1099511628211ull;
            return (size_t)h;
        }
    };
    std::vector<int32_t> from(rules.size(), -1);
    std::unordered_map<std::vector<int64_t>, int32_t, KeyHash> first;
    std::vector<int32_t> name;
    std::vector<int64_t> key;
    for (size_t i = 0; i < rules.size(); ++i) {
        if (status[i] != kRuleOk || !comp[i].pairsLaunch || comp[i].pairItems <= 0) continue;
        const RuleIR& r = rules[i];
        if (!r.head.s.var || !r.head.o.var || r.head.s.v == r.head.o.v) continue;
        int32_t maxVar = std::max(r.head.s.v, r.head.o.v);
        for (const AtomIR& a : r.body) {
            if (a.s.var) maxVar = std::max(maxVar, a.s.v);
            if (a.o.var) maxVar = std::max(maxVar, a.o.v);
        }
        name.assign((size_t)maxVar + 1, -1);
        name[r.head.s.v] = 0;
        name[r.head.o.v] = 1;
        int32_t next = 2;
        auto enc = [&](const TermIR& t) -> int64_t {
            if (!t.var) return 2 * (int64_t)t.v + 1;
            if (name[t.v] < 0) name[t.v] = next++;
            return 2 * (int64_t)name[t.v];
        };
        std::vector<std::array<int64_t, 3>> atoms;
        for (const AtomIR& a : r.body) atoms.push_back({(int64_t)a.pterm, enc(a.s), enc(a.o)});
        std::sort(atoms.begin(), atoms.end());
        key.clear();
        for (const auto& a : atoms) key.insert(key.end(), a.begin(), a.end());
        auto it = first.find(key);
        if (it == first.end()) first.emplace(key, (int32_t)i);
        else from[i] = it->second;
    }
    return from;
}

bool batch_build(Context& C, Batch& B, const std::vector<RuleIR>& rules, const std::vector<int32_t>& parseStatus,
                 int32_t flags, hipStream_t st) {
    B.ctx = &C;
    B.n = (int32_t)rules.size();
    B.slots[0] = B.slots[1] = 1;
    B.flags = flags;
    B.comp.assign(rules.size(), CompiledRule{});
    B.status = parseStatus;
    const bool gated = !(flags & GPU_RULES_UNGATED);
    const size_t np = std::max<size_t>(2 * rules.size(), 1);
    std::vector<Plan> plans(np), wplans(np);
    std::vector<ConflictRefs> refs(np), wrefs(np);
    std::vector<Chunk> lists[2];
    uint64_t items[2] = {0, 0};
    struct Job { int ph; size_t k; int64_t n; bool fus; bool shared; };
    std::vector<Job> jobs;
    auto add = [&](int ph, size_t i, const Plan& P, const ConflictRefs& R, const Plan& W, const ConflictRefs& RW,
                   int64_t n, int32_t supPred, bool fus) {
        const size_t k = 2 * i + ph;
        plans[k] = P;
        wplans[k] = W;
        refs[k] = R;
        wrefs[k] = RW;
        plans[k].rule = wplans[k].rule = (int32_t)i;
        plans[k].supPred = wplans[k].supPred = ph == kPairs ? supPred : -1;
        B.slots[ph] = std::max({B.slots[ph], plan_slots(P, R), plan_slots(W, RW)});
        jobs.push_back(Job{ph, k, n, fus, false});
        items[ph] += (uint64_t)n;
    };
    // compile on host threads: rules are independent and the graph is only read
    {
        const size_t nr = rules.size();
        const unsigned hw = std::thread::hardware_concurrency();
        const size_t nt = std::min<size_t>(std::min<unsigned>(hw ? hw : 1u, 16u), std::max<size_t>(1, nr / 256));
        std::vector<std::exception_ptr> errs(nt);
        auto compile_range = [&](size_t t) {
            try {
                for (size_t i = t; i < nr; i += nt) {
                    if (B.status[i] != kRuleOk) continue;
                    B.comp[i] = compile_rule(rules[i], C.g, C.tune);
                    B.status[i] = B.comp[i].status;
                }
            } catch (...) {
                errs[t] = std::current_exception();
            }
        };
        std::vector<std::thread> pool;
        for (size_t t = 1; t < nt; ++t) pool.emplace_back(compile_range, t);
        compile_range(0);
        for (auto& th : pool) th.join();
        for (auto& e : errs)
            if (e) std::rethrow_exception(e);
    }
    B.ruleThr.assign(rules.size(), 1u);
    if (gated)
        for (size_t i = 0; i < rules.size(); ++i)
            if (B.status[i] == kRuleOk) B.ruleThr[i] = support_threshold(B.comp[i].headSize, B.thr);
    // rules with the body of an earlier rule take its body-pair count
    B.pairsFrom = C.shareBodies ? shared_pairs(rules, B.comp, B.status) : std::vector<int32_t>(rules.size(), -1);
    std::vector<std::vector<int32_t>> members(rules.size());
    B.nShared = 0;
    for (size_t i = 0; i < rules.size(); ++i)
        if (B.pairsFrom[i] >= 0) {
            members[B.pairsFrom[i]].push_back((int32_t)i);
            ++B.nShared;
        }
    // Rules whose support the pairs pass may count (a fused launch then skips
    // their support chunks): a search's own rule alone (Plan::supPred = its
    // head predicate), or, when rules share the search, every qualifying one
    // through a member list in supTab (supPred <= -2).
    // A gated rule whose threshold exceeds all its head triples cannot meet it
    // (its support counts head triples): it needs no body-pair count, and its
    // support comes from the support pass, an existence search per head
    // triple, not from a body-pair search. A search other rules share still
    // runs while one of them may meet its threshold.
    std::vector<uint8_t> noPairs(rules.size(), 0);
    if (gated)
        for (size_t i = 0; i < rules.size(); ++i)
            if (B.status[i] == kRuleOk &&
                (int64_t)B.ruleThr[i] > std::max<int64_t>(B.comp[i].headSize, B.comp[i].headSupport))
                noPairs[i] = 1;
    auto searched = [&](size_t i) {
        if (!noPairs[i]) return true;
        for (int32_t m : members[i])
            if (!noPairs[(size_t)m]) return true;
        return false;
    };
    auto fusable = [&](size_t i) {
        const CompiledRule& cr = B.comp[i];
        return B.status[i] == kRuleOk && !noPairs[i] && cr.supPred >= 0 && cr.supItems > 0;
    };
    std::vector<uint8_t> fused(rules.size(), 0);
    std::vector<int32_t> supField(rules.size(), -1);
    std::vector<int32_t> supTab(1, 0);   // word 0 unused: the buffer is never empty
    for (size_t i = 0; i < rules.size(); ++i) {
        if (B.status[i] != kRuleOk || B.pairsFrom[i] >= 0) continue;
        const CompiledRule& cr = B.comp[i];
        if (!cr.pairsLaunch || cr.pairItems <= 0 || !searched(i)) continue;
        std::vector<int32_t> list;
        if (fusable(i)) list.push_back((int32_t)i);
        for (int32_t m : members[i])
            if (fusable((size_t)m)) list.push_back(m);
        if (list.empty()) continue;
        for (int32_t r : list) fused[r] = 1;
        if (list.size() == 1 && list[0] == (int32_t)i) {
            supField[i] = cr.supPred;
        } else {
            supField[i] = -(int32_t)supTab.size() - 2;
            supTab.push_back((int32_t)list.size());
            for (int32_t r : list) {
                supTab.push_back(r);
                supTab.push_back(B.comp[r].supPred);
            }
        }
    }
    for (size_t i = 0; i < rules.size(); ++i) {
        if (B.status[i] != kRuleOk) continue;
        const CompiledRule& cr = B.comp[i];
        if (cr.supportLaunch && cr.supItems > 0)
            add(0, i, cr.sup, cr.supRefs, cr.supWave, cr.supWaveRefs, cr.supItems, -1, fused[i] != 0);
        if (B.pairsFrom[i] < 0 && cr.pairsLaunch && cr.pairItems > 0 && searched(i)) {
            add(1, i, cr.pairs, cr.pairsRefs, cr.pairsWave, cr.pairsWaveRefs, cr.pairItems, supField[i],
                supField[i] != -1);
            if (!members[i].empty()) jobs.back().shared = true;
        }
    }
    // Seeds per thread-mode work unit: 1024 (fewer plan loads and cursor
    // atomics) unless that leaves fewer than chunksPerWave units per
    // thread-mode wave, down to 64 -- a batch of few rules with few, costly
    // seeds each still spreads over the whole GPU.
    const uint64_t waves = (uint64_t)C.numCUs * (uint64_t)C.threadWavesPerCU;
    B.nFusable = 0;
    B.nFused = 0;
    for (size_t i = 0; i < rules.size(); ++i) B.nFused += fused[i];
    B.nSearches[0] = B.nSearches[1] = 0;
    for (const Job& j : jobs) ++B.nSearches[j.ph];
    for (int ph = 0; ph < 2; ++ph) {
        int64_t chunk = kChunkSeeds;
        while (chunk > kMinChunkSeeds && items[ph] / (uint64_t)chunk < (uint64_t)C.chunksPerWave * waves) chunk >>= 1;
        // support chunks of fusable rules go last (pass 1)
        for (int pass = 0; pass < (ph == kSupport ? 2 : 1); ++pass)
            for (const Job& j : jobs) {
                if (j.ph != ph || (ph == kSupport && (int)j.fus != pass)) continue;
                // a search whose count other rules share runs whatever its own rule's support
                const uint32_t gate = (ph == kPairs && gated && !j.shared)
                                          ? (B.ruleThr[j.k / 2] | (j.fus ? kGateFused : 0u))
                                          : 0u;
                for (int64_t b = 0; b < j.n; b += chunk) {
                    lists[ph].push_back(
                        Chunk{(uint32_t)j.k, (uint32_t)b, (uint32_t)std::min<int64_t>(chunk, j.n - b), gate});
                    if (ph == kSupport && j.fus) ++B.nFusable;
                }
            }
    }
    for (int ph = 0; ph < 2; ++ph)
        if (items[ph] > C.phaseSeedCap) {
            std::fprintf(stderr, "[gpu_rules] a batch phase with %llu seeds exceeds the heavy queue's 32-bit range; "
                                 "split the batch\n", (unsigned long long)items[ph]);
            B.seedsOverflow = true;
            return false;
        }
    std::vector<Chunk> all;
    uint64_t heavyTotal = 0;
    for (int ph = 0; ph < 2; ++ph) {
        B.chunkOff[ph] = (uint32_t)all.size();
        B.nChunks[ph] = (uint32_t)lists[ph].size();
        all.insert(all.end(), lists[ph].begin(), lists[ph].end());
        B.heavyOff[ph] = heavyTotal;
        B.heavyCap[ph] = (uint32_t)items[ph];
        heavyTotal += items[ph];
    }
    if (all.empty()) all.push_back(Chunk{0, 0, 0, 0});
    B.capRules = std::max<uint32_t>((uint32_t)rules.size(), 1);
    // stream-ordered buffers: a batch rebuilt on a serving slot grows them
    // without waiting for the batches of other slots
    if (!put_on(B.plans, plans, st) || !put_on(B.wplans, wplans, st) || !put_on(B.refs, refs, st) ||
        !put_on(B.wrefs, wrefs, st) || !put_on(B.chunks, all, st) || !put_on(B.supTab, supTab, st))
        return false;
    if (!B.counts.ensure_on((size_t)2 * B.capRules * 8, st) ||
        !B.heavy.ensure_on((size_t)std::max<uint64_t>(heavyTotal, 1) * sizeof(HeavyItem), st) ||
        !B.ctl.ensure_on(kCtlWords * 4, st))
        return false;
    if (B.hCountsCap < (size_t)2 * B.capRules) {   // pinned: grows seldom (hipHostFree waits for the device)
        if (B.hCounts) (void)hipHostFree(B.hCounts);
        B.hCounts = nullptr;
        const size_t cap = std::max<size_t>({(size_t)2 * B.capRules, 2 * B.hCountsCap, 8192});
        GR_HIP(hipHostMalloc((void**)&B.hCounts, cap * 8, hipHostMallocDefault));
        B.hCountsCap = cap;
    }
    if (!B.hCtl) GR_HIP(hipHostMalloc((void**)&B.hCtl, kCtlWords * 4, hipHostMallocDefault));
    if (!B.hGiant) GR_HIP(hipHostMalloc((void**)&B.hGiant, 8, hipHostMallocDefault));
    if (!B.hWaveStats) GR_HIP(hipHostMalloc((void**)&B.hWaveStats, 2 * kWaveStats * 8, hipHostMallocDefault));
    std::memset(B.hWaveStats, 0, 2 * kWaveStats * 8);
    B.hGiant[0] = B.hGiant[1] = 0;
    GR_HIP(hipStreamSynchronize(st));
    return true;
}

// The device pass of a batch (what the HIP graph captures): per phase, one
// persistent thread-mode kernel over all chunks, then one wave-mode kernel
// over the seeds it escalated. PAIRS chunks of rules whose support is 0 are
// skipped on the device when the batch is gated. With bitmap dedup the pairs
// phase also counts the support of fusable rules, whose support chunks are
// then left out (and their pairs chunks are not gated). A search shared by
// rules with equal bodies is never gated; each of its fusable rules' support
// goes to that rule's counter, so the phases never write the same counter.
hipError_t batch_enqueue(Context& C, Batch& B, const ExecTarget& t) {
    HipFirst err;
    hipStream_t st = t.st;
    uint32_t* ctl = B.ctl.as<uint32_t>();
    unsigned long long* counts = B.counts.as<unsigned long long>();
    const bool fuse = C.fuseActive(t.pairsDedup);
    const uint32_t ncPh[2] = {B.nChunks[0] - (fuse ? B.nFusable : 0u), B.nChunks[1]};
    // ungated: no pairs chunk reads a support count, so the phases may overlap
    const bool fork = t.st2 && t.giants2 && !t.ev && (B.flags & GPU_RULES_UNGATED) && ncPh[0] && ncPh[1];
    // one launch zeroes the counters and each phase's giant counters and
    // flags -- except a second phase's that share the first phase's giant
    // buffers, which are zeroed between the phases
    GiantBufs* const gbOf[2] = {t.giants, fork ? t.giants2 : t.giants};
    GiantArgs gas[2] = {};
    bool early[2] = {false, false};
    ZeroRanges z{};
    zero_add(z, counts, 4 * (int64_t)B.capRules);
    zero_add(z, ctl, kCtlWords);
    for (int ph = 0; ph < 2; ++ph) {
        early[ph] = ncPh[ph] && !(ph == kPairs && ncPh[0] && gbOf[1] == gbOf[0]);
        if (early[ph]) giant_args(*gbOf[ph], t.giantAfter, gas[ph], ph, z);
    }
    launch_zero(z, st);
    if (t.ev) err(hipEventRecord(t.ev[0], st), "hipEventRecord");
    if (fork) {
        err(hipEventRecord(t.forkEv, st), "hipEventRecord(fork)");
        err(hipStreamWaitEvent(t.st2, t.forkEv, 0), "hipStreamWaitEvent(fork)");
    }
    const hipStream_t st0 = st;
    for (int ph = 0; ph < 2; ++ph) {
        const hipStream_t st = (fork && ph == kPairs) ? t.st2 : st0;
        GiantBufs* gbp = gbOf[ph];
        const uint32_t nc = ncPh[ph];
        const bool fusePh = fuse && ph == kPairs;
        uint32_t* c = ctl + ph * 4;
        if (nc) {
            ThreadArgs ta{B.plans.as<Plan>(), B.refs.as<ConflictRefs>(), B.chunks.as<Chunk>() + B.chunkOff[ph], nc,
                          B.heavyCap[ph], c + 0, counts, B.heavy.as<HeavyItem>() + B.heavyOff[ph], c + 1,
                          fusePh ? 1 : 0, B.supTab.as<int32_t>()};
            const int blocks = std::min<int>(scale_grid(C.threadBlocks(), t.gridPercent), (int)((nc + 3) / 4));
            launch_thread_batch(ph, B.slots[ph], ta, C.dg, std::max(blocks, 1), st);
        }
        if (t.ev) err(hipEventRecord(t.ev[1 + 2 * ph], st), "hipEventRecord");
        if (nc) {
            const bool sort = ph == kPairs && t.pairsDedup == kDedupSort && t.sort;
            const int hoBits = bits_for((uint64_t)C.g.nE);
            SortSink sink{};
            if (sort) sink = SortSink{t.sort->keys[0].as<unsigned long long>(), ctl_emitted(ctl, ph), t.sort->cap, hoBits};
            WaveArgs wa{B.wplans.as<Plan>(), B.wrefs.as<ConflictRefs>(), B.heavy.as<HeavyItem>() + B.heavyOff[ph],
                        c + 1, B.heavyCap[ph], fusePh ? 1u : 0u, c + 2, counts, t.dedup, t.rowWords, sink};
            wa.supTab = B.supTab.as<int32_t>();
            wa.engine = C.waveEngine;
            const int mode = ph == kPairs ? t.pairsDedup : kDedupRows;
            if (!early[ph]) giant_prepare(*gbp, t.giantAfter, gas[ph], st, ph);
            wa.giant = gas[ph];
            const int wb = scale_grid(t.pairsDedup == kDedupLds ? C.waveBlocksFor(B.slots[ph]) : t.waveBlocks,
                                      t.gridPercent);
            launch_wave_batch(ph, B.slots[ph], mode, wa, C.dg, wb, st);
            if (t.giantAfter > 0) {
                launch_giants(ph, B.slots[ph], mode, wa, C.dg, wb, st);
                err(hipMemcpyAsync(B.hGiant + ph, wa.giant.count, 4, hipMemcpyDeviceToHost, st), "giant count copy");
            }
            err(hipMemcpyAsync(B.hWaveStats + ph * kWaveStats, wa.giant.stats, kWaveStats * 8, hipMemcpyDeviceToHost,
                               st), "wave stats copy");
            if (sort) {
                if (t.ev) err(hipEventRecord(t.ev[6], st), "hipEventRecord");
                SortJob j{{t.sort->keys[0].as<unsigned long long>(), t.sort->keys[1].as<unsigned long long>()},
                          ctl_emitted(ctl, ph), t.sort->cap, sort_max_tiles(t.sort->cap), t.sort->hist.as<uint32_t>(),
                          hoBits + bits_for(B.heavyCap[ph]), hoBits, B.heavy.as<HeavyItem>() + B.heavyOff[ph],
                          B.wplans.as<Plan>(), counts, t.sortBlocks};
                launch_sort_unique(j, st);
            }
        }
        if (t.ev) err(hipEventRecord(t.ev[2 + 2 * ph], st), "hipEventRecord");
    }
    if (fork) {
        err(hipEventRecord(t.joinEv, t.st2), "hipEventRecord(join)");
        err(hipStreamWaitEvent(st0, t.joinEv, 0), "hipStreamWaitEvent(join)");
    }
    err(hipMemcpyAsync(B.hCounts, counts, (size_t)2 * B.capRules * 8, hipMemcpyDeviceToHost, st0), "counts copy");
    err(hipMemcpyAsync(B.hCtl, ctl, kCtlWords * 4, hipMemcpyDeviceToHost, st0), "control copy");
    err(hipGetLastError(), "kernel launch");
    if (!err.ok()) log_err(err.what, err.e);
    return err.e;
}

int64_t batch_emitted(const Batch& B) {
    return B.hCtl ? (int64_t)*ctl_emitted(B.hCtl, kPairs) : 0;
}

bool batch_heavy_ok(const Batch& B) {
    if (!B.hCtl) return true;
    for (int ph = 0; ph < 2; ++ph)
        if (heavy_count_action(B.hCtl[4 * ph + 1], B.heavyCap[ph]) != kCountsOk) {
            std::fprintf(stderr, "[gpu_rules] phase %d: %u seeds escalated, beyond the heavy queue's %u\n", ph,
                         B.hCtl[4 * ph + 1], B.heavyCap[ph]);
            return false;
        }
    return true;
}

void batch_collect(Batch& B, GpuRuleResult64* out) {
    B.lastHeavy = (int64_t)B.hCtl[1] + B.hCtl[5];
    B.lastEmitted = batch_emitted(B);
    const bool gated = !(B.flags & GPU_RULES_UNGATED);
    for (int32_t i = 0; i < B.n; ++i) {
        GpuRuleResult64& r = out[i];
        std::memset(&r, 0, sizeof(r));
        const CompiledRule& cr = B.comp[i];
        r.status = B.status[i];
        r.headSize = cr.headSize;
        r.headSupport = cr.headSupport;
        r.bodyPairs = -1;
        if (r.status != kRuleOk) continue;
        r.support = cr.supportLaunch ? (int64_t)B.hCounts[2 * i] : 0;
        const size_t src = B.pairsFrom.empty() || B.pairsFrom[i] < 0 ? (size_t)i : (size_t)B.pairsFrom[i];
        if (!gated || r.support >= (int64_t)B.ruleThr[i])
            r.bodyPairs = cr.pairsLaunch ? (int64_t)B.hCounts[2 * src + 1] : 0;
        r.bodySize = r.support > 0 && r.bodyPairs > 0 ? r.bodyPairs : 0;
        finish_measures(r);
    }
}

bool batch_run(Context& C, Batch& B, GpuRuleResult64* out) {
    const bool graphMode = C.useGraph;
    SortBufs* sb = nullptr;
    const int mode = C.pairsDedup();
    if (mode == kDedupSort && B.nChunks[kPairs]) {
        if (!C.sortb.ensure(1, kSortCapMin, C.sortCapMax)) return false;
        sb = &C.sortb;
    }
    for (;;) {
        B.hGiant[0] = B.hGiant[1] = 0;
        if (graphMode) {
            // re-capture if the context's scratch or launch shape changed since
            const void* sp = sb ? sb->keys[0].p : nullptr;
            const uint32_t sc = sb ? sb->cap : 0;
            if (B.exec && (B.capturedDedup != C.dedup.p || B.capturedBlocks[0] != C.threadBlocks() ||
                           B.capturedBlocks[1] != C.waveBlocks() || B.capturedSort != sp || B.capturedSortCap != sc ||
                           B.capturedMode != mode || B.capturedGiants != C.giants.rows.p ||
                           B.capturedGiantAfter != C.giantAfter || B.capturedOverlap != (int)C.overlap ||
                           B.capturedEngine != C.waveEngine ||
                           B.capturedFuse != (int)C.fuseActive(mode))) {
                (void)hipGraphExecDestroy(B.exec);
                (void)hipGraphDestroy(B.graph);
                B.exec = nullptr;
                B.graph = nullptr;
            }
            if (!B.exec) {
                B.capturedDedup = C.dedup.p;
                B.capturedBlocks[0] = C.threadBlocks();
                B.capturedBlocks[1] = C.waveBlocks();
                B.capturedSort = sp;
                B.capturedSortCap = sc;
                B.capturedMode = mode;
                B.capturedGiants = C.giants.rows.p;
                B.capturedGiantAfter = C.giantAfter;
                B.capturedEngine = C.waveEngine;
                B.capturedOverlap = (int)C.overlap;
                B.capturedFuse = (int)C.fuseActive(mode);
                GR_HIP(hipStreamBeginCapture(C.st, hipStreamCaptureModeThreadLocal));
                ExecTarget t{C.st, C.dedup.as<uint32_t>(), C.rowWords, C.waveBlocks(), nullptr, sb,
                             C.sortBlocks(), mode, &C.giants, C.giantAfter};
                if (C.overlap) {
                    t.st2 = C.st2;
                    t.forkEv = C.forkEv;
                    t.joinEv = C.joinEv;
                    t.giants2 = &C.giants2;
                }
                const hipError_t ee = batch_enqueue(C, B, t);
                GR_HIP(hipStreamEndCapture(C.st, &B.graph));
                if (ee != hipSuccess) {
                    (void)hipGraphDestroy(B.graph);
                    B.graph = nullptr;
                    return false;
                }
                GR_HIP(hipGraphInstantiate(&B.exec, B.graph, nullptr, nullptr, 0));
            }
            GR_HIP(hipEventRecord(C.ev[0], C.st));
            GR_HIP(hipGraphLaunch(B.exec, C.st));
            GR_HIP(hipEventRecord(C.ev[5], C.st));
            GR_HIP(hipStreamSynchronize(C.st));
            float f = 0;
            (void)hipEventElapsedTime(&f, C.ev[0], C.ev[5]);
            B.lastMs = f;
            B.lastSupMs = B.lastPairMs = B.lastSortMs = 0;
            for (double& k : B.kernelMs) k = 0;
        } else {
            GR_HIP(batch_enqueue(C, B, ExecTarget{C.st, C.dedup.as<uint32_t>(), C.rowWords, C.waveBlocks(), C.ev, sb,
                                                  C.sortBlocks(), mode, &C.giants, C.giantAfter}));
            GR_HIP(hipEventRecord(C.ev[5], C.st));
            GR_HIP(hipStreamSynchronize(C.st));
            float t[5] = {0, 0, 0, 0, 0};
            for (int k = 0; k < 4; ++k) (void)hipEventElapsedTime(&t[k], C.ev[k], C.ev[k + 1]);
            (void)hipEventElapsedTime(&t[4], C.ev[0], C.ev[5]);
            B.lastSortMs = 0;
            if (sb) {   // pairs wave kernel alone, then the sort-unique pass
                float w = 0, so = 0;
                (void)hipEventElapsedTime(&w, C.ev[3], C.ev[6]);
                (void)hipEventElapsedTime(&so, C.ev[6], C.ev[4]);
                t[3] = w;
                B.lastSortMs = so;
            }
            for (int k = 0; k < 4; ++k) B.kernelMs[k] = t[k];
            B.lastSupMs = t[0] + t[1];
            B.lastPairMs = t[2] + t[3] + (float)B.lastSortMs;
            B.lastMs = t[4];
        }
        GR_HIP(hipGetLastError());
        if (!batch_heavy_ok(B)) return false;
        if (sb) {   // keys overflowed: larger buffer and run again, or fail past the largest one
            const uint64_t emitted = (uint64_t)batch_emitted(B);
            const int32_t act = sort_count_action(emitted, sb->cap, C.sortCapMax);
            if (act == kCountsFail) {
                std::fprintf(stderr, "[gpu_rules] body-pair keys of one batch (%llu) exceed the sort capacity %u\n",
                             (unsigned long long)emitted, sb->cap);
                return false;
            }
            if (act == kCountsRegrow) {
                if (!sb->ensure(emitted, kSortCapMin, C.sortCapMax)) return false;
                continue;
            }
        }
        break;
    }
    batch_collect(B, out);
    C.lastMs = B.lastMs;
    C.lastSupMs = B.lastSupMs;
    C.lastPairMs = B.lastPairMs;
    C.lastSortMs = B.lastSortMs;
    C.lastHeavy = B.lastHeavy;
    C.lastEmitted = sb ? B.lastEmitted : 0;
    C.lastGiants = (int64_t)B.hGiant[0] + B.hGiant[1];
    C.lastFused = C.fuseActive(mode) ? B.nFused : 0;
    C.lastShared = B.nShared;
    // fused support chunks do not run: those rules' support is counted by the pairs searches
    C.lastSearches[0] = B.nSearches[0] - (C.fuseActive(mode) ? B.nFused : 0);
    C.lastSearches[1] = B.nSearches[1];
    for (int k = 0; k < 2 * kWaveStats; ++k) C.lastWaveStats[k] = B.hWaveStats[k];
    for (int k = 0; k < 4; ++k) C.lastKernelMs[k] = B.kernelMs[k];
    return true;
}

// ------------------------------------------------------------ handles

Context* as_ctx(void* h) {
    if (!h || static_cast<Handle*>(h)->magic != kMagicCtx) return nullptr;
    return static_cast<Context*>(h);
}

const HostGraph* as_graph(void* h) {
    if (!h) return nullptr;
    uint32_t m = static_cast<Handle*>(h)->magic;
    if (m == kMagicCtx) return &static_cast<Context*>(h)->g;
    if (m == kMagicHost) return &static_cast<HostHandle*>(h)->g;
    return nullptr;
}

TermResolver* as_resolver(void* h) {
    if (!h) return nullptr;
    uint32_t m = static_cast<Handle*>(h)->magic;
    if (m == kMagicCtx) return static_cast<Context*>(h)->tr.get();
    if (m == kMagicHost) return static_cast<HostHandle*>(h)->tr.get();
    return nullptr;
}

// The adjacency arrays of g from its triples: with the GPU's sort when
// device_build_wanted (falling back to the host build if it cannot run),
// else on the host. Both give the same arrays.
void build_arrays(HostGraph& g, std::vector<int32_t>& S, std::vector<int32_t>& P, std::vector<int32_t>& O,
                  int64_t nTerms) {
    if (device_build_wanted((int64_t)S.size())) {
        HostGraph d;
        std::string err;
        if (build_graph_device(d, S.data(), P.data(), O.data(), (int64_t)S.size(), nTerms, err)) {
            // the device build fills everything but what parsing produced
            d.terms = std::move(g.terms);
            d.prefixes = std::move(g.prefixes);
            g = std::move(d);
            std::vector<int32_t>().swap(S);
            std::vector<int32_t>().swap(P);
            std::vector<int32_t>().swap(O);
            return;
        }
        std::fprintf(stderr, "[gpu_rules] device build not used (%s); building on the host\n", err.c_str());
    }
    build_graph(g, S, P, O, nTerms);
}

Context* new_context_from(HostGraph&& g) {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) {
        std::fprintf(stderr, "[gpu_rules] no GPU device available\n");
        return nullptr;
    }
    auto* C = new Context();
    C->magic = kMagicCtx;
    C->g = std::move(g);
    C->tr.reset(new TermResolver(C->g));
    C->rowStats = row_stats(C->g);
    // GPU_RULES_WAVE_ENGINE (0 or 1) picks the wave-mode engine of new contexts
    if (const char* we = std::getenv("GPU_RULES_WAVE_ENGINE"))
        if ((we[0] == '0' || we[0] == '1') && we[1] == 0) C->waveEngine = we[0] - '0';
    // HBM estimate, as the reference checks before uploading (GPU_03/gpu_rules_lib.cu:222-279)
    size_t freeB = 0, totalB = 0;
    if (hipMemGetInfo(&freeB, &totalB) == hipSuccess) {
        const size_t need = (size_t)2 * C->g.P() * C->g.nBlocks() * 16 + (size_t)C->g.T() * 16 +
                            (size_t)(C->g.keysS.size() + C->g.keysO.size()) * 4 + (size_t)2048 * (C->g.nE / 8 + 4) +
                            (64u << 20);
        if (need > freeB) {
            std::fprintf(stderr, "[gpu_rules] not enough HBM: need ~%zu MiB, %zu MiB free (of %zu MiB)\n",
                         need >> 20, freeB >> 20, totalB >> 20);
            delete C;
            return nullptr;
        }
    }
    if (!C->upload()) {
        delete C;
        return nullptr;
    }
    return C;
}

}  // namespace gr

using namespace gr;

// ===================================================================== ABI

extern "C" {

const char* gpu_rules_version(void) { return "gpu_rules 1.0 (gfx950)"; }

int gpu_query_vram(size_t* freeBytesOut, size_t* totalBytesOut) {
    if (!freeBytesOut || !totalBytesOut) return -1;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) return -1;
    return hipMemGetInfo(freeBytesOut, totalBytesOut) == hipSuccess ? 0 : -1;
}

void* gpu_load_graph(const char* ttlPath) {
    if (!ttlPath) return nullptr;
    try {
        HostGraph g;
        std::string err;
        std::vector<int32_t> S, P, O;
        if (!parse_turtle(ttlPath, g, S, P, O, err)) {
            std::fprintf(stderr, "[gpu_rules] failed to parse TTL: %s (%s)\n", ttlPath, err.c_str());
            return nullptr;
        }
        build_arrays(g, S, P, O, (int64_t)g.terms.size());
        return new_context_from(std::move(g));
    } catch (const std::exception& e) {
        std::fprintf(stderr, "[gpu_rules] gpu_load_graph error: %s\n", e.what());
        return nullptr;
    }
}

void* gpu_load_graph_triples(const int32_t* s, const int32_t* p, const int32_t* o, int64_t n, int64_t nTerms) {
    if ((!s || !p || !o) && n > 0) return nullptr;
    try {
        for (int64_t i = 0; i < n; ++i) {
            if (s[i] < 0 || p[i] < 0 || o[i] < 0 || s[i] >= nTerms || p[i] >= nTerms || o[i] >= nTerms) {
                std::fprintf(stderr, "[gpu_rules] triple %lld has a term id outside [0, %lld)\n", (long long)i,
                             (long long)nTerms);
                return nullptr;
            }
        }
        HostGraph g;
        std::vector<int32_t> S(s, s + n), P(p, p + n), O(o, o + n);
        build_arrays(g, S, P, O, nTerms);
        return new_context_from(std::move(g));
    } catch (const std::exception& e) {
        std::fprintf(stderr, "[gpu_rules] gpu_load_graph_triples error: %s\n", e.what());
        return nullptr;
    }
}

void* gpu_load_snapshot(const char* path) {
    if (!path) return nullptr;
    try {
        HostGraph g;
        std::string err;
        if (!load_snapshot(path, g, err)) {
            std::fprintf(stderr, "[gpu_rules] failed to load snapshot: %s (%s)\n", path, err.c_str());
            return nullptr;
        }
        return new_context_from(std::move(g));
    } catch (const std::exception& e) {
        std::fprintf(stderr, "[gpu_rules] gpu_load_snapshot error: %s\n", e.what());
        return nullptr;
    }
}

int gpu_save_snapshot(void* handle, const char* path) {
    const HostGraph* g = as_graph(handle);
    if (!g || !path) return -1;
    try {
        std::string err;
        if (!save_snapshot(*g, path, err)) {
            std::fprintf(stderr, "[gpu_rules] failed to save snapshot: %s\n", err.c_str());
            return -1;
        }
        return 0;
    } catch (const std::exception& e) {
        std::fprintf(stderr, "[gpu_rules] gpu_save_snapshot error: %s\n", e.what());
        return -1;
    }
}

void gpu_free_graph(void* ctx) {
    Context* C = as_ctx(ctx);
    if (!C) return;
    delete C->lastBatch;
    C->lastBatch = nullptr;
    delete C;
}

static int eval_one_abi(Context* C, const RuleIR& r, GpuRuleResult* out, OneScratch S) {
    GpuRuleResult64 res;
    double supMs = 0, confMs = 0;
    if (!eval_rule_plain(*C, S, r, true, Thresholds{}, res, supMs, confMs)) return -1;
    out->support = (int)res.support;
    out->headSize = (int)res.headSize;
    out->headSupport = (int)res.headSupport;
    out->bodySize = (int)res.bodySize;
    // measures from the int fields, as FinalRule::setMeasures/setConfidence do
    // (GPU_03/FinalRule.cpp:19-40)
    out->headCoverage = out->headSize > 0 ? (double)out->support / (double)out->headSize : 0.0;
    out->confidence = out->bodySize > 0 ? (double)out->support / (double)out->bodySize : 0.0;
    out->supportMs = supMs;
    out->confidenceMs = confMs;
    return 0;
}

int gpu_evaluate_rule(void* ctx, const char* ruleText, GpuRuleResult* out) {
    Context* C = as_ctx(ctx);
    if (!C || !ruleText || !out) return -1;
    std::memset(out, 0, sizeof(*out));
    try {
        RuleIR r = parse_rule_text(ruleText, *C->tr);
        return eval_one_abi(C, r, out, context_scratch(*C));
    } catch (const std::exception& e) {
        std::fprintf(stderr, "[gpu_rules] gpu_evaluate_rule error: %s\n", e.what());
        return -1;
    }
}

int gpu_evaluate_rule_noalloc(void* ctx, const char* ruleText, GpuRuleResult* out) {
    Context* C = as_ctx(ctx);
    if (!C || !ruleText || !out) return -1;
    std::memset(out, 0, sizeof(*out));
    char* s = C->noallocReady ? static_cast<char*>(noalloc_scratch(C->device)) : nullptr;
    if (!s) {
        std::fprintf(stderr, "[gpu_rules] gpu_evaluate_rule_noalloc: no device scratch on device %d\n", C->device);
        return -1;
    }
    try {
        RuleIR r = parse_rule_text(ruleText, *C->tr);
        OneScratch S{s, reinterpret_cast<HeavyItem*>(s + kNoAllocHeader),
                     std::min<uint64_t>(kNoAllocMaxSeeds, C->noallocMaxSeeds), false};
        return eval_one_abi(C, r, out, S);
    } catch (const std::exception& e) {
        std::fprintf(stderr, "[gpu_rules] gpu_evaluate_rule_noalloc error: %s\n", e.what());
        return -1;
    }
}

int gpu_evaluate_rule_json(void* ctx, const char* ruleJson, GpuRuleResult* out) {
    Context* C = as_ctx(ctx);
    if (!C || !ruleJson || !out) return -1;
    std::memset(out, 0, sizeof(*out));
    try {
        RuleIR r = parse_rule_json(ruleJson, *C->tr);
        return eval_one_abi(C, r, out, context_scratch(*C));
    } catch (const std::exception& e) {
        std::fprintf(stderr, "[gpu_rules] gpu_evaluate_rule_json error: %s\n", e.what());
        return -1;
    }
}

static int eval_batch(Context* C, std::vector<RuleIR>& rules, std::vector<int32_t>& st, int32_t flags,
                      GpuRuleResult64* out, const Thresholds& thr = Thresholds{}) {
    if (flags & GPU_RULES_PER_RULE) {
        const bool gated = !(flags & GPU_RULES_UNGATED);
        double tot = 0, ts = 0, tp = 0;
        int64_t heavy = 0;
        C->lastRuleSupMs.assign(rules.size(), 0.0);
        C->lastRuleConfMs.assign(rules.size(), 0.0);
        for (size_t i = 0; i < rules.size(); ++i) {
            if (st[i] != kRuleOk) {
                std::memset(&out[i], 0, sizeof(out[i]));
                out[i].status = st[i];
                out[i].bodyPairs = -1;
                continue;
            }
            double a = 0, b = 0;
            const auto t0 = std::chrono::steady_clock::now();
            if (!eval_rule_plain(*C, context_scratch(*C), rules[i], gated, thr, out[i], a, b)) return -1;
            const auto t1 = std::chrono::steady_clock::now();
            out[i].latencyUs = (int32_t)std::min<int64_t>(
                std::chrono::duration_cast<std::chrono::microseconds>(t1 - t0).count(), INT32_MAX);
            ts += a; tp += b; tot += a + b;
            C->lastRuleSupMs[i] = a;
            C->lastRuleConfMs[i] = b;
            heavy += C->lastHeavy;
        }
        C->lastMs = tot; C->lastSupMs = ts; C->lastPairMs = tp; C->lastHeavy = heavy;
        return 0;
    }
    delete C->lastBatch;
    C->lastBatch = nullptr;
    // One batch; or, when a phase's seeds exceed the heavy queue's range (a
    // host check before any device work), its two halves, recursively. The
    // last-pass figures then add up over the parts.
    double ms[4] = {0, 0, 0, 0};
    int64_t heavy = 0, giants = 0, fused = 0, shared = 0, searches[2] = {0, 0};
    int parts = 0;
    std::function<bool(size_t, size_t)> part = [&](size_t lo, size_t hi) -> bool {
        std::vector<RuleIR> rs(rules.begin() + (ptrdiff_t)lo, rules.begin() + (ptrdiff_t)hi);
        std::vector<int32_t> ss(st.begin() + (ptrdiff_t)lo, st.begin() + (ptrdiff_t)hi);
        auto* B = new Batch();
        B->magic = kMagicBatch;
        B->thr = thr;
        if (!batch_build(*C, *B, rs, ss, flags, C->st)) {
            const bool split = B->seedsOverflow && hi - lo > 1;
            delete B;
            const size_t mid = lo + (hi - lo) / 2;
            return split && part(lo, mid) && part(mid, hi);
        }
        if (!batch_run(*C, *B, out + lo)) {
            delete B;
            return false;
        }
        delete C->lastBatch;
        C->lastBatch = B;
        ms[0] += C->lastMs; ms[1] += C->lastSupMs; ms[2] += C->lastPairMs; ms[3] += C->lastSortMs;
        heavy += C->lastHeavy; giants += C->lastGiants; fused += C->lastFused; shared += C->lastShared;
        searches[0] += C->lastSearches[0]; searches[1] += C->lastSearches[1];
        ++parts;
        return true;
    };
    if (!part(0, rules.size())) return -1;
    if (parts > 1) {
        C->lastMs = ms[0]; C->lastSupMs = ms[1]; C->lastPairMs = ms[2]; C->lastSortMs = ms[3];
        C->lastHeavy = heavy; C->lastGiants = giants; C->lastFused = fused; C->lastShared = shared;
        C->lastSearches[0] = searches[0]; C->lastSearches[1] = searches[1];
    }
    return 0;
}

int gpu_evaluate_rules(void* ctx, const GpuRuleDesc* rules, int32_t n, int32_t flags, GpuRuleResult64* out) {
    Context* C = as_ctx(ctx);
    if (!C || (!rules && n > 0) || (!out && n > 0) || n < 0) return -1;
    try {
        std::vector<RuleIR> rs(n);
        std::vector<int32_t> st(n, kRuleOk);
        for (int32_t i = 0; i < n; ++i) rs[i] = rule_from_words(rules[i].words);
        return eval_batch(C, rs, st, flags, out);
    } catch (const std::exception& e) {
        std::fprintf(stderr, "[gpu_rules] gpu_evaluate_rules error: %s\n", e.what());
        return -1;
    }
}

int gpu_evaluate_rules_text(void* ctx, const char* const* rules, int32_t n, int32_t isJson, int32_t flags,
                            GpuRuleResult64* out) {
    Context* C = as_ctx(ctx);
    if (!C || (!rules && n > 0) || (!out && n > 0) || n < 0) return -1;
    try {
        std::vector<RuleIR> rs(n);
        std::vector<int32_t> st(n, kRuleOk);
        for (int32_t i = 0; i < n; ++i) {
            try {
                if (!rules[i]) throw ParseError("null rule");
                rs[i] = isJson ? parse_rule_json(rules[i], *C->tr) : parse_rule_text(rules[i], *C->tr);
            } catch (const std::exception& e) {
                st[i] = kRuleError;
            }
        }
        return eval_batch(C, rs, st, flags, out);
    } catch (const std::exception& e) {
        std::fprintf(stderr, "[gpu_rules] gpu_evaluate_rules_text error: %s\n", e.what());
        return -1;
    }
}

// Thresholds of the *_min entry points: NaN, a negative or a > 1 head
// coverage, and an ungated batch are refused.
static bool thresholds_of(int64_t minSupport, double minHeadCoverage, int32_t flags, Thresholds& t) {
    if (flags & GPU_RULES_UNGATED) return false;
    if (!(minHeadCoverage >= 0.0 && minHeadCoverage <= 1.0)) return false;
    t.minSupport = minSupport;
    t.minHeadCoverage = minHeadCoverage;
    return true;
}

int gpu_evaluate_rules_min(void* ctx, const GpuRuleDesc* rules, int32_t n, int32_t flags, int64_t minSupport,
                           double minHeadCoverage, GpuRuleResult64* out) {
    Context* C = as_ctx(ctx);
    Thresholds t;
    if (!C || (!rules && n > 0) || (!out && n > 0) || n < 0 || !thresholds_of(minSupport, minHeadCoverage, flags, t))
        return -1;
    try {
        std::vector<RuleIR> rs(n);
        std::vector<int32_t> st(n, kRuleOk);
        for (int32_t i = 0; i < n; ++i) rs[i] = rule_from_words(rules[i].words);
        return eval_batch(C, rs, st, flags, out, t);
    } catch (const std::exception& e) {
        std::fprintf(stderr, "[gpu_rules] gpu_evaluate_rules_min error: %s\n", e.what());
        return -1;
    }
}

static void* batch_create(void* ctx, const GpuRuleDesc* rules, int32_t n, int32_t flags, const Thresholds& thr) {
    Context* C = as_ctx(ctx);
    if (!C || (!rules && n > 0) || n < 0) return nullptr;
    try {
        std::vector<RuleIR> rs(n);
        std::vector<int32_t> st(n, kRuleOk);
        for (int32_t i = 0; i < n; ++i) rs[i] = rule_from_words(rules[i].words);
        auto* B = new Batch();
        B->magic = kMagicBatch;
        B->thr = thr;
        if (!batch_build(*C, *B, rs, st, flags & GPU_RULES_UNGATED, C->st)) {
            delete B;
            return nullptr;
        }
        return B;
    } catch (const std::exception& e) {
        std::fprintf(stderr, "[gpu_rules] gpu_batch_create error: %s\n", e.what());
        return nullptr;
    }
}

void* gpu_batch_create(void* ctx, const GpuRuleDesc* rules, int32_t n, int32_t flags) {
    return batch_create(ctx, rules, n, flags, Thresholds{});
}

void* gpu_batch_create_min(void* ctx, const GpuRuleDesc* rules, int32_t n, int32_t flags, int64_t minSupport,
                           double minHeadCoverage) {
    Thresholds t;
    if (!thresholds_of(minSupport, minHeadCoverage, flags, t)) return nullptr;
    return batch_create(ctx, rules, n, flags, t);
}

int gpu_batch_run(void* batch, GpuRuleResult64* out) {
    if (!batch || !out || static_cast<Handle*>(batch)->magic != kMagicBatch) return -1;
    Batch* B = static_cast<Batch*>(batch);
    if (!B->ctx) return -1;
    return batch_run(*B->ctx, *B, out) ? 0 : -1;
}

void gpu_batch_free(void* batch) {
    if (batch && static_cast<Handle*>(batch)->magic == kMagicBatch) delete static_cast<Batch*>(batch);
}

int gpu_last_batch_stats(void* ctx, double* totalMs, double* supportMs, double* pairsMs, int64_t* heavySeeds) {
    Context* C = as_ctx(ctx);
    if (!C) return -1;
    if (totalMs) *totalMs = C->lastMs;
    if (supportMs) *supportMs = C->lastSupMs;
    if (pairsMs) *pairsMs = C->lastPairMs;
    if (heavySeeds) *heavySeeds = C->lastHeavy;
    return 0;
}

double gpu_copy_probe(void* ctx, int64_t bytes) {
    Context* C = as_ctx(ctx);
    if (!C || bytes < 16) return -1;
    DevBuf a, b;
    if (!a.ensure((size_t)bytes) || !b.ensure((size_t)bytes)) return -1;
    if (hipMemsetAsync(a.p, 1, (size_t)bytes, C->st) != hipSuccess) return -1;
    launch_copy(a.p, b.p, bytes, C->st);   // untimed first pass
    float best = -1;
    for (int rep = 0; rep < 5; ++rep) {     // every dispatch moves the same bytes: best of five
        (void)hipEventRecord(C->ev[0], C->st);
        launch_copy(a.p, b.p, bytes, C->st);
        (void)hipEventRecord(C->ev[1], C->st);
        if (hipStreamSynchronize(C->st) != hipSuccess) return -1;
        float ms = 0;
        (void)hipEventElapsedTime(&ms, C->ev[0], C->ev[1]);
        if (best < 0 || ms < best) best = ms;
    }
    a.release();
    b.release();
    return best;
}

int gpu_autotune(void* ctx, const GpuRuleDesc* rules, int32_t n, int32_t flags, double* log, int32_t logCap,
                 int32_t* logLen) {
    Context* C = as_ctx(ctx);
    if (!C || (!rules && n > 0) || n <= 0) return -1;
    try {
        std::vector<RuleIR> rs(n);
        std::vector<int32_t> st(n, kRuleOk);
        for (int32_t i = 0; i < n; ++i) rs[i] = rule_from_words(rules[i].words);
        const bool graphWas = C->useGraph;
        C->useGraph = true;
        int32_t nlog = 0;
        // one trial: compile with the current knobs, capture the HIP graph, best of 3 replays
        auto trial = [&](double& ms) -> bool {
            Batch B;
            B.magic = kMagicBatch;
            if (!batch_build(*C, B, rs, st, flags & GPU_RULES_UNGATED, C->st)) return false;
            std::vector<GpuRuleResult64> out(n);
            if (!batch_run(*C, B, out.data())) return false;
            double best = 1e30;
            for (int k = 0; k < 3; ++k) {
                if (!batch_run(*C, B, out.data())) return false;
                best = std::min(best, B.lastMs);
            }
            ms = best;
            if (log && nlog < logCap) {
                double* r = log + 7 * nlog;
                r[0] = C->tune.budget; r[1] = C->threadWavesPerCU; r[2] = C->waveWavesPerCU; r[3] = ms;
                r[4] = (double)B.lastHeavy; r[5] = C->pairsDedup(); r[6] = C->giantAfter;
            }
            ++nlog;
            return true;
        };
        // budget and thread-mode waves together (they interact: a small budget
        // pays only with many waves, and at few waves every budget ties), then
        // the other settings one at a time
        const int32_t budgets[] = {64, 128, 256, 1024, 4096, 16384, 65536};
        const int tw[] = {8, 16, 24, 32};
        const int ww[] = {4, 8, 16};
        constexpr int nB = 7, nT = 4;
        double grid[nB][nT];
        double bestMs = 1e30, ms = 0;
        for (int b = 0; b < nB; ++b)
            for (int q = 0; q < nT; ++q) {
                C->tune.budget = budgets[b];
                C->threadWavesPerCU = tw[q];
                if (!trial(ms)) return -1;
                grid[b][q] = ms;
                bestMs = std::min(bestMs, ms);
            }
        // the most thread-mode waves within 2% of the fastest (in a replay the
        // phases share the GPU, so neighbouring settings tie and the choice
        // would otherwise follow noise), then the fastest budget with them
        int bestB = -1, bestQ = -1;
        for (int q = nT - 1; q >= 0 && bestB < 0; --q)
            for (int b = 0; b < nB; ++b)
                if (grid[b][q] <= bestMs * 1.02 && (bestB < 0 || grid[b][q] < grid[bestB][bestQ])) {
                    bestB = b;
                    bestQ = q;
                }
        C->tune.budget = budgets[bestB];
        int bestTw = tw[bestQ];
        bestMs = grid[bestB][bestQ];
        C->threadWavesPerCU = bestTw;
        // thread-mode waves and work units per wave together: the unit size
        // follows their product (units of 64 to 1,024 seeds, at least
        // chunks_per_wave per wave), so one tuned alone fixes the other's
        // best value -- waves that just fill the resident slots may want
        // more units per wave than the grid above used
        int bestCpw = C->chunksPerWave;
        const int gridTw = bestTw, gridCpw = bestCpw;
        for (int q = 0; q < nT; ++q)
            for (int cpw : {1, 2, 4, 8, 16}) {
                if (tw[q] == gridTw && cpw == gridCpw) continue;   // timed in the grid
                C->threadWavesPerCU = tw[q];
                C->chunksPerWave = cpw;
                if (!trial(ms)) return -1;
                if (ms < bestMs) { bestMs = ms; bestTw = tw[q]; bestCpw = cpw; }
            }
        C->threadWavesPerCU = bestTw;
        C->chunksPerWave = bestCpw;
        int bestWw = C->waveWavesPerCU;
        for (int w : ww) {
            C->waveWavesPerCU = w;
            if (!C->ensure_dedup()) return -1;
            if (!trial(ms)) return -1;
            if (ms < bestMs) { bestMs = ms; bestWw = w; }
        }
        C->waveWavesPerCU = bestWw;
        if (!C->ensure_dedup()) return -1;
        // head-object bitmaps in HBM rows or in LDS (when they fit), unless fixed by the caller
        if (C->dedupMode == 0 && C->ldsFits && !C->sortActive()) {
            int bestMode = 0;
            for (int m : {1, 3}) {
                C->dedupMode = m;
                if (!C->ensure_dedup()) return -1;
                if (!trial(ms)) return -1;
                if (ms < bestMs) { bestMs = ms; bestMode = m; }
            }
            C->dedupMode = bestMode;
            if (!C->ensure_dedup()) return -1;
        }
        // steps before a wave-mode seed is cut into sub-items
        int bestGa = C->giantAfter;
        for (int gaf : {512, 1024, 2048, 4096, 8192}) {
            C->giantAfter = gaf;
            if (!trial(ms)) return -1;
            if (ms < bestMs) { bestMs = ms; bestGa = gaf; }
        }
        C->giantAfter = bestGa;
        // second pass over the budget and the thread-mode waves with every other
        // setting chosen (the grid tried them under the default work units,
        // wave-mode waves and giant threshold)
        for (int32_t b : budgets) {
            if (b == C->tune.budget) continue;
            const int32_t keep = C->tune.budget;
            C->tune.budget = b;
            if (!trial(ms)) return -1;
            if (ms < bestMs) bestMs = ms;
            else C->tune.budget = keep;
        }
        for (int q = 0; q < 4; ++q) {
            if (tw[q] == C->threadWavesPerCU) continue;
            const int keep = C->threadWavesPerCU;
            C->threadWavesPerCU = tw[q];
            if (!trial(ms)) return -1;
            if (ms < bestMs) bestMs = ms;
            else C->threadWavesPerCU = keep;
        }
        C->useGraph = graphWas;
        if (logLen) *logLen = std::min(nlog, logCap);
        return 0;
    } catch (const std::exception& e) {
        std::fprintf(stderr, "[gpu_rules] gpu_autotune error: %s\n", e.what());
        return -1;
    }
}

int gpu_last_dedup_stats(void* ctx, double* sortMs, int64_t* keysEmitted, int64_t* sortCap) {
    Context* C = as_ctx(ctx);
    if (!C) return -1;
    if (sortMs) *sortMs = C->lastSortMs;
    if (keysEmitted) *keysEmitted = C->lastEmitted;
    if (sortCap) *sortCap = C->sortb.cap;
    return 0;
}

int gpu_last_rule_times(void* ctx, double* supportMs, double* confidenceMs, int32_t n) {
    Context* C = as_ctx(ctx);
    if (!C || n < 0 || (n > 0 && (!supportMs || !confidenceMs))) return -1;
    const int32_t k = (int32_t)std::min<size_t>((size_t)n, C->lastRuleSupMs.size());
    for (int32_t i = 0; i < k; ++i) {
        supportMs[i] = C->lastRuleSupMs[i];
        confidenceMs[i] = C->lastRuleConfMs[i];
    }
    return k;
}

int gpu_last_kernel_ms(void* ctx, double* ms4) {
    Context* C = as_ctx(ctx);
    if (!C || !ms4) return -1;
    for (int k = 0; k < 4; ++k) ms4[k] = C->lastKernelMs[k];
    return 0;
}

int gpu_plan_traffic(void* handle, const GpuRuleDesc* rules, int32_t n, int64_t* out) {
    const HostGraph* g = as_graph(handle);
    if (!g || (!rules && n > 0) || (!out && n > 0)) return -1;
    try {
        Tuning t;
        // a context whose batches count fusable rules' support in the pairs pass is modelled that way
        // (and a context that shares body-pair searches between equal bodies, that way)
        const Context* C = as_ctx(handle);
        const bool fuse = C && C->fuseActive(C->pairsDedup());
        std::vector<CompiledRule> crs(std::max<int32_t>(n, 0));
        std::vector<RuleIR> irs(std::max<int32_t>(n, 0));
        std::vector<int32_t> status(std::max<int32_t>(n, 0));
        for (int32_t i = 0; i < n; ++i) {
            irs[i] = rule_from_words(rules[i].words);
            crs[i] = compile_rule(irs[i], *g, t);
            status[i] = crs[i].status;
        }
        const std::vector<int32_t> from =
            C && C->shareBodies ? shared_pairs(irs, crs, status) : std::vector<int32_t>(crs.size(), -1);
        // fused support, as a batch does it: a search's own rule and the
        // qualifying rules that share it
        auto fusable = [&](int32_t i) { return crs[i].status == kRuleOk && crs[i].supPred >= 0 && crs[i].supItems > 0; };
        std::vector<uint8_t> fusedRule(crs.size(), 0);
        std::vector<std::vector<int32_t>> heads(crs.size());
        if (fuse) {
            for (int32_t i = 0; i < n; ++i) {
                const int32_t src = from[i] < 0 ? i : from[i];
                if (crs[src].status != kRuleOk || !crs[src].pairsLaunch || crs[src].pairItems <= 0 || !fusable(i))
                    continue;
                fusedRule[i] = 1;
                heads[src].push_back(crs[i].supPred);
            }
        }
        std::vector<TrafficJob> jobs;
        std::vector<int64_t> slot;   // out index of each job
        for (int32_t i = 0; i < n; ++i) {
            CompiledRule& cr = crs[i];
            for (int k = 0; k < 4; ++k) out[4 * i + k] = 0;
            if (cr.status != kRuleOk) continue;
            const bool fus = fusedRule[i] != 0;
            cr.pairs.supPred = -1;
            if (cr.supportLaunch && !fus) {
                jobs.push_back(TrafficJob{&cr.sup, &cr.supRefs, cr.supItems});
                slot.push_back(4 * (int64_t)i);
            }
            if (cr.pairsLaunch && from[i] < 0) {
                jobs.push_back(TrafficJob{&cr.pairs, &cr.pairsRefs, cr.pairItems, heads[i].empty() ? nullptr : &heads[i]});
                slot.push_back(4 * (int64_t)i + 1);
            }
        }
        std::vector<TrafficEstimate> est;
        const unsigned hw = std::thread::hardware_concurrency();
        plan_traffic_many(jobs, *g, est, (int)std::min<unsigned>(hw ? hw : 1, 16));
        for (size_t j = 0; j < jobs.size(); ++j) {
            out[slot[j]] = est[j].bytes;
            out[slot[j] + 2] = est[j].nodes;
        }
        return 0;
    } catch (const std::exception& e) {
        std::fprintf(stderr, "[gpu_rules] gpu_plan_traffic error: %s\n", e.what());
        return -1;
    }
}

int gpu_set_tuning(void* ctx, const char* key, int64_t value) {
    Context* C = as_ctx(ctx);
    if (!C || !key) return -1;
    std::string k(key);
    if (k == "budget") C->tune.budget = (int32_t)std::max<int64_t>(1, value);
    else if (k == "thread_waves_per_cu") C->threadWavesPerCU = (int)std::max<int64_t>(1, value);
    else if (k == "wave_waves_per_cu") {
        C->waveWavesPerCU = (int)std::max<int64_t>(1, value);
        delete C->lastBatch;
        C->lastBatch = nullptr;
        if (!C->ensure_dedup()) return -1;
    } else if (k == "use_graph") C->useGraph = value != 0;
    else if (k == "overlap") C->overlap = value != 0;
    else if (k == "fuse_support") C->fuseSupport = value != 0;
    else if (k == "share_bodies") C->shareBodies = value != 0;
    else if (k == "chunks_per_wave") C->chunksPerWave = (int)std::min<int64_t>(1024, std::max<int64_t>(1, value));
    else if (k == "dedup") {
        if (value < 0 || value > 3) return -1;
        C->dedupMode = (int)value;
        delete C->lastBatch;
        C->lastBatch = nullptr;
        if (!C->ensure_dedup()) return -1;
    } else if (k == "wave_engine") {
        // wave-mode search: 0 pool of entries, 1 level-synchronous
        if (value < 0 || value > 1) return -1;
        C->waveEngine = (int)value;
        delete C->lastBatch;
        C->lastBatch = nullptr;
        if (!C->ensure_dedup()) return -1;
    } else if (k == "giant_after") {
        C->giantAfter = (int)std::max<int64_t>(0, value);
    } else if (k == "per_rule_giant_work") {
        // per-rule launches cut giant seeds for rules of at least this estimated work (0: every rule)
        C->perRuleGiantWork = (double)std::max<int64_t>(0, value);
    } else if (k == "phase_seed_cap") {
        // seeds one phase of a batch may hold (the heavy queue's range at most); one-shot
        // evaluations split a larger batch in halves, gpu_batch_create refuses it
        if (value < 1 || (uint64_t)value > kMaxPhaseSeeds) return -1;
        C->phaseSeedCap = (uint64_t)value;
    } else if (k == "serve_heavy_work") {
        // servers started afterwards: estimated work above which a request runs in a heavy batch (0: off)
        C->serveHeavyWork = (double)std::max<int64_t>(0, value);
    } else if (k == "serve_heavy_percent") {
        if (value < 1 || value > 100) return -1;
        C->serveHeavyPercent = (int)value;
    } else if (k == "sort_cap") {
        // start the sort-unique buffers at this many keys (they still grow on overflow)
        if (value < kSortTile || value > (int64_t)C->sortCapMax) return -1;
        C->sortb.release();
        if (!C->sortb.ensure((uint64_t)value, (uint64_t)value, C->sortCapMax)) return -1;
    } else if (k == "noalloc_max_seeds") {
        // seeds per launch the allocation-free entry point accepts (at most what its scratch holds)
        if (value < 1 || (uint64_t)value > kNoAllocMaxSeeds) return -1;
        C->noallocMaxSeeds = (uint64_t)value;
    } else if (k == "sort_cap_max") {
        // largest key capacity the sort-unique buffers may grow to; a pass
        // that emits more keys fails (default and upper limit 2^31)
        if (value < kSortTile || value > (int64_t)kSortCapMax) return -1;
        C->sortCapMax = (uint32_t)value;
        if (C->sortb.cap > C->sortCapMax) C->sortb.release();
    } else return -1;
    if (C->lastBatch) { delete C->lastBatch; C->lastBatch = nullptr; }
    return 0;
}

int64_t gpu_get_tuning(void* ctx, const char* key) {
    Context* C = as_ctx(ctx);
    if (!C || !key) return -1;
    std::string k(key);
    if (k == "budget") return C->tune.budget;
    if (k == "thread_waves_per_cu") return C->threadWavesPerCU;
    if (k == "wave_waves_per_cu") return C->waveWavesPerCU;
    if (k == "use_graph") return C->useGraph ? 1 : 0;
    if (k == "overlap") return C->overlap ? 1 : 0;
    if (k == "fuse_support") return C->fuseSupport ? 1 : 0;
    if (k == "share_bodies") return C->shareBodies ? 1 : 0;
    if (k == "last_shared") return C->lastShared;
    if (k == "last_support_searches") return C->lastSearches[0];
    if (k == "last_pair_searches") return C->lastSearches[1];
    if (k == "chunks_per_wave") return C->chunksPerWave;
    if (k == "fuse_active") return C->fuseActive(C->pairsDedup()) ? 1 : 0;
    if (k == "num_cus") return C->numCUs;
    if (k == "lds_per_cu") return (int64_t)C->ldsPerCU;
    if (k == "lds_per_block") return (int64_t)C->ldsPerBlock;
    if (k == "wave_lds_static") return (int64_t)wave_static_lds(4, C->waveEngine);
    if (k == "wave_lds_blocks_per_cu") return wave_lds_blocks_per_cu(4, (size_t)C->rowWords * 4, C->waveEngine);
    if (k == "wave_slots") return C->waveSlots;
    if (k == "dedup") return C->dedupMode;
    if (k == "sort_active") return C->sortActive() ? 1 : 0;
    if (k == "pairs_dedup") return C->pairsDedup();
    if (k == "giant_after") return C->giantAfter;
    if (k == "wave_engine") return C->waveEngine;
    if (k == "serve_heavy_work") return (int64_t)C->serveHeavyWork;
    if (k == "per_rule_giant_work") return (int64_t)C->perRuleGiantWork;
    if (k == "phase_seed_cap") return (int64_t)C->phaseSeedCap;
    if (k == "serve_heavy_percent") return C->serveHeavyPercent;
    if (k == "last_giants") return C->lastGiants;
    if (k == "last_fused") return C->lastFused;
    // wave-mode counters of the last batch: wave_<phase>_<seeds|steps|donations|leaf_draws|draws|busy_lane_steps|draw_lanes>
    static const char* kStatNames[kWaveStats] = {"seeds", "steps", "donations", "leaf_draws", "draws",
                                                 "busy_lane_steps", "draw_lanes", "cycles_take", "cycles_setup",
                                                 "cycles_leaf_draw", "cycles_draw", "cycles_own", "cycles_step_end",
                                                 "cycles_seed_end", "cycles_sub_items"};
    for (int ph = 0; ph < 2; ++ph)
        for (int q = 0; q < kWaveStats; ++q)
            if (k == std::string(ph ? "wave_pairs_" : "wave_support_") + kStatNames[q])
                return (int64_t)C->lastWaveStats[ph * kWaveStats + q];
    if (k == "sort_cap") return C->sortb.cap;
    if (k == "sort_cap_max") return C->sortCapMax;
    if (k == "noalloc_max_seeds") return (int64_t)C->noallocMaxSeeds;
    return -1;
}

// ---------------------------------------------------------- host handles

void* gpu_host_load_turtle(const char* ttlPath) {
    if (!ttlPath) return nullptr;
    try {
        auto* H = new HostHandle();
        H->magic = kMagicHost;
        std::string err;
        if (!load_turtle(ttlPath, H->g, err)) {
            std::fprintf(stderr, "[gpu_rules] failed to parse TTL: %s (%s)\n", ttlPath, err.c_str());
            delete H;
            return nullptr;
        }
        H->tr.reset(new TermResolver(H->g));
        return H;
    } catch (const std::exception& e) {
        std::fprintf(stderr, "[gpu_rules] gpu_host_load_turtle error: %s\n", e.what());
        return nullptr;
    }
}

void* gpu_host_load_triples(const int32_t* s, const int32_t* p, const int32_t* o, int64_t n, int64_t nTerms) {
    if ((!s || !p || !o) && n > 0) return nullptr;
    try {
        for (int64_t i = 0; i < n; ++i) {
            if (s[i] < 0 || p[i] < 0 || o[i] < 0 || s[i] >= nTerms || p[i] >= nTerms || o[i] >= nTerms) {
                std::fprintf(stderr, "[gpu_rules] triple %lld has a term id outside [0, %lld)\n", (long long)i,
                             (long long)nTerms);
                return nullptr;
            }
        }
        auto* H = new HostHandle();
        H->magic = kMagicHost;
        std::vector<int32_t> S(s, s + n), P(p, p + n), O(o, o + n);
        build_graph(H->g, S, P, O, nTerms);
        H->tr.reset(new TermResolver(H->g));
        return H;
    } catch (const std::exception& e) {
        std::fprintf(stderr, "[gpu_rules] gpu_host_load_triples error: %s\n", e.what());
        return nullptr;
    }
}

void* gpu_host_load_snapshot(const char* path) {
    if (!path) return nullptr;
    try {
        auto* H = new HostHandle();
        H->magic = kMagicHost;
        std::string err;
        if (!load_snapshot(path, H->g, err)) {
            std::fprintf(stderr, "[gpu_rules] failed to load snapshot: %s (%s)\n", path, err.c_str());
            delete H;
            return nullptr;
        }
        H->tr.reset(new TermResolver(H->g));
        return H;
    } catch (const std::exception& e) {
        std::fprintf(stderr, "[gpu_rules] gpu_host_load_snapshot error: %s\n", e.what());
        return nullptr;
    }
}

void gpu_host_free(void* host) {
    if (host && static_cast<Handle*>(host)->magic == kMagicHost) delete static_cast<HostHandle*>(host);
}

int gpu_host_graph_info(void* handle, int64_t* nTerms, int32_t* nPredicates, int64_t* nTriples) {
    const HostGraph* g = as_graph(handle);
    if (!g) return -1;
    if (nTerms) *nTerms = g->nE;
    if (nPredicates) *nPredicates = g->P();
    if (nTriples) *nTriples = g->T();
    return 0;
}

int64_t gpu_host_term_id(void* handle, const char* term) {
    const HostGraph* g = as_graph(handle);
    if (!g || !term) return -1;
    return g->terms.find(term);
}

int64_t gpu_host_term_string(void* handle, int64_t id, char* buf, int64_t cap) {
    TermResolver* tr = as_resolver(handle);
    if (!tr || !buf || cap < 1 || id < 0 || id > INT32_MAX) return -1;
    const std::string* s = tr->name((int32_t)id);
    if (!s) return -1;
    const size_t k = std::min<size_t>(s->size(), (size_t)cap - 1);
    std::memcpy(buf, s->data(), k);
    buf[k] = '\0';
    return (int64_t)s->size();
}

int gpu_host_parse_rule(void* handle, const char* rule, int32_t isJson, int32_t* words) {
    TermResolver* tr = as_resolver(handle);
    if (!tr || !rule || !words) return -1;
    try {
        RuleIR r = isJson ? parse_rule_json(rule, *tr) : parse_rule_text(rule, *tr);
        if (r.body.size() > (size_t)kMaxBody) return -1;
        rule_to_words(r, words);
        return 0;
    } catch (const std::exception&) {
        return -1;
    }
}

int32_t gpu_host_count_action(int32_t which, int64_t count, int64_t cap, int64_t capMax) {
    if (count < 0 || cap < 0 || capMax < 0) return -1;
    if (which == 0) return sort_count_action((uint64_t)count, (uint64_t)cap, (uint64_t)capMax);
    if (which == 1) return heavy_count_action((uint64_t)count, (uint64_t)cap);
    return -1;
}

int64_t gpu_host_support_threshold(int64_t headSize, int64_t minSupport, double minHeadCoverage) {
    Thresholds t;
    if (headSize < 0 || !thresholds_of(minSupport, minHeadCoverage, 0, t)) return -1;
    return (int64_t)support_threshold(headSize, t);
}

int gpu_host_rules_work(void* handle, const int32_t* words, int32_t n, double* work) {
    const HostGraph* g = as_graph(handle);
    if (!g || n < 0 || (n > 0 && (!words || !work))) return -1;
    const RowStats st = row_stats(*g);
    Tuning t;
    for (int32_t i = 0; i < n; ++i)
        work[i] = estimate_work(compile_rule(rule_from_words(words + (size_t)i * GPU_RULE_WORDS), *g, t), *g, st);
    return 0;
}

int gpu_host_body_groundings(void* handle, const int32_t* words, int32_t n, int64_t* out) {
    const HostGraph* g = as_graph(handle);
    if (!g || n < 0 || (n > 0 && (!words || !out))) return -1;
    try {
        std::atomic<int32_t> next{0};
        std::atomic<bool> failed{false};
        auto work = [&]() {
            try {
                for (;;) {
                    const int32_t i = next.fetch_add(1);
                    if (i >= n) break;
                    out[i] = body_groundings(rule_from_words(words + (size_t)i * GPU_RULE_WORDS), *g);
                }
            } catch (...) {
                failed = true;
            }
        };
        const unsigned hw = std::thread::hardware_concurrency();
        const int nt = (int)std::max(1u, std::min(hw ? hw : 1u, 16u));
        std::vector<std::thread> pool;
        for (int k = 1; k < nt; ++k) pool.emplace_back(work);
        work();
        for (auto& th : pool) th.join();
        return failed ? -1 : 0;
    } catch (const std::exception& e) {
        std::fprintf(stderr, "[gpu_rules] gpu_host_body_groundings error: %s\n", e.what());
        return -1;
    }
}

int64_t gpu_host_refine_rules(const int32_t* parents, int32_t n, const int32_t* preds, int32_t nPreds,
                              int32_t maxAtoms, int32_t* out, int64_t cap) {
    try {
        return refine_rules(parents, n, preds, nPreds, maxAtoms, out, cap);
    } catch (const std::exception& e) {
        std::fprintf(stderr, "[gpu_rules] gpu_host_refine_rules error: %s\n", e.what());
        return -1;
    }
}

int gpu_host_plan_info(void* handle, const int32_t* words, int64_t* info) {
    const HostGraph* g = as_graph(handle);
    if (!g || !words || !info) return -1;
    Tuning t;
    CompiledRule cr = compile_rule(rule_from_words(words), *g, t);
    info[0] = cr.headSize;
    info[1] = cr.headSupport;
    info[2] = cr.status;
    info[3] = cr.supportLaunch;
    info[4] = cr.pairsLaunch;
    info[5] = cr.supItems;
    info[6] = cr.pairItems;
    return 0;
}

}  // extern "C"
