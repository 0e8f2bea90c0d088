// serve.cpp -- continuous batching of rule requests over a resident graph.
//
// A scheduler thread owns all device work of a server. Requests are queued
// by gpu_serve_submit from any thread. Whenever an execution slot is free,
// the scheduler takes the pending requests (up to maxBatch, in arrival
// order), admits as many as fit the HBM budget left by the batches in flight
// (seed queues, plans, chunk lists, counters), compiles them into one batch
// and launches its device pass on the slot's own stream with the slot's own
// dedup rows -- so batches from different slots overlap on the device while
// new requests keep arriving. Finished slots are found by event query; their
// results go to a completion queue read by gpu_serve_poll.
//
// Requests whose estimated search work (estimate_work) passes the context's
// serveHeavyWork are taken out of the light batches into batches of their
// own: at most one such heavy batch is in flight, launched on
// serveHeavyPercent of the grid, so a request of seconds delays neither the
// light requests it would have shared a batch with nor the light batches
// that run beside it.
//
// Counts are exact integers computed per rule, so a rule's result does not
// depend on which batch or slot it ran in.
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cstdlib>
#include <cstring>
#include <deque>
#include <mutex>
#include <thread>

#include "engine_internal.hpp"

namespace gr {

namespace {

constexpr uint32_t kMagicServer = 0x47525331;   // "GRS1"
constexpr int64_t kHeavyOverdueNs = 20'000'000;  // a heavy request waits at most this long for a spare slot

int64_t mono_ns() {
    return std::chrono::duration_cast<std::chrono::nanoseconds>(
               std::chrono::steady_clock::now().time_since_epoch())
        .count();
}

struct Request {
    uint64_t tag;
    GpuRuleDesc rule;
    int64_t submitNs;
};

// A request with its compiled rule and estimated work (the scheduler's own).
struct Compiled {
    Request r;
    RuleIR ir;
    CompiledRule cr;
    double work = 0;
};

struct Slot {
    hipStream_t st = nullptr;
    hipEvent_t done = nullptr;
    DevBuf dedup;
    int64_t rowWords = 0;
    int waveSlots = 0;
    SortBufs sortb;          // sort-unique dedup path (when the context uses it)
    GiantBufs giants;        // giant-seed pass
    bool sort = false;
    int mode = kDedupRows;   // DedupMode of the body-pair wave-mode kernel
    Batch batch;
    std::vector<Request> reqs;
    std::vector<GpuRuleResult64> out;
    int64_t reserved = 0;
    int64_t admitNs = 0;
    bool busy = false;
    bool heavy = false;      // the batch in flight is a heavy one (reduced grid)
    bool retired = false;    // its buffers could not be regrown: no further batches
    ~Slot() {
        if (st) (void)hipStreamSynchronize(st);
        batch.free_device();   // stream-ordered on st: before the stream goes
        if (st) (void)hipStreamSynchronize(st);
        dedup.release();
        sortb.release();
        giants.release();
        if (done) (void)hipEventDestroy(done);
        if (st) (void)hipStreamDestroy(st);
    }
};

// Bytes a request needs on the device while in flight: seed queue entries
// for both phases, two plans and two conflict tables per phase, chunk
// descriptors, counters.
int64_t request_bytes(const CompiledRule& cr) {
    int64_t items = 0, chunks = 0;
    if (cr.supportLaunch) {
        items += cr.supItems;
        chunks += (cr.supItems + kMinChunkSeeds - 1) / kMinChunkSeeds;
    }
    if (cr.pairsLaunch) {
        items += cr.pairItems;
        chunks += (cr.pairItems + kMinChunkSeeds - 1) / kMinChunkSeeds;
    }
    return items * (int64_t)sizeof(HeavyItem) + chunks * (int64_t)sizeof(Chunk) +
           4 * (int64_t)(sizeof(Plan) + sizeof(ConflictRefs)) + 16;
}

struct Server : Handle {
    Context* C = nullptr;
    int32_t maxBatch = 1024;
    int32_t flags = 0;
    int64_t budget = 0;      // bytes for in-flight batches
    int64_t reservedNow = 0;
    int64_t slotBytes = 0;   // HBM each slot holds for its whole life (dedup, sort and giant buffers)
    std::vector<std::unique_ptr<Slot>> slots;
    std::mutex mu;
    std::condition_variable cv;      // new requests / stop
    std::condition_variable doneCv;  // new completions
    std::deque<Request> pending;
    std::deque<Compiled> heavyPending;   // scheduler thread only
    double heavyWork = 0;                // 0: no heavy batches
    int heavyPercent = 100;
    std::deque<GpuServeCompletion> done;
    std::thread th;
    std::atomic<bool> stopping{false};
    std::atomic<bool> failed{false};
    // statistics
    std::atomic<int64_t> nSubmitted{0}, nAdmitted{0}, nCompleted{0}, nBatches{0}, nRefused{0}, nDeferred{0};
    std::atomic<int64_t> maxBatchSeen{0}, inFlightPeak{0}, nHeavy{0}, nHeavyBatches{0};

    bool init(int32_t maxInFlight, int64_t memBudget) {
        // Each slot needs a hardware queue of its own: streams beyond
        // GPU_MAX_HW_QUEUES (HIP's default 4) share queues, and a slot whose
        // queue another slot's batch holds waits for it.
        const char* q = std::getenv("GPU_MAX_HW_QUEUES");
        const int queues = q && std::atoi(q) > 0 ? std::atoi(q) : 4;
        if (maxInFlight + 2 > queues)
            std::fprintf(stderr, "[gpu_rules] serve: %d slots and the context's 2 streams exceed GPU_MAX_HW_QUEUES=%d; "
                                 "slots sharing a hardware queue wait for each other (set GPU_MAX_HW_QUEUES >= %d "
                                 "before the process starts)\n", maxInFlight, queues, maxInFlight + 2);
        if (memBudget <= 0) {
            size_t f = 0, t = 0;
            if (hipMemGetInfo(&f, &t) != hipSuccess) return false;
            memBudget = (int64_t)(f / 2);
        }
        budget = memBudget;
        heavyWork = C->serveHeavyWork;
        heavyPercent = C->serveHeavyPercent;
        slotBytes = 0;
        for (int i = 0; i < maxInFlight; ++i) {
            auto S = std::make_unique<Slot>();
            GR_HIP(hipStreamCreateWithFlags(&S->st, hipStreamNonBlocking));
            GR_HIP(hipEventCreateWithFlags(&S->done, hipEventDisableTiming));
            S->rowWords = C->rowWords;
            S->waveSlots = C->waveSlots;
            S->mode = C->pairsDedup();
            S->sort = S->mode == kDedupSort;
            // only the HBM-row mode uses the rows; one keeps the kernels' pointer valid otherwise
            size_t bytes = (size_t)(S->mode == kDedupRows ? S->waveSlots : 1) * S->rowWords * 4;
            if (!S->dedup.ensure(bytes)) return false;
            GR_HIP(hipMemset(S->dedup.p, 0, bytes));
            if (S->sort) {
                if (!S->sortb.ensure(1, kSortCapMin, C->sortCapMax)) return false;
                bytes += S->sortb.bytes();
            }
            if (!S->giants.ensure(S->rowWords)) return false;
            bytes += S->giants.bytes();
            S->batch.magic = kMagicBatch;
            S->batch.ctx = C;
            budget -= (int64_t)bytes;
            slotBytes = (int64_t)bytes;
            slots.push_back(std::move(S));
        }
        if (budget <= 0) {
            std::fprintf(stderr, "[gpu_rules] serve: HBM budget smaller than the slots' dedup rows\n");
            return false;
        }
        th = std::thread([this] { loop(); });
        return true;
    }

    void complete(Request& r, const GpuRuleResult64& res, int64_t admitNs, int32_t batchSize) {
        GpuServeCompletion c;
        std::memset(&c, 0, sizeof(c));
        c.tag = r.tag;
        c.result = res;
        c.submitNs = r.submitNs;
        c.admitNs = admitNs;
        c.completeNs = mono_ns();
        c.batchSize = batchSize;
        std::lock_guard<std::mutex> g(mu);
        done.push_back(c);
        nCompleted++;
    }

    // Finished slots -> completion queue. A batch whose pass failed, whose
    // escalated seeds passed the heavy queue, or whose body-pair keys passed
    // the largest sort capacity completes with status kRuleError for every
    // request; a slot whose sort buffers could not be regrown is retired.
    int reap() {
        int busy = 0;
        for (auto& S : slots) {
            if (!S->busy) continue;
            hipError_t q = hipEventQuery(S->done);
            if (q == hipErrorNotReady) { ++busy; continue; }
            bool bad = false;
            if (q != hipSuccess) {
                log_err("serve: batch failed", q);
                bad = true;
            } else if (!batch_heavy_ok(S->batch)) {
                bad = true;
            } else if (S->sort) {
                const uint64_t emitted = (uint64_t)batch_emitted(S->batch);
                const int32_t act = sort_count_action(emitted, S->sortb.cap, C->sortCapMax);
                if (act == kCountsRegrow) {
                    // body-pair keys overflowed the slot's sort buffers: grow them and run the batch again
                    const size_t was = S->sortb.bytes();
                    if (S->sortb.ensure(emitted, kSortCapMin, C->sortCapMax)) {
                        budget -= (int64_t)(S->sortb.bytes() - was);
                        if (launch(*S)) {
                            ++busy;
                            continue;
                        }
                    } else {
                        std::fprintf(stderr, "[gpu_rules] serve: sort buffers of a slot could not grow; slot retired\n");
                        S->retired = true;
                    }
                    bad = true;
                } else if (act == kCountsFail) {
                    std::fprintf(stderr, "[gpu_rules] serve: body-pair keys of a batch (%llu) exceed the sort "
                                         "capacity %u\n", (unsigned long long)emitted, S->sortb.cap);
                    bad = true;
                }
            }
            if (bad) failed = true;
            S->out.resize(S->reqs.size());
            batch_collect(S->batch, S->out.data());
            for (size_t i = 0; i < S->reqs.size(); ++i) {
                if (bad) S->out[i].status = kRuleError;
                complete(S->reqs[i], S->out[i], S->admitNs, (int32_t)S->reqs.size());
            }
            S->reqs.clear();
            reservedNow -= S->reserved;
            S->reserved = 0;
            S->busy = false;
            S->heavy = false;
            doneCv.notify_all();
        }
        return busy;
    }

    // The slot's batch -> its stream, with the slot's dedup rows or sort
    // buffers. False if an enqueued call failed (logged).
    bool launch(Slot& S) {
        const hipError_t e = batch_enqueue(*C, S.batch, ExecTarget{S.st, S.dedup.as<uint32_t>(), S.rowWords,
                                                                   std::max(1, S.waveSlots), nullptr,
                                                                   S.sort ? &S.sortb : nullptr, C->sortBlocks(),
                                                                   S.mode, &S.giants, C->giantAfter,
                                                                   S.heavy ? heavyPercent : 100});
        const hipError_t r = hipEventRecord(S.done, S.st);
        if (r != hipSuccess) log_err("serve: hipEventRecord", r);
        return e == hipSuccess && r == hipSuccess;
    }

    // Completes every pending request with `status` (no usable slot is left).
    void fail_pending(int32_t status) {
        std::deque<Request> all;
        {
            std::lock_guard<std::mutex> g(mu);
            all.swap(pending);
        }
        for (auto& c : heavyPending) all.push_back(c.r);
        heavyPending.clear();
        for (auto& r : all) {
            GpuRuleResult64 res;
            std::memset(&res, 0, sizeof(res));
            res.status = status;
            res.bodyPairs = -1;
            complete(r, res, mono_ns(), 0);
        }
    }

    // Fails a request that can never fit the HBM budget (status 3).
    void refuse(Request& r) {
        GpuRuleResult64 res;
        std::memset(&res, 0, sizeof(res));
        res.status = 3;
        res.bodyPairs = -1;
        complete(r, res, mono_ns(), 0);
        nRefused++;
    }

    void fail_all(std::vector<Compiled>& xs) {
        for (auto& x : xs) {
            GpuRuleResult64 res;
            std::memset(&res, 0, sizeof(res));
            res.status = kRuleError;
            res.bodyPairs = -1;
            complete(x.r, res, mono_ns(), 0);
        }
    }

    // Compiled requests -> a batch on slot S. False if nothing was launched.
    bool launch_batch(Slot* S, std::vector<Compiled>& admitted, int64_t need, bool heavy) {
        std::vector<RuleIR> rules;
        std::vector<int32_t> status;
        rules.reserve(admitted.size());
        for (auto& x : admitted) {
            rules.push_back(x.ir);
            status.push_back(kRuleOk);
        }
        S->heavy = heavy;
        if (!batch_build(*C, S->batch, rules, status, flags, S->st)) {
            failed = true;
            S->heavy = false;
            fail_all(admitted);
            return false;
        }
        if (!launch(*S)) {   // nothing to wait for: the batch's requests fail now
            failed = true;
            S->heavy = false;
            (void)hipStreamSynchronize(S->st);
            fail_all(admitted);
            return false;
        }
        S->reqs.clear();
        for (auto& x : admitted) S->reqs.push_back(x.r);
        S->reserved = need;
        S->admitNs = mono_ns();
        S->busy = true;
        reservedNow += need;
        nAdmitted += (int64_t)S->reqs.size();
        nBatches++;
        if (heavy) nHeavyBatches++;
        if ((int64_t)S->reqs.size() > maxBatchSeen) maxBatchSeen = (int64_t)S->reqs.size();
        int64_t inflight = 0;
        for (auto& s : slots) inflight += s->busy ? 1 : 0;
        if (inflight > inFlightPeak) inFlightPeak = inflight;
        return true;
    }

    // Pending requests -> a batch on a free slot. Returns true if launched.
    // A heavy batch (from heavyPending) goes first when none is in flight and
    // another slot stays free for light batches (or nothing light waits,
    // there is one slot only, or the oldest heavy request has waited
    // kHeavyOverdueNs); light batches divert heavy requests into heavyPending
    // as they compile them.
    bool admit() {
        Slot* S = nullptr;
        int usable = 0, freeSlots = 0;
        bool heavyInFlight = false;
        for (auto& s : slots) {
            if (s->retired) continue;
            ++usable;
            if (s->busy) heavyInFlight |= s->heavy;
            else {
                ++freeSlots;
                if (!S) S = s.get();
            }
        }
        if (usable == 0) {   // every slot retired: nothing can run any more
            failed = true;
            fail_pending(kRuleError);
            return false;
        }
        if (!S) return false;
        bool lightWaiting;
        {
            std::lock_guard<std::mutex> g(mu);
            lightWaiting = !pending.empty();
        }
        const bool idle = reservedNow == 0;
        // a heavy batch from heavyPending, when the policy allows one now
        auto try_heavy = [&](bool light) -> bool {
            if (heavyInFlight || heavyPending.empty()) return false;
            const bool overdue = mono_ns() - heavyPending.front().r.submitNs > kHeavyOverdueNs;
            if (!(freeSlots >= 2 || !light || usable == 1 || overdue)) return false;
            std::vector<Compiled> admitted;
            int64_t need = 0;
            while (!heavyPending.empty() && (int32_t)admitted.size() < maxBatch) {
                const int64_t b = request_bytes(heavyPending.front().cr);
                if (reservedNow + need + b > budget) {
                    if (idle && admitted.empty()) {   // can never fit: refuse it
                        refuse(heavyPending.front().r);
                        heavyPending.pop_front();
                        continue;
                    }
                    break;   // wait for memory to free up
                }
                need += b;
                admitted.push_back(std::move(heavyPending.front()));
                heavyPending.pop_front();
            }
            return !admitted.empty() && launch_batch(S, admitted, need, true);
        };
        if (try_heavy(lightWaiting)) return true;
        std::vector<Request> take;
        {
            std::lock_guard<std::mutex> g(mu);
            while (!pending.empty() && (int32_t)take.size() < maxBatch) {
                take.push_back(pending.front());
                pending.pop_front();
            }
        }
        if (take.empty()) return false;
        std::vector<Compiled> admitted;
        std::vector<Request> back;
        int64_t need = 0;
        for (auto& r : take) {
            if (!back.empty()) { back.push_back(r); continue; }
            Compiled x;
            x.r = r;
            x.ir = rule_from_words(r.rule.words);
            x.cr = compile_rule(x.ir, C->g, C->tune);
            if (heavyWork > 0) {
                x.work = estimate_work(x.cr, C->g, C->rowStats);
                if (x.work > heavyWork) {
                    nHeavy++;
                    heavyPending.push_back(std::move(x));
                    continue;
                }
            }
            const int64_t b = request_bytes(x.cr);
            if (reservedNow + need + b > budget) {
                if (idle && admitted.empty()) {   // can never fit: refuse it
                    refuse(r);
                    continue;
                }
                back.push_back(r);   // wait for memory to free up
                continue;
            }
            need += b;
            admitted.push_back(std::move(x));
        }
        if (!back.empty()) {
            nDeferred += (int64_t)back.size();
            std::lock_guard<std::mutex> g(mu);
            for (auto it = back.rbegin(); it != back.rend(); ++it) pending.push_front(*it);
        }
        if (admitted.empty()) {   // every request taken was heavy (or must wait): the heavy queue's turn
            bool lightLeft;
            {
                std::lock_guard<std::mutex> g(mu);
                lightLeft = !pending.empty();
            }
            return try_heavy(lightLeft);
        }
        return launch_batch(S, admitted, need, false);
    }

    void loop() {
        for (;;) {
            const int busy = reap();
            bool launched = false;
            while (admit()) launched = true;
            bool havePending;
            {
                std::lock_guard<std::mutex> g(mu);
                havePending = !pending.empty() || !heavyPending.empty();
            }
            if (stopping && !havePending && busy == 0) {
                bool anyBusy = false;
                for (auto& s : slots) anyBusy |= s->busy;
                if (!anyBusy) break;
            }
            if (launched) continue;
            if (busy > 0 || havePending) {
                std::this_thread::sleep_for(std::chrono::microseconds(20));
            } else {
                std::unique_lock<std::mutex> lk(mu);
                cv.wait_for(lk, std::chrono::milliseconds(2), [this] { return !pending.empty() || stopping.load(); });
            }
        }
    }
};

Server* as_server(void* h) {
    if (!h || static_cast<Handle*>(h)->magic != kMagicServer) return nullptr;
    return static_cast<Server*>(h);
}

}  // namespace
}  // namespace gr

using namespace gr;

extern "C" {

int64_t gpu_monotonic_ns(void) { return mono_ns(); }

void* gpu_serve_start(void* ctx, int32_t maxBatch, int32_t maxInFlight, int64_t memBudgetBytes, int32_t flags) {
    if (!ctx || static_cast<Handle*>(ctx)->magic != kMagicCtx) return nullptr;
    auto* S = new Server();
    S->magic = kMagicServer;
    S->C = static_cast<Context*>(ctx);
    S->maxBatch = maxBatch > 0 ? maxBatch : 1024;
    S->flags = flags & GPU_RULES_UNGATED;
    if (!S->init(maxInFlight > 0 ? maxInFlight : 3, memBudgetBytes)) {
        delete S;
        return nullptr;
    }
    return S;
}

int gpu_serve_submit(void* srv, const GpuRuleDesc* rules, const uint64_t* tags, int32_t n) {
    Server* S = as_server(srv);
    if (!S || (!rules && n > 0) || (!tags && n > 0) || n < 0) return -1;
    if (S->stopping) return -1;
    const int64_t now = mono_ns();
    {
        std::lock_guard<std::mutex> g(S->mu);
        for (int32_t i = 0; i < n; ++i) S->pending.push_back(Request{tags[i], rules[i], now});
    }
    S->nSubmitted += n;
    S->cv.notify_one();
    return n;
}

int gpu_serve_poll(void* srv, GpuServeCompletion* out, int32_t maxOut, int32_t waitMs) {
    Server* S = as_server(srv);
    if (!S || (!out && maxOut > 0)) return -1;
    std::unique_lock<std::mutex> lk(S->mu);
    if (S->done.empty() && waitMs > 0)
        S->doneCv.wait_for(lk, std::chrono::milliseconds(waitMs), [S] { return !S->done.empty(); });
    int32_t k = 0;
    while (k < maxOut && !S->done.empty()) {
        out[k++] = S->done.front();
        S->done.pop_front();
    }
    return k;
}

int gpu_serve_stats(void* srv, int64_t* stats) {
    Server* S = as_server(srv);
    if (!S || !stats) return -1;
    stats[0] = S->nSubmitted;
    stats[1] = S->nAdmitted;
    stats[2] = S->nCompleted;
    stats[3] = S->nBatches;
    stats[4] = S->nRefused;
    stats[5] = S->nDeferred;
    stats[6] = S->maxBatchSeen;
    stats[7] = S->inFlightPeak;
    return S->failed ? 1 : 0;
}

int gpu_serve_heavy_stats(void* srv, int64_t* heavyRequests, int64_t* heavyBatches) {
    Server* S = as_server(srv);
    if (!S) return -1;
    if (heavyRequests) *heavyRequests = S->nHeavy;
    if (heavyBatches) *heavyBatches = S->nHeavyBatches;
    return 0;
}

int64_t gpu_serve_slot_bytes(void* srv) {
    Server* S = as_server(srv);
    return S ? S->slotBytes : -1;
}

void gpu_serve_stop(void* srv) {
    Server* S = as_server(srv);
    if (!S) return;
    S->stopping = true;
    S->cv.notify_all();
    if (S->th.joinable()) S->th.join();
    delete S;
}

}  // extern "C"
