/* gpu_rules_api.h -- C ABI of libgpu_rules.so, the MI355X (gfx950) rule
 * support / confidence counting library.
 *
 * The first five declarations are the reference's interface, unchanged in
 * name, signature, struct layout and return conventions, so the reference's
 * JNA binding (GPU_03/GpuRulesLib.scala:23-58, Native.load("gpu_rules"))
 * binds this library as it is. Everything after them is additive.
 *
 * Self-contained as plain C (the reference header relies on the includer
 * for size_t; this one includes <stddef.h> and <stdint.h>).
 */
#ifndef GPU_RULES_API_H
#define GPU_RULES_API_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Result of one rule (replaces GPU_03/gpu_rules_api.h:21-30; same fields,
 * order and 56-byte layout). Counts are the library's exact 64-bit counts
 * truncated to int, as the reference does (GPU_03/gpu_rules_lib.cu:435,525). */
typedef struct {
    int    support;       /* head triples whose body has an injective grounding        */
    int    headSize;      /* distinct triples of the head predicate                     */
    int    headSupport;   /* head groundings, constant-aware                            */
    double headCoverage;  /* support / headSize                                         */
    int    bodySize;      /* distinct (headS, headO) body bindings; 0 when support == 0 */
    double confidence;    /* support / bodySize, 0 when bodySize == 0                   */
    double supportMs;     /* device time of the support launches (ms)                   */
    double confidenceMs;  /* device time of the body-size launches (ms)                 */
} GpuRuleResult;

/* Replaces GPU_03/gpu_rules_api.h:43 (impl. GPU_03/gpu_rules_lib.cu:246-325).
 * Parses a Turtle file, builds the per-predicate SPO/POS adjacency and keeps
 * it resident in HBM. Returns an opaque context, or NULL (message on stderr
 * prefixed "[gpu_rules]") if the file cannot be read, no GPU is present or
 * HBM is short. */
void* gpu_load_graph(const char* ttlPath);

/* Replaces GPU_03/gpu_rules_api.h:59 (impl. GPU_03/gpu_rules_lib.cu:330-543).
 * Evaluates one rule line "( ?a <p> ?c ) ^ ... => ( ?a <q> ?b )" with one
 * support launch and (if support > 0) one body-size launch, each followed by
 * a device sync and a device-to-host copy. *out is zeroed first. Returns -1
 * if an argument is NULL or the rule cannot be parsed, else 0 -- including
 * rules with an unsupported shape (body outside 1..8 atoms, > 16 variables),
 * which report headSize/headSupport and zero measures. Not thread-safe per
 * context. */
int gpu_evaluate_rule(void* ctx, const char* ruleText, GpuRuleResult* out);

/* Replaces GPU_03/gpu_rules_api.h:72 (impl. GPU_03/gpu_rules_lib.cu:550-565).
 * Same as gpu_evaluate_rule for one rule given as an RDFRules JSON object
 * {"body":[{"subject":{"type":..,"value":..},"predicate":..,"object":{..}}],"head":{..}}. */
int gpu_evaluate_rule_json(void* ctx, const char* ruleJson, GpuRuleResult* out);

/* Replaces GPU_03/gpu_rules_api.h:79 (impl. GPU_03/gpu_rules_lib.cu:570-574).
 * Releases the context and all its device memory. NULL-safe. */
void gpu_free_graph(void* ctx);

/* Replaces GPU_03/gpu_rules_api.h:87 (impl. GPU_03/gpu_rules_lib.cu:237-241).
 * Free and total HBM of the current device; 0 on success, -1 if no device. */
int gpu_query_vram(size_t* freeBytesOut, size_t* totalBytesOut);

/* ------------------------------------------------------------------------
 * Additive API (no counterpart in the reference).
 * ------------------------------------------------------------------------ */

/* No counterpart in the reference. For callers that cannot allocate device
 * memory -- code under HIP graph capture, tight serving or mining loops,
 * harnesses that forbid hipMalloc: gpu_evaluate_rule with the same
 * arguments, results and return conventions, but its kernels' scratch
 * (counters, control words, conflict tables and the heavy-seed queue) is a
 * fixed 16 MiB device array that every GPU holds from library load to
 * process end, so a call allocates nothing (gpu_evaluate_rule grows its
 * queue on demand). Returns -1 also when a launch of the rule has more seeds
 * than the array holds (2,097,088: the rule's head triples, or its body
 * atom's distinct head subjects) or than the context's "noalloc_max_seeds",
 * or when sort-unique dedup is in use and its key buffers are missing or
 * too small (they are never grown here). Calls on one GPU share the array:
 * calls that overlap on one GPU, from any context, race. */
int gpu_evaluate_rule_noalloc(void* ctx, const char* ruleText, GpuRuleResult* out);

/* Structured rule: words[0] = number of body atoms, words[1..3] = head
 * (s, p, o), words[4+3j .. 6+3j] = body atom j (s, p, o), j < 8. A term
 * >= 0 is a term id of the graph; a term < 0 is variable -(t+1). */
#define GPU_RULE_WORDS 28
typedef struct {
    int32_t words[GPU_RULE_WORDS];
} GpuRuleDesc;

/* Exact 64-bit result of one rule in a batch. */
typedef struct {
    int64_t support;
    int64_t headSize;
    int64_t headSupport;
    int64_t bodySize;      /* distinct body pairs, 0 when support == 0 (reference convention) */
    int64_t bodyPairs;     /* distinct body pairs regardless of support; -1 if not computed   */
    double  headCoverage;
    double  confidence;
    int32_t status;        /* 0 ok, 1 unsupported shape, 2 parse error or failed evaluation
                              (serving: device error, counts beyond the library's buffers),
                              3 refused by the serving memory budget */
    int32_t latencyUs;     /* GPU_RULES_PER_RULE: host wall time of this rule's launches (us) */
} GpuRuleResult64;

/* Batch flags. */
#define GPU_RULES_UNGATED   1  /* also count body pairs of rules with support 0 (bodyPairs) */
#define GPU_RULES_PER_RULE  2  /* evaluate rule by rule with the per-rule launch sequence    */

/* Loads a graph from term-id triples: subject, predicate and object ids in
 * [0, nTerms). Predicates and entities share the id space, as in the
 * reference's IdMapper. Rule text can then only name terms by id-less
 * syntax, so such graphs are used with the structured entry points. */
void* gpu_load_graph_triples(const int32_t* s, const int32_t* p, const int32_t* o, int64_t n,
                             int64_t nTerms);

/* Binary graph snapshots. No counterpart in the reference, whose loader
 * parses the Turtle file and sorts the triples on every load
 * (GPU_03/gpu_rules_lib.cu:246-325, GPU_03/RdfIndexes.cpp:146-415).
 * gpu_save_snapshot writes the built graph of a context or host handle --
 * the predicate-grouped sorted triples in both directions, the term
 * dictionary in id order and the Turtle prefix lines -- to `path`; 0 or -1.
 * gpu_load_snapshot maps such a file, checks it (sizes, term ids, sort
 * order, both directions holding the same triples) and returns a context
 * exactly like gpu_load_graph on the original graph (same term ids, same
 * rule resolution, same results), or NULL. */
int gpu_save_snapshot(void* handle, const char* path);
void* gpu_load_snapshot(const char* path);

/* Evaluates n structured rules in one batched pass (persistent thread-mode
 * kernel over all rules' seeds, wave-mode kernel over escalated seeds,
 * captured in a HIP graph and replayed while the batch shape is unchanged).
 * Returns 0, or -1 on a NULL argument / device error. */
int gpu_evaluate_rules(void* ctx, const GpuRuleDesc* rules, int32_t n, int32_t flags, GpuRuleResult64* out);

/* Same for rule lines (isJson = 0) or JSON objects (isJson = 1). A rule that
 * fails to parse gets status 2; the others are evaluated. */
int gpu_evaluate_rules_text(void* ctx, const char* const* rules, int32_t n, int32_t isJson, int32_t flags,
                            GpuRuleResult64* out);

/* A compiled batch kept resident in HBM (plans, seed chunks, counters), so
 * the same rules can be evaluated again without host work: each
 * gpu_batch_run replays the batch's HIP graph (or launches eagerly when the
 * context's "use_graph" is 0) and fills out[n]. The batch must be freed
 * before its context. Flags as for gpu_evaluate_rules (GPU_RULES_UNGATED). */
void* gpu_batch_create(void* ctx, const GpuRuleDesc* rules, int32_t n, int32_t flags);
int gpu_batch_run(void* batch, GpuRuleResult64* out);
void gpu_batch_free(void* batch);

/* Support-pruned evaluation for a rule miner's refinement loop. No
 * counterpart in the reference's ABI: the reference evaluates every rule
 * fully and leaves pruning to the miner, whose support computation stops a
 * rule once it cannot reach minSupport (Rewritten_CPP/SupportCounting.cpp:
 * 66-90 restates RDFRules' computeSupport with minSupport = 1). Here the
 * miner's thresholds go to the device: every rule's support is counted
 * exactly, and body pairs only for rules with support >= minSupport and
 * headCoverage = support / headSize >= minHeadCoverage (the same double
 * quotient as the result's headCoverage); thresholds below 1 act as 1, the
 * gate of every gated batch. A rule below them keeps status 0 and its exact
 * support, headSize, headSupport and headCoverage, with bodyPairs = -1,
 * bodySize = 0 and confidence = 0. A rule whose support the body-pair pass
 * counts itself ("fuse_support"), or whose body-pair search other rules
 * with equal bodies share, has its body pairs searched whatever its support:
 * they are reported only if it meets the thresholds. A rule whose support
 * threshold exceeds all its head triples cannot meet it and gets no
 * body-pair search (unless a rule sharing that search may meet its own), so
 * a support-only evaluation (minSupport above every head size) runs the
 * support pass alone. flags: 0 or
 * GPU_RULES_PER_RULE; GPU_RULES_UNGATED, a NaN, negative or > 1
 * minHeadCoverage return -1 (NULL). */
int gpu_evaluate_rules_min(void* ctx, const GpuRuleDesc* rules, int32_t n, int32_t flags, int64_t minSupport,
                           double minHeadCoverage, GpuRuleResult64* out);
void* gpu_batch_create_min(void* ctx, const GpuRuleDesc* rules, int32_t n, int32_t flags, int64_t minSupport,
                           double minHeadCoverage);

/* Device time of the last gpu_evaluate_rules* call (ms) and its phases. */
int gpu_last_batch_stats(void* ctx, double* totalMs, double* supportMs, double* pairsMs, int64_t* heavySeeds);

/* Per rule of the last GPU_RULES_PER_RULE evaluation, the device time of its
 * support launches and of its body-size launches (ms, HIP events around
 * each rule's launches, as GPU_03/bench_gpu.cu:540-679 times each rule;
 * 0 where a launch did not run). Fills up to n entries of each array and
 * returns how many, or -1 on a bad argument. */
int gpu_last_rule_times(void* ctx, double* supportMs, double* confidenceMs, int32_t n);

/* Device time of the four batched kernels of the last eager (non-graph)
 * batch: support thread mode, support wave mode, pairs thread mode, pairs
 * wave mode (ms). Zero after a graph replay. */
int gpu_last_kernel_ms(void* ctx, double* ms4);

/* Sort-unique dedup of the last evaluation: device time of the radix sort +
 * unique pass (eager batches; 0 after a graph replay), keys the wave-mode
 * kernels emitted, and the buffers' current key capacity. 0 or -1. The
 * counterpart of the reference's overflow path (Thrust sort + unique,
 * GPU_03/gpu_rules_lib.cu:42-51,508-511), here an additive symbol. */
int gpu_last_dedup_stats(void* ctx, double* sortMs, int64_t* keysEmitted, int64_t* sortCap);

/* Host model of the HBM bytes each rule's searches read (the same plans,
 * walked serially): out[4*i] support bytes, out[4*i+1] body-pair bytes,
 * out[4*i+2], out[4*i+3] search levels entered. Works on a context or a
 * host handle; does no device work. On a context with fuse_support active,
 * a rule whose support the body-pair pass counts has no support bytes and
 * the body-pair bytes include the support lookups; with share_bodies, a rule
 * that takes another rule's body-pair count has no body-pair bytes. */
int gpu_plan_traffic(void* handle, const GpuRuleDesc* rules, int32_t n, int64_t* out);

/* Measurement yardstick: copies `bytes` (multiple of 16) between two fresh
 * device buffers with a plain streaming kernel (k_copy_u32): one untimed
 * pass, then the best device time in ms of five timed passes (-1 on error).
 * Used to scale hardware byte counters. */
double gpu_copy_probe(void* ctx, int64_t bytes);

/* Autotuner: times the given rules (a sample of the expected workload)
 * under candidate settings -- thread-mode work budget per seed (which
 * decides how much work stays one-lane-per-seed and how much moves to
 * wave64 mode), thread-mode waves per CU and work units per wave, wave-mode
 * waves per CU, head-object bitmaps
 * in HBM rows or LDS (auto dedup only), and the steps before a wave-mode
 * seed is cut into sub-items -- budget and thread-mode waves as a grid,
 * then thread-mode waves and work units per wave as a grid, the others one
 * at a time, then budget and waves again -- and keeps the
 * fastest in the context. If `log` is given it receives up to logCap rows of
 * (budget, thread waves/CU, wave waves/CU, ms, escalated seeds, pairs dedup
 * mode, giant_after); *logLen gets the row count. 0 or -1. */
int gpu_autotune(void* ctx, const GpuRuleDesc* rules, int32_t n, int32_t flags, double* log, int32_t logCap,
                 int32_t* logLen);

/* Tuning knobs: "budget" (thread-mode candidates per seed before a seed is
 * moved to wave mode), "thread_waves_per_cu", "wave_waves_per_cu",
 * "giant_after" (wave-mode steps before a seed is cut into sub-items that
 * all waves share; 0 never),
 * "share_bodies" (0/1, default 1: rules of a batch whose bodies are equal
 * up to a renaming that fixes the head variables share one body-pair
 * search, which also counts each one's fused support),
 * "chunks_per_wave" (thread-mode work units per thread-mode wave a batch
 * aims for, default 2; units hold 64 to 1,024 seeds),
 * "use_graph" (0/1), "overlap" (0/1, default 1: graph replays of ungated
 * batches run the support and body-pair phases side by side on two
 * streams), "fuse_support" (0/1, default 1: with bitmap dedup, batches
 * count the support of rules with distinct head variables and no body atom
 * that can ground to the head triple in the body-pair pass -- a head triple
 * (s, o) is supported iff (s, o) is a distinct body pair -- instead of a
 * support pass of their own; in gated batches such rules' body pairs are
 * then searched even where the support turns out 0), "dedup" (body-pair
 * dedup of wave-mode seeds; the
 * bitmaps index head objects by rank among the keys of one (direction,
 * predicate): 0 auto -- LDS bitmaps when one fits a block and at least 8
 * wave-mode waves per CU still fit, else HBM bitmap rows unless a row per CU
 * exceeds 8 GiB, else sort-unique; 1 HBM bitmap rows; 2 sort-unique; 3 LDS
 * bitmaps where they fit, else rows), "sort_cap" (initial key capacity of the
 * sort-unique buffers, >= 4096; they grow on overflow), "noalloc_max_seeds"
 * (seeds per launch gpu_evaluate_rule_noalloc accepts, 1 .. 2,097,088,
 * default all its scratch holds), "sort_cap_max" (the
 * largest capacity they may grow to, 4096 .. 2^31, default 2^31: a pass that
 * emits more keys fails -- gpu_evaluate_rules* and gpu_batch_run return -1,
 * served requests complete with status 2), "phase_seed_cap" (seeds one
 * phase of a batch may hold, 1 .. 2^32 - 16, default the largest: the
 * gpu_evaluate_rules* calls evaluate a larger set of rules in halves,
 * recursively, gpu_batch_create* return NULL). Returns 0, or -1 for
 * an unknown key or value. gpu_get_tuning also reads "num_cus",
 * "wave_slots", "sort_active" (0/1), "sort_cap", "pairs_dedup" (the
 * DedupMode in use: 0 HBM rows, 1 sort, 2 LDS), "fuse_active" (0/1: the
 * dedup in use lets fuse_support act), "last_giants" (seeds cut into
 * sub-items by the last batch), "last_fused" (rules whose support the last
 * batch counted in its body-pair pass), "last_shared" (rules of the last
 * batch that took another rule's body-pair count), "last_support_searches"
 * and "last_pair_searches" (support and body-pair plans the last batch ran:
 * rules whose support the pairs pass counts, and rules that share another
 * rule's body-pair search, need none of their own). */
int gpu_set_tuning(void* ctx, const char* key, int64_t value);
int64_t gpu_get_tuning(void* ctx, const char* key);

/* ---- serving: continuous batching over a resident graph ----
 *
 * A server owns a scheduler thread and maxInFlight execution slots (stream +
 * dedup rows each). Whenever a slot is free it takes the pending requests in
 * arrival order (up to maxBatch), admits those that fit the HBM budget left
 * by the batches in flight (memBudgetBytes; <= 0 means half the free HBM),
 * and launches them as one batch; batches in different slots overlap on the
 * device. A request that cannot fit even on an idle device completes with
 * status 3. Requests whose estimated search work (gpu_host_rules_work)
 * passes the context's "serve_heavy_work" tuning (taken at start; 0: off)
 * run in heavy batches of their own -- one in flight at a time, on
 * "serve_heavy_percent" of the grid -- so light requests are neither
 * batched with them nor kept off the device by them. Results are exact per
 * rule, independent of batching. While a server runs, its context must not
 * be used by other calls; stop it before gpu_free_graph. */
typedef struct {
    uint64_t tag;              /* the caller's request id */
    GpuRuleResult64 result;
    int64_t submitNs;          /* gpu_monotonic_ns() clock */
    int64_t admitNs;           /* batch launched */
    int64_t completeNs;        /* result available to gpu_serve_poll */
    int32_t batchSize;         /* requests in the batch it ran in */
    int32_t pad;
} GpuServeCompletion;

void* gpu_serve_start(void* ctx, int32_t maxBatch, int32_t maxInFlight, int64_t memBudgetBytes, int32_t flags);
/* Queues n requests; returns n, or -1 if the server is stopping. */
int gpu_serve_submit(void* srv, const GpuRuleDesc* rules, const uint64_t* tags, int32_t n);
/* Moves up to maxOut completions into out, waiting up to waitMs for the
 * first; returns how many. */
int gpu_serve_poll(void* srv, GpuServeCompletion* out, int32_t maxOut, int32_t waitMs);
/* stats[8]: submitted, admitted, completed, batches, refused, deferred for
 * memory, largest batch, peak batches in flight. Returns 0, 1 if a batch
 * failed on the device, -1 on a bad handle. */
int gpu_serve_stats(void* srv, int64_t* stats);
/* Requests routed to heavy batches and heavy batches launched so far;
 * 0, or -1 on a bad handle. No counterpart in the reference. */
int gpu_serve_heavy_stats(void* srv, int64_t* heavyRequests, int64_t* heavyBatches);
/* HBM one execution slot holds for the server's life (bitmap rows, sort and
 * giant-seed buffers), outside the per-request budget; -1 on a bad handle. */
int64_t gpu_serve_slot_bytes(void* srv);
/* Finishes queued and in-flight work, joins the scheduler, frees the server. */
void gpu_serve_stop(void* srv);
/* The monotonic clock the completion timestamps use (ns). */
int64_t gpu_monotonic_ns(void);

/* ---- host-side introspection (no device work; usable without a GPU) ---- */

/* Parses a Turtle file into a host-only graph handle. */
void* gpu_host_load_turtle(const char* ttlPath);
/* Host-only graph handle from term-id triples (as gpu_load_graph_triples). */
void* gpu_host_load_triples(const int32_t* s, const int32_t* p, const int32_t* o, int64_t n, int64_t nTerms);
/* Host-only graph handle from a snapshot (as gpu_load_snapshot). */
void* gpu_host_load_snapshot(const char* path);
void gpu_host_free(void* host);
/* Term count, predicate count and distinct triple count of a graph handle
 * (host handle or full context). */
int gpu_host_graph_info(void* handle, int64_t* nTerms, int32_t* nPredicates, int64_t* nTriples);
/* Term id of a term string as the Turtle reader stores it (IRIs without
 * brackets), or -1. */
int64_t gpu_host_term_id(void* handle, const char* term);
/* The string of term `id` as the Turtle reader stores it (IRIs without
 * brackets; constants a parsed rule introduced included), written
 * NUL-terminated into buf (at most cap bytes, cap >= 1). Returns its full
 * length, or -1 for an unknown id or a bad argument. */
int64_t gpu_host_term_string(void* handle, int64_t id, char* buf, int64_t cap);
/* Parses a rule line (isJson = 0) or JSON object into GPU_RULE_WORDS words.
 * Unknown constants get ids >= the graph's term count. 0 or -1. */
int gpu_host_parse_rule(void* handle, const char* rule, int32_t isJson, int32_t* words);
/* The host's decision on a pass's device counters, as the library takes it
 * after every pass: which = 0, keys the sort-unique path emitted (count)
 * against the buffers' capacity and the largest capacity allowed -- 0
 * counted, 1 grow the buffers and run the pass again, 2 fail; which = 1,
 * seeds escalated to wave mode against the heavy queue's capacity -- 0 or 2.
 * -1 on a negative argument or unknown `which`. Both counters are 64-bit
 * or bounded on the device, so they cannot wrap. */
int32_t gpu_host_count_action(int32_t which, int64_t count, int64_t cap, int64_t capMax);

/* The support gate of gpu_evaluate_rules_min for a rule of headSize head
 * triples: the smallest support that meets minSupport and minHeadCoverage
 * (at least 1; 2^31 - 1 when none can), or -1 for a threshold the entry
 * points refuse. A host check with no counterpart in the reference. */
int64_t gpu_host_support_threshold(int64_t headSize, int64_t minSupport, double minHeadCoverage);

/* Head figures and plan facts of a structured rule: info[0] headSize,
 * info[1] headSupport, info[2] status, info[3] support launched (0/1),
 * info[4] body pairs launched (0/1), info[5] support seeds, info[6] pair seeds. */
int gpu_host_plan_info(void* handle, const int32_t* words, int64_t* info);
/* Estimated search work of n structured rules (GPU_RULE_WORDS words each):
 * candidates their support and body-pair searches enumerate, from per-
 * (direction, predicate) row statistics -- the figure the server schedules
 * by (order of magnitude, not a bound); 0 for rules not counted on the
 * device. 0 or -1. No counterpart in the reference: it is for callers that
 * schedule rule batches. */
int gpu_host_rules_work(void* handle, const int32_t* words, int32_t n, double* work);
/* Injective body groundings of n structured rules: the (head subject, head
 * object) pairs the reference's body-size kernel emits before it
 * deduplicates them (GPU_03/gpu_kernels.cuh:218-317), counted on the host
 * threads (a search whose last atom's row is counted, not enumerated); -1
 * for a rule with constants, more than 8 body atoms or a head variable
 * missing from its body. For measurement: this library never enumerates
 * them. 0 or -1. */
int gpu_host_body_groundings(void* handle, const int32_t* words, int32_t n, int64_t* out);
/* One refinement level of an AMIE-style rule miner (no counterpart in the
 * reference, which leaves mining to its Scala caller). A rule p(?a, ?b) <=
 * body over variables is a key record of GPU_RULE_KEY_WORDS int32: head
 * predicate, atom count n, then n atoms (predicate, subject, object) with
 * variables numbered 0 (?a), 1 (?b), 2, 3, ..., the rest zero; canonical when
 * its atoms, over the numberings of the non-head variables, form the
 * smallest sorted sequence. Writes the canonical keys of every refinement of
 * the n parents -- one more atom over the nPreds body predicates, closing (two
 * variables of the rule) or dangling (one variable and a fresh one), in
 * either orientation, no repeated atom or copy of the head atom, at most
 * maxAtoms atoms, and no rule left with more variables occurring once than
 * the remaining atoms can close -- sorted, without repeats, up to cap
 * records into out. Returns their number (more than cap: call again with
 * room), or -1 on a bad argument. */
#define GPU_RULE_KEY_WORDS 26
int64_t gpu_host_refine_rules(const int32_t* parents, int32_t n, const int32_t* preds, int32_t nPreds,
                              int32_t maxAtoms, int32_t* out, int64_t cap);

/* Library version string. */
const char* gpu_rules_version(void);

#ifdef __cplusplus
}
#endif

#endif /* GPU_RULES_API_H */
