/*  GpuRulesBatchLib.scala -- JNA binding of libgpu_rules.so's batched,
 *  serving and allocation-free entry points (include/gpu_rules_api.h), for
 *  a miner that already binds the four per-rule functions through the
 *  reference's GpuRulesLib (GPU_03/GpuRulesLib.scala:46-58). It adds the
 *  JSON entry point that binding leaves out, batched evaluation with exact
 *  64-bit counts, resident batches, the continuous-batching server and the
 *  allocation-free per-rule call.
 *
 *  Dependency: "net.java.dev.jna" % "jna" % "5.14.0" (as the reference).
 *  The library is found as "gpu_rules" on jna.library.path.
 *
 *  Not compiled in this repository (no JVM here). tests/test_abi_host.py
 *  checks every declaration below against the C header (names, argument
 *  counts, pointer vs value arguments) and the structure field orders
 *  against the C layouts, which the ctypes binding (lib.py) pins with
 *  sizeof/offset checks.
 */

import com.sun.jna.{Library, Memory, Native, Pointer, Structure}
import com.sun.jna.ptr.{DoubleByReference, IntByReference, LongByReference}
import java.util

// GpuRuleResult as the reference binding declares it (56 bytes; field order
// is the C order).
class GpuRuleResultJ extends Structure {
  var support: Int = 0
  var headSize: Int = 0
  var headSupport: Int = 0
  var headCoverage: Double = 0.0
  var bodySize: Int = 0
  var confidence: Double = 0.0
  var supportMs: Double = 0.0
  var confidenceMs: Double = 0.0
  override def getFieldOrder: util.List[String] = util.Arrays.asList(
    "support", "headSize", "headSupport", "headCoverage", "bodySize", "confidence", "supportMs", "confidenceMs")
}

// GpuRuleResult64: one rule of a batch, exact 64-bit counts (64 bytes).
class GpuRuleResult64 extends Structure {
  var support: Long = 0L
  var headSize: Long = 0L
  var headSupport: Long = 0L
  var bodySize: Long = 0L      // distinct body pairs, 0 when support == 0
  var bodyPairs: Long = 0L     // distinct body pairs whatever the support; -1 if not computed
  var headCoverage: Double = 0.0
  var confidence: Double = 0.0
  var status: Int = 0          // 0 ok, 1 unsupported shape, 2 parse error or failed evaluation, 3 refused
  var latencyUs: Int = 0
  override def getFieldOrder: util.List[String] = util.Arrays.asList(
    "support", "headSize", "headSupport", "bodySize", "bodyPairs", "headCoverage", "confidence", "status",
    "latencyUs")
}

// GpuServeCompletion: one finished serving request (104 bytes).
class GpuServeCompletion extends Structure {
  var tag: Long = 0L
  var result: GpuRuleResult64 = new GpuRuleResult64()
  var submitNs: Long = 0L
  var admitNs: Long = 0L
  var completeNs: Long = 0L
  var batchSize: Int = 0
  var pad: Int = 0
  override def getFieldOrder: util.List[String] = util.Arrays.asList(
    "tag", "result", "submitNs", "admitNs", "completeNs", "batchSize", "pad")
}

trait GpuRulesBatchNative extends Library {
  // int gpu_evaluate_rule_json(void* ctx, const char* ruleJson, GpuRuleResult* out)
  def gpu_evaluate_rule_json(ctx: Pointer, ruleJson: String, out: GpuRuleResultJ): Int
  // int gpu_evaluate_rule_noalloc(void* ctx, const char* ruleText, GpuRuleResult* out)
  def gpu_evaluate_rule_noalloc(ctx: Pointer, ruleText: String, out: GpuRuleResultJ): Int
  // int gpu_evaluate_rules_text(void* ctx, const char* const* rules, int32_t n, int32_t isJson,
  //                             int32_t flags, GpuRuleResult64* out)
  def gpu_evaluate_rules_text(ctx: Pointer, rules: Array[String], n: Int, isJson: Int, flags: Int,
                              out: GpuRuleResult64): Int
  // int gpu_evaluate_rules(void* ctx, const GpuRuleDesc* rules, int32_t n, int32_t flags, GpuRuleResult64* out)
  def gpu_evaluate_rules(ctx: Pointer, rules: Pointer, n: Int, flags: Int, out: GpuRuleResult64): Int
  // void* gpu_batch_create(void* ctx, const GpuRuleDesc* rules, int32_t n, int32_t flags)
  def gpu_batch_create(ctx: Pointer, rules: Pointer, n: Int, flags: Int): Pointer
  // int gpu_evaluate_rules_min(void* ctx, const GpuRuleDesc* rules, int32_t n, int32_t flags,
  //                            int64_t minSupport, double minHeadCoverage, GpuRuleResult64* out)
  def gpu_evaluate_rules_min(ctx: Pointer, rules: Pointer, n: Int, flags: Int, minSupport: Long,
                             minHeadCoverage: Double, out: GpuRuleResult64): Int
  // void* gpu_batch_create_min(void* ctx, const GpuRuleDesc* rules, int32_t n, int32_t flags,
  //                            int64_t minSupport, double minHeadCoverage)
  def gpu_batch_create_min(ctx: Pointer, rules: Pointer, n: Int, flags: Int, minSupport: Long,
                           minHeadCoverage: Double): Pointer
  // int64_t gpu_host_support_threshold(int64_t headSize, int64_t minSupport, double minHeadCoverage)
  def gpu_host_support_threshold(headSize: Long, minSupport: Long, minHeadCoverage: Double): Long
  // int64_t gpu_host_refine_rules(const int32_t* parents, int32_t n, const int32_t* preds, int32_t nPreds,
  //                               int32_t maxAtoms, int32_t* out, int64_t cap)
  def gpu_host_refine_rules(parents: Pointer, n: Int, preds: Array[Int], nPreds: Int, maxAtoms: Int, out: Pointer,
                            cap: Long): Long
  // int gpu_batch_run(void* batch, GpuRuleResult64* out)
  def gpu_batch_run(batch: Pointer, out: GpuRuleResult64): Int
  // void gpu_batch_free(void* batch)
  def gpu_batch_free(batch: Pointer): Unit
  // int gpu_host_parse_rule(void* handle, const char* rule, int32_t isJson, int32_t* words)
  def gpu_host_parse_rule(handle: Pointer, rule: String, isJson: Int, words: Array[Int]): Int
  // int gpu_last_batch_stats(void* ctx, double* totalMs, double* supportMs, double* pairsMs, int64_t* heavySeeds)
  def gpu_last_batch_stats(ctx: Pointer, totalMs: DoubleByReference, supportMs: DoubleByReference,
                           pairsMs: DoubleByReference, heavySeeds: LongByReference): Int
  // int gpu_set_tuning(void* ctx, const char* key, int64_t value)
  def gpu_set_tuning(ctx: Pointer, key: String, value: Long): Int
  // int64_t gpu_get_tuning(void* ctx, const char* key)
  def gpu_get_tuning(ctx: Pointer, key: String): Long
  // int gpu_autotune(void* ctx, const GpuRuleDesc* rules, int32_t n, int32_t flags, double* log,
  //                  int32_t logCap, int32_t* logLen)
  def gpu_autotune(ctx: Pointer, rules: Pointer, n: Int, flags: Int, log: Pointer, logCap: Int,
                   logLen: IntByReference): Int
  // void* gpu_serve_start(void* ctx, int32_t maxBatch, int32_t maxInFlight, int64_t memBudgetBytes, int32_t flags)
  def gpu_serve_start(ctx: Pointer, maxBatch: Int, maxInFlight: Int, memBudgetBytes: Long, flags: Int): Pointer
  // int gpu_serve_submit(void* srv, const GpuRuleDesc* rules, const uint64_t* tags, int32_t n)
  def gpu_serve_submit(srv: Pointer, rules: Pointer, tags: Array[Long], n: Int): Int
  // int gpu_serve_poll(void* srv, GpuServeCompletion* out, int32_t maxOut, int32_t waitMs)
  def gpu_serve_poll(srv: Pointer, out: GpuServeCompletion, maxOut: Int, waitMs: Int): Int
  // int gpu_serve_stats(void* srv, int64_t* stats)
  def gpu_serve_stats(srv: Pointer, stats: Array[Long]): Int
  // void gpu_serve_stop(void* srv)
  def gpu_serve_stop(srv: Pointer): Unit
}

// Scala-side wrapper over a context the caller loaded with gpu_load_graph
// (through the reference binding's GpuRulesLib, or its own gpu_load_graph call).
class GpuRulesBatchLib(val ctx: Pointer) {
  import GpuRulesBatchLib._

  val lib: GpuRulesBatchNative = Native.load("gpu_rules", classOf[GpuRulesBatchNative])

  /** One rule given as an RDFRules JSON object (throws on a non-zero return, as evaluateRule does). */
  def evaluateRuleJson(ruleJson: String): GpuRuleResultJ = {
    val out = new GpuRuleResultJ()
    val rc = lib.gpu_evaluate_rule_json(ctx, ruleJson, out)
    if (rc != 0) throw new RuntimeException(s"gpu_evaluate_rule_json error code $rc for: $ruleJson")
    out
  }

  /** One rule line without any device allocation in the call (fixed per-device scratch). */
  def evaluateRuleNoAlloc(ruleText: String): GpuRuleResultJ = {
    val out = new GpuRuleResultJ()
    val rc = lib.gpu_evaluate_rule_noalloc(ctx, ruleText, out)
    if (rc != 0) throw new RuntimeException(s"gpu_evaluate_rule_noalloc error code $rc for: $ruleText")
    out
  }

  /** Many rule lines (or JSON objects) in one batched pass; results in input order. */
  def evaluateRules(rules: Seq[String], json: Boolean = false, ungated: Boolean = false): Array[GpuRuleResult64] = {
    if (rules.isEmpty) return Array.empty
    val out = new GpuRuleResult64().toArray(rules.length).asInstanceOf[Array[GpuRuleResult64]]
    val rc = lib.gpu_evaluate_rules_text(ctx, rules.toArray, rules.length, if (json) 1 else 0,
                                         if (ungated) Ungated else 0, out(0))
    if (rc != 0) throw new RuntimeException(s"gpu_evaluate_rules_text error code $rc")
    out.foreach(_.read())
    out
  }

  /** A refinement level of a miner: support of every rule, confidence only for the rules with
   *  support >= minSupport and headCoverage >= minHeadCoverage (bodyPairs -1 for the others). */
  def evaluateRulesMin(rules: Seq[String], minSupport: Long, minHeadCoverage: Double,
                       json: Boolean = false): Array[GpuRuleResult64] = {
    if (rules.isEmpty) return Array.empty
    val out = new GpuRuleResult64().toArray(rules.length).asInstanceOf[Array[GpuRuleResult64]]
    val rc = lib.gpu_evaluate_rules_min(ctx, ruleDescs(rules, json), rules.length, 0, minSupport, minHeadCoverage,
                                        out(0))
    if (rc != 0) throw new RuntimeException(s"gpu_evaluate_rules_min error code $rc")
    out.foreach(_.read())
    out
  }

  /** Structured rules (GPU_RULE_WORDS ints each) of rule lines, for the batch and serving calls. */
  def ruleDescs(rules: Seq[String], json: Boolean = false): Memory = {
    val mem = new Memory(math.max(1L, rules.length.toLong * RuleWords * 4))
    val words = new Array[Int](RuleWords)
    rules.zipWithIndex.foreach { case (r, i) =>
      if (lib.gpu_host_parse_rule(ctx, r, if (json) 1 else 0, words) != 0)
        throw new RuntimeException(s"cannot parse rule: $r")
      mem.write(i.toLong * RuleWords * 4, words, 0, RuleWords)
    }
    mem
  }

  /** A batch kept resident in HBM and replayed as a HIP graph on every run(). */
  final class ResidentBatch(rules: Seq[String], ungated: Boolean, minSupport: Long = 1L,
                            minHeadCoverage: Double = 0.0) extends AutoCloseable {
    private val descs = ruleDescs(rules)
    private val h =
      if (minSupport <= 1L && minHeadCoverage == 0.0) lib.gpu_batch_create(ctx, descs, rules.length, if (ungated) Ungated else 0)
      else lib.gpu_batch_create_min(ctx, descs, rules.length, 0, minSupport, minHeadCoverage)
    if (h == null) throw new RuntimeException("gpu_batch_create failed")
    private val out = new GpuRuleResult64().toArray(math.max(1, rules.length)).asInstanceOf[Array[GpuRuleResult64]]
    def run(): Array[GpuRuleResult64] = {
      if (lib.gpu_batch_run(h, out(0)) != 0) throw new RuntimeException("gpu_batch_run failed")
      out.foreach(_.read())
      out.take(rules.length)
    }
    override def close(): Unit = lib.gpu_batch_free(h)
  }

  /** Continuous-batching server: submit rules with tags, poll completions. */
  final class Server(maxBatch: Int = 2048, maxInFlight: Int = 3, memBudgetBytes: Long = 0L,
                     ungated: Boolean = false) extends AutoCloseable {
    private val h = lib.gpu_serve_start(ctx, maxBatch, maxInFlight, memBudgetBytes, if (ungated) Ungated else 0)
    if (h == null) throw new RuntimeException("gpu_serve_start failed")
    def submit(rules: Seq[String], tags: Seq[Long]): Int =
      lib.gpu_serve_submit(h, ruleDescs(rules), tags.toArray, rules.length)
    def poll(maxOut: Int = 1024, waitMs: Int = 10): Array[GpuServeCompletion] = {
      val buf = new GpuServeCompletion().toArray(maxOut).asInstanceOf[Array[GpuServeCompletion]]
      val k = lib.gpu_serve_poll(h, buf(0), maxOut, waitMs)
      if (k < 0) throw new RuntimeException("gpu_serve_poll failed")
      buf.take(k).map { c => c.read(); c }
    }
    /** submitted, admitted, completed, batches, refused, deferred for memory, largest batch, peak in flight */
    def stats(): Array[Long] = {
      val s = new Array[Long](8)
      lib.gpu_serve_stats(h, s)
      s
    }
    override def close(): Unit = lib.gpu_serve_stop(h)
  }
}

object GpuRulesBatchLib {
  val RuleWords = 28      // GPU_RULE_WORDS
  val Ungated = 1         // GPU_RULES_UNGATED
  val PerRule = 2         // GPU_RULES_PER_RULE
}
