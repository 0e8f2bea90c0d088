// gpu_rules_counter -- command-line rule counter over libgpu_rules.so that
// writes the output of the reference's CPU counter
// (CPU_01_nonparallel/main.cpp:64-175), so a caller that runs that program
// as a subprocess and parses its --json array (the Scala miner's CPU
// fallback) can run this one instead, with the counts computed on the GPU.
//
//   gpu_rules_counter <graph.ttl> <rules.txt|rules.json> [--json] [--per-rule]
//
// Rules: one rule line per non-empty line, or, for a file ending in .json,
// a top-level JSON array of RDFRules rule objects. Every rule is parsed and
// evaluated through the C ABI (gpu_evaluate_rules_text), in one batched pass
// or, with --per-rule, rule by rule with the per-rule launch sequence.
//
// Output (as CPU_01_nonparallel): with --json a JSON array on stdout, one
// object per rule in input order -- ruleId (1-based), rule (variables as
// ?k numbered by first appearance, terms as the graph stores them),
// support, headSize, headSupport, headCoverage, bodySize, confidence,
// supportMs, confidenceMs -- and "total_ms=" on stderr; without it one
// human-readable line pair per rule. Numbers are printed with the default
// stream format, so equal values print equally. bodySize is the number of
// distinct body pairs whatever the support (CPU_01 computes it for every
// rule); headCoverage and confidence are the IEEE quotients of the printed
// integers, 0 where the divisor is 0 (CPU_01_nonparallel/FinalRule.cpp:19-40).
// The times are device times: per rule with --per-rule, else the batch's
// support-phase and pair-phase times spread evenly over its rules (their
// sums are the batch's).
//
// Exit status 1 (message on stderr) if the graph or rule file cannot be
// read, a rule does not parse, or evaluation fails -- where CPU_01 ends on
// an exception.
#include <chrono>
#include <cstdint>
#include <cstdio>
#include <fstream>
#include <iostream>
#include <sstream>
#include <string>
#include <vector>

#include "gpu_rules_api.h"

namespace {

std::string trim(const std::string& s) {
    const size_t a = s.find_first_not_of(" \t\r\n");
    if (a == std::string::npos) return "";
    const size_t b = s.find_last_not_of(" \t\r\n");
    return s.substr(a, b - a + 1);
}

bool ends_with(const std::string& s, const std::string& t) {
    return s.size() >= t.size() && s.compare(s.size() - t.size(), t.size(), t) == 0;
}

// Top-level objects of the array in `text` (brace-balanced, strings skipped).
bool json_objects(const std::string& text, std::vector<std::string>& out) {
    const size_t a = text.find('['), b = text.rfind(']');
    if (a == std::string::npos || b == std::string::npos || b < a) return false;
    int depth = 0;
    bool str = false, esc = false;
    size_t start = 0;
    for (size_t i = a + 1; i < b; ++i) {
        const char c = text[i];
        if (str) {
            if (esc) esc = false;
            else if (c == '\\') esc = true;
            else if (c == '"') str = false;
            continue;
        }
        if (c == '"') str = true;
        else if (c == '{') {
            if (depth++ == 0) start = i;
        } else if (c == '}') {
            if (--depth == 0) out.push_back(text.substr(start, i - start + 1));
            if (depth < 0) return false;
        }
    }
    return depth == 0 && !str;
}

std::string json_escape(const std::string& s) {
    std::string out;
    for (char c : s) {
        if (c == '"') out += "\\\"";
        else if (c == '\\') out += "\\\\";
        else out += c;
    }
    return out;
}

std::string term_string(void* ctx, int32_t t) {
    if (t < 0) return "?" + std::to_string(-(t + 1));
    std::vector<char> buf(256);
    int64_t n = gpu_host_term_string(ctx, t, buf.data(), (int64_t)buf.size());
    if (n >= (int64_t)buf.size()) {
        buf.resize((size_t)n + 1);
        n = gpu_host_term_string(ctx, t, buf.data(), (int64_t)buf.size());
    }
    return n < 0 ? std::string("?") : std::string(buf.data(), (size_t)n);
}

// "( s p o ) ^ ... => ( s p o )" from the structured words of a parsed rule.
std::string rule_string(void* ctx, const int32_t* w) {
    std::string out;
    const int nb = w[0];
    for (int j = 0; j < nb; ++j) {
        const int32_t* a = w + 4 + 3 * j;
        out += "( " + term_string(ctx, a[0]) + " " + term_string(ctx, a[1]) + " " + term_string(ctx, a[2]) + " )";
        if (j + 1 < nb) out += " ^ ";
    }
    out += " => ( " + term_string(ctx, w[1]) + " " + term_string(ctx, w[2]) + " " + term_string(ctx, w[3]) + " )";
    return out;
}

}  // namespace

int main(int argc, char** argv) {
    std::string ttl, rulesFile;
    bool json = false, perRule = false;
    int positional = 0;
    for (int i = 1; i < argc; ++i) {
        const std::string arg = argv[i];
        if (arg == "--json") json = true;
        else if (arg == "--per-rule") perRule = true;
        else if (positional == 0) { ttl = arg; ++positional; }
        else if (positional == 1) { rulesFile = arg; ++positional; }
    }
    if (positional < 2) {
        std::fprintf(stderr, "usage: %s <graph.ttl> <rules.txt|rules.json> [--json] [--per-rule]\n", argv[0]);
        return 1;
    }
    const auto t0 = std::chrono::steady_clock::now();
    void* ctx = gpu_load_graph(ttl.c_str());
    if (!ctx) {
        std::cerr << "Failed to parse TTL file: " << ttl << "\n";
        return 1;
    }
    std::ifstream in(rulesFile);
    if (!in) {
        std::cerr << "Error: Cannot open rule file: " << rulesFile << "\n";
        gpu_free_graph(ctx);
        return 1;
    }
    const bool isJson = ends_with(rulesFile, ".json");
    std::vector<std::string> rules;
    if (isJson) {
        std::stringstream ss;
        ss << in.rdbuf();
        if (!json_objects(ss.str(), rules)) {
            std::cerr << "Error: JSON rule file must contain a top-level array\n";
            gpu_free_graph(ctx);
            return 1;
        }
    } else {
        std::string line;
        while (std::getline(in, line)) {
            const std::string c = trim(line);
            if (!c.empty()) rules.push_back(c);
        }
    }
    const int32_t n = (int32_t)rules.size();
    std::vector<int32_t> words((size_t)n * GPU_RULE_WORDS);
    std::vector<const char*> ptrs((size_t)n);
    for (int32_t i = 0; i < n; ++i) {
        ptrs[i] = rules[i].c_str();
        if (gpu_host_parse_rule(ctx, rules[i].c_str(), isJson ? 1 : 0, words.data() + (size_t)i * GPU_RULE_WORDS) != 0) {
            std::cerr << "Error: cannot parse rule " << (i + 1) << ": " << rules[i] << "\n";
            gpu_free_graph(ctx);
            return 1;
        }
    }
    std::vector<GpuRuleResult64> out((size_t)n);
    const int32_t flags = GPU_RULES_UNGATED | (perRule ? GPU_RULES_PER_RULE : 0);
    if (n > 0 && gpu_evaluate_rules_text(ctx, ptrs.data(), n, isJson ? 1 : 0, flags, out.data()) != 0) {
        std::cerr << "Error: evaluation failed\n";
        gpu_free_graph(ctx);
        return 1;
    }
    std::vector<double> supMs((size_t)n, 0.0), confMs((size_t)n, 0.0);
    if (n > 0 && perRule) {
        gpu_last_rule_times(ctx, supMs.data(), confMs.data(), n);
    } else if (n > 0) {
        double total = 0, sup = 0, pairs = 0;
        int64_t heavy = 0;
        gpu_last_batch_stats(ctx, &total, &sup, &pairs, &heavy);
        if (sup <= 0 && pairs <= 0) pairs = total;   // graph replays time the whole pass
        for (int32_t i = 0; i < n; ++i) {
            supMs[i] = sup / n;
            confMs[i] = pairs / n;
        }
    }
    const double totalMs =
        std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    if (json) {
        std::cerr << "total_ms=" << totalMs << "\n";
        std::cout << "[\n";
    }
    for (int32_t i = 0; i < n; ++i) {
        const GpuRuleResult64& r = out[i];
        const int support = (int)r.support, headSize = (int)r.headSize, headSupport = (int)r.headSupport;
        const int bodySize = r.bodyPairs > 0 ? (int)r.bodyPairs : 0;
        const double headCoverage = headSize > 0 ? (double)support / (double)headSize : 0.0;
        const double confidence = bodySize > 0 ? (double)support / (double)bodySize : 0.0;
        const std::string text = rule_string(ctx, words.data() + (size_t)i * GPU_RULE_WORDS);
        if (json) {
            std::cout << "  {\n"
                      << "    \"ruleId\": " << (i + 1) << ",\n"
                      << "    \"rule\": \"" << json_escape(text) << "\",\n"
                      << "    \"support\": " << support << ",\n"
                      << "    \"headSize\": " << headSize << ",\n"
                      << "    \"headSupport\": " << headSupport << ",\n"
                      << "    \"headCoverage\": " << headCoverage << ",\n"
                      << "    \"bodySize\": " << bodySize << ",\n"
                      << "    \"confidence\": " << confidence << ",\n"
                      << "    \"supportMs\": " << supMs[i] << ",\n"
                      << "    \"confidenceMs\": " << confMs[i] << "\n"
                      << "  }" << (i + 1 < n ? "," : "") << "\n";
        } else {
            std::cout << "Rule " << (i + 1) << ": " << text << "\n"
                      << "  support=" << support << "  headCoverage=" << headCoverage << "  bodySize=" << bodySize
                      << "  confidence=" << confidence << "  supportMs=" << supMs[i]
                      << "  confidenceMs=" << confMs[i] << "\n";
        }
    }
    if (json) std::cout << "]\n";
    else std::cout << "\nRules evaluated: " << n << ", total " << totalMs << " ms\n";
    gpu_free_graph(ctx);
    return 0;
}
