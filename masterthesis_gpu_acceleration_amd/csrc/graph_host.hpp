// graph_host.hpp -- host-side knowledge graph: term dictionary, Turtle
// reader and the predicate-grouped adjacency arrays uploaded to HBM.
#pragma once
#include <algorithm>
#include <cstdint>
#include <string>
#include <unordered_map>
#include <utility>
#include <vector>

namespace gr {

// Dense term ids in first-seen order (GPU_03/RdfIndexes.cpp:11-28 assigns
// ids the same way: subject, predicate, object of each triple in turn).
struct TermTable {
    std::unordered_map<std::string, int32_t> ids;
    std::vector<std::string> strs;

    int32_t intern(const std::string& s);
    int32_t find(const std::string& s) const;
    size_t size() const { return strs.size(); }
};

struct HostGraph {
    TermTable terms;                                        // empty for id-loaded graphs
    std::unordered_map<std::string, std::string> prefixes;  // prefix lines of the Turtle file
    int64_t nE = 0;                                         // row space (term ids at load)

    std::vector<int32_t> predTerm;                 // dense predicate -> term id
    std::unordered_map<int32_t, int32_t> predOf;   // term id -> dense predicate

    // Triples grouped by predicate; unique. Entries of predicate p occupy
    // [predOff[p], predOff[p+1]) in every array below.
    std::vector<int32_t> predOff;
    std::vector<int32_t> spoKey, spoVal;   // sorted by (p, s, o): key = s, val = o
    std::vector<int32_t> posKey, posVal;   // sorted by (p, o, s): key = o, val = s
    std::vector<int32_t> keysS, keysSOff;  // distinct subjects per predicate
    std::vector<int32_t> keysO, keysOOff;  // distinct objects per predicate

    int32_t P() const { return (int32_t)predTerm.size(); }
    int64_t nBlocks() const { return (std::max<int64_t>(nE, 1) + 63) / 64; }   // key-index blocks per (dir, pred)
    int64_t T() const { return (int64_t)spoKey.size(); }
    int32_t nnz(int32_t p) const { return predOff[p + 1] - predOff[p]; }
    int32_t nkeys(int dir, int32_t p) const {
        const auto& o = dir == 0 ? keysSOff : keysOOff;
        return o[p + 1] - o[p];
    }
    // Entry range [begin,end) of key k in direction dir for predicate p.
    std::pair<int32_t, int32_t> row(int dir, int32_t p, int32_t k) const;
    bool has(int32_t p, int32_t s, int32_t o) const;
    int32_t dense_pred(int32_t term) const {
        auto it = predOf.find(term);
        return it == predOf.end() ? -1 : it->second;
    }
};

// Raw triples (term ids) -> grouped, deduplicated arrays. `n_terms` sizes
// the row space; every id must be below it.
void build_graph(HostGraph& g, std::vector<int32_t>& s, std::vector<int32_t>& p,
                 std::vector<int32_t>& o, int64_t n_terms);

// Reads a Turtle file into g (terms, prefixes, arrays). Returns false and
// sets err if the file cannot be read.
bool load_turtle(const std::string& path, HostGraph& g, std::string& err);

// The parsing half of load_turtle: terms and prefixes into g, the triples'
// term ids into S, P, O (the arrays are left to build).
bool parse_turtle(const std::string& path, HostGraph& g, std::vector<int32_t>& S, std::vector<int32_t>& P,
                  std::vector<int32_t>& O, std::string& err);

// Prefix declarations as the rule parser sees them: line-based scan for
// "@prefix name: <iri>" / "PREFIX name: <iri>" (GPU_03/RuleParser.cpp:51-96).
std::unordered_map<std::string, std::string> scan_prefix_lines(const std::string& text);

}  // namespace gr
