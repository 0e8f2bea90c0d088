// rules_host.hpp -- rule IR and the two rule front ends (text line, JSON).
#pragma once
#include <cstdint>
#include <stdexcept>
#include <string>
#include <unordered_map>
#include <vector>

#include "graph_host.hpp"
#include "plan.hpp"

namespace gr {

struct TermIR {
    bool var = true;
    int32_t v = -1;   // variable number, or constant term id (ids >= nE never match)
};

struct AtomIR {
    TermIR s;
    int32_t pterm = -1;  // predicate term id (may be unknown to the graph)
    TermIR o;
};

struct RuleIR {
    AtomIR head;
    std::vector<AtomIR> body;
    int32_t nvars = 0;
};

struct ParseError : std::runtime_error {
    using std::runtime_error::runtime_error;
};

// Term resolution shared by both front ends. Unknown constants get ids
// past the graph's term space, in first-seen order, kept for the context's
// lifetime (the reference interns them too: GPU_03/RuleParser.cpp:246-248).
class TermResolver {
public:
    explicit TermResolver(const HostGraph& g) : g_(g) {}
    int32_t resolve(const std::string& expanded);
    std::string expand(const std::string& token) const;  // <iri>, prefix:local, literals
    const std::string* name(int32_t id) const;

private:
    const HostGraph& g_;
    std::unordered_map<std::string, int32_t> extra_;
    std::vector<std::string> extraNames_;
};

// "( s p o ) ^ ... => ( s p o ) [| comment]" (GPU_03/RuleParser.cpp:251-322).
// Variables are numbered by first appearance: body atoms subject then
// object, then the head. Throws ParseError.
RuleIR parse_rule_text(const std::string& line, TermResolver& tr);

// RDFRules JSON object {"body":[{subject,predicate,object}...],"head":{...}}
// (GPU_03/RuleParser.cpp:296-375, body_test_data/rules_to_json.py:9-12).
RuleIR parse_rule_json(const std::string& json, TermResolver& tr);

// Structured rule record: [nBody, head(s,p,o), body[8](s,p,o)]; a term >= 0
// is a constant term id, < 0 is variable -(t+1).
RuleIR rule_from_words(const int32_t* w);
void rule_to_words(const RuleIR& r, int32_t* w);

}  // namespace gr
