// Probe of the reference's front end (Turtle reader + rule parser), linked
// against GPU_03/RdfIndexes.cpp, RuleParser.cpp and FinalRule.cpp where they
// lie (see oracle/ref.mk). Test infrastructure only: it prints, for a Turtle
// file and a rule file,
//   T <id> <term>                       every term in id order (\\, \\n, \\r, \\t escaped)
//   N <terms> <predicates> <triples>    graph totals
//   R <nbody> <hs> <hp> <ho> (<s> <p> <o>)*   each rule, terms as
//                                       v<id> (variable) or c<id> (constant)
#include <fstream>
#include <iostream>
#include <string>

#include "RdfIndexes.hpp"
#include "RuleParser.hpp"

// Terms may contain newlines or tabs (long literals): print them escaped.
static std::string esc(const std::string& s) {
    std::string o;
    for (char c : s) {
        if (c == '\\') o += "\\\\";
        else if (c == '\n') o += "\\n";
        else if (c == '\r') o += "\\r";
        else if (c == '\t') o += "\\t";
        else o += c;
    }
    return o;
}

static std::string term(const Term& t) { return (t.isVariable() ? "v" : "c") + std::to_string(t.value); }

int main(int argc, char** argv) {
    if (argc < 2) return 2;
    RdfIndexes idx;
    if (!idx.parseTurtleFile(argv[1])) return 1;
    long triples = 0;
    for (const auto& kv : idx.predIndexes) triples += kv.second.totalPairs();
    const size_t nTerms = idx.mapper.size();
    RuleParser parser(idx, argv[1]);
    if (argc > 2) {
        std::ifstream in(argv[2]);
        std::string line;
        while (std::getline(in, line)) {
            if (line.empty()) continue;
            try {
                FinalRule r = parser.parseRuleLine(line);
                std::cout << "R " << r.body.size() << " " << term(r.head.subject) << " c" << r.head.predicate << " "
                          << term(r.head.object);
                for (const auto& a : r.body)
                    std::cout << " " << term(a.subject) << " c" << a.predicate << " " << term(a.object);
                std::cout << "\n";
            } catch (const std::exception&) {
                std::cout << "R error\n";
            }
        }
    }
    for (size_t i = 0; i < nTerms; ++i) std::cout << "T " << i << " " << esc(idx.mapper.getValue((int)i)) << "\n";
    std::cout << "N " << nTerms << " " << idx.predIndexes.size() << " " << triples << "\n";
    return 0;
}
