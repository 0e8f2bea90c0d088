// graph_host.cpp -- term dictionary, Turtle reader, adjacency build.
//
// The Turtle reader accepts the subset the reference's reader accepts
// (GPU_03/RdfIndexes.cpp:146-415) and produces the same term strings, so
// rule constants written as <iri>, prefix:local, "literal" resolve to the
// same terms: IRIs without brackets, prefixed names expanded, the keyword
// `a` as the rdf:type IRI, literals as "\"text\"" plus "@lang" or "^^type",
// blank nodes "_:label", anonymous "[]" as "_:anonN", bare numbers quoted.
#include "graph_host.hpp"

#include <algorithm>
#include <cctype>
#include <cstring>
#include <fstream>
#include <numeric>
#include <sstream>

namespace gr {

int32_t TermTable::intern(const std::string& s) {
    auto it = ids.find(s);
    if (it != ids.end()) return it->second;
    int32_t id = (int32_t)strs.size();
    ids.emplace(s, id);
    strs.push_back(s);
    return id;
}

int32_t TermTable::find(const std::string& s) const {
    auto it = ids.find(s);
    return it == ids.end() ? -1 : it->second;
}

std::pair<int32_t, int32_t> HostGraph::row(int dir, int32_t p, int32_t k) const {
    const std::vector<int32_t>& key = dir == 0 ? spoKey : posKey;
    auto b = key.begin() + predOff[p], e = key.begin() + predOff[p + 1];
    auto lo = std::lower_bound(b, e, k);
    auto hi = std::upper_bound(lo, e, k);
    return {(int32_t)(lo - key.begin()), (int32_t)(hi - key.begin())};
}

bool HostGraph::has(int32_t p, int32_t s, int32_t o) const {
    auto r = row(0, p, s);
    return std::binary_search(spoVal.begin() + r.first, spoVal.begin() + r.second, o);
}

namespace {

struct Trip { int32_t p, k, v; };

void group(const std::vector<Trip>& t, int32_t P, std::vector<int32_t>& off, std::vector<int32_t>& key,
           std::vector<int32_t>& val, std::vector<int32_t>& keys, std::vector<int32_t>& keysOff) {
    off.assign(P + 1, 0);
    keysOff.assign(P + 1, 0);
    key.resize(t.size());
    val.resize(t.size());
    keys.clear();
    for (size_t i = 0; i < t.size(); ++i) {
        key[i] = t[i].k;
        val[i] = t[i].v;
        off[t[i].p + 1]++;
        if (i == 0 || t[i].p != t[i - 1].p || t[i].k != t[i - 1].k) {
            keys.push_back(t[i].k);
            keysOff[t[i].p + 1]++;
        }
    }
    std::partial_sum(off.begin(), off.end(), off.begin());
    std::partial_sum(keysOff.begin(), keysOff.end(), keysOff.begin());
}

}  // namespace

void build_graph(HostGraph& g, std::vector<int32_t>& s, std::vector<int32_t>& p,
                 std::vector<int32_t>& o, int64_t n_terms) {
    g.nE = n_terms;
    g.predTerm.clear();
    g.predOf.clear();
    std::vector<Trip> t(s.size());
    for (size_t i = 0; i < s.size(); ++i) {
        auto it = g.predOf.find(p[i]);
        int32_t dp;
        if (it == g.predOf.end()) {
            dp = (int32_t)g.predTerm.size();
            g.predOf.emplace(p[i], dp);
            g.predTerm.push_back(p[i]);
        } else {
            dp = it->second;
        }
        t[i] = {dp, s[i], o[i]};
    }
    std::vector<int32_t>().swap(s);
    std::vector<int32_t>().swap(p);
    std::vector<int32_t>().swap(o);
    auto lt = [](const Trip& a, const Trip& b) {
        return a.p != b.p ? a.p < b.p : a.k != b.k ? a.k < b.k : a.v < b.v;
    };
    auto eq = [](const Trip& a, const Trip& b) { return a.p == b.p && a.k == b.k && a.v == b.v; };
    std::sort(t.begin(), t.end(), lt);
    t.erase(std::unique(t.begin(), t.end(), eq), t.end());
    int32_t P = (int32_t)g.predTerm.size();
    group(t, P, g.predOff, g.spoKey, g.spoVal, g.keysS, g.keysSOff);
    for (auto& x : t) std::swap(x.k, x.v);
    std::sort(t.begin(), t.end(), lt);
    std::vector<int32_t> off2;
    group(t, P, off2, g.posKey, g.posVal, g.keysO, g.keysOOff);
}

// ------------------------------------------------------------------ Turtle

namespace {

const char* kRdfType = "http://www.w3.org/1999/02/22-rdf-syntax-ns#type";
const char* kXsdBool = "http://www.w3.org/2001/XMLSchema#boolean";

inline bool name_char(unsigned char c, bool allow_high) {
    return std::isalnum(c) || c == '_' || c == '-' || c == '.' || (allow_high && c >= 128);
}

class TurtleReader {
public:
    TurtleReader(const std::string& text, HostGraph& g, std::vector<int32_t>& S,
                 std::vector<int32_t>& P, std::vector<int32_t>& O)
        : t_(text), g_(g), S_(S), P_(P), O_(O) {}

    void run() {
        for (;;) {
            ws();
            if (eof()) break;
            if (cur() == '@') {
                ++i_;
                std::string kw;
                while (!eof() && std::isalpha((unsigned char)cur())) kw += t_[i_++];
                if (kw == "prefix") {
                    prefix_decl();
                } else {  // @base and other directives are skipped to their '.'
                    while (!eof() && cur() != '.') ++i_;
                    if (!eof()) ++i_;
                }
                continue;
            }
            if (sparql_prefix()) continue;
            std::string subj = term();
            if (subj.empty()) { ++i_; continue; }
            predicate_objects(subj);
            ws();
            if (!eof() && cur() == '.') ++i_;
        }
    }

private:
    const std::string& t_;
    HostGraph& g_;
    std::vector<int32_t>& S_;
    std::vector<int32_t>& P_;
    std::vector<int32_t>& O_;
    std::unordered_map<std::string, std::string> pfx_;
    size_t i_ = 0;
    int anon_ = 0;

    bool eof() const { return i_ >= t_.size(); }
    char cur() const { return t_[i_]; }

    void ws() {
        while (!eof()) {
            char c = cur();
            if (std::isspace((unsigned char)c)) ++i_;
            else if (c == '#') { while (!eof() && cur() != '\n') ++i_; }
            else break;
        }
    }

    bool sparql_prefix() {
        if (t_.size() - i_ < 6) return false;
        for (int k = 0; k < 6; ++k)
            if (std::tolower((unsigned char)t_[i_ + k]) != "prefix"[k]) return false;
        if (t_.size() - i_ > 6 && !std::isspace((unsigned char)t_[i_ + 6])) return false;
        i_ += 6;
        prefix_decl();
        return true;
    }

    void prefix_decl() {
        ws();
        std::string name;
        while (!eof() && cur() != ':') {
            if (!std::isspace((unsigned char)cur())) name += cur();
            ++i_;
        }
        if (!eof()) ++i_;
        ws();
        std::string iri;
        if (!eof() && cur() == '<') iri = iriref();
        pfx_[name] = iri;
        ws();
        if (!eof() && cur() == '.') ++i_;
    }

    std::string iriref() {
        ++i_;
        size_t b = i_;
        while (!eof() && cur() != '>') ++i_;
        std::string r = t_.substr(b, i_ - b);
        if (!eof()) ++i_;
        return r;
    }

    // Reads a name made of name characters; trailing dots belong to the
    // statement, not the name.
    std::string name(bool allow_high) {
        size_t b = i_;
        while (!eof() && name_char((unsigned char)cur(), allow_high)) ++i_;
        while (i_ > b && t_[i_ - 1] == '.') --i_;
        return t_.substr(b, i_ - b);
    }

    std::string expand(const std::string& prefix, const std::string& local) {
        auto it = pfx_.find(prefix);
        return it != pfx_.end() ? it->second + local : prefix + ":" + local;
    }

    std::string quoted() {
        char q = t_[i_++];
        bool triple = i_ + 1 < t_.size() && t_[i_] == q && t_[i_ + 1] == q;
        if (triple) i_ += 2;
        std::string v;
        while (!eof()) {
            if (triple) {
                if (i_ + 2 < t_.size() && t_[i_] == q && t_[i_ + 1] == q && t_[i_ + 2] == q) { i_ += 3; break; }
            } else {
                if (cur() == q) { ++i_; break; }
                if (cur() == '\n') break;
            }
            if (cur() == '\\' && i_ + 1 < t_.size()) {
                char e = t_[i_ + 1];
                i_ += 2;
                switch (e) {
                    case 'n': v += '\n'; break;
                    case 't': v += '\t'; break;
                    case 'r': v += '\r'; break;
                    case '\\': v += '\\'; break;
                    case '"': v += '"'; break;
                    case '\'': v += '\''; break;
                    default: v += '\\'; v += e; break;
                }
            } else {
                v += t_[i_++];
            }
        }
        return v;
    }

    std::string term() {
        ws();
        if (eof()) return "";
        char c = cur();
        if (c == '<') return iriref();
        if (c == '_' && i_ + 1 < t_.size() && t_[i_ + 1] == ':') {
            i_ += 2;
            return "_:" + name(false);
        }
        if (c == '[') {
            ++i_;
            ws();
            if (!eof() && cur() == ']') ++i_;
            return "_:anon" + std::to_string(++anon_);
        }
        if (c == '"' || c == '\'') {
            std::string lit = "\"" + quoted() + "\"";
            ws();
            if (!eof() && cur() == '@') {
                ++i_;
                size_t b = i_;
                while (!eof() && (std::isalnum((unsigned char)cur()) || cur() == '-')) ++i_;
                lit += "@" + t_.substr(b, i_ - b);
            } else if (i_ + 1 < t_.size() && cur() == '^' && t_[i_ + 1] == '^') {
                i_ += 2;
                lit += "^^" + term();
            }
            return lit;
        }
        if (c == ':') {
            ++i_;
            return expand_default(name(true));
        }
        if (std::isalpha((unsigned char)c) || c == '_') {
            std::string id = name(true);
            if (!eof() && cur() == ':') {
                ++i_;
                return expand(id, name(true));
            }
            if (id == "a") return kRdfType;
            if (id == "true") return std::string("\"true\"^^") + kXsdBool;
            if (id == "false") return std::string("\"false\"^^") + kXsdBool;
            return id;
        }
        if (std::isdigit((unsigned char)c) || c == '+' || c == '-') {
            size_t b = i_;
            while (!eof()) {
                char d = cur();
                if (std::isdigit((unsigned char)d) || d == '.' || d == 'e' || d == 'E' || d == '+' || d == '-') ++i_;
                else break;
            }
            return "\"" + t_.substr(b, i_ - b) + "\"";
        }
        return "";
    }

    std::string expand_default(const std::string& local) {
        auto it = pfx_.find("");
        return it != pfx_.end() ? it->second + local : ":" + local;
    }

    void emit(const std::string& s, const std::string& p, const std::string& o) {
        // ids in subject, predicate, object order (GPU_03/RdfIndexes.cpp:172-178)
        S_.push_back(g_.terms.intern(s));
        P_.push_back(g_.terms.intern(p));
        O_.push_back(g_.terms.intern(o));
    }

    void predicate_objects(const std::string& subj) {
        for (;;) {
            ws();
            if (eof() || cur() == '.' || cur() == ']') break;
            if (cur() == ';') {
                ++i_;
                ws();
                if (!eof() && (cur() == '.' || cur() == ']')) break;
                continue;
            }
            std::string pred = term();
            if (pred.empty()) break;
            for (;;) {
                std::string obj = term();
                if (!obj.empty()) emit(subj, pred, obj);
                ws();
                if (!eof() && cur() == ',') { ++i_; continue; }
                break;
            }
            ws();
            if (!eof() && cur() == ';') { ++i_; continue; }
            break;
        }
    }
};

std::string trim(const std::string& s) {
    size_t b = 0, e = s.size();
    while (b < e && std::isspace((unsigned char)s[b])) ++b;
    while (e > b && std::isspace((unsigned char)s[e - 1])) --e;
    return s.substr(b, e - b);
}

}  // namespace

std::unordered_map<std::string, std::string> scan_prefix_lines(const std::string& text) {
    std::unordered_map<std::string, std::string> out;
    std::istringstream in(text);
    std::string line;
    while (std::getline(in, line)) {
        std::string t = trim(line);
        if (t.empty()) continue;
        std::string low = t;
        for (auto& ch : low) ch = (char)std::tolower((unsigned char)ch);
        size_t start;
        if (low.compare(0, 7, "@prefix") == 0) start = 7;
        else if (low.compare(0, 6, "prefix") == 0) start = 6;
        else continue;
        size_t colon = t.find(':', start);
        if (colon == std::string::npos) continue;
        std::string nm = trim(t.substr(start, colon - start));
        size_t lt = t.find('<', colon);
        if (lt == std::string::npos) continue;
        size_t gt = t.find('>', lt);
        if (gt == std::string::npos) continue;
        std::string iri = t.substr(lt + 1, gt - lt - 1);
        if (!nm.empty() && !iri.empty()) out[nm] = iri;
    }
    return out;
}

bool parse_turtle(const std::string& path, HostGraph& g, std::vector<int32_t>& S, std::vector<int32_t>& P,
                  std::vector<int32_t>& O, std::string& err) {
    std::ifstream f(path, std::ios::binary);
    if (!f) {
        err = "cannot open " + path;
        return false;
    }
    std::string text((std::istreambuf_iterator<char>(f)), std::istreambuf_iterator<char>());
    TurtleReader(text, g, S, P, O).run();
    g.prefixes = scan_prefix_lines(text);
    return true;
}

bool load_turtle(const std::string& path, HostGraph& g, std::string& err) {
    std::vector<int32_t> S, P, O;
    if (!parse_turtle(path, g, S, P, O, err)) return false;
    build_graph(g, S, P, O, (int64_t)g.terms.size());
    return true;
}

}  // namespace gr
