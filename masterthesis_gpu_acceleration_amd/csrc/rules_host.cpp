// rules_host.cpp -- rule text and JSON front ends.
#include "rules_host.hpp"

#include <cctype>
#include <sstream>

namespace gr {

namespace {

std::string trim(const std::string& s) {
    size_t b = 0, e = s.size();
    while (b < e && std::isspace((unsigned char)s[b])) ++b;
    while (e > b && std::isspace((unsigned char)s[e - 1])) --e;
    return s.substr(b, e - b);
}

struct VarNumbering {
    std::unordered_map<std::string, int32_t> ids;
    TermIR get(const std::string& name) {
        auto it = ids.find(name);
        if (it != ids.end()) return {true, it->second};
        int32_t id = (int32_t)ids.size();
        ids.emplace(name, id);
        return {true, id};
    }
};

}  // namespace

std::string TermResolver::expand(const std::string& tok) const {
    if (tok.empty()) return tok;
    if (tok.front() == '<' && tok.back() == '>') return tok.substr(1, tok.size() - 2);
    if (tok.front() == '?' || tok.front() == '"') return tok;
    size_t c = tok.find(':');
    if (c == std::string::npos) return tok;
    auto it = g_.prefixes.find(tok.substr(0, c));
    if (it != g_.prefixes.end()) return it->second + tok.substr(c + 1);
    return tok;
}

int32_t TermResolver::resolve(const std::string& expanded) {
    int32_t id = g_.terms.find(expanded);
    if (id >= 0) return id;
    auto it = extra_.find(expanded);
    if (it != extra_.end()) return it->second;
    int32_t nid = (int32_t)(g_.terms.size() + extraNames_.size());
    if (nid < g_.nE) nid = (int32_t)g_.nE + (int32_t)extraNames_.size();
    extra_.emplace(expanded, nid);
    extraNames_.push_back(expanded);
    return nid;
}

const std::string* TermResolver::name(int32_t id) const {
    if (id >= 0 && (size_t)id < g_.terms.size()) return &g_.terms.strs[id];
    int64_t k = (int64_t)id - std::max<int64_t>((int64_t)g_.terms.size(), g_.nE);
    if (k >= 0 && (size_t)k < extraNames_.size()) return &extraNames_[k];
    return nullptr;
}

namespace {

TermIR text_term(const std::string& tok, VarNumbering& vars, TermResolver& tr) {
    std::string t = trim(tok);
    if (t.empty()) throw ParseError("empty term");
    if (t.front() == '?') return vars.get(t);
    return {false, tr.resolve(tr.expand(t))};
}

AtomIR text_atom(const std::string& text, VarNumbering& vars, TermResolver& tr, const char* what) {
    std::string s = trim(text);
    if (s.size() < 2 || s.front() != '(' || s.back() != ')')
        throw ParseError(std::string("invalid ") + what + " atom: " + text);
    std::istringstream in(trim(s.substr(1, s.size() - 2)));
    std::string a, p, o;
    in >> a >> p >> o;
    if (a.empty() || p.empty() || o.empty()) throw ParseError(std::string("invalid ") + what + " atom: " + text);
    AtomIR at;
    at.s = text_term(a, vars, tr);
    at.o = text_term(o, vars, tr);
    at.pterm = tr.resolve(tr.expand(p));
    return at;
}

}  // namespace

RuleIR parse_rule_text(const std::string& line, TermResolver& tr) {
    std::string t = trim(line);
    if (t.empty()) throw ParseError("empty rule line");
    size_t bar = t.find('|');
    if (bar != std::string::npos) t = trim(t.substr(0, bar));
    size_t arrow = t.find("=>");
    if (arrow == std::string::npos) throw ParseError("rule line missing => : " + line);
    std::string body = trim(t.substr(0, arrow));
    std::string head = trim(t.substr(arrow + 2));

    // split the body at '^' outside parentheses
    std::vector<std::string> parts;
    std::string cur;
    int depth = 0;
    for (char ch : body) {
        if (ch == '(') ++depth;
        else if (ch == ')') --depth;
        if (ch == '^' && depth == 0) {
            std::string x = trim(cur);
            if (!x.empty()) parts.push_back(x);
            cur.clear();
        } else {
            cur += ch;
        }
    }
    if (!trim(cur).empty()) parts.push_back(trim(cur));

    VarNumbering vars;
    RuleIR r;
    for (auto& a : parts) r.body.push_back(text_atom(a, vars, tr, "body"));
    r.head = text_atom(head, vars, tr, "head");
    r.nvars = (int32_t)vars.ids.size();
    return r;
}

// ------------------------------------------------------------------- JSON
// A small scanner over the export format: it looks fields up by key the way
// RDFRules' exporter writes them (one object per atom, "type"/"value" per
// term) and honours string quoting when matching braces.

namespace {

size_t skip_string(const std::string& s, size_t i) {  // i at opening quote
    for (++i; i < s.size(); ++i) {
        if (s[i] == '\\') { ++i; continue; }
        if (s[i] == '"') return i + 1;
    }
    return s.size();
}

// Balanced {...} starting at s[i] == '{'.
std::string object_at(const std::string& s, size_t i) {
    if (i >= s.size() || s[i] != '{') return "";
    int depth = 0;
    for (size_t j = i; j < s.size();) {
        char c = s[j];
        if (c == '"') { j = skip_string(s, j); continue; }
        if (c == '{') ++depth;
        if (c == '}' && --depth == 0) return s.substr(i, j - i + 1);
        ++j;
    }
    return "";
}

// Position just after `"key"` followed by ':' at nesting depth 1 of obj.
size_t field(const std::string& obj, const std::string& key) {
    int depth = 0;
    for (size_t j = 0; j < obj.size();) {
        char c = obj[j];
        if (c == '"') {
            size_t e = skip_string(obj, j);
            if (depth == 1 && obj.compare(j + 1, key.size(), key) == 0 && j + 1 + key.size() + 1 == e) {
                size_t k = e;
                while (k < obj.size() && std::isspace((unsigned char)obj[k])) ++k;
                if (k < obj.size() && obj[k] == ':') {
                    ++k;
                    while (k < obj.size() && std::isspace((unsigned char)obj[k])) ++k;
                    return k;
                }
            }
            j = e;
            continue;
        }
        if (c == '{' || c == '[') ++depth;
        if (c == '}' || c == ']') --depth;
        ++j;
    }
    return std::string::npos;
}

std::string string_field(const std::string& obj, const std::string& key) {
    size_t k = field(obj, key);
    if (k == std::string::npos || k >= obj.size() || obj[k] != '"') return "";
    size_t e = skip_string(obj, k);
    std::string raw = obj.substr(k + 1, e - k - 2);
    std::string out;
    for (size_t i = 0; i < raw.size(); ++i) {
        if (raw[i] == '\\' && i + 1 < raw.size()) { out += raw[++i]; continue; }
        out += raw[i];
    }
    return out;
}

TermIR json_term(const std::string& obj, VarNumbering& vars, TermResolver& tr) {
    std::string type = string_field(obj, "type");
    std::string value = string_field(obj, "value");
    if (value.empty()) throw ParseError("JSON term missing 'value': " + obj);
    if (type == "variable") return vars.get(value);
    return {false, tr.resolve(tr.expand(value))};
}

AtomIR json_atom(const std::string& obj, VarNumbering& vars, TermResolver& tr) {
    size_t ks = field(obj, "subject"), ko = field(obj, "object");
    if (ks == std::string::npos || ko == std::string::npos)
        throw ParseError("JSON atom missing subject/object: " + obj);
    std::string pred = string_field(obj, "predicate");
    if (pred.empty()) throw ParseError("JSON atom missing 'predicate': " + obj);
    AtomIR a;
    a.s = json_term(object_at(obj, ks), vars, tr);
    a.o = json_term(object_at(obj, ko), vars, tr);
    a.pterm = tr.resolve(tr.expand(pred));
    return a;
}

}  // namespace

RuleIR parse_rule_json(const std::string& json, TermResolver& tr) {
    size_t b = json.find('{');
    std::string top = object_at(json, b == std::string::npos ? json.size() : b);
    if (top.empty()) throw ParseError("JSON rule is not an object");
    size_t kb = field(top, "body");
    if (kb == std::string::npos || top[kb] != '[') throw ParseError("JSON rule missing 'body' array");
    size_t kh = field(top, "head");
    if (kh == std::string::npos) throw ParseError("JSON rule missing 'head'");
    VarNumbering vars;
    RuleIR r;
    // body atoms: the objects directly inside the array
    int depth = 0;
    for (size_t j = kb; j < top.size();) {
        char c = top[j];
        if (c == '"') { j = skip_string(top, j); continue; }
        if (c == '[') ++depth;
        else if (c == ']') { if (--depth == 0) break; }
        else if (c == '{' && depth == 1) {
            std::string o = object_at(top, j);
            r.body.push_back(json_atom(o, vars, tr));
            j += o.size();
            continue;
        }
        ++j;
    }
    r.head = json_atom(object_at(top, kh), vars, tr);
    r.nvars = (int32_t)vars.ids.size();
    return r;
}

// ---------------------------------------------------------- word records

RuleIR rule_from_words(const int32_t* w) {
    RuleIR r;
    std::unordered_map<int32_t, int32_t> vm;
    auto term = [&](int32_t t) -> TermIR {
        if (t >= 0) return {false, t};
        auto it = vm.find(t);
        if (it != vm.end()) return {true, it->second};
        int32_t id = (int32_t)vm.size();
        vm.emplace(t, id);
        return {true, id};
    };
    int32_t nb = w[0];
    int32_t n = nb < 0 ? 0 : (nb > kMaxBody ? kMaxBody : nb);
    for (int32_t j = 0; j < n; ++j) {
        const int32_t* a = w + 4 + 3 * j;
        AtomIR at;
        at.s = term(a[0]);
        at.o = term(a[2]);
        at.pterm = a[1];
        r.body.push_back(at);
    }
    r.head.s = term(w[1]);
    r.head.o = term(w[3]);
    r.head.pterm = w[2];
    r.nvars = (int32_t)vm.size();
    if (nb > kMaxBody) r.body.resize(kMaxBody + 1);  // marks an over-long body for the skip rule
    return r;
}

void rule_to_words(const RuleIR& r, int32_t* w) {
    for (int k = 0; k < 28; ++k) w[k] = 0;
    auto enc = [](const TermIR& t) { return t.var ? -(t.v + 1) : t.v; };
    w[0] = (int32_t)r.body.size();
    w[1] = enc(r.head.s);
    w[2] = r.head.pterm;
    w[3] = enc(r.head.o);
    for (size_t j = 0; j < r.body.size() && j < (size_t)kMaxBody; ++j) {
        w[4 + 3 * j] = enc(r.body[j].s);
        w[5 + 3 * j] = r.body[j].pterm;
        w[6 + 3 * j] = enc(r.body[j].o);
    }
}

}  // namespace gr
