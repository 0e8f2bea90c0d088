// refine.cpp -- one refinement level of a rule miner, on the host.
//
// A rule p(?a, ?b) <= body over variables is a key record of kKeyWords
// int32: head predicate, atom count n, then n atoms (predicate, subject,
// object) with variables numbered 0 (head subject), 1 (head object), 2, 3,
// ...; the rest zero. A key is canonical when, over the numberings of the
// non-head variables, its atoms form the smallest sorted sequence: rules that
// differ only in atom order and non-head variable names have one canonical
// key. The refinements of a rule add one atom over the given body predicates:
// a closing atom (two variables of the rule) or a dangling atom (one variable
// and a fresh one), in either orientation -- never a repeated atom or the head
// atom, and none that leaves more variables occurring once than the atoms
// still allowed can close (an atom closes at most two). This is the
// refinement step of AMIE-style miners such as RDFRules, the caller the
// reference's JNA binding serves (GPU_03/GpuRulesLib.scala); the reference
// itself has no miner.
#include <algorithm>
#include <array>
#include <atomic>
#include <cstdint>
#include <cstring>
#include <thread>
#include <vector>

#include "plan.hpp"
#include "refine.hpp"

namespace gr {

namespace {

constexpr int kKeyWords = 2 + 3 * kMaxBody;
using Key = std::array<int32_t, kKeyWords>;
using Atom = std::array<int32_t, 3>;

// The canonical key of n atoms over variable numbers.
Key canonical_key(int32_t hp, const Atom* atoms, int n) {
    int32_t others[2 * kMaxBody];
    int no = 0;
    for (int i = 0; i < n; ++i)
        for (int t = 1; t <= 2; ++t) {
            const int32_t v = atoms[i][(size_t)t];
            if (v > 1 && std::find(others, others + no, v) == others + no) others[no++] = v;
        }
    std::sort(others, others + no);
    int32_t perm[2 * kMaxBody];
    for (int i = 0; i < no; ++i) perm[i] = 2 + i;
    Atom cur[kMaxBody], best[kMaxBody];
    bool have = false;
    do {
        for (int i = 0; i < n; ++i) {
            cur[i] = atoms[i];
            for (int t = 1; t <= 2; ++t) {
                const int32_t v = atoms[i][(size_t)t];
                if (v > 1) cur[i][(size_t)t] = perm[std::find(others, others + no, v) - others];
            }
        }
        std::sort(cur, cur + n);
        if (!have || std::lexicographical_compare(cur, cur + n, best, best + n)) {
            std::copy(cur, cur + n, best);
            have = true;
        }
    } while (std::next_permutation(perm, perm + no));
    Key k{};
    k[0] = hp;
    k[1] = n;
    for (int i = 0; i < n; ++i)
        for (int t = 0; t < 3; ++t) k[(size_t)(2 + 3 * i + t)] = best[i][(size_t)t];
    return k;
}

void refine_one(const int32_t* rec, const int32_t* preds, int32_t nPreds, int32_t maxAtoms, std::vector<Key>& out) {
    const int32_t hp = rec[0];
    const int n = rec[1];
    if (n < 0 || n >= maxAtoms || n >= kMaxBody) return;
    Atom atoms[kMaxBody];
    int32_t cnt[2 * kMaxBody + 2] = {};
    cnt[0] = cnt[1] = 1;
    int32_t nv = 2;
    for (int i = 0; i < n; ++i) {
        atoms[i] = Atom{rec[2 + 3 * i], rec[3 + 3 * i], rec[4 + 3 * i]};
        for (int t = 1; t <= 2; ++t) {
            const int32_t v = atoms[i][(size_t)t];
            if (v < 0 || v >= 2 * kMaxBody + 2) return;   // not a key record
            ++cnt[v];
            nv = std::max(nv, v + 1);
        }
    }
    const int32_t fresh = nv;
    const int left = maxAtoms - n - 1;   // atoms still allowed after this one
    auto add = [&](int32_t x, int32_t y) {
        int32_t c2[2 * kMaxBody + 2];
        std::memcpy(c2, cnt, sizeof(c2));
        ++c2[x];
        ++c2[y];
        int once = 0;
        for (int32_t v = 0; v <= fresh && v < 2 * kMaxBody + 2; ++v) once += c2[v] == 1;
        if (once > 2 * left) return;
        for (int32_t j = 0; j < nPreds; ++j) {
            const Atom a{preds[j], x, y};
            if (x == 0 && y == 1 && a[0] == hp) continue;   // the head atom
            bool dup = false;
            for (int i = 0; i < n && !dup; ++i) dup = atoms[i] == a;
            if (dup) continue;
            atoms[n] = a;
            out.push_back(canonical_key(hp, atoms, n + 1));
        }
    };
    for (int32_t x = 0; x < nv; ++x)
        for (int32_t y = 0; y < nv; ++y)
            if (x != y) add(x, y);                          // closing
    if (fresh < 2 * kMaxBody + 2)
        for (int32_t x = 0; x < nv; ++x) {                  // dangling
            add(x, fresh);
            add(fresh, x);
        }
}

}  // namespace

int64_t refine_rules(const int32_t* parents, int32_t n, const int32_t* preds, int32_t nPreds, int32_t maxAtoms,
                     int32_t* out, int64_t cap) {
    if (n < 0 || nPreds < 0 || maxAtoms < 1 || maxAtoms > kMaxBody || (n > 0 && !parents) ||
        (nPreds > 0 && !preds) || cap < 0 || (cap > 0 && !out))
        return -1;
    const unsigned hw = std::thread::hardware_concurrency();
    const int nt = (int)std::max(1u, std::min({hw ? hw : 1u, 16u, (unsigned)std::max<int32_t>(1, n / 64)}));
    std::vector<std::vector<Key>> parts((size_t)nt);
    std::atomic<int32_t> next{0};
    auto work = [&](int t) {
        for (;;) {
            const int32_t i = next.fetch_add(64);
            if (i >= n) break;
            for (int32_t k = i; k < std::min(n, i + 64); ++k)
                refine_one(parents + (size_t)k * kKeyWords, preds, nPreds, maxAtoms, parts[(size_t)t]);
        }
        std::sort(parts[(size_t)t].begin(), parts[(size_t)t].end());
        parts[(size_t)t].erase(std::unique(parts[(size_t)t].begin(), parts[(size_t)t].end()), parts[(size_t)t].end());
    };
    std::vector<std::thread> pool;
    for (int t = 1; t < nt; ++t) pool.emplace_back(work, t);
    work(0);
    for (auto& th : pool) th.join();
    std::vector<Key> all;
    for (auto& p : parts) all.insert(all.end(), p.begin(), p.end());
    std::sort(all.begin(), all.end());
    all.erase(std::unique(all.begin(), all.end()), all.end());
    const int64_t m = (int64_t)all.size();
    for (int64_t i = 0; i < std::min(m, cap); ++i) std::memcpy(out + i * kKeyWords, all[(size_t)i].data(), sizeof(Key));
    return m;
}

}  // namespace gr
