// snapshot.cpp -- binary graph snapshot (see snapshot.hpp).
//
// Layout, little-endian, every section 64-byte aligned:
//   Header (256 bytes): magic "GRSNAP1\0", version, byte-order mark, the
//     counts (nE, T, P, terms, prefixes) and the offset and length in bytes
//     of each section;
//   predTerm[P], predOff[P+1], spoKey[T], spoVal[T], posKey[T], posVal[T]
//     (int32; the HostGraph arrays as built: triples grouped by predicate,
//     SPO sorted by (s, o), POS by (o, s) within a predicate);
//   termOff[terms+1] (int64) and the term bytes (term i is
//     bytes[termOff[i], termOff[i+1]); ids are dense in first-seen order);
//   prefixes: per entry, u32 name length, name, u32 IRI length, IRI.
// The distinct-key lists (keysS/keysO with their offsets) are derived from
// the sorted arrays on load.
#include "snapshot.hpp"

#include <fcntl.h>
#include <sys/mman.h>
#include <sys/stat.h>
#include <unistd.h>

#include <cstdio>
#include <cstring>

namespace gr {

namespace {

constexpr char kMagic[8] = {'G', 'R', 'S', 'N', 'A', 'P', '1', '\0'};
constexpr uint32_t kVersion = 1;
constexpr uint32_t kByteOrder = 0x01020304u;

enum Section { kPredTerm, kPredOff, kSpoKey, kSpoVal, kPosKey, kPosVal, kTermOff, kTermBytes, kPrefixes, kSections };

struct Header {
    char magic[8];
    uint32_t version;
    uint32_t byteOrder;
    int64_t nE, T, P, nTerms, nPrefixes;
    int64_t off[kSections];
    int64_t len[kSections];
    uint8_t pad[256 - 16 - 5 * 8 - 2 * kSections * 8];
};
static_assert(sizeof(Header) == 256, "snapshot header is 256 bytes");

int64_t align64(int64_t x) { return (x + 63) / 64 * 64; }

// Order-independent hash of a triple, summed over a direction's arrays:
// both directions must give the same sum.
uint64_t mix(uint32_t p, uint32_t s, uint32_t o) {
    uint64_t h = ((uint64_t)p << 42) ^ ((uint64_t)s << 21) ^ (uint64_t)o;
    h ^= h >> 33;
    h *= 0xff51afd7ed558ccdULL;
    h ^= h >> 33;
    h *= 0xc4ceb9fe1a85ec53ULL;
    h ^= h >> 33;
    return h;
}

struct Mapping {
    void* p = MAP_FAILED;
    size_t n = 0;
    int fd = -1;
    ~Mapping() {
        if (p != MAP_FAILED) munmap(p, n);
        if (fd >= 0) close(fd);
    }
};

// Distinct keys per predicate from one direction's sorted key array.
void derive_keys(const std::vector<int32_t>& predOff, const std::vector<int32_t>& key, std::vector<int32_t>& keys,
                 std::vector<int32_t>& keysOff) {
    const size_t P = predOff.size() - 1;
    keys.clear();
    keysOff.assign(P + 1, 0);
    for (size_t p = 0; p < P; ++p) {
        for (int32_t e = predOff[p]; e < predOff[p + 1]; ++e)
            if (e == predOff[p] || key[e] != key[e - 1]) keys.push_back(key[e]);
        keysOff[p + 1] = (int32_t)keys.size();
    }
}

}  // namespace

bool save_snapshot(const HostGraph& g, const std::string& path, std::string& err) {
    Header h;
    std::memset(&h, 0, sizeof(h));
    std::memcpy(h.magic, kMagic, 8);
    h.version = kVersion;
    h.byteOrder = kByteOrder;
    h.nE = g.nE;
    h.T = g.T();
    h.P = g.P();
    h.nTerms = (int64_t)g.terms.size();
    h.nPrefixes = (int64_t)g.prefixes.size();
    std::vector<int64_t> termOff((size_t)h.nTerms + 1, 0);
    for (int64_t i = 0; i < h.nTerms; ++i) termOff[i + 1] = termOff[i] + (int64_t)g.terms.strs[i].size();
    std::string pre;
    for (const auto& kv : g.prefixes) {
        for (const std::string* s : {&kv.first, &kv.second}) {
            const uint32_t n = (uint32_t)s->size();
            pre.append(reinterpret_cast<const char*>(&n), 4);
            pre.append(*s);
        }
    }
    h.len[kPredTerm] = h.P * 4;
    h.len[kPredOff] = (h.P + 1) * 4;
    h.len[kSpoKey] = h.len[kSpoVal] = h.len[kPosKey] = h.len[kPosVal] = h.T * 4;
    h.len[kTermOff] = (h.nTerms + 1) * 8;
    h.len[kTermBytes] = termOff.back();
    h.len[kPrefixes] = (int64_t)pre.size();
    int64_t at = (int64_t)sizeof(Header);
    for (int s = 0; s < kSections; ++s) {
        h.off[s] = at;
        at = align64(at + h.len[s]);
    }
    FILE* f = std::fopen(path.c_str(), "wb");
    if (!f) {
        err = "cannot write " + path;
        return false;
    }
    bool ok = std::fwrite(&h, sizeof(h), 1, f) == 1;
    int64_t pos = (int64_t)sizeof(Header);
    auto put = [&](int s, const void* data) {
        if (!ok) return;
        static const char zeros[64] = {};
        if (h.off[s] > pos) ok = std::fwrite(zeros, 1, (size_t)(h.off[s] - pos), f) == (size_t)(h.off[s] - pos);
        pos = h.off[s];
        if (ok && h.len[s] > 0) ok = std::fwrite(data, 1, (size_t)h.len[s], f) == (size_t)h.len[s];
        pos += h.len[s];
    };
    std::string bytes;
    bytes.reserve((size_t)termOff.back());
    for (const std::string& t : g.terms.strs) bytes += t;
    put(kPredTerm, g.predTerm.data());
    put(kPredOff, g.predOff.data());
    put(kSpoKey, g.spoKey.data());
    put(kSpoVal, g.spoVal.data());
    put(kPosKey, g.posKey.data());
    put(kPosVal, g.posVal.data());
    put(kTermOff, termOff.data());
    put(kTermBytes, bytes.data());
    put(kPrefixes, pre.data());
    ok = (std::fclose(f) == 0) && ok;
    if (!ok) err = "write failed: " + path;
    return ok;
}

bool load_snapshot(const std::string& path, HostGraph& g, std::string& err) {
    Mapping m;
    m.fd = open(path.c_str(), O_RDONLY);
    if (m.fd < 0) {
        err = "cannot open " + path;
        return false;
    }
    struct stat st;
    if (fstat(m.fd, &st) != 0 || st.st_size < (off_t)sizeof(Header)) {
        err = "not a snapshot (too short)";
        return false;
    }
    m.n = (size_t)st.st_size;
    m.p = mmap(nullptr, m.n, PROT_READ, MAP_PRIVATE, m.fd, 0);
    if (m.p == MAP_FAILED) {
        err = "cannot map " + path;
        return false;
    }
    const char* base = static_cast<const char*>(m.p);
    Header h;
    std::memcpy(&h, base, sizeof(h));
    if (std::memcmp(h.magic, kMagic, 8) != 0 || h.version != kVersion || h.byteOrder != kByteOrder) {
        err = "not a snapshot of this version";
        return false;
    }
    if (h.nE < 0 || h.nE > INT32_MAX || h.T < 0 || h.T > INT32_MAX / 2 || h.P < 0 || h.P > INT32_MAX - 1 ||
        h.nTerms < 0 || h.nTerms > h.nE || h.nPrefixes < 0) {
        err = "snapshot counts out of range";
        return false;
    }
    const int64_t want[kSections] = {h.P * 4, (h.P + 1) * 4, h.T * 4, h.T * 4, h.T * 4, h.T * 4,
                                     (h.nTerms + 1) * 8, -1, -1};
    for (int s = 0; s < kSections; ++s) {
        if (h.off[s] < (int64_t)sizeof(Header) || h.len[s] < 0 || h.off[s] > (int64_t)m.n ||
            h.len[s] > (int64_t)m.n - h.off[s] || h.off[s] % 8 != 0 || (want[s] >= 0 && h.len[s] != want[s])) {
            err = "snapshot section out of bounds";
            return false;
        }
    }
    auto i32 = [&](int s) { return reinterpret_cast<const int32_t*>(base + h.off[s]); };
    const int32_t P = (int32_t)h.P;
    const int64_t T = h.T;
    g = HostGraph();
    g.nE = h.nE;
    g.predTerm.assign(i32(kPredTerm), i32(kPredTerm) + P);
    g.predOff.assign(i32(kPredOff), i32(kPredOff) + P + 1);
    if (g.predOff[0] != 0 || g.predOff[P] != T) {
        err = "snapshot predicate offsets inconsistent";
        return false;
    }
    for (int32_t p = 0; p < P; ++p) {
        if (g.predOff[p + 1] < g.predOff[p] || g.predTerm[p] < 0 || g.predTerm[p] >= h.nE ||
            !g.predOf.emplace(g.predTerm[p], p).second) {
            err = "snapshot predicate table inconsistent";
            return false;
        }
    }
    g.spoKey.assign(i32(kSpoKey), i32(kSpoKey) + T);
    g.spoVal.assign(i32(kSpoVal), i32(kSpoVal) + T);
    g.posKey.assign(i32(kPosKey), i32(kPosKey) + T);
    g.posVal.assign(i32(kPosVal), i32(kPosVal) + T);
    uint64_t hashSpo = 0, hashPos = 0;
    for (int d = 0; d < 2; ++d) {
        const std::vector<int32_t>& key = d == 0 ? g.spoKey : g.posKey;
        const std::vector<int32_t>& val = d == 0 ? g.spoVal : g.posVal;
        uint64_t& hash = d == 0 ? hashSpo : hashPos;
        for (int32_t p = 0; p < P; ++p) {
            for (int32_t e = g.predOff[p]; e < g.predOff[p + 1]; ++e) {
                if (key[e] < 0 || key[e] >= h.nE || val[e] < 0 || val[e] >= h.nE) {
                    err = "snapshot term id outside the row space";
                    return false;
                }
                if (e > g.predOff[p] && (key[e] < key[e - 1] || (key[e] == key[e - 1] && val[e] <= val[e - 1]))) {
                    err = "snapshot rows not sorted";
                    return false;
                }
                hash += d == 0 ? mix((uint32_t)p, (uint32_t)key[e], (uint32_t)val[e])
                               : mix((uint32_t)p, (uint32_t)val[e], (uint32_t)key[e]);
            }
        }
    }
    if (hashSpo != hashPos) {
        err = "snapshot SPO and POS arrays hold different triples";
        return false;
    }
    derive_keys(g.predOff, g.spoKey, g.keysS, g.keysSOff);
    derive_keys(g.predOff, g.posKey, g.keysO, g.keysOOff);
    // terms
    const int64_t* toff = reinterpret_cast<const int64_t*>(base + h.off[kTermOff]);
    const char* tb = base + h.off[kTermBytes];
    if (toff[0] != 0 || toff[h.nTerms] != h.len[kTermBytes]) {
        err = "snapshot term table inconsistent";
        return false;
    }
    g.terms.strs.reserve((size_t)h.nTerms);
    g.terms.ids.reserve((size_t)h.nTerms);
    for (int64_t i = 0; i < h.nTerms; ++i) {
        if (toff[i + 1] < toff[i]) {
            err = "snapshot term table inconsistent";
            return false;
        }
        g.terms.strs.emplace_back(tb + toff[i], (size_t)(toff[i + 1] - toff[i]));
        if (!g.terms.ids.emplace(g.terms.strs.back(), (int32_t)i).second) {
            err = "snapshot term table repeats a term";
            return false;
        }
    }
    // prefixes
    const char* pp = base + h.off[kPrefixes];
    int64_t left = h.len[kPrefixes];
    for (int64_t i = 0; i < h.nPrefixes; ++i) {
        std::string parts[2];
        for (std::string& part : parts) {
            uint32_t n = 0;
            if (left < 4) {
                err = "snapshot prefixes truncated";
                return false;
            }
            std::memcpy(&n, pp, 4);
            pp += 4;
            left -= 4;
            if ((int64_t)n > left) {
                err = "snapshot prefixes truncated";
                return false;
            }
            part.assign(pp, n);
            pp += n;
            left -= n;
        }
        g.prefixes[parts[0]] = parts[1];
    }
    return true;
}

}  // namespace gr
