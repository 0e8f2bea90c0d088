// graph_device.hip -- the grouped adjacency arrays of a graph built with the
// GPU's radix sort instead of two host sorts.
//
// The host maps predicates to dense ids in first-seen order (the order the
// host build uses) and uploads the triples; one kernel packs each triple as
// two 64-bit keys, (p, s, o) and (p, o, s), with p in the high bits; the LSD
// radix sort of sort_unique.hip sorts both key arrays; the host reads them
// back and, in one pass each, drops repeated triples and splits the keys
// into the predicate offsets and the SPO / POS key and value arrays --
// exactly the arrays build_graph produces.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "graph_device.hpp"
#include "sort_unique.hpp"

namespace gr {

namespace {

__global__ void k_triple_keys(const int32_t* __restrict__ s, const int32_t* __restrict__ p,
                              const int32_t* __restrict__ o, int64_t n, int bits, unsigned long long* spo,
                              unsigned long long* pos) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const unsigned long long hp = (unsigned long long)(uint32_t)p[i] << (2 * bits);
        const unsigned long long ss = (uint32_t)s[i], oo = (uint32_t)o[i];
        spo[i] = hp | (ss << bits) | oo;
        pos[i] = hp | (oo << bits) | ss;
    }
}

struct DevMem {
    void* p = nullptr;
    ~DevMem() {
        if (p) (void)hipFree(p);
    }
    bool alloc(size_t bytes) { return hipMalloc(&p, bytes < 1 ? 1 : bytes) == hipSuccess; }
    template <class T>
    T* as() const { return static_cast<T*>(p); }
};

#define GD_CHECK(x)                                                             \
    do {                                                                        \
        hipError_t e_ = (x);                                                    \
        if (e_ != hipSuccess) {                                                 \
            err = std::string(#x) + ": " + hipGetErrorString(e_);               \
            return false;                                                       \
        }                                                                       \
    } while (0)

// Sorted keys (repeats adjacent) -> predicate offsets, keys, values.
void split(const std::vector<unsigned long long>& k, int bits, int32_t P, std::vector<int32_t>& off,
           std::vector<int32_t>& key, std::vector<int32_t>& val) {
    const unsigned long long mask = (1ull << bits) - 1ull;
    off.assign((size_t)P + 1, 0);
    key.clear();
    val.clear();
    key.reserve(k.size());
    val.reserve(k.size());
    for (size_t i = 0; i < k.size(); ++i) {
        if (i > 0 && k[i] == k[i - 1]) continue;
        off[(size_t)(k[i] >> (2 * bits)) + 1]++;
        key.push_back((int32_t)((k[i] >> bits) & mask));
        val.push_back((int32_t)(k[i] & mask));
    }
    for (int32_t p = 0; p < P; ++p) off[p + 1] += off[p];
}

void distinct_keys(const std::vector<int32_t>& off, const std::vector<int32_t>& key, std::vector<int32_t>& keys,
                   std::vector<int32_t>& keysOff) {
    const size_t P = off.size() - 1;
    keys.clear();
    keysOff.assign(P + 1, 0);
    for (size_t p = 0; p < P; ++p) {
        for (int32_t e = off[p]; e < off[p + 1]; ++e)
            if (e == off[p] || key[e] != key[e - 1]) keys.push_back(key[e]);
        keysOff[p + 1] = (int32_t)keys.size();
    }
}

}  // namespace

bool build_graph_device(HostGraph& g, const int32_t* s, const int32_t* p, const int32_t* o, int64_t n,
                        int64_t nTerms, std::string& err) {
    if (n < 0 || n > (int64_t)UINT32_MAX / 2 || nTerms < 1 || nTerms > INT32_MAX) {
        err = "triple count out of range for the device build";
        return false;
    }
    // dense predicates in first-seen order (a table over the term space when
    // it is small, else the hash map)
    g = HostGraph();
    g.nE = nTerms;
    std::vector<int32_t> dp((size_t)n);
    std::vector<int32_t> dense(nTerms <= ((int64_t)1 << 24) ? (size_t)nTerms : 0, -1);
    for (int64_t i = 0; i < n; ++i) {
        int32_t d;
        if (!dense.empty()) {
            d = dense[(size_t)p[i]];
        } else {
            auto it = g.predOf.find(p[i]);
            d = it == g.predOf.end() ? -1 : it->second;
        }
        if (d < 0) {
            d = (int32_t)g.predTerm.size();
            g.predTerm.push_back(p[i]);
            g.predOf.emplace(p[i], d);
            if (!dense.empty()) dense[(size_t)p[i]] = d;
        }
        dp[(size_t)i] = d;
    }
    const int32_t P = g.P();
    const int bits = bits_for((uint64_t)nTerms);
    const int keyBits = 2 * bits + bits_for((uint64_t)std::max<int32_t>(P, 1));
    if (keyBits > 64) {
        err = "term ids too wide for 64-bit triple keys";
        return false;
    }
    const uint32_t cap = (uint32_t)std::max<int64_t>(n, 1);
    DevMem ds, dd, dO, k0, k1, k2, k3, hist, cnt;
    if (!ds.alloc((size_t)n * 4) || !dd.alloc((size_t)n * 4) || !dO.alloc((size_t)n * 4) || !k0.alloc((size_t)cap * 8) ||
        !k1.alloc((size_t)cap * 8) || !k2.alloc((size_t)cap * 8) || !k3.alloc((size_t)cap * 8) ||
        !hist.alloc(sort_hist_words(cap) * 4) || !cnt.alloc(8)) {
        err = "not enough device memory for the device build";
        return false;
    }
    // uploads complete before the kernels start (synchronous copies)
    const unsigned long long n64 = (unsigned long long)n;
    if (n > 0) {
        GD_CHECK(hipMemcpy(ds.p, s, (size_t)n * 4, hipMemcpyHostToDevice));
        GD_CHECK(hipMemcpy(dd.p, dp.data(), (size_t)n * 4, hipMemcpyHostToDevice));
        GD_CHECK(hipMemcpy(dO.p, o, (size_t)n * 4, hipMemcpyHostToDevice));
    }
    GD_CHECK(hipMemcpy(cnt.p, &n64, 8, hipMemcpyHostToDevice));
    hipStream_t st = nullptr;
    GD_CHECK(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
    struct StreamGuard {
        hipStream_t s;
        ~StreamGuard() { (void)hipStreamDestroy(s); }
    } guard{st};
    if (n > 0) {
        const int64_t blocks = std::min<int64_t>((n + 255) / 256, 4096);
        hipLaunchKernelGGL(k_triple_keys, dim3((unsigned)blocks), dim3(256), 0, st, ds.as<int32_t>(), dd.as<int32_t>(),
                           dO.as<int32_t>(), n, bits, k0.as<unsigned long long>(), k2.as<unsigned long long>());
        GD_CHECK(hipGetLastError());
    }
    int dev = 0, cus = 256;
    if (hipGetDevice(&dev) == hipSuccess) {
        hipDeviceProp_t prop;
        if (hipGetDeviceProperties(&prop, dev) == hipSuccess && prop.multiProcessorCount > 0) cus = prop.multiProcessorCount;
    }
    unsigned long long* const spoKeys[2] = {k0.as<unsigned long long>(), k1.as<unsigned long long>()};
    unsigned long long* const posKeys[2] = {k2.as<unsigned long long>(), k3.as<unsigned long long>()};
    int a = 0, b = 0;
    if (n > 0) {
        a = launch_radix_sort(spoKeys, cnt.as<unsigned long long>(), cap, hist.as<uint32_t>(), keyBits, cus * 4, st);
        b = launch_radix_sort(posKeys, cnt.as<unsigned long long>(), cap, hist.as<uint32_t>(), keyBits, cus * 4, st);
        GD_CHECK(hipGetLastError());
    }
    GD_CHECK(hipStreamSynchronize(st));
    std::vector<unsigned long long> hs((size_t)n), hp((size_t)n);
    if (n > 0) {
        GD_CHECK(hipMemcpy(hs.data(), spoKeys[a], (size_t)n * 8, hipMemcpyDeviceToHost));
        GD_CHECK(hipMemcpy(hp.data(), posKeys[b], (size_t)n * 8, hipMemcpyDeviceToHost));
    }
    std::vector<int32_t>().swap(dp);
    std::vector<int32_t> off2;
    split(hs, bits, P, g.predOff, g.spoKey, g.spoVal);
    split(hp, bits, P, off2, g.posKey, g.posVal);
    if (off2 != g.predOff) {
        err = "device build: SPO and POS groupings differ";
        return false;
    }
    distinct_keys(g.predOff, g.spoKey, g.keysS, g.keysSOff);
    distinct_keys(g.predOff, g.posKey, g.keysO, g.keysOOff);
    return true;
}

bool device_build_wanted(int64_t n) {
    const char* e = std::getenv("GPU_RULES_DEVICE_BUILD");
    if (e && e[0] == '0') return false;
    if (e && e[0] == '1') return true;
    return n >= (int64_t)1 << 20;
}

}  // namespace gr
