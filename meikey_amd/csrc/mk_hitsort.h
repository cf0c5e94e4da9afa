// Host-side ordering of exact hits by (search, index).
//
// The confirm kernel appends hits in atomic order; the caller's list is ordered by search,
// then index, so a search's first record is its lowest exact hit (the one index that
// Engine::launch, src/service/gpu.rs:825, hands back per launch). A counting sort
// groups the hits by search; each group is then put in index order by an LSD radix sort
// (11-bit digits) of index - the group's smallest index, with only as many passes as the
// group's index range needs (three for a 2^26-key search, none for a single hit). The
// groups are independent, so a large hit list sorts them on a few host threads, each
// taking a run of searches with about the same number of hits.
#pragma once

#include <algorithm>
#include <cstdint>
#include <system_error>
#include <thread>
#include <vector>

// Host threads for n items of a sort or gather: one below 2^16 items, at most 8.
inline uint32_t mk_host_threads(uint32_t n) {
    if (n < (1u << 16)) return 1;
    const uint32_t hw = std::thread::hardware_concurrency();
    return std::max(1u, std::min(8u, hw));
}

// Runs f(t) for t in [0, T) on T threads (t = 0 on the caller's); where a thread cannot be
// started its part runs on the caller's thread.
template <class F>
void mk_host_parallel(uint32_t T, F f) {
    std::vector<std::thread> th;
    std::vector<uint32_t> left;
    for (uint32_t t = 1; t < T; t++) {
        try {
            th.emplace_back(f, t);
        } catch (const std::system_error&) {
            left.push_back(t);
        }
    }
    f(0u);
    for (uint32_t t : left) f(t);
    for (auto& x : th) x.join();
}

// Buffers of the sort, kept by the caller between calls (fresh ones cost their page
// faults on every call).
struct mk_sort_pool {
    std::vector<uint64_t> key, key2;
    std::vector<uint32_t> order, tmp, start;
};

// Returns order[0..n): order[i] is the position in hs of the i-th hit in (search, index)
// order. H has fields `search` (uint32_t) and `index` (uint64_t).
template <class H>
const uint32_t* mk_sort_hits(const H* hs, uint32_t n, mk_sort_pool& p) {
    p.order.resize(n);
    if (n == 0) return p.order.data();
    uint32_t ns = 0;
    for (uint32_t i = 0; i < n; i++) ns = std::max<uint32_t>(ns, hs[i].search + 1);
    p.start.assign((size_t)ns + 1, 0);
    for (uint32_t i = 0; i < n; i++) p.start[hs[i].search + 1]++;
    for (uint32_t s = 0; s < ns; s++) p.start[s + 1] += p.start[s];
    p.key.resize(n);
    p.key2.resize(n);
    p.tmp.resize(n);
    uint64_t* const key = p.key.data();
    uint32_t* const order = p.order.data();
    {
        std::vector<uint32_t> at(p.start.begin(), p.start.end() - 1);
        for (uint32_t i = 0; i < n; i++) {
            const uint32_t j = at[hs[i].search]++;
            order[j] = i;
            key[j] = hs[i].index;
        }
    }
    // each group [start[s], start[s + 1]) is sorted in place; groups touch disjoint parts
    // of key, key2, order and tmp
    auto sort_groups = [&](uint32_t s0, uint32_t s1) {
        for (uint32_t s = s0; s < s1; s++) {
            const uint32_t b = p.start[s], e = p.start[s + 1];
            if (e - b < 2) continue;
            uint64_t lo = ~0ull, hi = 0;
            for (uint32_t i = b; i < e; i++) {
                lo = std::min(lo, key[i]);
                hi = std::max(hi, key[i]);
            }
            for (uint32_t i = b; i < e; i++) key[i] -= lo;
            uint64_t *k = key, *k2 = p.key2.data();
            uint32_t *o = order, *o2 = p.tmp.data();
            const uint64_t range = hi - lo;
            for (int sh = 0; sh < 64 && (range >> sh) != 0; sh += 11) {
                uint32_t cnt[2049] = {0};
                for (uint32_t i = b; i < e; i++) cnt[((k[i] >> sh) & 2047u) + 1]++;
                for (int d = 0; d < 2048; d++) cnt[d + 1] += cnt[d];
                for (uint32_t i = b; i < e; i++) {
                    const uint32_t j = b + cnt[(k[i] >> sh) & 2047u]++;
                    k2[j] = k[i];
                    o2[j] = o[i];
                }
                std::swap(k, k2);
                std::swap(o, o2);
            }
            if (o != order) std::copy(o + b, o + e, order + b);
        }
    };
    const uint32_t T = std::min(mk_host_threads(n), ns);
    if (T <= 1) {
        sort_groups(0, ns);
        return order;
    }
    // thread t takes the searches whose first hit lies in [n t / T, n (t + 1) / T)
    std::vector<uint32_t> cut(T + 1);
    for (uint32_t t = 0; t <= T; t++) {
        const uint64_t at = (uint64_t)n * t / T;
        cut[t] = t == T ? ns : (uint32_t)(std::lower_bound(p.start.begin(), p.start.begin() + ns, at) - p.start.begin());
    }
    mk_host_parallel(T, [&](uint32_t t) { sort_groups(cut[t], cut[t + 1]); });
    return order;
}
