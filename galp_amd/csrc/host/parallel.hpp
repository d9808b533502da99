// parallel.hpp -- minimal std::thread parallel-for for the host-side graph
// builders (no OpenMP runtime in libgalp, so it never clashes with a
// caller's).  GALP_HOST_THREADS overrides the worker count (default: the
// hardware concurrency, at most 16).
#pragma once
#include <algorithm>
#include <atomic>
#include <cstdint>
#include <cstdlib>
#include <thread>
#include <vector>

namespace galp {

inline int host_threads() {
    if (const char *s = std::getenv("GALP_HOST_THREADS")) {
        int t = std::atoi(s);
        if (t > 0) return t;
    }
    unsigned hc = std::thread::hardware_concurrency();
    return (int)std::max(1u, std::min(hc, 16u));
}

// Static split of [b, e) into contiguous chunks, one per worker.
template <class F>
void parallel_for(int64_t b, int64_t e, F &&f) {
    int64_t len = e - b;
    int nt = host_threads();
    if (len <= 0) return;
    if (nt == 1 || len < 4096) {
        for (int64_t i = b; i < e; ++i) f(i);
        return;
    }
    std::vector<std::thread> th;
    for (int t = 0; t < nt; ++t) {
        int64_t lo = b + len * t / nt, hi = b + len * (t + 1) / nt;
        th.emplace_back([lo, hi, &f] {
            for (int64_t i = lo; i < hi; ++i) f(i);
        });
    }
    for (auto &x : th) x.join();
}

// Dynamic chunks of `grain` items handed out by an atomic cursor.
template <class F>
void parallel_for_dynamic(int64_t b, int64_t e, int64_t grain, F &&f) {
    int nt = host_threads();
    if (e <= b) return;
    if (nt == 1 || e - b <= grain) {
        for (int64_t i = b; i < e; ++i) f(i);
        return;
    }
    std::atomic<int64_t> cur(b);
    std::vector<std::thread> th;
    for (int t = 0; t < nt; ++t) {
        th.emplace_back([&] {
            for (;;) {
                int64_t lo = cur.fetch_add(grain);
                if (lo >= e) break;
                int64_t hi = std::min(e, lo + grain);
                for (int64_t i = lo; i < hi; ++i) f(i);
            }
        });
    }
    for (auto &x : th) x.join();
}

}  // namespace galp
