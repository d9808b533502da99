// cpu_lp.cpp -- the serial CPU label propagation behind the CLI's serial key
// (method*5+policy outside 0..10), replacing LabelPropagator::run
// (label_propagation/include/label_propagator.h:37-82, selected by the
// `default:` case of label_propagation/src/main.cu:200-206).
//
// It computes the same synchronous update as the GPU methods (labels(t)
// from labels(t-1) only; the tie rule argmax_L (count, last position), i.e.
// the `max_count <= c` scan of hybrid_cpu_worker.cuh:183; degree-0 vertices
// keep their label; the n*1e-5 stop rule), so its labels equal theirs.  The
// reference's serial path differs on purpose in two points that make it
// irreproducible: it updates in place in a shuffled order
// (label_propagator.h:50,68-71) and flips a rand() coin on ties (:61-62).
//
// Counting: one open-addressing table (label -> count) per vertex, sized to
// the power of two at or above twice the degree and cleared through the
// list of slots it used, so a row costs O(degree).  The row is scanned in
// CSR order and a label whose running count reaches the maximum so far takes
// the lead, which is the argmax of (count, last position).

#include <algorithm>
#include <chrono>
#include <cstring>
#include <new>
#include <string>
#include <vector>

#include "galp_host.h"

namespace galp {
int host_fail(int code, const std::string &msg);
}

namespace {

struct RowTable {
    std::vector<int32_t> key;    // -1: empty
    std::vector<int32_t> cnt;
    std::vector<uint32_t> used;  // slots filled by the current row
    void reserve_for(int64_t deg) {
        size_t cap = 16;
        while (cap < (size_t)deg * 2) cap <<= 1;
        if (cap > key.size()) {
            key.assign(cap, -1);
            cnt.assign(cap, 0);
        }
    }
};

// New label of v: the label whose (count, last position) is largest.
int32_t update_row(const int64_t *off, const int32_t *nbr, const int32_t *prev, int64_t v, RowTable &t) {
    const int64_t b = off[v], e = off[v + 1];
    if (b == e) return prev[v];
    t.reserve_for(e - b);
    const uint32_t mask = (uint32_t)t.key.size() - 1;
    int32_t best = prev[v];
    int32_t best_cnt = 0;
    for (int64_t k = b; k < e; ++k) {
        const int32_t l = prev[nbr[k]];
        uint32_t s = ((uint32_t)l * 0x9E3779B1u) & mask;
        while (t.key[s] != -1 && t.key[s] != l) s = (s + 1) & mask;
        if (t.key[s] == -1) {
            t.key[s] = l;
            t.cnt[s] = 0;
            t.used.push_back(s);
        }
        const int32_t c = ++t.cnt[s];
        if (c >= best_cnt) {
            best_cnt = c;
            best = l;
        }
    }
    for (uint32_t s : t.used) t.key[s] = -1;
    t.used.clear();
    return best;
}

}  // namespace

extern "C" {

GALP_API int galp_cpu_lp_run(const galp_host_csr *g, int32_t niter, int32_t *labels_out, double *seconds,
                             int32_t *iterations) {
    if (!g || !labels_out || niter < 0 || g->n < 0 || (g->n > 0 && !g->offsets))
        return galp::host_fail(GALP_EINVAL, "galp_cpu_lp_run: invalid arguments");
    const int64_t n = g->n;
    const auto t0 = std::chrono::steady_clock::now();
    std::vector<int32_t> a, b;
    try {
        a.resize((size_t)n);
        b.resize((size_t)n);
    } catch (const std::bad_alloc &) {
        return galp::host_fail(GALP_ENOMEM, "galp_cpu_lp_run: out of host memory");
    }
    for (int64_t v = 0; v < n; ++v) a[(size_t)v] = (int32_t)v;
    RowTable t;
    int32_t done = 0;
    for (int32_t it = 0; it < niter; ++it) {
        int64_t changed = 0;
        for (int64_t v = 0; v < n; ++v) {
            const int32_t l = update_row(g->offsets, g->neighbors, a.data(), v, t);
            b[(size_t)v] = l;
            changed += (l != a[(size_t)v]);
        }
        std::swap(a, b);
        ++done;
        if ((double)changed <= (double)n * 1e-5) break;
    }
    if (n) std::memcpy(labels_out, a.data(), sizeof(int32_t) * (size_t)n);
    if (seconds) seconds[0] = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    if (iterations) *iterations = done;
    return GALP_OK;
}

}  // extern "C"
