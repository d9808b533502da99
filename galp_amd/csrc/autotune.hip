// autotune.hip -- picks each degree class's kernel configuration by timing.
//
// No reference counterpart (the reference fixes its block sizes at compile
// time, lfht.cuh:127-165).  For every degree class with vertices, hubs
// included, each of its configurations is launched `reps` times on the
// graph's current labels and timed with HIP events; the fastest is kept.
// The configurations are the method's: launch_class in lp_hash.hip
// (workgroup size x slots per lane and blocks per CU; for hubs the tile and
// merge workgroup sizes) and sort_class in lp_sort.hip (workgroup size x
// keys per thread, radix or bitonic, grids).  The choice never changes
// labels (every configuration computes the same rule), which is how it is
// tested.  Last, where an iteration may run its class kernels on three
// streams, whole iterations are timed on one stream and on three; the
// choice by size stays unless the other is more than 3 % faster (the
// overlap helps while the label array stays in the Infinity Cache and can
// hurt once it does not).  Choices can be cached in a
// text file, one line per class (class 12 = the stream choice, 1 or 3):
//   <arch> <method> <class> <log2 vertices> <log2 edges> <option>
// and are reused for a graph whose class has the same arch, method and sizes.

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <fstream>
#include <sstream>
#include <string>
#include <vector>

#include "internal.hpp"

namespace galp {

constexpr int kTuneOptions = 3;
constexpr int kStreamsLine = NUM_BUCKETS;   // the cache line of the stream choice

// The LP kernels of one in-core iteration (no bookkeeping kernel).
static int one_iteration(galp_graph *g) {
    return g->opts.method == GALP_METHOD_HASH ? launch_hash_iteration_incore(g)
                                              : launch_sort_iteration(g, full_view(g));
}

static int ilog2(int64_t x) { return x > 0 ? 63 - __builtin_clzll((unsigned long long)x) : -1; }

struct CacheLine {
    std::string arch;
    int method, cls, lv, le, opt;
};

static std::vector<CacheLine> read_cache(const char *path) {
    std::vector<CacheLine> out;
    if (!path) return out;
    std::ifstream f(path);
    std::string line;
    while (std::getline(f, line)) {
        std::istringstream is(line);
        CacheLine c;
        if (!(is >> c.arch >> c.method >> c.cls >> c.lv >> c.le >> c.opt)) continue;
        if ((c.cls >= 0 && c.cls < NUM_BUCKETS && c.opt >= 0 && c.opt < kTuneOptions) ||
            (c.cls == kStreamsLine && (c.opt == 1 || c.opt == 3)))
            out.push_back(c);
    }
    return out;
}

}  // namespace galp

using namespace galp;

extern "C" {

GALP_API int galp_graph_set_tuning(galp_graph *g, const int32_t *options) {
    if (!g || !options) return GALP_EINVAL;
    for (int b = 0; b < NUM_BUCKETS; ++b)
        if (options[b] < 0 || options[b] >= kTuneOptions) return fail(g->ctx, GALP_EINVAL, "tuning option out of range");
    for (int b = 0; b < NUM_BUCKETS; ++b) g->tune[b] = (int8_t)options[b];
    invalidate_captured(g);
    return GALP_OK;
}

GALP_API int galp_graph_autotune(galp_graph *g, int32_t reps, const char *cache_path, int32_t *chosen) {
    if (!g || reps < 1) return GALP_EINVAL;
    const bool hash = g->opts.method == GALP_METHOD_HASH;
    if ((hash && g->opts.variant != 1) || g->opts.policy != GALP_POLICY_INCORE || g->ooc)
        return fail(g->ctx, GALP_EUNSUPPORTED, "the autotuner tunes policy 0 (in-core), and method 1 with -l 1");
    {   // tune on labels[i] = i, the first iteration's (and heaviest) workload
        const int rc0 = galp_lp_init(g);
        if (rc0 != GALP_OK) return rc0;
    }
    const std::string arch = g->ctx->prop.gcnArchName;
    std::vector<CacheLine> cache = read_cache(cache_path);
    std::vector<CacheLine> fresh;
    hipEvent_t e0, e1;
    if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess)
        return fail(g->ctx, GALP_EHIP, "autotune events");
    int rc = GALP_OK;
    const bool prof = g->prof.on;
    const int32_t dispatch = g->opts.dispatch;
    g->prof.on = false;
    g->opts.dispatch = GALP_DISPATCH_PLAIN;   // time each class over all of its vertices
    for (int b = 0; b <= B_HUB && rc == GALP_OK; ++b) {
        const BucketRange &r = g->bucket[b];
        if (!r.nv) continue;
        const int lv = ilog2(r.nv), le = ilog2(r.edges);
        bool hit = false;
        for (const CacheLine &c : cache)
            if (c.arch == arch && c.method == g->opts.method && c.cls == b && c.lv == lv && c.le == le) {
                g->tune[b] = (int8_t)c.opt;
                hit = true;
            }
        if (hit) continue;
        float best = 1e30f;
        int bo = 0;
        for (int o = 0; o < kTuneOptions && rc == GALP_OK; ++o) {
            g->tune[b] = (int8_t)o;
            rc = hash ? launch_hash_class(g, b) : launch_sort_class(g, b);   // warm-up launch
            if (rc != GALP_OK) break;
            (void)hipEventRecord(e0, g->stream);
            for (int k = 0; k < reps && rc == GALP_OK; ++k)
                rc = hash ? launch_hash_class(g, b) : launch_sort_class(g, b);
            (void)hipEventRecord(e1, g->stream);
            if (hipEventSynchronize(e1) != hipSuccess) rc = fail(g->ctx, GALP_EHIP, "autotune timing");
            float ms = 0.f;
            if (rc == GALP_OK) (void)hipEventElapsedTime(&ms, e0, e1);
            if (rc == GALP_OK && ms < best) {
                best = ms;
                bo = o;
            }
        }
        g->tune[b] = (int8_t)bo;
        fresh.push_back(CacheLine{arch, g->opts.method, b, lv, le, bo});
    }
    if (rc == GALP_OK && concurrent_allowed(g)) {   // whole iterations: one stream or three
        const int lv = ilog2(g->n), le = ilog2(g->m);
        int mode = 0;
        for (const CacheLine &c : cache)
            if (c.arch == arch && c.method == g->opts.method && c.cls == kStreamsLine && c.lv == lv && c.le == le)
                mode = c.opt;
        if (!mode) {
            // the two choices timed alternately, best of three rounds each;
            // the size rule's choice stays unless the other is faster by
            // more than 3 % (on labels[i] = i the two are often within noise,
            // while later iterations favour the size rule)
            g->streams = 0;
            const int def = concurrent_classes(g) ? 3 : 1;
            float best[2] = {1e30f, 1e30f};
            for (int round = 0; round < 3 && rc == GALP_OK; ++round)
                for (int q = 0; q < 2 && rc == GALP_OK; ++q) {
                    g->streams = (int8_t)(q ? 3 : 1);
                    rc = one_iteration(g);   // warm-up
                    if (rc != GALP_OK) break;
                    (void)hipEventRecord(e0, g->stream);
                    for (int k = 0; k < reps && rc == GALP_OK; ++k) rc = one_iteration(g);
                    (void)hipEventRecord(e1, g->stream);
                    if (hipEventSynchronize(e1) != hipSuccess) rc = fail(g->ctx, GALP_EHIP, "autotune timing");
                    float ms = 0.f;
                    if (rc == GALP_OK) (void)hipEventElapsedTime(&ms, e0, e1);
                    best[q] = std::min(best[q], ms);
                }
            const float td = best[def == 3], to = best[def != 3];
            mode = to < 0.97f * td ? (def == 3 ? 1 : 3) : def;
            if (rc == GALP_OK) fresh.push_back(CacheLine{arch, g->opts.method, kStreamsLine, lv, le, mode});
        }
        g->streams = (int8_t)mode;
    }
    g->prof.on = prof;
    g->opts.dispatch = dispatch;
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    invalidate_captured(g);
    if (rc != GALP_OK) return rc;
    // the timed launches wrote next labels and the change counter: start over
    rc = galp_lp_init(g);
    if (rc != GALP_OK) return rc;
    if (cache_path && !fresh.empty()) {
        std::ofstream f(cache_path, std::ios::app);
        for (const CacheLine &c : fresh)
            f << c.arch << ' ' << c.method << ' ' << c.cls << ' ' << c.lv << ' ' << c.le << ' ' << c.opt << '\n';
    }
    if (chosen) {
        for (int b = 0; b < NUM_BUCKETS; ++b) chosen[b] = g->tune[b];
        chosen[kStreamsLine] = g->streams;
        for (int b = kStreamsLine + 1; b < 16; ++b) chosen[b] = 0;
    }
    return GALP_OK;
}

}  // extern "C"
