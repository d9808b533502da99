// autotune.hip -- picks each degree class's kernel configuration by timing.
//
// No reference counterpart (the reference fixes its block sizes at compile
// time, lfht.cuh:127-165).  Each round sets every untuned class to option 0,
// then 1, then 2, and times a real run of kTuneIters iterations from
// labels[i] = i under the graph's own dispatch (learned when it has a
// network), reading every class's device time from per-kernel events; each
// class keeps its fastest option over the rounds.  Timing whole runs, not
// single launches on the first iteration's labels, weighs the converged
// iterations, where rows change path and most of a run's time goes.  The
// configurations are the method's: launch_class in lp_hash.hip (workgroup
// size x slots per lane and blocks per CU; for hubs the tile and merge
// workgroup sizes) and sort_class in lp_sort.hip (workgroup size x keys per
// thread, the sort algorithm, grids).  The choice never changes labels (every
// configuration computes the same rule), which is how it is tested.  Last,
// where an iteration may run its class kernels on three streams, whole runs
// are timed on one stream and on three; the choice by size stays unless the
// other is more than 3 % faster (the overlap helps while the label array
// stays in the Infinity Cache and can hurt once it does not).  Choices can
// be cached in a text file, one line per class (class 12 = the stream
// choice, 1 or 3):
//   <arch> <method> <class> <log2 vertices> <log2 edges> <option>
// and are reused for a graph whose class has the same arch, method and sizes.
// Under learned dispatch, in core, the tuner then judges the network against
// the plain dispatcher iteration by iteration (the per-iteration mask,
// galp_graph.learn_mask), which is not cached: it depends on the network.

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

constexpr int kTuneIters = 10;   // iterations of each timed run

// Per-class device time (ms) of a run of `niter` iterations from labels[i]
// = i under the current options, from per-kernel events (which also keep the
// run on one stream).  The caller's profile totals are kept aside and put
// back.
static int run_class_times(galp_graph *g, int niter, double *t) {
    const std::vector<KernelAgg> saved = g->prof.aggs;
    for (KernelAgg &a : g->prof.aggs) {
        a.launches = 0;
        a.total_ms = 0.0;
        a.alg_bytes = 0.0;
    }
    const bool on = g->prof.on;
    g->prof.on = true;
    galp_run_result res{};
    const int rc = galp_lp_run(g, niter, 0, nullptr, &res);
    g->prof.on = on;
    for (auto &pl : g->prof.pending) {   // a failed run may leave some
        if (pl.e0) g->prof.pool.push_back(pl.e0);
        if (pl.e1) g->prof.pool.push_back(pl.e1);
    }
    g->prof.pending.clear();
    for (int b = 0; b < NUM_BUCKETS; ++b) t[b] = 0.0;
    for (const KernelAgg &a : g->prof.aggs) {
        int b = -1;
        for (int c = 0; c < NUM_BUCKETS; ++c)
            if (a.name == kBucketName[c]) b = c;
        if (a.name == "hub_tile" || a.name == "hub_merge" || a.name == "hub_final" || a.name == "hub_sort") b = B_HUB;
        if (b >= 0) t[b] += a.total_ms;
    }
    g->prof.aggs = saved;   // no launch is pending after the run
    return rc;
}

// Every option of every class marked in `need`, timed together in whole
// runs, `reps` rounds; each class keeps its fastest option.
static int tune_class_options(galp_graph *g, int reps, const bool *need) {
    double best[NUM_BUCKETS];
    int8_t bo[NUM_BUCKETS];
    for (int b = 0; b < NUM_BUCKETS; ++b) {
        best[b] = 1e300;
        bo[b] = g->tune[b];
    }
    int rc = GALP_OK;
    for (int round = 0; round < reps && rc == GALP_OK; ++round)
        for (int o = 0; o < kTuneOptions && rc == GALP_OK; ++o) {
            for (int b = 0; b < NUM_BUCKETS; ++b)
                if (need[b]) g->tune[b] = (int8_t)o;
            double t[NUM_BUCKETS];
            rc = run_class_times(g, kTuneIters, t);
            for (int b = 0; b < NUM_BUCKETS && rc == GALP_OK; ++b)
                if (need[b] && t[b] < best[b]) {
                    best[b] = t[b];
                    bo[b] = (int8_t)o;
                }
        }
    for (int b = 0; b < NUM_BUCKETS; ++b)
        if (need[b]) g->tune[b] = bo[b];
    return rc;
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

GALP_API int galp_graph_set_learned_iterations(galp_graph *g, uint64_t mask) {
    if (!g) return GALP_EINVAL;
    if (g->opts.dispatch != GALP_DISPATCH_LEARNED) return fail(g->ctx, GALP_ESTATE, "not a learned-dispatch graph");
    g->learn_mask = mask;
    invalidate_captured(g);
    return GALP_OK;
}

GALP_API int galp_graph_autotune(galp_graph *g, int32_t reps, const char *cache_path, int32_t *chosen) {
    if (!g || reps < 1) return GALP_EINVAL;
    const bool hash = g->opts.method == GALP_METHOD_HASH;
    if ((hash && g->opts.variant != 1) || g->opts.policy != GALP_POLICY_INCORE || g->ooc)
        return fail(g->ctx, GALP_EUNSUPPORTED, "the autotuner tunes policy 0 (in-core), and method 1 with -l 1");
    const std::string arch = g->ctx->prop.gcnArchName;
    std::vector<CacheLine> cache = read_cache(cache_path);
    std::vector<CacheLine> fresh;
    // the runs below use the graph's own dispatch (learned needs its network)
    const int32_t dispatch = g->opts.dispatch;
    if (dispatch == GALP_DISPATCH_LEARNED && !(g->learn && g->learn->H)) g->opts.dispatch = GALP_DISPATCH_PLAIN;
    const int32_t capture = g->opts.graph_capture;
    g->opts.graph_capture = 0;
    invalidate_captured(g);
    int rc = GALP_OK;
    bool need[NUM_BUCKETS] = {};
    bool any = false;
    for (int b = 0; b <= B_HUB; ++b) {
        const BucketRange &r = g->bucket[b];
        if (!r.nv) continue;
        const int lv = ilog2(r.nv), le = ilog2(r.edges);
        bool hit = false;
        for (const CacheLine &c : cache)
            if (c.arch == arch && c.method == g->opts.method && c.cls == b && c.lv == lv && c.le == le) {
                g->tune[b] = (int8_t)c.opt;
                hit = true;
            }
        need[b] = !hit;
        any = any || need[b];
    }
    if (any) rc = tune_class_options(g, reps, need);
    if (rc == GALP_OK && concurrent_allowed(g)) {   // whole runs: one stream or three
        const int lv = ilog2(g->n), le = ilog2(g->m);
        int mode = 0;
        for (const CacheLine &c : cache)
            if (c.arch == arch && c.method == g->opts.method && c.cls == kStreamsLine && c.lv == lv && c.le == le)
                mode = c.opt;
        if (!mode) {
            // the two choices alternately, best of `reps` rounds each; the
            // choice by size stays unless the other is faster by more than 3 %
            g->streams = 0;
            const int def = concurrent_classes(g) ? 3 : 1;
            double tbest[2] = {1e300, 1e300};
            for (int round = 0; round < reps && rc == GALP_OK; ++round)
                for (int q = 0; q < 2 && rc == GALP_OK; ++q) {
                    g->streams = (int8_t)(q ? 3 : 1);
                    galp_run_result res{};
                    rc = galp_lp_run(g, kTuneIters, 0, nullptr, &res);
                    if (rc == GALP_OK) tbest[q] = std::min(tbest[q], res.device_ms);
                }
            const double td = tbest[def == 3], to = tbest[def != 3];
            mode = to < 0.97 * td ? (def == 3 ? 1 : 3) : def;
            if (rc == GALP_OK) fresh.push_back(CacheLine{arch, g->opts.method, kStreamsLine, lv, le, mode});
        }
        g->streams = (int8_t)mode;
    }
    if (rc == GALP_OK && g->opts.dispatch == GALP_DISPATCH_LEARNED && !g->ooc) {
        // the judge: iteration by iteration, routing by the network against
        // routing by the degree classes alone (the plain dispatcher), on
        // whole iterations (device time, all streams) of the same run -- the
        // labels, hence every iteration's work, are the same either way.  An
        // iteration keeps the network where it is faster by more than 2 %
        // (below that, repeated timings of one iteration disagree: on LFR-8M
        // a mask of near-ties ran 2.5 % slower than plain); iterations past
        // the timed ones follow the last of them.
        double t[2][kTuneIters];
        for (int q = 0; q < 2; ++q)
            for (int i = 0; i < kTuneIters; ++i) t[q][i] = 1e300;
        int nit = kTuneIters;
        for (int round = 0; round < reps && rc == GALP_OK; ++round)
            for (int q = 0; q < 2 && rc == GALP_OK; ++q) {
                g->learn_mask = q ? ~0ull : 0ull;   // q = 1: the network
                rc = galp_lp_init(g);
                for (int i = 0; i < kTuneIters && rc == GALP_OK; ++i) {
                    galp_step_stats st{};
                    rc = galp_lp_step(g, &st);
                    if (rc == GALP_OK) t[q][i] = std::min(t[q][i], st.device_ms);
                    if (rc == GALP_OK && st.stopped) {
                        nit = std::min(nit, i + 1);
                        break;
                    }
                }
            }
        uint64_t mask = 0;
        for (int i = 0; i < nit; ++i)
            if (t[1][i] < 0.98 * t[0][i]) mask |= 1ull << i;
        if (nit > 0 && ((mask >> (nit - 1)) & 1ull)) mask |= ~0ull << nit;
        g->learn_mask = rc == GALP_OK ? mask : ~0ull;
        // the class options were timed on the rows the network left them;
        // iterations now routed by the classes alone give them every row
        // of their degree range: time the options again under the mask
        if (rc == GALP_OK && any && mask != ~0ull) rc = tune_class_options(g, reps, need);
    }
    for (int b = 0; b < NUM_BUCKETS && rc == GALP_OK; ++b)
        if (need[b]) {
            const BucketRange &r = g->bucket[b];
            fresh.push_back(CacheLine{arch, g->opts.method, b, ilog2(r.nv), ilog2(r.edges), g->tune[b]});
        }
    g->opts.dispatch = dispatch;
    g->opts.graph_capture = capture;
    invalidate_captured(g);
    if (rc != GALP_OK) return rc;
    // the timed runs moved the labels on: start over
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
        if (g->opts.dispatch == GALP_DISPATCH_LEARNED) {   // the judge's mask
            chosen[kStreamsLine + 1] = (int32_t)(uint32_t)g->learn_mask;
            chosen[kStreamsLine + 2] = (int32_t)(uint32_t)(g->learn_mask >> 32);
        }
    }
    return GALP_OK;
}

}  // extern "C"
