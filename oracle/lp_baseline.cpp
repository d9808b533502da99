/*
 * lp_baseline.cpp -- the CPU baseline: the oracle's synchronous LP update
 * restated with the reference CPU path's own data structures.
 *
 * TEST INFRASTRUCTURE ONLY (timed as bench.py's cpu_baseline and checked
 * against lp_oracle.c by the tests); the product never links it.
 *
 * Structure followed (reference paths relative to the reference checkout):
 *   - a std::map<label, count> histogram per vertex, the running count
 *     compared after every insert: label_propagation/include/
 *     label_propagator.h:52-66 and src/common/hybrid_cpu_worker.cuh:173-195;
 *   - the comparator `max_count <= c` and degree-0 behaviour of
 *     hybrid_cpu_worker.cuh:176-183;
 *   - OpenMP over the vertices with schedule(dynamic, 1), as the CPU worker
 *     of hybrid_cpu_worker.cuh:123-129 splits its batches.
 * Made synchronous (reads prev, writes next) like lp_oracle.c, so its labels
 * equal the oracle's and the GPU path's bit for bit.
 */
#include <stdint.h>

#include <map>

#ifdef _OPENMP
#include <omp.h>
#endif

extern "C" {

/* One synchronous iteration over the vertices vlist[0..nv) (all of [0, n)
 * when vlist is NULL, nv then being n): next[v] for each of them.
 * Returns the number whose label changed.  nthreads <= 0: all OpenMP
 * threads; 1: serial. */
int64_t baseline_map_iterate(int64_t n, const int64_t *off, const int32_t *nbr, const int32_t *prev,
                             int32_t *next, const int64_t *vlist, int64_t nv, int nthreads)
{
    (void)n;
    int64_t changed = 0;
#ifdef _OPENMP
    if (nthreads <= 0) nthreads = omp_get_max_threads();
#else
    nthreads = 1;
#endif
#pragma omp parallel for schedule(dynamic, 1) num_threads(nthreads) reduction(+ : changed)
    for (int64_t i = 0; i < nv; ++i) {
        const int64_t v = vlist ? vlist[i] : i;
        std::map<int32_t, int> label_count;
        int32_t max_label = prev[v];
        int max_count = 0;
        for (int64_t k = off[v]; k < off[v + 1]; ++k) {
            const int32_t label = prev[nbr[k]];
            const int c = ++label_count[label];
            if (max_count <= c) {
                max_count = c;
                max_label = label;
            }
        }
        next[v] = max_label;
        changed += (max_label != prev[v]);
    }
    return changed;
}

int baseline_max_threads(void)
{
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

}  // extern "C"
