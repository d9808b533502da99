/*
 * galp_host.h -- host-side graph input for the galp LP path (no device calls).
 *
 * These functions build the CSR that galp_graph_upload() (galp.h) consumes.
 * They replace the reference's CSRGraph constructor and loaders
 * (label_propagation/include/graph.h:58-188) and add the seeded synthetic
 * generators the benchmarks run on.  All functions are thread-safe, never
 * throw, return 0 on success or a negative galp_status, and leave a message
 * for galp_host_last_error().
 */
#ifndef GALP_HOST_H
#define GALP_HOST_H

#include <stdint.h>
#include <stddef.h>
#include "galp_status.h"

#ifdef __cplusplus
extern "C" {
#endif

#ifndef GALP_API
#define GALP_API __attribute__((visibility("default")))
#endif

/* A CSR graph owned by the library.  offsets has n+1 entries, neighbors m.
 * Row v is neighbors[offsets[v] .. offsets[v+1]), in the order that fixes the
 * tie rule (file order for loaded graphs, ascending for generated ones).
 * Replaces CSRGraph<V,E>::{n, m, neighbors, offsets} (graph.h:41-42,160-161). */
typedef struct galp_host_csr {
    int64_t n;
    int64_t m;
    int64_t *offsets;
    int32_t *neighbors;
} galp_host_csr;

/* Loads ".txt"/".dat" (one "u v" edge per line, u < v, symmetrised as the
 * text loader of graph.h:166-188 does: row u gets v, row v gets u, in line
 * order) or ".bin" (int32 pairs holding both directions, graph.h:119-150).
 * Replaces CSRGraph(const std::string &filename) (graph.h:65-112).
 * Differences (each deliberate): vertex ids are kept, so an id with no edges
 * gets an empty row and n = max id + 1 (the reference's .bin loader opens a
 * new row per source run, shifting every later vertex after a gap); text
 * lines that do not start with two integers (SNAP '#' comments) are skipped;
 * a .bin whose size is not a multiple of 8 bytes is rejected (GALP_EIO). */
GALP_API int galp_load_graph(const char *path, galp_host_csr **out);

/* Writes g as a ".bin" file: int32 (u, v) pairs, rows in order, both
 * directions present -- the format graph.h:119-150 reads and
 * utils/edgelist2bin.cpp:68-96 (the `_simple` variant) writes. */
GALP_API int galp_save_bin(const char *path, const galp_host_csr *g);

/* Planted partition graph.  Vertex v belongs to community v / community_size
 * (the last community may be smaller).  Each vertex draws k_intra partners
 * uniformly from its own community without itself and k_inter partners
 * uniformly from the other communities; both directions are stored,
 * duplicates removed, rows sorted ascending.  Deterministic in seed
 * (SplitMix64 streams, Steele/Lea/Flood 2014).  No reference counterpart:
 * the reference ships no generator; this makes the synthetic inputs. */
GALP_API int galp_gen_planted_partition(int64_t n, int64_t community_size,
                                        int32_t k_intra, int32_t k_inter,
                                        uint64_t seed, galp_host_csr **out);

/* A batch of `count` independent planted-partition graphs as one CSR, their
 * disjoint union: graph k owns vertices [vbase[k], vbase[k+1]).  Graph k's
 * parameters come from the SplitMix64 stream (seed, k): n ~ U[n_lo, n_hi],
 * community size ~ U[comm_lo, comm_hi], degree target d ~ U[deg_lo, deg_hi];
 * it is galp_gen_planted_partition(n, community, (d + 1) / 2,
 * max(1, round(inter_frac * d / 2)), seed_k), so intra degree ~ d and inter
 * degree ~ inter_frac * d after symmetrisation.  vbase holds count + 1
 * entries; spec (optional) receives count rows of (n, community, k_intra,
 * k_inter, seed_k).  Graphs are built in parallel, one per host thread.
 * No reference counterpart (the reference runs one graph per process). */
GALP_API int galp_gen_batch_planted(int64_t count, int64_t n_lo, int64_t n_hi,
                                    int64_t comm_lo, int64_t comm_hi,
                                    int32_t deg_lo, int32_t deg_hi,
                                    double inter_frac, uint64_t seed,
                                    galp_host_csr **out, int64_t *vbase,
                                    int64_t *spec);

/* R-MAT / Kronecker graph (Chakrabarti et al. 2004; Graph500 parameters
 * without noise): edge_factor * 2^scale samples with quadrant probabilities
 * a, b, c, d = 1-a-b-c; self-loops dropped, symmetrised, deduplicated, vertex
 * ids permuted by a seeded Fisher-Yates shuffle, rows sorted ascending.
 * No reference counterpart (synthetic input). */
GALP_API int galp_gen_rmat(int32_t scale, int32_t edge_factor, double a,
                           double b, double c, uint64_t seed,
                           galp_host_csr **out);

/* LFR-style benchmark graph (Lancichinetti, Fortunato, Radicchi 2008):
 * power-law degrees (exponent gamma, mean avg_deg, cap max_deg), power-law
 * community sizes (exponent beta, in [cmin, cmax]), mixing mu; intra and
 * inter stubs matched by a seeded configuration model, loops and multi-edges
 * dropped, rows sorted ascending.  If community_out is non-NULL it receives
 * an n-entry int32 array (free with galp_host_free) of ground-truth
 * community ids.  No reference counterpart (synthetic input). */
GALP_API int galp_gen_lfr(int64_t n, double gamma, double avg_deg,
                          int32_t max_deg, double beta, int32_t cmin,
                          int32_t cmax, double mu, uint64_t seed,
                          galp_host_csr **out, int32_t **community_out);

/* Takes ownership of nothing: copies the arrays into a library-owned CSR. */
GALP_API int galp_host_csr_from_arrays(int64_t n, int64_t m,
                                       const int64_t *offsets,
                                       const int32_t *neighbors,
                                       galp_host_csr **out);

/* Clustering quality of labels[n] against a ground-truth file holding one
 * cluster per line (whitespace-separated vertex ids): normalised mutual
 * information, F-measure and adjusted Rand index, with the formulas of
 * compute_nmi / compute_f_measure / compute_ari
 * (label_propagation/include/accuracy.h:49-319), which main.cu:223-238 calls
 * for the optional `test` argument.  Like compute_nmi (accuracy.h:156) it
 * prints "<#truth clusters>, <#found clusters>" to stdout. */
GALP_API int galp_accuracy(const char *truth_file, const int32_t *labels,
                           int64_t n, double *nmi, double *fmeasure,
                           double *ari);

/* The serial CPU path: labels[i] = i, then up to niter synchronous
 * iterations with the n*1e-5 stop rule, on one host thread; the final labels
 * go to labels_out[n], the wall time to *seconds and the iterations run to
 * *iterations (both may be NULL).  Replaces LabelPropagator::run
 * (label_propagation/include/label_propagator.h:37-82), which main.cu:200-206
 * selects for method*5+policy outside 0..10.  Same update and tie rule as the
 * GPU methods (galp.h), so the labels equal theirs; the reference's serial
 * path instead updates in place in a shuffled order and flips a coin on ties
 * (label_propagator.h:50,61-62). */
GALP_API int galp_cpu_lp_run(const galp_host_csr *g, int32_t niter,
                             int32_t *labels_out, double *seconds,
                             int32_t *iterations);

GALP_API void galp_host_csr_free(galp_host_csr *g);
GALP_API void galp_host_free(void *p);

/* Message of the last failing galp_* call on this thread ("" if none). */
GALP_API const char *galp_host_last_error(void);

#ifdef __cplusplus
}
#endif

#endif /* GALP_HOST_H */
