/*
 * lp_oracle.c -- CPU oracle for galp's label-propagation (LP) update.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product (libgalp, the galp CLI,
 * the Python bindings) links, loads or calls this file.  Only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg use it, and only
 * as the parity reference / the timed CPU baseline, never as the thing being measured.
 *
 * What it restates (reference paths relative to the reference checkout):
 *   - loop structure of the serial CPU LP: label_propagation/include/
 *     label_propagator.h:37-82 (labels[i]=i at 44-47, per-vertex histogram
 *     at 54-58, change counting at 68-71, the `nupdates <= n*1e-5` break at
 *     73-75);
 *   - the deterministic per-vertex comparator `max_count <= c` and the
 *     degree-0 behaviour of label_propagation/src/common/
 *     hybrid_cpu_worker.cuh:173-195 (CPUWorker::update_label);
 *   - the counter definition of label_propagation/src/dpp/kernel.cuh:95-119
 *     (number of vertices whose label changed).
 * Deliberate differences from label_propagator.h, fixed for bit-exactness:
 *   - synchronous (Jacobi) update: every vertex reads `prev`, writes `next`
 *     (the serial path updates in place in a shuffled order, :50,68-71);
 *   - no coin flip on ties: the `<=` rule of hybrid_cpu_worker.cuh:183
 *     (the serial path flips rand() on ==, label_propagator.h:61-62).
 * The histogram is a dense counter array indexed by label instead of the
 * reference's std::map<V,int>; both yield the same running counts, which is
 * all the comparator reads.
 *
 * Pinning (tests/test_oracle.py): hand-derived per-iteration labels of
 * datasets/sample.txt (tests/golden/sample_labels.json), hand-built tie
 * micro-cases, the closed form argmax_L (count(L), last(L)) computed by an
 * independent brute force, renumbering equivariance, degree-0 invariance.
 */
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <math.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* One vertex: the scan loop of hybrid_cpu_worker.cuh:177-186.  `cnt` is a
 * zeroed dense array indexed by label; it is left zeroed on return. */
static int32_t update_one(int64_t v, const int64_t *off, const int32_t *nbr,
                          const int32_t *prev, int32_t *cnt)
{
    int32_t max_label = prev[v];          /* degree 0 keeps its label (:176) */
    int32_t max_count = 0;
    for (int64_t e = off[v]; e < off[v + 1]; ++e) {
        int32_t label = prev[nbr[e]];
        int32_t c = ++cnt[label];
        if (max_count <= c) {             /* hybrid_cpu_worker.cuh:183 */
            max_count = c;
            max_label = label;
        }
    }
    for (int64_t e = off[v]; e < off[v + 1]; ++e)
        cnt[prev[nbr[e]]] = 0;
    return max_label;
}

/* One synchronous LP iteration over vertices [0, n).  Labels must lie in
 * [0, nlabels).  Returns #{v : next[v] != prev[v]} (dpp/kernel.cuh:95-119).
 * nthreads <= 0 uses the OpenMP default; 1 runs serially. */
int64_t oracle_lp_iterate(int64_t n, const int64_t *off, const int32_t *nbr,
                          const int32_t *prev, int32_t *next, int64_t nlabels,
                          int nthreads)
{
    int64_t changed = 0;
#ifdef _OPENMP
    if (nthreads <= 0) nthreads = omp_get_max_threads();
#else
    nthreads = 1;
#endif
#pragma omp parallel num_threads(nthreads) reduction(+:changed)
    {
        int32_t *cnt = (int32_t *)calloc((size_t)nlabels, sizeof(int32_t));
#pragma omp for schedule(dynamic, 256)
        for (int64_t v = 0; v < n; ++v) {
            int32_t l = update_one(v, off, nbr, prev, cnt);
            next[v] = l;
            changed += (l != prev[v]);
        }
        free(cnt);
    }
    return changed;
}

/* The run loop of label_propagator.h:37-82 / lfht.cuh:73-112 made synchronous:
 * labels[i] = i, then up to `niter` iterations, stopping after an iteration
 * whose change count is <= n * 1e-5 (lfht.cuh:97, dpp.cuh:88).
 * per_iter (niter*n, may be NULL) receives the labels after each executed
 * iteration, counts (niter, may be NULL) the change counts.
 * Returns the number of iterations executed. */
int oracle_lp_run(int64_t n, const int64_t *off, const int32_t *nbr, int niter,
                  int32_t *labels_out, int32_t *per_iter, int64_t *counts,
                  int nthreads)
{
    int32_t *a = (int32_t *)malloc(sizeof(int32_t) * (size_t)(n > 0 ? n : 1));
    int32_t *b = (int32_t *)malloc(sizeof(int32_t) * (size_t)(n > 0 ? n : 1));
    for (int64_t i = 0; i < n; ++i) a[i] = (int32_t)i;
    int done = 0;
    for (int it = 0; it < niter; ++it) {
        int64_t c = oracle_lp_iterate(n, off, nbr, a, b, n, nthreads);
        int32_t *t = a; a = b; b = t;
        if (per_iter) memcpy(per_iter + (size_t)it * (size_t)n, a, sizeof(int32_t) * (size_t)n);
        if (counts) counts[it] = c;
        ++done;
        if ((double)c <= (double)n * 1e-5) break;
    }
    memcpy(labels_out, a, sizeof(int32_t) * (size_t)n);
    free(a);
    free(b);
    return done;
}

/* Number of distinct neighbour labels per vertex under `prev` (the quantity
 * the learned dispatcher predicts).  No reference counterpart; pinned by a
 * brute force over Python sets in the tests. */
void oracle_distinct_counts(int64_t n, const int64_t *off, const int32_t *nbr,
                            const int32_t *prev, int64_t nlabels, int32_t *out)
{
    uint8_t *seen = (uint8_t *)calloc((size_t)nlabels, 1);
    for (int64_t v = 0; v < n; ++v) {
        int32_t d = 0;
        for (int64_t e = off[v]; e < off[v + 1]; ++e) {
            int32_t l = prev[nbr[e]];
            if (!seen[l]) { seen[l] = 1; ++d; }
        }
        for (int64_t e = off[v]; e < off[v + 1]; ++e) seen[prev[nbr[e]]] = 0;
        out[v] = d;
    }
    free(seen);
}

/* Static per-vertex features of the learned dispatcher: degree, mean and max
 * neighbour degree.  mean = (float)((double)sum / (double)deg), 0 for deg 0;
 * the sum is an exact integer, so the rounding is fixed by this definition.
 * No reference counterpart (the learned dispatcher is new). */
void oracle_vertex_features(int64_t n, const int64_t *off, const int32_t *nbr,
                            float *deg_out, float *mean_out, float *max_out)
{
    for (int64_t v = 0; v < n; ++v) {
        int64_t d = off[v + 1] - off[v];
        int64_t sum = 0, mx = 0;
        for (int64_t e = off[v]; e < off[v + 1]; ++e) {
            int64_t u = nbr[e];
            int64_t du = off[u + 1] - off[u];
            sum += du;
            if (du > mx) mx = du;
        }
        deg_out[v] = (float)d;
        mean_out[v] = d ? (float)((double)sum / (double)d) : 0.0f;
        max_out[v] = (float)mx;
    }
}

/* Vertex-aligned out-of-core batches of fewer than B edges each, the cut rule
 * of label_propagation/src/common/outofcore.cuh:38-60 (binary search for the
 * first vertex whose offset reaches batch_offset, cut one vertex earlier).
 * The reference loops forever when a single vertex has degree >= B (the cut
 * stops advancing); this restatement cuts that vertex into a batch of its own
 * instead.  Writes at most max_out boundaries into bbs; returns the count
 * (#batches + 1), or -1 if max_out is too small. */
int64_t oracle_batch_boundaries(int64_t n, const int64_t *off, int64_t B,
                                int64_t *bbs, int64_t max_out)
{
    int64_t k = 0;
    int64_t m = off[n];
    if (k >= max_out) return -1;
    bbs[k++] = 0;
    int64_t batch_offset = B;
    int64_t last = 0;
    while (batch_offset < m) {
        int64_t left = 0, right = n + 1;
        while (left < right) {
            int64_t mid = (left + right) / 2;
            if (off[mid] < batch_offset) left = mid + 1; else right = mid;
        }
        int64_t cut = left - 1;
        if (cut <= last) cut = last + 1;      /* hub of degree >= B */
        if (cut >= n) break;
        if (k >= max_out) return -1;
        bbs[k++] = cut;
        last = cut;
        batch_offset = off[cut] + B;
    }
    if (k >= max_out) return -1;
    bbs[k++] = n;
    return k;
}

/* ------------------------------------------------------------------------
 * MLP 4 -> H -> H -> 1 (ReLU, ReLU, linear) in float64, the CPU restatement
 * of the learned dispatcher's network.  No reference counterpart.  Pinned by
 * central finite differences in the tests.
 * Parameter layout (row-major, all float64 here):
 *   W1[H][4], b1[H], W2[H][H], b2[H], W3[H], b3[1]
 * Loss: mean over the batch of 0.5 * (y_hat - y)^2.
 * ---------------------------------------------------------------------- */
static void mlp_offsets(int H, int64_t *o)
{
    o[0] = 0;                 /* W1 */
    o[1] = o[0] + (int64_t)H * 4;   /* b1 */
    o[2] = o[1] + H;          /* W2 */
    o[3] = o[2] + (int64_t)H * H;   /* b2 */
    o[4] = o[3] + H;          /* W3 */
    o[5] = o[4] + H;          /* b3 */
    o[6] = o[5] + 1;          /* total */
}

int64_t oracle_mlp_num_params(int H)
{
    int64_t o[7];
    mlp_offsets(H, o);
    return o[6];
}

/* Forward over a batch; writes predictions to yhat (may be NULL) and returns
 * the loss if y is non-NULL (else 0).  If grad is non-NULL it receives the
 * gradient of the loss with respect to every parameter. */
double oracle_mlp_loss_grad(int H, const double *p, int64_t B, const double *x,
                            const double *y, double *yhat, double *grad)
{
    int64_t o[7];
    mlp_offsets(H, o);
    const double *W1 = p + o[0], *b1 = p + o[1], *W2 = p + o[2];
    const double *b2 = p + o[3], *W3 = p + o[4], *b3 = p + o[5];
    if (grad) memset(grad, 0, sizeof(double) * (size_t)o[6]);
    double *h1 = (double *)malloc(sizeof(double) * (size_t)H);
    double *h2 = (double *)malloc(sizeof(double) * (size_t)H);
    double *d2 = (double *)malloc(sizeof(double) * (size_t)H);
    double *d1 = (double *)malloc(sizeof(double) * (size_t)H);
    double loss = 0.0;
    for (int64_t s = 0; s < B; ++s) {
        const double *xs = x + s * 4;
        for (int j = 0; j < H; ++j) {
            double a = b1[j];
            for (int k = 0; k < 4; ++k) a += W1[j * 4 + k] * xs[k];
            h1[j] = a > 0 ? a : 0;
        }
        for (int j = 0; j < H; ++j) {
            double a = b2[j];
            for (int k = 0; k < H; ++k) a += W2[(int64_t)j * H + k] * h1[k];
            h2[j] = a > 0 ? a : 0;
        }
        double out = b3[0];
        for (int k = 0; k < H; ++k) out += W3[k] * h2[k];
        if (yhat) yhat[s] = out;
        if (!y) continue;
        double r = out - y[s];
        loss += 0.5 * r * r / (double)B;
        if (!grad) continue;
        double g = r / (double)B;
        grad[o[5]] += g;
        for (int k = 0; k < H; ++k) {
            grad[o[4] + k] += g * h2[k];
            d2[k] = h2[k] > 0 ? g * W3[k] : 0.0;
        }
        for (int j = 0; j < H; ++j) {
            grad[o[3] + j] += d2[j];
            for (int k = 0; k < H; ++k) grad[o[2] + (int64_t)j * H + k] += d2[j] * h1[k];
        }
        for (int k = 0; k < H; ++k) {
            double a = 0.0;
            for (int j = 0; j < H; ++j) a += W2[(int64_t)j * H + k] * d2[j];
            d1[k] = h1[k] > 0 ? a : 0.0;
        }
        for (int j = 0; j < H; ++j) {
            grad[o[1] + j] += d1[j];
            for (int k = 0; k < 4; ++k) grad[o[0] + j * 4 + k] += d1[j] * xs[k];
        }
    }
    free(h1); free(h2); free(d2); free(d1);
    return loss;
}
