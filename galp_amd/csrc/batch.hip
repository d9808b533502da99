// batch.hip -- many independent graphs in one run.
//
// No reference counterpart: the reference runs one graph per process.  The
// graphs are uploaded as one CSR, their disjoint union (graph k owns the
// vertices [vbase[k], vbase[k+1]), no edge crosses graphs), so one launch per
// degree class serves every graph and the 10-iteration loop can be captured
// once in a hipGraph.  What keeps each graph's labels equal to an
// independent run of it (shifted by vbase[k]) is its own stop rule: after
// every iteration an epilogue counts each graph's changes, freezes the graphs
// that stopped (their next labels are reset to prev), and raises the global
// stop flag when every graph has stopped.  The rule is the reference's
// (lfht.cuh:97, dpp.cuh:88) applied per graph: stop after an iteration with
// updated_k <= n_k * 1e-5.
// Under the hash method with plain dispatch, the graphs that stopped also
// leave the work: k_batch_lists builds, per degree class below the hubs,
// the list of the running graphs' vertices (lengths on the device) before
// the first iteration and again after each iteration in which a graph
// stopped, and every class kernel walks its list, so a frozen graph costs no
// gathers; the epilogue keeps its labels in both buffers.

#include <algorithm>
#include <vector>

#include "device_common.hpp"
#include "internal.hpp"

namespace galp {

// graph id of every vertex, in the internal numbering
__global__ void k_batch_gid(const int64_t *__restrict__ vbase, int64_t ngraphs, const int32_t *__restrict__ perm,
                            int64_t n, int32_t *__restrict__ gid) {
    for (int64_t u = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; u < n; u += (int64_t)gridDim.x * blockDim.x) {
        int64_t lo = 0, hi = ngraphs;   // last k with vbase[k] <= u
        while (hi - lo > 1) {
            const int64_t mid = (lo + hi) >> 1;
            if (vbase[mid] <= u) lo = mid;
            else hi = mid;
        }
        gid[perm ? perm[u] : u] = (int32_t)lo;
    }
}

// Flags an edge whose endpoints lie in different graphs.
template <typename E>
__global__ void k_batch_check(const E *__restrict__ off, const int32_t *__restrict__ nbr,
                              const int32_t *__restrict__ gid, int64_t n, int32_t *__restrict__ bad) {
    for (int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; v < n; v += (int64_t)gridDim.x * blockDim.x) {
        const int32_t k = gid[v];
        for (int64_t e = (int64_t)off[v]; e < (int64_t)off[v + 1]; ++e)
            if (gid[nbr[e]] != k) {
                atomicOr(bad, 1);
                return;
            }
    }
}

// The work lists: the members of classes G4..B4K (flat index over the
// classes: class b holds [f0[b], f0[b+1])) whose graph is still running, each
// appended to its class's list.  Appends are aggregated per block: a block
// takes kListChunk consecutive members at a time (kListItems per thread),
// ranks them per class in LDS, reserves each class's range with one device
// atomic and writes.  (One atomic per wave put ~6 million same-address
// atomics per iteration on a 400-million-vertex batch, serialised at the L2.)
struct ClassSpans {
    int64_t v0[NUM_BUCKETS];   // first id, or first index into vlist
    int64_t f0[NUM_BUCKETS + 1];
};

constexpr int kListItems = 16;
constexpr int64_t kListChunk = 256 * kListItems;

__global__ void __launch_bounds__(256)
k_batch_lists(ClassSpans cs, const int32_t *__restrict__ vlist, const int32_t *__restrict__ gid,
              const int32_t *__restrict__ gstop, int32_t *const *__restrict__ alist, int64_t *__restrict__ alen,
              const int32_t *__restrict__ stop) {
    __shared__ uint32_t scnt[NUM_BUCKETS];
    __shared__ int64_t sbase[NUM_BUCKETS];
    if (*stop || !alen[NUM_BUCKETS]) return;   // no graph stopped since the last rebuild
    const int64_t total = cs.f0[B_HUB];
    const unsigned long long lt = (1ull << lane_id()) - 1ull;
    for (int64_t c0 = (int64_t)blockIdx.x * kListChunk; c0 < total; c0 += (int64_t)gridDim.x * kListChunk) {
        if (threadIdx.x < NUM_BUCKETS) scnt[threadIdx.x] = 0u;
        __syncthreads();
        int32_t v[kListItems];
        int cls[kListItems];       // -1: not kept
        uint32_t rank[kListItems];   // within the block's share of the class
#pragma unroll
        for (int r = 0; r < kListItems; ++r) {
            const int64_t i = c0 + (int64_t)r * 256 + threadIdx.x;
            int b = 0;
            bool keep = false;
            v[r] = 0;
            if (i < total) {
                while (i >= cs.f0[b + 1]) ++b;
                const int64_t k = cs.v0[b] + (i - cs.f0[b]);
                v[r] = vlist ? vlist[k] : (int32_t)k;
                keep = gstop[gid[v[r]]] == 0;
            }
            cls[r] = keep ? b : -1;
            rank[r] = 0u;
            unsigned long long ball;
            while ((ball = __ballot(keep)) != 0ull) {
                const int leader = __ffsll((long long)ball) - 1;
                const int bl = __builtin_amdgcn_readlane(b, leader);
                const bool mine = keep && b == bl;
                const unsigned long long m = __ballot(mine);
                uint32_t at = 0u;
                if (lane_id() == leader) at = atomicAdd(&scnt[bl], (uint32_t)__popcll(m));
                at = (uint32_t)__shfl((int)at, leader);
                if (mine) rank[r] = at + (uint32_t)__popcll(m & lt);
                keep = keep && !mine;
            }
        }
        __syncthreads();
        if (threadIdx.x < NUM_BUCKETS && scnt[threadIdx.x])
            sbase[threadIdx.x] = (int64_t)atomicAdd((unsigned long long *)(alen + threadIdx.x),
                                                    (unsigned long long)scnt[threadIdx.x]);
        __syncthreads();
#pragma unroll
        for (int r = 0; r < kListItems; ++r)
            if (cls[r] >= 0) alist[cls[r]][sbase[cls[r]] + rank[r]] = v[r];
    }
}

// Per-vertex epilogue: frozen graphs keep prev; changes of the others are
// counted per graph (wave-aggregated: consecutive ids mostly share a graph).
__global__ void __launch_bounds__(256)
k_batch_epilogue(const int32_t *__restrict__ prev, int32_t *__restrict__ next, const int32_t *__restrict__ gid,
                 const int32_t *__restrict__ gstop, unsigned long long *__restrict__ gcount, int64_t n,
                 const int32_t *__restrict__ stop) {
    if (*stop) return;
    for (int64_t base = (int64_t)blockIdx.x * 256; base < n; base += (int64_t)gridDim.x * 256) {
        const int64_t v = base + threadIdx.x;
        bool pending = false;
        int32_t k = 0;
        if (v < n) {
            k = gid[v];
            const int32_t p = prev[v];
            if (gstop[k]) next[v] = p;
            else pending = next[v] != p;
        }
        unsigned long long ball;
        while ((ball = __ballot(pending)) != 0ull) {
            const int leader = __ffsll((long long)ball) - 1;
            const int32_t kl = __builtin_amdgcn_readlane(k, leader);
            const bool mine = pending && k == kl;
            const unsigned long long m = __ballot(mine);
            if (lane_id() == leader) atomicAdd(&gcount[kl], (unsigned long long)__popcll(m));
            pending = pending && !mine;
        }
    }
}

// Before a rebuild of the work lists: zero their lengths.  alen holds the
// NUM_BUCKETS lengths, then the rebuild flag (set when a graph stopped in
// the last iteration; the lists stand otherwise).
__global__ void k_batch_lists_begin(int64_t *__restrict__ alen, const int32_t *__restrict__ stop) {
    if (*stop || !alen[NUM_BUCKETS]) return;
    if (threadIdx.x < NUM_BUCKETS) alen[threadIdx.x] = 0;
}

// Per-graph stop rule and the run's bookkeeping (one block).  rebuild (may
// be null) is set to whether a graph stopped in this iteration.
__global__ void __launch_bounds__(1024)
k_batch_stop(unsigned long long *__restrict__ gcount, const int64_t *__restrict__ gn, int32_t *__restrict__ gstop,
             int32_t *__restrict__ giters, int64_t ngraphs, unsigned long long *__restrict__ counter,
             int32_t *__restrict__ stop, long long *__restrict__ iter_counts, int32_t *__restrict__ iters_done,
             int it, int64_t *__restrict__ rebuild) {
    __shared__ unsigned long long stotal;
    __shared__ int32_t sactive, sstopped;
    if (*stop) return;
    if (threadIdx.x == 0) {
        stotal = 0;
        sactive = 0;
        sstopped = 0;
    }
    __syncthreads();
    unsigned long long tot = 0;
    int active = 0, stopped = 0;
    for (int64_t k = threadIdx.x; k < ngraphs; k += 1024) {
        const unsigned long long c = gcount[k];
        gcount[k] = 0ull;
        if (gstop[k]) continue;
        tot += c;
        giters[k] = it + 1;
        if ((double)c <= (double)gn[k] * 1e-5) {
            gstop[k] = 1;
            stopped = 1;
        } else {
            active = 1;
        }
    }
    atomicAdd(&stotal, tot);
    if (active) sactive = 1;
    if (stopped) sstopped = 1;
    __syncthreads();
    if (threadIdx.x == 0) {
        iter_counts[it] = (long long)stotal;
        *iters_done = it + 1;
        *counter = 0ull;   // the class kernels' whole-graph count is not used in batched runs
        if (!sactive) *stop = 1;
        if (rebuild) *rebuild = sstopped;
    }
}

// rebuild (may be null): the work lists' rebuild flag, set for the first
// iteration.
__global__ void k_batch_reset(int32_t *__restrict__ gstop, int32_t *__restrict__ giters,
                              unsigned long long *__restrict__ gcount, int64_t ngraphs, int64_t *__restrict__ rebuild) {
    if (rebuild && blockIdx.x == 0 && threadIdx.x == 0) *rebuild = 1;
    for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < ngraphs; k += (int64_t)gridDim.x * blockDim.x) {
        gstop[k] = 0;
        giters[k] = 0;
        gcount[k] = 0ull;
    }
}

void batch_free(galp_graph *g) {
    BatchState *B = g->batch;
    if (!B) return;
    void *ptrs[] = {B->d_gid, B->d_gn, B->d_gcount, B->d_gstop, B->d_giters, B->d_vbase, B->d_alen, B->d_alist_dev};
    for (void *p : ptrs)
        if (p) (void)hipFree(p);
    for (int b = 0; b < NUM_BUCKETS; ++b)
        if (B->d_alist[b]) (void)hipFree(B->d_alist[b]);
    delete B;
    g->batch = nullptr;
}

int batch_reset(galp_graph *g) {
    BatchState *B = g->batch;
    if (!B) return GALP_OK;
    k_batch_reset<<<(int)std::min<int64_t>((B->ngraphs + 255) / 256, 1024), 256, 0, g->stream>>>(
        B->d_gstop, B->d_giters, B->d_gcount, B->ngraphs, B->lists ? B->d_alen + NUM_BUCKETS : nullptr);
    return hipGetLastError() == hipSuccess ? GALP_OK : fail(g->ctx, GALP_EHIP, "batch reset failed");
}

int launch_batch_lists(galp_graph *g) {
    BatchState *B = g->batch;
    if (!B || !B->lists) return GALP_OK;
    ClassSpans cs{};
    cs.f0[0] = 0;
    for (int b = 0; b < NUM_BUCKETS; ++b) {
        cs.v0[b] = g->bucket[b].v0;
        cs.f0[b + 1] = cs.f0[b] + (b < B_HUB ? g->bucket[b].nv : 0);
    }
    k_batch_lists_begin<<<1, 64, 0, g->stream>>>(B->d_alen, g->d_stop);
    const int64_t total = cs.f0[B_HUB];
    const int grid = (int)std::max<int64_t>(1, std::min<int64_t>((total + kListChunk - 1) / kListChunk,
                                                                 (int64_t)g->ctx->prop.multiProcessorCount * 8));
    k_batch_lists<<<grid, 256, 0, g->stream>>>(cs, g->d_vlist, B->d_gid, B->d_gstop, B->d_alist_dev, B->d_alen,
                                               g->d_stop);
    const hipError_t e = hipGetLastError();
    return e == hipSuccess ? GALP_OK : fail(g->ctx, GALP_EHIP, std::string("batch lists: ") + hipGetErrorString(e));
}

int launch_batch_end(galp_graph *g) {
    BatchState *B = g->batch;
    const int32_t *prev = g->d_lab[g->cur];
    int32_t *next = g->d_lab[g->cur ^ 1];
    const int grid = (int)std::max<int64_t>(1, std::min<int64_t>((g->n + 255) / 256,
                                                                 (int64_t)g->ctx->prop.multiProcessorCount * 8));
    k_batch_epilogue<<<grid, 256, 0, g->stream>>>(prev, next, B->d_gid, B->d_gstop, B->d_gcount, g->n, g->d_stop);
    k_batch_stop<<<1, 1024, 0, g->stream>>>(B->d_gcount, B->d_gn, B->d_gstop, B->d_giters, B->ngraphs, g->d_counter,
                                            g->d_stop, g->d_iter_counts, g->d_iters_done, g->iter % kMaxIters,
                                            B->lists ? B->d_alen + NUM_BUCKETS : nullptr);
    const hipError_t e = hipGetLastError();
    return e == hipSuccess ? GALP_OK : fail(g->ctx, GALP_EHIP, std::string("batch epilogue: ") + hipGetErrorString(e));
}

}  // namespace galp

using namespace galp;

extern "C" {

GALP_API int galp_graph_set_batch(galp_graph *g, int64_t ngraphs, const int64_t *vbase) {
    if (!g) return GALP_EINVAL;
    if (ngraphs < 1 || !vbase) return fail(g->ctx, GALP_EINVAL, "need ngraphs >= 1 and ngraphs + 1 vertex bases");
    if (vbase[0] != 0 || vbase[ngraphs] != g->n) return fail(g->ctx, GALP_EINVAL, "vbase must run from 0 to n");
    for (int64_t k = 0; k < ngraphs; ++k)
        if (vbase[k + 1] < vbase[k]) return fail(g->ctx, GALP_EINVAL, "vbase must be non-decreasing");
    if (g->ooc) return fail(g->ctx, GALP_EUNSUPPORTED, "batched runs are in-core");
    invalidate_captured(g);
    batch_free(g);
    BatchState *B = new (std::nothrow) BatchState();
    if (!B) return fail(g->ctx, GALP_ENOMEM, "host allocation failed");
    g->batch = B;
    B->ngraphs = ngraphs;
    std::vector<int64_t> gn((size_t)ngraphs);
    for (int64_t k = 0; k < ngraphs; ++k) {
        gn[(size_t)k] = vbase[k + 1] - vbase[k];
        B->max_n = std::max(B->max_n, gn[(size_t)k]);
    }
    const size_t n = (size_t)std::max<int64_t>(g->n, 1), ng = (size_t)ngraphs;
    if (hipMalloc((void **)&B->d_gid, 4 * n) != hipSuccess || hipMalloc((void **)&B->d_gn, 8 * ng) != hipSuccess ||
        hipMalloc((void **)&B->d_gcount, 8 * ng) != hipSuccess || hipMalloc((void **)&B->d_gstop, 4 * ng) != hipSuccess ||
        hipMalloc((void **)&B->d_giters, 4 * ng) != hipSuccess ||
        hipMalloc((void **)&B->d_vbase, 8 * (ng + 1)) != hipSuccess) {
        batch_free(g);
        return fail(g->ctx, GALP_ENOMEM, "batch state");
    }
    g->device_bytes += (int64_t)(4 * n + 24 * ng + 8);
    if (hipMemcpy(B->d_gn, gn.data(), 8 * ng, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(B->d_vbase, vbase, 8 * (ng + 1), hipMemcpyHostToDevice) != hipSuccess) {
        batch_free(g);
        return fail(g->ctx, GALP_EHIP, "batch upload failed");
    }
    const int grid = (int)std::max<int64_t>(1, std::min<int64_t>((g->n + 255) / 256, 4096));
    k_batch_gid<<<grid, 256, 0, g->stream>>>(B->d_vbase, ngraphs, g->d_perm, g->n, B->d_gid);
    int32_t *d_bad = nullptr;
    int32_t bad = 0;
    if (hipMalloc((void **)&d_bad, 4) != hipSuccess || hipMemcpy(d_bad, &bad, 4, hipMemcpyHostToDevice) != hipSuccess) {
        if (d_bad) (void)hipFree(d_bad);
        batch_free(g);
        return fail(g->ctx, GALP_ENOMEM, "batch check flag");
    }
    if (g->n) {
        if (g->off64)
            k_batch_check<int64_t><<<grid, 256, 0, g->stream>>>((const int64_t *)g->d_off, g->d_nbr, B->d_gid, g->n, d_bad);
        else
            k_batch_check<int32_t><<<grid, 256, 0, g->stream>>>((const int32_t *)g->d_off, g->d_nbr, B->d_gid, g->n, d_bad);
    }
    const bool ok = hipGetLastError() == hipSuccess && hipStreamSynchronize(g->stream) == hipSuccess &&
                    hipMemcpy(&bad, d_bad, 4, hipMemcpyDeviceToHost) == hipSuccess;
    (void)hipFree(d_bad);
    if (!ok) {
        batch_free(g);
        return fail(g->ctx, GALP_EHIP, "batch set-up kernels failed");
    }
    if (bad) {
        batch_free(g);
        return fail(g->ctx, GALP_EINVAL, "an edge joins two graphs of the batch");
    }
    if (g->opts.method == GALP_METHOD_HASH && g->opts.dispatch == GALP_DISPATCH_PLAIN) {
        // work lists of the running graphs' vertices, one per class below the hubs
        bool okl = hipMalloc((void **)&B->d_alen, 8 * (NUM_BUCKETS + 1)) == hipSuccess &&
                   hipMalloc((void **)&B->d_alist_dev, sizeof(int32_t *) * NUM_BUCKETS) == hipSuccess;
        for (int b = 0; b < B_HUB && okl; ++b)
            okl = hipMalloc((void **)&B->d_alist[b], 4 * (size_t)std::max<int64_t>(1, g->bucket[b].nv)) == hipSuccess;
        if (okl)
            okl = hipMemcpy(B->d_alist_dev, B->d_alist, sizeof(int32_t *) * NUM_BUCKETS, hipMemcpyHostToDevice) ==
                  hipSuccess;
        if (!okl) {
            (void)hipGetLastError();
            batch_free(g);
            return fail(g->ctx, GALP_ENOMEM, "batch work lists");
        }
        // the lane-group pipeline reads entry 0 of a list even when it is
        // empty (clamped addresses): every entry starts as vertex 0
        for (int b = 0; b < B_HUB; ++b) {
            const int rc = dev_fill_u32(g, (uint32_t *)B->d_alist[b], std::max<int64_t>(1, g->bucket[b].nv), 0u);
            if (rc != GALP_OK) {
                batch_free(g);
                return rc;
            }
        }
        int64_t members = 0;
        for (int b = 0; b < B_HUB; ++b) members += g->bucket[b].nv;
        g->device_bytes += 4 * members + 8 * (NUM_BUCKETS + 1);
        B->lists = true;
    }
    return batch_reset(g);
}

GALP_API int galp_batch_iterations(galp_graph *g, int32_t *iters_out) {
    if (!g || !iters_out) return GALP_EINVAL;
    if (!g->batch) return fail(g->ctx, GALP_ESTATE, "not a batched graph");
    if (hipMemcpy(iters_out, g->batch->d_giters, 4 * (size_t)g->batch->ngraphs, hipMemcpyDeviceToHost) != hipSuccess)
        return fail(g->ctx, GALP_EHIP, "readback failed");
    return GALP_OK;
}

}  // extern "C"
