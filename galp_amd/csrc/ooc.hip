// ooc.hip -- out-of-core policies: the neighbour array stays in pinned host
// memory and streams to the GPU in vertex-aligned batches every iteration.
//
// Replaces the reference's out-of-core propagators:
//   OutOfCore::compute_batch_boundaries  label_propagation/src/common/outofcore.cuh:38-60
//   OutOfCore::transfer_batch / transfer_next_batch   outofcore.cuh:63-97
//   SyncLP::iterate                     src/common/sync.cuh:43-59   (policy 1)
//   AsyncLP::iterate, swap_buffers      src/common/async.cuh:66-107 (policy 2)
// Design: offsets and both label arrays live in HBM; only the neighbour ids
// move.  A batch is [bv0, bv1) with fewer than 2^b edges (one vertex of
// degree >= 2^b gets a batch of its own: the reference's cut loop would not
// advance there).  For each batch the GPU copies the slice into one of two
// device buffers, rebases the batch's offsets into a per-buffer offset array
// (off[v] - off[bv0]), and runs the method's kernels on the batch's share of
// each degree class and of the hub lists.  Every batch reads labels(t-1)
// and writes labels(t), so the labels equal the in-core ones for any 2^b.
//   policy 1: copy and kernels in order on the graph's stream;
//   policy 2: copies on a second stream, double-buffered: the copy of batch
//             b+1 overlaps the kernels of batch b; events order each buffer's
//             reuse after the kernels that read it.

#include <algorithm>
#include <cstring>
#include <vector>

#include "device_common.hpp"
#include "internal.hpp"
#include "parallel.hpp"

namespace galp {

template <typename E>
__global__ void __launch_bounds__(256)
k_rebase(const E *__restrict__ off, E *__restrict__ offb, int64_t v0, int64_t v1) {
    const E base = off[v0];
    for (int64_t v = v0 + (int64_t)blockIdx.x * 256 + threadIdx.x; v <= v1; v += (int64_t)gridDim.x * 256)
        offb[v] = off[v] - base;
}

// The cut rule of outofcore.cuh:38-60 with the degree >= B case cut off
// into a batch of its own (restated by the oracle's oracle_batch_boundaries).
static std::vector<int64_t> batch_boundaries(const std::vector<int64_t> &off, int64_t B) {
    const int64_t n = (int64_t)off.size() - 1, m = off[(size_t)n];
    std::vector<int64_t> bbs{0};
    int64_t batch_offset = B, last = 0;
    while (batch_offset < m) {
        const int64_t left = std::lower_bound(off.begin(), off.end(), batch_offset) - off.begin();
        int64_t cut = left - 1;
        if (cut <= last) cut = last + 1;
        if (cut >= n) break;
        bbs.push_back(cut);
        last = cut;
        batch_offset = off[(size_t)cut] + B;
    }
    bbs.push_back(n);
    return bbs;
}

void ooc_free(galp_graph *g) {
    OocState *O = g->ooc;
    if (!O) return;
    for (int k = 0; k < 2; ++k) {
        if (O->d_buf[k]) (void)hipFree(O->d_buf[k]);
        if (O->d_offb[k]) (void)hipFree(O->d_offb[k]);
        if (O->ev_copied[k]) (void)hipEventDestroy(O->ev_copied[k]);
        if (O->ev_free[k]) (void)hipEventDestroy(O->ev_free[k]);
    }
    if (O->copy_stream) (void)hipStreamDestroy(O->copy_stream);
    if (O->h_nbr) (void)hipHostFree(O->h_nbr);
    delete O;
    g->ooc = nullptr;
}

int ooc_setup(galp_graph *g, const galp_csr_view *c, const std::vector<int64_t> &offi,
              const std::vector<int32_t> &perm, const std::vector<int32_t> &iperm,
              const std::vector<int32_t> &vlist, const std::vector<int32_t> &hub_vertices,
              const std::vector<int32_t> &hub_tile0, const std::vector<int64_t> &hub_item0) {
    OocState *O = new (std::nothrow) OocState();
    if (!O) return fail(g->ctx, GALP_ENOMEM, "host allocation failed");
    g->ooc = O;
    const int64_t n = g->n, m = g->m;
    // the neighbour array, in the internal numbering, in pinned host memory
    if (hipHostMalloc((void **)&O->h_nbr, 4 * (size_t)std::max<int64_t>(m, 1), hipHostMallocDefault) != hipSuccess) {
        O->h_nbr = nullptr;
        return fail(g->ctx, GALP_ENOMEM, "pinned host buffer for the neighbour array");
    }
    const int32_t *src = c->neighbors;
    if (g->reorder) {
        const bool o64 = c->offsets_are_64 != 0;
        parallel_for(0, n, [&](int64_t k) {
            const int64_t v = iperm[(size_t)k];
            const int64_t b = o64 ? ((const int64_t *)c->offsets)[v] : ((const int32_t *)c->offsets)[v];
            const int64_t d = offi[(size_t)k + 1] - offi[(size_t)k];
            int32_t *dst = O->h_nbr + offi[(size_t)k];
            for (int64_t j = 0; j < d; ++j) dst[j] = perm[(size_t)src[b + j]];
        });
    } else {
        const int64_t chunk = 1 << 22;
        parallel_for(0, (m + chunk - 1) / chunk, [&](int64_t k) {
            const int64_t lo = k * chunk, hi = std::min(m, lo + chunk);
            std::memcpy(O->h_nbr + lo, src + lo, 4 * (size_t)(hi - lo));
        });
    }
    // batches and their share of every class and of the hub lists
    const int64_t B = (int64_t)1 << g->opts.buffer_pow;
    O->bbs = batch_boundaries(offi, B);
    const int64_t nb = (int64_t)O->bbs.size() - 1;
    int64_t maxe = 0;
    O->views.resize((size_t)std::max<int64_t>(nb, 0));
    O->e0.resize((size_t)std::max<int64_t>(nb, 0));
    O->e1.resize((size_t)std::max<int64_t>(nb, 0));
    auto cut_list = [&](int64_t lo, int64_t hi, int64_t s0, int64_t cnt, int64_t &k0, int64_t &k1) {
        if (g->reorder) {   // class = ids [s0, s0 + cnt)
            k0 = std::min(std::max(lo, s0), s0 + cnt);
            k1 = std::min(std::max(hi, s0), s0 + cnt);
        } else {            // class = vlist[s0, s0 + cnt), sorted by id
            auto b = vlist.begin() + s0, e = vlist.begin() + s0 + cnt;
            k0 = s0 + (std::lower_bound(b, e, (int32_t)lo) - b);
            k1 = s0 + (std::lower_bound(b, e, (int32_t)std::min<int64_t>(hi, INT32_MAX)) - b);
        }
    };
    for (int64_t b = 0; b < nb; ++b) {
        const int64_t lo = O->bbs[(size_t)b], hi = O->bbs[(size_t)b + 1];
        O->e0[(size_t)b] = offi[(size_t)lo];
        O->e1[(size_t)b] = offi[(size_t)hi];
        maxe = std::max(maxe, O->e1[(size_t)b] - O->e0[(size_t)b]);
        IterView &w = O->views[(size_t)b];
        for (int cl = 0; cl < NUM_BUCKETS; ++cl) {
            const BucketRange &r = g->bucket[cl];
            int64_t k0, k1;
            cut_list(lo, hi, r.v0, r.nv, k0, k1);
            w.cls[cl].v0 = k0;
            w.cls[cl].nv = k1 - k0;
            w.cls[cl].edges = 0;
        }
        // hubs: ascending ids in both numberings
        w.h0 = std::lower_bound(hub_vertices.begin(), hub_vertices.end(), (int32_t)lo) - hub_vertices.begin();
        w.h1 = std::lower_bound(hub_vertices.begin(), hub_vertices.end(), (int32_t)std::min<int64_t>(hi, INT32_MAX)) -
               hub_vertices.begin();
        const int64_t nh = (int64_t)hub_vertices.size();
        w.t0 = w.h0 < nh ? hub_tile0[(size_t)w.h0] : g->nhub_tiles;
        w.t1 = w.h1 < nh ? hub_tile0[(size_t)w.h1] : g->nhub_tiles;
        w.i0 = hub_item0[(size_t)w.h0];
        w.i1 = hub_item0[(size_t)w.h1];
    }
    O->bufcap = std::max<int64_t>(maxe, 1);
    const size_t esz = g->off64 ? 8 : 4;
    for (int k = 0; k < 2; ++k) {
        if (hipMalloc((void **)&O->d_buf[k], 4 * (size_t)O->bufcap) != hipSuccess ||
            hipMalloc(&O->d_offb[k], esz * (size_t)(n + 1)) != hipSuccess)
            return fail(g->ctx, GALP_ENOMEM, "out-of-core batch buffers");
        g->device_bytes += (int64_t)(4 * (size_t)O->bufcap + esz * (size_t)(n + 1));
        if (hipEventCreateWithFlags(&O->ev_copied[k], hipEventDisableTiming) != hipSuccess ||
            hipEventCreateWithFlags(&O->ev_free[k], hipEventDisableTiming) != hipSuccess)
            return fail(g->ctx, GALP_EHIP, "out-of-core events");
    }
    if (g->opts.policy == GALP_POLICY_OOC_ASYNC &&
        hipStreamCreateWithFlags(&O->copy_stream, hipStreamNonBlocking) != hipSuccess)
        return fail(g->ctx, GALP_EHIP, "out-of-core copy stream");
    return GALP_OK;
}

template <typename E>
static int ooc_round(galp_graph *g) {
    OocState *O = g->ooc;
    const int64_t nb = (int64_t)O->views.size();
    const bool async = g->opts.policy == GALP_POLICY_OOC_ASYNC;
    for (int64_t b = 0; b < nb; ++b) {
        const int k = async ? (int)(b & 1) : 0;
        const int64_t e0 = O->e0[(size_t)b], ne = O->e1[(size_t)b] - e0;
        const int64_t lo = O->bbs[(size_t)b], hi = O->bbs[(size_t)b + 1];
        hipStream_t cs = async ? O->copy_stream : g->stream;
        if (async && O->used[k]) {   // the kernels that read this buffer are done
            if (hipStreamWaitEvent(cs, O->ev_free[k], 0) != hipSuccess)
                return fail(g->ctx, GALP_EHIP, "out-of-core wait");
        }
        PendingLaunch pl;
        if (!async) prof_begin(g, "h2d", 4.0 * (double)ne, &pl);
        if (ne && hipMemcpyAsync(O->d_buf[k], O->h_nbr + e0, 4 * (size_t)ne, hipMemcpyHostToDevice, cs) != hipSuccess)
            return fail(g->ctx, GALP_EHIP, "out-of-core batch copy");
        if (!async) prof_end(g, &pl);
        if (async) {
            if (hipEventRecord(O->ev_copied[k], cs) != hipSuccess ||
                hipStreamWaitEvent(g->stream, O->ev_copied[k], 0) != hipSuccess)
                return fail(g->ctx, GALP_EHIP, "out-of-core copy event");
        }
        const int64_t nv = hi - lo + 1;
        const int grid = (int)std::max<int64_t>(1, std::min<int64_t>((nv + 255) / 256, 1024));
        k_rebase<E><<<grid, 256, 0, g->stream>>>((const E *)g->d_off, (E *)O->d_offb[k], lo, hi);
        IterView w = O->views[(size_t)b];
        w.off = O->d_offb[k];
        w.nbr = O->d_buf[k];
        if (g->opts.method == GALP_METHOD_HASH && g->opts.dispatch == GALP_DISPATCH_LEARNED) {
            const int rcf = launch_learned_front_batch(g, w, lo, hi);   // this batch's lists and candidates
            if (rcf != GALP_OK) return rcf;
        }
        const int rc = g->opts.method == GALP_METHOD_HASH ? launch_hash_iteration(g, w)
                                                            : launch_sort_iteration(g, w);
        if (rc != GALP_OK) return rc;
        if (async) {
            if (hipEventRecord(O->ev_free[k], g->stream) != hipSuccess)
                return fail(g->ctx, GALP_EHIP, "out-of-core event");
            O->used[k] = true;
        }
    }
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) return fail(g->ctx, GALP_EHIP, std::string("out-of-core launch: ") + hipGetErrorString(e));
    return GALP_OK;
}

// Learned dispatch set-up: every batch is copied (synchronously, before any
// iteration) and its big vertices' features computed from its rows.
int ooc_features(galp_graph *g) {
    OocState *O = g->ooc;
    const int64_t nb = (int64_t)O->views.size();
    for (int64_t b = 0; b < nb; ++b) {
        const int64_t e0 = O->e0[(size_t)b], ne = O->e1[(size_t)b] - e0;
        const int64_t lo = O->bbs[(size_t)b], hi = O->bbs[(size_t)b + 1];
        if (ne && hipMemcpyAsync(O->d_buf[0], O->h_nbr + e0, 4 * (size_t)ne, hipMemcpyHostToDevice, g->stream) !=
                      hipSuccess)
            return fail(g->ctx, GALP_EHIP, "out-of-core batch copy");
        const int grid = (int)std::max<int64_t>(1, std::min<int64_t>((hi - lo + 256) / 256, 1024));
        if (g->off64)
            k_rebase<int64_t><<<grid, 256, 0, g->stream>>>((const int64_t *)g->d_off, (int64_t *)O->d_offb[0], lo, hi);
        else
            k_rebase<int32_t><<<grid, 256, 0, g->stream>>>((const int32_t *)g->d_off, (int32_t *)O->d_offb[0], lo, hi);
        const int rc = learn_features_view(g, O->d_offb[0], O->d_buf[0], lo, hi);
        if (rc != GALP_OK) return rc;
        // the next copy reuses the buffer
        if (hipStreamSynchronize(g->stream) != hipSuccess) return fail(g->ctx, GALP_EHIP, "feature pass failed");
    }
    return GALP_OK;
}

int launch_ooc_iteration(galp_graph *g) {
    if (!g->ooc) return fail(g->ctx, GALP_ESTATE, "internal: out-of-core state missing");
    return g->off64 ? ooc_round<int64_t>(g) : ooc_round<int32_t>(g);
}

}  // namespace galp
