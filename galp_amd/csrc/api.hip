// api.hip -- the galp C-ABI (include/galp.h): contexts, graph upload and the
// LP driver.
//
// Replaces the host side of the reference's propagators:
//   main.cu:123-128       device warm-up           -> galp_open
//   incore.cuh:34-73      init_gmem/transfer_data  -> galp_graph_upload
//   lfht.cuh:73-112,      run(): init, iterate, count, print, break, copy
//   dpp.cuh:64-102                                 -> galp_lp_run / galp_lp_step
//   label_propagator.h:22-28  get_labels           -> galp_lp_get_labels
// Every HIP status is checked and mapped to a galp_status with a message.

#include <algorithm>
#include <chrono>
#include <cstdio>
#include <cstring>
#include <new>
#include <string>
#include <thread>
#include <vector>

#include "galp.h"
#include "internal.hpp"

namespace galp {

int fail(galp_ctx *ctx, int code, const std::string &msg) {
    if (ctx) ctx->err = msg;
    return code;
}

#define GALP_HIP(ctx, call)                                                                     \
    do {                                                                                        \
        hipError_t _e = (call);                                                                 \
        if (_e != hipSuccess)                                                                   \
            return fail((ctx), GALP_EHIP, std::string(#call) + ": " + hipGetErrorString(_e));   \
    } while (0)

#define GALP_TRY(expr)                \
    do {                              \
        int _rc = (expr);             \
        if (_rc != GALP_OK) return _rc; \
    } while (0)

int Profiler::find(const char *name) {
    for (size_t i = 0; i < aggs.size(); ++i)
        if (aggs[i].name == name) return (int)i;
    aggs.push_back(KernelAgg{});
    aggs.back().name = name;
    return (int)aggs.size() - 1;
}

void prof_begin(galp_graph *g, const char *name, double alg_bytes, PendingLaunch *pl) {
    if (!g->prof.on) return;
    auto take = [&]() {
        hipEvent_t e;
        if (!g->prof.pool.empty()) {
            e = g->prof.pool.back();
            g->prof.pool.pop_back();
        } else if (hipEventCreate(&e) != hipSuccess) {
            e = nullptr;
        }
        return e;
    };
    pl->agg = g->prof.find(name);
    pl->alg_bytes = alg_bytes;
    pl->dev_bytes = nullptr;
    pl->e0 = take();
    pl->e1 = take();
    if (pl->e0) (void)hipEventRecord(pl->e0, g->stream);
}

void prof_begin_dev(galp_graph *g, const char *name, const int64_t *dev_bytes, PendingLaunch *pl) {
    prof_begin(g, name, 0.0, pl);
    if (g->prof.on) pl->dev_bytes = dev_bytes;
}

void prof_end(galp_graph *g, PendingLaunch *pl) {
    if (!g->prof.on) return;
    if (pl->e1) (void)hipEventRecord(pl->e1, g->stream);
    g->prof.pending.push_back(*pl);
}

// After the stream is synchronised: fold pending launch times into totals.
static void prof_collect(galp_graph *g) {
    for (auto &pl : g->prof.pending) {
        float ms = 0.f;
        if (pl.e0 && pl.e1 && hipEventElapsedTime(&ms, pl.e0, pl.e1) == hipSuccess) {
            KernelAgg &a = g->prof.aggs[(size_t)pl.agg];
            a.launches += 1;
            a.total_ms += ms;
            a.alg_bytes += pl.alg_bytes;
            int64_t db = 0;
            if (pl.dev_bytes && hipMemcpy(&db, pl.dev_bytes, 8, hipMemcpyDeviceToHost) == hipSuccess)
                a.alg_bytes += (double)db;
        }
        if (pl.e0) g->prof.pool.push_back(pl.e0);
        if (pl.e1) g->prof.pool.push_back(pl.e1);
    }
    g->prof.pending.clear();
}

// Degree class of a vertex; rows longer than hub_min go to the hub path.
static int bucket_of(int64_t d, int64_t hub_min) {
    if (d <= 0) return -1;
    if (d > hub_min) return B_HUB;
    for (int b = 0; b < NUM_BUCKETS; ++b)
        if (d <= kBucketHi[b]) return b;
    return NUM_BUCKETS - 1;
}

// -l of method 1 (main.cu:98-100): which rows take the tiled hub path.
//   2 (small LDS tables + spill, lfht/kernel.cuh:141-190): rows > 512;
//   3 (load-balanced tiles, lfht/kernel.cuh:193-304): rows > 1024;
//   0, 1 and method 0: rows > 4096.
static int64_t hub_min_of(const galp_opts &o) {
    if (o.method == GALP_METHOD_HASH && o.variant == 2) return 512;
    if (o.method == GALP_METHOD_HASH && o.variant == 3) return 1024;
    return kBucketHi[B_B4K];
}


static void dev_free_all(galp_graph *g) {
    learn_free(g);
    sort_free(g);
    ooc_free(g);
    batch_free(g);
    if (g->gexec) (void)hipGraphExecDestroy(g->gexec);
    g->gexec = nullptr;
    void *ptrs[] = {g->d_off, g->d_nbr, g->d_lab[0], g->d_lab[1], g->d_perm, g->d_iperm, g->d_vlist,
                    g->d_out, g->d_hub_vid, g->d_hub_nparts, g->d_item_hub, g->d_item_part, g->d_item_meta,
                    g->d_hub_best, g->d_merge_big, g->d_hub_tile0, g->d_hub_ntiles, g->d_tile_hub,
                    g->d_tile_idx_off,
                    g->d_tile_idx, g->d_out_key, g->d_out_ent, g->d_counter, g->d_stop, g->d_iters_done, g->d_iter_counts, g->d_err};
    for (void *p : ptrs)
        if (p) (void)hipFree(p);
    if (g->h_pinned) (void)hipHostFree(g->h_pinned);
    for (auto e : g->prof.pool) (void)hipEventDestroy(e);
    for (auto &pl : g->prof.pending) {
        if (pl.e0) (void)hipEventDestroy(pl.e0);
        if (pl.e1) (void)hipEventDestroy(pl.e1);
    }
    for (int k = 0; k < 2; ++k) {
        if (g->side[k]) (void)hipStreamDestroy(g->side[k]);
        if (g->side_fork[k]) (void)hipEventDestroy(g->side_fork[k]);
        if (g->side_join[k]) (void)hipEventDestroy(g->side_join[k]);
    }
    if (g->stream) (void)hipStreamDestroy(g->stream);
}

// Host validation of a borrowed CSR view (O(n + m), multithreaded).
static int validate(galp_ctx *ctx, const galp_csr_view *c, std::vector<int64_t> &off) {
    if (c->n < 0 || c->m < 0 || c->n > INT32_MAX)
        return fail(ctx, GALP_EINVAL, "n must be in [0, 2^31) and m >= 0");
    if (!c->offsets) return fail(ctx, GALP_EINVAL, "offsets is NULL");
    if (c->m > 0 && !c->neighbors) return fail(ctx, GALP_EINVAL, "neighbors is NULL");
    const int64_t n = c->n;
    off.resize((size_t)n + 1);
    for (int64_t i = 0; i <= n; ++i)
        off[(size_t)i] = c->offsets_are_64 ? ((const int64_t *)c->offsets)[i]
                                           : (int64_t)((const int32_t *)c->offsets)[i];
    if (off[0] != 0 || off[(size_t)n] != c->m)
        return fail(ctx, GALP_EINVAL, "offsets[0] must be 0 and offsets[n] must equal m");
    for (int64_t i = 0; i < n; ++i)
        if (off[(size_t)i + 1] < off[(size_t)i]) return fail(ctx, GALP_EINVAL, "offsets are not monotone");
    // neighbour range check, split over host threads
    const int nt = (int)std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    std::vector<int> bad(nt, 0);
    std::vector<std::thread> th;
    const int64_t m = c->m;
    for (int t = 0; t < nt; ++t)
        th.emplace_back([&, t] {
            int64_t lo = m * t / nt, hi = m * (t + 1) / nt;
            for (int64_t e = lo; e < hi; ++e) {
                int32_t u = c->neighbors[e];
                if (u < 0 || (int64_t)u >= n) {
                    bad[t] = 1;
                    return;
                }
            }
        });
    for (auto &x : th) x.join();
    for (int b : bad)
        if (b) return fail(ctx, GALP_EINVAL, "a neighbour id is outside [0, n)");
    return GALP_OK;
}

template <typename E>
static int upload_offsets(galp_graph *g, const std::vector<int64_t> &off) {
    std::vector<E> tmp(off.size());
    for (size_t i = 0; i < off.size(); ++i) tmp[i] = (E)off[i];
    E *d = nullptr;
    GALP_TRY(dev_alloc(g, &d, (int64_t)off.size()));
    g->d_off = d;
    GALP_HIP(g->ctx, hipMemcpy(d, tmp.data(), sizeof(E) * tmp.size(), hipMemcpyHostToDevice));
    return GALP_OK;
}

static int upload_impl(galp_graph *g, const galp_csr_view *c) {
    auto t0 = std::chrono::steady_clock::now();
    galp_ctx *ctx = g->ctx;
    std::vector<int64_t> off;
    GALP_TRY(validate(ctx, c, off));
    const int64_t n = c->n, m = c->m;
    g->n = n;
    g->m = m;
    g->off64 = m > (int64_t)INT32_MAX || g->opts.offsets64 != 0;
    g->reorder = g->opts.reorder != 0;

    std::vector<int64_t> deg((size_t)n);
    int64_t maxdeg = 0;
    for (int64_t v = 0; v < n; ++v) {
        deg[(size_t)v] = off[(size_t)v + 1] - off[(size_t)v];
        maxdeg = std::max(maxdeg, deg[(size_t)v]);
    }

    // ---- dispatch plan -------------------------------------------------
    std::vector<int32_t> iperm, perm;          // reorder
    std::vector<int64_t> off_int;              // offsets in the internal numbering
    std::vector<int32_t> vlist;                // no reorder
    std::vector<int32_t> hub_vertices;         // internal ids of hubs, in class order
    if (g->reorder) {
        iperm.resize((size_t)n);
        perm.resize((size_t)n);
        if (maxdeg <= 8 * n + 1024) {
            std::vector<int64_t> pos((size_t)maxdeg + 2, 0);
            for (int64_t v = 0; v < n; ++v) pos[(size_t)(maxdeg - deg[(size_t)v]) + 1]++;
            for (int64_t d = 0; d <= maxdeg; ++d) pos[(size_t)d + 1] += pos[(size_t)d];
            for (int64_t v = 0; v < n; ++v) {
                int64_t k = pos[(size_t)(maxdeg - deg[(size_t)v])]++;
                iperm[(size_t)k] = (int32_t)v;
            }
        } else {
            for (int64_t v = 0; v < n; ++v) iperm[(size_t)v] = (int32_t)v;
            std::stable_sort(iperm.begin(), iperm.end(),
                             [&](int32_t a, int32_t b) { return deg[(size_t)a] > deg[(size_t)b]; });
        }
        for (int64_t k = 0; k < n; ++k) perm[(size_t)iperm[(size_t)k]] = (int32_t)k;
        off_int.resize((size_t)n + 1);
        off_int[0] = 0;
        for (int64_t k = 0; k < n; ++k) off_int[(size_t)k + 1] = off_int[(size_t)k] + deg[(size_t)iperm[(size_t)k]];
        for (int b = 0; b < NUM_BUCKETS; ++b) g->bucket[b] = BucketRange{};
        for (int64_t k = 0; k < n; ++k) {
            int b = bucket_of(deg[(size_t)iperm[(size_t)k]], hub_min_of(g->opts));
            if (b < 0) continue;
            BucketRange &r = g->bucket[b];
            if (r.nv == 0) r.v0 = k;
            r.nv++;
            r.edges += deg[(size_t)iperm[(size_t)k]];
        }
        for (int64_t k = 0; k < g->bucket[B_HUB].nv; ++k) hub_vertices.push_back((int32_t)(g->bucket[B_HUB].v0 + k));
    } else {
        std::vector<int64_t> cnt(NUM_BUCKETS, 0);
        for (int64_t v = 0; v < n; ++v) {
            int b = bucket_of(deg[(size_t)v], hub_min_of(g->opts));
            if (b >= 0) cnt[(size_t)b]++;
        }
        int64_t acc = 0;
        for (int b = 0; b < NUM_BUCKETS; ++b) {
            g->bucket[b] = BucketRange{};
            g->bucket[b].v0 = acc;
            acc += cnt[(size_t)b];
        }
        vlist.resize((size_t)acc);
        std::vector<int64_t> cur(NUM_BUCKETS);
        for (int b = 0; b < NUM_BUCKETS; ++b) cur[(size_t)b] = g->bucket[b].v0;
        for (int64_t v = 0; v < n; ++v) {
            int b = bucket_of(deg[(size_t)v], hub_min_of(g->opts));
            if (b < 0) continue;
            vlist[(size_t)cur[(size_t)b]++] = (int32_t)v;
            g->bucket[b].nv++;
            g->bucket[b].edges += deg[(size_t)v];
            if (b == B_HUB) hub_vertices.push_back((int32_t)v);
        }
    }
    const std::vector<int64_t> &offi = g->reorder ? off_int : off;

    // big vertices (classes W128..HUB) for the learned dispatcher, hubs first
    g->big_vid.clear();
    g->big_deg.clear();
    g->big_off.clear();
    for (int32_t x : hub_vertices) {
        g->big_vid.push_back(x);
        g->big_deg.push_back(offi[(size_t)x + 1] - offi[(size_t)x]);
        g->big_off.push_back(offi[(size_t)x]);
    }
    for (int b = B_W128; b < B_HUB; ++b)
        for (int64_t k = 0; k < g->bucket[b].nv; ++k) {
            const int32_t x = g->reorder ? (int32_t)(g->bucket[b].v0 + k) : vlist[(size_t)(g->bucket[b].v0 + k)];
            g->big_vid.push_back(x);
            g->big_deg.push_back(offi[(size_t)x + 1] - offi[(size_t)x]);
            g->big_off.push_back(offi[(size_t)x]);
        }

    // hubs: phase-1 tiles and phase-2 (hub, label partition) items
    std::vector<int32_t> hub_nparts, item_hub, item_part, hub_tile0, hub_ntiles, tile_hub;
    std::vector<int64_t> tile_idx_off;
    std::vector<HubItem> item_meta;
    int64_t ntile_idx = 0;
    for (size_t h = 0; h < hub_vertices.size(); ++h) {
        const int32_t x = hub_vertices[h];
        const int64_t d = offi[(size_t)x + 1] - offi[(size_t)x];
        // partitions are capped at kHubMaxParts (the tile kernels' offsets):
        // beyond kHubPart * kHubMaxParts edges a partition may hold more
        // distinct labels than a merge table, and the merges then take it in
        // several rounds
        const int64_t part = g->opts.method == GALP_METHOD_SORT ? kHubPartSort : kHubPart;
        const int32_t P = (int32_t)std::min<int64_t>(kHubMaxParts, (d + part - 1) / part);
        const int32_t nt = (int32_t)((d + kHubTile - 1) / kHubTile);
        hub_nparts.push_back(P);
        hub_tile0.push_back((int32_t)tile_hub.size());
        hub_ntiles.push_back(nt);
        const int64_t idx0 = ntile_idx;
        for (int32_t t = 0; t < nt; ++t) {
            tile_hub.push_back((int32_t)h);
            tile_idx_off.push_back(ntile_idx);
            ntile_idx += P + 1;
        }
        for (int32_t p = 0; p < P; ++p) {
            item_hub.push_back((int32_t)h);
            item_part.push_back(p);
            item_meta.push_back(HubItem{(int32_t)h, p, hub_tile0.back(), nt, P, 0, idx0});
        }
    }
    g->nhubs = (int64_t)hub_vertices.size();
    g->nhub_items = (int64_t)item_hub.size();
    g->nhub_tiles = (int64_t)tile_hub.size();
    g->ntile_idx = ntile_idx;

    // ---- memory check (check_gpu_mem, main.cu:51-64) --------------------
    {
        const int64_t esz = g->off64 ? 8 : 4;
        const bool ooc = g->opts.policy == GALP_POLICY_OOC_SYNC || g->opts.policy == GALP_POLICY_OOC_ASYNC;
        // out-of-core: two batch buffers and two rebased offset arrays instead of the neighbour array
        const int64_t resident = ooc ? 2 * 4 * std::min<int64_t>(m, (int64_t)1 << g->opts.buffer_pow) +
                                           2 * esz * (n + 1)
                                     : 4 * m;
        int64_t need = resident + esz * (n + 1) + 4 * 3 * n + 32 * g->nhubs + 4 * g->nhub_tiles + 4096;
        // the hub plan of both methods: merge items, tile outputs and partition
        // offsets; the sort method's list of large merge items
        need += (12 + (int64_t)sizeof(HubItem)) * g->nhub_items + (8 + 8 * kHubTile) * g->nhub_tiles +
                4 * ntile_idx;
        if (g->reorder) need += 8 * n + (ooc ? 0 : 4 * m + 8 * (n + 1));  // perm arrays + transient original CSR
        else need += 4 * n;
        size_t fr = 0, tot = 0;
        GALP_HIP(ctx, hipMemGetInfo(&fr, &tot));
        int64_t limit = g->opts.mem_budget > 0 ? std::min<int64_t>(g->opts.mem_budget, (int64_t)fr) : (int64_t)fr;
        if (need > limit)
            return fail(ctx, GALP_ENOMEM, "not enough GPU memory: need " + std::to_string(need) +
                                              " bytes, available " + std::to_string(limit));
    }

    const bool ooc = g->opts.policy == GALP_POLICY_OOC_SYNC || g->opts.policy == GALP_POLICY_OOC_ASYNC;
    GALP_HIP(ctx, hipStreamCreateWithFlags(&g->stream, hipStreamNonBlocking));
    if (!ooc)
        for (int k = 0; k < 2; ++k) {
            GALP_HIP(ctx, hipStreamCreateWithFlags(&g->side[k], hipStreamNonBlocking));
            GALP_HIP(ctx, hipEventCreateWithFlags(&g->side_fork[k], hipEventDisableTiming));
            GALP_HIP(ctx, hipEventCreateWithFlags(&g->side_join[k], hipEventDisableTiming));
        }
    if (!ooc) GALP_TRY(dev_alloc(g, &g->d_nbr, m));
    GALP_TRY(dev_alloc(g, &g->d_lab[0], n));
    GALP_TRY(dev_alloc(g, &g->d_lab[1], n));
    GALP_TRY(dev_alloc(g, &g->d_counter, 1));
    GALP_TRY(dev_alloc(g, &g->d_stop, 1));
    GALP_TRY(dev_alloc(g, &g->d_iters_done, 1));
    GALP_TRY(dev_alloc(g, &g->d_err, 1));
    GALP_TRY(dev_alloc(g, &g->d_iter_counts, kMaxIters));
    GALP_HIP(ctx, hipHostMalloc((void **)&g->h_pinned, 64, hipHostMallocDefault));

    if (g->off64) GALP_TRY(upload_offsets<int64_t>(g, offi));
    else GALP_TRY(upload_offsets<int32_t>(g, offi));

    if (g->reorder) {
        GALP_TRY(dev_alloc(g, &g->d_perm, n));
        GALP_TRY(dev_alloc(g, &g->d_iperm, n));
        GALP_TRY(dev_alloc(g, &g->d_out, n));
        GALP_HIP(ctx, hipMemcpy(g->d_perm, perm.data(), 4 * (size_t)n, hipMemcpyHostToDevice));
        GALP_HIP(ctx, hipMemcpy(g->d_iperm, iperm.data(), 4 * (size_t)n, hipMemcpyHostToDevice));
    }
    if (g->reorder && !ooc) {
        int64_t *d_off_old = nullptr;
        int32_t *d_nbr_old = nullptr;
        const int64_t before = g->device_bytes;
        GALP_TRY(dev_alloc(g, &d_off_old, n + 1));
        int rc = dev_alloc(g, &d_nbr_old, m);
        if (rc != GALP_OK) {
            (void)hipFree(d_off_old);
            return rc;
        }
        hipError_t e1 = hipMemcpy(d_off_old, off.data(), 8 * (size_t)(n + 1), hipMemcpyHostToDevice);
        hipError_t e2 = m ? hipMemcpy(d_nbr_old, c->neighbors, 4 * (size_t)m, hipMemcpyHostToDevice) : hipSuccess;
        rc = (e1 == hipSuccess && e2 == hipSuccess) ? dev_permute_rows(g, d_off_old, d_nbr_old, g->d_nbr)
                                                    : fail(ctx, GALP_EHIP, "upload of the original CSR failed");
        hipError_t e3 = hipStreamSynchronize(g->stream);
        (void)hipFree(d_off_old);
        (void)hipFree(d_nbr_old);
        g->device_bytes = before;
        if (rc != GALP_OK) return rc;
        if (e3 != hipSuccess) return fail(ctx, GALP_EHIP, std::string("permute: ") + hipGetErrorString(e3));
    } else if (!g->reorder) {
        if (m && !ooc) GALP_HIP(ctx, hipMemcpy(g->d_nbr, c->neighbors, 4 * (size_t)m, hipMemcpyHostToDevice));
        if (!vlist.empty()) {
            GALP_TRY(dev_alloc(g, &g->d_vlist, (int64_t)vlist.size()));
            GALP_HIP(ctx, hipMemcpy(g->d_vlist, vlist.data(), 4 * vlist.size(), hipMemcpyHostToDevice));
        }
    }

    if (g->nhubs) {
        GALP_TRY(dev_alloc(g, &g->d_hub_vid, g->nhubs));
        GALP_TRY(dev_alloc(g, &g->d_hub_nparts, g->nhubs));
        GALP_TRY(dev_alloc(g, &g->d_hub_best, g->nhubs));
        GALP_TRY(dev_alloc(g, &g->d_merge_big, g->nhub_items));
        auto up = [&](void *dst, const void *src, size_t bytes) {
            return hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice);
        };
        GALP_HIP(ctx, up(g->d_hub_vid, hub_vertices.data(), 4 * hub_vertices.size()));
        GALP_HIP(ctx, up(g->d_hub_nparts, hub_nparts.data(), 4 * hub_nparts.size()));
        // row-order tiles of kHubTile edges
        GALP_TRY(dev_alloc(g, &g->d_hub_tile0, g->nhubs));
        GALP_TRY(dev_alloc(g, &g->d_hub_ntiles, g->nhubs));
        GALP_TRY(dev_alloc(g, &g->d_tile_hub, g->nhub_tiles));
        GALP_TRY(dev_alloc(g, &g->d_tile_idx_off, g->nhub_tiles));
        GALP_TRY(dev_alloc(g, &g->d_tile_idx, ntile_idx));
        GALP_TRY(dev_alloc(g, &g->d_out_key, g->nhub_tiles * kHubTile));
        GALP_TRY(dev_alloc(g, &g->d_out_ent, g->nhub_tiles * kHubTile));
        GALP_HIP(ctx, up(g->d_hub_tile0, hub_tile0.data(), 4 * hub_tile0.size()));
        GALP_HIP(ctx, up(g->d_hub_ntiles, hub_ntiles.data(), 4 * hub_ntiles.size()));
        GALP_HIP(ctx, up(g->d_tile_hub, tile_hub.data(), 4 * tile_hub.size()));
        GALP_HIP(ctx, up(g->d_tile_idx_off, tile_idx_off.data(), 8 * tile_idx_off.size()));
        // merge items (hub, label partition) of both methods' hub paths
        GALP_TRY(dev_alloc(g, &g->d_item_hub, g->nhub_items));
        GALP_TRY(dev_alloc(g, &g->d_item_part, g->nhub_items));
        GALP_HIP(ctx, up(g->d_item_hub, item_hub.data(), 4 * item_hub.size()));
        GALP_HIP(ctx, up(g->d_item_part, item_part.data(), 4 * item_part.size()));
        GALP_TRY(dev_alloc(g, &g->d_item_meta, g->nhub_items));
        GALP_HIP(ctx, up(g->d_item_meta, item_meta.data(), sizeof(HubItem) * item_meta.size()));
        GALP_TRY(dev_fill_u32(g, (uint32_t *)g->d_hub_best, 2 * g->nhubs, 0u));
        GALP_TRY(dev_fill_u32(g, (uint32_t *)g->d_merge_big, g->nhub_items, 0u));
    }
    if (ooc) {
        std::vector<int64_t> hub_item0(hub_nparts.size() + 1, 0);
        for (size_t h = 0; h < hub_nparts.size(); ++h) hub_item0[h + 1] = hub_item0[h] + hub_nparts[h];
        GALP_TRY(ooc_setup(g, c, offi, perm, iperm, vlist, hub_vertices, hub_tile0, hub_item0));
    }
    if (g->opts.method == GALP_METHOD_SORT) GALP_TRY(sort_setup(g));
    GALP_HIP(ctx, hipStreamSynchronize(g->stream));
    if (g->opts.dispatch == GALP_DISPATCH_LEARNED) GALP_TRY(learn_setup(g));
    g->preprocess_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    return GALP_OK;
}

IterView full_view(const galp_graph *g) {
    IterView w;
    w.off = g->d_off;
    w.nbr = g->d_nbr;
    for (int b = 0; b < NUM_BUCKETS; ++b) w.cls[b] = g->bucket[b];
    w.h0 = 0;
    w.h1 = g->nhubs;
    w.t0 = 0;
    w.t1 = g->nhub_tiles;
    w.i0 = 0;
    w.i1 = g->nhub_items;
    return w;
}

bool concurrent_allowed(const galp_graph *g) {
    return !g->ooc && !g->prof.on && g->side[0] && g->side[1] && g->opts.variant != 0 &&
           g->opts.policy != GALP_POLICY_LI;
}

bool concurrent_classes(const galp_graph *g) {
    if (!concurrent_allowed(g) || g->streams == 1) return false;
    const int64_t span = (g->batch && !g->reorder) ? g->batch->max_n : g->n;   // ids a gather can reach
    return g->streams == 3 || (g->m >= kConcurrentMinEdges && span <= kConcurrentMaxVertices);
}

int side_begin(galp_graph *g, int k, hipStream_t *main_stream) {
    *main_stream = g->stream;
    GALP_HIP(g->ctx, hipEventRecord(g->side_fork[k], g->stream));
    GALP_HIP(g->ctx, hipStreamWaitEvent(g->side[k], g->side_fork[k], 0));
    g->stream = g->side[k];
    return GALP_OK;
}

int side_end(galp_graph *g, int k, hipStream_t main_stream) {
    g->stream = main_stream;
    GALP_HIP(g->ctx, hipEventRecord(g->side_join[k], g->side[k]));
    return GALP_OK;
}

int side_join(galp_graph *g, int k) {
    GALP_HIP(g->ctx, hipStreamWaitEvent(g->stream, g->side_join[k], 0));
    return GALP_OK;
}

static int launch_iteration(galp_graph *g) {
    if (g->batch) GALP_TRY(launch_batch_lists(g));   // drops the graphs that stopped
    switch (g->opts.method) {
    case GALP_METHOD_HASH:
        if (g->opts.dispatch == GALP_DISPATCH_LEARNED) {
            if (!g->learn || !g->learn->H)
                return fail(g->ctx, GALP_ESTATE, "learned dispatch needs a network: call galp_graph_set_mlp");
        }
        // out of core, the learned front runs per batch
        if (g->ooc) GALP_TRY(launch_ooc_iteration(g));
        else GALP_TRY(launch_hash_iteration_incore(g));
        if (g->learn) GALP_TRY(learn_after_iteration(g));
        break;
    case GALP_METHOD_SORT:
        if (g->ooc) GALP_TRY(launch_ooc_iteration(g));
        else GALP_TRY(launch_sort_iteration(g, full_view(g)));
        break;
    default:
        return fail(g->ctx, GALP_EUNSUPPORTED, "method " + std::to_string(g->opts.method) + " is not built");
    }
    return g->batch ? launch_batch_end(g) : launch_iter_end(g);
}

// The iteration loop of galp_lp_run as one hipGraph: captured on the
// graph's stream the first time a run of `niter` iterations is asked for,
// replayed afterwards.  The captured launches hold the same arguments the
// eager loop passes (ping-pong buffers, iteration slots); every run starts
// from galp_lp_init, so the replay starts from the state the capture saw.
// The loop records no memset node: every reset is a kernel.
static int run_captured(galp_graph *g, int niter) {
    if (!g->gexec || g->gexec_niter != niter) {
        if (g->gexec) (void)hipGraphExecDestroy(g->gexec);
        g->gexec = nullptr;
        const int cur0 = g->cur;
        const int dcur0 = g->learn ? g->learn->dcur : 0;
        GALP_HIP(g->ctx, hipStreamBeginCapture(g->stream, hipStreamCaptureModeThreadLocal));
        int rc = GALP_OK;
        for (int i = 0; i < niter && rc == GALP_OK; ++i) {
            g->iter = i;
            rc = launch_iteration(g);
            g->cur ^= 1;
        }
        hipGraph_t graph = nullptr;
        const hipError_t e = hipStreamEndCapture(g->stream, &graph);
        g->cur = cur0;
        if (g->learn) g->learn->dcur = dcur0;
        if (rc != GALP_OK) {
            if (graph) (void)hipGraphDestroy(graph);
            return rc;
        }
        if (e != hipSuccess) return fail(g->ctx, GALP_EHIP, std::string("graph capture: ") + hipGetErrorString(e));
        const hipError_t ei = hipGraphInstantiate(&g->gexec, graph, nullptr, nullptr, 0);
        (void)hipGraphDestroy(graph);
        if (ei != hipSuccess) {
            g->gexec = nullptr;
            return fail(g->ctx, GALP_EHIP, std::string("graph instantiate: ") + hipGetErrorString(ei));
        }
        g->gexec_niter = niter;
    }
    GALP_HIP(g->ctx, hipGraphLaunch(g->gexec, g->stream));
    if (g->learn && (niter & 1)) g->learn->dcur ^= 1;   // as the eager loop leaves it
    return GALP_OK;
}

static int check_err(galp_graph *g) {
    GALP_HIP(g->ctx, hipMemcpyAsync(g->h_pinned + 8, g->d_err, 4, hipMemcpyDeviceToHost, g->stream));
    GALP_HIP(g->ctx, hipStreamSynchronize(g->stream));
    if (g->h_pinned[8] != 0)
        return fail(g->ctx, GALP_EHIP, "kernel reported an internal table overflow (code " +
                                           std::to_string(g->h_pinned[8]) + ")");
    return GALP_OK;
}

}  // namespace galp

using namespace galp;

extern "C" {

GALP_API void galp_opts_default(galp_opts *o) {
    if (!o) return;
    std::memset(o, 0, sizeof(*o));
    o->method = GALP_METHOD_HASH;
    o->policy = GALP_POLICY_INCORE;
    o->variant = 1;
    o->buffer_pow = 24;
    o->dispatch = GALP_DISPATCH_PLAIN;
    o->reorder = 1;
    o->profile = 0;
    o->mem_budget = 0;
}

GALP_API int galp_open(int device, galp_ctx **out) {
    if (!out) return GALP_EINVAL;
    *out = nullptr;
    galp_ctx *ctx = new (std::nothrow) galp_ctx();
    if (!ctx) return GALP_ENOMEM;
    int count = 0;
    hipError_t e = hipGetDeviceCount(&count);
    if (e != hipSuccess || count == 0 || device < 0 || device >= count) {
        delete ctx;
        return GALP_EHIP;
    }
    if (hipSetDevice(device) != hipSuccess || hipGetDeviceProperties(&ctx->prop, device) != hipSuccess ||
        hipFree(nullptr) != hipSuccess) {
        delete ctx;
        return GALP_EHIP;
    }
    ctx->device = device;
    if (mlp_scratch_setup(ctx, device) != GALP_OK) {
        delete ctx;
        return GALP_EHIP;
    }
    *out = ctx;
    return GALP_OK;
}

GALP_API int galp_close(galp_ctx *ctx) {
    delete ctx;
    return GALP_OK;
}

GALP_API const char *galp_last_error(galp_ctx *ctx) { return ctx ? ctx->err.c_str() : "null context"; }

GALP_API int galp_device_query(galp_ctx *ctx, galp_device_info *out) {
    if (!ctx || !out) return GALP_EINVAL;
    std::memset(out, 0, sizeof(*out));
    std::snprintf(out->name, sizeof(out->name), "%s", ctx->prop.name);
    std::snprintf(out->arch, sizeof(out->arch), "%s", ctx->prop.gcnArchName);
    out->compute_units = ctx->prop.multiProcessorCount;
    out->wavefront_size = ctx->prop.warpSize;
    out->lds_per_block = (int64_t)ctx->prop.sharedMemPerBlock;
    out->total_mem = (int64_t)ctx->prop.totalGlobalMem;
    out->l2_bytes = (int64_t)ctx->prop.l2CacheSize;
    return GALP_OK;
}

GALP_API int galp_graph_upload(galp_ctx *ctx, const galp_csr_view *csr, const galp_opts *opts,
                               galp_graph **out) {
    if (!ctx || !csr || !out) return ctx ? fail(ctx, GALP_EINVAL, "null argument") : GALP_EINVAL;
    *out = nullptr;
    galp_opts o;
    if (opts) o = *opts; else galp_opts_default(&o);
    if (o.method != GALP_METHOD_SORT && o.method != GALP_METHOD_HASH)
        return fail(ctx, GALP_EINVAL, "method must be 0 (sort) or 1 (hash)");
    if (o.policy == 3) return fail(ctx, GALP_EUNSUPPORTED, "policy 3 (CPU-GPU hybrid) is not part of the one-GPU build");
    if (o.policy == 5 || (o.policy == 4 && o.method == GALP_METHOD_HASH))
        return fail(ctx, GALP_EUNSUPPORTED, "multi-GPU policies are not part of the one-GPU build");
    if (o.policy != GALP_POLICY_INCORE && o.policy != GALP_POLICY_LI && o.policy != GALP_POLICY_OOC_SYNC &&
        o.policy != GALP_POLICY_OOC_ASYNC)
        return fail(ctx, GALP_EUNSUPPORTED, "policy " + std::to_string(o.policy) + " is not built");
    if ((o.policy == GALP_POLICY_OOC_SYNC || o.policy == GALP_POLICY_OOC_ASYNC) &&
        (o.buffer_pow < 10 || o.buffer_pow > 34))
        return fail(ctx, GALP_EINVAL, "buffer_pow (-b) must be in [10, 34] for the out-of-core policies");
    if ((o.policy == GALP_POLICY_OOC_SYNC || o.policy == GALP_POLICY_OOC_ASYNC) && o.dispatch == GALP_DISPATCH_LEARNED &&
        !o.reorder)
        return fail(ctx, GALP_EUNSUPPORTED, "learned dispatch out-of-core needs the degree renumbering (reorder)");
    if (o.dispatch == GALP_DISPATCH_LEARNED && (o.method != GALP_METHOD_HASH || o.variant != 1))
        return fail(ctx, GALP_EUNSUPPORTED, "learned dispatch routes method 1, -l 1");
    if (o.variant < 0 || o.variant > 3) return fail(ctx, GALP_EINVAL, "variant (-l) must be 0..3");
    if (hipSetDevice(ctx->device) != hipSuccess) return fail(ctx, GALP_EHIP, "hipSetDevice failed");
    galp_graph *g = new (std::nothrow) galp_graph();
    if (!g) return fail(ctx, GALP_ENOMEM, "host allocation failed");
    g->ctx = ctx;
    g->opts = o;
    g->prof.on = o.profile != 0;
    int rc;
    try {
        rc = upload_impl(g, csr);
    } catch (const std::bad_alloc &) {
        rc = fail(ctx, GALP_ENOMEM, "host allocation failed during upload");
    }
    if (rc != GALP_OK) {
        dev_free_all(g);
        delete g;
        return rc;
    }
    *out = g;
    return GALP_OK;
}

GALP_API int galp_graph_free(galp_graph *g) {
    if (!g) return GALP_OK;
    (void)hipStreamSynchronize(g->stream);
    dev_free_all(g);
    delete g;
    return GALP_OK;
}

GALP_API int galp_graph_query(galp_graph *g, galp_graph_info *out) {
    if (!g || !out) return GALP_EINVAL;
    std::memset(out, 0, sizeof(*out));
    out->n = g->n;
    out->m = g->m;
    out->offsets_are_64 = g->off64;
    out->nbuckets = NUM_BUCKETS;
    for (int b = 0; b < NUM_BUCKETS; ++b) {
        out->bucket_vertices[b] = g->bucket[b].nv;
        out->bucket_edges[b] = g->bucket[b].edges;
    }
    out->device_bytes = g->device_bytes;
    out->preprocess_ms = g->preprocess_ms;
    return GALP_OK;
}

GALP_API int galp_lp_init(galp_graph *g) {
    if (!g) return GALP_EINVAL;
    GALP_TRY(dev_init_labels(g));
    GALP_TRY(learn_reset_iteration(g));
    GALP_TRY(batch_reset(g));
    GALP_HIP(g->ctx, hipStreamSynchronize(g->stream));
    g->cur = 0;
    g->iter = 0;
    g->initialized = true;
    return GALP_OK;
}

GALP_API int galp_lp_set_labels(galp_graph *g, const int32_t *host_labels) {
    if (!g || (!host_labels && g->n)) return GALP_EINVAL;
    for (int64_t v = 0; v < g->n; ++v)
        if (host_labels[v] < 0 || (int64_t)host_labels[v] >= g->n)
            return fail(g->ctx, GALP_EINVAL, "labels must lie in [0, n)");
    GALP_TRY(dev_init_labels(g));   // resets the control words
    GALP_TRY(learn_reset_iteration(g));
    GALP_TRY(batch_reset(g));
    if (g->n) {
        int32_t *dst = g->d_lab[0];
        if (g->reorder) {
            GALP_HIP(g->ctx, hipMemcpyAsync(g->d_out, host_labels, 4 * (size_t)g->n, hipMemcpyHostToDevice, g->stream));
            GALP_TRY(dev_unmap_labels(g, g->d_out, dst));
        } else {
            GALP_HIP(g->ctx, hipMemcpyAsync(dst, host_labels, 4 * (size_t)g->n, hipMemcpyHostToDevice, g->stream));
        }
        // both buffers hold the state: vertices without neighbours are never rewritten
        GALP_HIP(g->ctx, hipMemcpyAsync(g->d_lab[1], dst, 4 * (size_t)g->n, hipMemcpyDeviceToDevice, g->stream));
    }
    GALP_HIP(g->ctx, hipStreamSynchronize(g->stream));
    g->cur = 0;
    g->iter = 0;
    g->initialized = true;
    return GALP_OK;
}

GALP_API int galp_lp_step(galp_graph *g, galp_step_stats *st) {
    if (!g) return GALP_EINVAL;
    if (!g->initialized) return fail(g->ctx, GALP_ESTATE, "galp_lp_step before galp_lp_init");
    hipEvent_t e0, e1;
    GALP_HIP(g->ctx, hipEventCreate(&e0));
    GALP_HIP(g->ctx, hipEventCreate(&e1));
    int rc = dev_clear_stop(g);
    if (rc == GALP_OK) {
        (void)hipEventRecord(e0, g->stream);
        rc = launch_iteration(g);
        (void)hipEventRecord(e1, g->stream);
    }
    const int idx = g->iter % kMaxIters;
    if (rc == GALP_OK) {
        hipError_t e = hipMemcpyAsync(g->h_pinned, g->d_iter_counts + idx, 8, hipMemcpyDeviceToHost, g->stream);
        g->h_pinned[9] = 0;
        if (e == hipSuccess && learned_now(g))
            e = hipMemcpyAsync(g->h_pinned + 9, g->learn->d_plist_n + kCtrOverflow, 4, hipMemcpyDeviceToHost,
                               g->stream);
        if (e == hipSuccess) e = hipMemcpyAsync(g->h_pinned + 2, g->d_stop, 4, hipMemcpyDeviceToHost, g->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(g->stream);
        if (e != hipSuccess) rc = fail(g->ctx, GALP_EHIP, std::string("step readback: ") + hipGetErrorString(e));
    }
    float ms = 0.f;
    if (rc == GALP_OK) (void)hipEventElapsedTime(&ms, e0, e1);
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (rc == GALP_OK) rc = check_err(g);
    prof_collect(g);
    if (rc != GALP_OK) return rc;
    long long cnt;
    std::memcpy(&cnt, g->h_pinned, 8);
    g->cur ^= 1;
    g->iter += 1;
    if (st) {
        st->updated = cnt;
        st->overflowed = g->h_pinned[9];
        st->device_ms = ms;
        st->iteration = g->iter;
        st->stopped = g->h_pinned[2] != 0;
    }
    return GALP_OK;
}

GALP_API int galp_lp_get_labels(galp_graph *g, int32_t *host_out) {
    if (!g || !host_out) return GALP_EINVAL;
    if (!g->initialized) return fail(g->ctx, GALP_ESTATE, "labels read before galp_lp_init");
    if (!g->n) return GALP_OK;
    const int32_t *src = g->d_lab[g->cur];
    if (g->reorder) {
        GALP_TRY(dev_map_labels(g, src, g->d_out));
        src = g->d_out;
    }
    GALP_HIP(g->ctx, hipMemcpyAsync(host_out, src, 4 * (size_t)g->n, hipMemcpyDeviceToHost, g->stream));
    GALP_HIP(g->ctx, hipStreamSynchronize(g->stream));
    return GALP_OK;
}

GALP_API int galp_lp_run(galp_graph *g, int32_t niter, int32_t verbose, int32_t *labels_out,
                         galp_run_result *res) {
    if (!g) return GALP_EINVAL;
    if (niter < 0 || niter > kMaxIters) return fail(g->ctx, GALP_EINVAL, "niter must be in [0, 4096]");
    using clk = std::chrono::steady_clock;
    GALP_TRY(galp_lp_init(g));
    auto t1 = clk::now();
    double dev_ms = 0.0;
    int done = 0;
    if (verbose) {
        for (int i = 0; i < niter; ++i) {
            auto ti = clk::now();
            galp_step_stats st;
            GALP_TRY(galp_lp_step(g, &st));
            double s = std::chrono::duration<double>(clk::now() - ti).count();
            std::printf("%d: %f (s), updated %d\n", i + 1, s, (int)st.updated);
            dev_ms += st.device_ms;
            ++done;
            if (st.stopped) break;
        }
        std::fflush(stdout);
    } else {
        hipEvent_t e0, e1;
        GALP_HIP(g->ctx, hipEventCreate(&e0));
        GALP_HIP(g->ctx, hipEventCreate(&e1));
        int rc = dev_clear_stop(g);
        (void)hipEventRecord(e0, g->stream);
        const int cur0 = g->cur;
        if (g->opts.graph_capture && !g->ooc && !g->prof.on && niter > 0) {
            rc = run_captured(g, niter);
        } else {
            for (int i = 0; i < niter && rc == GALP_OK; ++i) {
                g->iter = i;
                rc = launch_iteration(g);
                g->cur ^= 1;
            }
        }
        (void)hipEventRecord(e1, g->stream);
        if (rc == GALP_OK) {
            hipError_t e = hipMemcpyAsync(g->h_pinned + 4, g->d_iters_done, 4, hipMemcpyDeviceToHost, g->stream);
            if (e == hipSuccess) e = hipStreamSynchronize(g->stream);
            if (e != hipSuccess) rc = fail(g->ctx, GALP_EHIP, std::string("run readback: ") + hipGetErrorString(e));
        }
        float ms = 0.f;
        if (rc == GALP_OK) (void)hipEventElapsedTime(&ms, e0, e1);
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        if (rc == GALP_OK) rc = check_err(g);
        prof_collect(g);
        if (rc != GALP_OK) return rc;
        done = niter ? g->h_pinned[4] : 0;
        g->iter = done;
        g->cur = (cur0 + done) & 1;
        dev_ms = ms;
    }
    if (labels_out) GALP_TRY(galp_lp_get_labels(g, labels_out));
    GALP_HIP(g->ctx, hipDeviceSynchronize());
    double f1 = std::chrono::duration<double>(clk::now() - t1).count();
    if (res) {
        res->f1_s = f1;
        res->f2_s = f1 + g->preprocess_ms / 1000.0;
        res->device_ms = dev_ms;
        res->iterations = done;
    }
    return GALP_OK;
}

GALP_API int galp_profile_read(galp_graph *g, galp_kernel_stat *out, int32_t max, int32_t *count) {
    if (!g || !count) return GALP_EINVAL;
    *count = (int32_t)g->prof.aggs.size();
    for (int32_t i = 0; i < max && i < *count; ++i) {
        const KernelAgg &a = g->prof.aggs[(size_t)i];
        std::memset(&out[i], 0, sizeof(out[i]));
        std::snprintf(out[i].name, sizeof(out[i].name), "%s", a.name.c_str());
        out[i].launches = a.launches;
        out[i].total_ms = a.total_ms;
        out[i].alg_bytes = a.alg_bytes;
    }
    return GALP_OK;
}

GALP_API int galp_graph_dispatch_stats(galp_graph *g, int64_t *out, int32_t max, int32_t *count) {
    if (!g || !count) return GALP_EINVAL;
    *count = kNumCtrs;
    if (!g->learn) {
        for (int32_t i = 0; i < max && i < kNumCtrs; ++i) out[i] = 0;
        return GALP_OK;
    }
    int64_t tmp[kNumCtrs];
    GALP_HIP(g->ctx, hipMemcpy(tmp, g->learn->d_plist_n, sizeof(tmp), hipMemcpyDeviceToHost));
    for (int32_t i = 0; i < max && i < kNumCtrs; ++i) out[i] = tmp[i];
    return GALP_OK;
}

GALP_API int galp_profile_reset(galp_graph *g) {
    if (!g) return GALP_EINVAL;
    for (auto &a : g->prof.aggs) {
        a.launches = 0;
        a.total_ms = 0.0;
        a.alg_bytes = 0.0;
    }
    return GALP_OK;
}

}  // extern "C"
