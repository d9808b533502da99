// lp_hash.hip -- method 1 (hash-table LP) on gfx950.
//
// Replaces the lock-free hash LP of the reference:
//   update_lockfree<64>          label_propagation/src/lfht/kernel.cuh:51-92
//   _increment / GlobalHT / SharedHT  src/lfht/hashtable.cuh:18-79
//   LFHTBase::perform_lp         src/lfht/lfht.cuh:115-167
// Design (not a translation): the reference launches one 64-thread block per
// vertex against per-vertex slices of two global tables that it clears with
// two memsets every iteration.  Here the vertices are split into degree
// classes, each with its own kernel:
//   - degree <= 64: aligned groups of G = 4..64 lanes per vertex; counts and
//     last positions come from G lane shuffles, no table at all;
//   - 65..4096: one wave (or one 256-thread workgroup) per vertex with an
//     open-addressing table in LDS sized to the vertex's degree;
//   - hubs (> 4096): 2048-edge tiles aggregate in LDS, then merge (count add,
//     last max) into a per-hub global table; a finalize kernel takes the
//     argmax and clears the table for the next iteration (no memsets).
// All kernels are synchronous (read prev, write next) and apply the tie rule
// argmax (count, last position), so the result is independent of scheduling
// and of atomic order.

#include "device_common.hpp"
#include "internal.hpp"

namespace galp {

const int64_t kBucketHi[NUM_BUCKETS] = {4, 8, 16, 32, 64, 512, 4096, INT64_MAX};
const char *const kBucketName[NUM_BUCKETS] = {"g4", "g8", "g16", "g32", "g64", "wave", "block", "hub"};

__device__ __forceinline__ int ceil_log2_dev(uint32_t x) {
    return x <= 1 ? 0 : 32 - __clz((int)(x - 1));
}

// ---------------------------------------------------------------------------
// Degree <= G: one aligned group of G lanes per vertex.  Lane p holds the
// label of the p-th neighbour; a sweep of G shuffles gives, per lane, the
// count of its label and the last position holding it.  The group max of
// (count << 8 | last) is the tie rule; the winning label is read back from
// lane `last`.
template <int G, typename E>
__global__ void __launch_bounds__(256)
k_hash_group(const E *__restrict__ off, const int32_t *__restrict__ nbr,
             const int32_t *__restrict__ prev, int32_t *__restrict__ next,
             const int32_t *__restrict__ vlist, int64_t v0, int64_t nv,
             unsigned long long *__restrict__ counter, const int32_t *__restrict__ stop) {
    __shared__ uint32_t scratch[256 / kWave];
    if (*stop) return;
    const int p = lane_id() & (G - 1);
    const int64_t per_block = blockDim.x / G;
    uint32_t changed = 0;
    for (int64_t base = (int64_t)blockIdx.x * per_block; base < nv;
         base += (int64_t)gridDim.x * per_block) {
        const int64_t i = base + threadIdx.x / G;
        const bool vok = i < nv;
        int64_t v = 0;
        E b = 0, e = 0;
        if (vok) {
            v = vlist ? (int64_t)vlist[v0 + i] : v0 + i;
            b = off[v];
            e = off[v + 1];
        }
        const int deg = (int)(e - b);
        const bool ok = p < deg;
        const int32_t L = ok ? prev[nbr[b + p]] : -1;
        uint32_t cnt = 0;
        int last = 0;
#pragma unroll
        for (int j = 0; j < G; ++j) {
            const int32_t Lj = __shfl(L, j, G);
            const bool eq = (Lj == L) & (j < deg);
            cnt += eq;
            last = eq ? j : last;
        }
        uint32_t key = ok ? ((cnt << 8) | (uint32_t)last) : 0u;
        key = group_max<G>(key);
        const int32_t win = __shfl(L, (int)(key & 0xffu), G);
        if (ok && p == 0) {
            const int32_t old = prev[v];
            next[v] = win;
            changed += (win != old);
        }
    }
    block_count_add(changed, counter, scratch);
}

// ---------------------------------------------------------------------------
// 64 < degree <= T*EPT: one block of T threads per vertex, LDS table of
// 2^ceil(log2(2*deg)) slots (load factor <= 1/2).  Pass 1 inserts (CAS on the
// key) and counts (atomic add); the block max of the returned counts is the
// winning count C.  Pass 2 re-reads each edge's slot from registers and keeps
// the largest position whose label has count C: that position's label is
// argmax (count, last).
template <int T, int EPT, int CAPBITS, typename E>
__global__ void __launch_bounds__(T)
k_hash_lds(const E *__restrict__ off, const int32_t *__restrict__ nbr,
           const int32_t *__restrict__ prev, int32_t *__restrict__ next,
           const int32_t *__restrict__ vlist, int64_t v0, int64_t nv,
           unsigned long long *__restrict__ counter, const int32_t *__restrict__ stop,
           int32_t *__restrict__ err) {
    constexpr int CAP = 1 << CAPBITS;
    __shared__ uint32_t skey[CAP];
    __shared__ uint32_t scnt[CAP];
    __shared__ uint32_t sred[T / kWave];
    __shared__ unsigned long long sred64[T / kWave];
    if (*stop) return;
    uint32_t changed = 0;
    for (int64_t i = blockIdx.x; i < nv; i += gridDim.x) {
        const int64_t v = vlist ? (int64_t)vlist[v0 + i] : v0 + i;
        const E b = off[v], e = off[v + 1];
        const int deg = (int)(e - b);
        const int bits = max(6, ceil_log2_dev((uint32_t)(2 * deg)));
        if (bits > CAPBITS || deg > T * EPT) {   // outside this class: never planned here
            if (threadIdx.x == 0) atomicOr(err, 8);
            continue;
        }
        const uint32_t mask = (1u << bits) - 1u;
        for (int s = threadIdx.x; s <= (int)mask; s += T) {
            skey[s] = kEmpty;
            scnt[s] = 0;
        }
        __syncthreads();
        int32_t lab[EPT];
        uint32_t slot[EPT];
        uint32_t maxc = 0;
#pragma unroll
        for (int k = 0; k < EPT; ++k) {
            const int pos = k * T + (int)threadIdx.x;
            lab[k] = -1;
            slot[k] = 0;
            if (pos < deg) {
                const uint32_t L = (uint32_t)prev[nbr[b + pos]];
                lab[k] = (int32_t)L;
                uint32_t h = hash_slot(L, bits);
                for (uint32_t probe = 0;; ++probe) {
                    if (probe > mask) {   // impossible: at most deg keys in 2*deg slots
                        atomicOr(err, 1);
                        break;
                    }
                    const uint32_t cur = skey[h];
                    if (cur == L) break;
                    if (cur == kEmpty) {
                        const uint32_t old = atomicCAS(&skey[h], kEmpty, L);
                        if (old == kEmpty || old == L) break;
                    }
                    h = (h + 1) & mask;
                }
                slot[k] = h;
                const uint32_t c = atomicAdd(&scnt[h], 1u) + 1u;
                maxc = max(maxc, c);
            }
        }
        __syncthreads();
        maxc = block_max(maxc, sred);
        unsigned long long best = 0;
#pragma unroll
        for (int k = 0; k < EPT; ++k) {
            const int pos = k * T + (int)threadIdx.x;
            if (pos < deg && scnt[slot[k]] == maxc)
                best = ((unsigned long long)(pos + 1) << 32) | (uint32_t)lab[k];
        }
        best = block_max(best, sred64);
        if (threadIdx.x == 0 && deg > 0) {
            const int32_t w = (int32_t)(uint32_t)best;
            const int32_t old = prev[v];
            next[v] = w;
            changed += (w != old);
        }
        __syncthreads();
    }
    block_count_add(changed, counter, sred);
}

// ---------------------------------------------------------------------------
// Hubs: tile pass.  A block aggregates one tile of <= kHubTile edges in an
// LDS table of (key, count, last position in the row), then merges each
// occupied slot into the hub's global table with device-scope atomics.
template <int T, typename E>
__global__ void __launch_bounds__(T)
k_hub_tile(const E *__restrict__ off, const int32_t *__restrict__ nbr,
           const int32_t *__restrict__ prev, const HubTile *__restrict__ tiles, int64_t ntiles,
           const int64_t *__restrict__ hub_toff, const int32_t *__restrict__ hub_tbits,
           uint32_t *__restrict__ gkey, uint32_t *__restrict__ gcnt, uint32_t *__restrict__ glast,
           const int32_t *__restrict__ stop, int32_t *__restrict__ err) {
    constexpr int CAP = 1 << kHubTileCapBits;
    constexpr int EPT = kHubTile / T;
    __shared__ uint32_t skey[CAP];
    __shared__ uint32_t scnt[CAP];
    __shared__ uint32_t slast[CAP];
    if (*stop) return;
    for (int64_t t = blockIdx.x; t < ntiles; t += gridDim.x) {
        const HubTile tl = tiles[t];
        const int64_t rowb = (int64_t)off[tl.vtx];
        const int len = (int)(tl.e1 - tl.e0);
        const int bits = max(6, ceil_log2_dev((uint32_t)(2 * len)));
        if (bits > kHubTileCapBits || len > kHubTile) {
            if (threadIdx.x == 0) atomicOr(err, 16);
            continue;
        }
        const uint32_t mask = (1u << bits) - 1u;
        for (int s = threadIdx.x; s <= (int)mask; s += T) {
            skey[s] = kEmpty;
            scnt[s] = 0;
            slast[s] = 0;
        }
        __syncthreads();
#pragma unroll
        for (int k = 0; k < EPT; ++k) {
            const int pos = k * T + (int)threadIdx.x;
            if (pos < len) {
                const int64_t ea = tl.e0 + pos;
                const uint32_t L = (uint32_t)prev[nbr[ea]];
                uint32_t h = hash_slot(L, bits);
                for (uint32_t probe = 0;; ++probe) {
                    if (probe > mask) {
                        atomicOr(err, 2);
                        break;
                    }
                    const uint32_t cur = skey[h];
                    if (cur == L) break;
                    if (cur == kEmpty) {
                        const uint32_t old = atomicCAS(&skey[h], kEmpty, L);
                        if (old == kEmpty || old == L) break;
                    }
                    h = (h + 1) & mask;
                }
                atomicAdd(&scnt[h], 1u);
                atomicMax(&slast[h], (uint32_t)(ea - rowb));
            }
        }
        __syncthreads();
        const int64_t gb = hub_toff[tl.hub];
        const int gbits = hub_tbits[tl.hub];
        const uint32_t gmask = (1u << gbits) - 1u;
        for (int s = threadIdx.x; s <= (int)mask; s += T) {
            const uint32_t L = skey[s];
            if (L == kEmpty) continue;
            uint32_t h = hash_slot(L, gbits);
            for (uint32_t probe = 0;; ++probe) {
                if (probe > gmask) {
                    atomicOr(err, 4);
                    break;
                }
                const uint32_t cur = load_coherent(&gkey[gb + h]);
                if (cur == L) break;
                if (cur == kEmpty) {
                    const uint32_t old = atomicCAS(&gkey[gb + h], kEmpty, L);
                    if (old == kEmpty || old == L) break;
                }
                h = (h + 1) & gmask;
            }
            atomicAdd(&gcnt[gb + h], scnt[s]);
            atomicMax(&glast[gb + h], slast[s]);
        }
        __syncthreads();
    }
}

// Hubs: finalize.  One block per hub scans its global table for the max of
// (count << 32 | last), writes the winning key, and clears the table.
template <int T>
__global__ void __launch_bounds__(T)
k_hub_final(const int32_t *__restrict__ hub_vid, int64_t nhubs,
            const int64_t *__restrict__ hub_toff, const int32_t *__restrict__ hub_tbits,
            uint32_t *__restrict__ gkey, uint32_t *__restrict__ gcnt, uint32_t *__restrict__ glast,
            const int32_t *__restrict__ prev, int32_t *__restrict__ next,
            unsigned long long *__restrict__ counter, const int32_t *__restrict__ stop) {
    __shared__ unsigned long long sred64[T / kWave];
    __shared__ uint32_t sred[T / kWave];
    __shared__ uint32_t swin;
    if (*stop) return;
    uint32_t changed = 0;
    for (int64_t h = blockIdx.x; h < nhubs; h += gridDim.x) {
        const int64_t gb = hub_toff[h];
        const int64_t cap = (int64_t)1 << hub_tbits[h];
        unsigned long long best = 0;
        uint32_t bestkey = 0;
        for (int64_t s = threadIdx.x; s < cap; s += T) {
            const uint32_t k = gkey[gb + s];
            if (k != kEmpty) {
                const unsigned long long pk =
                    ((unsigned long long)gcnt[gb + s] << 32) | (unsigned long long)glast[gb + s];
                if (pk > best) {
                    best = pk;
                    bestkey = k;
                }
                gkey[gb + s] = kEmpty;
                gcnt[gb + s] = 0;
                glast[gb + s] = 0;
            }
        }
        const unsigned long long top = block_max(best, sred64);
        if (best == top && top != 0) swin = bestkey;   // (count, last) is unique per label
        __syncthreads();
        if (threadIdx.x == 0) {
            const int32_t v = hub_vid[h];
            const int32_t w = (int32_t)swin;
            const int32_t old = prev[v];
            next[v] = w;
            changed += (w != old);
        }
        __syncthreads();
    }
    block_count_add(changed, counter, sred);
}

__global__ void k_iter_end(unsigned long long *counter, int32_t *stop, long long *iter_counts,
                           int32_t *iters_done, int it, double thr) {
    if (*stop) return;
    const unsigned long long c = *counter;
    iter_counts[it] = (long long)c;
    *iters_done = it + 1;
    *counter = 0ull;
    if ((double)c <= thr) *stop = 1;
}

// ---------------------------------------------------------------------------
// Host launchers.

static int grid_for(const galp_graph *g, int64_t work_items, int blocks_per_cu) {
    const int64_t cap = (int64_t)g->ctx->prop.multiProcessorCount * blocks_per_cu;
    int64_t b = work_items < cap ? work_items : cap;
    return (int)(b < 1 ? 1 : b);
}

template <typename E>
static int hash_iteration(galp_graph *g) {
    const E *off = (const E *)g->d_off;
    const int32_t *prev = g->d_lab[g->cur];
    int32_t *next = g->d_lab[g->cur ^ 1];
    hipStream_t s = g->stream;
    PendingLaunch pl;
    auto bytes = [&](const BucketRange &r) {
        // offsets (two reads per vertex, one line-shared), neighbours, gathered
        // labels, own previous label, next label
        return (double)sizeof(E) * (double)(r.nv + 1) + 8.0 * (double)r.edges + 8.0 * (double)r.nv;
    };
#define GALP_GROUP(G, BID)                                                                      \
    {                                                                                           \
        const BucketRange &r = g->bucket[BID];                                                  \
        if (r.nv) {                                                                             \
            prof_begin(g, "hash_group" #G, bytes(r), &pl);                                      \
            const int64_t per_block = 256 / G;                                                  \
            k_hash_group<G, E><<<grid_for(g, (r.nv + per_block - 1) / per_block, 16), 256, 0, s>>>( \
                off, g->d_nbr, prev, next, g->d_vlist, r.v0, r.nv, g->d_counter, g->d_stop);    \
            prof_end(g, &pl);                                                                   \
        }                                                                                       \
    }
    GALP_GROUP(4, B_G4)
    GALP_GROUP(8, B_G8)
    GALP_GROUP(16, B_G16)
    GALP_GROUP(32, B_G32)
    GALP_GROUP(64, B_G64)
#undef GALP_GROUP
    {
        const BucketRange &r = g->bucket[B_WAVE];
        if (r.nv) {
            prof_begin(g, "hash_wave", bytes(r), &pl);
            k_hash_lds<64, 8, 10, E><<<grid_for(g, r.nv, 20), 64, 0, s>>>(
                off, g->d_nbr, prev, next, g->d_vlist, r.v0, r.nv, g->d_counter, g->d_stop, g->d_err);
            prof_end(g, &pl);
        }
    }
    {
        const BucketRange &r = g->bucket[B_BLOCK];
        if (r.nv) {
            prof_begin(g, "hash_block", bytes(r), &pl);
            k_hash_lds<256, 16, 13, E><<<grid_for(g, r.nv, 2), 256, 0, s>>>(
                off, g->d_nbr, prev, next, g->d_vlist, r.v0, r.nv, g->d_counter, g->d_stop, g->d_err);
            prof_end(g, &pl);
        }
    }
    if (g->nhubs) {
        const BucketRange &r = g->bucket[B_HUB];
        prof_begin(g, "hash_hub_tile", 4.0 * (double)r.edges * 2.0, &pl);
        k_hub_tile<256, E><<<grid_for(g, g->ntiles, 2), 256, 0, s>>>(
            off, g->d_nbr, prev, g->d_tiles, g->ntiles, g->d_hub_toff, g->d_hub_tbits, g->d_gkey,
            g->d_gcnt, g->d_glast, g->d_stop, g->d_err);
        prof_end(g, &pl);
        prof_begin(g, "hash_hub_final", 16.0 * (double)r.nv, &pl);
        k_hub_final<1024><<<grid_for(g, g->nhubs, 2), 1024, 0, s>>>(
            g->d_hub_vid, g->nhubs, g->d_hub_toff, g->d_hub_tbits, g->d_gkey, g->d_gcnt, g->d_glast,
            prev, next, g->d_counter, g->d_stop);
        prof_end(g, &pl);
    }
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return fail(g->ctx, GALP_EHIP, std::string("hash kernel launch: ") + hipGetErrorString(e));
    return GALP_OK;
}

int launch_hash_iteration(galp_graph *g) {
    return g->off64 ? hash_iteration<int64_t>(g) : hash_iteration<int32_t>(g);
}

int launch_iter_end(galp_graph *g) {
    k_iter_end<<<1, 1, 0, g->stream>>>(g->d_counter, g->d_stop, g->d_iter_counts, g->d_iters_done,
                                       g->iter % kMaxIters, (double)g->n * 1e-5);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return fail(g->ctx, GALP_EHIP, std::string("iter_end launch: ") + hipGetErrorString(e));
    return GALP_OK;
}

}  // namespace galp
