// learned.hip -- the learned dispatcher of the hash method.
//
// Each iteration, a dispatch kernel runs the MLP (mlp.hpp) on every vertex
// of degree > 64 ("big" vertices) and routes it:
//   - predicted distinct labels <= cand_thresh: the candidate path -- one
//     wave per tile of <= kCandTile edges keeps up to kCandK (label, count,
//     last position) candidates in the lanes of three registers and peels
//     each 64-edge chunk label by label with ballots: no LDS table, so the
//     workgroup's LDS holds
//     the kCandCache hottest labels (the lowest ids after the degree-
//     descending renumbering) and gathers that hit them skip L2;
//   - otherwise: the plain degree-class kernels of lp_hash.hip, through
//     per-class vertex lists built here.
// A tile that meets more than kCandK distinct labels sends its vertex to the
// plain class list in the same iteration (hub tiles: the hub to the plain
// hub path), so a misprediction costs time, never labels.  Every path
// records the number of distinct labels it saw, the network's fourth input
// for the next iteration.  The prediction also picks kernel variants: rows
// predicted high-distinct (more than degree / kHdDiv labels) go to lists
// whose table kernels insert lane by lane, and the lane-group kernels switch
// to their sweep-only resolver when most big vertices are high-distinct.
// Out of core, each batch runs its own dispatch and candidate path over its
// rebased rows (launch_learned_front_batch); rows of several candidate tiles
// stay on the class and hub paths there.
// No reference counterpart: the reference picks a kernel by -l only
// (lfht.cuh:127-165).

#include <algorithm>
#include <type_traits>
#include <cstring>
#include <vector>

#include "device_common.hpp"
#include "internal.hpp"
#include "mlp.hpp"
#include "pipeline.hpp"
#include "partsort.hpp"

namespace galp {

// log2-scaled network inputs from the raw features and the previous count
__device__ __forceinline__ void mlp_inputs(float deg, float mean, float mx, int32_t prevd, float *x) {
    x[0] = __log2f(1.f + deg);
    x[1] = __log2f(1.f + mean);
    x[2] = __log2f(1.f + mx);
    x[3] = prevd >= 0 ? __log2f(1.f + (float)prevd) : -1.f;
}


// ---------------------------------------------------------------------------
// Static features: degree, mean and max neighbour degree (one wave per big
// vertex, once at upload).  mean = (float)((double)sum / deg), the exact
// integer sum rounded once.  offrow locates the rows in nbr (an out-of-core
// batch's rebased offsets), off gives every vertex's degree.
template <typename E>
__global__ void __launch_bounds__(256)
k_features(const E *__restrict__ offrow, const int32_t *__restrict__ nbr, const E *__restrict__ off,
           const int32_t *__restrict__ vlist, int64_t v0, int64_t nv, float *__restrict__ fdeg,
           float *__restrict__ fmean, float *__restrict__ fmax) {
    const int64_t wave = ((int64_t)blockIdx.x * 256 + threadIdx.x) / kWave;
    const int64_t nw = (int64_t)gridDim.x * (256 / kWave);
    for (int64_t i = wave; i < nv; i += nw) {
        const int64_t v = vlist ? (int64_t)vlist[v0 + i] : v0 + i;
        const int64_t b = (int64_t)offrow[v], d = (int64_t)offrow[v + 1] - b;
        unsigned long long sum = 0, mx = 0;
        for (int64_t j = lane_id(); j < d; j += kWave) {
            const int64_t u = nbr[b + j];
            const unsigned long long du = (unsigned long long)((int64_t)off[u + 1] - (int64_t)off[u]);
            sum += du;
            mx = du > mx ? du : mx;
        }
#pragma unroll
        for (int s = 32; s > 0; s >>= 1) {
            sum += __shfl_xor(sum, s);
            const unsigned long long o = __shfl_xor(mx, s);
            mx = o > mx ? o : mx;
        }
        if (lane_id() == 0) {
            fdeg[v] = (float)d;
            fmean[v] = d ? (float)((double)sum / (double)d) : 0.f;
            fmax[v] = (float)mx;
        }
    }
}

// ---------------------------------------------------------------------------
// Dispatch.  Wave-aggregated appends: one atomic per wave and list.
__device__ __forceinline__ int64_t wave_append(int64_t *ctr, bool want) {
    const unsigned long long ball = __ballot(want);
    if (!ball) return -1;
    const int leader = __ffsll((long long)ball) - 1;
    unsigned long long base = 0;
    if (lane_id() == leader) base = atomicAdd((unsigned long long *)ctr, (unsigned long long)__popcll(ball));
    base = __shfl(base, leader);
    return want ? (int64_t)base + __popcll(ball & ((1ull << lane_id()) - 1ull)) : -1;
}

// Zeroes ctr[0, count); count <= blockDim.x.  Every caller passes the
// number of words its buffer holds.
static_assert(kNumCtrs <= 64, "k_reset_ctrs runs one 64-thread block");
__global__ void k_reset_ctrs(int64_t *ctr, int count) {
    if ((int)threadIdx.x < count) ctr[threadIdx.x] = 0;
}

// Clears the dispatch's lists and counters.  first == 0 (out-of-core batches
// after the first of an iteration): the iteration's statistics -- overflows,
// vertices routed to candidates, candidate-path bytes -- keep adding up.
__global__ void k_reset_front(int64_t *ctr, int64_t *cand_bytes, int first) {
    const int c = threadIdx.x;
    if (c < kNumCtrs && (first || (c != kCtrOverflow && c != kCtrCandVerts))) ctr[c] = 0;
    if (c == 0 && first) *cand_bytes = 0;
}

// One block routes kDispChunk consecutive big vertices.  Appends are
// aggregated per block: the block counts its entries per list in LDS, takes
// one global atomicAdd per list, then hands out positions from LDS cursors.
// (Same-address global atomics serialise at the L2; one per wave and list
// cost ~100 us per iteration on RMAT-22, one per block and list does not.)
constexpr int kDispChunk = 256;   // one vertex per thread
// codes: [0, NUM_BUCKETS) peeled class list, [NUM_BUCKETS, 2 NUM_BUCKETS)
// high-distinct class list, then candidate tile, multi-tile candidate, hub
enum : uint8_t { kCodeCand = 2 * NUM_BUCKETS, kCodeCandMt, kCodeHub };
// LDS list slots: [0, 2 NUM_BUCKETS) class lists, then kCandClasses tile
// classes, candidate vertices, plain hubs, multi-tile candidates
constexpr int kSlotTiles = 2 * NUM_BUCKETS;
constexpr int kSlotCandV = kSlotTiles + kCandClasses;
constexpr int kSlotHubs = kSlotCandV + 1;
constexpr int kSlotMt = kSlotHubs + 1;
constexpr int kSlotHdAll = kSlotMt + 1;
constexpr int kSlots = kSlotHdAll + 1;

template <int H, typename E>
__global__ void __launch_bounds__(256)
k_dispatch(const E *__restrict__ off, const int32_t *__restrict__ vlist, int64_t big_v0, int64_t big_n,
           bool allow_mt, const float *__restrict__ params, const float *__restrict__ fdeg,
           const float *__restrict__ fmean, const float *__restrict__ fmax, const int32_t *__restrict__ dprev,
           float cand_thresh, int64_t nhubs,
           const int32_t *__restrict__ mt_of, int32_t *__restrict__ mt_cand, int32_t *__restrict__ hub_sel,
           int32_t *__restrict__ cw_v, int64_t *__restrict__ cw_e0, int32_t *__restrict__ cw_len,
           int32_t *__restrict__ cw_pslot, const int64_t *__restrict__ cw_base, int32_t *const *__restrict__ plist,
           int32_t *__restrict__ hub_hd, int64_t *__restrict__ ctr, int64_t *__restrict__ cand_bytes,
           const int32_t *__restrict__ stop) {
    __shared__ uint8_t scode[kDispChunk];
    __shared__ int32_t sv[kDispChunk];
    __shared__ int64_t sb[kDispChunk], sd[kDispChunk];
    __shared__ unsigned long long scount[kSlots];
    __shared__ unsigned long long sbytes;
    if (*stop) return;
    if (threadIdx.x < kSlots) scount[threadIdx.x] = 0;
    if (threadIdx.x == 0) sbytes = 0;
    __syncthreads();
    const int64_t i0 = (int64_t)blockIdx.x * kDispChunk;
    const int nloc = (int)min((int64_t)kDispChunk, big_n - i0);
    // phase 1: predict, classify, count
    for (int k = threadIdx.x; k < nloc; k += 256) {
        const int64_t i = i0 + k;
        const int64_t v = vlist ? (int64_t)vlist[big_v0 + i] : big_v0 + i;
        const int64_t b = (int64_t)off[v];
        const int64_t d = (int64_t)off[v + 1] - b;
        float x[4];
        mlp_inputs(fdeg[v], fmean[v], fmax[v], dprev[v], x);
        // weights through the scalar cache: wave-uniform addresses become
        // scalar loads and SGPR operands of the FMAs (no LDS traffic)
        const float y = mlp_forward<H>(params, x[0], x[1], x[2], x[3]);
        const float dist = exp2f(y) - 1.f;            // predicted distinct labels
        // rows of several tiles use the fixed tile list, which holds in-core
        // edge offsets: out of core they stay on the class and hub paths
        const bool cand = dist <= cand_thresh && (allow_mt || d <= kCandTile);
        const bool hd = dist * (float)kHdDiv > (float)d;
        const int cls = bucket_of_deg(d);
        const int list = hd ? NUM_BUCKETS + cls : cls;
        const int mt = d > kCandTile ? mt_of[v] : -1;
        uint8_t code;
        if (mt >= 0) {
            mt_cand[mt] = cand ? 1 : 0;
            if (mt < nhubs) {
                hub_sel[mt] = cand ? 0 : 1;
                hub_hd[mt] = (!cand && hd) ? 1 : 0;
                if (!cand) atomicAdd(&scount[kSlotHubs], 1ull);
                // the hub class has no list of its own: its high-distinct
                // slot counts the hubs of the hub path's high-distinct launch
                if (!cand && hd) atomicAdd(&scount[NUM_BUCKETS + B_HUB], 1ull);
            }
            code = cand ? (uint8_t)kCodeCandMt : cls < B_HUB ? (uint8_t)list : (uint8_t)kCodeHub;
        } else {
            code = cand ? (uint8_t)kCodeCand : (uint8_t)list;
        }
        if (code < 2 * NUM_BUCKETS) atomicAdd(&scount[code], 1ull);
        if (hd) atomicAdd(&scount[kSlotHdAll], 1ull);
        if (code == kCodeCand) atomicAdd(&scount[kSlotTiles + cand_class(d)], 1ull);
        if (code == kCodeCandMt) atomicAdd(&scount[kSlotMt], 1ull);
        if (cand) {
            atomicAdd(&scount[kSlotCandV], 1ull);
            // neighbour ids and gathered labels, offsets, own old and new label
            atomicAdd(&sbytes, (unsigned long long)(8 * d + 2 * (int64_t)sizeof(E) + 8));
        }
        scode[k] = code;
        sv[k] = (int32_t)v;
        sb[k] = b;
        sd[k] = d;
    }
    __syncthreads();
    // phase 2: one global atomic per non-empty list; LDS cursors start at its base
    if (threadIdx.x < kSlots) {
        const int c = threadIdx.x;
        const int gc = c < NUM_BUCKETS ? c
                       : c < 2 * NUM_BUCKETS ? kCtrHd + (c - NUM_BUCKETS)
                       : c < kSlotCandV ? kCtrCandTiles + (c - kSlotTiles)
                       : c == kSlotCandV ? kCtrCandVerts : c == kSlotHubs ? kCtrPlainHubs
                       : c == kSlotMt ? kCtrCandMt : kCtrHdAll;
        const unsigned long long cnt = scount[c];
        scount[c] = cnt ? atomicAdd((unsigned long long *)(ctr + gc), cnt) : 0ull;
        if (c >= kSlotTiles && c < kSlotCandV) scount[c] += (unsigned long long)cw_base[c - kSlotTiles];
    }
    if (threadIdx.x == 64 && sbytes) atomicAdd((unsigned long long *)cand_bytes, sbytes);
    __syncthreads();
    // phase 3: write the entries (the tiles of multi-tile vertices are a fixed
    // list made at set-up; the candidate kernel skips those not selected)
    for (int k = threadIdx.x; k < nloc; k += 256) {
        const uint8_t code = scode[k];
        const int32_t v = sv[k];
        if (code < 2 * NUM_BUCKETS) {
            const unsigned long long q = atomicAdd(&scount[code], 1ull);
            plist[code][q] = v;
        } else if (code == kCodeCand) {
            const unsigned long long q = atomicAdd(&scount[kSlotTiles + cand_class(sd[k])], 1ull);
            cw_v[q] = v;
            cw_e0[q] = sb[k];
            cw_len[q] = (int32_t)sd[k];
            cw_pslot[q] = -1;
        }
    }
}

// ---------------------------------------------------------------------------
// Candidate path.  One wave per tile (pipelined over the wave's tiles);
// candidates live in wave-uniform registers.  POS: the tile's edges come
// from a part-sorted copy with their row positions alongside (the fixed
// tiles of multi-tile rows in core), so a label's last position is the max
// over its lanes instead of its highest lane's.
template <int EPT, typename E, bool POS = false>
struct CandPolicy {
    struct Item {
        bool valid;
        int len;
        int64_t e0;
        int32_t v;
        int32_t pslot;
        uint32_t row0;    // position of the tile's first edge in the row
        int32_t old;
    };
    const E *off;
    const int32_t *prev;
    int32_t *next;
    const int32_t *cw_v, *cw_len, *cw_pslot;
    const int64_t *cw_e0;
    int64_t ntiles;
    uint32_t *pkey, *pcnt, *plast;
    int32_t *pn;
    int32_t *dist_out;
    int32_t *const *plist;
    int64_t *ctr;
    const int32_t *cache;   // LDS copy of labels [0, hot)
    int32_t hot;
    const int32_t *tile_mt;   // fixed multi-tile list: tile -> multi-tile vertex (else null)
    const int32_t *mt_sel;    // multi-tile vertex routed to candidates this iteration
    uint32_t changed;

    __device__ Item item(int64_t t) const {
        Item it{false, 0, 0, 0, -1, 0, 0};
        if (t < ntiles) {
            it.v = cw_v[t];
            it.e0 = cw_e0[t];
            it.len = cw_len[t];
            it.pslot = cw_pslot[t];
            if constexpr (!POS) it.row0 = (uint32_t)(it.e0 - (int64_t)off[it.v]);
            it.old = prev[it.v];
            it.valid = true;
            if (tile_mt && !mt_sel[tile_mt[t]]) it.len = 0;   // left to the class kernels
        }
        return it;
    }
    __device__ int64_t edge0(const Item &it) const { return it.e0; }

    // lab holds prev[u] for cold neighbours; hot ones are re-read from LDS
    // (the pipeline's gather still issued a load, to prev[0] for them, which
    // all such lanes share, so it costs one line instead of one per lane)
    // Candidates live in lanes: lane j < ncand holds candidate j's label,
    // count and last position.  Each 64-edge chunk is peeled: one round per
    // distinct label present in the chunk (ballot of the equal lanes gives
    // its count and last position), looked up among the candidates with one
    // more ballot.  A converged row costs one round per chunk.
    // hid: per element the neighbour id when it is hot (< hot, so 16 bits),
    // else 0xffff; two elements per word (element k in half k & 1 of word
    // k / 2), which keeps the kernel within 128 VGPRs at 1024 threads.
    static constexpr int HW = (EPT + 1) / 2;
    __device__ void process(const Item &it, const int32_t (&lab)[EPT], const uint32_t (&hid)[HW]) {
        uint32_t none[EPT];
        process(it, lab, hid, none);
    }
    // rpos: the row positions of the item's elements (POS only)
    __device__ void process(const Item &it, const int32_t (&lab)[EPT], const uint32_t (&hid)[HW],
                            const uint32_t (&rpos)[EPT]) {
        if (it.len == 0) return;
        const int lane = lane_id();
        int32_t ckv = -1;          // candidate label of this lane (-1: none)
        uint32_t ccv = 0, clv = 0;
        int ncand = 0;             // uniform
        bool overflow = false;     // uniform
#pragma unroll
        for (int k = 0; k < EPT; ++k) {
            if (k * kWave >= it.len || overflow) break;   // uniform
            const int pos = k * kWave + lane;
            const uint32_t h = (hid[k >> 1] >> (16 * (k & 1))) & 0xffffu;
            const int32_t L = pos < it.len ? (h != 0xffffu ? cache[h] : lab[k]) : -1;
            const uint32_t rowpos = it.row0 + (uint32_t)(k * kWave);
            unsigned long long ball = __builtin_amdgcn_ballot_w64(L >= 0);
            while (ball) {
                const int32_t Lk = __builtin_amdgcn_readlane(L, __ffsll((long long)ball) - 1);
                const unsigned long long m = __builtin_amdgcn_ballot_w64(L == Lk);
                ball &= ~m;
                const uint32_t c = (uint32_t)__popcll(m);
                uint32_t lastp;
                if constexpr (POS) lastp = wave_max(L == Lk ? rpos[k] : 0u);
                else lastp = rowpos + (uint32_t)(63 - __clzll((long long)m));
                const unsigned long long hit = __builtin_amdgcn_ballot_w64(ckv == Lk);
                if (hit) {
                    if (lane == __ffsll((long long)hit) - 1) {
                        ccv += c;
                        clv = lastp;
                    }
                } else {
                    if (ncand == kCandK) {
                        overflow = true;
                        break;
                    }
                    if (lane == ncand) {
                        ckv = Lk;
                        ccv = c;
                        clv = lastp;
                    }
                    ++ncand;
                }
            }
        }
        if (it.pslot >= 0) {   // tile of a multi-tile vertex: partial candidates
            if (lane == 0) pn[it.pslot] = overflow ? -1 : ncand;
            if (!overflow && lane < ncand) {
                const int64_t e = (int64_t)it.pslot * kCandK + lane;
                pkey[e] = (uint32_t)ckv;
                pcnt[e] = ccv;
                plast[e] = clv;
            }
            return;
        }
        if (overflow) {   // the plain class kernel of this degree takes the vertex
            if (lane == 0) {
                const int c = bucket_of_deg(it.len);
                const unsigned long long q = atomicAdd((unsigned long long *)(ctr + c), 1ull);
                plist[c][q] = it.v;
                atomicAdd((unsigned long long *)(ctr + kCtrOverflow), 1ull);
            }
            return;
        }
        const unsigned long long val = lane < ncand ? (((unsigned long long)ccv << 32) | clv) : 0ull;
        const unsigned long long top = wave_max(val);
        const unsigned long long wl = __builtin_amdgcn_ballot_w64(lane < ncand && val == top);
        const int32_t w = __builtin_amdgcn_readlane(ckv, __ffsll((long long)wl) - 1);
        if (lane == 0) {
            next[it.v] = w;
            changed += (w != it.old);
            if (dist_out) dist_out[it.v] = ncand;
        }
    }
};

// The candidate path's own pipeline: like pipelined(), plus the neighbour
// ids of each item are kept for the LDS label cache.
template <int EPT, class Policy>
__device__ __forceinline__ void cand_pipeline(Policy &pol, const int32_t *__restrict__ nbr,
                                              const int32_t *__restrict__ prev, int64_t first, int64_t stride,
                                              int32_t hot) {
    using Item = typename Policy::Item;
    const int lane = lane_id();
    int64_t i = first;
    auto load_nbr = [&](int32_t (&u)[EPT], const Item &it) {
        const int64_t e0 = it.valid ? it.e0 : 0;
#pragma unroll
        for (int k = 0; k < EPT; ++k) {
            const int pos = k * kWave + lane;
            const bool ok = it.valid && pos < it.len;
            u[k] = __builtin_nontemporal_load(&nbr[ok ? e0 + pos : 0]);
        }
    };
    auto gather = [&](int32_t (&lab)[EPT], uint32_t (&ids)[(EPT + 1) / 2], const int32_t (&u)[EPT],
                      const Item &it) {
#pragma unroll
        for (int k = 0; k < EPT; ++k) {
            const int pos = k * kWave + lane;
            const bool ok = it.valid && pos < it.len;
            const uint32_t h = ok && u[k] < hot ? (uint32_t)u[k] : 0xffffu;
            if (k & 1) ids[k >> 1] |= h << 16;
            else ids[k >> 1] = h;
            lab[k] = prev[(ok && u[k] >= hot) ? u[k] : 0];
        }
    };
    Item s0 = pol.item(i), s1 = pol.item(i + stride), s2 = pol.item(i + 2 * stride);
    int32_t u[EPT], labA[EPT], labB[EPT];
    uint32_t idA[(EPT + 1) / 2], idB[(EPT + 1) / 2];   // hot ids, two per word
    load_nbr(u, s0);
    gather(labA, idA, u, s0);
    load_nbr(u, s1);
    while (s0.valid) {
        gather(labB, idB, u, s1);
        load_nbr(u, s2);
        Item s3 = pol.item(i + 3 * stride);
        pol.process(s0, labA, idA);
        s0 = s1; s1 = s2; s2 = s3; i += stride;
        if (!s0.valid) break;
        gather(labA, idA, u, s1);
        load_nbr(u, s2);
        s3 = pol.item(i + 3 * stride);
        pol.process(s0, labB, idB);
        s0 = s1; s1 = s2; s2 = s3; i += stride;
    }
}

// cand_pipeline for POS tiles: the row positions (pos_src) are loaded with
// each item's labels, into the same ping-pong slot.
template <int EPT, class Policy>
__device__ __forceinline__ void cand_pipeline_pos(Policy &pol, const int32_t *__restrict__ nbr,
                                                  const uint32_t *__restrict__ pos_src,
                                                  const int32_t *__restrict__ prev, int64_t first, int64_t stride,
                                                  int32_t hot) {
    using Item = typename Policy::Item;
    const int lane = lane_id();
    int64_t i = first;
    auto load_nbr = [&](int32_t (&u)[EPT], const Item &it) {
        const int64_t e0 = it.valid ? it.e0 : 0;
#pragma unroll
        for (int k = 0; k < EPT; ++k) {
            const int pos = k * kWave + lane;
            const bool ok = it.valid && pos < it.len;
            u[k] = __builtin_nontemporal_load(&nbr[ok ? e0 + pos : 0]);
        }
    };
    auto gather = [&](int32_t (&lab)[EPT], uint32_t (&ids)[(EPT + 1) / 2], uint32_t (&ps)[EPT], const int32_t (&u)[EPT],
                      const Item &it) {
        const int64_t e0 = it.valid ? it.e0 : 0;
#pragma unroll
        for (int k = 0; k < EPT; ++k) {
            const int pos = k * kWave + lane;
            const bool ok = it.valid && pos < it.len;
            const uint32_t h = ok && u[k] < hot ? (uint32_t)u[k] : 0xffffu;
            if (k & 1) ids[k >> 1] |= h << 16;
            else ids[k >> 1] = h;
            lab[k] = prev[(ok && u[k] >= hot) ? u[k] : 0];
            ps[k] = __builtin_nontemporal_load(&pos_src[ok ? e0 + pos : 0]);
        }
    };
    Item s0 = pol.item(i), s1 = pol.item(i + stride), s2 = pol.item(i + 2 * stride);
    int32_t u[EPT], labA[EPT], labB[EPT];
    uint32_t idA[(EPT + 1) / 2], idB[(EPT + 1) / 2];   // hot ids, two per word
    uint32_t pA[EPT], pB[EPT];
    load_nbr(u, s0);
    gather(labA, idA, pA, u, s0);
    load_nbr(u, s1);
    while (s0.valid) {
        gather(labB, idB, pB, u, s1);
        load_nbr(u, s2);
        Item s3 = pol.item(i + 3 * stride);
        pol.process(s0, labA, idA, pA);
        s0 = s1; s1 = s2; s2 = s3; i += stride;
        if (!s0.valid) break;
        gather(labA, idA, pA, u, s1);
        load_nbr(u, s2);
        s3 = pol.item(i + 3 * stride);
        pol.process(s0, labB, idB, pB);
        s0 = s1; s1 = s2; s2 = s3; i += stride;
    }
}

// Candidate work lists, walked in turn by every wave of one launch (one
// LDS cache fill): the fixed tiles of multi-tile vertices, then single-tile
// vertices by length class.
struct CandLists {
    const int32_t *v[kCandClasses + 1], *len[kCandClasses + 1], *pslot[kCandClasses + 1];
    const int64_t *e0[kCandClasses + 1];
    const int32_t *tile_mt;   // fixed list: tile -> multi-tile vertex
    int64_t nfixed;
    // in core, the fixed tiles cover a part-sorted copy of the multi-tile
    // rows (POS tiles of at most kCandPartTile edges); null out of core
    const int32_t *fnbr;
    const uint32_t *fpos;
};

// One block per CU: the LDS cache takes 152 KB of the CU's 160 KB.  1024
// threads (16 waves) hold the kernel to 128 VGPRs, which it fits with the
// hot ids packed two to a register.  Unpacked it needed 142 and spilled 41
// to scratch at 1024 threads; 768 threads (12 waves, no spill) ran as fast
// at RMAT-26 but 17 % slower at RMAT-22 (418 against 357 µs per launch),
// and 512 threads (8 waves) 16 % slower there.
constexpr int kCandThreads = 1024;

template <typename E>
__global__ void __launch_bounds__(kCandThreads)
k_cand(const E *__restrict__ off, const int32_t *__restrict__ nbr, const int32_t *__restrict__ prev,
       int32_t *__restrict__ next, CandLists cl, const int32_t *__restrict__ mt_sel, uint32_t *__restrict__ pkey,
       uint32_t *__restrict__ pcnt, uint32_t *__restrict__ plast, int32_t *__restrict__ pn,
       int32_t *__restrict__ dist_out, int32_t *const *__restrict__ plist, int64_t *__restrict__ ctr, int32_t hot,
       unsigned long long *__restrict__ counter, const int32_t *__restrict__ stop) {
    __shared__ int32_t cache[kCandCache];
    if (*stop) return;
    const int64_t nmt = ctr[kCtrCandMt];
    int64_t nt[kCandClasses];
    int64_t any = nmt;
#pragma unroll
    for (int c = 0; c < kCandClasses; ++c) any += (nt[c] = ctr[kCtrCandTiles + c]);
    if (any == 0) return;
    for (int i = threadIdx.x; i < hot; i += kCandThreads) cache[i] = prev[i];
    __syncthreads();
    constexpr int kW = kCandThreads / kWave;
    const int64_t wave = (int64_t)blockIdx.x * kW + __builtin_amdgcn_readfirstlane((int)(threadIdx.x / kWave));
    const int64_t nwaves = (int64_t)gridDim.x * kW;
    uint32_t changed = 0;
    auto run = [&](auto ept, int li, int64_t ntiles, const int32_t *tmt) {
        constexpr int EPT = decltype(ept)::value;
        if (ntiles == 0) return;
        CandPolicy<EPT, E> pol{off,      prev,      next,  cl.v[li], cl.len[li], cl.pslot[li], cl.e0[li],
                               ntiles,   pkey,      pcnt,  plast,    pn,         dist_out,     plist,
                               ctr,      cache,     hot,   tmt,      mt_sel,     0u};
        cand_pipeline<EPT>(pol, nbr, prev, wave, nwaves, hot);
        changed += pol.changed;
    };
    constexpr int kTileEPT = kCandTile / kWave;   // slots per lane of a full tile
    if (nmt && cl.fpos) {   // fixed POS tiles, in their part-major order
        constexpr int EPT = kCandPartTile / kWave;
        constexpr int li = kCandClasses;
        CandPolicy<EPT, E, true> pol{off,    prev,  next,  cl.v[li], cl.len[li], cl.pslot[li], cl.e0[li],
                                     cl.nfixed, pkey, pcnt, plast,   pn,         dist_out,     plist,
                                     ctr,    cache, hot,   cl.tile_mt, mt_sel,   0u};
        cand_pipeline_pos<EPT>(pol, cl.fnbr, cl.fpos, prev, wave, nwaves, hot);
        changed += pol.changed;
    } else if (nmt) {
        run(std::integral_constant<int, kTileEPT>{}, kCandClasses, cl.nfixed, cl.tile_mt);
    }
    if (kCandTile > 512) run(std::integral_constant<int, kTileEPT>{}, 3, nt[3], nullptr);
    run(std::integral_constant<int, 8>{}, 2, nt[2], nullptr);
    run(std::integral_constant<int, 4>{}, 1, nt[1], nullptr);
    run(std::integral_constant<int, 2>{}, 0, nt[0], nullptr);
    const uint32_t s = wave_sum_u32(changed);
    if (lane_id() == 0 && s) atomicAdd(counter, (unsigned long long)s);
}

constexpr int kMergeU = 8;   // entry batches per load round in k_cand_merge

// Multi-tile vertices on the candidate path: merge the tiles' partial
// candidates, one wave per vertex.  The merged labels live in lanes (lane j
// holds label j), their summed counts and largest positions in the wave's
// LDS slots.  The tiles' entries come 64 at a time: each finds its label
// among the merged ones (a broadcast of each merged label), adds its count
// and positions by LDS atomics; entries whose label is new are then placed
// one at a time (a ballot finds an equal label placed meanwhile).  A tile
// that overflowed, or more than 64 distinct labels in all, hands the vertex
// to the plain path of its class (hubs: the plain hub path), which runs
// after this kernel.  Counts add and positions take the max, so the merged
// (count, last) of every label is exact.
template <typename E>
__global__ void __launch_bounds__(256)
k_cand_merge(const E *__restrict__ off, const int32_t *__restrict__ mt_vid, int64_t nmt, int64_t nhubs,
             const int32_t *__restrict__ mt_cand, int32_t *__restrict__ hub_sel,
             const int32_t *__restrict__ mt_ctile0, const int32_t *__restrict__ mt_nctiles,
             int32_t *const *__restrict__ plist, const uint32_t *__restrict__ pkey, const uint32_t *__restrict__ pcnt,
             const uint32_t *__restrict__ plast, const int32_t *__restrict__ pn, const int32_t *__restrict__ prev,
             int32_t *__restrict__ next, int32_t *__restrict__ dist_out, unsigned long long *__restrict__ counter,
             int64_t *__restrict__ ctr, const int32_t *__restrict__ stop) {
    __shared__ uint32_t s_cnt[256 / kWave][kWave], s_last[256 / kWave][kWave];
    if (*stop || ctr[kCtrCandMt] == 0) return;
    const int lane = lane_id();
    const int wv = (int)(threadIdx.x / kWave);
    uint32_t *scnt = s_cnt[wv], *slast = s_last[wv];
    const int64_t wave = (int64_t)blockIdx.x * (256 / kWave) + wv;
    const int64_t nw = (int64_t)gridDim.x * (256 / kWave);
    uint32_t changed = 0;
    for (int64_t h = wave; h < nmt; h += nw) {
        if (!mt_cand[h]) continue;   // uniform per wave
        const int64_t t0 = mt_ctile0[h], nt = mt_nctiles[h];
        const int64_t eend = (t0 + nt) * kCandK;
        bool bad = false;
        int32_t mkey = -1;   // merged label of this lane (-1: none)
        int nm = 0;          // merged labels so far (uniform)
        scnt[lane] = 0u;
        slast[lane] = 0u;
        __builtin_amdgcn_wave_barrier();   // cleared before any lane adds
        for (int64_t eb = t0 * kCandK; eb < eend && !bad; eb += kMergeU * kWave) {
            // kMergeU batches of 64 entries in flight: one wave walks a
            // hub's thousands of tiles, so it waits on load latency
            int32_t bk[kMergeU];
            uint32_t bc[kMergeU], bl[kMergeU];
            int32_t bn[kMergeU];
#pragma unroll
            for (int u = 0; u < kMergeU; ++u) {
                const int64_t e = eb + u * kWave + lane;
                const int64_t ec = e < eend ? e : eend - 1;
                bn[u] = pn[ec / kCandK];
                bk[u] = (int32_t)pkey[ec];
                bc[u] = pcnt[ec];
                bl[u] = plast[ec];
            }
            bool over = false;   // a tile that overflowed
#pragma unroll
            for (int u = 0; u < kMergeU; ++u) over = over || (eb + u * kWave + lane < eend && bn[u] < 0);
            if (__ballot(over)) {
                bad = true;
                break;
            }
#pragma unroll
            for (int u = 0; u < kMergeU && !bad; ++u) {
                const int64_t e = eb + u * kWave + lane;
                const bool ok = e < eend && (int)(e % kCandK) < bn[u];
                const int32_t ik = ok ? bk[u] : -1;
                const uint32_t ic = bc[u], il = bl[u];
                int slot = -1;
                for (int m = 0; m < nm; ++m) slot = __builtin_amdgcn_readlane(mkey, m) == ik ? m : slot;
                if (ok && slot >= 0) {
                    atomicAdd(&scnt[slot], ic);
                    atomicMax(&slast[slot], il);
                }
                unsigned long long fresh = __ballot(ok && slot < 0);
                while (fresh) {
                    const int j = __ffsll((long long)fresh) - 1;
                    fresh &= fresh - 1ull;
                    const int32_t k = __builtin_amdgcn_readlane(ik, j);
                    const uint32_t c = (uint32_t)__builtin_amdgcn_readlane((int)ic, j);
                    const uint32_t l = (uint32_t)__builtin_amdgcn_readlane((int)il, j);
                    const unsigned long long hit = __ballot(mkey == k);   // placed earlier in this batch
                    if (hit) {
                        if (lane == __ffsll((long long)hit) - 1) {
                            atomicAdd(&scnt[lane], c);
                            atomicMax(&slast[lane], l);
                        }
                    } else if (nm == kWave) {
                        bad = true;   // more labels than lanes
                        break;
                    } else {
                        if (lane == nm) {
                            mkey = k;
                            atomicAdd(&scnt[lane], c);
                            atomicMax(&slast[lane], l);
                        }
                        ++nm;
                    }
                }
            }
        }
        __builtin_amdgcn_wave_barrier();
        const uint32_t mcnt = scnt[lane], mlast = slast[lane];
        __builtin_amdgcn_wave_barrier();   // read before the next vertex clears the slots
        const int64_t v = mt_vid[h];
        if (bad) {
            if (lane == 0) {
                if (h < nhubs) {
                    hub_sel[h] = 1;
                    atomicAdd((unsigned long long *)(ctr + kCtrPlainHubs), 1ull);
                } else {
                    const int c = bucket_of_deg((int64_t)off[v + 1] - (int64_t)off[v]);
                    const unsigned long long q = atomicAdd((unsigned long long *)(ctr + c), 1ull);
                    plist[c][q] = (int32_t)v;
                }
                atomicAdd((unsigned long long *)(ctr + kCtrOverflow), 1ull);
            }
            continue;
        }
        const unsigned long long val = mkey >= 0 ? (((unsigned long long)mcnt << 32) | mlast) : 0ull;
        const unsigned long long top = wave_max(val);
        const unsigned long long wl = __ballot(mkey >= 0 && val == top);
        const int32_t w = __builtin_amdgcn_readlane(mkey, __ffsll((long long)wl) - 1);
        if (lane == 0 && wl) {
            next[v] = w;
            changed += (w != prev[v]);
            if (dist_out) dist_out[v] = nm;
        }
    }
    if (lane == 0 && changed) atomicAdd(counter, (unsigned long long)changed);
}

// Training-sample log: for every big vertex, (inputs from the previous
// count, log2(1 + this iteration's count)).
template <typename E>
__global__ void __launch_bounds__(256)
k_log_samples(const int32_t *__restrict__ vlist, int64_t big_v0, int64_t big_n, const float *__restrict__ fdeg,
              const float *__restrict__ fmean, const float *__restrict__ fmax, const int32_t *__restrict__ dprev,
              const int32_t *__restrict__ dcur, float *__restrict__ lx, float *__restrict__ ly, int64_t cap,
              int64_t *__restrict__ ln, const int32_t *__restrict__ stop) {
    if (*stop) return;
    for (int64_t base = (int64_t)blockIdx.x * 256; base < big_n; base += (int64_t)gridDim.x * 256) {
        const int64_t i = base + threadIdx.x;
        const bool ok = i < big_n;
        const int64_t pos = wave_append(ln, ok);
        if (ok && pos < cap) {
            const int64_t v = vlist ? (int64_t)vlist[big_v0 + i] : big_v0 + i;
            float x[4];
            mlp_inputs(fdeg[v], fmean[v], fmax[v], dprev[v], x);
            for (int k = 0; k < 4; ++k) lx[4 * pos + k] = x[k];
            ly[pos] = __log2f(1.f + (float)dcur[v]);
        }
    }
}

__global__ void k_fill_i32(int32_t *p, int64_t n, int32_t val) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        p[i] = val;
}

__global__ void k_gather_feat(const float *__restrict__ f, const int32_t *__restrict__ perm, int64_t n,
                              float *__restrict__ out) {
    for (int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; v < n; v += (int64_t)gridDim.x * blockDim.x)
        out[v] = f[perm ? perm[v] : v];
}

// ---------------------------------------------------------------------------
// Host side.

static int grid_cap(const galp_graph *g, int64_t work, int per_cu) {
    const int64_t cap = (int64_t)g->ctx->prop.multiProcessorCount * per_cu;
    return (int)std::max<int64_t>(1, std::min(work, cap));
}

template <class T>
static int lalloc(galp_graph *g, T **p, int64_t count) {
    *p = nullptr;
    if (count <= 0) count = 1;
    hipError_t e = hipMalloc((void **)p, sizeof(T) * (size_t)count);
    if (e != hipSuccess) {
        *p = nullptr;
        (void)hipGetLastError();
        return fail(g->ctx, GALP_ENOMEM, "learned-dispatch allocation failed");
    }
    g->device_bytes += (int64_t)(sizeof(T) * (size_t)count);
    return GALP_OK;
}

#define GALP_LTRY(expr)                 \
    do {                                \
        int _rc = (expr);               \
        if (_rc != GALP_OK) return _rc; \
    } while (0)

void learn_free(galp_graph *g) {
    LearnState *L = g->learn;
    if (!L) return;
    void *ptrs[] = {L->d_params,  L->d_fdeg,       L->d_fmean,    L->d_fmax,    L->d_dist[0],  L->d_dist[1],
                    L->d_cw_v,    L->d_cw_len,     L->d_cw_pslot, L->d_cw_e0,   L->d_mt_vid,   L->d_mt_of,
                    L->d_mt_ctile0, L->d_mt_nctiles, L->d_mt_cand, L->d_pn,     L->d_pkey,     L->d_pcnt,
                    L->d_plast,   L->d_hub_sel,    L->d_hub_distinct, L->d_plist, L->d_plist_n, L->d_log_x,
                    L->d_log_y,   L->d_log_n, L->d_cand_bytes, L->d_cw_base,
                    L->d_mt_tv,   L->d_mt_te0,     L->d_mt_tlen,  L->d_mt_tslot, L->d_mt_tmt,
                    L->d_hub_hd,  L->d_pnbr,       L->d_ppos};
    for (void *p : ptrs)
        if (p) (void)hipFree(p);
    for (int b = 0; b < 2 * NUM_BUCKETS; ++b)
        if (L->plist[b]) (void)hipFree(L->plist[b]);
    delete L;
    g->learn = nullptr;
}

// In core: the part-sorted copy of the multi-tile rows (L->d_pnbr,
// L->d_ppos; mt-major, compact) and its POS tiles of at most kCandPartTile
// edges (partsort.hpp): slots mt-major, the list sorted by first part.
static int fixed_part_tiles(galp_graph *g, LearnState *L, const std::vector<int32_t> &mt_vid,
                            const std::vector<int64_t> &mt_deg, std::vector<int32_t> &ctile0,
                            std::vector<int32_t> &nctiles, std::vector<int32_t> &tv, std::vector<int64_t> &te0,
                            std::vector<int32_t> &tlen, std::vector<int32_t> &tslot, std::vector<int32_t> &tmt) {
    const int64_t nr = (int64_t)mt_vid.size();
    std::vector<int64_t> roff((size_t)nr + 1, 0);
    for (int64_t r = 0; r < nr; ++r) roff[(size_t)r + 1] = roff[(size_t)r] + mt_deg[(size_t)r];
    GALP_LTRY(lalloc(g, &L->d_pnbr, roff.back()));
    GALP_LTRY(lalloc(g, &L->d_ppos, roff.back()));
    int32_t *d_rows = nullptr, *d_count = nullptr;
    int64_t *d_roff = nullptr;
    std::vector<int32_t> count((size_t)(nr * kNbrParts));
    hipError_t e = hipMalloc((void **)&d_rows, 4 * (size_t)nr);
    if (e == hipSuccess) e = hipMalloc((void **)&d_roff, 8 * (size_t)(nr + 1));
    if (e == hipSuccess) e = hipMalloc((void **)&d_count, 4 * count.size());
    if (e == hipSuccess) e = hipMemcpy(d_rows, mt_vid.data(), 4 * (size_t)nr, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(d_roff, roff.data(), 8 * roff.size(), hipMemcpyHostToDevice);
    if (e == hipSuccess) {
        const int grid = (int)((nr + 3) / 4);
        if (g->off64)
            k_rows_partsort<int64_t><<<grid, 256, 0, g->stream>>>((const int64_t *)g->d_off, g->d_nbr, d_rows, nr,
                                                                  d_roff, L->d_pnbr, L->d_ppos, d_count);
        else
            k_rows_partsort<int32_t><<<grid, 256, 0, g->stream>>>((const int32_t *)g->d_off, g->d_nbr, d_rows, nr,
                                                                  d_roff, L->d_pnbr, L->d_ppos, d_count);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipStreamSynchronize(g->stream);
    if (e == hipSuccess) e = hipMemcpy(count.data(), d_count, 4 * count.size(), hipMemcpyDeviceToHost);
    if (d_rows) (void)hipFree(d_rows);
    if (d_roff) (void)hipFree(d_roff);
    if (d_count) (void)hipFree(d_count);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        return fail(g->ctx, GALP_EHIP, std::string("multi-tile part sort: ") + hipGetErrorString(e));
    }
    std::vector<int32_t> part;
    ctile0.assign((size_t)nr, 0);
    nctiles.assign((size_t)nr, 0);
    part_tiles(nr, roff, count, kCandPartTile, [&](int64_t r, int64_t e0, int64_t len, int p) {
        if (nctiles[(size_t)r]++ == 0) ctile0[(size_t)r] = (int32_t)tv.size();
        tslot.push_back((int32_t)tv.size());
        tv.push_back(mt_vid[(size_t)r]);
        te0.push_back(e0);
        tlen.push_back((int32_t)len);
        tmt.push_back((int32_t)r);
        part.push_back(p);
    });
    std::vector<size_t> ord(tv.size());
    for (size_t i = 0; i < ord.size(); ++i) ord[i] = i;
    std::stable_sort(ord.begin(), ord.end(), [&](size_t a, size_t b) { return part[a] < part[b]; });
    auto permute = [&](auto &v) {
        auto c = v;
        for (size_t i = 0; i < ord.size(); ++i) v[i] = c[ord[i]];
    };
    permute(tv);
    permute(te0);
    permute(tlen);
    permute(tslot);
    permute(tmt);
    return GALP_OK;
}

// Allocates the learned-dispatch state and computes the static features.
int learn_setup(galp_graph *g) {
    if (g->learn) return GALP_OK;
    if (g->ooc && !g->reorder)   // a batch's big vertices are then one id range
        return fail(g->ctx, GALP_EUNSUPPORTED, "learned dispatch out-of-core needs the degree renumbering (reorder)");
    LearnState *L = new (std::nothrow) LearnState();
    if (!L) return fail(g->ctx, GALP_ENOMEM, "host allocation failed");
    g->learn = L;
    const int64_t n = g->n;
    // big vertices: one contiguous range of ids (degree-descending
    // renumbering: hubs first) or of the class-ordered vertex lists
    int64_t big_n = 0;
    for (int b = B_W128; b <= B_HUB; ++b) big_n += g->bucket[b].nv;
    L->big_n = big_n;
    L->big_v0 = g->reorder ? 0 : g->bucket[B_W128].v0;
    GALP_LTRY(lalloc(g, &L->d_fdeg, n));
    GALP_LTRY(lalloc(g, &L->d_fmean, n));
    GALP_LTRY(lalloc(g, &L->d_fmax, n));
    GALP_LTRY(lalloc(g, &L->d_dist[0], n));
    GALP_LTRY(lalloc(g, &L->d_dist[1], n));
    GALP_LTRY(lalloc(g, &L->d_plist_n, kNumCtrs));
    GALP_LTRY(lalloc(g, &L->d_cand_bytes, kMaxIters));
    GALP_LTRY(lalloc(g, &L->d_log_n, 1));
    for (int b = B_W128; b < B_HUB; ++b) {
        GALP_LTRY(lalloc(g, &L->plist[b], g->bucket[b].nv));
        GALP_LTRY(lalloc(g, &L->plist[NUM_BUCKETS + b], g->bucket[b].nv));
    }
    GALP_LTRY(lalloc(g, &L->d_plist, 2 * NUM_BUCKETS));
    if (hipMemcpy(L->d_plist, L->plist, sizeof(L->plist), hipMemcpyHostToDevice) != hipSuccess)
        return fail(g->ctx, GALP_EHIP, "class list table upload failed");
    // multi-tile vertices: hubs first (g->big_vid lists hubs first)
    std::vector<int32_t> mt_vid, ctile0, nctiles, mt_of((size_t)n, -1);
    std::vector<int64_t> mt_deg, mt_off;
    for (size_t k = 0; k < g->big_vid.size(); ++k) {
        const int64_t d = g->big_deg[k];
        if (d <= kCandTile) continue;
        mt_of[(size_t)g->big_vid[k]] = (int32_t)mt_vid.size();
        mt_vid.push_back(g->big_vid[k]);
        mt_deg.push_back(d);
        mt_off.push_back(g->big_off[k]);
    }
    // the fixed tile list: vertex, first edge, length, partial slot,
    // multi-tile index per tile; slots are consecutive per vertex (the
    // merge's ranges), the list itself is in processing order
    std::vector<int32_t> tv, tlen, tslot, tmt;
    std::vector<int64_t> te0;
    int64_t nct = 0;
    if (!g->ooc && !mt_vid.empty()) {
        // in core: POS tiles over a part-sorted copy of the rows, listed
        // part-major so the device gathers from one part's lines at a time
        GALP_LTRY(fixed_part_tiles(g, L, mt_vid, mt_deg, ctile0, nctiles, tv, te0, tlen, tslot, tmt));
        nct = (int64_t)tv.size();
    } else {
        // out of core (only row-order tiles of kCandTile edges over the batch
        // buffers; the dispatch leaves multi-tile rows to the class paths)
        for (size_t r = 0; r < mt_vid.size(); ++r) {
            ctile0.push_back((int32_t)nct);
            int32_t k = 0;
            for (int64_t t = 0; t * kCandTile < mt_deg[r]; ++t, ++k) {
                tslot.push_back((int32_t)tv.size());
                tv.push_back(mt_vid[r]);
                te0.push_back(mt_off[r] + t * kCandTile);
                tlen.push_back((int32_t)std::min<int64_t>(kCandTile, mt_deg[r] - t * kCandTile));
                tmt.push_back((int32_t)r);
            }
            nctiles.push_back(k);
            nct += k;
        }
    }
    for (int64_t h = 0; h < g->nhubs; ++h)   // hub h is multi-tile vertex h
        if (mt_of[(size_t)g->big_vid[(size_t)h]] != (int32_t)h)
            return fail(g->ctx, GALP_EINVAL, "internal: hub order");
    L->nmt = (int64_t)mt_vid.size();
    L->nctiles = nct;
    {   // tile-class segments: single-tile vertices by length, multi-tile tiles in the last
        int64_t segn[kCandClasses] = {};
        for (size_t k = 0; k < g->big_deg.size(); ++k)
            if (g->big_deg[k] <= kCandTile) segn[cand_class(g->big_deg[k])]++;
        L->cw_base[0] = 0;
        for (int c = 0; c < kCandClasses; ++c) L->cw_base[c + 1] = L->cw_base[c] + segn[c];
    }
    L->cw_cap = L->cw_base[kCandClasses];
    GALP_LTRY(lalloc(g, &L->d_cw_base, kCandClasses + 1));
    if (hipMemcpy(L->d_cw_base, L->cw_base, sizeof(L->cw_base), hipMemcpyHostToDevice) != hipSuccess)
        return fail(g->ctx, GALP_EHIP, "work-list segments upload failed");
    GALP_LTRY(lalloc(g, &L->d_cw_v, L->cw_cap));
    GALP_LTRY(lalloc(g, &L->d_cw_len, L->cw_cap));
    GALP_LTRY(lalloc(g, &L->d_cw_pslot, L->cw_cap));
    GALP_LTRY(lalloc(g, &L->d_cw_e0, L->cw_cap));
    GALP_LTRY(lalloc(g, &L->d_mt_vid, L->nmt));
    GALP_LTRY(lalloc(g, &L->d_mt_of, n));
    GALP_LTRY(lalloc(g, &L->d_mt_ctile0, L->nmt));
    GALP_LTRY(lalloc(g, &L->d_mt_nctiles, L->nmt));
    GALP_LTRY(lalloc(g, &L->d_mt_cand, L->nmt));
    {   // the fixed tile list
        GALP_LTRY(lalloc(g, &L->d_mt_tv, nct));
        GALP_LTRY(lalloc(g, &L->d_mt_te0, nct));
        GALP_LTRY(lalloc(g, &L->d_mt_tlen, nct));
        GALP_LTRY(lalloc(g, &L->d_mt_tslot, nct));
        GALP_LTRY(lalloc(g, &L->d_mt_tmt, nct));
        bool ok = true;
        if (nct) {
            ok = hipMemcpy(L->d_mt_tv, tv.data(), 4 * tv.size(), hipMemcpyHostToDevice) == hipSuccess &&
                 hipMemcpy(L->d_mt_te0, te0.data(), 8 * te0.size(), hipMemcpyHostToDevice) == hipSuccess &&
                 hipMemcpy(L->d_mt_tlen, tlen.data(), 4 * tlen.size(), hipMemcpyHostToDevice) == hipSuccess &&
                 hipMemcpy(L->d_mt_tslot, tslot.data(), 4 * tslot.size(), hipMemcpyHostToDevice) == hipSuccess &&
                 hipMemcpy(L->d_mt_tmt, tmt.data(), 4 * tmt.size(), hipMemcpyHostToDevice) == hipSuccess;
        }
        if (!ok || (int64_t)tv.size() != nct) return fail(g->ctx, GALP_EHIP, "multi-tile list upload failed");
    }
    GALP_LTRY(lalloc(g, &L->d_pn, nct));
    GALP_LTRY(lalloc(g, &L->d_pkey, nct * kCandK));
    GALP_LTRY(lalloc(g, &L->d_pcnt, nct * kCandK));
    GALP_LTRY(lalloc(g, &L->d_plast, nct * kCandK));
    GALP_LTRY(lalloc(g, &L->d_hub_sel, g->nhubs));
    GALP_LTRY(lalloc(g, &L->d_hub_hd, g->nhubs));
    GALP_LTRY(lalloc(g, &L->d_hub_distinct, g->nhubs));
    auto up = [&](void *dst, const void *src, size_t bytes) {
        return bytes == 0 || hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice) == hipSuccess;
    };
    if (!up(L->d_mt_vid, mt_vid.data(), 4 * mt_vid.size()) || !up(L->d_mt_ctile0, ctile0.data(), 4 * ctile0.size()) ||
        !up(L->d_mt_nctiles, nctiles.data(), 4 * nctiles.size()) || !up(L->d_mt_of, mt_of.data(), 4 * mt_of.size()))
        return fail(g->ctx, GALP_EHIP, "multi-tile tables upload failed");
    const int gd = grid_cap(g, (n + 255) / 256, 32);
    k_fill_i32<<<gd, 256, 0, g->stream>>>(L->d_dist[0], n, -1);
    k_fill_i32<<<gd, 256, 0, g->stream>>>(L->d_dist[1], n, -1);
    k_fill_i32<<<gd, 256, 0, g->stream>>>((int32_t *)L->d_fdeg, n, 0);   // 0.0f: features of small vertices
    k_fill_i32<<<gd, 256, 0, g->stream>>>((int32_t *)L->d_fmean, n, 0);
    k_fill_i32<<<gd, 256, 0, g->stream>>>((int32_t *)L->d_fmax, n, 0);
    k_fill_i32<<<grid_cap(g, (L->nmt + 255) / 256, 8), 256, 0, g->stream>>>(L->d_mt_cand, L->nmt, 0);
    k_fill_i32<<<1, 256, 0, g->stream>>>(L->d_hub_sel, g->nhubs, 1);
    k_fill_i32<<<1, 256, 0, g->stream>>>(L->d_hub_hd, g->nhubs, 0);
    k_fill_i32<<<1, 256, 0, g->stream>>>((int32_t *)L->d_hub_distinct, g->nhubs, 0);
    k_reset_ctrs<<<1, 64, 0, g->stream>>>(L->d_log_n, 1);   // d_log_n holds one word
    if (big_n && g->ooc) {
        GALP_LTRY(ooc_features(g));   // the rows stream through the batch buffers
    } else if (big_n) {
        const int gf = grid_cap(g, (big_n + 3) / 4, 64);
        if (g->off64)
            k_features<int64_t><<<gf, 256, 0, g->stream>>>((const int64_t *)g->d_off, g->d_nbr,
                                                           (const int64_t *)g->d_off, g->d_vlist, L->big_v0, big_n,
                                                           L->d_fdeg, L->d_fmean, L->d_fmax);
        else
            k_features<int32_t><<<gf, 256, 0, g->stream>>>((const int32_t *)g->d_off, g->d_nbr,
                                                           (const int32_t *)g->d_off, g->d_vlist, L->big_v0, big_n,
                                                           L->d_fdeg, L->d_fmean, L->d_fmax);
    }
    if (hipGetLastError() != hipSuccess || hipStreamSynchronize(g->stream) != hipSuccess)
        return fail(g->ctx, GALP_EHIP, "feature kernels failed");
    return GALP_OK;
}

int learn_features_view(galp_graph *g, const void *offrow, const int32_t *nbr, int64_t lo, int64_t hi) {
    LearnState *L = g->learn;
    const int64_t v1 = std::min(hi, L->big_n);   // renumbered: big vertices are ids [0, big_n)
    if (lo >= v1) return GALP_OK;
    const int gf = grid_cap(g, (v1 - lo + 3) / 4, 64);
    if (g->off64)
        k_features<int64_t><<<gf, 256, 0, g->stream>>>((const int64_t *)offrow, nbr, (const int64_t *)g->d_off,
                                                       nullptr, lo, v1 - lo, L->d_fdeg, L->d_fmean, L->d_fmax);
    else
        k_features<int32_t><<<gf, 256, 0, g->stream>>>((const int32_t *)offrow, nbr, (const int32_t *)g->d_off,
                                                       nullptr, lo, v1 - lo, L->d_fdeg, L->d_fmean, L->d_fmax);
    return hipGetLastError() == hipSuccess ? GALP_OK : fail(g->ctx, GALP_EHIP, "feature kernel launch failed");
}

int learn_reset_iteration(galp_graph *g) {
    LearnState *L = g->learn;
    if (L) L->dcur = 0;
    if (L) {
        const int gd = grid_cap(g, (g->n + 255) / 256, 32);
        k_fill_i32<<<gd, 256, 0, g->stream>>>(L->d_dist[0], g->n, -1);
        k_fill_i32<<<1, 256, 0, g->stream>>>(L->d_hub_sel, g->nhubs, 1);
        k_fill_i32<<<grid_cap(g, (L->nmt + 255) / 256, 8), 256, 0, g->stream>>>(L->d_mt_cand, L->nmt, 0);
        k_reset_ctrs<<<1, 64, 0, g->stream>>>(L->d_plist_n, kNumCtrs);
    }
    return hipGetLastError() == hipSuccess ? GALP_OK : fail(g->ctx, GALP_EHIP, "reset kernels failed");
}

// Big vertices [v0, v0 + nv) of the big list, rows located by off.
template <int H, typename E>
static void launch_dispatch(galp_graph *g, LearnState *L, int32_t *const *dpl, const E *off, int64_t v0,
                            int64_t nv) {
    if (!nv) return;
    k_dispatch<H, E><<<(int)((nv + kDispChunk - 1) / kDispChunk), 256, 0, g->stream>>>(
        off, g->d_vlist, v0, nv, !g->ooc, L->d_params, L->d_fdeg, L->d_fmean, L->d_fmax,
        L->d_dist[L->dcur], L->cand_thresh, g->nhubs, L->d_mt_of, L->d_mt_cand, L->d_hub_sel, L->d_cw_v,
        L->d_cw_e0, L->d_cw_len, L->d_cw_pslot, L->d_cw_base, dpl, L->d_hub_hd, L->d_plist_n,
        L->d_cand_bytes + g->iter % kMaxIters,
        g->d_stop);
}

// The dispatch and the candidate path over big vertices [v0, v0 + nv) (an
// index range of the big list), rows located by off / nbr: the whole graph
// in core, one batch's rebased offsets and buffer out of core.  The counters
// and lists are reset even when the range is empty, since the class kernels
// that follow read them.
template <typename E>
static int learned_front(galp_graph *g, const E *off, const int32_t *nbr, int64_t v0, int64_t nv, bool first,
                         bool merge) {
    LearnState *L = g->learn;
    int32_t *const *dpl = L->d_plist;
    const int32_t *prev = g->d_lab[g->cur];
    int32_t *next = g->d_lab[g->cur ^ 1];
    PendingLaunch pl;
    int64_t *cb = L->d_cand_bytes + g->iter % kMaxIters;
    k_reset_front<<<1, 64, 0, g->stream>>>(L->d_plist_n, cb, first ? 1 : 0);
    L->front_n = nv;
    if (!nv) return hipGetLastError() == hipSuccess ? GALP_OK : fail(g->ctx, GALP_EHIP, "reset launch failed");
    prof_begin(g, "dispatch", 0.0, &pl);
    switch (L->H) {
    case 8: launch_dispatch<8, E>(g, L, dpl, off, v0, nv); break;
    case 16: launch_dispatch<16, E>(g, L, dpl, off, v0, nv); break;
    default: launch_dispatch<32, E>(g, L, dpl, off, v0, nv); break;
    }
    prof_end(g, &pl);
    // the dispatch counts the candidate path's bytes per iteration (out of
    // core they add up over the batches, so no per-launch figure there)
    if (g->ooc) prof_begin(g, "cand", 0.0, &pl);
    else prof_begin_dev(g, "cand", cb, &pl);
    const int32_t hot = g->reorder ? (int32_t)std::min<int64_t>(kCandCache, g->n) : 0;
    CandLists cl{};
    for (int c = 0; c < kCandClasses; ++c) {
        const int64_t b = L->cw_base[c];
        cl.v[c] = L->d_cw_v + b;
        cl.len[c] = L->d_cw_len + b;
        cl.pslot[c] = L->d_cw_pslot + b;
        cl.e0[c] = L->d_cw_e0 + b;
    }
    cl.v[kCandClasses] = L->d_mt_tv;
    cl.len[kCandClasses] = L->d_mt_tlen;
    cl.pslot[kCandClasses] = L->d_mt_tslot;
    cl.e0[kCandClasses] = L->d_mt_te0;
    cl.tile_mt = L->d_mt_tmt;
    cl.nfixed = L->nctiles;
    cl.fnbr = g->ooc ? nullptr : L->d_pnbr;
    cl.fpos = g->ooc ? nullptr : L->d_ppos;
    k_cand<E><<<g->ctx->prop.multiProcessorCount, kCandThreads, 0, g->stream>>>(
        off, nbr, prev, next, cl, L->d_mt_cand, L->d_pkey, L->d_pcnt, L->d_plast, L->d_pn, distinct_out(g), dpl,
        L->d_plist_n, hot, g->d_counter, g->d_stop);
    prof_end(g, &pl);
    if (merge && L->nmt && !g->ooc) {
        prof_begin(g, "cand_merge", 0.0, &pl);
        k_cand_merge<E><<<grid_cap(g, (L->nmt + 3) / 4, 8), 256, 0, g->stream>>>(
            (const E *)g->d_off, L->d_mt_vid, L->nmt, g->nhubs, L->d_mt_cand, L->d_hub_sel, L->d_mt_ctile0,
            L->d_mt_nctiles, dpl, L->d_pkey, L->d_pcnt, L->d_plast, L->d_pn, prev, next, distinct_out(g),
            g->d_counter, L->d_plist_n, g->d_stop);
        prof_end(g, &pl);
    }
    return hipGetLastError() == hipSuccess ? GALP_OK : fail(g->ctx, GALP_EHIP, "learned dispatch launch failed");
}

int launch_learned_front(galp_graph *g, bool merge) {
    LearnState *L = g->learn;
    return g->off64
               ? learned_front<int64_t>(g, (const int64_t *)g->d_off, g->d_nbr, L->big_v0, L->big_n, true, merge)
               : learned_front<int32_t>(g, (const int32_t *)g->d_off, g->d_nbr, L->big_v0, L->big_n, true, merge);
}

template <typename E>
static int learned_merge(galp_graph *g) {
    LearnState *L = g->learn;
    if (!L->nmt || g->ooc || !L->front_n) return GALP_OK;
    PendingLaunch pl;
    prof_begin(g, "cand_merge", 0.0, &pl);
    k_cand_merge<E><<<grid_cap(g, (L->nmt + 3) / 4, 8), 256, 0, g->stream>>>(
        (const E *)g->d_off, L->d_mt_vid, L->nmt, g->nhubs, L->d_mt_cand, L->d_hub_sel, L->d_mt_ctile0,
        L->d_mt_nctiles, L->d_plist, L->d_pkey, L->d_pcnt, L->d_plast, L->d_pn, g->d_lab[g->cur],
        g->d_lab[g->cur ^ 1], distinct_out(g), g->d_counter, L->d_plist_n, g->d_stop);
    prof_end(g, &pl);
    return hipGetLastError() == hipSuccess ? GALP_OK : fail(g->ctx, GALP_EHIP, "candidate merge launch failed");
}

int launch_learned_merge(galp_graph *g) {
    return g->off64 ? learned_merge<int64_t>(g) : learned_merge<int32_t>(g);
}

int launch_learned_front_batch(galp_graph *g, const IterView &w, int64_t lo, int64_t hi) {
    LearnState *L = g->learn;
    const int64_t v1 = std::min(hi, L->big_n);   // renumbered: big vertices are ids [0, big_n)
    const int64_t nv = lo < v1 ? v1 - lo : 0;
    return g->off64 ? learned_front<int64_t>(g, (const int64_t *)w.off, w.nbr, lo, nv, lo == 0, true)
                    : learned_front<int32_t>(g, (const int32_t *)w.off, w.nbr, lo, nv, lo == 0, true);
}

// After the iteration's LP kernels: log samples (if enabled), swap counts.
int learn_after_iteration(galp_graph *g) {
    LearnState *L = g->learn;
    if (!L) return GALP_OK;
    if (L->log_cap && L->big_n) {
        k_log_samples<int32_t><<<grid_cap(g, (L->big_n + 255) / 256, 8), 256, 0, g->stream>>>(
            g->d_vlist, L->big_v0, L->big_n, L->d_fdeg, L->d_fmean, L->d_fmax, L->d_dist[L->dcur],
            L->d_dist[L->dcur ^ 1], L->d_log_x, L->d_log_y, L->log_cap, L->d_log_n, g->d_stop);
    }
    L->dcur ^= 1;
    return hipGetLastError() == hipSuccess ? GALP_OK : fail(g->ctx, GALP_EHIP, "log kernel failed");
}

}  // namespace galp

using namespace galp;

extern "C" {

GALP_API int galp_graph_set_mlp(galp_graph *g, int32_t hidden, const float *params, float cand_thresh) {
    if (!g || !params) return GALP_EINVAL;
    if (hidden != 8 && hidden != 16 && hidden != 32) return fail(g->ctx, GALP_EINVAL, "hidden must be 8, 16 or 32");
    if (!(cand_thresh >= 0.f)) return fail(g->ctx, GALP_EINVAL, "cand_thresh must be >= 0");
    int rc = learn_setup(g);
    if (rc != GALP_OK) return rc;
    invalidate_captured(g);
    LearnState *L = g->learn;
    if (!L->d_params) {
        if (hipMalloc((void **)&L->d_params, 4 * (size_t)mlp_num_params(32)) != hipSuccess)
            return fail(g->ctx, GALP_ENOMEM, "parameter buffer");
    }
    if (hipMemcpy(L->d_params, params, 4 * (size_t)mlp_num_params(hidden), hipMemcpyHostToDevice) != hipSuccess)
        return fail(g->ctx, GALP_EHIP, "parameter upload failed");
    L->H = hidden;
    L->cand_thresh = cand_thresh;
    return GALP_OK;
}

GALP_API int galp_graph_log_enable(galp_graph *g, int64_t capacity) {
    if (!g || capacity < 0) return GALP_EINVAL;
    int rc = learn_setup(g);
    if (rc != GALP_OK) return rc;
    invalidate_captured(g);
    LearnState *L = g->learn;
    if (L->d_log_x) (void)hipFree(L->d_log_x);
    if (L->d_log_y) (void)hipFree(L->d_log_y);
    L->d_log_x = L->d_log_y = nullptr;
    L->log_cap = 0;
    if (capacity) {
        if (hipMalloc((void **)&L->d_log_x, 16 * (size_t)capacity) != hipSuccess ||
            hipMalloc((void **)&L->d_log_y, 4 * (size_t)capacity) != hipSuccess)
            return fail(g->ctx, GALP_ENOMEM, "log buffer");
        L->log_cap = capacity;
    }
    int64_t zero = 0;
    if (hipMemcpy(L->d_log_n, &zero, 8, hipMemcpyHostToDevice) != hipSuccess)
        return fail(g->ctx, GALP_EHIP, "log reset failed");
    return GALP_OK;
}

GALP_API int galp_graph_log_read(galp_graph *g, float *x, float *y, int64_t max, int64_t *count) {
    if (!g || !count || !g->learn || !g->learn->d_log_n) return GALP_EINVAL;
    LearnState *L = g->learn;
    int64_t nlog = 0;
    if (hipMemcpy(&nlog, L->d_log_n, 8, hipMemcpyDeviceToHost) != hipSuccess)
        return fail(g->ctx, GALP_EHIP, "log readback failed");
    nlog = std::min(nlog, L->log_cap);
    *count = nlog;
    const int64_t k = std::min(nlog, max);
    if (k > 0 && x && y) {
        if (hipMemcpy(x, L->d_log_x, 16 * (size_t)k, hipMemcpyDeviceToHost) != hipSuccess ||
            hipMemcpy(y, L->d_log_y, 4 * (size_t)k, hipMemcpyDeviceToHost) != hipSuccess)
            return fail(g->ctx, GALP_EHIP, "log readback failed");
    }
    return GALP_OK;
}

GALP_API int galp_graph_features(galp_graph *g, float *deg, float *mean, float *mx) {
    if (!g || !deg || !mean || !mx) return GALP_EINVAL;
    int rc = learn_setup(g);
    if (rc != GALP_OK) return rc;
    LearnState *L = g->learn;
    float *tmp = nullptr;
    if (hipMalloc((void **)&tmp, 4 * (size_t)std::max<int64_t>(g->n, 1)) != hipSuccess)
        return fail(g->ctx, GALP_ENOMEM, "feature staging");
    float *srcs[3] = {L->d_fdeg, L->d_fmean, L->d_fmax};
    float *dsts[3] = {deg, mean, mx};
    for (int k = 0; k < 3 && g->n; ++k) {
        k_gather_feat<<<grid_cap(g, (g->n + 255) / 256, 32), 256, 0, g->stream>>>(srcs[k], g->d_perm, g->n, tmp);
        if (hipMemcpyAsync(dsts[k], tmp, 4 * (size_t)g->n, hipMemcpyDeviceToHost, g->stream) != hipSuccess ||
            hipStreamSynchronize(g->stream) != hipSuccess) {
            (void)hipFree(tmp);
            return fail(g->ctx, GALP_EHIP, "feature readback failed");
        }
    }
    (void)hipFree(tmp);
    return GALP_OK;
}

}  // extern "C"
