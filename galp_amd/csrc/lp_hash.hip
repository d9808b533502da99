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
//     last positions come from lane exchanges, no table at all;
//   - 65..4096: one wave (or one workgroup of up to 512 threads) per vertex
//     with an open-addressing table in LDS at load factor <= 1/2;
//   - hubs (> 4096): tiles of 2048 edges counted in LDS, each writing its
//     distinct labels grouped by hash partition; a merge per (hub,
//     partition) sums them in LDS, and a 64-bit atomicMax of (count, last)
//     per merge picks the winner: no device-scope table at all.
// Table-based kernels run a software pipeline over their vertices: the
// labels of the next vertex and the neighbour ids of the one after are in
// flight while the current vertex is counted, so the dependent
// offsets -> neighbours -> labels chain is hidden.  Tables are cleared
// lazily (each lane resets the slots it used), never with a memset.
// All kernels are synchronous (read prev, write next) and apply the tie rule
// argmax (count, last position), independent of scheduling and atomic order.

#include "device_common.hpp"
#include "internal.hpp"
#include "pipeline.hpp"
#include "group_kernel.hpp"

namespace galp {

const int64_t kBucketHi[NUM_BUCKETS] = {bucket_hi(0), bucket_hi(1), bucket_hi(2),  bucket_hi(3),
                                        bucket_hi(4), bucket_hi(5), bucket_hi(6),  bucket_hi(7),
                                        bucket_hi(8), bucket_hi(9), bucket_hi(10), bucket_hi(11)};
static_assert(bucket_hi(0) == 4 && bucket_hi(4) == 64 && bucket_hi(5) == 128 && bucket_hi(10) == 4096,
              "degree classes");
const char *const kBucketName[NUM_BUCKETS] = {"g4",   "g8",   "g16",  "g32",  "g64",  "w128",
                                              "w256", "w512", "b1k",  "b2k",  "b4k",  "hub"};

// ---------------------------------------------------------------------------
// Degree <= G: one aligned group of G lanes per vertex; lane p holds the
// label of the p-th neighbour (or -1).  The group's max of (count << 8 |
// last) is found in two stages:
//   peel:  up to PEEL rounds; in each round every group takes its smallest
//          unresolved label (a min over lane moves) and one ballot gives that
//          label's count (popcount) and last position (highest set bit), so
//          the labels of a converged neighbourhood are settled in a round
//          each, whatever the order;
//   sweep: lanes still unresolved (a group with more than PEEL distinct
//          labels) count the lanes of their group holding their label; the
//          max over those lanes of (count << 8 | lane) is the max over their
//          labels of (count << 8 | last), since a label's highest lane is its
//          last position.
// A peeled label takes all of its lanes with it, so the sweep's counts
// among unresolved lanes are full counts.  The result is the same for every
// lane of the group.
template <int G, int PEEL>
__device__ __forceinline__ uint32_t group_best(int32_t L) {
    const int lane = lane_id();
    const int gbase = lane & ~(G - 1);
    bool active = L >= 0;
    uint32_t best = 0;
#pragma unroll
    for (int r = 0; r < PEEL; ++r) {
        const unsigned long long ball = __ballot(active);
        if (ball == 0) break;   // wave-uniform
        int32_t Lk;
        if constexpr (G == kWave) Lk = __builtin_amdgcn_readlane(L, __ffsll((long long)ball) - 1);
        else Lk = group_min<G>(active ? L : INT32_MAX);
        const bool eq = active && (L == Lk);
        const unsigned long long b = __ballot(eq);
        uint32_t cnt, last;
        bool any;
        if constexpr (G == kWave) {
            any = b != 0ull;
            cnt = (uint32_t)__popcll(b);
            last = (uint32_t)(63 - __clzll((long long)(b | 1ull)));
        } else {
            const uint32_t mg = (uint32_t)(b >> gbase) & (uint32_t)((1ull << G) - 1ull);
            any = mg != 0u;
            cnt = (uint32_t)__popc(mg);
            last = (uint32_t)(31 - __clz((int)(mg | 1u)));
        }
        best = any ? max(best, (cnt << 8) | last) : best;
        active = active && !eq;
    }
    if (__ballot(active)) {
        const int32_t La = active ? L : -1;
        uint32_t cnt = 0;
        if constexpr (G == kWave) {
#pragma unroll
            for (int j = 0; j < kWave; ++j) cnt += (uint32_t)(__builtin_amdgcn_readlane(La, j) == La);
        } else {
            cnt = group_count_eq<G>(La);
        }
        best = max(best, group_max<G>(active ? ((cnt << 8) | (uint32_t)(lane & (G - 1))) : 0u));
    }
    return best;
}

struct HashGroupResolve {
    template <int G>
    __device__ __forceinline__ static uint32_t best(int32_t L) { return group_best<G, (G < 4 ? G : 4)>(L); }
};
// High-distinct iterations: the pairwise sweep alone (peel rounds resolve one
// lane each when every label is rare).
struct HashGroupSweep {
    template <int G>
    __device__ __forceinline__ static uint32_t best(int32_t L) { return group_best<G, 0>(L); }
};

// Open-addressing insert into an LDS table of 2^BITS slots; returns the slot
// and sets *created when this lane claimed an empty slot.  The caller
// guarantees fewer keys than slots; the probe bound only guards the GPU.
template <int BITS>
__device__ __forceinline__ uint32_t lds_insert(uint32_t *skey, uint32_t L, bool *created, int32_t *err) {
    constexpr uint32_t MASK = (1u << BITS) - 1u;
    uint32_t h = hash_slot(L, BITS);
    *created = false;
    for (uint32_t probe = 0; probe <= MASK; ++probe) {
        const uint32_t cur = skey[h];
        if (cur == L) return h;
        if (cur == kEmpty) {
            const uint32_t old = atomicCAS(&skey[h], kEmpty, L);
            if (old == kEmpty) {
                *created = true;
                return h;
            }
            if (old == L) return h;
        }
        h = (h + 1) & MASK;
    }
    atomicOr(err, 1);
    return h;
}

// lds_insert for callers that recover from a full table: returns kTableFull
// (no slot taken) instead of flagging an error.
constexpr uint32_t kTableFull = 0xffffffffu;
template <int BITS>
__device__ __forceinline__ uint32_t lds_try_insert(uint32_t *skey, uint32_t L, bool *created) {
    constexpr uint32_t MASK = (1u << BITS) - 1u;
    uint32_t h = hash_slot(L, BITS);
    *created = false;
    for (uint32_t probe = 0; probe <= MASK; ++probe) {
        const uint32_t cur = skey[h];
        if (cur == L) return h;
        if (cur == kEmpty) {
            const uint32_t old = atomicCAS(&skey[h], kEmpty, L);
            if (old == kEmpty) {
                *created = true;
                return h;
            }
            if (old == L) return h;
        }
        h = (h + 1) & MASK;
    }
    return kTableFull;
}

// Inserts one 64-lane chunk of labels (L < 0: no edge) into an LDS table,
// adding each label's count (and, if LAST, the max of its positions `pos`).
// Up to PEEL labels are committed by a single leader lane after a ballot --
// in a converged neighbourhood one or two labels fill the whole chunk, and
// 64 lanes hitting one counter would serialise -- then the remaining lanes
// insert one by one.  Returns the lane's slot; *own marks lanes that wrote
// the slot (they reset it afterwards), *creator the one lane per table that
// claimed it; *maxc tracks the largest count seen.
template <int BITS, int PEEL, bool LAST>
__device__ __forceinline__ uint32_t lds_insert_chunk(uint32_t *skey, uint32_t *scnt, uint32_t *slast,
                                                     int32_t L, uint32_t pos, bool *own,
                                                     bool *creator, uint32_t *maxc, int32_t *err) {
    const int lane = lane_id();
    bool active = L >= 0;
    uint32_t slot = 0;
    *own = false;
    *creator = false;
#pragma unroll
    for (int r = 0; r < PEEL; ++r) {
        const unsigned long long ball = __ballot(active);
        if (ball == 0) break;
        const int leader = __ffsll((long long)ball) - 1;
        const int32_t Lk = __builtin_amdgcn_readlane(L, leader);
        const bool eq = active && (L == Lk);
        const unsigned long long m = __ballot(eq);
        const uint32_t lastpos = LAST ? (uint32_t)__builtin_amdgcn_readlane((int)pos, 63 - __clzll((long long)m)) : 0u;
        uint32_t sl = 0;
        if (lane == leader) {
            bool created;
            const uint32_t c = (uint32_t)__popcll(m);
            sl = lds_insert<BITS>(skey, (uint32_t)Lk, &created, err);
            *creator = created;
            *maxc = max(*maxc, atomicAdd(&scnt[sl], c) + c);
            if (LAST) atomicMax(&slast[sl], lastpos);
            *own = true;
        }
        sl = (uint32_t)__builtin_amdgcn_readlane((int)sl, leader);
        if (eq) slot = sl;
        active = active && !eq;
    }
    if (active) {
        bool created;
        slot = lds_insert<BITS>(skey, (uint32_t)L, &created, err);
        *creator = created;
        *maxc = max(*maxc, atomicAdd(&scnt[slot], 1u) + 1u);
        if (LAST) atomicMax(&slast[slot], pos);
        *own = true;
    }
    return slot;
}

// ---------------------------------------------------------------------------
// 64 < degree <= T*EPT: one block of T threads per vertex, LDS table of
// 2^CAPBITS = 2*T*EPT slots.  Pass 1 inserts (CAS on the key) and counts
// (atomic add); the block max of the returned counts is the winning count C.
// Pass 2 keeps the largest position whose label has count C: that
// position's label is argmax (count, last).  Each lane then resets the slots
// it used, leaving the table empty for the next vertex.
template <int T, int EPT, int CAPBITS, typename E, int PEEL>
struct LdsVertexPolicy {
    struct Item {
        bool valid;
        int len;
        int64_t v;
        int64_t b;
        int32_t old;
    };
    const E *off;
    const int32_t *prev;
    int32_t *next;
    const int32_t *vlist;
    int64_t v0, nv;
    uint32_t *skey, *scnt, *sred;
    unsigned long long *sred64;
    int32_t *err;
    int32_t *distinct;   // if set: distinct labels per vertex (learned dispatch feature)
    uint32_t changed;

    __device__ Item item(int64_t i) const {
        Item it{false, 0, 0, 0, 0};
        if (i < nv) {
            it.v = vlist ? (int64_t)vlist[v0 + i] : v0 + i;
            it.b = (int64_t)off[it.v];
            it.len = (int)((int64_t)off[it.v + 1] - it.b);
            it.old = prev[it.v];
            it.valid = true;
            if (it.len > T * EPT) {   // never planned into this class
                if (threadIdx.x == 0) atomicOr(err, 8);
                it.len = 0;
            }
        }
        return it;
    }
    __device__ int64_t edge0(const Item &it) const { return it.b; }

    __device__ void process(const Item &it, const int32_t (&lab)[EPT]) {
        if (it.len == 0) return;
        uint32_t slot[EPT];
        bool own[EPT], creator[EPT];
        uint32_t maxc = 0;
        __syncthreads();
#pragma unroll
        for (int k = 0; k < EPT; ++k) {
            const int pos = k * T + (int)threadIdx.x;
            slot[k] = lds_insert_chunk<CAPBITS, PEEL, false>(skey, scnt, nullptr, pos < it.len ? lab[k] : -1,
                                                          (uint32_t)pos, &own[k], &creator[k], &maxc, err);
        }
        __syncthreads();
        maxc = block_max<T / kWave>(maxc, sred);
        unsigned long long best = 0;
#pragma unroll
        for (int k = 0; k < EPT; ++k) {
            const int pos = k * T + (int)threadIdx.x;
            if (pos < it.len && scnt[slot[k]] == maxc)
                best = ((unsigned long long)(pos + 1) << 32) | (uint32_t)lab[k];
        }
        best = block_max<T / kWave>(best, sred64);
        uint32_t ndist = 0;
        if (distinct) {
#pragma unroll
            for (int k = 0; k < EPT; ++k) ndist += creator[k];
            ndist = block_sum<T / kWave>(ndist, sred);
        }
#pragma unroll
        for (int k = 0; k < EPT; ++k) {
            if (own[k]) {
                skey[slot[k]] = kEmpty;
                scnt[slot[k]] = 0;
            }
        }
        if (threadIdx.x == 0) {
            const int32_t w = (int32_t)(uint32_t)best;
            next[it.v] = w;
            changed += (w != it.old);
            if (distinct) distinct[it.v] = (int32_t)ndist;
        }
    }
};

// Two vertex lists in one launch: list A with PEEL_A leader-committed peel
// rounds per 64-edge chunk before the lane-by-lane inserts, then list B with
// none (the high-distinct rows of learned dispatch; nvB = 0 otherwise).
// Peel rounds pay where a few labels repeat across a chunk; each phase is its
// own pipeline, so no per-row branch sits between the loads in flight.
template <int T, int EPT, int CAPBITS, typename E, int PEEL_A>
__global__ void __launch_bounds__(T)
k_hash_lds(const E *__restrict__ off, const int32_t *__restrict__ nbr,
           const int32_t *__restrict__ prev, int32_t *__restrict__ next,
           const int32_t *__restrict__ vlist, int64_t v0, int64_t nv, const int64_t *__restrict__ nv_dev,
           const int32_t *__restrict__ vlistB, const int64_t *__restrict__ nvB_dev,
           unsigned long long *__restrict__ counter, const int32_t *__restrict__ stop,
           int32_t *__restrict__ err, int32_t *__restrict__ distinct) {
    static_assert((1 << CAPBITS) >= 2 * T * EPT, "load factor must stay <= 1/2");
    constexpr int CAP = 1 << CAPBITS;
    __shared__ uint32_t skey[CAP];
    __shared__ uint32_t scnt[CAP];
    __shared__ uint32_t sred[T / kWave];
    __shared__ unsigned long long sred64[T / kWave];
    if (*stop) return;
    if (nv_dev) nv = *nv_dev;   // lists built on the device this iteration
    const int64_t nvB = nvB_dev ? *nvB_dev : 0;
    if ((int64_t)blockIdx.x >= nv && (int64_t)blockIdx.x >= nvB) return;
    for (int s = threadIdx.x; s < CAP; s += T) {
        skey[s] = kEmpty;
        scnt[s] = 0;
    }
    uint32_t changed = 0;
    if ((int64_t)blockIdx.x < nv) {   // uniform per block
        LdsVertexPolicy<T, EPT, CAPBITS, E, PEEL_A> pol{off, prev, next, vlist, v0, nv, skey, scnt, sred, sred64,
                                                        err, distinct, 0u};
        pipelined<T, EPT>(pol, nbr, prev, (int64_t)blockIdx.x, (int64_t)gridDim.x, (int)threadIdx.x);
        changed += pol.changed;
    }
    if ((int64_t)blockIdx.x < nvB) {
        __syncthreads();
        LdsVertexPolicy<T, EPT, CAPBITS, E, 0> pol{off, prev, next, vlistB, 0, nvB, skey, scnt, sred, sred64, err,
                                                   distinct, 0u};
        pipelined<T, EPT>(pol, nbr, prev, (int64_t)blockIdx.x, (int64_t)gridDim.x, (int)threadIdx.x);
        changed += pol.changed;
    }
    __syncthreads();
    block_count_add<T / kWave>(changed, counter, sred);
}

// ---------------------------------------------------------------------------
// Hubs (degree > 4096): two phases, every edge gathered once.
//  1. tile: a block aggregates one kHubTile-edge tile of a hub row in LDS
//     (key, count, last position in the row), counting-sorts the tile's
//     distinct labels by label partition (P_h = ceil(deg / kHubPart)
//     partitions of a hash of the label) and writes them as a compact list
//     plus P_h + 1 partition offsets;
//  2. merge: block (h, p) folds partition p of every tile of hub h into an
//     LDS table -- no two blocks share a label, so there is no device-scope
//     table and no atomics on it -- and applies the tie rule within the
//     partition; one 64-bit atomicMax of (count << 32 | last) per block
//     picks the hub's winner, which is the label at position `last`.

// HD: the launch for high-distinct hubs (lane-by-lane inserts, no peel
// rounds); with a hub_hd list, each launch takes the hubs of its kind.
template <int T, int EPT, typename E, bool HD>
struct HubTilePolicy {
    struct Item {
        bool valid;
        int len;
        int64_t e0;       // absolute index of the tile's first edge
        uint32_t row0;    // position of that edge in its row
        uint32_t P;       // partitions of the hub
        int64_t idx;      // offset of the tile's partition offsets
        int64_t t;        // tile id
    };
    const E *off;
    const int32_t *tile_hub;
    const int32_t *hub_tile0;
    const int32_t *hub_vid;
    const int32_t *hub_nparts;
    const int64_t *tile_idx_off;
    int64_t ntiles;
    uint32_t *skey, *scnt, *slast, *hist;
    uint32_t *out_key, *out_ent;
    int32_t *tile_idx;
    int32_t *err;
    const int32_t *hub_sel;   // null: every hub; else only hubs with a non-zero flag
    int *sscan;               // T / kWave ints of LDS for the block scan
    const int32_t *hub_hd;    // null: every hub; else only hubs whose flag equals HD

    __device__ Item item(int64_t t) const {
        Item it{false, 0, 0, 0, 1, 0, t};
        if (t < ntiles) {
            const int h = tile_hub[t];
            const int64_t v = hub_vid[h];
            const int64_t rb = (int64_t)off[v], re = (int64_t)off[v + 1];
            it.row0 = (uint32_t)((t - (int64_t)hub_tile0[h]) * (int64_t)kHubTile);   // tiles of a hub are consecutive
            it.e0 = rb + it.row0;
            it.len = (int)min((int64_t)kHubTile, re - it.e0);
            if (hub_sel && !hub_sel[h]) it.len = 0;   // processed as a no-op
            if (hub_hd && (hub_hd[h] != 0) != HD) it.len = 0;   // the other launch's hub
            it.P = (uint32_t)hub_nparts[h];
            it.idx = tile_idx_off[t];
            it.valid = true;
        }
        return it;
    }
    __device__ int64_t edge0(const Item &it) const { return it.e0; }

    // inclusive scan of the P + 1 partition counts in hist, each thread a run
    // of consecutive ones; hist[p] becomes the end of partition p - 1, and the
    // scanned counts are the tile's partition offsets
    __device__ void scan_parts(const Item &it) {
        const uint32_t np = it.P + 1;
        const uint32_t per = (np + T - 1) / T;
        const uint32_t p0 = (uint32_t)threadIdx.x * per;
        int run = 0;
        for (uint32_t p = p0; p < p0 + per && p < np; ++p) run += (int)hist[p];
        int acc = block_exclusive_scan<T / kWave>(run, sscan);   // barriers inside
        for (uint32_t p = p0; p < p0 + per && p < np; ++p) {
            acc += (int)hist[p];
            hist[p] = (uint32_t)acc;
            tile_idx[it.idx + p] = acc;
        }
    }

    // The table is clean (keys empty, counts and positions 0) when a tile
    // starts and again when it ends.  In between, the tile's distinct labels
    // move to registers and the table is reset, so the count array serves as
    // the partition histogram (counts are 0 after the reset, so it starts
    // cleared); it is zeroed again at the end.
    __device__ void process(const Item &it, const int32_t (&lab)[EPT]) {
        if (it.len == 0) return;   // hub left to the candidate path: the merge skips it too
        uint32_t slot[EPT];
        bool own[EPT], creator[EPT];
        uint32_t maxc = 0;
#pragma unroll
        for (int k = 0; k < EPT; ++k) {
            const int pos = k * T + (int)threadIdx.x;
            slot[k] = lds_insert_chunk<kHubTileCapBits, HD ? 0 : 3, true>(
                skey, scnt, slast, pos < it.len ? lab[k] : -1, it.row0 + (uint32_t)pos, &own[k], &creator[k], &maxc,
                err);
        }
        __syncthreads();
        uint32_t key[EPT], cnt[EPT], last[EPT], part[EPT];
#pragma unroll
        for (int k = 0; k < EPT; ++k)
            if (creator[k]) {
                key[k] = skey[slot[k]];
                cnt[k] = scnt[slot[k]];
                last[k] = slast[slot[k]];
                part[k] = hub_partition(key[k], it.P);
            }
        __syncthreads();
#pragma unroll
        for (int k = 0; k < EPT; ++k)
            if (own[k]) {
                skey[slot[k]] = kEmpty;
                scnt[slot[k]] = 0;
                slast[slot[k]] = 0;
            }
        __syncthreads();
        // counting sort of the tile's distinct labels by partition
#pragma unroll
        for (int k = 0; k < EPT; ++k)
            if (creator[k]) atomicAdd(&hist[part[k] + 1], 1u);
        __syncthreads();
        scan_parts(it);
        __syncthreads();
        const int64_t ob = it.t * (int64_t)kHubTile;
#pragma unroll
        for (int k = 0; k < EPT; ++k)
            if (creator[k]) {
                const uint32_t q = atomicAdd(&hist[part[k]], 1u);
                out_key[ob + q] = key[k];
                out_ent[ob + q] = hub_entry(cnt[k], last[k] - it.row0);
            }
        __syncthreads();
        for (uint32_t p = threadIdx.x; p <= it.P; p += T) hist[p] = 0;
        __syncthreads();
    }
};

template <int T, typename E, bool HD>
__global__ void __launch_bounds__(T)
k_hub_tile(const E *__restrict__ off, const int32_t *__restrict__ nbr, const int32_t *__restrict__ prev,
           const int32_t *__restrict__ tile_hub, const int32_t *__restrict__ hub_tile0,
           const int32_t *__restrict__ hub_vid, const int32_t *__restrict__ hub_nparts,
           const int64_t *__restrict__ tile_idx_off, int64_t tbase, int64_t ntiles,
           uint32_t *__restrict__ out_key, uint32_t *__restrict__ out_ent,
           int32_t *__restrict__ tile_idx, const int32_t *__restrict__ stop, int32_t *__restrict__ err,
           const int32_t *__restrict__ hub_sel, const int64_t *__restrict__ nsel,
           const int32_t *__restrict__ hub_hd, const int64_t *__restrict__ nhd) {
    constexpr int CAP = 1 << kHubTileCapBits;
    constexpr int EPT = kHubTile / T;
    // keys | counts | positions; the partition histogram (P + 1 entries)
    // reuses the counts and, for P = kHubMaxParts, the first position
    static_assert(kHubMaxParts + 1 <= 2 * CAP, "partition histogram fits in the table");
    __shared__ uint32_t tab[3 * CAP];
    __shared__ int sscan[T / kWave];
    if (*stop) return;
    if (nsel) {   // learned dispatch: nsel hubs on this path, nhd of them high-distinct
        const int64_t ns = *nsel, nh = nhd ? *nhd : 0;
        if ((hub_hd ? (HD ? nh : ns - nh) : ns) == 0) return;   // none for this launch
    }
    for (int s = threadIdx.x; s < CAP; s += T) {
        tab[s] = kEmpty;
        tab[CAP + s] = 0;
        tab[2 * CAP + s] = 0;
    }
    __syncthreads();
    HubTilePolicy<T, EPT, E, HD> pol{off,      tile_hub, hub_tile0,     hub_vid,  hub_nparts, tile_idx_off,
                                     ntiles,   tab,      tab + CAP,     tab + 2 * CAP,        tab + CAP,
                                     out_key,  out_ent,  tile_idx,      err,      hub_sel,    sscan,
                                     hub_hd};
    pipelined<T, EPT>(pol, nbr, prev, tbase + (int64_t)blockIdx.x, (int64_t)gridDim.x, (int)threadIdx.x);
}

// Merge items with at most kHmSmall entries -- the usual case once labels
// have settled, when a partition holds about one entry per tile -- by one
// wave each: the wave walks the item's entries 64 at a time and peels each
// batch label by label (one ballot per distinct label, a wave sum of its
// counts and a wave max of its last positions), keeping up to 64 distinct
// labels in lanes (lane j: label, count, last position of candidate j).
// Items with more entries, or more than 64 distinct labels, are flagged for
// the block merge (k_hub_merge), every other item's flag is cleared (hubs
// outside the launch's selection included); waves need no barrier.  Each
// wave loads its next item's record while it merges the current one.
constexpr int kHmSmall = 1024;
constexpr int kHmSmallT = 256;

__global__ void __launch_bounds__(kHmSmallT)
k_hub_merge_small(const HubItem *__restrict__ items, int64_t ibase, int64_t nitems,
                  const int32_t *__restrict__ tile_idx, const uint32_t *__restrict__ out_key,
                  const uint32_t *__restrict__ out_ent, unsigned long long *__restrict__ hub_best,
                  const int32_t *__restrict__ hub_sel, uint32_t *__restrict__ hub_distinct,
                  const int64_t *__restrict__ nsel, int32_t *__restrict__ big, const int32_t *__restrict__ stop) {
    constexpr int NW = kHmSmallT / kWave;
    __shared__ int32_t slo[NW][kWave], sstart[NW][kWave];
    if (*stop || (nsel && *nsel == 0)) return;
    const int lane = lane_id(), w = (int)(threadIdx.x / kWave);
    const int64_t iend = ibase + nitems, stride = (int64_t)gridDim.x * NW;
    int64_t it = ibase + (int64_t)blockIdx.x * NW + w;
    HubItem nx{};   // the wave's next item, loaded one item ahead
    if (it < iend) nx = items[it];
    for (; it < iend; it += stride) {
        const HubItem m = nx;
        if (it + stride < iend) nx = items[it + stride];
        const int h = m.hub;
        if (hub_sel && !hub_sel[h]) {   // uniform per wave; the block merge skips it too
            if (lane == 0) big[it] = 0;
            continue;
        }
        const int p = m.part;
        const int64_t t0 = m.tile0, nt = m.ntiles, pstride = (int64_t)m.nparts + 1;
        // slice bounds of partition p in each tile, one tile per lane; those
        // of the first 64 tiles stay in registers for the pass below
        int lo0 = 0, len0 = 0, total = 0;
        for (int64_t c0 = t0; c0 < t0 + nt; c0 += kWave) {
            const int64_t t = c0 + lane;
            int lo = 0, len = 0;
            if (t < t0 + nt) {
                const int64_t ib = m.idx0 + (t - t0) * pstride + p;
                lo = tile_idx[ib];
                len = tile_idx[ib + 1] - lo;
            }
            if (c0 == t0) {
                lo0 = lo;
                len0 = len;
            }
            total += (int)wave_sum_u32((uint32_t)len);
        }
        if (total == 0) {
            if (lane == 0) big[it] = 0;
            continue;
        }
        bool over = total > kHmSmall;   // uniform
        int32_t ckv = -1;   // candidate label of this lane (-1: none)
        uint32_t ccv = 0, clv = 0;
        int ncand = 0;      // uniform
        for (int64_t c0 = t0; c0 < t0 + nt && !over; c0 += kWave) {
            const int64_t t = c0 + lane;
            int lo = lo0, len = len0;
            if (c0 != t0) {
                lo = 0;
                len = 0;
                if (t < t0 + nt) {
                    const int64_t ib = m.idx0 + (t - t0) * pstride + p;
                    lo = tile_idx[ib];
                    len = tile_idx[ib + 1] - lo;
                }
            }
            int st = len;   // exclusive scan of the slice lengths
#pragma unroll
            for (int sft = 1; sft < kWave; sft <<= 1) {
                const int y = __shfl_up(st, sft);
                if (lane >= sft) st += y;
            }
            const int ctot = __shfl(st, kWave - 1);
            st -= len;
            slo[w][lane] = lo;
            sstart[w][lane] = st;
            __builtin_amdgcn_wave_barrier();
            const int nc = (int)min((int64_t)kWave, t0 + nt - c0);
            for (int i0 = 0; i0 < ctot && !over; i0 += kWave) {
                const int i = i0 + lane;
                int32_t L = -1;
                uint32_t c = 0, last = 0;
                if (i < ctot) {
                    int a = 0, b = nc;   // last tile with start <= i
                    while (b - a > 1) {
                        const int mid = (a + b) >> 1;
                        if (sstart[w][mid] <= i) a = mid;
                        else b = mid;
                    }
                    const int64_t e = (c0 + a) * (int64_t)kHubTile + slo[w][a] + (i - sstart[w][a]);
                    const uint32_t ent = out_ent[e];
                    L = (int32_t)out_key[e];
                    c = hub_entry_count(ent);
                    last = (uint32_t)((c0 + a - t0) * (int64_t)kHubTile) + hub_entry_pos(ent);
                }
                unsigned long long ball = __ballot(L >= 0);
                while (ball) {
                    const int32_t Lk = __builtin_amdgcn_readlane(L, __ffsll((long long)ball) - 1);
                    const bool mine = L == Lk;
                    ball &= ~__ballot(mine);
                    const uint32_t cs = wave_sum_u32(mine ? c : 0u);
                    const uint32_t lm = wave_max(mine ? last : 0u);
                    const unsigned long long hit = __ballot(ckv == Lk);
                    if (hit) {
                        if (lane == __ffsll((long long)hit) - 1) {
                            ccv += cs;
                            clv = lm > clv ? lm : clv;
                        }
                    } else {
                        if (ncand == kWave) {
                            over = true;
                            break;
                        }
                        if (lane == ncand) {
                            ckv = Lk;
                            ccv = cs;
                            clv = lm;
                        }
                        ++ncand;
                    }
                }
            }
            __builtin_amdgcn_wave_barrier();   // the slice bounds are rewritten by the next chunk
        }
        if (lane == 0) big[it] = over ? 1 : 0;
        if (over) continue;   // the block merge takes it
        const unsigned long long best = wave_max(lane < ncand ? (((unsigned long long)ccv << 32) | clv) : 0ull);
        if (lane == 0) {
            if (best) atomicMax(&hub_best[h], best);
            if (hub_distinct) atomicAdd(&hub_distinct[h], (uint32_t)ncand);
        }
    }
}

// LIST: keep a list of the slots created (argmax and reset visit only them);
// without it they scan the whole table, which frees 16 KB of LDS per block.
// A block takes its flagged items one after another, so each item's global
// loads are issued ahead of their use: the next item's record when an item
// starts, the next item's first slice bounds once this item's entry loads
// are in flight, and the entries EPT per thread at a time.
// Without the list, three 512-thread blocks fit a CU's LDS; their 24 waves
// need at most 80 VGPRs each (six waves per SIMD).
template <int T, bool LIST>
__global__ void __launch_bounds__(T) __attribute__((amdgpu_waves_per_eu(!LIST && T == 512 ? 6 : 1)))
k_hub_merge(const HubItem *__restrict__ items, int64_t ibase, int64_t nitems, const int32_t *__restrict__ big,
            const int32_t *__restrict__ tile_idx, const uint32_t *__restrict__ out_key,
            const uint32_t *__restrict__ out_ent, unsigned long long *__restrict__ hub_best,
            const int32_t *__restrict__ stop, int32_t *__restrict__ err, uint32_t *__restrict__ hub_distinct,
            const int64_t *__restrict__ nsel) {
    constexpr int CAP = 1 << kHubMergeCapBits;
    constexpr int EPT = 4;   // entries per thread and pass
    __shared__ uint32_t skey[CAP];
    __shared__ uint32_t scnt[CAP];
    __shared__ uint32_t slast[CAP];
    __shared__ uint32_t slist[LIST ? CAP : 1];
    __shared__ uint32_t nlist;
    __shared__ int32_t slo[T], shi[T];
    __shared__ int32_t sscan_w[T / kWave];
    __shared__ int32_t stotal;
    __shared__ unsigned long long sred64[T / kWave];
    __shared__ int32_t sfull;   // a label found the table full in this round
    __shared__ uint32_t sbig[T / 32];   // flags of the current chunk of items
    if (*stop || (nsel && *nsel == 0)) return;
    for (int s = threadIdx.x; s < CAP; s += T) {
        skey[s] = kEmpty;
        scnt[s] = 0;
        slast[s] = 0;
    }
    if (threadIdx.x == 0) nlist = 0;
    if (threadIdx.x == 0) sfull = 0;
    // items go round-robin over the blocks (item ibase + blockIdx.x + j *
    // gridDim.x is the block's j-th), T of them per chunk; the flags of a
    // chunk are gathered into an LDS bit mask (one flag load per thread
    // instead of a dependent load per item).  The wave merge has cleared the
    // flag of every item outside this launch's hub selection.
    const int64_t iend = ibase + nitems;
    for (int64_t j0 = 0; ibase + blockIdx.x + j0 * gridDim.x < iend; j0 += T) {
    __syncthreads();   // the previous chunk's reads of sbig are done
    if (threadIdx.x < T / 32) sbig[threadIdx.x] = 0u;
    __syncthreads();
    {
        const int64_t ij = ibase + blockIdx.x + (j0 + threadIdx.x) * gridDim.x;
        if (ij < iend && big[ij]) atomicOr(&sbig[threadIdx.x / 32], 1u << (threadIdx.x % 32));
    }
    __syncthreads();
    // the first flagged item of the chunk at or after q (T: none); uniform,
    // so item records load through the scalar cache into SGPRs
    auto next_q = [&](int q) {
        while (q < T) {
            const uint32_t word = sbig[q / 32] >> (q % 32);
            if (word) return __builtin_amdgcn_readfirstlane(q + __ffs((int)word) - 1);
            q = (q | 31) + 1;
        }
        return T;
    };
    int q = next_q(0);
    HubItem nx{};
    if (q < T) nx = items[ibase + blockIdx.x + (j0 + q) * gridDim.x];
    int plo = 0, phi = 0;   // bounds of the next item's slice in tile tile0 + threadIdx.x
    bool pre = false;       // plo / phi hold them
    while (q < T) {
        const HubItem m = nx;
        const int qn = next_q(q + 1);
        if (qn < T) nx = items[ibase + blockIdx.x + (j0 + qn) * gridDim.x];
        bool have = pre;   // this item's first slice bounds are in plo / phi
        pre = false;
        const int h = m.hub, p = m.part;
        const int64_t t0 = m.tile0, nt = m.ntiles, pstride = (int64_t)m.nparts + 1;
        // A partition with more distinct labels than the table holds (an
        // adversarial label set, or a hub whose partitions were capped at
        // kHubMaxParts) is merged again in R rounds, round r taking the
        // labels whose sub-round hash is r; R doubles until every round fits.
        uint32_t R = 1;
        unsigned long long item_best = 0;
        uint32_t item_nl = 0;
        for (uint32_t r = 0; r < R;) {
        // tiles in chunks of T: the chunk's slice bounds are loaded one tile
        // per thread (no chain of dependent index loads), then the block
        // strides over the chunk's slices concatenated (an exclusive scan of
        // the slice lengths maps a flat index back to its tile), so every
        // thread has an entry even when each slice holds a handful
        for (int64_t c0 = t0; c0 < t0 + nt; c0 += T) {
            __syncthreads();
            const int64_t t = c0 + threadIdx.x;
            int lo = 0, len = 0;
            if (have) {   // uniform: the first chunk of the item's first round
                lo = plo;
                len = phi - plo;
                have = false;
            } else if (t < t0 + nt) {
                const int64_t ib = m.idx0 + (t - t0) * pstride + p;
                lo = tile_idx[ib];
                len = tile_idx[ib + 1] - lo;
            }
            const int start = block_exclusive_scan<T / kWave>(len, sscan_w);   // barriers inside
            slo[threadIdx.x] = lo;
            shi[threadIdx.x] = start;   // first flat index of tile c0 + threadIdx.x
            if (threadIdx.x == T - 1) stotal = start + len;
            __syncthreads();
            const int total = stotal;
            const int nc = (int)min((int64_t)T, t0 + nt - c0);
            for (int i0 = 0; i0 < total; i0 += T * EPT) {   // uniform trip count: ballots below
                uint32_t key[EPT], ent[EPT], tb[EPT];   // key kEmpty: no entry
#pragma unroll
                for (int k = 0; k < EPT; ++k) {
                    const int i = i0 + k * T + (int)threadIdx.x;
                    key[k] = kEmpty;
                    ent[k] = 0u;
                    tb[k] = 0u;
                    if (i < total) {
                        int a = 0, b = nc;   // last tile with start <= i
                        while (b - a > 1) {
                            const int mid = (a + b) >> 1;
                            if (shi[mid] <= i) a = mid;
                            else b = mid;
                        }
                        const int64_t e = (c0 + a) * (int64_t)kHubTile + slo[a] + (i - shi[a]);
                        key[k] = out_key[e];
                        ent[k] = out_ent[e];
                        tb[k] = (uint32_t)((c0 + a - t0) * (int64_t)kHubTile);
                    }
                }
                if (!pre && qn < T) {   // uniform: once per item, behind its first entry loads
                    plo = 0;
                    phi = 0;
                    if ((int64_t)threadIdx.x < (int64_t)nx.ntiles) {
                        const int64_t ib =
                            nx.idx0 + (int64_t)threadIdx.x * ((int64_t)nx.nparts + 1) + nx.part;
                        plo = tile_idx[ib];
                        phi = tile_idx[ib + 1];
                    }
                    pre = true;
                }
#pragma unroll
                for (int k = 0; k < EPT; ++k) {
                    bool created = false;
                    uint32_t sl = 0;
                    if (key[k] != kEmpty && (R == 1 || sub_round(key[k], R) == r)) {
                        sl = lds_try_insert<kHubMergeCapBits>(skey, key[k], &created);
                        if (sl == kTableFull) {
                            sfull = 1;
                        } else {
                            atomicAdd(&scnt[sl], hub_entry_count(ent[k]));
                            atomicMax(&slast[sl], tb[k] + hub_entry_pos(ent[k]));
                        }
                    }
                    // one list append per wave
                    const unsigned long long ball = LIST ? __ballot(created) : 0ull;
                    if (ball) {
                        const int leader = __ffsll((long long)ball) - 1;
                        uint32_t base = 0;
                        if (lane_id() == leader) base = atomicAdd(&nlist, (uint32_t)__popcll(ball));
                        base = (uint32_t)__shfl((int)base, leader);
                        if (created) slist[base + __popcll(ball & ((1ull << lane_id()) - 1ull))] = sl;
                    }
                }
            }
        }
        __syncthreads();
        // argmax and reset over the slots created (cost follows the number of
        // distinct labels, not the table size)
        unsigned long long best = 0;
        uint32_t nl = 0;
        if constexpr (LIST) {
            nl = nlist;
            for (uint32_t i = threadIdx.x; i < nl; i += T) {
                const uint32_t sl = slist[i];
                const unsigned long long pk = ((unsigned long long)scnt[sl] << 32) | slast[sl];
                best = pk > best ? pk : best;
                skey[sl] = kEmpty;
                scnt[sl] = 0;
                slast[sl] = 0;
            }
        } else {
            for (int sl = threadIdx.x; sl < CAP; sl += T)
                if (skey[sl] != kEmpty) {
                    const unsigned long long pk = ((unsigned long long)scnt[sl] << 32) | slast[sl];
                    best = pk > best ? pk : best;
                    ++nl;
                    skey[sl] = kEmpty;
                    scnt[sl] = 0;
                    slast[sl] = 0;
                }
        }
        __syncthreads();
        const bool full = sfull != 0;   // uniform: read after the barrier
        best = block_max<T / kWave>(best, sred64);   // barriers inside: every read above is done
        if constexpr (!LIST) {
            if (hub_distinct) nl = block_sum<T / kWave>(nl, (uint32_t *)sscan_w);   // uniform branch
        }
        if (threadIdx.x == 0) {
            nlist = 0;
            sfull = 0;
        }
        if (full) {   // start the item over in twice as many rounds
            if (R >= (1u << kHubMergeCapBits)) {
                if (threadIdx.x == 0) atomicOr(err, 1);
                break;
            }
            R *= 2;
            r = 0;
            item_best = 0;
            item_nl = 0;
            continue;
        }
        item_best = best > item_best ? best : item_best;
        item_nl += nl;
        ++r;
        }
        if (threadIdx.x == 0 && item_best) atomicMax(&hub_best[h], item_best);
        if (threadIdx.x == 0 && hub_distinct) atomicAdd(&hub_distinct[h], item_nl);
        q = qn;
    }
    }
}

// Hubs: finalize.  The winning label is the one at position `last` of the
// row; resets hub_best for the next iteration.
template <typename E>
__global__ void __launch_bounds__(256)
k_hub_final(const E *__restrict__ off, const int32_t *__restrict__ nbr, const int32_t *__restrict__ hub_vid,
            int64_t nhubs, unsigned long long *__restrict__ hub_best, const int32_t *__restrict__ hub_sel,
            uint32_t *__restrict__ hub_distinct, int32_t *__restrict__ distinct, const int32_t *__restrict__ prev,
            int32_t *__restrict__ next, unsigned long long *__restrict__ counter,
            const int32_t *__restrict__ stop, const int64_t *__restrict__ nsel) {
    __shared__ uint32_t sred[256 / kWave];
    if (*stop || (nsel && *nsel == 0)) return;
    uint32_t changed = 0;
    for (int64_t h = (int64_t)blockIdx.x * 256 + threadIdx.x; h < nhubs; h += (int64_t)gridDim.x * 256) {
        if (hub_sel && !hub_sel[h]) continue;
        const int64_t v = hub_vid[h];
        const unsigned long long pk = hub_best[h];
        const uint32_t last = (uint32_t)pk;
        const int32_t w = prev[nbr[(int64_t)off[v] + last]];
        next[v] = w;
        changed += (w != prev[v]);
        hub_best[h] = 0ull;
        if (hub_distinct) {
            if (distinct) distinct[v] = (int32_t)hub_distinct[h];
            hub_distinct[h] = 0u;
        }
    }
    block_count_add<4>(changed, counter, sred);
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
static double class_bytes(const BucketRange &r) {
    // offsets, neighbour ids, gathered labels, own previous label, next label
    return (double)sizeof(E) * (double)(r.nv + 1) + 8.0 * (double)r.edges + 8.0 * (double)r.nv;
}

template <int G, typename E>
static void launch_group(galp_graph *g, const IterView &w, int bid, const int32_t *prev, int32_t *next,
                         int blocks_per_cu = 8) {
    const BucketRange &r = w.cls[bid];
    if (!r.nv) return;
    PendingLaunch pl;
    prof_begin(g, kBucketName[bid], class_bytes<E>(r), &pl);
    const int64_t per_block = 256 / G;
    const int64_t blocks = (r.nv + per_block - 1) / per_block;
    // learned dispatch: the sweep-only resolver in iterations where most big
    // vertices are predicted high-distinct
    const bool learned = learned_now(g);
    const int64_t *hd = learned ? g->learn->d_plist_n + kCtrHdAll : nullptr;
    const int64_t hd_total = learned ? g->learn->front_n : 0;   // big vertices of this dispatch
    if (batch_lists(g))   // the running graphs' vertices of the class
        k_group<G, HashGroupResolve, HashGroupSweep, true, E><<<grid_for(g, blocks, blocks_per_cu), 256, 0, g->stream>>>(
            (const E *)w.off, w.nbr, prev, next, g->batch->d_alist[bid], 0, r.nv, g->batch->d_alen + bid,
            g->d_counter, g->d_stop, hd, hd_total);
    else if (g->d_vlist)
        k_group<G, HashGroupResolve, HashGroupSweep, true, E><<<grid_for(g, blocks, blocks_per_cu), 256, 0, g->stream>>>(
            (const E *)w.off, w.nbr, prev, next, g->d_vlist, r.v0, r.nv, nullptr, g->d_counter, g->d_stop, hd,
            hd_total);
    else
        k_group<G, HashGroupResolve, HashGroupSweep, false, E><<<grid_for(g, blocks, blocks_per_cu), 256, 0,
                                                                 g->stream>>>(
            (const E *)w.off, w.nbr, prev, next, nullptr, r.v0, r.nv, nullptr, g->d_counter, g->d_stop, hd, hd_total);
    prof_end(g, &pl);
}

template <int T, int EPT, int CAPBITS, typename E>
static void launch_lds(galp_graph *g, const IterView &w, int bid, const int32_t *prev, int32_t *next,
                       int blocks_per_cu) {
    const BucketRange &r = w.cls[bid];
    if (!r.nv) return;
    PendingLaunch pl;
    prof_begin(g, kBucketName[bid], class_bytes<E>(r), &pl);
    int32_t *dist = distinct_out(g);
    if (batch_lists(g))   // the running graphs' vertices of the class
        k_hash_lds<T, EPT, CAPBITS, E, 3><<<grid_for(g, r.nv, blocks_per_cu), T, 0, g->stream>>>(
            (const E *)w.off, w.nbr, prev, next, g->batch->d_alist[bid], 0, r.nv, g->batch->d_alen + bid, nullptr,
            nullptr, g->d_counter, g->d_stop, g->d_err, dist);
    else if (learned_now(g))   // this iteration's two lists for the class
        k_hash_lds<T, EPT, CAPBITS, E, 3><<<grid_for(g, r.nv, blocks_per_cu), T, 0, g->stream>>>(
            (const E *)w.off, w.nbr, prev, next, g->learn->plist[bid], 0, r.nv, g->learn->d_plist_n + bid,
            g->learn->plist[NUM_BUCKETS + bid], g->learn->d_plist_n + kCtrHd + bid, g->d_counter, g->d_stop,
            g->d_err, dist);
    else
        k_hash_lds<T, EPT, CAPBITS, E, 3><<<grid_for(g, r.nv, blocks_per_cu), T, 0, g->stream>>>(
            (const E *)w.off, w.nbr, prev, next, g->d_vlist, r.v0, r.nv, nullptr, nullptr, nullptr, g->d_counter,
            g->d_stop, g->d_err, dist);
    prof_end(g, &pl);
}

// Hub rows: tile (TT threads per 2048-edge tile), merge (TM threads per
// (hub, partition) item), final.  The tile holds 48 KB of LDS per block and
// the merge 68 KB with its slot list, 52 KB without (three or two blocks per
// CU), so block size and list set the waves per CU.
template <typename E, int TT, int TM, bool LIST>
static void launch_hubs(galp_graph *g, const IterView &w, const int32_t *prev, int32_t *next) {
    if (w.h1 <= w.h0) return;
    const int64_t nh = w.h1 - w.h0, nt = w.t1 - w.t0, ni = w.i1 - w.i0;
    double hub_edges = 0.0;
    if (w.h0 == 0 && w.h1 == g->nhubs) hub_edges = (double)g->bucket[B_HUB].edges;
    const int32_t *sel = learned_now(g) ? g->learn->d_hub_sel : nullptr;
    uint32_t *hdist = distinct_out(g) ? g->learn->d_hub_distinct : nullptr;
    const int64_t *nsel = sel ? g->learn->d_plist_n + kCtrPlainHubs : nullptr;
    const int64_t *nhd = sel ? g->learn->d_plist_n + kCtrHd + B_HUB : nullptr;
    PendingLaunch pl;
    prof_begin(g, "hub_tile", 8.0 * hub_edges, &pl);
    // tiles [t0, t1): the pipeline starts at t0 + blockIdx.x, stops at t1
    k_hub_tile<TT, E, false><<<grid_for(g, nt, 3), TT, 0, g->stream>>>(
        (const E *)w.off, w.nbr, prev, g->d_tile_hub, g->d_hub_tile0, g->d_hub_vid, g->d_hub_nparts,
        g->d_tile_idx_off, w.t0, w.t1, g->d_out_key, g->d_out_ent, g->d_tile_idx, g->d_stop,
        g->d_err, sel, nsel, sel ? g->learn->d_hub_hd : nullptr, nhd);
    if (sel)   // learned dispatch: the hubs predicted high-distinct
        k_hub_tile<TT, E, true><<<grid_for(g, nt, 3), TT, 0, g->stream>>>(
            (const E *)w.off, w.nbr, prev, g->d_tile_hub, g->d_hub_tile0, g->d_hub_vid, g->d_hub_nparts,
            g->d_tile_idx_off, w.t0, w.t1, g->d_out_key, g->d_out_ent, g->d_tile_idx, g->d_stop,
            g->d_err, sel, nsel, g->learn->d_hub_hd, nhd);
    prof_end(g, &pl);
    prof_begin(g, "hub_merge", 0.0, &pl);
    k_hub_merge_small<<<grid_for(g, (ni + kHmSmallT / kWave - 1) / (kHmSmallT / kWave), 8), kHmSmallT, 0,
                        g->stream>>>(g->d_item_meta, w.i0, ni, g->d_tile_idx, g->d_out_key, g->d_out_ent,
                                     g->d_hub_best, sel, hdist, nsel, g->d_merge_big, g->d_stop);
    k_hub_merge<TM, LIST><<<grid_for(g, ni, TM == 256 || !LIST ? 3 : 2), TM, 0, g->stream>>>(
        g->d_item_meta, w.i0, ni, g->d_merge_big, g->d_tile_idx, g->d_out_key, g->d_out_ent, g->d_hub_best,
        g->d_stop, g->d_err, hdist, nsel);
    prof_end(g, &pl);
    prof_begin(g, "hub_final", (double)sizeof(E) * 2.0 * (double)nh + 16.0 * (double)nh, &pl);
    k_hub_final<E><<<grid_for(g, (nh + 255) / 256, 4), 256, 0, g->stream>>>(
        (const E *)w.off, w.nbr, g->d_hub_vid + w.h0, nh, g->d_hub_best + w.h0, sel ? sel + w.h0 : nullptr,
        hdist ? hdist + w.h0 : nullptr, distinct_out(g), prev, next, g->d_counter,
        g->d_stop, nsel);
    prof_end(g, &pl);
}

// The kernel configuration of each degree class: option 0 is the default,
// 1 and 2 the alternatives the autotuner times (workgroup size x slots per
// lane, and blocks per CU for the grid; for lane-group classes the grid
// only; for hubs the tile and merge workgroup sizes).  Every option computes
// the same labels.
template <typename E>
static void launch_class(galp_graph *g, const IterView &w, int b, const int32_t *prev, int32_t *next) {
    const int c = g->tune[b];
    switch (b) {
    case B_HUB:
        if (c == 1) launch_hubs<E, 512, 512, true>(g, w, prev, next);
        else if (c == 2) launch_hubs<E, 256, 256, true>(g, w, prev, next);
        else launch_hubs<E, 512, 512, false>(g, w, prev, next);
        break;
    case B_B4K:
        if (c == 1) launch_lds<1024, 4, 13, E>(g, w, b, prev, next, 1);
        else if (c == 2) launch_lds<256, 16, 13, E>(g, w, b, prev, next, 4);
        else launch_lds<512, 8, 13, E>(g, w, b, prev, next, 2);
        break;
    case B_B2K:
        if (c == 1) launch_lds<256, 8, 12, E>(g, w, b, prev, next, 4);
        else if (c == 2) launch_lds<128, 16, 12, E>(g, w, b, prev, next, 8);
        else launch_lds<512, 4, 12, E>(g, w, b, prev, next, 3);
        break;
    case B_B1K:
        if (c == 1) launch_lds<128, 8, 11, E>(g, w, b, prev, next, 8);
        else if (c == 2) launch_lds<64, 16, 11, E>(g, w, b, prev, next, 16);
        else launch_lds<256, 4, 11, E>(g, w, b, prev, next, 6);
        break;
    case B_W512:
        if (c == 1) launch_lds<128, 4, 10, E>(g, w, b, prev, next, 12);
        else if (c == 2) launch_lds<256, 2, 10, E>(g, w, b, prev, next, 8);
        else launch_lds<64, 8, 10, E>(g, w, b, prev, next, 16);
        break;
    case B_W256:
        if (c == 1) launch_lds<128, 2, 9, E>(g, w, b, prev, next, 16);
        else if (c == 2) launch_lds<64, 4, 9, E>(g, w, b, prev, next, 12);
        else launch_lds<64, 4, 9, E>(g, w, b, prev, next, 24);
        break;
    case B_W128:
        if (c == 1) launch_lds<128, 1, 8, E>(g, w, b, prev, next, 16);
        else if (c == 2) launch_lds<64, 2, 8, E>(g, w, b, prev, next, 16);
        else launch_lds<64, 2, 8, E>(g, w, b, prev, next, 32);
        break;
    default: {   // lane groups: blocks per CU of the persistent grid
        // the defaults are the autotuner's usual picks on RMAT, LFR and
        // batched planted partitions: 4 per CU for G4, 16 for G64, else 8
        const int def = b == B_G4 ? 4 : b == B_G64 ? 16 : 8;
        const int alt1 = b == B_G64 ? 8 : 16, alt2 = b == B_G4 ? 8 : 4;
        const int bpc = c == 1 ? alt1 : c == 2 ? alt2 : def;
        if (b == B_G64) launch_group<64, E>(g, w, b, prev, next, bpc);
        else if (b == B_G32) launch_group<32, E>(g, w, b, prev, next, bpc);
        else if (b == B_G16) launch_group<16, E>(g, w, b, prev, next, bpc);
        else if (b == B_G8) launch_group<8, E>(g, w, b, prev, next, bpc);
        else launch_group<4, E>(g, w, b, prev, next, bpc);
    }
    }
}

int launch_hash_class(galp_graph *g, int b) {
    const IterView w = full_view(g);
    const int32_t *prev = g->d_lab[g->cur];
    int32_t *next = g->d_lab[g->cur ^ 1];
    if (g->off64) launch_class<int64_t>(g, w, b, prev, next);
    else launch_class<int32_t>(g, w, b, prev, next);
    const hipError_t e = hipGetLastError();
    return e == hipSuccess ? GALP_OK : fail(g->ctx, GALP_EHIP, std::string("class launch: ") + hipGetErrorString(e));
}

template <typename E>
static int hash_iteration(galp_graph *g, const IterView &w) {
    const int32_t *prev = g->d_lab[g->cur];
    int32_t *next = g->d_lab[g->cur ^ 1];
    launch_class<E>(g, w, B_HUB, prev, next);   // largest work first
    if (g->opts.variant == 0) {
        // -l 0: no degree classes below the hubs -- every row through the
        // block-per-vertex table kernel, as the reference's naive variant
        // gives every vertex the same treatment
        for (int b = B_B4K; b >= B_G4; --b) launch_lds<512, 8, 13, E>(g, w, b, prev, next, 2);
        hipError_t e = hipGetLastError();
        if (e != hipSuccess) return fail(g->ctx, GALP_EHIP, std::string("hash kernel launch: ") + hipGetErrorString(e));
        return GALP_OK;
    }
    for (int b = B_B4K; b >= B_G4; --b) launch_class<E>(g, w, b, prev, next);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return fail(g->ctx, GALP_EHIP, std::string("hash kernel launch: ") + hipGetErrorString(e));
    return GALP_OK;
}

int launch_hash_iteration(galp_graph *g, const IterView &w) {
    return g->off64 ? hash_iteration<int64_t>(g, w) : hash_iteration<int32_t>(g, w);
}

// One in-core iteration.  Every class kernel reads prev and writes next only
// at its own rows, so the classes are independent once their vertex lists
// are final.  The lane-group classes (degree <= 64) go to side stream 0: at
// the start of the iteration, or under learned dispatch after the dispatch
// (which sets their resolver flag) and the candidate kernel (one
// 1024-thread block per CU holding 152 KB of LDS, launched first so that it
// finds its CUs free).  The LDS classes go to side stream 1 after the
// candidate merge, which appends the rows it gives up to their lists; the
// hub path stays on the graph's stream beside them.
template <typename E>
static int hash_iteration_incore(galp_graph *g) {
    const IterView w = full_view(g);
    const bool learned = learned_now(g);
    if (!concurrent_classes(g)) {
        if (learned) {
            const int rc = launch_learned_front(g, true);
            if (rc != GALP_OK) return rc;
        }
        return hash_iteration<E>(g, w);
    }
    const int32_t *prev = g->d_lab[g->cur];
    int32_t *next = g->d_lab[g->cur ^ 1];
    int rc = learned ? launch_learned_front(g, false) : GALP_OK;
    hipStream_t ms = nullptr;
    if (rc == GALP_OK) rc = side_begin(g, 0, &ms);
    if (rc != GALP_OK) return rc;
    else
        for (int b = B_G64; b >= B_G4; --b) launch_class<E>(g, w, b, prev, next);
    rc = side_end(g, 0, ms);
    if (rc == GALP_OK && learned) rc = launch_learned_merge(g);
    if (rc == GALP_OK) rc = side_begin(g, 1, &ms);
    if (rc != GALP_OK) return rc;
    for (int b = B_B4K; b >= B_W128; --b) launch_class<E>(g, w, b, prev, next);
    rc = side_end(g, 1, ms);
    if (rc != GALP_OK) return rc;
    launch_class<E>(g, w, B_HUB, prev, next);
    if ((rc = side_join(g, 0)) != GALP_OK || (rc = side_join(g, 1)) != GALP_OK) return rc;
    const hipError_t e = hipGetLastError();
    return e == hipSuccess ? GALP_OK : fail(g->ctx, GALP_EHIP, std::string("hash kernel launch: ") + hipGetErrorString(e));
}

int launch_hash_iteration_incore(galp_graph *g) {
    return g->off64 ? hash_iteration_incore<int64_t>(g) : hash_iteration_incore<int32_t>(g);
}

int launch_iter_end(galp_graph *g) {
    k_iter_end<<<1, 1, 0, g->stream>>>(g->d_counter, g->d_stop, g->d_iter_counts, g->d_iters_done,
                                       g->iter % kMaxIters, (double)g->n * 1e-5);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return fail(g->ctx, GALP_EHIP, std::string("iter_end launch: ") + hipGetErrorString(e));
    return GALP_OK;
}

}  // namespace galp
