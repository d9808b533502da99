// lp_sort.hip -- method 0 (sort / segmented reduce) on gfx950.
//
// Replaces the data-parallel-primitive pipeline of the reference:
//   DPPBase::perform_lp          label_propagation/src/dpp/dpp.cuh:104-138
//   gather_labels                src/dpp/kernel.cuh:27-37
//   SegmentedSorter::run         src/common/segmented_sort.cuh:35-64 (moderngpu)
//   find_boundaries .. compute_label_weights   src/dpp/kernel.cuh:41-92
//   SegmentedReducer::apply      src/common/segmented_reduce.cuh:559-626
//   update_labels                src/dpp/kernel.cuh:95-119
//   compute_max_labels (li)      src/dpp/kernel.cuh:135-206, incore_li.cuh:43-87
// Design (not a translation): the reference sorts all m gathered labels with
// a global segmented merge sort and then runs five m-sized passes (flags,
// scan, scatter, weights, segmented arg-max).  Here each degree class sorts
// where its rows fit:
//   - degree <= 64: a bitonic sort of (label, position) across the group's
//     lanes (register shuffles), then run ends and run starts from ballots;
//   - 65..1024: a bitonic sort of (label, position) in one wave's
//     registers per vertex (2..16 keys per lane; lane exchanges by DPP,
//     ds_swizzle and permutes), the LDS bitonic network as the autotuner's
//     alternative; 1025..4096: an LDS LSD radix sort of the labels (8-bit digits, stable
//     multisplit by ballots, positions as payload); run lengths by binary
//     search for each run's first element;
//   - hubs (> 4096): each 2048-edge tile sorted in LDS and run-length encoded
//     into (label, count, last) entries bucketed by label partition, then one
//     wave (at most 1024 entries) or one block per (hub, partition) sorts that
//     partition's entries of every tile and reduces each run; one 64-bit
//     atomicMax of (count, last) per item.
// The sort key is (label, position in the row), so every sort is stable in
// the row order and the last element of a run holds the run's largest
// position: the winner argmax (count, last) equals the hash method's.

#include <algorithm>
#include <vector>

#include "device_common.hpp"
#include "group_kernel.hpp"
#include "internal.hpp"
#include "pipeline.hpp"

namespace galp {

// ---------------------------------------------------------------------------
// Degree <= G.  Bitonic sort (Batcher 1968) of the G keys (label, lane)
// across the group with lane exchanges; -1 (no edge) sorts last as
// 0xffffffff.  Then lane p is a run end if the next key differs, a run start
// if the previous one does; the run's length comes from the highest start
// at or below p (one ballot), its last position is the sorted lane's
// original index.
struct SortGroupResolve {
    template <int G>
    __device__ __forceinline__ static uint32_t best(int32_t L) {
        const int lane = lane_id();
        const int p = lane & (G - 1);
        uint32_t key = L >= 0 ? (uint32_t)L : 0xffffffffu;
        uint32_t pos = (uint32_t)p;
#pragma unroll
        for (int k = 2; k <= G; k <<= 1) {
#pragma unroll
            for (int j = k >> 1; j > 0; j >>= 1) {
                const uint32_t ok = (uint32_t)__shfl_xor((int)key, j, G);
                const uint32_t op = (uint32_t)__shfl_xor((int)pos, j, G);
                const bool up = (p & k) == 0;
                const bool lower = (p & j) == 0;
                const bool other_less = ok < key || (ok == key && op < pos);
                const bool take = (lower == up) ? other_less : !other_less;
                key = take ? ok : key;
                pos = take ? op : pos;
            }
        }
        const uint32_t knext = (uint32_t)__shfl_down((int)key, 1, G);
        const uint32_t kprev = (uint32_t)__shfl_up((int)key, 1, G);
        const bool valid = key != 0xffffffffu;
        const bool is_end = valid && (p == G - 1 || knext != key);
        const bool is_start = valid && (p == 0 || kprev != key);
        const unsigned long long starts = __ballot(is_start);
        const unsigned long long upto = lane == 63 ? ~0ull : ((2ull << lane) - 1ull);
        const int s = 63 - __clzll((long long)((starts & upto) | 1ull));   // my run's start (in my group when valid)
        const uint32_t cnt = (uint32_t)(lane - s + 1);
        return group_max<G>(is_end ? ((cnt << 8) | pos) : 0u);
    }
};

// ---------------------------------------------------------------------------
// Rows of the smaller classes (65..1024): one block of T threads per vertex,
// a bitonic network in LDS (fewer, cheaper steps than radix passes at these
// sizes).  The (label,
// position) keys go to LDS, padded with all-ones to the next power of two,
// and a bitonic network sorts them; each run end finds its run's first
// element by binary search (lower bound of its label), giving the count; the
// block max of (count << 32 | last) picks the winner, whose label the
// thread that found it writes.
template <int T, int EPT, typename E>
struct LdsBitonicPolicy {
    static constexpr int N = T * EPT;
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
    unsigned long long *skey;
    unsigned long long *sred64;
    int32_t *err;
    uint32_t changed;

    __device__ Item item(int64_t i) const {
        Item it{false, 0, 0, 0, 0};
        if (i < nv) {
            it.v = vlist ? (int64_t)vlist[v0 + i] : v0 + i;
            it.b = (int64_t)off[it.v];
            it.len = (int)((int64_t)off[it.v + 1] - it.b);
            it.old = prev[it.v];
            it.valid = true;
            if (it.len > N) {   // never planned into this class
                if (threadIdx.x == 0) atomicOr(err, 8);
                it.len = 0;
            }
        }
        return it;
    }
    __device__ int64_t edge0(const Item &it) const { return it.b; }

    __device__ void process(const Item &it, const int32_t (&lab)[EPT]) {
        if (it.len == 0) return;
        const int len = it.len;
        int n2 = 1;
        while (n2 < len) n2 <<= 1;
        const int tid = (int)threadIdx.x;
        __syncthreads();   // the previous vertex's readers are done
#pragma unroll
        for (int k = 0; k < EPT; ++k) {
            const int pos = k * T + tid;
            if (pos < n2)
                skey[pos] = pos < len ? (((unsigned long long)(uint32_t)lab[k] << 32) | (uint32_t)pos) : ~0ull;
        }
        __syncthreads();
        for (int kk = 2; kk <= n2; kk <<= 1) {
            for (int j = kk >> 1; j > 0; j >>= 1) {
                for (int q = tid; q < (n2 >> 1); q += T) {
                    const int i = ((q & ~(j - 1)) << 1) | (q & (j - 1));
                    const int l = i + j;
                    const unsigned long long a = skey[i], b = skey[l];
                    const bool up = (i & kk) == 0;
                    if ((a > b) == up) {
                        skey[i] = b;
                        skey[l] = a;
                    }
                }
                __syncthreads();
            }
        }
        unsigned long long best = 0;
        int32_t bestlab = 0;
        for (int q = tid; q < len; q += T) {
            const uint32_t key = (uint32_t)(skey[q] >> 32);
            const bool end = q == len - 1 || (uint32_t)(skey[q + 1] >> 32) != key;
            if (end) {
                int lo = 0, hi = q;
                while (lo < hi) {
                    const int mid = (lo + hi) >> 1;
                    if ((uint32_t)(skey[mid] >> 32) < key) lo = mid + 1;
                    else hi = mid;
                }
                const unsigned long long val =
                    ((unsigned long long)(q - lo + 1) << 32) | (uint32_t)skey[q];
                if (val > best) {
                    best = val;
                    bestlab = (int32_t)key;
                }
            }
        }
        const unsigned long long top = block_max<T / kWave>(best, sred64);
        if (best == top) {   // unique: positions are distinct
            next[it.v] = bestlab;
            changed += (bestlab != it.old);
        }
    }
};

template <int T, int EPT, typename E>
__global__ void __launch_bounds__(T)
k_sort_bitonic(const E *__restrict__ off, const int32_t *__restrict__ nbr, const int32_t *__restrict__ prev,
           int32_t *__restrict__ next, const int32_t *__restrict__ vlist, int64_t v0, int64_t nv,
           unsigned long long *__restrict__ counter, const int32_t *__restrict__ stop, int32_t *__restrict__ err) {
    __shared__ unsigned long long skey[T * EPT];
    __shared__ unsigned long long sred64[T / kWave];
    __shared__ uint32_t sred[T / kWave];
    if (*stop) return;
    LdsBitonicPolicy<T, EPT, E> pol{off, prev, next, vlist, v0, nv, skey, sred64, err, 0u};
    pipelined<T, EPT>(pol, nbr, prev, (int64_t)blockIdx.x, (int64_t)gridDim.x, (int)threadIdx.x);
    __syncthreads();
    block_count_add<T / kWave>(pol.changed, counter, sred);
}

// ---------------------------------------------------------------------------
// Rows of 65..4096 edges sorted in registers, one block of NW waves per
// vertex.  Thread t holds elements t*EPT .. t*EPT+EPT-1 of a bitonic network
// (Batcher 1968) over N = 64*NW*EPT keys (label << 32 | position in the
// row), padded with all-ones.  A compare-exchange at distance j < EPT stays
// in the thread; up to 64*EPT it pairs lane l with lane l ^ (j / EPT) of the
// same wave through DPP moves (xor 1, 2, 4, 8), ds_swizzle (16) or a permute
// (32); only the distances that cross waves (the last one or two of each of
// the final merge stages) go through LDS.  Sorted, a run start is an
// element whose predecessor holds another label; an exclusive max-scan of
// each thread's last start gives every run end its run's first index, hence
// the count; the run's last element holds its largest position (positions
// break ties in the key).  The block max of (count << 32 | last) picks the
// winner.
template <int X>
__device__ __forceinline__ unsigned long long lane_xor_u64(unsigned long long v) {
    const uint32_t lo = (uint32_t)lane_xor<X>((int32_t)(uint32_t)v);
    const uint32_t hi = (uint32_t)lane_xor<X>((int32_t)(uint32_t)(v >> 32));
    return ((unsigned long long)hi << 32) | lo;
}

// Compare-exchange steps j = J, J/2, .., 1 of merge stage KK; sx: NW*64*EPT
// words of LDS for the steps that cross waves.
template <int NW, int EPT, int KK, int J>
__device__ __forceinline__ void reg_bitonic_steps(unsigned long long (&key)[EPT], int tid,
                                                  unsigned long long *sx) {
    constexpr int T = NW * kWave;
    if constexpr (J < EPT) {
#pragma unroll
        for (int k = 0; k < EPT; ++k) {
            if ((k & J) == 0) {
                const int q = k | J;
                const bool up = ((tid * EPT + k) & KK) == 0;
                const unsigned long long a = key[k], b = key[q];
                const bool sw = (a > b) == up;
                key[k] = sw ? b : a;
                key[q] = sw ? a : b;
            }
        }
    } else {
        constexpr int X = J / EPT;   // the partner thread is tid ^ X
        const bool lower = (tid & X) == 0;
        if constexpr (X >= kWave) {
#pragma unroll
            for (int k = 0; k < EPT; ++k) sx[k * T + tid] = key[k];
            __syncthreads();
        }
#pragma unroll
        for (int k = 0; k < EPT; ++k) {
            const unsigned long long a = key[k];
            unsigned long long p;
            if constexpr (X >= kWave) p = sx[k * T + (tid ^ X)];
            else p = lane_xor_u64<X>(a);
            const bool up = ((tid * EPT + k) & KK) == 0;
            const unsigned long long lo = a < p ? a : p, hi = a < p ? p : a;
            key[k] = lower == up ? lo : hi;
        }
        if constexpr (X >= kWave) __syncthreads();   // read before the next write
    }
    if constexpr (J > 1) reg_bitonic_steps<NW, EPT, KK, J / 2>(key, tid, sx);
}

template <int NW, int EPT, int KK = 2>
__device__ __forceinline__ void reg_bitonic(unsigned long long (&key)[EPT], int tid, unsigned long long *sx) {
    reg_bitonic_steps<NW, EPT, KK, KK / 2>(key, tid, sx);
    if constexpr (KK < NW * kWave * EPT) reg_bitonic<NW, EPT, KK * 2>(key, tid, sx);
}

template <int NW, int EPT, typename E>
struct RegBitonicPolicy {
    static constexpr int T = NW * kWave;
    static constexpr int N = T * EPT;
    using Item = typename LdsBitonicPolicy<T, EPT, E>::Item;
    const E *off;
    const int32_t *prev;
    int32_t *next;
    const int32_t *vlist;
    int64_t v0, nv;
    unsigned long long *sx;   // N words (NW > 1)
    uint32_t *sedge;          // 2*T: each thread's first and last label (NW > 1)
    int *sw;                  // NW
    unsigned long long *sred64;   // NW
    int32_t *err;
    uint32_t changed;

    __device__ Item item(int64_t i) const {
        Item it{false, 0, 0, 0, 0};
        if (i < nv) {
            it.v = vlist ? (int64_t)vlist[v0 + i] : v0 + i;
            it.b = (int64_t)off[it.v];
            it.len = (int)((int64_t)off[it.v + 1] - it.b);
            it.old = prev[it.v];
            it.valid = true;
            if (it.len > N) {   // never planned into this class
                if (threadIdx.x == 0) atomicOr(err, 8);
                it.len = 0;
            }
        }
        return it;
    }
    __device__ int64_t edge0(const Item &it) const { return it.b; }

    __device__ void process(const Item &it, const int32_t (&lab)[EPT]) {
        if (it.len == 0) return;   // uniform per block
        const int tid = (int)threadIdx.x;
        const int lane = lane_id();
        unsigned long long key[EPT];
#pragma unroll
        for (int k = 0; k < EPT; ++k) {
            const int pos = k * T + tid;   // the pipeline's slot of lab[k]
            key[k] = pos < it.len ? (((unsigned long long)(uint32_t)lab[k] << 32) | (uint32_t)pos) : ~0ull;
        }
        reg_bitonic<NW, EPT>(key, tid, sx);
        constexpr uint32_t kPad = 0xffffffffu;   // labels are < 2^31
        uint32_t L[EPT];
#pragma unroll
        for (int k = 0; k < EPT; ++k) L[k] = (uint32_t)(key[k] >> 32);
        uint32_t before, after;   // thread tid - 1's last element, tid + 1's first
        if constexpr (NW == 1) {
            before = (uint32_t)__shfl_up((int)L[EPT - 1], 1);
            // every lane shuffles (a shuffle under a condition reads 0 from
            // lanes that skipped it), then the last lane takes the pad
            const uint32_t dn = (uint32_t)__shfl_down((int)L[0], 1);
            after = lane == kWave - 1 ? kPad : dn;
        } else {
            sedge[tid] = L[0];
            sedge[T + tid] = L[EPT - 1];
            __syncthreads();
            before = tid ? sedge[T + tid - 1] : kPad;
            after = tid < T - 1 ? sedge[tid + 1] : kPad;
        }
        int ls = -1;   // this thread's last run start
#pragma unroll
        for (int k = 0; k < EPT; ++k) {
            const int i = tid * EPT + k;
            const uint32_t pl = k ? L[k - 1] : before;
            if (L[k] != kPad && (i == 0 || pl != L[k])) ls = i;
        }
        int inc = ls;   // inclusive max-scan over the wave's lanes
#pragma unroll
        for (int sft = 1; sft < kWave; sft <<= 1) {
            const int y = __shfl_up(inc, sft);
            if (lane >= sft) inc = inc > y ? inc : y;
        }
        int cur = __shfl_up(inc, 1);   // start of the run open at my first element
        if (lane == 0) cur = -1;
        if constexpr (NW > 1) {   // and over the earlier waves
            const int w = tid / kWave;
            if (lane == kWave - 1) sw[w] = inc;
            __syncthreads();
#pragma unroll
            for (int q = 0; q < NW; ++q)
                if (q < w) cur = cur > sw[q] ? cur : sw[q];
        }
        unsigned long long best = 0;
        uint32_t blab = 0;
#pragma unroll
        for (int k = 0; k < EPT; ++k) {
            const int i = tid * EPT + k;
            const uint32_t pl = k ? L[k - 1] : before;
            if (L[k] != kPad && (i == 0 || pl != L[k])) cur = i;
            const uint32_t nl = k < EPT - 1 ? L[k + 1] : after;
            if (L[k] != kPad && nl != L[k]) {
                const unsigned long long val =
                    ((unsigned long long)(uint32_t)(i - cur + 1) << 32) | (uint32_t)key[k];
                if (val > best) {
                    best = val;
                    blab = L[k];
                }
            }
        }
        if constexpr (NW == 1) {
            const unsigned long long top = wave_max(best);
            const unsigned long long wl = __ballot(best == top && best != 0ull);
            const int32_t w = __builtin_amdgcn_readlane((int)blab, __ffsll((long long)wl) - 1);
            if (lane == 0) {
                next[it.v] = w;
                changed += (w != it.old);
            }
        } else {
            const unsigned long long top = block_max<NW>(best, sred64);   // barriers inside
            if (best == top && best != 0ull) {   // unique: positions are distinct
                next[it.v] = (int32_t)blab;
                changed += ((int32_t)blab != it.old);
            }
        }
    }
};

template <int NW, int EPT, typename E>
__global__ void __launch_bounds__(NW * kWave)
k_sort_reg(const E *__restrict__ off, const int32_t *__restrict__ nbr, const int32_t *__restrict__ prev,
           int32_t *__restrict__ next, const int32_t *__restrict__ vlist, int64_t v0, int64_t nv,
           unsigned long long *__restrict__ counter, const int32_t *__restrict__ stop, int32_t *__restrict__ err) {
    constexpr int T = NW * kWave;
    __shared__ unsigned long long sx[NW > 1 ? T * EPT : 1];
    __shared__ uint32_t sedge[2 * T];
    __shared__ int sw[NW];
    __shared__ unsigned long long sred64[NW];
    __shared__ uint32_t sred[NW];
    if (*stop) return;
    RegBitonicPolicy<NW, EPT, E> pol{off, prev, next, vlist, v0, nv, sx, sedge, sw, sred64, err, 0u};
    pipelined<T, EPT>(pol, nbr, prev, (int64_t)blockIdx.x, (int64_t)gridDim.x, (int)threadIdx.x);
    if constexpr (NW > 1) __syncthreads();
    block_count_add<NW>(pol.changed, counter, sred);
}

// ---------------------------------------------------------------------------
// Rows of B2K and B4K (1025..4096): one block of T threads per vertex, an LSD radix sort
// of the row's labels in LDS (8-bit digits, `passes` of them), each pass a
// stable multisplit: wave w owns the contiguous elements [w*M, (w+1)*M) (M =
// N / waves) and walks them in 64-element chunks in order; 8 ballots give
// each lane the lanes of its chunk with the same digit, whose popcount
// below it is its rank; per-wave digit histograms, one scan over (digit,
// wave) and per-wave running bases make the pass stable across the block.
// Positions ride along as the payload, so equal labels stay in row order.
// Then each run end finds its run's first element by binary search; the
// block max of (count << 32 | last position) picks the winner.
// Lanes of this wave whose 8-bit digit equals this lane's, among `valid`.
__device__ __forceinline__ unsigned long long digit_peers8(uint32_t d, bool valid) {
    unsigned long long peers = __ballot(valid);
#pragma unroll
    for (int b = 0; b < 8; ++b) {
        const bool bit = (d >> b) & 1u;
        const unsigned long long bb = __ballot(bit);
        peers &= bit ? bb : ~bb;
    }
    return peers;
}

// Stable LSD radix sort of skey[0, len) by `passes` 8-bit digits, with the
// 16-bit payload spos[0, len) riding along, in LDS; each array has two
// halves of N entries (ping-pong).  All T threads of the block call it;
// every wave owns the contiguous elements [w*M, (w+1)*M) (M = N / waves),
// 8 ballots give each lane the lanes of its 64-element chunk with the same
// digit, per-wave digit histograms (whist: waves x 256) and one scan over
// (digit, wave) make each pass stable across the block.  `passes` is the
// cap; the pass count is that of the largest key present (digits above it
// are 0 everywhere, and a pass over them would keep the order): converged
// rows hold mostly small labels after the degree-descending renumbering.
// Returns the half (0 or 1) that holds the sorted keys; ends with a barrier.
template <int T, int N>
__device__ int lds_radix_sort(uint32_t *skey, uint16_t *spos, uint32_t *whist, int len, int passes) {
    constexpr int NW = T / kWave;
    constexpr int M = N / NW;        // elements per wave
    constexpr int CH = M / kWave;    // 64-element chunks per wave
    static_assert(M % kWave == 0, "whole chunks per wave");
    const int tid = (int)threadIdx.x;
    const int lane = lane_id();
    const int w = tid / kWave;
    const unsigned long long lt = (1ull << lane) - 1ull;
    {
        __syncthreads();   // the caller's keys are in place
        uint32_t mx = 0;
        for (int j = tid; j < len; j += T) mx = mx > skey[j] ? mx : skey[j];
        mx = wave_max(mx);
        if (lane == 0) whist[w * 256] = mx;
        __syncthreads();
#pragma unroll 1
        for (int ww = 0; ww < NW; ++ww) mx = mx > whist[ww * 256] ? mx : whist[ww * 256];
        const int need = mx ? (32 - __clz((int)mx) + 7) / 8 : 0;
        passes = need < passes ? need : passes;
        __syncthreads();   // read before the first pass clears the rows
    }
    int src = 0;
#pragma unroll 1   // rolled loops: unrolled, they held 160 VGPRs in the hub tile kernel
    for (int p = 0; p < passes; ++p) {
        const int shift = 8 * p;
        const uint32_t *ks = skey + src * N;
        const uint16_t *ps = spos + src * N;
        uint32_t *kd = skey + (src ^ 1) * N;
        uint16_t *pd = spos + (src ^ 1) * N;
        for (int i = lane; i < 256; i += kWave) whist[w * 256 + i] = 0;
        __syncthreads();   // keys written; histogram rows clear
        // A: per-wave digit counts (the lowest lane of each peer group adds)
#pragma unroll 1
        for (int c = 0; c < CH; ++c) {
            const int j = w * M + c * kWave + lane;
            const bool valid = j < len;
            const uint32_t d = valid ? (ks[j] >> shift) & 0xffu : 0u;
            const unsigned long long peers = digit_peers8(d, valid);
            if (valid && (peers & lt) == 0) whist[w * 256 + d] += (uint32_t)__popcll(peers);
            if (w * M + c * kWave >= len) break;   // uniform per wave
        }
        __syncthreads();
        // bases: digit-major exclusive scan over (digit, wave)
        if (w == 0) {
            uint32_t tot[4];
            uint32_t run = 0;
#pragma unroll
            for (int q = 0; q < 4; ++q) {   // lane owns digits 4*lane .. 4*lane+3
                uint32_t t = 0;
#pragma unroll 1
                for (int ww = 0; ww < NW; ++ww) t += whist[ww * 256 + 4 * lane + q];
                tot[q] = t;
                run += t;
            }
            uint32_t inc = run;
#pragma unroll
            for (int sft = 1; sft < kWave; sft <<= 1) {
                const uint32_t y = (uint32_t)__shfl_up((int)inc, sft);
                if (lane >= sft) inc += y;
            }
            uint32_t base = inc - run;
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const int d = 4 * lane + q;
                uint32_t b2 = base;
#pragma unroll 1
                for (int ww = 0; ww < NW; ++ww) {
                    const uint32_t cnt = whist[ww * 256 + d];
                    whist[ww * 256 + d] = b2;
                    b2 += cnt;
                }
                base += tot[q];
            }
        }
        __syncthreads();
        // B: stable scatter, each wave in chunk order with its running bases
#pragma unroll 1
        for (int c = 0; c < CH; ++c) {
            if (w * M + c * kWave >= len) break;   // uniform per wave
            const int j = w * M + c * kWave + lane;
            const bool valid = j < len;
            const uint32_t key = valid ? ks[j] : 0u;
            const uint32_t d = (key >> shift) & 0xffu;
            const unsigned long long peers = digit_peers8(d, valid);
            const uint32_t base = valid ? whist[w * 256 + d] : 0u;
            if (valid) {
                const uint32_t dst = base + (uint32_t)__popcll(peers & lt);
                kd[dst] = key;
                pd[dst] = ps[j];
            }
            __builtin_amdgcn_wave_barrier();
            if (valid && (peers & lt) == 0) whist[w * 256 + d] = base + (uint32_t)__popcll(peers);
            __builtin_amdgcn_wave_barrier();
        }
        src ^= 1;
        __syncthreads();
    }
    return src;
}

template <int T, int EPT, typename E>
struct LdsSortPolicy {
    static constexpr int N = T * EPT;
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
    uint32_t *skey;          // 2 x N
    uint16_t *spos;          // 2 x N
    uint32_t *whist;         // NW x 256 (counts, then running bases)
    unsigned long long *sred64;
    int32_t *err;
    int passes;
    uint32_t changed;

    __device__ Item item(int64_t i) const {
        Item it{false, 0, 0, 0, 0};
        if (i < nv) {
            it.v = vlist ? (int64_t)vlist[v0 + i] : v0 + i;
            it.b = (int64_t)off[it.v];
            it.len = (int)((int64_t)off[it.v + 1] - it.b);
            it.old = prev[it.v];
            it.valid = true;
            if (it.len > N) {   // never planned into this class
                if (threadIdx.x == 0) atomicOr(err, 8);
                it.len = 0;
            }
        }
        return it;
    }
    __device__ int64_t edge0(const Item &it) const { return it.b; }

    __device__ void process(const Item &it, const int32_t (&lab)[EPT]) {
        if (it.len == 0) return;
        const int len = it.len;
        const int tid = (int)threadIdx.x;
        __syncthreads();   // the previous vertex's readers are done
#pragma unroll
        for (int k = 0; k < EPT; ++k) {
            const int pos = k * T + tid;
            if (pos < len) {
                skey[pos] = (uint32_t)lab[k];
                spos[pos] = (uint16_t)pos;
            }
        }
        const int src = lds_radix_sort<T, N>(skey, spos, whist, len, passes);
        const uint32_t *kf = skey + src * N;
        const uint16_t *pf = spos + src * N;
        unsigned long long best = 0;
        int32_t bestlab = 0;
        for (int q = tid; q < len; q += T) {
            const uint32_t key = kf[q];
            const bool end = q == len - 1 || kf[q + 1] != key;
            if (end) {
                int lo = 0, hi = q;
                while (lo < hi) {
                    const int mid = (lo + hi) >> 1;
                    if (kf[mid] < key) lo = mid + 1;
                    else hi = mid;
                }
                const unsigned long long val = ((unsigned long long)(q - lo + 1) << 32) | (uint32_t)pf[q];
                if (val > best) {
                    best = val;
                    bestlab = (int32_t)key;
                }
            }
        }
        const unsigned long long top = block_max<T / kWave>(best, sred64);
        if (best == top) {   // unique: positions are distinct
            next[it.v] = bestlab;
            changed += (bestlab != it.old);
        }
    }
};

template <int T, int EPT, typename E>
__global__ void __launch_bounds__(T)
k_sort_lds(const E *__restrict__ off, const int32_t *__restrict__ nbr, const int32_t *__restrict__ prev,
           int32_t *__restrict__ next, const int32_t *__restrict__ vlist, int64_t v0, int64_t nv,
           unsigned long long *__restrict__ counter, const int32_t *__restrict__ stop, int32_t *__restrict__ err,
           int passes) {
    static_assert(T * EPT <= 65536, "positions are 16-bit");
    static_assert((T * EPT / (T / kWave)) % kWave == 0, "whole chunks per wave");
    __shared__ uint32_t skey[2 * T * EPT];
    __shared__ uint16_t spos[2 * T * EPT];
    __shared__ uint32_t whist[(T / kWave) * 256];
    __shared__ unsigned long long sred64[T / kWave];
    __shared__ uint32_t sred[T / kWave];
    if (*stop) return;
    LdsSortPolicy<T, EPT, E> pol{off, prev, next, vlist, v0, nv, skey, spos, whist, sred64, err, passes, 0u};
    pipelined<T, EPT>(pol, nbr, prev, (int64_t)blockIdx.x, (int64_t)gridDim.x, (int)threadIdx.x);
    __syncthreads();
    block_count_add<T / kWave>(pol.changed, counter, sred);
}

// ---------------------------------------------------------------------------
// Hubs (degree > 4096), in the hub plan both methods share (tiles of
// kHubTile edges, P_h label partitions per hub -- for this method
// min(kHubMaxParts, ceil(deg / kHubPartSort)) -- and merge items (hub,
// partition)); every hub edge crosses HBM as a neighbour id and a
// gathered label once, and the tiles' run-length output (8 bytes per
// distinct label of a tile: the label and a hub_entry word) once more each
// way.
//  1. tile (k_hs_tile_wave): one wave sorts one tile's (partition, label,
//     position) keys in registers, turns each run into (label, count, last
//     position) -- the runs come out bucketed by label partition -- and
//     writes them with the P_h + 1 partition offsets, the same layout as
//     the hash method's tiles;
//  2. merge: items with at most kHsSmall entries are merged by one wave
//     each (k_hs_merge_small), the others are flagged for k_hs_merge: a block
//     per (hub, partition) stages that
//     partition's entries of every tile in LDS, tiles in row order, sorts
//     them by label (stable radix sort, entry index as payload) and reduces
//     each run: count = the sum of its counts (a block prefix sum), last =
//     the last entry's position (each tile adds at most one entry per label
//     and tiles are in row order, so that is the largest).  More entries than
//     the LDS holds are merged in rounds: each round's reduced runs stay at
//     the front and the next entries are appended behind them.  One 64-bit
//     atomicMax of (count << 32 | last) per block picks the hub's winner;
//  3. final (k_hs_final): the winner is the label at position `last`.
constexpr int kHsMergeT = 1024;
constexpr int kHsMergeCap = 4096;                 // entries staged per round
constexpr uint32_t kMergePad = 0xffffffffu;       // staged in place of another round's labels

// Wave-wide exclusive scans (sum, max) of one value per lane.
__device__ __forceinline__ uint32_t wave_excl_sum(uint32_t x) {
    const int lane = lane_id();
    uint32_t inc = x;
#pragma unroll
    for (int sft = 1; sft < kWave; sft <<= 1) {
        const uint32_t y = (uint32_t)__shfl_up((int)inc, sft);
        if (lane >= sft) inc += y;
    }
    return inc - x;
}

__device__ __forceinline__ uint32_t wave_excl_max(uint32_t x) {
    const int lane = lane_id();
    uint32_t inc = x;
#pragma unroll
    for (int sft = 1; sft < kWave; sft <<= 1) {
        const uint32_t y = (uint32_t)__shfl_up((int)inc, sft);
        if (lane >= sft) inc = inc > y ? inc : y;
    }
    const uint32_t ex = (uint32_t)__shfl_up((int)inc, 1);
    return lane ? ex : 0u;
}

// The tile phase with one wave per tile and no LDS: the wave sorts the
// tile's keys (partition << 43 | label << 11 | position in the tile, padded
// with all-ones) with the register bitonic network, 32 per lane, so runs of
// one label come out grouped by label partition, partitions in order; a run
// end gives the entry (label, hub_entry(run length, its last position)),
// written at its rank among the tile's run ends, and the partition offsets
// are the ranks where the partition changes.  Waves share nothing, so there
// is no barrier and no LDS.  (A block-wide LDS radix sort of the tile, four
// 8-bit passes of three barriers each plus a counting sort by partition,
// took 34 ms per iteration at RMAT-26 where this takes a fraction.)
constexpr int kHsWaveEPT = kHubTile / kWave;   // 32
constexpr int kHsWaveT = 256;

__device__ __forceinline__ unsigned long long shfl_up_u64(unsigned long long v) {
    const uint32_t lo = (uint32_t)__shfl_up((int)(uint32_t)v, 1);
    const uint32_t hi = (uint32_t)__shfl_up((int)(uint32_t)(v >> 32), 1);
    return ((unsigned long long)hi << 32) | lo;
}
__device__ __forceinline__ unsigned long long shfl_down_u64(unsigned long long v) {
    const uint32_t lo = (uint32_t)__shfl_down((int)(uint32_t)v, 1);
    const uint32_t hi = (uint32_t)__shfl_down((int)(uint32_t)(v >> 32), 1);
    return ((unsigned long long)hi << 32) | lo;
}

template <typename E>
__global__ void __launch_bounds__(kHsWaveT)
k_hs_tile_wave(const E *__restrict__ off, const int32_t *__restrict__ nbr, const int32_t *__restrict__ prev,
               const int32_t *__restrict__ tile_hub, const int32_t *__restrict__ hub_tile0,
               const int32_t *__restrict__ hub_vid, const int32_t *__restrict__ hub_nparts,
               const int64_t *__restrict__ tile_idx_off, int64_t tbase, int64_t ntiles,
               uint32_t *__restrict__ out_key, uint32_t *__restrict__ out_ent, int32_t *__restrict__ tile_idx,
               const int32_t *__restrict__ stop) {
    constexpr int EPT = kHsWaveEPT;
    constexpr int NW = kHsWaveT / kWave;
    static_assert(kHubMaxParts <= (1 << 12) && kHubTile == (1 << 11), "key layout: 12 + 31 + 11 bits");
    if (*stop) return;
    const int lane = lane_id();
    const int64_t wid = (int64_t)blockIdx.x * NW + threadIdx.x / kWave;
    for (int64_t t = tbase + wid; t < ntiles; t += (int64_t)gridDim.x * NW) {
        const int h = tile_hub[t];
        const int64_t v = hub_vid[h];
        const int64_t rb = (int64_t)off[v], re = (int64_t)off[v + 1];
        const int64_t row0 = (t - (int64_t)hub_tile0[h]) * (int64_t)kHubTile;
        const int64_t e0 = rb + row0;
        const int len = (int)min((int64_t)kHubTile, re - e0);
        const uint32_t P = (uint32_t)hub_nparts[h];
        unsigned long long key[EPT];
        {
            int32_t u[EPT];
#pragma unroll
            for (int k = 0; k < EPT; ++k) {
                const int pos = k * kWave + lane;
                u[k] = __builtin_nontemporal_load(&nbr[pos < len ? e0 + pos : e0]);
            }
#pragma unroll
            for (int k = 0; k < EPT; ++k) {
                const int pos = k * kWave + lane;
                const uint32_t L = (uint32_t)prev[u[k]];
                key[k] = pos < len ? ((unsigned long long)hub_partition(L, P) << 43 |
                                      (unsigned long long)L << 11 | (unsigned long long)pos)
                                   : ~0ull;
            }
        }
        reg_bitonic<1, EPT>(key, lane, nullptr);
        // element e = lane * EPT + k; g = partition and label (key >> 11)
        const unsigned long long before = shfl_up_u64(key[EPT - 1]);
        const unsigned long long after = shfl_down_u64(key[0]);
        // run starts and ends of this lane's elements as bit masks (bit k:
        // element k), so the write loop below keeps no 64-bit neighbours
        uint32_t smask = 0, emask = 0;
        int maxpart = -1;   // partition of this lane's last run end
#pragma unroll
        for (int k = 0; k < EPT; ++k) {
            if (key[k] == ~0ull) continue;
            const unsigned long long g = key[k] >> 11;
            const unsigned long long gp = (k ? key[k - 1] : before) >> 11;
            const unsigned long long gn = (k + 1 < EPT ? key[k + 1] : after) >> 11;
            if ((lane == 0 && k == 0) || g != gp) smask |= 1u << k;
            if ((lane == kWave - 1 && k == EPT - 1) || g != gn) {
                emask |= 1u << k;
                maxpart = (int)(key[k] >> 43);
            }
        }
        const int last_start = smask ? lane * EPT + 31 - __clz((int)smask) : -1;
        const uint32_t nend = (uint32_t)__popc(emask);
        // carried from the lanes before: their last run start, the number of
        // run ends and the partition of the last one
        int start = (int)wave_excl_max((uint32_t)(last_start + 1)) - 1;
        uint32_t q = wave_excl_sum(nend);
        int pp = (int)wave_excl_max((uint32_t)(maxpart + 1)) - 1;
        const uint32_t total = (uint32_t)__shfl((int)(q + nend), kWave - 1);
        const int64_t ob = t * (int64_t)kHubTile, idx = tile_idx_off[t];
#pragma unroll
        for (int k = 0; k < EPT; ++k) {
            if ((smask >> k) & 1u) start = lane * EPT + k;
            if ((emask >> k) & 1u) {
                const int pt = (int)(key[k] >> 43);
                for (int p = pp + 1; p <= pt; ++p) tile_idx[idx + p] = (int32_t)q;   // partitions starting here
                pp = pt;
                out_key[ob + q] = (uint32_t)((key[k] >> 11) & 0xffffffffull);
                out_ent[ob + q] = hub_entry((uint32_t)(lane * EPT + k - start + 1), (uint32_t)(key[k] & (kHubTile - 1)));
                ++q;
            }
        }
        // the partitions after the last run end start at the tile's end
        const int last_part = (int)wave_max((uint32_t)(maxpart + 1)) - 1;
        for (int p = last_part + 1 + lane; p <= (int)P; p += kWave) tile_idx[idx + p] = (int32_t)total;
    }
}

// Block inclusive prefix sum of x[0, len) into y[0, len) (len <= T * EPT):
// thread t sums its EPT consecutive elements, one block scan of the sums.
template <int T, int EPT>
__device__ __forceinline__ void block_prefix_sum(const uint32_t *x, uint32_t *y, int len, int *sscan) {
    const int tid = (int)threadIdx.x;
    const int b = tid * EPT;
    uint32_t v[EPT];
    int sum = 0;
#pragma unroll
    for (int k = 0; k < EPT; ++k) {
        v[k] = b + k < len ? x[b + k] : 0u;
        sum += (int)v[k];
    }
    uint32_t acc = (uint32_t)block_exclusive_scan<T / kWave>(sum, sscan);   // barriers inside
#pragma unroll
    for (int k = 0; k < EPT; ++k) {
        acc += v[k];
        if (b + k < len) y[b + k] = acc;
    }
}

// Merge of one round of staged entries: keys in skey[0, len) (first half),
// their counts and last positions in scnt / slast at the same slots.  Sorts
// by label (entry slot as payload), reduces each run to (label, summed
// count, last entry's position) and writes the runs back, in label order, to
// slots [0, runs).  Returns the number of runs (uniform).
template <int T, int N>
__device__ int hs_merge_round(uint32_t *skey, uint16_t *spos, uint32_t *whist, uint32_t *scnt, uint32_t *slast,
                              int len, int passes, int *sscan, int *snruns) {
    constexpr int EPT = N / T;
    const int tid = (int)threadIdx.x;
    for (int i = tid; i < len; i += T) spos[i] = (uint16_t)i;
    // lds_radix_sort's first barrier orders these writes before its reads
    const int src = lds_radix_sort<T, N>(skey, spos, whist, len, passes);
    const uint32_t *kf = skey + src * N;
    const uint16_t *pf = spos + src * N;
    uint32_t *csort = skey + (src ^ 1) * N;   // free half: counts in sorted order, then their prefix sums
    for (int q = tid; q < len; q += T) csort[q] = scnt[pf[q]];
    __syncthreads();
    block_prefix_sum<T, EPT>(csort, csort, len, sscan);
    __syncthreads();
    // each thread: the run ends among its EPT consecutive sorted slots
    uint32_t rk[EPT], rc[EPT], rl[EPT];
    bool end[EPT];
    int mine = 0;
#pragma unroll
    for (int k = 0; k < EPT; ++k) {
        const int q = tid * EPT + k;
        end[k] = false;
        // a run of pads (labels of another round) is dropped: it sorts last
        if (q < len && kf[q] != kMergePad && (q == len - 1 || kf[q + 1] != kf[q])) {
            const uint32_t key = kf[q];
            int lo = 0, hi = q;
            while (lo < hi) {
                const int mid = (lo + hi) >> 1;
                if (kf[mid] < key) lo = mid + 1;
                else hi = mid;
            }
            end[k] = true;
            rk[k] = key;
            rc[k] = csort[q] - (lo ? csort[lo - 1] : 0u);
            rl[k] = slast[pf[q]];
            ++mine;
        }
    }
    int r = block_exclusive_scan<T / kWave>(mine, sscan);   // barriers inside: every read above is done
    if (tid == T - 1) *snruns = r + mine;
#pragma unroll
    for (int k = 0; k < EPT; ++k)
        if (end[k]) {
            skey[r] = rk[k];
            scnt[r] = rc[k];
            slast[r] = rl[k];
            ++r;
        }
    __syncthreads();
    return *snruns;
}

// Small merge items: one wave per item (hub, partition) whose tiles hold at
// most kHsSmall entries in it -- the usual case once labels have settled,
// when a partition gets about one entry per tile.  Larger items are flagged
// for the block merge (k_hs_merge).  The wave stages the item's entries in
// its LDS slice (flat index = tile order), sorts (label << 32 | flat index)
// with the register bitonic network, and reduces each run: its count is the
// difference of the inclusive prefix sums of the counts at the run's end and
// before its start (an exclusive max-scan carries the latter, prefix sums
// being non-decreasing), its last position that of its last entry (the
// stable order puts the latest tile last).
constexpr int kHsSmall = 1024;
constexpr int kHsSmallT = 256;

struct HsWaveLds {
    uint32_t key[kHsSmall], cnt[kHsSmall], last[kHsSmall];
    int32_t lo[kWave], start[kWave];
};

// max over the item's runs of (count << 32 | last); n <= 64 * EPT entries
template <int EPT>
__device__ unsigned long long hs_wave_reduce(const HsWaveLds &w, int n) {
    const int lane = lane_id();
    constexpr uint32_t kPad = 0xffffffffu;   // labels are < 2^31
    unsigned long long key[EPT];
#pragma unroll
    for (int k = 0; k < EPT; ++k) {
        const int i = lane * EPT + k;
        key[k] = i < n ? (((unsigned long long)w.key[i] << 32) | (uint32_t)i) : ~0ull;
    }
    reg_bitonic<1, EPT>(key, lane, nullptr);
    uint32_t L[EPT], P[EPT], slot[EPT];
    uint32_t run = 0;
#pragma unroll
    for (int k = 0; k < EPT; ++k) {
        L[k] = (uint32_t)(key[k] >> 32);
        slot[k] = (uint32_t)key[k];
        run += L[k] != kPad ? w.cnt[slot[k]] : 0u;
        P[k] = run;   // inclusive, within the lane for now
    }
    const uint32_t base = wave_excl_sum(run);
    // every lane shuffles: in `c ? shfl : x` only the lanes with c would, and
    // a lane reading from one that skipped it gets 0
    const uint32_t up = (uint32_t)__shfl_up((int)L[EPT - 1], 1);
    const uint32_t dn = (uint32_t)__shfl_down((int)L[0], 1);
    const uint32_t before = lane ? up : kPad;
    const uint32_t after = lane < kWave - 1 ? dn : kPad;
    // the prefix sum before each run start; their lane max is carried on
    uint32_t lmax = 0;
#pragma unroll
    for (int k = 0; k < EPT; ++k) {
        P[k] += base;
        const uint32_t pv = k ? L[k - 1] : before;
        const uint32_t c = L[k] != kPad ? w.cnt[slot[k]] : 0u;
        if (L[k] != kPad && L[k] != pv) lmax = lmax > P[k] - c ? lmax : P[k] - c;
    }
    uint32_t cur = wave_excl_max(lmax);
    unsigned long long best = 0;
#pragma unroll
    for (int k = 0; k < EPT; ++k) {
        if (L[k] == kPad) continue;
        const uint32_t pv = k ? L[k - 1] : before;
        const uint32_t c = w.cnt[slot[k]];
        if (L[k] != pv) cur = cur > P[k] - c ? cur : P[k] - c;
        const uint32_t nx = k + 1 < EPT ? L[k + 1] : after;
        if (L[k] != nx) {
            const unsigned long long pk = ((unsigned long long)(P[k] - cur) << 32) | w.last[slot[k]];
            best = pk > best ? pk : best;
        }
    }
    return wave_max(best);
}

__global__ void __launch_bounds__(kHsSmallT)
k_hs_merge_small(const int32_t *__restrict__ item_hub, const int32_t *__restrict__ item_part, int64_t ibase,
                 int64_t nitems, const int32_t *__restrict__ hub_tile0, const int32_t *__restrict__ hub_ntiles,
                 const int64_t *__restrict__ tile_idx_off, const int32_t *__restrict__ tile_idx,
                 const uint32_t *__restrict__ out_key, const uint32_t *__restrict__ out_ent,
                 unsigned long long *__restrict__ hub_best, int32_t *__restrict__ big,
                 const int32_t *__restrict__ stop) {
    constexpr int NW = kHsSmallT / kWave;
    __shared__ HsWaveLds sw[NW];
    if (*stop) return;
    const int lane = lane_id();
    HsWaveLds &w = sw[threadIdx.x / kWave];
    const int64_t wid = (int64_t)blockIdx.x * NW + threadIdx.x / kWave;
    for (int64_t it = ibase + wid; it < ibase + nitems; it += (int64_t)gridDim.x * NW) {
        const int h = item_hub[it];
        const int p = item_part[it];
        const int64_t t0 = hub_tile0[h], nt = hub_ntiles[h];
        int total = 0;
        for (int64_t c0 = t0; c0 < t0 + nt; c0 += kWave) {
            const int64_t t = c0 + lane;
            int len = 0;
            if (t < t0 + nt) {
                const int64_t ib = tile_idx_off[t];
                len = tile_idx[ib + p + 1] - tile_idx[ib + p];
            }
            total += (int)wave_sum_u32((uint32_t)len);
        }
        if (lane == 0) big[it] = total > kHsSmall ? 1 : 0;
        if (total == 0 || total > kHsSmall) continue;   // nothing, or the block merge takes it
        int n = 0;
        for (int64_t c0 = t0; c0 < t0 + nt; c0 += kWave) {
            const int64_t t = c0 + lane;
            int lo = 0, len = 0;
            if (t < t0 + nt) {
                const int64_t ib = tile_idx_off[t];
                lo = tile_idx[ib + p];
                len = tile_idx[ib + p + 1] - lo;
            }
            const int st = (int)wave_excl_sum((uint32_t)len);
            const int ctot = __shfl(st + len, kWave - 1);
            w.lo[lane] = lo;
            w.start[lane] = st;
            __builtin_amdgcn_wave_barrier();
            const int nc = (int)min((int64_t)kWave, t0 + nt - c0);
            for (int i = lane; i < ctot; i += kWave) {
                int a = 0, b = nc;   // last tile with start <= i
                while (b - a > 1) {
                    const int mid = (a + b) >> 1;
                    if (w.start[mid] <= i) a = mid;
                    else b = mid;
                }
                const int64_t e = (c0 + a) * (int64_t)kHubTile + w.lo[a] + (i - w.start[a]);
                const uint32_t ent = out_ent[e];
                w.key[n + i] = out_key[e];
                w.cnt[n + i] = hub_entry_count(ent);
                w.last[n + i] = (uint32_t)((c0 + a - t0) * (int64_t)kHubTile) + hub_entry_pos(ent);
            }
            n += ctot;
            __builtin_amdgcn_wave_barrier();
        }
        unsigned long long best;
        if (n <= kWave) best = hs_wave_reduce<1>(w, n);
        else if (n <= 2 * kWave) best = hs_wave_reduce<2>(w, n);
        else if (n <= 4 * kWave) best = hs_wave_reduce<4>(w, n);
        else if (n <= 8 * kWave) best = hs_wave_reduce<8>(w, n);
        else best = hs_wave_reduce<16>(w, n);
        if (lane == 0 && best) atomicMax(&hub_best[h], best);
        __builtin_amdgcn_wave_barrier();   // the slice is rewritten by the next item
    }
}

// One block per merge item (hub, partition), grid-stride over the items.
template <int T>
__global__ void __launch_bounds__(T)
k_hs_merge(const int32_t *__restrict__ item_hub, const int32_t *__restrict__ item_part,
           int64_t ibase, int64_t nitems, const int32_t *__restrict__ big,
           const int32_t *__restrict__ hub_tile0, const int32_t *__restrict__ hub_ntiles,
           const int64_t *__restrict__ tile_idx_off, const int32_t *__restrict__ tile_idx,
           const uint32_t *__restrict__ out_key, const uint32_t *__restrict__ out_ent,
           unsigned long long *__restrict__ hub_best, const int32_t *__restrict__ stop, int32_t *__restrict__ err,
           int passes) {
    constexpr int N = kHsMergeCap;
    __shared__ uint32_t skey[2 * N];
    __shared__ uint16_t spos[2 * N];
    __shared__ uint32_t scnt[N], slast[N];
    __shared__ uint32_t whist[(T / kWave) * 256];
    __shared__ int32_t slo[T], sstart[T];
    __shared__ int sscan[T / kWave];
    __shared__ int stotal, snruns;
    __shared__ unsigned long long sred64[T / kWave];
    __shared__ uint32_t sbig[T / 32];   // flags of the current chunk of items
    if (*stop) return;
    const int tid = (int)threadIdx.x;
    // the flagged items of each chunk of T of the block's items, gathered
    // into an LDS bit mask (one flag load per thread instead of a dependent
    // load per item)
    // items go round-robin over the blocks (item ibase + blockIdx.x + j *
    // gridDim.x is the block's j-th), T of them per chunk
    const int64_t iend = ibase + nitems;
    for (int64_t j0 = 0; ibase + blockIdx.x + j0 * gridDim.x < iend; j0 += T) {
    __syncthreads();   // the previous chunk's reads of sbig are done
    if (tid < T / 32) sbig[tid] = 0u;
    __syncthreads();
    {
        const int64_t ij = ibase + blockIdx.x + (j0 + tid) * gridDim.x;
        if (ij < iend && big[ij]) atomicOr(&sbig[tid / 32], 1u << (tid % 32));
    }
    __syncthreads();
    for (int qq = 0; qq < T; ++qq) {
        const uint32_t word = sbig[qq / 32] >> (qq % 32);   // uniform
        if (word == 0u) {   // no flagged item left in this word
            qq |= 31;
            continue;
        }
        qq += __ffs((int)word) - 1;
        const int64_t it = ibase + blockIdx.x + (j0 + qq) * gridDim.x;
        const int h = item_hub[it];
        const int p = item_part[it];
        const int64_t t0 = hub_tile0[h], nt = hub_ntiles[h];
        // More distinct labels than N in the partition: merged again in R
        // rounds, round r staging the labels whose sub-round hash is r (the
        // others as pads); R doubles until every round fits.
        uint32_t R = 1;
        unsigned long long item_best = 0;
        for (uint32_t r = 0; r < R;) {
        int held = 0;   // staged slots: reduced runs first, then new entries (uniform)
        bool overflow = false;
        for (int64_t c0 = t0; c0 < t0 + nt && !overflow; c0 += T) {
            // this chunk of T tiles: slice bounds one tile per thread, a scan
            // of the slice lengths maps a flat index back to its tile
            __syncthreads();
            const int64_t t = c0 + tid;
            int lo = 0, len = 0;
            if (t < t0 + nt) {
                const int64_t ib = tile_idx_off[t];
                lo = tile_idx[ib + p];
                len = tile_idx[ib + p + 1] - lo;
            }
            const int start = block_exclusive_scan<T / kWave>(len, sscan);   // barriers inside
            slo[tid] = lo;
            sstart[tid] = start;
            if (tid == T - 1) stotal = start + len;
            __syncthreads();
            const int total = stotal;
            const int nc = (int)min((int64_t)T, t0 + nt - c0);
            for (int i0 = 0; i0 < total;) {
                if (held == N) {   // full: merge what is staged
                    held = hs_merge_round<T, N>(skey, spos, whist, scnt, slast, held, passes, sscan, &snruns);
                    if (held == N) {   // N distinct labels in this round: no room to progress
                        overflow = true;
                        break;
                    }
                }
                const int take = min(N - held, total - i0);
                for (int j = tid; j < take; j += T) {
                    const int i = i0 + j;
                    int a = 0, b = nc;   // last tile with start <= i
                    while (b - a > 1) {
                        const int mid = (a + b) >> 1;
                        if (sstart[mid] <= i) a = mid;
                        else b = mid;
                    }
                    const int64_t e = (c0 + a) * (int64_t)kHubTile + slo[a] + (i - sstart[a]);
                    const uint32_t key = out_key[e];
                    const bool mine = R == 1 || sub_round(key, R) == r;
                    const uint32_t w = out_ent[e];
                    skey[held + j] = mine ? key : kMergePad;
                    scnt[held + j] = mine ? hub_entry_count(w) : 0u;
                    slast[held + j] = (uint32_t)((c0 + a - t0) * (int64_t)kHubTile) + hub_entry_pos(w);
                }
                held += take;
                i0 += take;
                __syncthreads();
            }
        }
        if (overflow) {   // start the item over in twice as many rounds
            if (R >= 4096u) {
                if (tid == 0) atomicOr(err, 16);
                break;
            }
            R *= 2;
            r = 0;
            item_best = 0;
            continue;
        }
        if (held > 0) held = hs_merge_round<T, N>(skey, spos, whist, scnt, slast, held, passes, sscan, &snruns);
        unsigned long long best = 0;
        for (int q = tid; q < held; q += T) {
            const unsigned long long pk = ((unsigned long long)scnt[q] << 32) | slast[q];
            best = pk > best ? pk : best;
        }
        best = block_max<T / kWave>(best, sred64);
        item_best = best > item_best ? best : item_best;
        ++r;
        }
        if (tid == 0 && item_best) atomicMax(&hub_best[h], item_best);
    }
    }
}

// Hub winners: the label at position `last` of the row; resets hub_best.
template <typename E>
__global__ void __launch_bounds__(256)
k_hs_final(const E *__restrict__ off, const int32_t *__restrict__ nbr, const int32_t *__restrict__ hub_vid,
           int64_t nhubs, unsigned long long *__restrict__ hub_best, const int32_t *__restrict__ prev,
           int32_t *__restrict__ next, unsigned long long *__restrict__ counter, const int32_t *__restrict__ stop) {
    __shared__ uint32_t sred[4];
    if (*stop) return;
    uint32_t changed = 0;
    for (int64_t base = (int64_t)blockIdx.x * 256; base < nhubs; base += (int64_t)gridDim.x * 256) {
        const int64_t h = base + threadIdx.x;
        if (h < nhubs) {
            const int64_t v = hub_vid[h];
            const unsigned long long b = hub_best[h];
            hub_best[h] = 0ull;
            const int32_t w = prev[nbr[(int64_t)off[v] + (uint32_t)b]];
            next[v] = w;
            changed += (w != prev[v]);
        }
    }
    block_count_add<4>(changed, counter, sred);
}

// ---------------------------------------------------------------------------
// Host side.

static int grid_for(const galp_graph *g, int64_t work_items, int blocks_per_cu) {
    const int64_t cap = (int64_t)g->ctx->prop.multiProcessorCount * blocks_per_cu;
    int64_t b = work_items < cap ? work_items : cap;
    return (int)(b < 1 ? 1 : b);
}

template <class T>
static int salloc(galp_graph *g, T **p, int64_t count) {
    *p = nullptr;
    if (count <= 0) count = 1;
    if (hipMalloc((void **)p, sizeof(T) * (size_t)count) != hipSuccess) {
        *p = nullptr;
        (void)hipGetLastError();
        return fail(g->ctx, GALP_ENOMEM, "sort-method allocation failed");
    }
    g->device_bytes += (int64_t)(sizeof(T) * (size_t)count);
    return GALP_OK;
}

void sort_free(galp_graph *g) {
    delete g->sort;
    g->sort = nullptr;
}

// The radix passes labels < n need; the hub path's buffers belong to the hub
// plan both methods share (api.hip).
int sort_setup(galp_graph *g) {
    if (g->sort) return GALP_OK;
    SortState *S = new (std::nothrow) SortState();
    if (!S) return fail(g->ctx, GALP_ENOMEM, "host allocation failed");
    g->sort = S;
    const int bits = g->n > 1 ? 64 - __builtin_clzll((unsigned long long)(g->n - 1)) : 1;
    S->passes = (bits + 7) / 8;
    return GALP_OK;
}

template <typename E>
static double class_bytes(const BucketRange &r) {
    // offsets, neighbour ids, gathered labels, own previous label, next label
    return (double)sizeof(E) * (double)(r.nv + 1) + 8.0 * (double)r.edges + 8.0 * (double)r.nv;
}

template <int G, typename E>
static void sort_group(galp_graph *g, const IterView &w, int bid, const int32_t *prev, int32_t *next,
                       int blocks_per_cu = 8) {
    const BucketRange &r = w.cls[bid];
    if (!r.nv) return;
    PendingLaunch pl;
    prof_begin(g, kBucketName[bid], class_bytes<E>(r), &pl);
    const int64_t per_block = 256 / G;
    const int64_t blocks = (r.nv + per_block - 1) / per_block;
    if (g->d_vlist)
        k_group<G, SortGroupResolve, SortGroupResolve, true, E><<<grid_for(g, blocks, blocks_per_cu), 256, 0,
                                                                  g->stream>>>(
            (const E *)w.off, w.nbr, prev, next, g->d_vlist, r.v0, r.nv, nullptr, g->d_counter, g->d_stop, nullptr, 0);
    else
        k_group<G, SortGroupResolve, SortGroupResolve, false, E><<<grid_for(g, blocks, blocks_per_cu), 256, 0,
                                                                   g->stream>>>(
            (const E *)w.off, w.nbr, prev, next, nullptr, r.v0, r.nv, nullptr, g->d_counter, g->d_stop, nullptr, 0);
    prof_end(g, &pl);
}

// RADIX: the LSD radix sort (large rows) or the bitonic network (smaller).
template <int T, int EPT, typename E, bool RADIX>
static void sort_lds(galp_graph *g, const IterView &w, int bid, const int32_t *prev, int32_t *next,
                     int blocks_per_cu) {
    const BucketRange &r = w.cls[bid];
    if (!r.nv) return;
    PendingLaunch pl;
    prof_begin(g, kBucketName[bid], class_bytes<E>(r), &pl);
    if constexpr (RADIX)
        k_sort_lds<T, EPT, E><<<grid_for(g, r.nv, blocks_per_cu), T, 0, g->stream>>>(
            (const E *)w.off, w.nbr, prev, next, g->d_vlist, r.v0, r.nv, g->d_counter, g->d_stop, g->d_err,
            g->sort->passes);
    else
        k_sort_bitonic<T, EPT, E><<<grid_for(g, r.nv, blocks_per_cu), T, 0, g->stream>>>(
            (const E *)w.off, w.nbr, prev, next, g->d_vlist, r.v0, r.nv, g->d_counter, g->d_stop, g->d_err);
    prof_end(g, &pl);
}

template <int NW, int EPT, typename E>
static void sort_reg(galp_graph *g, const IterView &w, int bid, const int32_t *prev, int32_t *next,
                     int blocks_per_cu) {
    const BucketRange &r = w.cls[bid];
    if (!r.nv) return;
    PendingLaunch pl;
    prof_begin(g, kBucketName[bid], class_bytes<E>(r), &pl);
    k_sort_reg<NW, EPT, E><<<grid_for(g, r.nv, blocks_per_cu), NW * kWave, 0, g->stream>>>(
        (const E *)w.off, w.nbr, prev, next, g->d_vlist, r.v0, r.nv, g->d_counter, g->d_stop, g->d_err);
    prof_end(g, &pl);
}

// Hubs [h0, h1) of a view, with their tiles [t0, t1) and merge items
// [i0, i1): tile sorts, partition merges, winners.  tile_per_cu and
// merge_per_cu size the grids (the autotuner's options).
template <typename E>
static int sort_hubs(galp_graph *g, const IterView &w, const int32_t *prev, int32_t *next, int tile_per_cu = 2,
                     int merge_per_cu = 1) {
    const int64_t nt = w.t1 - w.t0, ni = w.i1 - w.i0, nh = w.h1 - w.h0;
    double hub_edges = 0.0;
    if (w.h0 == 0 && w.h1 == g->nhubs) hub_edges = (double)g->bucket[B_HUB].edges;
    PendingLaunch pl;
    prof_begin(g, "hub_sort", 8.0 * hub_edges, &pl);
    k_hs_tile_wave<E><<<grid_for(g, (nt + kHsWaveT / kWave - 1) / (kHsWaveT / kWave), tile_per_cu), kHsWaveT, 0,
                        g->stream>>>((const E *)w.off, w.nbr, prev, g->d_tile_hub, g->d_hub_tile0, g->d_hub_vid,
                                     g->d_hub_nparts, g->d_tile_idx_off, w.t0, w.t1, g->d_out_key, g->d_out_ent,
                                     g->d_tile_idx, g->d_stop);
    const SortState *S = g->sort;
    k_hs_merge_small<<<grid_for(g, (ni + kHsSmallT / kWave - 1) / (kHsSmallT / kWave), 3), kHsSmallT, 0,
                       g->stream>>>(g->d_item_hub, g->d_item_part, w.i0, ni, g->d_hub_tile0, g->d_hub_ntiles,
                                    g->d_tile_idx_off, g->d_tile_idx, g->d_out_key, g->d_out_ent, g->d_hub_best,
                                    g->d_merge_big, g->d_stop);
    k_hs_merge<kHsMergeT><<<grid_for(g, ni, merge_per_cu), kHsMergeT, 0, g->stream>>>(
        g->d_item_hub, g->d_item_part, w.i0, ni, g->d_merge_big, g->d_hub_tile0, g->d_hub_ntiles, g->d_tile_idx_off,
        g->d_tile_idx, g->d_out_key, g->d_out_ent, g->d_hub_best, g->d_stop, g->d_err, S->passes);
    k_hs_final<E><<<grid_for(g, (nh + 255) / 256, 4), 256, 0, g->stream>>>(
        (const E *)w.off, w.nbr, g->d_hub_vid + w.h0, nh, g->d_hub_best + w.h0, prev, next, g->d_counter, g->d_stop);
    prof_end(g, &pl);
    return GALP_OK;
}

// The sort method's configuration of each degree class (g->tune[b]): option
// 0 is the default, 1 and 2 the alternatives the autotuner times --
// workgroup size x keys per thread, the radix sort or the bitonic network,
// blocks per CU; for lane groups the grid; for hubs the blocks per CU of the
// tile grid (2 or 1 four-wave blocks; at 224 VGPRs two fit) and of the
// merge grid (1 or 2).  Every option computes
// the same labels.
template <typename E>
static int sort_class(galp_graph *g, const IterView &w, int b, const int32_t *prev, int32_t *next) {
    const int c = g->tune[b];
    switch (b) {
    case B_HUB:
        if (w.h1 > w.h0) return sort_hubs<E>(g, w, prev, next, c == 1 ? 1 : 2, c == 2 ? 2 : 1);
        break;
    case B_B4K:
        if (c == 1) sort_reg<4, 16, E>(g, w, b, prev, next, 2);
        else if (c == 2) sort_lds<256, 16, E, true>(g, w, b, prev, next, 4);
        else sort_lds<512, 8, E, true>(g, w, b, prev, next, 2);
        break;
    case B_B2K:
        if (c == 1) sort_lds<512, 4, E, true>(g, w, b, prev, next, 2);
        else if (c == 2) sort_lds<128, 16, E, true>(g, w, b, prev, next, 8);
        else sort_lds<256, 8, E, true>(g, w, b, prev, next, 4);
        break;
    case B_B1K:
        if (c == 1) sort_lds<128, 8, E, false>(g, w, b, prev, next, 8);
        else if (c == 2) sort_lds<128, 8, E, true>(g, w, b, prev, next, 8);
        else sort_reg<1, 16, E>(g, w, b, prev, next, 16);
        break;
    case B_W512:
        if (c == 1) sort_lds<64, 8, E, false>(g, w, b, prev, next, 16);
        else if (c == 2) sort_lds<64, 8, E, true>(g, w, b, prev, next, 16);
        else sort_reg<1, 8, E>(g, w, b, prev, next, 16);
        break;
    case B_W256:
        if (c == 1) sort_lds<64, 4, E, false>(g, w, b, prev, next, 24);
        else if (c == 2) sort_lds<64, 4, E, true>(g, w, b, prev, next, 24);
        else sort_reg<1, 4, E>(g, w, b, prev, next, 24);
        break;
    case B_W128:
        if (c == 1) sort_lds<64, 2, E, false>(g, w, b, prev, next, 32);
        else if (c == 2) sort_lds<64, 2, E, true>(g, w, b, prev, next, 32);
        else sort_reg<1, 2, E>(g, w, b, prev, next, 32);
        break;
    default: {
        const int bpc = c == 1 ? 16 : c == 2 ? 4 : 8;
        if (b == B_G64) sort_group<64, E>(g, w, b, prev, next, bpc);
        else if (b == B_G32) sort_group<32, E>(g, w, b, prev, next, bpc);
        else if (b == B_G16) sort_group<16, E>(g, w, b, prev, next, bpc);
        else if (b == B_G8) sort_group<8, E>(g, w, b, prev, next, bpc);
        else sort_group<4, E>(g, w, b, prev, next, bpc);
    }
    }
    return GALP_OK;
}

int launch_sort_class(galp_graph *g, int b) {
    const IterView w = full_view(g);
    const int32_t *prev = g->d_lab[g->cur];
    int32_t *next = g->d_lab[g->cur ^ 1];
    const int rc = g->off64 ? sort_class<int64_t>(g, w, b, prev, next) : sort_class<int32_t>(g, w, b, prev, next);
    if (rc != GALP_OK) return rc;
    const hipError_t e = hipGetLastError();
    return e == hipSuccess ? GALP_OK : fail(g->ctx, GALP_EHIP, std::string("class launch: ") + hipGetErrorString(e));
}

// In core, the classes are independent (each reads prev and writes next at
// its own rows): the lane-group classes run on side stream 0, the LDS
// classes on side stream 1, the hub path on the graph's stream.
template <typename E>
static int sort_iteration_concurrent(galp_graph *g, const IterView &w, const int32_t *prev, int32_t *next) {
    hipStream_t ms = nullptr;
    int rc = side_begin(g, 0, &ms);
    if (rc != GALP_OK) return rc;
    for (int b = B_G64; b >= B_G4 && rc == GALP_OK; --b) rc = sort_class<E>(g, w, b, prev, next);
    const int rc0 = side_end(g, 0, ms);
    if (rc != GALP_OK || rc0 != GALP_OK) return rc != GALP_OK ? rc : rc0;
    if ((rc = side_begin(g, 1, &ms)) != GALP_OK) return rc;
    for (int b = B_B4K; b >= B_W128 && rc == GALP_OK; --b) rc = sort_class<E>(g, w, b, prev, next);
    const int rc1 = side_end(g, 1, ms);
    if (rc != GALP_OK || rc1 != GALP_OK) return rc != GALP_OK ? rc : rc1;
    if ((rc = sort_class<E>(g, w, B_HUB, prev, next)) != GALP_OK) return rc;
    if ((rc = side_join(g, 0)) != GALP_OK || (rc = side_join(g, 1)) != GALP_OK) return rc;
    const hipError_t e = hipGetLastError();
    return e == hipSuccess ? GALP_OK : fail(g->ctx, GALP_EHIP, std::string("sort kernel launch: ") + hipGetErrorString(e));
}

template <typename E>
static int sort_iteration(galp_graph *g, const IterView &w) {
    const int32_t *prev = g->d_lab[g->cur];
    int32_t *next = g->d_lab[g->cur ^ 1];
    if (!g->ooc && concurrent_classes(g)) return sort_iteration_concurrent<E>(g, w, prev, next);
    if (g->opts.policy == GALP_POLICY_LI) {
        // "load-imbalanced" ablation: every class <= 4096 through the
        // 512-thread block-per-vertex kernel, as the reference's li path
        // runs one block per vertex whatever its degree.
        if (w.h1 > w.h0) {
            const int rc = sort_hubs<E>(g, w, prev, next);
            if (rc != GALP_OK) return rc;
        }
        for (int b = B_B4K; b >= B_G4; --b) sort_lds<512, 8, E, true>(g, w, b, prev, next, 2);
    } else {
        for (int b = B_HUB; b >= B_G4; --b) {   // largest work first
            const int rc = sort_class<E>(g, w, b, prev, next);
            if (rc != GALP_OK) return rc;
        }
    }
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return fail(g->ctx, GALP_EHIP, std::string("sort kernel launch: ") + hipGetErrorString(e));
    return GALP_OK;
}

int launch_sort_iteration(galp_graph *g, const IterView &w) {
    if (!g->sort) return fail(g->ctx, GALP_ESTATE, "internal: sort state missing");
    return g->off64 ? sort_iteration<int64_t>(g, w) : sort_iteration<int32_t>(g, w);
}

}  // namespace galp
