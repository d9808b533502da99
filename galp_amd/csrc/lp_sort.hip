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
//   - hubs (> 4096): a segmented LSD radix sort (8-bit digits, as many passes
//     as the label width needs) of the hub edges in a compact buffer, then a
//     per-tile run-length pass; tiles reduce (count, last) with one 64-bit
//     atomicMax per tile per hub.
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
            after = lane == kWave - 1 ? kPad : (uint32_t)__shfl_down((int)L[0], 1);
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

template <int T, int EPT, typename E>
struct LdsSortPolicy {
    static constexpr int N = T * EPT;
    static constexpr int NW = T / kWave;
    static constexpr int M = N / NW;        // elements per wave
    static constexpr int CH = M / kWave;    // 64-element chunks per wave
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
        const int lane = lane_id();
        const int w = tid / kWave;
        const unsigned long long lt = (1ull << lane) - 1ull;
        __syncthreads();   // the previous vertex's readers are done
#pragma unroll
        for (int k = 0; k < EPT; ++k) {
            const int pos = k * T + tid;
            if (pos < len) {
                skey[pos] = (uint32_t)lab[k];
                spos[pos] = (uint16_t)pos;
            }
        }
        int src = 0;
        for (int p = 0; p < passes; ++p) {
            const int shift = 8 * p;
            const uint32_t *ks = skey + src * N;
            const uint16_t *ps = spos + src * N;
            uint32_t *kd = skey + (src ^ 1) * N;
            uint16_t *pd = spos + (src ^ 1) * N;
            for (int i = lane; i < 256; i += kWave) whist[w * 256 + i] = 0;
            __syncthreads();   // keys written; histogram rows clear
            // A: per-wave digit counts (the lowest lane of each peer group adds)
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
// Hubs.  The edges of hub h occupy [hoff[h], hoff[h+1]) of a compact key
// buffer, in kHubTile-edge tiles (the hub path's tile list: tiles of a hub
// are consecutive).  Each LSD pass p sorts by bits [8p, 8p+8) of the label:
//   hist:    per tile, the 256-digit histogram, stored hub-major, then
//            digit-major, then tile: idx = tile0[h]*256 + d*ntiles[h] + tl;
//   scan:    one exclusive scan over that array gives, for every (hub,
//            digit, tile), the tile's first output slot -- hub-major order
//            keeps every hub inside its own range, so the sort is segmented;
//   scatter: per tile, a stable multisplit: wave w owns elements
//            [512w, 512w+512) in order; 8 ballots give each lane the lanes
//            of its chunk with the same digit, whose popcount below it is
//            its rank; per-wave running counters in LDS carry the ranks
//            across chunks.
// The payload is the element's position in its row.  Pass 0 gathers the
// labels (prev[nbr[e]]) into the key buffer and takes the position from the
// index.
constexpr int kSortTileT = 256;
constexpr int kSortEPT = kHubTile / kSortTileT;   // 8
static_assert(kHubTile == 2048 && kSortEPT == 8, "hub tile layout");

struct HubTile {
    int h;
    int64_t tl;      // tile index within its hub
    int64_t c0;      // first compact slot of the tile
    int len;         // elements in the tile
    int64_t hbase;   // tile0[h] * 256
    int64_t nth;     // tiles of the hub
};

__device__ __forceinline__ HubTile hub_tile_at(int64_t t, const int32_t *tile_hub, const int32_t *hub_tile0,
                                               const int32_t *hub_ntiles, const int64_t *hoff) {
    HubTile x;
    x.h = tile_hub[t];
    const int64_t t0 = hub_tile0[x.h];
    x.tl = t - t0;
    x.nth = hub_ntiles[x.h];
    x.hbase = t0 * 256;
    const int64_t hb = hoff[x.h], he = hoff[x.h + 1];
    x.c0 = hb + x.tl * kHubTile;
    x.len = (int)min((int64_t)kHubTile, he - x.c0);
    return x;
}

// Pass-0 histogram; also gathers the labels into the key buffer.
template <typename E>
__global__ void __launch_bounds__(kSortTileT)
k_hs_gather_hist(const E *__restrict__ off, const int32_t *__restrict__ nbr, const int32_t *__restrict__ prev,
                 const int32_t *__restrict__ hub_vid, const int32_t *__restrict__ tile_hub,
                 const int32_t *__restrict__ hub_tile0, const int32_t *__restrict__ hub_ntiles,
                 const int64_t *__restrict__ hoff, int64_t tbase, int64_t ntiles, uint32_t *__restrict__ keys,
                 uint32_t *__restrict__ hist, const int32_t *__restrict__ stop) {
    __shared__ uint32_t sh[256];
    if (*stop) return;
    for (int64_t t = tbase + blockIdx.x; t < ntiles; t += gridDim.x) {
        const HubTile x = hub_tile_at(t, tile_hub, hub_tile0, hub_ntiles, hoff);
        const int64_t rb = (int64_t)off[hub_vid[x.h]] + x.tl * kHubTile;
        sh[threadIdx.x] = 0;
        __syncthreads();
        int32_t u[kSortEPT];
#pragma unroll
        for (int k = 0; k < kSortEPT; ++k) {
            const int j = k * kSortTileT + (int)threadIdx.x;
            u[k] = __builtin_nontemporal_load(&nbr[rb + (j < x.len ? j : 0)]);
        }
#pragma unroll
        for (int k = 0; k < kSortEPT; ++k) {
            const int j = k * kSortTileT + (int)threadIdx.x;
            const uint32_t L = (uint32_t)prev[u[k]];
            if (j < x.len) {
                keys[x.c0 + j] = L;
                atomicAdd(&sh[L & 0xffu], 1u);
            }
        }
        __syncthreads();
        hist[x.hbase + (int64_t)threadIdx.x * x.nth + x.tl] = sh[threadIdx.x];
        __syncthreads();
    }
}

// Histogram of digit `shift` for passes >= 1.
__global__ void __launch_bounds__(kSortTileT)
k_hs_hist(const int32_t *__restrict__ tile_hub, const int32_t *__restrict__ hub_tile0,
          const int32_t *__restrict__ hub_ntiles, const int64_t *__restrict__ hoff, int64_t tbase, int64_t ntiles,
          const uint32_t *__restrict__ keys, int shift, uint32_t *__restrict__ hist,
          const int32_t *__restrict__ stop) {
    __shared__ uint32_t sh[256];
    if (*stop) return;
    for (int64_t t = tbase + blockIdx.x; t < ntiles; t += gridDim.x) {
        const HubTile x = hub_tile_at(t, tile_hub, hub_tile0, hub_ntiles, hoff);
        sh[threadIdx.x] = 0;
        __syncthreads();
        uint32_t kv[kSortEPT];
#pragma unroll
        for (int k = 0; k < kSortEPT; ++k) {
            const int j = k * kSortTileT + (int)threadIdx.x;
            kv[k] = keys[x.c0 + (j < x.len ? j : 0)];
        }
#pragma unroll
        for (int k = 0; k < kSortEPT; ++k) {
            const int j = k * kSortTileT + (int)threadIdx.x;
            if (j < x.len) atomicAdd(&sh[(kv[k] >> shift) & 0xffu], 1u);
        }
        __syncthreads();
        hist[x.hbase + (int64_t)threadIdx.x * x.nth + x.tl] = sh[threadIdx.x];
        __syncthreads();
    }
}

// Exclusive scan of n uint32 into int64, in three launches: per-chunk sums,
// a one-block scan of the sums, per-chunk scans from their bases.
constexpr int kScanT = 256;
constexpr int kScanPer = 16;
constexpr int kScanChunk = kScanT * kScanPer;   // 4096

template <int NW>
__device__ __forceinline__ int64_t block_exclusive_sum(int64_t x, int64_t *sw, int64_t *total) {
    const int lane = lane_id();
    const int w = (int)(threadIdx.x / kWave);
    int64_t inc = x;
#pragma unroll
    for (int s = 1; s < kWave; s <<= 1) {
        const int64_t y = __shfl_up(inc, s);
        if (lane >= s) inc += y;
    }
    if (lane == kWave - 1) sw[w] = inc;
    __syncthreads();
    int64_t before = 0, all = 0;
#pragma unroll
    for (int i = 0; i < NW; ++i) {
        if (i < w) before += sw[i];
        all += sw[i];
    }
    __syncthreads();
    *total = all;
    return before + inc - x;
}

__global__ void __launch_bounds__(kScanT)
k_scan_reduce(const uint32_t *__restrict__ in, int64_t n, int64_t *__restrict__ bsum, const int32_t *__restrict__ stop) {
    __shared__ unsigned long long sred64[kScanT / kWave];
    if (*stop) return;
    const int64_t base = (int64_t)blockIdx.x * kScanChunk;
    unsigned long long s = 0;
#pragma unroll
    for (int k = 0; k < kScanPer; ++k) {
        const int64_t i = base + (int64_t)k * kScanT + threadIdx.x;
        if (i < n) s += in[i];
    }
#pragma unroll
    for (int o = kWave / 2; o > 0; o >>= 1) s += __shfl_xor(s, o);
    if (lane_id() == 0) sred64[threadIdx.x / kWave] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned long long t = 0;
        for (int i = 0; i < kScanT / kWave; ++i) t += sred64[i];
        bsum[blockIdx.x] = (int64_t)t;
    }
}

__global__ void __launch_bounds__(1024)
k_scan_top(int64_t *__restrict__ bsum, int64_t nb, int64_t base, const int32_t *__restrict__ stop) {
    __shared__ int64_t sw[1024 / kWave];
    if (*stop) return;
    const int64_t per = (nb + 1023) / 1024;
    const int64_t b0 = (int64_t)threadIdx.x * per;
    int64_t s = 0;
    for (int64_t i = b0; i < b0 + per && i < nb; ++i) s += bsum[i];
    int64_t total;
    int64_t run = base + block_exclusive_sum<1024 / kWave>(s, sw, &total);
    for (int64_t i = b0; i < b0 + per && i < nb; ++i) {
        const int64_t v = bsum[i];
        bsum[i] = run;
        run += v;
    }
}

__global__ void __launch_bounds__(kScanT)
k_scan_down(const uint32_t *__restrict__ in, int64_t n, const int64_t *__restrict__ bsum,
            int64_t *__restrict__ out, const int32_t *__restrict__ stop) {
    __shared__ int64_t sw[kScanT / kWave];
    if (*stop) return;
    const int64_t base = (int64_t)blockIdx.x * kScanChunk;
    // thread t scans the kScanPer consecutive entries [base + t*kScanPer, ...)
    const int64_t i0 = base + (int64_t)threadIdx.x * kScanPer;
    uint32_t v[kScanPer];
    int64_t s = 0;
#pragma unroll
    for (int k = 0; k < kScanPer; ++k) {
        v[k] = i0 + k < n ? in[i0 + k] : 0u;
        s += v[k];
    }
    int64_t total;
    int64_t run = bsum[blockIdx.x] + block_exclusive_sum<kScanT / kWave>(s, sw, &total);
#pragma unroll
    for (int k = 0; k < kScanPer; ++k) {
        if (i0 + k < n) out[i0 + k] = run;
        run += v[k];
    }
}

// Lanes of this wave whose 8-bit digit equals this lane's, among `valid`.
__device__ __forceinline__ unsigned long long digit_peers(uint32_t d, bool valid) {
    unsigned long long peers = __ballot(valid);
#pragma unroll
    for (int b = 0; b < 8; ++b) {
        const bool bit = (d >> b) & 1u;
        const unsigned long long bb = __ballot(bit);
        peers &= bit ? bb : ~bb;
    }
    return peers;
}

// Stable scatter of one pass.  FIRST: source payload is the position itself.
template <bool FIRST>
__global__ void __launch_bounds__(kSortTileT)
k_hs_scatter(const int32_t *__restrict__ tile_hub, const int32_t *__restrict__ hub_tile0,
             const int32_t *__restrict__ hub_ntiles, const int64_t *__restrict__ hoff, int64_t tbase, int64_t ntiles,
             const uint32_t *__restrict__ ksrc, const uint32_t *__restrict__ psrc, int shift,
             const int64_t *__restrict__ scan, uint32_t *__restrict__ kdst, uint32_t *__restrict__ pdst,
             const int32_t *__restrict__ stop) {
    constexpr int NW = kSortTileT / kWave;   // 4 waves, 512 elements each
    constexpr int CH = kHubTile / NW / kWave;   // 8 chunks of 64 per wave
    __shared__ uint32_t whist[NW][256];
    __shared__ int64_t wbase[NW][256];
    if (*stop) return;
    const int lane = lane_id();
    const int w = (int)(threadIdx.x / kWave);
    const unsigned long long lt = (1ull << lane) - 1ull;
    for (int64_t t = tbase + blockIdx.x; t < ntiles; t += gridDim.x) {
        const HubTile x = hub_tile_at(t, tile_hub, hub_tile0, hub_ntiles, hoff);
        for (int i = threadIdx.x; i < NW * 256; i += kSortTileT) (&whist[0][0])[i] = 0;
        uint32_t kv[CH], pv[CH];
#pragma unroll
        for (int c = 0; c < CH; ++c) {
            const int j = w * (CH * kWave) + c * kWave + lane;
            const int jc = j < x.len ? j : 0;
            kv[c] = ksrc[x.c0 + jc];
            pv[c] = FIRST ? (uint32_t)(x.tl * kHubTile + j) : psrc[x.c0 + jc];
        }
        __syncthreads();
        unsigned long long peers[CH];
#pragma unroll
        for (int c = 0; c < CH; ++c) {
            const int j = w * (CH * kWave) + c * kWave + lane;
            const bool valid = j < x.len;
            const uint32_t d = (kv[c] >> shift) & 0xffu;
            peers[c] = digit_peers(d, valid);
            if (valid && (peers[c] & lt) == 0) whist[w][d] += (uint32_t)__popcll(peers[c]);   // lowest peer
        }
        __syncthreads();
        {
            const int d = (int)threadIdx.x;   // 256 threads = 256 digits
            int64_t b = scan[x.hbase + (int64_t)d * x.nth + x.tl];
#pragma unroll
            for (int ww = 0; ww < NW; ++ww) {
                wbase[ww][d] = b;
                b += whist[ww][d];
            }
        }
        __syncthreads();
#pragma unroll
        for (int c = 0; c < CH; ++c) {
            const int j = w * (CH * kWave) + c * kWave + lane;
            const bool valid = j < x.len;
            const uint32_t d = (kv[c] >> shift) & 0xffu;
            const int64_t base = wbase[w][d];
            if (valid) {
                const int64_t dst = base + __popcll(peers[c] & lt);
                kdst[dst] = kv[c];
                pdst[dst] = pv[c];
            }
            __builtin_amdgcn_wave_barrier();
            if (valid && (peers[c] & lt) == 0) wbase[w][d] = base + __popcll(peers[c]);
            __builtin_amdgcn_wave_barrier();
        }
        __syncthreads();
    }
}

// Run lengths of the sorted hub rows.  Thread t of a tile owns the
// kSortEPT consecutive slots [t*8, t*8+8); a block max-scan of run starts
// gives each run end its start; a run that began in an earlier tile finds
// its start by binary search over the hub's sorted keys.  One atomicMax of
// the tile's best (count << 32 | last) per tile.
__global__ void __launch_bounds__(kSortTileT)
k_hs_runs(const int32_t *__restrict__ tile_hub, const int32_t *__restrict__ hub_tile0,
          const int32_t *__restrict__ hub_ntiles, const int64_t *__restrict__ hoff, int64_t tbase, int64_t ntiles,
          const uint32_t *__restrict__ keys, const uint32_t *__restrict__ pos,
          unsigned long long *__restrict__ hub_best, const int32_t *__restrict__ stop) {
    constexpr int NW = kSortTileT / kWave;
    __shared__ int64_t sw[NW];
    __shared__ unsigned long long sred64[NW];
    if (*stop) return;
    for (int64_t t = tbase + blockIdx.x; t < ntiles; t += gridDim.x) {
        const HubTile x = hub_tile_at(t, tile_hub, hub_tile0, hub_ntiles, hoff);
        const int64_t hb = hoff[x.h], he = hoff[x.h + 1];
        const int j0 = (int)threadIdx.x * kSortEPT;
        uint32_t kv[kSortEPT + 1];
        int64_t last_start = -1;   // hub-local index of the last run start in my slots
#pragma unroll
        for (int k = 0; k <= kSortEPT; ++k) {
            const int64_t g = x.c0 + j0 + k;
            kv[k] = g < he ? keys[g] : 0xffffffffu;
        }
        const uint32_t kbefore = (x.c0 + j0 > hb) ? keys[x.c0 + j0 - 1] : 0xffffffffu;
#pragma unroll
        for (int k = 0; k < kSortEPT; ++k) {
            const int j = j0 + k;
            const uint32_t kp = k == 0 ? kbefore : kv[k - 1];
            const bool start = j < x.len && (x.c0 + j == hb || kp != kv[k]);
            if (start) last_start = x.c0 + j - hb;
        }
        // exclusive max-scan of last_start over threads (wave scan + waves)
        int64_t inc = last_start;
        const int lane = lane_id();
        const int w = (int)(threadIdx.x / kWave);
#pragma unroll
        for (int s = 1; s < kWave; s <<= 1) {
            const int64_t y = __shfl_up(inc, s);
            if (lane >= s) inc = inc > y ? inc : y;
        }
        if (lane == kWave - 1) sw[w] = inc;
        __syncthreads();
        int64_t before = -1;
#pragma unroll
        for (int i = 0; i < NW; ++i)
            if (i < w) before = before > sw[i] ? before : sw[i];
        __syncthreads();
        int64_t ex = __shfl_up(inc, 1);
        if (lane == 0) ex = -1;
        int64_t cur = ex > before ? ex : before;   // start of the run open at my first slot
        unsigned long long best = 0;
#pragma unroll
        for (int k = 0; k < kSortEPT; ++k) {
            const int j = j0 + k;
            if (j >= x.len) break;
            const uint32_t kp = k == 0 ? kbefore : kv[k - 1];
            const int64_t gl = x.c0 + j - hb;   // hub-local index
            if (gl == 0 || kp != kv[k]) cur = gl;
            const bool end = (x.c0 + j + 1 == he) || kv[k + 1] != kv[k];
            if (end) {
                int64_t st = cur;
                if (st < 0) {   // the run began in an earlier tile
                    int64_t lo = 0, hi = x.c0 - hb;
                    while (lo < hi) {
                        const int64_t mid = (lo + hi) >> 1;
                        if (keys[hb + mid] < kv[k]) lo = mid + 1;
                        else hi = mid;
                    }
                    st = lo;
                }
                const unsigned long long val = ((unsigned long long)(gl - st + 1) << 32) | pos[x.c0 + j];
                best = val > best ? val : best;
            }
        }
        best = block_max<NW>(best, sred64);
        if (threadIdx.x == 0 && best) atomicMax(&hub_best[x.h], best);
        __syncthreads();
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
    SortState *S = g->sort;
    if (!S) return;
    void *ptrs[] = {S->d_hoff, S->d_kA, S->d_kB, S->d_pA, S->d_pB, S->d_hist, S->d_scan, S->d_bsum};
    for (void *p : ptrs)
        if (p) (void)hipFree(p);
    delete S;
    g->sort = nullptr;
}

int sort_setup(galp_graph *g, const std::vector<int64_t> &hub_deg) {
    if (g->sort) return GALP_OK;
    SortState *S = new (std::nothrow) SortState();
    if (!S) return fail(g->ctx, GALP_ENOMEM, "host allocation failed");
    g->sort = S;
    const int bits = g->n > 1 ? 64 - __builtin_clzll((unsigned long long)(g->n - 1)) : 1;
    S->passes = (bits + 7) / 8;
    std::vector<int64_t> hoff(hub_deg.size() + 1, 0);
    for (size_t h = 0; h < hub_deg.size(); ++h) hoff[h + 1] = hoff[h] + hub_deg[h];
    S->eh = hoff.back();
    S->hoff_host = hoff;
    if (!g->nhubs) return GALP_OK;
    S->nscan = g->nhub_tiles * 256;
    S->nbsum = (S->nscan + kScanChunk - 1) / kScanChunk;
    int rc;
    if ((rc = salloc(g, &S->d_hoff, (int64_t)hoff.size())) || (rc = salloc(g, &S->d_kA, S->eh)) ||
        (rc = salloc(g, &S->d_kB, S->eh)) || (rc = salloc(g, &S->d_pA, S->eh)) || (rc = salloc(g, &S->d_pB, S->eh)) ||
        (rc = salloc(g, &S->d_hist, S->nscan)) || (rc = salloc(g, &S->d_scan, S->nscan)) ||
        (rc = salloc(g, &S->d_bsum, S->nbsum)))
        return rc;
    if (hipMemcpy(S->d_hoff, hoff.data(), 8 * hoff.size(), hipMemcpyHostToDevice) != hipSuccess)
        return fail(g->ctx, GALP_EHIP, "hub offsets upload failed");
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
            (const E *)w.off, w.nbr, prev, next, g->d_vlist, r.v0, r.nv, g->d_counter, g->d_stop, nullptr, 0);
    else
        k_group<G, SortGroupResolve, SortGroupResolve, false, E><<<grid_for(g, blocks, blocks_per_cu), 256, 0,
                                                                   g->stream>>>(
            (const E *)w.off, w.nbr, prev, next, nullptr, r.v0, r.nv, g->d_counter, g->d_stop, nullptr, 0);
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

// Hubs [h0, h1) with tiles [t0, t1): the histogram, scan and scatter of every
// pass cover just those tiles; the scan starts at the hubs' first compact
// slot, so the output lands at the same places as in a whole-graph pass.
template <typename E>
static int sort_hubs(galp_graph *g, const IterView &w, const int32_t *prev, int32_t *next, int tiles_per_cu = 8) {
    SortState *S = g->sort;
    const int64_t t0 = w.t0, t1 = w.t1, nt = t1 - t0, nh = w.h1 - w.h0;
    const int gt = grid_for(g, nt, tiles_per_cu);
    const int64_t nscan = nt * 256;
    const int nbs = (int)((nscan + kScanChunk - 1) / kScanChunk);
    uint32_t *hist = S->d_hist + t0 * 256;
    int64_t *scan = S->d_scan + t0 * 256;
    double hub_edges = 0.0;
    if (w.h0 == 0 && w.h1 == g->nhubs) hub_edges = (double)g->bucket[B_HUB].edges;
    PendingLaunch pl;
    prof_begin(g, "hub_sort", 8.0 * hub_edges, &pl);
    k_hs_gather_hist<E><<<gt, kSortTileT, 0, g->stream>>>((const E *)w.off, w.nbr, prev, g->d_hub_vid,
                                                          g->d_tile_hub, g->d_hub_tile0, g->d_hub_ntiles, S->d_hoff,
                                                          t0, t1, S->d_kA, S->d_hist, g->d_stop);
    uint32_t *ks = S->d_kA, *ps = S->d_pA, *kd = S->d_kB, *pd = S->d_pB;
    for (int p = 0; p < S->passes; ++p) {
        if (p > 0)
            k_hs_hist<<<gt, kSortTileT, 0, g->stream>>>(g->d_tile_hub, g->d_hub_tile0, g->d_hub_ntiles, S->d_hoff, t0,
                                                        t1, ks, 8 * p, S->d_hist, g->d_stop);
        k_scan_reduce<<<nbs, kScanT, 0, g->stream>>>(hist, nscan, S->d_bsum, g->d_stop);
        k_scan_top<<<1, 1024, 0, g->stream>>>(S->d_bsum, nbs, S->hoff_host[(size_t)w.h0], g->d_stop);
        k_scan_down<<<nbs, kScanT, 0, g->stream>>>(hist, nscan, S->d_bsum, scan, g->d_stop);
        if (p == 0)
            k_hs_scatter<true><<<gt, kSortTileT, 0, g->stream>>>(g->d_tile_hub, g->d_hub_tile0, g->d_hub_ntiles,
                                                                 S->d_hoff, t0, t1, ks, nullptr, 0, S->d_scan, kd, pd,
                                                                 g->d_stop);
        else
            k_hs_scatter<false><<<gt, kSortTileT, 0, g->stream>>>(g->d_tile_hub, g->d_hub_tile0, g->d_hub_ntiles,
                                                                  S->d_hoff, t0, t1, ks, ps, 8 * p, S->d_scan, kd,
                                                                  pd, g->d_stop);
        std::swap(ks, kd);
        std::swap(ps, pd);
    }
    k_hs_runs<<<gt, kSortTileT, 0, g->stream>>>(g->d_tile_hub, g->d_hub_tile0, g->d_hub_ntiles, S->d_hoff, t0, t1,
                                                 ks, ps, g->d_hub_best, g->d_stop);
    k_hs_final<E><<<grid_for(g, (nh + 255) / 256, 4), 256, 0, g->stream>>>(
        (const E *)w.off, w.nbr, g->d_hub_vid + w.h0, nh, g->d_hub_best + w.h0, prev, next, g->d_counter, g->d_stop);
    prof_end(g, &pl);
    return GALP_OK;
}

// The sort method's configuration of each degree class (g->tune[b]): option
// 0 is the default, 1 and 2 the alternatives the autotuner times --
// workgroup size x keys per thread, the radix sort or the bitonic network,
// blocks per CU; for lane groups the grid; for hubs the tiles per CU of the
// histogram and scatter grids.  Every option computes the same labels.
template <typename E>
static int sort_class(galp_graph *g, const IterView &w, int b, const int32_t *prev, int32_t *next) {
    const int c = g->tune[b];
    switch (b) {
    case B_HUB:
        if (w.h1 > w.h0) return sort_hubs<E>(g, w, prev, next, c == 1 ? 4 : c == 2 ? 16 : 8);
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
