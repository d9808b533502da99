// device_common.hpp -- wave64 helpers shared by the gfx950 LP kernels.
//
// Everything here is written for CDNA wavefronts of 64 lanes: 64-bit ballot
// masks, lane-index shuffles and reductions across full waves or aligned
// power-of-two lane groups.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace galp {

constexpr int kWave = 64;
constexpr uint32_t kEmpty = 0xffffffffu;   // labels are vertex ids < 2^31

// Fibonacci (multiplicative) hashing into a power-of-two table of 2^bits
// slots (Knuth, TAOCP vol. 3, 6.4).  Labels are arbitrary vertex ids, and the
// top bits of the product spread consecutive ids across the table.
__device__ __forceinline__ uint32_t hash_slot(uint32_t key, int bits) {
    return (key * 0x9E3779B1u) >> (32 - bits);
}

// Label partition of a hub's label (P partitions): a multiplicative hash
// scaled to [0, P), so a hub's distinct labels spread evenly over its
// partitions whatever their values.  Both methods' hub paths use it.
__device__ __forceinline__ uint32_t hub_partition(uint32_t L, uint32_t P) {
    return (uint32_t)(((unsigned long long)(L * 0x85EBCA77u) * P) >> 32);
}

// Round of a label when one hub partition is merged in R rounds (R a power
// of two, at most 4096): bits of a third multiplicative hash, so labels
// that share a partition spread over the rounds.
__device__ __forceinline__ uint32_t sub_round(uint32_t L, uint32_t R) {
    return ((L * 0xC2B2AE3Du) >> 20) & (R - 1u);
}

__device__ __forceinline__ int lane_id() { return (int)(threadIdx.x & (kWave - 1)); }

// ---------------------------------------------------------------------------
// Lane moves inside aligned power-of-two lane groups.  Within a DPP row (16
// lanes) they are VALU data-parallel-primitive moves, not LDS permutes:
//   quad_perm [1,0,3,2], [2,3,0,1], [3,2,1,0]: lane i <- i ^ 1, i ^ 2, i ^ 3;
//   row_half_mirror: i <- i ^ 7 (8-lane halves); row_mirror: i <- i ^ 15.
// Lane i ^ 16 comes from ds_swizzle (bit mode: and 31, or 0, xor 16), which
// needs no address register; lane i ^ 32 from a permute.  All lanes of the
// wave must be active (DPP reads the source lane's register).
constexpr int kDppQuadXor1 = 0xB1;
constexpr int kDppQuadXor2 = 0x4E;
constexpr int kDppQuadXor3 = 0x1B;
constexpr int kDppRowMirror = 0x140;
constexpr int kDppRowHalfMirror = 0x141;
constexpr int kSwizzleXor16 = 0x401F;

template <int CTRL>
__device__ __forceinline__ int32_t dpp_mov(int32_t x) {
    return __builtin_amdgcn_update_dpp(0, x, CTRL, 0xF, 0xF, false);
}

// x of lane (lane ^ J), J a power of two below 64.
template <int J>
__device__ __forceinline__ int32_t lane_xor(int32_t x) {
    if constexpr (J == 1) return dpp_mov<kDppQuadXor1>(x);
    else if constexpr (J == 2) return dpp_mov<kDppQuadXor2>(x);
    else if constexpr (J == 4) return dpp_mov<kDppQuadXor3>(dpp_mov<kDppRowHalfMirror>(x));   // ^7 then ^3
    else if constexpr (J == 8) return dpp_mov<kDppRowHalfMirror>(dpp_mov<kDppRowMirror>(x));  // ^15 then ^7
    else if constexpr (J == 16) return __builtin_amdgcn_ds_swizzle(x, kSwizzleXor16);
    else return __shfl_xor(x, J);
}

// Max over an aligned group of G lanes (G a power of two, G <= 64).
template <int G, typename T>
__device__ __forceinline__ T group_max(T x) {
    if constexpr (sizeof(T) == 4) {   // butterfly of lane moves
        auto step = [&](int32_t y) {
            const T o = (T)y;
            x = x > o ? x : o;
        };
        if constexpr (G >= 64) step(lane_xor<32>((int32_t)x));
        if constexpr (G >= 32) step(lane_xor<16>((int32_t)x));
        if constexpr (G >= 16) step(lane_xor<8>((int32_t)x));
        if constexpr (G >= 8) step(lane_xor<4>((int32_t)x));
        if constexpr (G >= 4) step(lane_xor<2>((int32_t)x));
        if constexpr (G >= 2) step(lane_xor<1>((int32_t)x));
        return x;
    } else {
#pragma unroll
        for (int s = G / 2; s > 0; s >>= 1) {
            T y = __shfl_xor(x, s, G);
            x = x > y ? x : y;
        }
        return x;
    }
}

// Min over an aligned group of G <= 32 lanes.
template <int G>
__device__ __forceinline__ int32_t group_min(int32_t x) {
    if constexpr (G >= 32) x = min(x, lane_xor<16>(x));
    if constexpr (G >= 16) x = min(x, lane_xor<8>(x));
    if constexpr (G >= 8) x = min(x, lane_xor<4>(x));
    if constexpr (G >= 4) x = min(x, lane_xor<2>(x));
    if constexpr (G >= 2) x = min(x, lane_xor<1>(x));
    return x;
}

// Lanes i ^ k, k < K (K = 1, 2, 4, 8, 16 within a row), whose y equals x.
template <int K>
__device__ __forceinline__ uint32_t count_eq_row(int32_t x, int32_t y) {
    uint32_t c = (uint32_t)(y == x);
    if constexpr (K >= 2) c += (uint32_t)(dpp_mov<kDppQuadXor1>(y) == x);
    if constexpr (K >= 4) {
        c += (uint32_t)(dpp_mov<kDppQuadXor2>(y) == x);
        c += (uint32_t)(dpp_mov<kDppQuadXor3>(y) == x);
    }
    if constexpr (K >= 8) c += count_eq_row<4>(x, dpp_mov<kDppRowHalfMirror>(y));   // k = 7, 6, 5, 4
    if constexpr (K >= 16) c += count_eq_row<8>(x, dpp_mov<kDppRowMirror>(y));      // k = 15 .. 8
    return c;
}

// Lanes of this lane's aligned group of G <= 32 (itself included) whose x
// equals this lane's.
template <int G>
__device__ __forceinline__ uint32_t group_count_eq(int32_t x) {
    uint32_t c = count_eq_row<(G < 16 ? G : 16)>(x, x);
    if constexpr (G >= 32) c += count_eq_row<16>(x, lane_xor<16>(x));
    return c;
}

template <typename T>
__device__ __forceinline__ T wave_max(T x) { return group_max<kWave, T>(x); }

__device__ __forceinline__ uint32_t wave_sum_u32(uint32_t x) {
#pragma unroll
    for (int s = kWave / 2; s > 0; s >>= 1) x += __shfl_xor(x, s, kWave);
    return x;
}

// Block-wide max over NW waves; every thread gets the result.  The wave
// count is a template argument: blockDim is read from the dispatch packet
// with a vector load, which would drain every load in flight.
// `scratch` holds NW entries.  Contains barriers when NW > 1.
template <int NW, typename T>
__device__ __forceinline__ T block_max(T x, T *scratch) {
    x = wave_max(x);
    if constexpr (NW == 1) {
        return x;
    } else {
        const int w = (int)(threadIdx.x / kWave);
        if (lane_id() == 0) scratch[w] = x;
        __syncthreads();
        T r = scratch[0];
#pragma unroll
        for (int i = 1; i < NW; ++i) r = r > scratch[i] ? r : scratch[i];
        __syncthreads();
        return r;
    }
}

// Block-wide sum over NW waves; every thread gets the result.  Contains
// barriers when NW > 1.
template <int NW>
__device__ __forceinline__ uint32_t block_sum(uint32_t x, uint32_t *scratch) {
    x = wave_sum_u32(x);
    if constexpr (NW == 1) {
        return x;
    } else {
        const int w = (int)(threadIdx.x / kWave);
        if (lane_id() == 0) scratch[w] = x;
        __syncthreads();
        uint32_t r = 0;
#pragma unroll
        for (int i = 0; i < NW; ++i) r += scratch[i];
        __syncthreads();
        return r;
    }
}

// Block-wide exclusive prefix sum of one int per thread (NW waves); `scratch`
// holds NW entries.  Contains barriers.
template <int NW>
__device__ __forceinline__ int block_exclusive_scan(int x, int *scratch) {
    const int lane = lane_id();
    const int w = (int)(threadIdx.x / kWave);
    int inc = x;
#pragma unroll
    for (int s = 1; s < kWave; s <<= 1) {
        const int y = __shfl_up(inc, s);
        if (lane >= s) inc += y;
    }
    if (lane == kWave - 1) scratch[w] = inc;
    __syncthreads();
    int before = 0;
#pragma unroll
    for (int i = 0; i < NW; ++i)
        if (i < w) before += scratch[i];
    __syncthreads();
    return before + inc - x;
}

// Adds a per-thread count into *counter with one atomic per block of NW waves.
template <int NW>
__device__ __forceinline__ void block_count_add(uint32_t mine, unsigned long long *counter,
                                                uint32_t *scratch) {
    uint32_t s = wave_sum_u32(mine);
    if constexpr (NW == 1) {
        if (lane_id() == 0 && s) atomicAdd(counter, (unsigned long long)s);
    } else {
        const int w = (int)(threadIdx.x / kWave);
        __syncthreads();
        if (lane_id() == 0) scratch[w] = s;
        __syncthreads();
        if (threadIdx.x == 0) {
            unsigned long long t = 0;
#pragma unroll
            for (int i = 0; i < NW; ++i) t += scratch[i];
            if (t) atomicAdd(counter, t);
        }
    }
}

// Loads that bypass the per-CU vector L1, for table words other workgroups
// update with device-scope atomics.
__device__ __forceinline__ uint32_t load_coherent(const uint32_t *p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

}  // namespace galp
