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

__device__ __forceinline__ int lane_id() { return (int)(threadIdx.x & (kWave - 1)); }

// Max over an aligned group of G lanes (G a power of two, G <= 64).
template <int G, typename T>
__device__ __forceinline__ T group_max(T x) {
#pragma unroll
    for (int s = G / 2; s > 0; s >>= 1) {
        T y = __shfl_xor(x, s, G);
        x = x > y ? x : y;
    }
    return x;
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
