// partsort.hpp -- rows stored grouped by neighbour part (set-up kernel).
#pragma once

#include "device_common.hpp"
#include "internal.hpp"

namespace galp {

// The part of a neighbour id: its 128-byte label line modulo kNbrParts.
__device__ __forceinline__ int nbr_part(int32_t u) { return (int)(((uint32_t)u >> kPartLineBits) & (kNbrParts - 1)); }

// One wave per listed row r (vertex rows[r]) writes the row into its compact
// range [roff[r], roff[r+1]) of pnbr grouped by neighbour part, stable (each
// part in row order), with each edge's position in the row in ppos, and the
// row's edge count per part in pcount[r * kNbrParts + p].  Per 64-edge chunk,
// ballots over the part bits give each lane its rank among the chunk's lanes
// of its part; per-part running bases in LDS carry the ranks across chunks.
template <typename E>
__global__ void __launch_bounds__(256)
k_rows_partsort(const E *__restrict__ off, const int32_t *__restrict__ nbr, const int32_t *__restrict__ rows,
                int64_t nrows, const int64_t *__restrict__ roff, int32_t *__restrict__ pnbr,
                uint32_t *__restrict__ ppos, int32_t *__restrict__ pcount) {
    static_assert(kNbrParts <= kWave && (kNbrParts & (kNbrParts - 1)) == 0, "parts fit one wave");
    __shared__ uint32_t sbase[256 / kWave][kNbrParts];
    const int w = (int)(threadIdx.x / kWave), lane = lane_id();
    const int64_t r = (int64_t)blockIdx.x * (256 / kWave) + w;
    if (r >= nrows) return;   // the whole wave: no block barrier below
    const int64_t v = rows[r];
    const int64_t rb = (int64_t)off[v], d = (int64_t)off[v + 1] - rb, ob = roff[r];
    if (lane < kNbrParts) sbase[w][lane] = 0;
    __builtin_amdgcn_wave_barrier();
    for (int64_t e = lane; e < d; e += kWave) atomicAdd(&sbase[w][nbr_part(nbr[rb + e])], 1u);
    __builtin_amdgcn_wave_barrier();
    const uint32_t c = lane < kNbrParts ? sbase[w][lane] : 0u;
    if (lane < kNbrParts) pcount[r * kNbrParts + lane] = (int32_t)c;
    uint32_t inc = c;
#pragma unroll
    for (int sft = 1; sft < kWave; sft <<= 1) {
        const uint32_t y = (uint32_t)__shfl_up((int)inc, sft);
        if (lane >= sft) inc += y;
    }
    __builtin_amdgcn_wave_barrier();
    if (lane < kNbrParts) sbase[w][lane] = inc - c;
    __builtin_amdgcn_wave_barrier();
    const unsigned long long lt = (1ull << lane) - 1ull;
    for (int64_t e0 = 0; e0 < d; e0 += kWave) {
        const int64_t e = e0 + lane;
        const bool valid = e < d;
        const int32_t u = valid ? nbr[rb + e] : 0;
        const int pt = nbr_part(u);
        unsigned long long peers = __ballot(valid);
#pragma unroll
        for (int b = 0; (1 << b) < kNbrParts; ++b) {
            const bool bit = (pt >> b) & 1;
            const unsigned long long bb = __ballot(bit);
            peers &= bit ? bb : ~bb;
        }
        const uint32_t base = sbase[w][pt];
        if (valid) {
            const int64_t dst = ob + (int64_t)base + __popcll(peers & lt);
            pnbr[dst] = u;
            ppos[dst] = (uint32_t)e;
        }
        __builtin_amdgcn_wave_barrier();
        if (valid && (peers & lt) == 0) sbase[w][pt] = base + (uint32_t)__popcll(peers);
        __builtin_amdgcn_wave_barrier();
    }
}

// Tiles of part-sorted rows: per row, its parts in order; a part of more
// than `tile` edges is cut into `tile`-edge chunks, smaller parts are
// grouped while a tile holds at most `tile` edges (consecutive parts are
// adjacent in the compact copy).  emit(row, e0, len, part) gets each tile
// (part = its first part), rows in order.
template <class Emit>
void part_tiles(int64_t nrows, const std::vector<int64_t> &roff, const std::vector<int32_t> &count, int64_t tile,
                Emit &&emit) {
    for (int64_t r = 0; r < nrows; ++r) {
        int64_t at = roff[(size_t)r], open_e0 = 0, open_len = 0;
        int open_part = -1;
        for (int p = 0; p < kNbrParts; ++p) {
            int64_t c = count[(size_t)(r * kNbrParts + p)];
            if (c == 0) continue;
            if (open_len > 0 && open_len + c > tile) {
                emit(r, open_e0, open_len, open_part);
                open_len = 0;
            }
            while (c > tile) {   // full chunks of a big part
                emit(r, at, tile, p);
                at += tile;
                c -= tile;
            }
            if (open_len == 0) {
                open_e0 = at;
                open_part = p;
            }
            open_len += c;
            at += c;
        }
        if (open_len > 0) emit(r, open_e0, open_len, open_part);
    }
}

}  // namespace galp
