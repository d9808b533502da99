// graph_dev.hip -- device-side graph preparation and label bookkeeping.
//
// Replaces InCoreLP::transfer_data (label_propagation/src/common/
// incore.cuh:59-73), initialize_labels (src/dpp/kernel.cuh:15-22) and the
// final label copy of LFHTBase::run (src/lfht/lfht.cuh:103).  The renumbering
// kernel is new: rows are rewritten in degree-descending order so every
// degree class is one contiguous id range and the most-gathered labels (the
// hubs') share cache lines.  Each row keeps its neighbour order, which keeps
// the tie rule's last positions, so labels mapped back through the inverse
// permutation equal the labels of the original numbering.

#include <algorithm>

#include "device_common.hpp"
#include "internal.hpp"

namespace galp {

__global__ void k_iota2(int32_t *a, int32_t *b, int64_t n) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x) {
        a[i] = (int32_t)i;
        b[i] = (int32_t)i;
    }
}

__global__ void k_fill_u32(uint32_t *p, int64_t n, uint32_t val) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        p[i] = val;
}

__global__ void k_reset_control(unsigned long long *counter, int32_t *stop, int32_t *iters_done,
                                int32_t *err, int reset_done) {
    *counter = 0ull;
    *stop = 0;
    if (reset_done) {
        *iters_done = 0;
        *err = 0;
    }
}

// One wave per internal row k: copy row iperm[k] of the original CSR,
// mapping every neighbour through perm, in the original order.
template <typename E>
__global__ void k_permute_rows(const int64_t *__restrict__ off_old, const int32_t *__restrict__ nbr_old,
                               const int32_t *__restrict__ iperm, const int32_t *__restrict__ perm,
                               const E *__restrict__ off_new, int32_t *__restrict__ nbr_new, int64_t n) {
    const int64_t wave = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) / kWave;
    const int64_t nwaves = ((int64_t)gridDim.x * blockDim.x) / kWave;
    const int lane = lane_id();
    for (int64_t k = wave; k < n; k += nwaves) {
        const int64_t v = iperm[k];
        const int64_t b = off_old[v], d = off_old[v + 1] - b;
        const int64_t nb = (int64_t)off_new[k];
        for (int64_t j = lane; j < d; j += kWave) nbr_new[nb + j] = perm[nbr_old[b + j]];
    }
}

// out[v] = iperm[labels[perm[v]]]: internal labels back to original ids.
__global__ void k_map_labels(const int32_t *__restrict__ lab, const int32_t *__restrict__ perm,
                             const int32_t *__restrict__ iperm, int32_t *__restrict__ out, int64_t n) {
    for (int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; v < n;
         v += (int64_t)gridDim.x * blockDim.x)
        out[v] = iperm[lab[perm[v]]];
}

// lab[perm[v]] = perm[in[v]]: original labels into the internal numbering.
__global__ void k_unmap_labels(const int32_t *__restrict__ in, const int32_t *__restrict__ perm,
                               int32_t *__restrict__ lab, int64_t n) {
    for (int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; v < n;
         v += (int64_t)gridDim.x * blockDim.x)
        lab[perm[v]] = perm[in[v]];
}

static int grid1d(int64_t n, int threads) {
    int64_t b = (n + threads - 1) / threads;
    if (b > 65536) b = 65536;
    return (int)(b < 1 ? 1 : b);
}

int dev_init_labels(galp_graph *g) {
    if (g->n) k_iota2<<<grid1d(g->n, 256), 256, 0, g->stream>>>(g->d_lab[0], g->d_lab[1], g->n);
    k_reset_control<<<1, 1, 0, g->stream>>>(g->d_counter, g->d_stop, g->d_iters_done, g->d_err, 1);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return fail(g->ctx, GALP_EHIP, std::string("init kernels: ") + hipGetErrorString(e));
    return GALP_OK;
}

int dev_clear_stop(galp_graph *g) {
    k_reset_control<<<1, 1, 0, g->stream>>>(g->d_counter, g->d_stop, g->d_iters_done, g->d_err, 0);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return fail(g->ctx, GALP_EHIP, std::string("reset kernel: ") + hipGetErrorString(e));
    return GALP_OK;
}

int dev_fill_u32(galp_graph *g, uint32_t *p, int64_t n, uint32_t val) {
    if (n) k_fill_u32<<<grid1d(n, 256), 256, 0, g->stream>>>(p, n, val);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return fail(g->ctx, GALP_EHIP, std::string("fill kernel: ") + hipGetErrorString(e));
    return GALP_OK;
}

int dev_permute_rows(galp_graph *g, const int64_t *d_off_old, const int32_t *d_nbr_old,
                     int32_t *d_nbr_new) {
    if (!g->n) return GALP_OK;
    const int64_t blocks = std::min<int64_t>((g->n + 3) / 4, 65536);
    if (g->off64)
        k_permute_rows<int64_t><<<(int)blocks, 256, 0, g->stream>>>(
            d_off_old, d_nbr_old, g->d_iperm, g->d_perm, (const int64_t *)g->d_off, d_nbr_new, g->n);
    else
        k_permute_rows<int32_t><<<(int)blocks, 256, 0, g->stream>>>(
            d_off_old, d_nbr_old, g->d_iperm, g->d_perm, (const int32_t *)g->d_off, d_nbr_new, g->n);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return fail(g->ctx, GALP_EHIP, std::string("permute kernel: ") + hipGetErrorString(e));
    return GALP_OK;
}

int dev_map_labels(galp_graph *g, const int32_t *lab, int32_t *out) {
    if (g->n) k_map_labels<<<grid1d(g->n, 256), 256, 0, g->stream>>>(lab, g->d_perm, g->d_iperm, out, g->n);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return fail(g->ctx, GALP_EHIP, std::string("map kernel: ") + hipGetErrorString(e));
    return GALP_OK;
}

int dev_unmap_labels(galp_graph *g, const int32_t *in, int32_t *lab) {
    if (g->n) k_unmap_labels<<<grid1d(g->n, 256), 256, 0, g->stream>>>(in, g->d_perm, lab, g->n);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return fail(g->ctx, GALP_EHIP, std::string("unmap kernel: ") + hipGetErrorString(e));
    return GALP_OK;
}

}  // namespace galp
