// probe.hip -- the gather ceiling of a graph's access pattern (diagnostic).
//
// No reference counterpart.  Every LP kernel reads each neighbour id once
// (sequentially, non-temporal) and gathers that neighbour's label (random);
// counting and label writes come on top.  k_gather_probe does only the first
// two, with the same load types and the same graph layout (renumbering
// included), as many loads in flight per thread as it can hold, and no
// output but a checksum.  Its rate is the best an LP kernel can reach on this
// graph's gathers, so an LP kernel's rate is best read as a fraction of it.

#include "device_common.hpp"
#include "internal.hpp"

namespace galp {

constexpr int kProbeThreads = 256;
constexpr int kProbeUnroll = 8;   // gathers in flight per thread

// Thread t of the grid takes edges base + k*T + t (k < kProbeUnroll) of each
// chunk of kProbeUnroll*T edges, so each k is one coalesced read of ids.
__global__ void __launch_bounds__(kProbeThreads)
k_gather_probe(const int32_t *__restrict__ nbr, const int32_t *__restrict__ lab, int64_t m,
               unsigned long long *__restrict__ sum) {
    const int64_t nthreads = (int64_t)gridDim.x * kProbeThreads;
    const int64_t t = (int64_t)blockIdx.x * kProbeThreads + threadIdx.x;
    unsigned long long s = 0;
    for (int64_t base = 0; base < m; base += kProbeUnroll * nthreads) {
        int32_t u[kProbeUnroll];
#pragma unroll
        for (int k = 0; k < kProbeUnroll; ++k) {
            const int64_t e = base + k * nthreads + t;
            u[k] = __builtin_nontemporal_load(&nbr[e < m ? e : 0]);
        }
#pragma unroll
        for (int k = 0; k < kProbeUnroll; ++k) {
            const int64_t e = base + k * nthreads + t;
            const uint32_t L = (uint32_t)lab[u[k]];
            s += e < m ? L : 0u;
        }
    }
#pragma unroll
    for (int o = kWave / 2; o > 0; o >>= 1) s += __shfl_xor(s, o);
    if (lane_id() == 0 && s) atomicAdd(sum, s);
}

}  // namespace galp

using namespace galp;

extern "C" {

GALP_API int galp_graph_gather_probe(galp_graph *g, int32_t reps, double *ms_per_pass, uint64_t *checksum) {
    if (!g || reps < 1 || !ms_per_pass) return GALP_EINVAL;
    if (g->ooc || !g->d_nbr) return fail(g->ctx, GALP_EUNSUPPORTED, "the gather probe runs in core");
    *ms_per_pass = 0.0;
    if (checksum) *checksum = 0;
    if (g->m == 0) return GALP_OK;
    unsigned long long *d_sum = nullptr;
    if (hipMalloc((void **)&d_sum, sizeof(unsigned long long)) != hipSuccess) {
        (void)hipGetLastError();
        return fail(g->ctx, GALP_ENOMEM, "gather probe: device word");
    }
    hipEvent_t e0 = nullptr, e1 = nullptr;
    int rc = GALP_OK;
    if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess)
        rc = fail(g->ctx, GALP_EHIP, "gather probe: events");
    const int grid = g->ctx->prop.multiProcessorCount * 8;   // 32 waves per CU
    const int32_t *lab = g->d_lab[g->cur];
    if (rc == GALP_OK) rc = dev_fill_u32(g, (uint32_t *)d_sum, 2, 0u);
    if (rc == GALP_OK) {
        // one pass for the checksum (and warm-up), then `reps` timed passes
        k_gather_probe<<<grid, kProbeThreads, 0, g->stream>>>(g->d_nbr, lab, g->m, d_sum);
        unsigned long long host = 0;
        if (hipMemcpyAsync(&host, d_sum, sizeof(host), hipMemcpyDeviceToHost, g->stream) != hipSuccess ||
            hipStreamSynchronize(g->stream) != hipSuccess)
            rc = fail(g->ctx, GALP_EHIP, "gather probe: checksum");
        if (checksum) *checksum = host;
    }
    if (rc == GALP_OK) {
        (void)hipEventRecord(e0, g->stream);
        for (int r = 0; r < reps; ++r)
            k_gather_probe<<<grid, kProbeThreads, 0, g->stream>>>(g->d_nbr, lab, g->m, d_sum);
        (void)hipEventRecord(e1, g->stream);
        float ms = 0.f;
        if (hipEventSynchronize(e1) != hipSuccess || hipEventElapsedTime(&ms, e0, e1) != hipSuccess)
            rc = fail(g->ctx, GALP_EHIP, "gather probe: timing");
        *ms_per_pass = (double)ms / reps;
    }
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    (void)hipFree(d_sum);
    return rc;
}

}  // extern "C"
