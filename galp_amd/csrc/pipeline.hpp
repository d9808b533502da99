// pipeline.hpp -- software pipeline over work items for the LP kernels.
#pragma once

#include "device_common.hpp"

namespace galp {

// ---------------------------------------------------------------------------
// Software pipeline shared by the table kernels.  A "work item" is a vertex
// (LDS kernels), a hub tile or a candidate-path tile; a team of T lanes
// (lane index `tid`) walks items first, first + stride, ...; Policy supplies
//   struct Item { bool valid; int len; ... };
//   Item item(int64_t i)               -- uniform per team (scalar loads)
//   int64_t edge0(const Item &)        -- absolute index of the item's edge 0
//   void process(const Item &, const int32_t (&lab)[EPT])
// The pipeline keeps: labels of item t (being processed), neighbour ids of
// item t+1, descriptor of item t+2; a period-2 unroll ping-pongs the label
// registers so no in-flight load is ever copied.
template <int T, int EPT, class Policy>
__device__ __forceinline__ void pipelined(Policy &pol, const int32_t *__restrict__ nbr,
                                          const int32_t *__restrict__ prev, int64_t first, int64_t stride,
                                          int tid) {
    using Item = typename Policy::Item;
    int64_t i = first;
    // Every lane issues every load (from a clamped, valid address) and
    // selects afterwards: a load under a lane condition becomes a branch,
    // and the compiler then drains all outstanding loads at the join.
    // Loaded values are kept raw: masking them right after the load would
    // wait for it.  Slots past an item's length hold junk that process()
    // never reads (it checks pos < len).
    auto load_nbr = [&](int32_t (&u)[EPT], const Item &it) {
        const int64_t e0 = it.valid ? pol.edge0(it) : 0;
#pragma unroll
        for (int k = 0; k < EPT; ++k) {
            const int pos = k * T + tid;
            const bool ok = it.valid && pos < it.len;
            u[k] = __builtin_nontemporal_load(&nbr[ok ? e0 + pos : 0]);
        }
    };
    auto gather = [&](int32_t (&lab)[EPT], const int32_t (&u)[EPT], const Item &it) {
#pragma unroll
        for (int k = 0; k < EPT; ++k) {
            const int pos = k * T + tid;
            const bool ok = it.valid && pos < it.len;
            lab[k] = prev[ok ? u[k] : 0];
        }
    };
    Item s0 = pol.item(i), s1 = pol.item(i + stride), s2 = pol.item(i + 2 * stride);
    int32_t u[EPT], labA[EPT], labB[EPT];
    load_nbr(u, s0);
    gather(labA, u, s0);
    load_nbr(u, s1);
    while (s0.valid) {
        gather(labB, u, s1);    // labels of item t+1
        load_nbr(u, s2);        // neighbours of item t+2
        Item s3 = pol.item(i + 3 * stride);
        pol.process(s0, labA);  // item t
        s0 = s1; s1 = s2; s2 = s3; i += stride;
        if (!s0.valid) break;
        gather(labA, u, s1);
        load_nbr(u, s2);
        s3 = pol.item(i + 3 * stride);
        pol.process(s0, labB);
        s0 = s1; s1 = s2; s2 = s3; i += stride;
    }
}

}  // namespace galp
