// group_kernel.hpp -- the degree <= 64 driver shared by both LP methods.
#pragma once

#include "device_common.hpp"

namespace galp {

// Degree <= G: one aligned group of G lanes per vertex; lane p holds the
// label of the p-th neighbour (-1 past the degree).  Resolve::best<G>(L)
// returns the group's max of (count << 8 | last), the same in every lane of
// the group, `last` being a lane index (a CSR position in the row), so the
// winner is the label of lane `last`.
// One wave holds 64/G vertices per round; rounds are software-pipelined:
// while round t is resolved, the labels of round t+1 and the neighbour ids
// of round t+2 are loading, and the offsets of round t+3.
template <int G, class Resolve, bool VLIST, typename E>
__device__ __forceinline__ uint32_t group_rounds(const E *__restrict__ off, const int32_t *__restrict__ nbr,
                                                 const int32_t *__restrict__ prev, int32_t *__restrict__ next,
                                                 const int32_t *__restrict__ vlist, int64_t v0, int64_t nv) {
    const int p = lane_id() & (G - 1);
    constexpr int64_t per_block = 256 / G;
    const int64_t stride = (int64_t)gridDim.x * per_block;
    // Raw offsets stay in the pipeline registers until the next stage uses
    // them (any arithmetic on a just-loaded value would wait for it), and
    // every load is issued by every lane from a clamped address, so no load
    // sits under a branch (see pipelined()).
    struct St {
        int64_t v;
        E ob, oe;     // off[v], off[v+1], raw
        bool ok;      // i < nv
    };
    auto state = [&](int64_t i) {
        St s;
        s.ok = i < nv;
        const int64_t ic = s.ok ? i : 0;
        if constexpr (VLIST) s.v = (int64_t)vlist[v0 + ic];
        else s.v = v0 + ic;
        s.ob = off[s.v];
        s.oe = off[s.v + 1];
        return s;
    };
    auto has = [&](const St &s) { return s.ok && (int64_t)p < (int64_t)(s.oe - s.ob); };
    auto ldn = [&](const St &s) -> int32_t {
        return __builtin_nontemporal_load(&nbr[has(s) ? (int64_t)s.ob + p : 0]);
    };
    auto gat = [&](int32_t u, const St &s) -> int32_t { return prev[has(s) ? u : 0]; };
    uint32_t changed = 0;
    // the winner lane (best's low byte) holds the new label and writes it
    auto resolve = [&](const St &s, int32_t Lraw, int32_t old) {
        const int32_t L = has(s) ? Lraw : -1;
        const uint32_t best = Resolve::template best<G>(L);
        if (p == (int)(best & 0xffu) && s.ok && s.oe != s.ob) {
            next[s.v] = L;
            changed += (L != old);
        }
    };
    int64_t base = (int64_t)blockIdx.x * per_block;   // uniform round index
    int64_t i = base + threadIdx.x / G;               // this lane's vertex index
    // Four state slots rotate through the stages; the loop is unrolled by
    // the rotation period so no register holding an in-flight load is ever
    // copied (a copy would wait for it).
    St sa = state(i), sb = state(i + stride), sc = state(i + 2 * stride), sd;
    int32_t u = ldn(sa);
    int32_t LA = gat(u, sa), oA = prev[sa.v], LB, oB;
    u = ldn(sb);
#define GALP_GROUP_STAGE(S0, S1, S2, S3, LC, OC, LN, ON) \
    LN = gat(u, S1);                                     \
    ON = prev[S1.v];                                     \
    u = ldn(S2);                                         \
    S3 = state(i + 3 * stride);                          \
    resolve(S0, LC, OC);                                 \
    i += stride;                                         \
    base += stride;                                      \
    if (base >= nv) break;
    while (base < nv) {
        GALP_GROUP_STAGE(sa, sb, sc, sd, LA, oA, LB, oB)
        GALP_GROUP_STAGE(sb, sc, sd, sa, LB, oB, LA, oA)
        GALP_GROUP_STAGE(sc, sd, sa, sb, LA, oA, LB, oB)
        GALP_GROUP_STAGE(sd, sa, sb, sc, LB, oB, LA, oA)
    }
#undef GALP_GROUP_STAGE
    return changed;
}

// Alt: the resolver of iterations where most rows are high-distinct (the
// learned dispatcher's count of such big vertices, *hd_count, exceeds half of
// hd_total); chosen once per launch, so each choice is its own pipeline.
// nv_dev, when given, holds the list's length (written on the device by an
// earlier kernel; nv is then only its capacity).
template <int G, class Resolve, class Alt, bool VLIST, typename E>
__global__ void __launch_bounds__(256)
k_group(const E *__restrict__ off, const int32_t *__restrict__ nbr,
        const int32_t *__restrict__ prev, int32_t *__restrict__ next,
        const int32_t *__restrict__ vlist, int64_t v0, int64_t nv, const int64_t *__restrict__ nv_dev,
        unsigned long long *__restrict__ counter, const int32_t *__restrict__ stop,
        const int64_t *__restrict__ hd_count, int64_t hd_total) {
    __shared__ uint32_t scratch[256 / kWave];
    if (*stop) return;
    if (nv_dev) nv = *nv_dev;
    const bool alt = hd_count && 2 * (*hd_count) > hd_total;   // uniform
    const uint32_t changed = alt ? group_rounds<G, Alt, VLIST, E>(off, nbr, prev, next, vlist, v0, nv)
                                 : group_rounds<G, Resolve, VLIST, E>(off, nbr, prev, next, vlist, v0, nv);
    block_count_add<4>(changed, counter, scratch);
}

}  // namespace galp
