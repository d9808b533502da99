// graph_host.cpp -- CSR construction, file loaders and seeded generators for
// the galp LP path.  Host code only (no HIP calls); exported through
// include/galp_host.h.
//
// Loaders restate label_propagation/include/graph.h:58-188 (neighbour order
// = file order, text edges symmetrised per line); the generators produce the
// synthetic planted-partition, R-MAT and LFR-style inputs of the benchmarks.

#include "galp_host.h"
#include "galp_status.h"
#include "parallel.hpp"

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <new>
#include <numeric>
#include <sstream>
#include <string>
#include <vector>

namespace galp {

thread_local std::string g_host_err;

int host_fail(int code, const std::string &msg) {
    g_host_err = msg;
    return code;
}

// ---------------------------------------------------------------- RNG ----
// SplitMix64 (Steele, Lea, Flood: "Fast splittable pseudorandom number
// generators", OOPSLA 2014).  One independent stream per (seed, item).
static inline uint64_t mix64(uint64_t z) {
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
    return z ^ (z >> 31);
}
struct SplitMix64 {
    uint64_t s;
    SplitMix64(uint64_t seed, uint64_t item)
        : s(mix64(seed ^ 0x6a09e667f3bcc909ULL) ^ mix64(item + 0x9e3779b97f4a7c15ULL)) {}
    uint64_t next() {
        s += 0x9e3779b97f4a7c15ULL;
        return mix64(s);
    }
    // uniform integer in [0, k) by the multiply-shift map (k >= 1)
    uint64_t below(uint64_t k) {
        return (uint64_t)(((unsigned __int128)next() * (unsigned __int128)k) >> 64);
    }
    double uniform() { return (double)(next() >> 11) * (1.0 / 9007199254740992.0); }
};

// ------------------------------------------------------------- CSR -------
galp_host_csr *csr_alloc(int64_t n, int64_t m) {
    galp_host_csr *g = (galp_host_csr *)std::calloc(1, sizeof(galp_host_csr));
    if (!g) return nullptr;
    g->n = n;
    g->m = m;
    g->offsets = (int64_t *)std::malloc(sizeof(int64_t) * (size_t)(n + 1));
    g->neighbors = (int32_t *)std::malloc(sizeof(int32_t) * (size_t)(m > 0 ? m : 1));
    if (!g->offsets || !g->neighbors) {
        std::free(g->offsets);
        std::free(g->neighbors);
        std::free(g);
        return nullptr;
    }
    return g;
}

// Directed pairs (both directions already present) -> CSR with rows sorted
// ascending and duplicates removed.  Self-loops are dropped.
static int csr_from_pairs(int64_t n, const std::vector<uint32_t> &src,
                          const std::vector<uint32_t> &dst, galp_host_csr **out) {
    const int64_t np = (int64_t)src.size();
    std::vector<int64_t> cnt(n + 1, 0);
    {
        std::atomic<int64_t> *acnt = reinterpret_cast<std::atomic<int64_t> *>(cnt.data());
        parallel_for(0, np, [&](int64_t i) {
            if (src[i] != dst[i]) acnt[src[i]].fetch_add(1, std::memory_order_relaxed);
        });
    }
    std::vector<int64_t> off(n + 1, 0);
    for (int64_t v = 0; v < n; ++v) off[v + 1] = off[v] + cnt[v];
    std::vector<int32_t> tmp((size_t)std::max<int64_t>(off[n], 1));
    {
        std::vector<int64_t> cur(off.begin(), off.end() - 1);
        std::atomic<int64_t> *acur = reinterpret_cast<std::atomic<int64_t> *>(cur.data());
        parallel_for(0, np, [&](int64_t i) {
            if (src[i] == dst[i]) return;
            int64_t p = acur[src[i]].fetch_add(1, std::memory_order_relaxed);
            tmp[p] = (int32_t)dst[i];
        });
    }
    // sort + unique each row; record new degrees
    parallel_for_dynamic(0, n, 1024, [&](int64_t v) {
        int32_t *b = tmp.data() + off[v];
        int32_t *e = tmp.data() + off[v + 1];
        std::sort(b, e);
        cnt[v] = std::unique(b, e) - b;
    });
    int64_t m = 0;
    for (int64_t v = 0; v < n; ++v) m += cnt[v];
    galp_host_csr *g = csr_alloc(n, m);
    if (!g) return host_fail(GALP_ENOMEM, "host allocation failed for CSR");
    g->offsets[0] = 0;
    for (int64_t v = 0; v < n; ++v) g->offsets[v + 1] = g->offsets[v] + cnt[v];
    parallel_for_dynamic(0, n, 4096, [&](int64_t v) {
        std::memcpy(g->neighbors + g->offsets[v], tmp.data() + off[v], sizeof(int32_t) * (size_t)cnt[v]);
    });
    *out = g;
    return GALP_OK;
}

// ----------------------------------------------------------- loaders -----
static int load_txt(const std::string &path, galp_host_csr **out) {
    std::ifstream ifs(path);
    if (!ifs) return host_fail(GALP_EIO, "cannot open " + path);
    // graph.h:166-188: per line "u v": adj[u] += v, adj[v] += u.
    std::vector<std::vector<int32_t>> adj;
    std::string s;
    while (std::getline(ifs, s)) {
        const char *p = s.c_str();
        char *q;
        long long u = std::strtoll(p, &q, 10);
        if (q == p) continue;
        p = q;
        long long v = std::strtoll(p, &q, 10);
        if (q == p) continue;
        if (u < 0 || v < 0 || u > INT32_MAX - 1 || v > INT32_MAX - 1)
            return host_fail(GALP_EIO, "vertex id out of int32 range in " + path);
        size_t need = (size_t)std::max(u, v) + 1;
        if (adj.size() < need) adj.resize(need);
        adj[(size_t)u].push_back((int32_t)v);
        adj[(size_t)v].push_back((int32_t)u);
    }
    int64_t n = (int64_t)adj.size(), m = 0;
    for (auto &r : adj) m += (int64_t)r.size();
    galp_host_csr *g = csr_alloc(n, m);
    if (!g) return host_fail(GALP_ENOMEM, "host allocation failed for CSR");
    int64_t cur = 0;
    g->offsets[0] = 0;
    for (int64_t u = 0; u < n; ++u) {
        for (int32_t v : adj[(size_t)u]) g->neighbors[cur++] = v;
        g->offsets[u + 1] = cur;
        std::vector<int32_t>().swap(adj[(size_t)u]);
    }
    *out = g;
    return GALP_OK;
}

static int load_bin(const std::string &path, galp_host_csr **out) {
    FILE *f = std::fopen(path.c_str(), "rb");
    if (!f) return host_fail(GALP_EIO, "cannot open " + path);
    std::fseek(f, 0, SEEK_END);
    long long bytes = std::ftell(f);
    std::fseek(f, 0, SEEK_SET);
    if (bytes < 0 || bytes % 8 != 0) {
        std::fclose(f);
        return host_fail(GALP_EIO, path + ": size is not a multiple of 8 bytes (int32 pairs)");
    }
    int64_t np = bytes / 8;
    std::vector<int32_t> buf((size_t)np * 2);
    if (np && std::fread(buf.data(), 8, (size_t)np, f) != (size_t)np) {
        std::fclose(f);
        return host_fail(GALP_EIO, "short read on " + path);
    }
    std::fclose(f);
    int64_t n = 0;
    for (int64_t i = 0; i < 2 * np; ++i) {
        if (buf[i] < 0) return host_fail(GALP_EIO, path + ": negative vertex id");
        n = std::max<int64_t>(n, (int64_t)buf[i] + 1);
    }
    // Rows grouped by source; within a source, file order (graph.h:129-147).
    galp_host_csr *g = csr_alloc(n, np);
    if (!g) return host_fail(GALP_ENOMEM, "host allocation failed for CSR");
    std::vector<int64_t> cur(n + 1, 0);
    for (int64_t i = 0; i < np; ++i) cur[buf[2 * i] + 1]++;
    for (int64_t v = 0; v < n; ++v) cur[v + 1] += cur[v];
    std::memcpy(g->offsets, cur.data(), sizeof(int64_t) * (size_t)(n + 1));
    for (int64_t i = 0; i < np; ++i) g->neighbors[cur[buf[2 * i]]++] = buf[2 * i + 1];
    *out = g;
    return GALP_OK;
}

// -------------------------------------------------------- generators ----
// The draws of vertex v of a planted-partition graph: kin partners from its
// own community (without v), kout from the other communities; emit(v, w)
// and emit(w, v) for each.
template <class Emit>
static inline void pp_vertex(int64_t v, int64_t n, int64_t cs, int64_t ncomm, int32_t kin, int32_t kout,
                             uint64_t seed, Emit &&emit) {
    SplitMix64 r(seed, (uint64_t)v);
    const int64_t c = v / cs, base = c * cs;
    const int64_t size = std::min(cs, n - base);
    for (int32_t i = 0; i < kin; ++i) {
        int64_t w = v;
        if (size > 1) {
            int64_t x = (int64_t)r.below((uint64_t)(size - 1));
            if (x >= v - base) ++x;
            w = base + x;
        }
        emit(v, w);
    }
    for (int32_t i = 0; i < kout; ++i) {
        int64_t w = v;
        if (ncomm > 1) {
            int64_t c2 = (int64_t)r.below((uint64_t)(ncomm - 1));
            if (c2 >= c) ++c2;
            int64_t b2 = c2 * cs, s2 = std::min(cs, n - b2);
            w = b2 + (int64_t)r.below((uint64_t)s2);
        }
        emit(v, w);
    }
}

static bool pp_params_ok(int64_t n, int64_t cs, int32_t kin, int32_t kout) {
    return n > 0 && n <= INT32_MAX && cs > 0 && kin >= 0 && kout >= 0;
}

static int gen_pp(int64_t n, int64_t cs, int32_t kin, int32_t kout, uint64_t seed,
                  galp_host_csr **out) {
    if (!pp_params_ok(n, cs, kin, kout)) return host_fail(GALP_EINVAL, "planted partition: bad parameters");
    const int64_t ncomm = (n + cs - 1) / cs;
    const int64_t per = (int64_t)kin + kout;
    std::vector<uint32_t> src((size_t)(2 * n * per)), dst((size_t)(2 * n * per));
    parallel_for(0, n, [&](int64_t v) {
        size_t k = (size_t)(2 * v * per);
        pp_vertex(v, n, cs, ncomm, kin, kout, seed, [&](int64_t a, int64_t b) {
            src[k] = (uint32_t)a; dst[k] = (uint32_t)b; ++k;
            src[k] = (uint32_t)b; dst[k] = (uint32_t)a; ++k;
        });
    });
    return csr_from_pairs(n, src, dst, out);
}

// One planted-partition graph built on the calling thread (the batch
// generator runs one graph per worker): the same graph as gen_pp, since rows
// are sorted and deduplicated whatever the order the pairs come in.
static void pp_serial(int64_t n, int64_t cs, int32_t kin, int32_t kout, uint64_t seed, std::vector<int64_t> &off,
                      std::vector<int32_t> &nbr) {
    const int64_t ncomm = (n + cs - 1) / cs;
    const int64_t per = (int64_t)kin + kout;
    std::vector<int32_t> src((size_t)(2 * n * per)), dst((size_t)(2 * n * per));
    size_t k = 0;
    for (int64_t v = 0; v < n; ++v)
        pp_vertex(v, n, cs, ncomm, kin, kout, seed, [&](int64_t a, int64_t b) {
            src[k] = (int32_t)a; dst[k] = (int32_t)b; ++k;
            src[k] = (int32_t)b; dst[k] = (int32_t)a; ++k;
        });
    std::vector<int64_t> cur((size_t)n + 1, 0);
    for (size_t i = 0; i < k; ++i)
        if (src[i] != dst[i]) cur[(size_t)src[i] + 1]++;
    for (int64_t v = 0; v < n; ++v) cur[(size_t)v + 1] += cur[(size_t)v];
    std::vector<int32_t> tmp((size_t)std::max<int64_t>(cur[(size_t)n], 1));
    std::vector<int64_t> pos(cur.begin(), cur.end() - 1);
    for (size_t i = 0; i < k; ++i)
        if (src[i] != dst[i]) tmp[(size_t)pos[(size_t)src[i]]++] = dst[i];
    std::vector<int32_t>().swap(src);
    std::vector<int32_t>().swap(dst);
    off.assign((size_t)n + 1, 0);
    nbr.clear();
    nbr.reserve((size_t)cur[(size_t)n]);
    for (int64_t v = 0; v < n; ++v) {
        int32_t *b = tmp.data() + cur[(size_t)v], *e = tmp.data() + cur[(size_t)v + 1];
        std::sort(b, e);
        e = std::unique(b, e);
        nbr.insert(nbr.end(), b, e);
        off[(size_t)v + 1] = (int64_t)nbr.size();
    }
    nbr.shrink_to_fit();
}

// Graph k of a batch: its size, community size, draws and seed, all from the
// stream (seed, k).
static void batch_spec(uint64_t seed, int64_t k, int64_t n_lo, int64_t n_hi, int64_t c_lo, int64_t c_hi, int32_t d_lo,
                       int32_t d_hi, double inter_frac, int64_t spec[5]) {
    SplitMix64 r(seed, (uint64_t)k);
    const int64_t n = n_lo + (int64_t)r.below((uint64_t)(n_hi - n_lo + 1));
    const int64_t cs = c_lo + (int64_t)r.below((uint64_t)(c_hi - c_lo + 1));
    const int64_t d = d_lo + (int64_t)r.below((uint64_t)(d_hi - d_lo + 1));
    spec[0] = n;
    spec[1] = cs;
    spec[2] = (d + 1) / 2;                                                        // draws: degree ~ d
    spec[3] = std::max<int64_t>(1, std::llround(inter_frac * (double)d / 2.0));   // degree ~ inter_frac * d
    spec[4] = (int64_t)(mix64(seed ^ 0x5851f42d4c957f2dULL) + (uint64_t)k);
}

static int gen_batch_pp(int64_t count, int64_t n_lo, int64_t n_hi, int64_t c_lo, int64_t c_hi, int32_t d_lo,
                        int32_t d_hi, double inter_frac, uint64_t seed, galp_host_csr **out, int64_t *vbase,
                        int64_t *spec_out) {
    if (count < 1 || n_lo < 1 || n_hi < n_lo || c_lo < 1 || c_hi < c_lo || d_lo < 0 || d_hi < d_lo ||
        !(inter_frac >= 0.0))
        return host_fail(GALP_EINVAL, "batch planted partition: bad parameters");
    std::vector<int64_t> spec((size_t)(5 * count));
    vbase[0] = 0;
    for (int64_t k = 0; k < count; ++k) {
        batch_spec(seed, k, n_lo, n_hi, c_lo, c_hi, d_lo, d_hi, inter_frac, &spec[(size_t)(5 * k)]);
        vbase[k + 1] = vbase[k] + spec[(size_t)(5 * k)];
    }
    if (vbase[count] > INT32_MAX) return host_fail(GALP_EINVAL, "batch: more than 2^31 - 1 vertices in all");
    std::vector<std::vector<int64_t>> offs((size_t)count);
    std::vector<std::vector<int32_t>> nbrs((size_t)count);
    parallel_for_dynamic(0, count, 1, [&](int64_t k) {
        const int64_t *sp = &spec[(size_t)(5 * k)];
        pp_serial(sp[0], sp[1], (int32_t)sp[2], (int32_t)sp[3], (uint64_t)sp[4], offs[(size_t)k], nbrs[(size_t)k]);
    });
    std::vector<int64_t> ebase((size_t)count + 1, 0);
    for (int64_t k = 0; k < count; ++k) ebase[(size_t)k + 1] = ebase[(size_t)k] + (int64_t)nbrs[(size_t)k].size();
    galp_host_csr *g = csr_alloc(vbase[count], ebase[(size_t)count]);
    if (!g) return host_fail(GALP_ENOMEM, "host allocation failed for the batch CSR");
    g->offsets[0] = 0;
    parallel_for_dynamic(0, count, 1, [&](int64_t k) {
        const std::vector<int64_t> &o = offs[(size_t)k];
        const int64_t vb = vbase[k], eb = ebase[(size_t)k];
        for (size_t v = 1; v < o.size(); ++v) g->offsets[vb + (int64_t)v] = eb + o[v];
        int32_t *dst = g->neighbors + eb;
        for (int32_t u : nbrs[(size_t)k]) *dst++ = u + (int32_t)vb;
        std::vector<int64_t>().swap(offs[(size_t)k]);
        std::vector<int32_t>().swap(nbrs[(size_t)k]);
    });
    if (spec_out) std::memcpy(spec_out, spec.data(), sizeof(int64_t) * spec.size());
    *out = g;
    return GALP_OK;
}

static std::vector<uint32_t> fisher_yates(int64_t n, uint64_t seed) {
    std::vector<uint32_t> p((size_t)n);
    std::iota(p.begin(), p.end(), 0u);
    SplitMix64 r(seed, 0xfeedULL);
    for (int64_t i = n - 1; i > 0; --i) {
        int64_t j = (int64_t)r.below((uint64_t)(i + 1));
        std::swap(p[(size_t)i], p[(size_t)j]);
    }
    return p;
}

static int gen_rmat(int32_t scale, int32_t ef, double a, double b, double c, uint64_t seed,
                    galp_host_csr **out) {
    if (scale < 1 || scale > 30 || ef < 1 || a < 0 || b < 0 || c < 0 || a + b + c > 1.0)
        return host_fail(GALP_EINVAL, "rmat: bad parameters");
    const int64_t n = (int64_t)1 << scale;
    const int64_t ns = (int64_t)ef * n;
    const uint32_t ta = (uint32_t)std::min(4294967295.0, a * 4294967296.0);
    const uint32_t tab = (uint32_t)std::min(4294967295.0, (a + b) * 4294967296.0);
    const uint32_t tabc = (uint32_t)std::min(4294967295.0, (a + b + c) * 4294967296.0);
    std::vector<uint32_t> perm = fisher_yates(n, seed);
    std::vector<uint32_t> src((size_t)(2 * ns)), dst((size_t)(2 * ns));
    parallel_for(0, ns, [&](int64_t i) {
        SplitMix64 r(seed, (uint64_t)i);
        uint32_t u = 0, v = 0;
        uint64_t bits = 0;
        for (int32_t l = 0; l < scale; ++l) {
            if ((l & 1) == 0) bits = r.next();
            uint32_t x = (uint32_t)(bits >> ((l & 1) ? 32 : 0));
            uint32_t bu = x >= tab, bv = (x >= ta && x < tab) || x >= tabc;
            u = (u << 1) | bu;
            v = (v << 1) | bv;
        }
        uint32_t pu = perm[u], pv = perm[v];
        src[(size_t)(2 * i)] = pu; dst[(size_t)(2 * i)] = pv;
        src[(size_t)(2 * i + 1)] = pv; dst[(size_t)(2 * i + 1)] = pu;
    });
    return csr_from_pairs(n, src, dst, out);
}

// Continuous power law on [lo, hi] with exponent g (density ~ x^-g).
static double pl_sample(double u, double lo, double hi, double g) {
    double e = 1.0 - g;
    double a = std::pow(lo, e), bb = std::pow(hi, e);
    return std::pow(a + u * (bb - a), 1.0 / e);
}
static double pl_mean(double lo, double hi, double g) {
    double e1 = 1.0 - g, e2 = 2.0 - g;
    return (e1 / e2) * (std::pow(hi, e2) - std::pow(lo, e2)) / (std::pow(hi, e1) - std::pow(lo, e1));
}

static int gen_lfr(int64_t n, double gamma, double avg, int32_t kmax, double beta, int32_t cmin,
                   int32_t cmax, double mu, uint64_t seed, galp_host_csr **out,
                   int32_t **comm_out) {
    if (n <= 1 || n > INT32_MAX || gamma <= 1.0 || gamma == 2.0 || avg < 1.0 || kmax < 2 ||
        beta <= 1.0 || beta == 2.0 || cmin < 2 || cmax < cmin || mu < 0 || mu > 1 || avg >= kmax)
        return host_fail(GALP_EINVAL, "lfr: bad parameters");
    // 1. degrees: continuous power law, lower cut solved so the mean is avg
    double lo = 1.0, hi = avg;
    for (int it = 0; it < 200; ++it) {
        double mid = 0.5 * (lo + hi);
        if (pl_mean(mid, (double)kmax, gamma) < avg) lo = mid; else hi = mid;
    }
    const double kmin = 0.5 * (lo + hi);
    std::vector<int32_t> deg((size_t)n);
    parallel_for(0, n, [&](int64_t v) {
        SplitMix64 r(seed, (uint64_t)v);
        double x = pl_sample(r.uniform(), kmin, (double)kmax, gamma);
        deg[(size_t)v] = (int32_t)std::max<long>(1, std::min<long>(kmax, std::lround(x)));
    });
    // 2. community sizes until they cover n
    std::vector<int64_t> csize;
    {
        SplitMix64 r(seed, 0xc0ffeeULL);
        int64_t tot = 0;
        while (tot < n) {
            int64_t s = std::lround(pl_sample(r.uniform(), (double)cmin, (double)cmax, beta));
            s = std::max<int64_t>(cmin, std::min<int64_t>(cmax, s));
            if (tot + s > n) {
                int64_t rest = n - tot;
                if (rest < cmin && !csize.empty()) csize.back() += rest; else csize.push_back(rest);
                tot = n;
                break;
            }
            csize.push_back(s);
            tot += s;
        }
    }
    const int64_t nc = (int64_t)csize.size();
    // 3. membership: a seeded shuffle of the community slots
    std::vector<int32_t> comm((size_t)n);
    {
        int64_t k = 0;
        for (int64_t c = 0; c < nc; ++c)
            for (int64_t i = 0; i < csize[(size_t)c]; ++i) comm[(size_t)k++] = (int32_t)c;
        SplitMix64 r(seed, 0xabcdefULL);
        for (int64_t i = n - 1; i > 0; --i) {
            int64_t j = (int64_t)r.below((uint64_t)(i + 1));
            std::swap(comm[(size_t)i], comm[(size_t)j]);
        }
    }
    // intra degree round((1-mu) k), capped at community size - 1
    std::vector<int32_t> kin((size_t)n);
    for (int64_t v = 0; v < n; ++v) {
        int64_t k = std::lround((1.0 - mu) * deg[(size_t)v]);
        kin[(size_t)v] = (int32_t)std::min<int64_t>(k, csize[(size_t)comm[(size_t)v]] - 1);
    }
    // 4. configuration model: intra stubs per community, inter stubs global
    std::vector<int64_t> coff((size_t)nc + 1, 0);
    for (int64_t v = 0; v < n; ++v) coff[(size_t)comm[(size_t)v] + 1] += kin[(size_t)v];
    for (int64_t c = 0; c < nc; ++c) coff[(size_t)c + 1] += coff[(size_t)c];
    std::vector<uint32_t> istub((size_t)coff[(size_t)nc]);
    {
        std::vector<int64_t> cur(coff.begin(), coff.end() - 1);
        for (int64_t v = 0; v < n; ++v)
            for (int32_t i = 0; i < kin[(size_t)v]; ++i)
                istub[(size_t)cur[(size_t)comm[(size_t)v]]++] = (uint32_t)v;
    }
    parallel_for_dynamic(0, nc, 1, [&](int64_t c) {
        SplitMix64 r(seed ^ 0x5151ULL, (uint64_t)c);
        uint32_t *s = istub.data() + coff[(size_t)c];
        int64_t len = coff[(size_t)c + 1] - coff[(size_t)c];
        for (int64_t i = len - 1; i > 0; --i) std::swap(s[i], s[(int64_t)r.below((uint64_t)(i + 1))]);
    });
    std::vector<uint32_t> ostub;
    for (int64_t v = 0; v < n; ++v)
        for (int32_t i = kin[(size_t)v]; i < deg[(size_t)v]; ++i) ostub.push_back((uint32_t)v);
    {
        SplitMix64 r(seed ^ 0x7777ULL, 1);
        for (int64_t i = (int64_t)ostub.size() - 1; i > 0; --i)
            std::swap(ostub[(size_t)i], ostub[(size_t)r.below((uint64_t)(i + 1))]);
    }
    std::vector<uint32_t> src, dst;
    src.reserve(istub.size() + ostub.size());
    dst.reserve(istub.size() + ostub.size());
    for (int64_t c = 0; c < nc; ++c)
        for (int64_t i = coff[(size_t)c]; i + 1 < coff[(size_t)c + 1]; i += 2) {
            uint32_t u = istub[(size_t)i], v = istub[(size_t)i + 1];
            src.push_back(u); dst.push_back(v);
            src.push_back(v); dst.push_back(u);
        }
    for (size_t i = 0; i + 1 < ostub.size(); i += 2) {
        uint32_t u = ostub[i], v = ostub[i + 1];
        if (comm[u] == comm[v]) continue;   // inter stubs must leave the community
        src.push_back(u); dst.push_back(v);
        src.push_back(v); dst.push_back(u);
    }
    int rc = csr_from_pairs(n, src, dst, out);
    if (rc == GALP_OK && comm_out) {
        *comm_out = (int32_t *)std::malloc(sizeof(int32_t) * (size_t)n);
        if (!*comm_out) {
            galp_host_csr_free(*out);
            *out = nullptr;
            return host_fail(GALP_ENOMEM, "host allocation failed");
        }
        std::memcpy(*comm_out, comm.data(), sizeof(int32_t) * (size_t)n);
    }
    return rc;
}

}  // namespace galp

using namespace galp;

extern "C" {

GALP_API const char *galp_host_last_error(void) { return g_host_err.c_str(); }

GALP_API void galp_host_csr_free(galp_host_csr *g) {
    if (!g) return;
    std::free(g->offsets);
    std::free(g->neighbors);
    std::free(g);
}

GALP_API void galp_host_free(void *p) { std::free(p); }

GALP_API int galp_host_csr_from_arrays(int64_t n, int64_t m, const int64_t *offsets,
                                       const int32_t *neighbors, galp_host_csr **out) {
    if (!out || n < 0 || m < 0 || !offsets || (m > 0 && !neighbors))
        return host_fail(GALP_EINVAL, "galp_host_csr_from_arrays: bad arguments");
    if (offsets[0] != 0 || offsets[n] != m)
        return host_fail(GALP_EINVAL, "galp_host_csr_from_arrays: offsets[0] must be 0 and offsets[n] == m");
    galp_host_csr *g = csr_alloc(n, m);
    if (!g) return host_fail(GALP_ENOMEM, "host allocation failed");
    std::memcpy(g->offsets, offsets, sizeof(int64_t) * (size_t)(n + 1));
    if (m) std::memcpy(g->neighbors, neighbors, sizeof(int32_t) * (size_t)m);
    *out = g;
    return GALP_OK;
}

GALP_API int galp_load_graph(const char *path, galp_host_csr **out) {
    if (!path || !out) return host_fail(GALP_EINVAL, "galp_load_graph: null argument");
    *out = nullptr;
    std::string p(path);
    auto dot = p.find_last_of('.');
    std::string ext = dot == std::string::npos ? "" : p.substr(dot);
    try {
        if (ext == ".txt" || ext == ".dat") return load_txt(p, out);
        return load_bin(p, out);
    } catch (const std::bad_alloc &) {
        return host_fail(GALP_ENOMEM, "host allocation failed while loading " + p);
    }
}

GALP_API int galp_save_bin(const char *path, const galp_host_csr *g) {
    if (!path || !g) return host_fail(GALP_EINVAL, "galp_save_bin: null argument");
    FILE *f = std::fopen(path, "wb");
    if (!f) return host_fail(GALP_EIO, std::string("cannot create ") + path);
    std::vector<int32_t> buf;
    for (int64_t v = 0; v < g->n; ++v) {
        buf.clear();
        for (int64_t e = g->offsets[v]; e < g->offsets[v + 1]; ++e) {
            buf.push_back((int32_t)v);
            buf.push_back(g->neighbors[e]);
        }
        if (!buf.empty() && std::fwrite(buf.data(), sizeof(int32_t), buf.size(), f) != buf.size()) {
            std::fclose(f);
            return host_fail(GALP_EIO, std::string("short write on ") + path);
        }
    }
    std::fclose(f);
    return GALP_OK;
}

GALP_API int galp_gen_planted_partition(int64_t n, int64_t community_size, int32_t k_intra,
                                        int32_t k_inter, uint64_t seed, galp_host_csr **out) {
    if (!out) return host_fail(GALP_EINVAL, "null out");
    *out = nullptr;
    try {
        return gen_pp(n, community_size, k_intra, k_inter, seed, out);
    } catch (const std::bad_alloc &) {
        return host_fail(GALP_ENOMEM, "host allocation failed (planted partition)");
    }
}

GALP_API int galp_gen_batch_planted(int64_t count, int64_t n_lo, int64_t n_hi, int64_t comm_lo, int64_t comm_hi,
                                    int32_t deg_lo, int32_t deg_hi, double inter_frac, uint64_t seed,
                                    galp_host_csr **out, int64_t *vbase, int64_t *spec) {
    if (!out || !vbase) return host_fail(GALP_EINVAL, "null out or vbase");
    *out = nullptr;
    try {
        return gen_batch_pp(count, n_lo, n_hi, comm_lo, comm_hi, deg_lo, deg_hi, inter_frac, seed, out, vbase, spec);
    } catch (const std::bad_alloc &) {
        return host_fail(GALP_ENOMEM, "host allocation failed (batch planted partition)");
    }
}

GALP_API int galp_gen_rmat(int32_t scale, int32_t edge_factor, double a, double b, double c,
                           uint64_t seed, galp_host_csr **out) {
    if (!out) return host_fail(GALP_EINVAL, "null out");
    *out = nullptr;
    try {
        return gen_rmat(scale, edge_factor, a, b, c, seed, out);
    } catch (const std::bad_alloc &) {
        return host_fail(GALP_ENOMEM, "host allocation failed (rmat)");
    }
}

GALP_API int galp_gen_lfr(int64_t n, double gamma, double avg_deg, int32_t max_deg, double beta,
                          int32_t cmin, int32_t cmax, double mu, uint64_t seed,
                          galp_host_csr **out, int32_t **community_out) {
    if (!out) return host_fail(GALP_EINVAL, "null out");
    *out = nullptr;
    if (community_out) *community_out = nullptr;
    try {
        return gen_lfr(n, gamma, avg_deg, max_deg, beta, cmin, cmax, mu, seed, out, community_out);
    } catch (const std::bad_alloc &) {
        return host_fail(GALP_ENOMEM, "host allocation failed (lfr)");
    }
}

}  // extern "C"
