// accuracy.cpp -- clustering quality of a label vector against a ground-truth
// file: NMI, F-measure and adjusted Rand index, with the formulas of
// label_propagation/include/accuracy.h (compute_nmi 49-158,
// compute_f_measure 161-236, compute_ari 239-319).
//
// Ground truth: one cluster per line, whitespace-separated vertex ids
// (accuracy.h:53-66).  Found clusters: vertices grouped by label.  The
// reference intersects every pair of clusters as std::sets
// (O(ca * cb * size)); here one pass over the ground-truth memberships builds
// the sparse contingency table N_ij, which gives the same sums.

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <sstream>
#include <string>
#include <unordered_map>
#include <vector>

#include "galp_host.h"

namespace galp {
int host_fail(int code, const std::string &msg);
}

extern "C" GALP_API int galp_accuracy(const char *truth_file, const int32_t *labels, int64_t n,
                                      double *nmi, double *fmeasure, double *ari) {
    using galp::host_fail;
    if (!truth_file || (!labels && n > 0) || n < 0) return host_fail(GALP_EINVAL, "galp_accuracy: bad arguments");
    std::ifstream ifs(truth_file);
    if (!ifs) return host_fail(GALP_EIO, std::string("cannot open ") + truth_file);

    // found clusters: labels mapped to 0..cb-1 in ascending label order
    std::vector<int32_t> sorted(labels, labels + n);
    std::sort(sorted.begin(), sorted.end());
    sorted.erase(std::unique(sorted.begin(), sorted.end()), sorted.end());
    const int64_t cb = (int64_t)sorted.size();
    std::vector<int64_t> cid((size_t)n);
    std::vector<double> csize((size_t)cb, 0.0);
    for (int64_t v = 0; v < n; ++v) {
        cid[(size_t)v] = std::lower_bound(sorted.begin(), sorted.end(), labels[v]) - sorted.begin();
        csize[(size_t)cid[(size_t)v]] += 1.0;
    }

    // sparse contingency: per ground-truth cluster i, counts over found j
    std::vector<std::unordered_map<int64_t, double>> N;
    std::vector<double> gsize;   // |gt_i| as a set (ids beyond n included, accuracy.h:214)
    std::string s;
    std::vector<long long> ids;
    while (std::getline(ifs, s)) {
        ids.clear();
        std::istringstream ss(s);
        long long v;
        while (ss >> v) ids.push_back(v);
        std::sort(ids.begin(), ids.end());
        ids.erase(std::unique(ids.begin(), ids.end()), ids.end());
        N.emplace_back();
        gsize.push_back((double)ids.size());
        for (long long x : ids)
            if (x >= 0 && x < n) N.back()[cid[(size_t)x]] += 1.0;
    }
    const int64_t ca = (int64_t)N.size();

    std::vector<double> Ni((size_t)ca, 0.0), Nj((size_t)cb, 0.0);
    double Nt = 0.0;
    for (int64_t i = 0; i < ca; ++i)
        for (auto &kv : N[(size_t)i]) {
            Ni[(size_t)i] += kv.second;
            Nj[(size_t)kv.first] += kv.second;
            Nt += kv.second;
        }

    // NMI (accuracy.h:92-157)
    double v1 = 0.0;
    for (int64_t i = 0; i < ca; ++i)
        for (auto &kv : N[(size_t)i]) {
            double nij = kv.second;
            v1 += nij * std::log2((nij * Nt) / (Ni[(size_t)i] * Nj[(size_t)kv.first]));
        }
    v1 *= -2.0;
    double v2 = 0.0;
    for (double x : Ni)
        if (x > 0.0) v2 += x * std::log2(x / Nt);
    for (double x : Nj)
        if (x > 0.0) v2 += x * std::log2(x / Nt);
    std::printf("%lld, %lld\n", (long long)ca, (long long)cb);   // accuracy.h:156
    if (nmi) *nmi = v1 / v2;

    // F-measure (accuracy.h:204-235)
    double f = 0.0;
    for (int64_t i = 0; i < ca; ++i) {
        double mx = 0.0;
        for (auto &kv : N[(size_t)i])
            mx = std::max(mx, 2.0 * kv.second / (gsize[(size_t)i] + csize[(size_t)kv.first]));
        f += mx * gsize[(size_t)i];
    }
    if (fmeasure) *fmeasure = n ? f / (double)n : 0.0;

    // ARI (accuracy.h:282-318)
    double a1 = 0.0, a2 = 0.0, a3 = 0.0;
    for (int64_t i = 0; i < ca; ++i)
        for (auto &kv : N[(size_t)i]) a1 += kv.second * (kv.second - 1.0) / 2.0;
    for (double x : Ni) a2 += x * (x - 1.0) / 2.0;
    for (double x : Nj) a3 += x * (x - 1.0) / 2.0;
    const double nn = (double)n;
    const double e = 2.0 * a2 * a3 / (nn * (nn - 1.0));
    if (ari) *ari = (a1 - e) / ((a2 + a3) / 2.0 - e);
    return GALP_OK;
}
