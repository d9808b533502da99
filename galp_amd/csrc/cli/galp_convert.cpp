// galp_convert.cpp -- writes the ".bin" graph format the galp program reads:
// int32 (u, v) pairs, rows in vertex order, BOTH directions of every edge
// (label_propagation/include/graph.h:119-150, README.md:84-88).  Replaces
// utils/edgelist2bin.cpp and utils/metis2bin.cpp.
//
//   galp_convert [--metis] [--keep-ids] [--file-order] [-o out.bin] input
//
// Edge lists (default): lines "u v" (tab or space separated; lines starting
// with '#' -- SNAP comments -- are skipped, edgelist2bin.cpp:24-26).  Row u
// receives v and row v receives u, in line order (edgelist2bin.cpp:32-33).
// By default the ids that occur are renumbered densely in ascending order
// (edgelist2bin.cpp:41-46) and every row is sorted ascending (:53), the
// normalisation the reference's converter applies; --keep-ids keeps the
// ids (n = max id + 1) and --file-order keeps each row in line order, which
// together give the plain variant of edgelist2bin.cpp:68-96.  Unlike the
// normalising variant, which keeps only x < y (edgelist2bin.cpp:58), both
// directions are written, as the binary format is specified to hold them.
//
// METIS (--metis): a header line, then line i lists the neighbours of vertex
// i (ids from 1), written in file order (metis2bin.cpp:13-101).  Vertices
// that have neighbours or are some vertex's neighbour are renumbered densely
// in ascending order, as metis2bin.cpp:19-44 does for a symmetric file; a
// vertex with neither is dropped.
//
// Output: "graph.bin" in the working directory unless -o is given, like the
// reference tools.  Prints "n m" of the written graph.  Exit status 1 on
// any error.

#include <algorithm>
#include <cerrno>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "galp_host.h"

namespace {

int usage() {
    std::fprintf(stderr, "usage: galp_convert [--metis] [--keep-ids] [--file-order] [-o out.bin] input\n");
    return 1;
}

// Parses the next integer from p (skipping blanks); false at end of line.
bool next_int(const char *&p, long long &x) {
    while (*p == ' ' || *p == '\t' || *p == '\r') ++p;
    if (!(*p == '-' || (*p >= '0' && *p <= '9'))) return false;
    char *end = nullptr;
    errno = 0;
    x = std::strtoll(p, &end, 10);
    if (end == p || errno) return false;
    p = end;
    return true;
}

struct Rows {
    std::vector<std::vector<int32_t>> adj;
    void add(int64_t u, int32_t v) {
        if ((int64_t)adj.size() <= u) adj.resize((size_t)u + 1);
        adj[(size_t)u].push_back(v);
    }
};

// Dense ranks of the ids with mark[id] set, ascending; -1 elsewhere.
std::vector<int32_t> dense_ids(const std::vector<uint8_t> &mark) {
    std::vector<int32_t> id(mark.size(), -1);
    int32_t k = 0;
    for (size_t i = 0; i < mark.size(); ++i)
        if (mark[i]) id[i] = k++;
    return id;
}

bool read_edgelist(FILE *f, bool keep_ids, Rows &rows, int64_t &n) {
    std::vector<std::pair<int32_t, int32_t>> edges;
    std::vector<uint8_t> seen;
    char *line = nullptr;
    size_t cap = 0;
    ssize_t len;
    int64_t maxid = -1;
    while ((len = getline(&line, &cap, f)) != -1) {
        if (len == 0 || line[0] == '#') continue;
        const char *p = line;
        long long u, v;
        if (!next_int(p, u) || !next_int(p, v)) continue;   // blank or malformed line
        if (u < 0 || v < 0 || u > INT32_MAX || v > INT32_MAX) {
            std::fprintf(stderr, "galp_convert: vertex id out of the int32 range\n");
            std::free(line);
            return false;
        }
        edges.emplace_back((int32_t)u, (int32_t)v);
        maxid = std::max<int64_t>(maxid, std::max(u, v));
    }
    std::free(line);
    n = maxid + 1;
    std::vector<int32_t> id;
    if (!keep_ids) {
        seen.assign((size_t)n, 0);
        for (auto &e : edges) seen[(size_t)e.first] = seen[(size_t)e.second] = 1;
        id = dense_ids(seen);
        n = 0;
        for (uint8_t s : seen) n += s;
    }
    rows.adj.assign((size_t)n, {});
    for (auto &e : edges) {
        const int32_t u = keep_ids ? e.first : id[(size_t)e.first];
        const int32_t v = keep_ids ? e.second : id[(size_t)e.second];
        rows.add(u, v);
        rows.add(v, u);
    }
    return true;
}

bool read_metis(FILE *f, Rows &rows, int64_t &n) {
    char *line = nullptr;
    size_t cap = 0;
    ssize_t len;
    if ((len = getline(&line, &cap, f)) == -1) {   // header
        std::free(line);
        std::fprintf(stderr, "galp_convert: empty METIS file\n");
        return false;
    }
    std::vector<std::vector<int32_t>> raw;   // raw[i]: neighbours of vertex i + 1, ids from 1
    while ((len = getline(&line, &cap, f)) != -1) {
        if (line[0] == '%') continue;   // METIS comment line
        raw.emplace_back();
        const char *p = line;
        long long j;
        while (next_int(p, j)) {
            if (j < 1 || j > INT32_MAX) {
                std::fprintf(stderr, "galp_convert: METIS neighbour id out of range\n");
                std::free(line);
                return false;
            }
            raw.back().push_back((int32_t)j);
        }
    }
    std::free(line);
    int64_t maxid = (int64_t)raw.size();
    for (auto &r : raw)
        for (int32_t j : r) maxid = std::max<int64_t>(maxid, j);
    std::vector<uint8_t> mark((size_t)maxid + 1, 0);
    for (size_t i = 0; i < raw.size(); ++i) {
        if (!raw[i].empty()) mark[i + 1] = 1;
        for (int32_t j : raw[i]) mark[(size_t)j] = 1;
    }
    const std::vector<int32_t> id = dense_ids(mark);
    n = 0;
    for (uint8_t s : mark) n += s;
    rows.adj.assign((size_t)n, {});
    for (size_t i = 0; i < raw.size(); ++i)
        for (int32_t j : raw[i]) rows.add(id[i + 1], id[(size_t)j]);
    return true;
}

}  // namespace

int main(int argc, char **argv) {
    bool metis = false, keep_ids = false, file_order = false;
    const char *out = "graph.bin";
    const char *in = nullptr;
    for (int i = 1; i < argc; ++i) {
        if (!std::strcmp(argv[i], "--metis")) metis = true;
        else if (!std::strcmp(argv[i], "--keep-ids")) keep_ids = true;
        else if (!std::strcmp(argv[i], "--file-order")) file_order = true;
        else if (!std::strcmp(argv[i], "-o") && i + 1 < argc) out = argv[++i];
        else if (argv[i][0] == '-' || in) return usage();
        else in = argv[i];
    }
    if (!in) return usage();
    FILE *f = std::fopen(in, "r");
    if (!f) {
        std::fprintf(stderr, "galp_convert: cannot open %s\n", in);
        return 1;
    }
    Rows rows;
    int64_t n = 0;
    const bool ok = metis ? read_metis(f, rows, n) : read_edgelist(f, keep_ids, rows, n);
    std::fclose(f);
    if (!ok) return 1;
    if (!metis && !file_order)
        for (auto &r : rows.adj) std::sort(r.begin(), r.end());
    n = std::max<int64_t>(n, (int64_t)rows.adj.size());
    rows.adj.resize((size_t)n);
    std::vector<int64_t> off((size_t)n + 1, 0);
    for (int64_t v = 0; v < n; ++v) off[(size_t)v + 1] = off[(size_t)v] + (int64_t)rows.adj[(size_t)v].size();
    std::vector<int32_t> nbr((size_t)off[(size_t)n]);
    for (int64_t v = 0; v < n; ++v)
        std::copy(rows.adj[(size_t)v].begin(), rows.adj[(size_t)v].end(), nbr.begin() + off[(size_t)v]);
    galp_host_csr csr{n, off[(size_t)n], off.data(), nbr.data()};
    if (galp_save_bin(out, &csr) != GALP_OK) {
        std::fprintf(stderr, "galp_convert: %s\n", galp_host_last_error());
        return 1;
    }
    std::printf("%lld %lld\n", (long long)n, (long long)csr.m);
    return 0;
}
