// internal.hpp -- host-side state behind the galp C-ABI (not installed).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string>
#include <vector>

#include "galp.h"

namespace galp {

// Degree classes of the plain dispatcher.  Each class is one contiguous id
// range after the degree-descending renumbering (or one vertex list without
// it) and one kernel configuration.
enum BucketId {
    B_G4 = 0,   // 1..4       lanes-per-vertex groups of 4
    B_G8,       // 5..8
    B_G16,      // 9..16
    B_G32,      // 17..32
    B_G64,      // 33..64     one full wave per vertex
    B_W128,     // 65..128    one wave per vertex, 256-slot LDS table
    B_W256,     // 129..256   one wave, 512 slots
    B_W512,     // 257..512   one wave, 1024 slots
    B_B1K,      // 513..1024  128 threads, 2048 slots
    B_B2K,      // 1025..2048 256 threads, 4096 slots
    B_B4K,      // 2049..4096 512 threads, 8192 slots
    B_HUB,      // > 4096     label-partitioned blocks, ceil(deg/2048) per hub
    NUM_BUCKETS
};
// Upper degree bound of each class (host and device).
__host__ __device__ constexpr int64_t bucket_hi(int b) {
    return b < 5 ? ((int64_t)4 << b) : b < B_HUB ? ((int64_t)128 << (b - 5)) : INT64_MAX;
}
__host__ __device__ inline int bucket_of_deg(int64_t d) {
    int b = 0;
    while (b < B_HUB && d > bucket_hi(b)) ++b;
    return b;
}
extern const int64_t kBucketHi[NUM_BUCKETS];
extern const char *const kBucketName[NUM_BUCKETS];

constexpr int kHubTileBits = 11;
constexpr int kHubTile = 1 << kHubTileBits;   // edges per hub tile (phase 1)
// A hub tile's entry for one label: its count in the tile (<= kHubTile, 12
// bits) and its last position within the tile, in one word; the merge adds
// the tile's first row position back.
__host__ __device__ constexpr uint32_t hub_entry(uint32_t count, uint32_t tile_pos) {
    return count << kHubTileBits | tile_pos;
}
__host__ __device__ constexpr uint32_t hub_entry_count(uint32_t w) { return w >> kHubTileBits; }
__host__ __device__ constexpr uint32_t hub_entry_pos(uint32_t w) { return w & (kHubTile - 1); }
constexpr int kHubTileCapBits = 12;   // LDS table of a tile (4096 slots)
constexpr int kHubPart = 2048;        // hub partitions: min(kHubMaxParts, ceil(deg / kHubPart))
constexpr int kHubMaxParts = 4096;
// The sort method cuts hubs finer, min(kHubMaxParts, ceil(deg / 512)): a
// partition then holds about 512 entries even when every label of the row
// is distinct (the first iterations), which one wave merges in registers.
constexpr int kHubPartSort = 512;
constexpr int kHubMergeCapBits = 12;  // LDS table of one partition (4096 slots)

// A merge item (hub, label partition) with what its merge reads first, in
// one 32-byte record: the hub's tiles are tile0 .. tile0 + ntiles - 1, each
// holding nparts + 1 partition offsets, consecutive from tile_idx[idx0].
struct alignas(16) HubItem {
    int32_t hub, part, tile0, ntiles;
    int32_t nparts, pad;
    int64_t idx0;
};

// Parts of the neighbour-id space: (u >> kPartLineBits) mod kNbrParts, the
// 128-byte label line of u modulo kNbrParts.  Rows stored grouped by part
// and tiles listed part-major let the device gather from one part's lines
// at a time (partsort.hpp, the candidate path's multi-tile rows).
constexpr int kNbrParts = 32;
constexpr int kPartLineBits = 5;   // 32 labels = one 128-byte line

struct BucketRange {
    int64_t v0 = 0;      // first renumbered id, or first index into vlist
    int64_t nv = 0;      // vertices in the class
    int64_t edges = 0;   // sum of their degrees
};


// What one round of LP kernels works on: the whole graph in-core, or one
// out-of-core batch (offsets rebased to the batch's neighbour buffer, each
// class and the hub lists cut to the batch's vertices).
struct IterView {
    const void *off = nullptr;      // int32 or int64 offsets (rebased for a batch)
    const int32_t *nbr = nullptr;   // neighbour ids (the batch buffer for a batch)
    BucketRange cls[NUM_BUCKETS];   // v0/nv per class (v0: id, or index into vlist)
    int64_t h0 = 0, h1 = 0;         // hubs [h0, h1)
    int64_t t0 = 0, t1 = 0;         // their tiles
    int64_t i0 = 0, i1 = 0;         // their merge items (hash method)
};

struct KernelAgg {
    std::string name;
    int64_t launches = 0;
    double total_ms = 0.0;
    double alg_bytes = 0.0;   // algorithmic bytes summed over launches
};

struct PendingLaunch {
    int agg;                  // index into Profiler::aggs
    hipEvent_t e0, e1;
    double alg_bytes;
    const int64_t *dev_bytes; // if set: algorithmic bytes counted on the device
};

struct Profiler {
    bool on = false;
    std::vector<KernelAgg> aggs;
    std::vector<PendingLaunch> pending;
    std::vector<hipEvent_t> pool;
    int find(const char *name);
};

// Candidate tiles come in length classes (<= 128, 256, 512, kCandTile
// edges), each walked by a kernel that loads just enough slots per lane.
constexpr int kCandClasses = 4;
__host__ __device__ constexpr int cand_class(int64_t len) {
    return len <= 128 ? 0 : len <= 256 ? 1 : len <= 512 ? 2 : 3;
}
constexpr int kCtrCandTiles = NUM_BUCKETS;                 // kCandClasses entries
constexpr int kCtrOverflow = NUM_BUCKETS + kCandClasses;
constexpr int kCtrCandVerts = kCtrOverflow + 1;
constexpr int kCtrPlainHubs = kCtrOverflow + 2;           // hubs left to the plain hub path
constexpr int kCtrCandMt = kCtrOverflow + 3;              // multi-tile vertices routed to candidates
constexpr int kCtrHd = kCtrOverflow + 4;                  // NUM_BUCKETS sizes of the high-distinct lists
                                                          // (B_HUB's: hubs of the high-distinct hub launch)
constexpr int kCtrHdAll = kCtrHd + NUM_BUCKETS;           // big vertices predicted high-distinct (any path)
constexpr int kNumCtrs = kCtrHdAll + 1;
// A vertex whose predicted distinct-label count exceeds its degree / kHdDiv
// is "high-distinct": its table kernel inserts lane by lane (no peel rounds,
// which only pay when a few labels repeat across a chunk).
constexpr int kHdDiv = 8;

// Candidate path of the learned dispatcher: a wave walks a tile of up to
// kCandTile edges keeping up to kCandK labels in registers.
constexpr int kCandTile = 1024;
constexpr int kCandK = 16;     // candidates per tile (lanes of the candidate table, <= 64)
constexpr int kCandCache = 38912;   // hottest labels held in LDS per workgroup (152 KB)
// In core, the fixed tiles of multi-tile rows hold at most kCandPartTile
// edges of a part-sorted copy of the rows (one part, or consecutive small
// parts), listed part-major.
constexpr int kCandPartTile = 512;

// State of the learned dispatcher (allocated when dispatch is learned or
// training samples are logged).  "Big" vertices are those of the classes
// W128..HUB, the only ones the network routes; big vertices of degree >
// kCandTile are "multi-tile" (mt): hubs first (mt index = hub index), then
// the others.
struct LearnState {
    int H = 0;                         // hidden width of the attached MLP (0: none)
    float *d_params = nullptr;
    float *d_fdeg = nullptr, *d_fmean = nullptr, *d_fmax = nullptr;   // raw static features
    int32_t *d_dist[2] = {nullptr, nullptr};   // distinct labels per vertex, double-buffered
    int dcur = 0;                      // d_dist[dcur]: previous iteration's counts
    int64_t big_v0 = 0, big_n = 0;     // big vertex i: vlist ? vlist[big_v0+i] : big_v0+i
    int64_t front_n = 0;               // big vertices the last dispatch routed (an out-of-core batch's share)
    float cand_thresh = 4.0f;          // predicted distinct labels routed to candidates
    // candidate work list (tiles)
    int64_t cw_cap = 0;
    int64_t cw_base[kCandClasses + 1] = {};   // tile-class segments of the work list
    int64_t *d_cw_base = nullptr;
    int32_t *d_cw_v = nullptr, *d_cw_len = nullptr, *d_cw_pslot = nullptr;
    int64_t *d_cw_e0 = nullptr;
    // multi-tile vertices and their partial candidates
    int64_t nmt = 0, nctiles = 0;
    int32_t *d_mt_vid = nullptr, *d_mt_of = nullptr;        // mt -> vertex, vertex -> mt or -1
    int32_t *d_mt_ctile0 = nullptr, *d_mt_nctiles = nullptr;
    int32_t *d_mt_cand = nullptr;      // mt vertex routed to candidates this iteration
    // fixed tile list of the mt vertices: vertex, first edge, length, partial slot, mt index
    int32_t *d_mt_tv = nullptr, *d_mt_tlen = nullptr, *d_mt_tslot = nullptr, *d_mt_tmt = nullptr;
    int64_t *d_mt_te0 = nullptr;
    int32_t *d_pnbr = nullptr;         // in core: multi-tile rows grouped by neighbour part (mt-major)
    uint32_t *d_ppos = nullptr;        // and the row position of each of their edges
    int32_t *d_pn = nullptr;           // candidates per tile, -1 on overflow
    uint32_t *d_pkey = nullptr, *d_pcnt = nullptr, *d_plast = nullptr;
    int32_t *d_hub_sel = nullptr;      // hub left to the plain hub path this iteration
    uint32_t *d_hub_distinct = nullptr;
    // plain class lists of this iteration (classes W128..B4K): [c] peeled
    // inserts, [NUM_BUCKETS + c] high-distinct rows (lane-by-lane inserts)
    int32_t *plist[2 * NUM_BUCKETS] = {};
    int32_t **d_plist = nullptr;       // device copy of plist
    int32_t *d_hub_hd = nullptr;       // hub on the plain hub path, predicted high-distinct
    // device counters: [0..NUM_BUCKETS) list sizes, then candidate tiles,
    // candidate overflows, vertices routed to candidates
    int64_t *d_plist_n = nullptr;
    int64_t *d_cand_bytes = nullptr;   // algorithmic bytes of the candidate path, per iteration slot
    // training-sample log: x[cap][4], y[cap]
    int64_t log_cap = 0;
    float *d_log_x = nullptr, *d_log_y = nullptr;
    int64_t *d_log_n = nullptr;
};
// State of the sort method: the 8-bit radix passes its LDS sorts make for
// labels < n.  Its hub path uses the hub plan both methods share.
struct SortState {
    int passes = 0;
};



}  // namespace galp

namespace galp {
// Out-of-core state (ooc.hip): the internal-numbering neighbour array in
// pinned host memory, the batch plan, two device batch buffers with their
// rebased offsets, and the copy stream and events of policy 2.
struct OocState {
    int32_t *h_nbr = nullptr;
    std::vector<int64_t> bbs;            // vertex boundaries, #batches + 1
    std::vector<int64_t> e0, e1;         // edge range of each batch
    std::vector<IterView> views;         // each batch's share of the classes and hubs
    int64_t bufcap = 0;                  // edges per device buffer
    int32_t *d_buf[2] = {nullptr, nullptr};
    void *d_offb[2] = {nullptr, nullptr};
    hipStream_t copy_stream = nullptr;
    hipEvent_t ev_copied[2] = {nullptr, nullptr}, ev_free[2] = {nullptr, nullptr};
    bool used[2] = {false, false};
};
}  // namespace galp

namespace galp {
// Batched runs (batch.hip): graph id per vertex (internal numbering), per
// graph vertex count, change counter, stop flag and iterations done.
struct BatchState {
    int64_t ngraphs = 0;
    int64_t max_n = 0;   // vertices of the largest graph
    int32_t *d_gid = nullptr;
    int64_t *d_gn = nullptr, *d_vbase = nullptr;
    unsigned long long *d_gcount = nullptr;
    int32_t *d_gstop = nullptr, *d_giters = nullptr;
    // Work lists (hash method, plain dispatch): per degree class G4..B4K,
    // the vertices of the graphs still running, rebuilt before the first
    // iteration and after every iteration in which a graph stopped
    // (k_batch_lists), with device-side lengths, which the class kernels walk
    // instead of the class vertex ranges.
    bool lists = false;
    int32_t *d_alist[NUM_BUCKETS] = {};
    int32_t **d_alist_dev = nullptr;   // device copy of d_alist
    int64_t *d_alen = nullptr;    // NUM_BUCKETS lengths, then the rebuild flag
};
}  // namespace galp

struct galp_ctx {
    int device = 0;
    hipDeviceProp_t prop{};
    std::string err;
};

struct galp_graph {
    galp_ctx *ctx = nullptr;
    galp_opts opts{};
    int64_t n = 0, m = 0;
    bool off64 = false;
    bool reorder = false;
    hipStream_t stream = nullptr;
    // Side streams for the independent parts of an in-core iteration (the
    // lane-group classes; the LDS classes), forked from and joined back to
    // `stream` by events.
    hipStream_t side[2] = {nullptr, nullptr};
    hipEvent_t side_fork[2] = {nullptr, nullptr}, side_join[2] = {nullptr, nullptr};

    // CSR in HBM (internal numbering when reorder is on)
    void *d_off = nullptr;          // int32 or int64, n+1
    int32_t *d_nbr = nullptr;       // m
    int32_t *d_lab[2] = {nullptr, nullptr};
    int cur = 0;                    // d_lab[cur] holds the current labels
    int32_t *d_perm = nullptr;      // original id -> internal id
    int32_t *d_iperm = nullptr;     // internal id -> original id
    int32_t *d_vlist = nullptr;     // vertex lists per class (no reorder)
    int32_t *d_out = nullptr;       // staging for get_labels (n)

    galp::BucketRange bucket[galp::NUM_BUCKETS];

    // hub path: phase-1 tiles, phase-2 items (hub, label partition)
    int64_t nhubs = 0, nhub_items = 0, nhub_tiles = 0, ntile_idx = 0;
    int32_t *d_hub_vid = nullptr;           // vertex id per hub
    int32_t *d_hub_nparts = nullptr;        // partitions per hub
    int32_t *d_hub_tile0 = nullptr;         // first tile of each hub
    int32_t *d_hub_ntiles = nullptr;        // tiles of each hub
    int32_t *d_tile_hub = nullptr;          // hub of each tile
    int64_t *d_tile_idx_off = nullptr;      // offset of each tile's P_h + 1 partition offsets
    int32_t *d_tile_idx = nullptr;          // partition offsets of every tile
    uint32_t *d_out_key = nullptr, *d_out_ent = nullptr;   // kHubTile per tile: label, hub_entry
    int32_t *d_item_hub = nullptr;          // hub of each merge item
    int32_t *d_item_part = nullptr;         // partition of each merge item
    galp::HubItem *d_item_meta = nullptr;        // the same, with the hub's tile range and offsets
    unsigned long long *d_hub_best = nullptr;  // max (count << 32 | last) per hub
    // per merge item, set by the wave merge in every iteration: 1 when the
    // item has more entries (or distinct labels) than one wave merges, and
    // the block merge takes it.  A flag per item rather than a list: one
    // append per item serialised ~400,000 same-address atomics per
    // iteration at RMAT-26 when every item was large.
    int32_t *d_merge_big = nullptr;

    // iteration control (device)
    unsigned long long *d_counter = nullptr;
    int32_t *d_stop = nullptr;
    int32_t *d_iters_done = nullptr;
    long long *d_iter_counts = nullptr;  // kMaxIters
    int32_t *d_err = nullptr;            // set by a kernel that hit an impossible state
    int32_t *h_pinned = nullptr;         // pinned staging for small readbacks
    int iter = 0;                        // iterations executed since init
    bool initialized = false;
    double preprocess_ms = 0.0;
    int64_t device_bytes = 0;

    galp::Profiler prof;
    galp::LearnState *learn = nullptr;
    galp::SortState *sort = nullptr;
    galp::OocState *ooc = nullptr;
    galp::BatchState *batch = nullptr;
    int8_t tune[galp::NUM_BUCKETS] = {};   // kernel configuration per degree class (autotuner)
    int8_t streams = 0;   // in-core class kernels: 0 by size, 1 one stream, 3 three (the autotuner's choice)
    // learned dispatch in core: bit i set = iteration i routes by the network,
    // clear = by the degree classes alone (the judge's choice, autotune.hip);
    // iterations from 63 on follow bit 63
    uint64_t learn_mask = ~0ull;
    hipGraphExec_t gexec = nullptr;    // captured iteration loop (opts.graph_capture)
    int gexec_niter = -1;
    std::vector<int32_t> big_vid;      // host: big vertices (classes W128..HUB), hubs first
    std::vector<int64_t> big_deg;
    std::vector<int64_t> big_off;      // row start of each big vertex
};

namespace galp {

constexpr int kMaxIters = 4096;

int fail(galp_ctx *ctx, int code, const std::string &msg);

// hipMalloc of count elements (at least one) under the graph's memory budget;
// counted in g->device_bytes.
template <class T>
int dev_alloc(galp_graph *g, T **p, int64_t count) {
    *p = nullptr;
    if (count <= 0) count = 1;
    size_t bytes = sizeof(T) * (size_t)count;
    if (g->opts.mem_budget > 0 && g->device_bytes + (int64_t)bytes > g->opts.mem_budget)
        return fail(g->ctx, GALP_ENOMEM,
                    "device memory budget exceeded: need " + std::to_string(g->device_bytes + (int64_t)bytes) +
                        " bytes, budget " + std::to_string(g->opts.mem_budget));
    hipError_t e = hipMalloc((void **)p, bytes);
    if (e != hipSuccess) {
        *p = nullptr;
        (void)hipGetLastError();
        return fail(g->ctx, GALP_ENOMEM, "hipMalloc of " + std::to_string(bytes) + " bytes failed: " +
                                             hipGetErrorString(e));
    }
    g->device_bytes += (int64_t)bytes;
    return GALP_OK;
}

// The whole graph as one view (in-core).
IterView full_view(const galp_graph *g);
// Kernel launches for one LP round of the hash method over a view (lp_hash.hip).
int launch_hash_iteration(galp_graph *g, const IterView &w);
// One class kernel of the hash method over the whole graph (autotuner).
int launch_hash_class(galp_graph *g, int b);
// Iteration bookkeeping: records the change count, sets the stop flag.
int launch_iter_end(galp_graph *g);
// In-core iterations with independent class kernels on the side streams
// (both methods).  Never out of core, with per-kernel profiling on, or for
// the -l 0 and li layouts; otherwise as the autotuner chose (g->streams), or
// by size: from kConcurrentMinEdges edges (below, an iteration's kernels are
// short and the fork and join cost more than the overlap gains: a
// 20000-vertex graph ran at half speed with them) while the labels one
// gather can reach fit in half of the 256 MB Infinity Cache
// (kConcurrentMaxVertices: at RMAT-26, 268 MB of labels, overlapping
// classes ran 3-6 % slower, at RMAT-22 and LFR-8M 2-7 % faster).  In a
// batch kept in its input numbering a gather stays inside its own graph, so
// the largest graph counts, not the batch.
constexpr int64_t kConcurrentMinEdges = (int64_t)1 << 24;
constexpr int64_t kConcurrentMaxVertices = (int64_t)1 << 25;
bool concurrent_allowed(const galp_graph *g);
bool concurrent_classes(const galp_graph *g);
// Launches made between side_begin(g, k) and side_end(g, k, s) go to side
// stream k, ordered after everything launched on the graph's stream so far;
// side_begin returns the graph's stream, which side_end restores.
// side_join(g, k) orders later launches on the graph's stream after them.
int side_begin(galp_graph *g, int k, hipStream_t *main_stream);
int side_end(galp_graph *g, int k, hipStream_t main_stream);
int side_join(galp_graph *g, int k);
// One in-core hash iteration with the learned front when dispatch is
// learned, classes on side streams when concurrent_classes(g).
int launch_hash_iteration_incore(galp_graph *g);

// Sort method (lp_sort.hip).
int sort_setup(galp_graph *g);
void sort_free(galp_graph *g);
int launch_sort_iteration(galp_graph *g, const IterView &w);
int launch_sort_class(galp_graph *g, int b);   // one class of the whole graph (the autotuner's unit)

// Out-of-core policies (ooc.hip).
int ooc_setup(galp_graph *g, const galp_csr_view *c, const std::vector<int64_t> &offi,
              const std::vector<int32_t> &perm, const std::vector<int32_t> &iperm,
              const std::vector<int32_t> &vlist, const std::vector<int32_t> &hub_vertices,
              const std::vector<int32_t> &hub_tile0, const std::vector<int64_t> &hub_item0);
void ooc_free(galp_graph *g);
int launch_ooc_iteration(galp_graph *g);
// Streams every batch through learn_features_view once (learned dispatch set-up).
int ooc_features(galp_graph *g);

// Looks up the allocation-free entry point's scratch array on `device`
// (current) once per device (mlp.hip).
int mlp_scratch_setup(galp_ctx *ctx, int device);

// Whether the iteration being launched routes by the network: learned
// dispatch, and in core the judge's mask for this iteration (out of core
// every batch routes by the network).
inline bool learned_now(const galp_graph *g) {
    if (g->opts.dispatch != GALP_DISPATCH_LEARNED || !g->learn) return false;
    if (g->ooc) return true;
    const int i = g->iter < 63 ? g->iter : 63;
    return ((g->learn_mask >> i) & 1ull) != 0;
}

// Where this iteration's paths record each big row's distinct-label count:
// the next iteration's dispatch input and the logged training target, so
// nowhere (null) when the next iteration routes by the degree classes alone
// and nothing is logged.
inline int32_t *distinct_out(const galp_graph *g) {
    const LearnState *L = g->learn;
    if (!L) return nullptr;
    const int i = g->iter + 1 < 63 ? g->iter + 1 : 63;
    const bool next_learned = g->opts.dispatch == GALP_DISPATCH_LEARNED &&
                              (g->ooc || ((g->learn_mask >> i) & 1ull) != 0);
    return next_learned || L->log_cap ? L->d_dist[L->dcur ^ 1] : nullptr;
}

// Drops the captured iteration loop after a change to what it launches.
inline void invalidate_captured(galp_graph *g) {
    if (g->gexec) (void)hipGraphExecDestroy(g->gexec);
    g->gexec = nullptr;
    g->gexec_niter = -1;
}

// Batched runs (batch.hip).
void batch_free(galp_graph *g);
int batch_reset(galp_graph *g);
int launch_batch_end(galp_graph *g);
// Rebuilds the batch's work lists from the graphs' stop flags (start of an
// iteration); a no-op without lists.
int launch_batch_lists(galp_graph *g);
inline bool batch_lists(const galp_graph *g) { return g->batch && g->batch->lists; }

// Learned dispatch (learned.hip).
int learn_setup(galp_graph *g);
void learn_free(galp_graph *g);
int learn_reset_iteration(galp_graph *g);
// The dispatch and the candidate path; with merge = false the multi-tile
// merge is left to launch_learned_merge.
int launch_learned_front(galp_graph *g, bool merge = true);
int launch_learned_merge(galp_graph *g);
// Out-of-core: the dispatch and candidate path of one batch (vertices
// [lo, hi), the batch's rebased offsets and neighbour buffer in w).
int launch_learned_front_batch(galp_graph *g, const IterView &w, int64_t lo, int64_t hi);
// Static features of the big vertices among [lo, hi) from a batch's rows.
int learn_features_view(galp_graph *g, const void *offrow, const int32_t *nbr, int64_t lo, int64_t hi);
int learn_after_iteration(galp_graph *g);

// Graph bookkeeping kernels (graph_dev.hip), all on g->stream.
int dev_init_labels(galp_graph *g);
int dev_clear_stop(galp_graph *g);
int dev_fill_u32(galp_graph *g, uint32_t *p, int64_t n, uint32_t val);
int dev_permute_rows(galp_graph *g, const int64_t *d_off_old, const int32_t *d_nbr_old,
                     int32_t *d_nbr_new);
int dev_map_labels(galp_graph *g, const int32_t *lab, int32_t *out);
int dev_unmap_labels(galp_graph *g, const int32_t *in, int32_t *lab);

// Profiling helpers (api.hip): bracket a launch with events when enabled.
void prof_begin(galp_graph *g, const char *name, double alg_bytes, PendingLaunch *pl);
void prof_end(galp_graph *g, PendingLaunch *pl);
// As prof_begin, with the launch's algorithmic bytes read from the device
// when the timings are collected.
void prof_begin_dev(galp_graph *g, const char *name, const int64_t *dev_bytes, PendingLaunch *pl);

}  // namespace galp
