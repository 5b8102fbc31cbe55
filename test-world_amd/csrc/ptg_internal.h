// Internal declarations shared by the host runtime and the HIP kernels.
#pragma once

#include <hip/hip_ext.h>
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <functional>
#include <string>
#include <unordered_map>
#include <vector>

#include "ptg.h"

namespace ptg {

// Records a failure detail for ptg_last_error() and returns `code`.
int fail(int code, const std::string &msg);
// Converts a HIP status into PTG_ERR_HIP with a message naming `what`.
int hip_fail(hipError_t e, const char *what);

#define PTG_HIP(call)                                            \
    do {                                                         \
        hipError_t _e = (call);                                  \
        if (_e != hipSuccess) return ::ptg::hip_fail(_e, #call); \
    } while (0)

constexpr int kBlock = 256;  // threads per workgroup (4 waves of 64)
constexpr int kWavesPerBlock = kBlock / 64;
constexpr uint32_t kChunkEdges = 2048;  // edges per workgroup task for hub vertices
constexpr uint32_t kRing = 4096;        // per-round counter blocks kept on the device
constexpr int kShards = 16;             // counters and queues are split 16 ways
constexpr uint32_t kQueueCap = 16384;   // per-shard queue capacity = max queued frontier
constexpr uint32_t kModeQueue = 0, kModeBitmap = 1;
constexpr uint32_t kDirTopDown = 0, kDirBottomUp = 1;

constexpr uint32_t kWaveChunk = 64;    // edges per chunk of the frontier chunk queue (one wave step)

// Description of the frontier consumed by round r (ring[r % kRing]) and the
// state carried into that round. Written by round r - 1 (and, for bitmap
// frontiers, the chunk queue by the pre-pass of round r): header fields by one
// thread, counts through sharded atomics (workgroup b uses shard b % 16).
// A frontier vertex of degree d is queued in the S queue (0 < d <= sd) or the
// chunk queue (d > sd, ceil(d / kWaveChunk) chunks).
struct RoundCtr {
    uint32_t qn[kShards];                   // S-queue entries per shard
    unsigned long long cpack[kShards];      // chunk queue per shard: (entries << 40) | chunks
    uint32_t mode;                          // representation (queue, bitmap or flags)
    uint32_t dir;                           // BFS: direction of the producing round
    uint32_t thr;                           // SSSP: near/far threshold of this frontier
    uint32_t any;                           // 1 if the producing round emitted anything
    unsigned long long seen;                // BFS: sum of out-degrees of earlier frontiers
    uint32_t far_any;                       // SSSP: far set non-empty after the producing round
    uint32_t need_heavy;                    // SSSP: the bucket's heavy edges are still to relax
    uint32_t lo;                            // SSSP: lower end of the current bucket
    uint32_t fmin_inv;                      // SSSP: ~(smallest value in this frontier); 0 = unknown
    uint32_t ndefer;                        // SSSP: 1 if the pre-pass deferred any vertex
    uint32_t ucount;                        // SSSP: 1 if the pre-pass filled uedges_sh
    uint32_t pulled;                        // SSSP: 1 if the round reading this entry pulled
    uint32_t swept;                         // SSSP: 1 if the round reading this entry swept every edge
    uint32_t pad[2];
    uint32_t found_sh[kShards];             // frontier vertices (any degree)
    uint32_t far_sh[kShards];               // SSSP: vertices added to / kept in the far set
    unsigned long long edges_sh[kShards];   // sum of out-degrees of the frontier
    unsigned long long uedges_sh[kShards];  // SSSP: in-degrees of the vertices that can still improve
    unsigned long long dedges_sh[kShards];  // SSSP: out-degrees of the deferred frontier vertices
};
static_assert(sizeof(RoundCtr) == 768, "RoundCtr layout");

// Counters of one frontier. Bins: S = lane per vertex, M = wave per vertex,
// L = edge chunks of kChunkEdges spread over workgroups.
struct BinCounters {
    uint32_t n_s;                 // entries in the S queue
    uint32_t n_m;                 // entries in the M queue
    unsigned long long l_pack;    // L queue: (entries << 40) | total chunks
    uint32_t found;               // vertices discovered / improved (any degree)
    uint32_t pad;
    unsigned long long edges;     // sum of out-degrees of `found`
    unsigned long long aux;       // algorithm specific
};
static_assert(sizeof(BinCounters) == 40, "BinCounters layout");

constexpr unsigned long long kLCountShift = 40;
constexpr unsigned long long kLChunkMask = (1ull << kLCountShift) - 1;

inline uint32_t l_entries(const BinCounters &c) { return (uint32_t)(c.l_pack >> kLCountShift); }
inline uint64_t bin_total(const BinCounters &c) { return (uint64_t)c.n_s + c.n_m + l_entries(c); }

// Queues of one frontier plus the counters they are filled against.
struct Frontier {
    uint32_t *qs;  // S queue (vertex ids)
    uint32_t *qm;  // M queue
    uint32_t *ql;  // L queue vertex ids
    uint32_t *lb;  // L queue first-chunk indices (non-decreasing with the slot)
    BinCounters *ctr;
};

// Queue storage without counters.
struct QueueSet {
    uint32_t *qs, *qm, *ql, *lb;
    Frontier with(BinCounters *c) const { return Frontier{qs, qm, ql, lb, c}; }
};

// Vertices of a CSR binned by degree once (CC hooking, PageRank pull).
struct StaticBins {
    QueueSet q{};
    uint32_t *zero = nullptr;     // degree-0 vertices
    BinCounters *d_ctr = nullptr; // counts of the bins on the device
    BinCounters h_ctr{};
    uint32_t n_zero = 0;
    uint32_t *mem = nullptr;
};

struct Csr {
    uint32_t *off = nullptr;  // V + 1
    uint32_t *col = nullptr;  // E
    StaticBins bins;
};

}  // namespace ptg

namespace ptg {
struct BatchState;   // batch.hip
struct StreamState;  // stream.hip
struct PrBins;       // pagerank_bins.hip
}

// The graph handle.
struct ptg_graph {
    uint32_t V = 0;
    uint64_t E = 0;
    int device = 0;
    int num_cus = 0;
    ptg_config cfg{};
    uint32_t grid = 0;  // persistent grid (workgroups)
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;

    // CSR in HBM: out-edges (+ SoA weights) and, when needed, in-edges.
    ptg::Csr out;
    uint32_t *d_w = nullptr;
    ptg::Csr in;             // aliases `out` on a symmetric graph
    bool in_is_out = false;
    bool have_in = false;

    // Query state, allocated once per graph and reused by every query.
    uint32_t *d_val = nullptr;          // V: levels / distances / labels
    uint32_t *d_bits[5] = {};           // bitmaps of nwords words: 3 rotating frontiers,
                                        // 2 scratch (symmetry-check flags)
    uint8_t *d_flags[2] = {};           // frontier byte flags (npairs * 64 bytes)
    uint8_t *d_far = nullptr;           // SSSP far-set byte flags
    uint32_t nwords = 0;                // 2 * npairs
    uint32_t npairs = 0;                // ceil(V / 64)
    uint32_t *d_queue[2] = {};          // sharded S queues (kShards x kQueueCap)
    uint32_t *d_cq[2] = {};             // sharded chunk queues: vertex ids (kShards x cq_cap)
    uint32_t *d_cb[2] = {};             // sharded chunk queues: first-chunk indices
    uint32_t cq_cap = 0;                // chunk-queue entries per shard
    uint32_t *d_qmem = nullptr;
    ptg::RoundCtr *d_ring = nullptr;    // kRing per-round counter blocks
    ptg::RoundCtr *h_ring = nullptr;    // pinned mirror
    unsigned long long *d_acc = nullptr; // 2 x kShards accumulators for statistics kernels
    unsigned long long *h_acc = nullptr; // pinned mirror of d_acc
    // Output delivery of the running query (unpermute and/or copy to the
    // caller's buffer), registered by the entry point. BFS and SSSP enqueue
    // it themselves (deliver_output) before their final wait; for the other
    // queries the entry point runs it after the query returns.
    std::function<int()> deliver;
    bool deliver_on_device = false;  // device-to-device only (cheap to repeat)
    bool delivered = false;

    // PageRank state.
    float *d_contrib[2] = {};
    unsigned long long *d_part = nullptr;   // PageRank dangling partials and the base value
    unsigned long long *d_lpart = nullptr;  // hub chunk sums (fixed point)
    uint32_t lchunks_cap = 0;
    ptg::PrBins *prb = nullptr;  // binned PageRank layout (pagerank_bins.hip), built on first use

    // Per-launch timing of the main kernel (PTG_FLAG_TIMING).
    std::vector<hipEvent_t> tev;  // pairs: [2i] before, [2i+1] after launch i
    uint32_t tev_used = 0;

    // Degree-ordered relabeling (relabel.cpp): device arrays are in new ids.
    bool relabeled = false;
    std::vector<uint32_t> h_perm;       // old id -> new id (host copy for sources)
    uint32_t *d_perm = nullptr;         // old id -> new id
    uint32_t *d_inv = nullptr;          // new id -> old id
    uint32_t *d_val2 = nullptr;         // V: results mapped back to old ids

    // SSSP light/heavy split of the weighted CSR for step split_delta
    // (sssp.hip): light = edges of weight < delta, heavy = the rest.
    uint32_t split_delta = 0;
    ptg::Csr light, heavy;
    uint32_t *d_wl = nullptr, *d_wh = nullptr;
    // SSSP deferred high-degree vertices of a round (bitmap of nwords words,
    // zero between rounds) and the automatic deferral window (0 = not
    // computed yet).
    uint32_t *d_defer = nullptr;
    uint2 *d_cw = nullptr;  // (column, weight) pairs for the default SSSP rounds, built on first use
    // (source, weight) pairs of every vertex's in-edges, in the offsets of
    // `out` (a symmetric graph's in-degrees are its out-degrees): the SSSP
    // pull rounds read them. Built on first use.
    uint2 *d_cwin = nullptr;
    // The long lists of the SSSP pull rounds (sssp.hip), built on first use:
    // their vertices, first 64-edge chunks and the entry of every chunk block.
    uint32_t *d_long_ids = nullptr, *d_long_base = nullptr, *d_long_block = nullptr;
    uint32_t n_long = 0, long_chunks = 0, pull_tile_deg = 0;
    bool long_built = false;
    uint32_t auto_window = 0;
    uint32_t wmin = 0, wmax = 0;  // smallest and largest edge weight (valid once auto_window is set)
    bool defer_dirty = false;  // the deferred bitmap may hold bits of a failed query

    // Batched multi-source state (batch.hip), allocated on the first batch.
    ptg::BatchState *batch = nullptr;
    // batched SSSP: the rounds of the next batch that pull, learned from the
    // last batch (kept when the batch state is released, e.g. by a server)
    uint8_t batch_pull_rounds[64] = {};
    bool batch_pull_learned = false;
    // PTG_FLAG_STREAM: host-resident edges and the device chunk cache (stream.hip).
    ptg::StreamState *streamed = nullptr;

    // Frontier size of every round of the last BFS / SSSP / label-propagation query.
    std::vector<uint64_t> frontiers;
    uint32_t pulled_rounds = 0;  // rounds of the last BFS / SSSP / label-propagation query that pulled
    uint64_t pulled_mask = 0;    // which of its rounds 0..63 pulled
    // SSSP: rounds 0..63 that pulled in the last two queries (all ones until
    // known) and the queries run, for the pull kernel's launch window.
    uint64_t pull_hist[2] = {~0ull, ~0ull};
    uint32_t pull_queries = 0;
    // The same for the binned sweep rounds (rounds of the last query that
    // swept, and those of the last two queries).
    uint32_t swept_rounds = 0;
    uint64_t swept_mask = 0;
    uint64_t sweep_hist[2] = {~0ull, ~0ull};
    // Rounds of the last BFS [0], SSSP [1] and label-propagation [2] query
    // (0 = none yet): the next query of the kind launches that many rounds + 1
    // before its first readback.
    uint32_t last_rounds[3] = {0, 0, 0};

    int symmetric = -1;  // -1 unknown, 0/1 once checked
    int wsymmetric = -1; // weights symmetric too: -1 unknown, 0/1 once checked
    uint64_t device_bytes = 0;                          // charged against the budget
    uint64_t create_bytes = 0;                          // device_bytes when creation finished
    std::unordered_map<void *, uint64_t> alloc_bytes;   // live allocations and their sizes
    bool weighted = false;
};

namespace ptg {

// Device allocation that is charged against the graph's budget.
int galloc(ptg_graph *g, void **p, uint64_t bytes);
void gfree(ptg_graph *g, void *p);
bool is_device_pointer(const void *p);

// Entries per shard of a chunk queue for V vertices and `grid` workgroups.
uint32_t chunk_queue_cap(uint32_t V, uint32_t grid);
// Workgroups of the per-round pre-pass for a round grid of `grid`: half of
// it (each workgroup then takes twice the tiles; 18% faster than the full
// grid and than a quarter of it at RMAT-22).
inline uint32_t prepare_grid(uint32_t grid) { return grid > 1 ? grid / 2 : 1u; }

// Host-built degree bins uploaded to the device.
int build_static_bins(ptg_graph *g, const uint32_t *h_off, uint32_t V, StaticBins *sb);
void free_static_bins(ptg_graph *g, StaticBins *sb);

// Kernel drivers (frontier.hip, pagerank.hip). `out` is a device pointer.
int fill_u32(ptg_graph *g, uint32_t *p, uint64_t n, uint32_t v);
int copy_bytes(void *dst, const void *src, uint64_t n, hipStream_t stream);
// Copy on the graph's own stream, then wait for it. The legacy default stream
// is never used: another thread may be capturing a graph replay meanwhile.
int copy_sync(ptg_graph *g, void *dst, const void *src, uint64_t bytes, hipMemcpyKind kind);
int run_bfs(ptg_graph *g, uint32_t src, uint32_t *out, ptg_stats *st);
int run_sssp(ptg_graph *g, uint32_t src, uint32_t *out, ptg_stats *st);
int run_cc(ptg_graph *g, uint32_t *out, ptg_stats *st);
int run_pagerank(ptg_graph *g, float damping, int iters, float *out, ptg_stats *st);
// Binned PageRank (pagerank_bins.hip): *done = false when the graph takes the
// in-list pull instead (pr_layout, or too many pairs for the automatic layout).
int run_pagerank_bins(ptg_graph *g, float damping, int iters, float *out, ptg_stats *st, bool *done);
bool pr_bins_wanted(const ptg_graph *g);
uint64_t pr_bins_need(const ptg_graph *g);
void release_pr_bins(ptg_graph *g);
void destroy_pr_bins(ptg_graph *g);
int pr_bins_info(const ptg_graph *g, uint64_t *pairs, uint32_t *chunks, uint32_t *rows);
// SSSP sweep rounds over the binned layout (pagerank_bins.hip). *ok = false
// when the graph cannot sweep (weights above 255, or the automatic layout
// would not pay); else the layout, with a weight byte per entry and a source
// per pair, is on the device. sssp_sweep_need: device bytes that call would
// still allocate.
int sssp_sweep_prepare(ptg_graph *g, bool *ok);
uint64_t sssp_sweep_need(const ptg_graph *g);
// Launches the two kernels of a possible sweep of round env.r behind its
// push (and pull) kernels; each returns at once unless the round sweeps
// (round_sweeps with the same factor and edge threshold).
struct RoundEnv;
struct LaunchTimer;
int sssp_sweep_launch(ptg_graph *g, const RoundEnv &env, uint32_t *val, uint32_t sent, float pull_factor,
                      uint64_t min_edges, LaunchTimer &timer);
// PageRank with its per-iteration buffers in a fixed per-GPU device array of
// kPrScratchBytes (pagerank.hip); allocates nothing.
constexpr uint64_t kPrScratchBytes = 64ull << 20;
int pr_scratch_setup(int device);
uint64_t pr_scratch_need(const ptg_graph *g);
int run_pagerank_noalloc(ptg_graph *g, float damping, int iters, float *out, ptg_stats *st);
int check_symmetric(ptg_graph *g, int *sym);
// Whether the weighted multiset of (v, u, w) equals that of (u, v, w), from
// the graph's (column, weight) pairs cw (once per graph).
int check_weight_symmetric(ptg_graph *g, const uint2 *cw, int *wsym);
int sssp_in_pairs(ptg_graph *g, const uint2 **cwin);
uint64_t sssp_in_pairs_need(const ptg_graph *g);
int ensure_in_edges(ptg_graph *g);
int reach_stats(ptg_graph *g, const uint32_t *val, uint32_t unreached, uint32_t *reached,
                uint64_t *edges);
// The same in three steps inside a query: zero d_acc (reach_reset_span, part
// of the query's reset), enqueue the count and its readback, and read the
// pinned copy after the stream has been waited for.
struct FillSpan {
    uint32_t *p;
    uint64_t n;
    uint32_t v;
};
FillSpan reach_reset_span(ptg_graph *g);
int reach_stats_launch(ptg_graph *g, const uint32_t *val, uint32_t unreached);
void reach_stats_read(const ptg_graph *g, uint32_t *reached, uint64_t *edges);
// Fills up to kMaxSpans word ranges with one launch.
constexpr uint32_t kMaxSpans = 10;
int fill_spans(ptg_graph *g, const FillSpan *spans, uint32_t k);
// Enqueues the registered output delivery, if any (it may be enqueued again
// later: the last one wins). A speculative call (the query may need more
// rounds) delivers only to device memory: a copy to the host is made once,
// after the last round.
int deliver_output(ptg_graph *g, bool speculative = false);
// Batched queries (batch.hip): `out` is a device buffer of n x V words in
// the caller's ids; batch_staging gives a library-owned one of 64 x V.
int run_batch_query(ptg_graph *g, bool sssp, const uint32_t *srcs, uint32_t n, uint32_t *out, ptg_stats *st);
int batch_staging(ptg_graph *g, uint32_t **out);
void destroy_batch(ptg_graph *g);
// Serving support: bytes a query kind would still allocate, release of the
// lazily allocated query state, and per-row checksums of n x V results.
uint64_t batch_state_need(const ptg_graph *g, bool sssp);
uint64_t pr_state_need(const ptg_graph *g);
uint64_t sssp_split_need(const ptg_graph *g);
void release_sssp_split(ptg_graph *g);
// Frees the in-edge pairs of the SSSP pull rounds (rebuilt on the next query that pulls).
void release_sssp_pull(ptg_graph *g);
void release_pr_state(ptg_graph *g);
int row_checksums(ptg_graph *g, hipStream_t stream, const uint32_t *rows, uint32_t n, unsigned long long *d_out);
// Streamed graphs (stream.hip).
int create_streamed(ptg_graph *g, const uint32_t *ptr, const uint32_t *cols, const uint32_t *weights,
                    const EdgeWithWeight *wedges);
int run_sssp_streamed(ptg_graph *g, uint32_t src, uint32_t *dist, ptg_stats *st);
// (Re)cuts a streamed graph into chunks of at most chunk_bytes and sizes its
// slot cache by the budget.
int stream_plan(ptg_graph *g, uint64_t chunk_bytes);
int run_cc_streamed(ptg_graph *g, uint32_t *label, ptg_stats *st);
// Union-find and label helpers of cc.hip, on the graph's stream.
int uf_start(ptg_graph *g, uint32_t *par);
int uf_finish(ptg_graph *g, uint32_t *par);
int init_keys(ptg_graph *g, uint32_t *val);
int count_roots(ptg_graph *g, const uint32_t *label, uint32_t *roots);
void destroy_streamed(ptg_graph *g);
int run_bfs_streamed(ptg_graph *g, uint32_t src, uint32_t *level, ptg_stats *st);
int run_pagerank_streamed(ptg_graph *g, float damping, int iters, float *ranks, ptg_stats *st);
int host_link_probe(uint64_t bytes, int reps, double *gbs);
int stream_info(const ptg_graph *g, uint32_t *chunks, uint32_t *slots, uint64_t *slot_bytes);
// dst[v] = src[perm[v]] for v < V (new-id results back to old ids).
int unpermute(ptg_graph *g, const uint32_t *src, uint32_t *dst);

// Brackets main-kernel launches with HIP events on the graph's stream when
// PTG_FLAG_TIMING is set; collect() sums the pairs after the stream is idle.
struct LaunchTimer {
    ptg_graph *g;
    bool on;
    uint32_t riders = 0;  // timed launches that belong to the round of the launch before
    explicit LaunchTimer(ptg_graph *gr);
    // Event pair for the next timed launch, or false when timing is off.
    bool next(hipEvent_t *start, hipEvent_t *stop);
    int collect(ptg_stats *st);

    // Launches `kernel`; when timing is on, the dispatch itself records its
    // start and stop (hipExtLaunchKernelGGL), which adds no stream packets.
    template <class K, class... A>
    void launch(K kernel, dim3 grid, dim3 block, A... args) {
        hipEvent_t a, b;
        if (next(&a, &b))
            hipExtLaunchKernelGGL(kernel, grid, block, 0, g->stream, a, b, 0, args...);
        else
            hipLaunchKernelGGL(kernel, grid, block, 0, g->stream, args...);
    }
    // Same, for a second kernel of the same round: its time adds to the
    // main kernel's, its launch does not count as another round.
    template <class K, class... A>
    void launch_rider(K kernel, dim3 grid, dim3 block, A... args) {
        const bool timed = on;
        if (timed) riders++;
        launch(kernel, grid, block, args...);
        if (timed && !on) riders--;  // timing switched itself off at its cap
    }
};

}  // namespace ptg
