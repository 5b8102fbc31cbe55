/*
 * ptg.h -- C-ABI of the MI355X (gfx950) graph hot path: BFS levels, SSSP
 * distances, connected-component labels and PageRank over a CSR graph.
 *
 * This header is the drop-in boundary for ptGraph's host contract
 * (ptGraph/gpu_kernels.cuh:26-124 kernels + ptGraph/main.cu:24-56 drivers):
 *   - CSR in the reference layout: `nodePointers` holds V entries (not V+1),
 *     the degree of the last vertex is E - nodePointers[V-1]
 *     (ptGraph/main.cu:257-264); columns are uint32 (`uint *edgeList`), weighted
 *     edges are the 8-byte AoS `EdgeWithWeight` (ptGraph/globals.cuh:2-5).
 *   - Outputs follow the reference's value conventions:
 *       BFS  level[src] = 1, level[v] = 1 + hops, unreached 0xFFFFFFFF
 *            (ptGraph/main.cu:1068-1076, :1745, :1793-1794);
 *       SSSP dist[src] = 1, dist[v] = 1 + path weight, unreached = sentinel,
 *            0xFFFFFFFE by default (ptGraph/main.cu:584-588) or 0xFFFFFFFF
 *            (ptGraph/main.cu:742-746);
 *       CC   label[v] = the smallest vertex id u with a path u -> v (the fixed
 *            point of ptGraph/gpu_kernels.cu:136-150 started from
 *            value[v] = v, ptGraph/main.cu:977-980); on a symmetric graph this
 *            is the minimum vertex id of v's connected component.
 *   - Errors are returned as negative codes; the library never exits
 *     (replaces gpuErrorcheck/exit, ptGraph/main.cu:15-22).
 *
 * Output pointers may be host or device memory; the library detects which.
 * A graph handle is used by one host thread at a time.
 */
#ifndef PTG_H
#define PTG_H

#include <stddef.h>
#include <stdint.h>

#if defined(__GNUC__)
#define PTG_API __attribute__((visibility("default")))
#else
#define PTG_API
#endif

#ifdef __cplusplus
extern "C" {
#endif

/* Bit-identical to ptGraph/globals.cuh:2-5. */
typedef struct EdgeWithWeight {
    uint32_t toNode;
    uint32_t weight;
} EdgeWithWeight;

/* Return codes. */
enum {
    PTG_OK = 0,
    PTG_ERR_NO_DEVICE = -1,     /* no HIP device / runtime not usable */
    PTG_ERR_INVALID_ARG = -2,   /* null pointer, bad size, source >= V ... */
    PTG_ERR_INVALID_GRAPH = -3, /* CSR fails validation */
    PTG_ERR_OUT_OF_MEMORY = -4, /* device or pinned allocation failed */
    PTG_ERR_BUDGET = -5,        /* does not fit the configured device budget */
    PTG_ERR_HIP = -6,           /* a HIP runtime call failed */
    PTG_ERR_UNSUPPORTED = -7,   /* operation not available for this graph */
    PTG_ERR_REJECTED = -8,      /* serving admission refused the request */
    PTG_ERR_IO = -9             /* file read/write failed or bad format */
};

/* Graph-creation flags (ptg_config.flags). */
#define PTG_FLAG_NO_VALIDATE 0x1u  /* skip the host-side CSR validation */
#define PTG_FLAG_TIMING 0x2u       /* time every main-kernel launch with HIP events */
#define PTG_FLAG_NO_RELABEL 0x4u   /* keep the caller's vertex order on the device */
/* Keep the edges in pinned host memory, cut into chunks of at most
 * chunk_bytes, and stream them into a device cache sized by
 * device_budget_bytes (replaces the reference's partition-swap path,
 * main.cu:1681-2059, and the Opt drivers' static prefix + overload split,
 * main.cu:222-440, :678-889, :1180-1418). Such a graph answers BFS, SSSP
 * (weighted graphs: the weights stream with the columns when SSSP needs
 * them), CC (one union-find pass over the chunks on a symmetric graph, label
 * propagation rounds otherwise) and PageRank (graphs symmetric as
 * multigraphs); batched queries return PTG_ERR_UNSUPPORTED. */
#define PTG_FLAG_STREAM 0x8u

typedef struct ptg_config {
    uint64_t device_budget_bytes; /* 0 = whole device; cap on graph + state */
    uint64_t chunk_bytes;         /* streamed CSR chunk size (partition path) */
    uint32_t sssp_sentinel;       /* unreached SSSP value: 0xFFFFFFFE or 0xFFFFFFFF */
    uint32_t grid_blocks;         /* persistent grid size in workgroups; 0 = auto */
    uint32_t small_degree;        /* degree <= small_degree: lane per vertex */
    uint32_t large_degree;        /* degree >  large_degree: split over workgroups */
    float td_to_bu;               /* BFS switches to bottom-up when m_f > m_u / td_to_bu */
    float bu_to_td;               /* BFS returns to top-down when n_f < V / bu_to_td */
    uint32_t sssp_delta;          /* near/far bucket step for SSSP (edges of weight < step are
                                     light); 0 = one unbounded bucket, no split */
    uint32_t flags;               /* PTG_FLAG_* */
    uint32_t batch_delta;         /* near/far step of batched SSSP; 0 = auto */
    uint32_t batch_grid;          /* workgroups of the batched BFS (and, unless batch_sssp_grid
                                     is set, SSSP) kernels; 0 = 4 x grid_blocks */
    float batch_td_to_bu;         /* batched BFS goes bottom-up when m_f > m_u / this; 0 = td_to_bu */
    uint32_t batch_chunk;         /* edges per static chunk of batched SSSP; 0 = auto (128) */
    uint32_t order_hub_degree;    /* vertex order on the device: 0 = by descending degree; d > 0 =
                                     vertices of degree > d by degree, then breadth-first from them */
    uint32_t sssp_defer_degree;   /* SSSP: a vertex of degree > this waits in the frontier while its
                                     distance >= the frontier minimum + sssp_defer_window;
                                     0 = auto (128), 0xFFFFFFFF = never */
    uint32_t sssp_defer_window;   /* 0 = auto (mean weight / 16, at least 1) */
    uint32_t stream_pinned;       /* PTG_FLAG_STREAM: the first stream_pinned chunks stay in their
                                     slots once loaded (at most slots - 2); in the default degree
                                     order they hold the highest-degree vertices, like the
                                     reference's degree-sorted static set (gpu_kernels.cu:83-96,
                                     :420-448, main.cu:2196-2394). 0 = least recently used only */
    uint32_t batch_sssp_grid;     /* workgroups of the batched SSSP kernels; 0 = batch_grid */
    float sssp_pull;              /* SSSP on a symmetric graph, one unbounded bucket: a large round
                                     pulls (every vertex that can still improve takes the minimum
                                     over its in-edges) when those vertices' in-edges x sssp_pull
                                     < the frontier's out-edges; the distances are the same.
                                     0 = auto (1.0), < 0 = never (the in-edge pairs, 8 bytes per
                                     edge, are then not built) */
    float batch_pull;             /* batched SSSP on a symmetric graph: round r of a batch pulls
                                     (every vertex takes, for every source, the minimum over its
                                     in-edges, one atomic per edge chunk instead of one per
                                     improving edge) when round r of the graph's previous batch
                                     expanded >= E x batch_pull edges and round r - 1 of this one
                                     >= E x batch_pull / 8; the distances are the same.
                                     0 = auto (0.25), < 0 = never */
    uint32_t pr_block;            /* threads per workgroup of the PageRank gather: 256, 512 or 1024
                                     (0 = 256). The hub contributions it stages in LDS grow with
                                     it (5088, 10176, 20352 vertices), so a CU holds the same 160
                                     KB of LDS; the ranks are the same bits for every value */
    uint32_t pr_layout;           /* PageRank on a resident graph: PTG_PR_AUTO (binned when the
                                     (source, destination partition) pairs come to at most half
                                     the edges, else the in-list pull), PTG_PR_PULL or
                                     PTG_PR_BINNED. Every layout gives the same bits */
    float sssp_sweep;             /* SSSP where pulls are possible, weights <= 255 and the binned
                                     layout pays (as for PageRank): a large round relaxes every
                                     edge through the binned layout (a byte of weight per edge,
                                     an LDS minimum per destination partition) when the round's
                                     push (its out-edges) and pull (in-edges x sssp_pull) would
                                     both scan more than E x sssp_sweep edges; the distances are
                                     the same. 0 = auto (0.25), < 0 = never (the layout is then
                                     not built for SSSP) */
} ptg_config;

/* ptg_config.pr_layout */
#define PTG_PR_AUTO 0u
#define PTG_PR_PULL 1u
#define PTG_PR_BINNED 2u

/* Per-query statistics (the reference prints `iter` and `nodeSum`,
 * ptGraph/main.cu:1109-1113; these are their counterparts). */
typedef struct ptg_stats {
    uint32_t iterations;      /* frontier expansions; BFS: eccentricity(src) + 1 */
    uint32_t reached;         /* vertices with a finite value (BFS/SSSP) or components (CC) */
    uint64_t edges_traversed; /* m_r: sum of out-degrees of reached vertices */
    uint64_t frontier_sum;    /* sum of frontier sizes (reference nodeSum) */
    uint64_t kernel_ns;       /* device time of the whole query (HIP events) */
    uint64_t host_link_bytes; /* bytes copied host -> device during the query */
    uint64_t host_link_ns;    /* time of those copies */
    uint64_t main_kernel_ns;  /* PTG_FLAG_TIMING: summed HIP-event time of the query's
                                 main kernel (frontier expansion / PageRank sweep; SSSP:
                                 the push, pull and sweep kernels of every round) */
    uint32_t main_launches;   /* PTG_FLAG_TIMING: launches of that kernel (SSSP: rounds) */
    uint32_t chunk_loads;     /* PTG_FLAG_STREAM: chunks copied from the host during the query */
    uint32_t chunk_hits;      /* PTG_FLAG_STREAM: chunk uses served from the device cache */
    uint32_t pull_rounds;     /* SSSP: rounds that pulled over in-edges (sssp_pull) */
    uint32_t sweep_rounds;    /* SSSP: rounds that swept every edge (sssp_sweep) */
} ptg_stats;

typedef struct ptg_graph ptg_graph;

/* Host CSR with V+1 offsets (offsets[V] == E); produced by the generators and
 * readers below, released with ptg_host_csr_free. `weights` may be NULL. */
typedef struct ptg_host_csr {
    uint32_t V;
    uint64_t E;
    uint32_t *offsets; /* V + 1 entries */
    uint32_t *cols;    /* E entries */
    uint32_t *weights; /* E entries or NULL */
} ptg_host_csr;

/* ---- runtime -------------------------------------------------------------- */
PTG_API int ptg_init(int device);             /* PTG_ERR_NO_DEVICE without a GPU */
PTG_API const char *ptg_strerror(int code);
PTG_API const char *ptg_last_error(void);     /* detail of the last failure on this thread */
PTG_API void ptg_config_default(ptg_config *cfg);
PTG_API int ptg_device_info(int *num_cus, uint64_t *total_mem, uint64_t *free_mem);

/* Device buffers for callers that keep outputs resident in HBM. */
PTG_API int ptg_device_alloc(uint64_t bytes, void **ptr);
PTG_API int ptg_device_free(void *ptr);
PTG_API int ptg_memcpy_d2h(void *dst, const void *src, uint64_t bytes);
PTG_API int ptg_memcpy_h2d(void *dst, const void *src, uint64_t bytes);
/* Device-to-device copy by the library's own streaming kernel (k_copy_bytes,
 * 16 B per lane when both pointers are 16-byte aligned); synchronous. Its
 * byte counters read exactly 2 x bytes of algorithmic traffic, which is what
 * calibrates hardware counters. */
PTG_API int ptg_memcpy_d2d(void *dst, const void *src, uint64_t bytes);

/* PTG_FLAG_STREAM graphs: number of chunks, device cache slots and bytes per
 * slot (PTG_ERR_UNSUPPORTED for a resident graph). */
PTG_API int ptg_graph_stream_info(const ptg_graph *g, uint32_t *chunks, uint32_t *slots, uint64_t *slot_bytes);
/* Pinned host -> device copy rate in GB/s (`reps` copies of `bytes`). */
PTG_API int ptg_host_link_probe(uint64_t bytes, int reps, double *gbs);

/* ---- graph ---------------------------------------------------------------- */
/* Checks the reference-layout CSR on the host: offsets non-decreasing,
 * nodePointers[V-1] <= E, every column < V, E <= UINT32_MAX. No GPU needed. */
PTG_API int ptg_validate_csr(uint32_t V, uint64_t E, const uint32_t *nodePointers,
                     const uint32_t *edgeList, const EdgeWithWeight *wedges);

/* Replaces the per-call setup of every ptGraph driver (degree pass
 * main.cu:257-264, H2D copies main.cu:1712-1714 / :1272-1275). Exactly one of
 * edgeList / wedges may be NULL; with wedges the graph carries weights. */
PTG_API int ptg_graph_create(uint32_t V, uint64_t E, const uint32_t *nodePointers,
                     const uint32_t *edgeList, const EdgeWithWeight *wedges,
                     const ptg_config *cfg, ptg_graph **out);
/* Same, from separate column and weight arrays (weights may be NULL). */
PTG_API int ptg_graph_create_soa(uint32_t V, uint64_t E, const uint32_t *nodePointers,
                         const uint32_t *cols, const uint32_t *weights,
                         const ptg_config *cfg, ptg_graph **out);
PTG_API void ptg_graph_destroy(ptg_graph *g);
/* Device bytes ptg_graph_create would allocate for this size (CSR + query
 * state); graph_create fails with PTG_ERR_BUDGET before allocating anything
 * when this exceeds cfg->device_budget_bytes. */
PTG_API uint64_t ptg_graph_estimate_bytes(uint32_t V, uint64_t E, int weighted, const ptg_config *cfg);
PTG_API uint64_t ptg_graph_device_bytes(const ptg_graph *g);
PTG_API int ptg_graph_is_symmetric(ptg_graph *g); /* 1, 0, or negative error */
/* Frontier size of every round of the last ptg_bfs / ptg_sssp / ptg_cc (label
 * propagation) on this graph, the per-iteration counterpart of the reference's
 * `nodeSum` (ptGraph/main.cu:1109-1113): copies up to `cap` into `sizes`,
 * sets *count to the number of rounds. */
PTG_API int ptg_graph_frontier_sizes(const ptg_graph *g, uint64_t *sizes, uint32_t cap, uint32_t *count);
/* Changes the graph's runtime flags. Only PTG_FLAG_TIMING may differ from the
 * flags the graph was created with (PTG_ERR_INVALID_ARG otherwise). No
 * reference counterpart: the reference times whole calls with the host clock. */
PTG_API int ptg_graph_set_flags(ptg_graph *g, uint32_t flags);
/* Replaces the graph's tuning: td_to_bu, bu_to_td, sssp_delta,
 * sssp_defer_degree, sssp_defer_window, batch_delta, batch_grid, batch_sssp_grid,
 * batch_td_to_bu, batch_chunk, batch_pull, pr_block, pr_layout, stream_pinned, PTG_FLAG_TIMING and, for a
 * streamed graph, chunk_bytes (the chunks are cut again and the slot cache
 * resized; PTG_ERR_BUDGET, with the previous size kept, when fewer than two
 * slots would fit). Every other field must equal the graph's
 * (PTG_ERR_INVALID_ARG otherwise): they shape allocations made at creation.
 * Results never depend on these values, only speed does. */
PTG_API int ptg_graph_set_config(ptg_graph *g, const ptg_config *cfg);
PTG_API int ptg_graph_get_config(const ptg_graph *g, ptg_config *cfg);

/* ---- queries ---------------------------------------------------------------- */
/* bfs_kernel / bfsCaculateIn* (gpu_kernels.cu:118-134, main.cu:1041-1123,
 * :1180-1418, :1420-1678, :1681-2059). */
PTG_API int ptg_bfs(ptg_graph *g, uint32_t src, uint32_t *levels_out, ptg_stats *st);
/* sssp_kernel / ssspCaculateIn* (gpu_kernels.cu:152-172, main.cu:563-638,
 * :678-889). Needs a weighted graph. */
PTG_API int ptg_sssp(ptg_graph *g, uint32_t src, uint32_t *dist_out, ptg_stats *st);
/* cc_kernel / ccCaculateIn* (gpu_kernels.cu:136-150, main.cu:961-1035, :222-440). */
PTG_API int ptg_cc(ptg_graph *g, uint32_t *labels_out, ptg_stats *st);
/* prCaculateInShare (main.cu:891-959) has no arithmetic in the reference; this
 * defines it: r0 = 1/V, r' = (1-d)/V + d * (sum_{u->v} r(u)/outdeg(u) + D/V),
 * D = rank mass of vertices with outdeg 0. FP32 ranks from exact 64-bit
 * fixed-point sums: the same bits for a resident or a streamed graph. */
PTG_API int ptg_pagerank(ptg_graph *g, float damping, int iterations, float *ranks_out,
                 ptg_stats *st);
/* ptg_pagerank that allocates nothing. No counterpart in the reference; it is
 * for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc). The
 * per-iteration buffers (2 x 4V bytes of contributions plus partials) live in
 * one fixed 64 MiB device array per GPU, so V is limited to about 8.3 M
 * (PTG_ERR_BUDGET beyond). ranks_out must be device memory
 * (PTG_ERR_INVALID_ARG otherwise); a streamed graph, or a graph that is not
 * symmetric and has not yet run ptg_pagerank with pr_layout PTG_PR_PULL
 * (the in-list pull builds its in-edges; the binned layout does not),
 * gives PTG_ERR_UNSUPPORTED. Per-launch timing is never recorded. Results are
 * bit-identical to ptg_pagerank's. Every call on one GPU shares that GPU's
 * array: calls that overlap on one GPU race. */
PTG_API int ptg_pagerank_noalloc(ptg_graph *g, float damping, int iterations, float *ranks_out,
                                 ptg_stats *st);

/* ---- batched multi-source queries ------------------------------------------ */
/* Up to 64 sources in one batch (no reference counterpart: the reference runs
 * one source per driver call, ptGraph/main.cu:537-556). Row k of the output
 * (n x V words, source-major) equals ptg_bfs / ptg_sssp from srcs[k]: same
 * values, same sentinels. The output may be host or device memory. Stats:
 * reached = reached (source, vertex) pairs, edges_traversed = sum over the
 * sources of m_r, iterations = levels / rounds run. */
PTG_API int ptg_bfs_batch(ptg_graph *g, const uint32_t *srcs, int n, uint32_t *levels_out, ptg_stats *st);
PTG_API int ptg_sssp_batch(ptg_graph *g, const uint32_t *srcs, int n, uint32_t *dist_out, ptg_stats *st);

/* ---- autotuner ------------------------------------------------------------- */
/* Times the graph's own queries from `srcs` (n >= 1; the first 64 form the
 * batch) over a small grid of tuning values, one knob at a time, keeps the
 * fastest device time for each and applies the result to the graph (also
 * returned in *best). `what` is a mask of PTG_TUNE_*; on a streamed graph
 * only PTG_TUNE_STREAM applies, on a resident one every other bit
 * (PTG_ERR_UNSUPPORTED otherwise). Results never depend on the tuning, only
 * speed does. No reference counterpart (the reference hard-codes 56 x 1024
 * threads, its partition fraction and fragment size, main.cu:66, :318-319). */
#define PTG_TUNE_BFS 0x1u         /* td_to_bu, bu_to_td */
#define PTG_TUNE_SSSP 0x2u        /* sssp_delta, sssp_defer_window, sssp_defer_degree, sssp_pull,
                                     sssp_sweep (unless set below 0) */
#define PTG_TUNE_BFS_BATCH 0x4u   /* batch_td_to_bu, batch_grid */
#define PTG_TUNE_SSSP_BATCH 0x8u  /* batch_delta, batch_sssp_grid, batch_chunk, batch_pull */
#define PTG_TUNE_PAGERANK 0x20u   /* pr_layout, then pr_block for the pull (timed on 3 iterations) */
#define PTG_TUNE_STREAM 0x10u     /* streamed graphs only: chunk_bytes, stream_pinned, timed on BFS from
                                     the sources plus 2 PageRank iterations (symmetric graphs) */
PTG_API int ptg_autotune(ptg_graph *g, const uint32_t *srcs, int n, uint32_t what, ptg_config *best);

/* ---- serving ----------------------------------------------------------------- */
/* A request loop over resident graphs (no reference counterpart: the
 * reference harness runs one query per process, ptGraph/main.cu:78-86).
 * Each graph gets a worker thread on its own stream, so work on different
 * graphs overlaps on the device. Continuous batching: whenever a graph's slot
 * is free, its pending BFS (or SSSP) requests, up to max_batch, run as one
 * batched launch; identical CC / PageRank requests on one graph share one
 * run. Admission by device memory: the query state a launch needs is charged
 * against state_budget_bytes; idle graphs' state is released
 * (least recently used first) to make room, else the launch waits for running
 * ones; requests that can never fit complete with PTG_ERR_REJECTED.
 * Every result equals the single query's (same values; checksum below). */
typedef struct ptg_server ptg_server;

typedef struct ptg_server_config {
    uint64_t state_budget_bytes; /* device bytes for lazily allocated query state; 0 = unlimited */
    uint32_t max_batch;          /* sources per BFS/SSSP launch, 1..64; 1 = plain per-query path */
    uint32_t max_inflight;       /* graphs executing at once; 0 = all */
    uint32_t coalesce;           /* share one run among identical CC / PageRank requests */
    uint32_t queue_cap;          /* pending requests per graph beyond which submit rejects; 0 = 1<<20 */
    uint32_t min_batch;          /* fewer pending BFS/SSSP requests than this run one at a time; 0 = 8 */
} ptg_server_config;

enum { PTG_Q_BFS = 0, PTG_Q_SSSP = 1, PTG_Q_CC = 2, PTG_Q_PAGERANK = 3 };

typedef struct ptg_request {
    uint32_t graph;      /* id from ptg_server_add_graph */
    uint32_t kind;       /* PTG_Q_* */
    uint32_t source;     /* BFS / SSSP */
    uint32_t iterations; /* PageRank */
    float damping;       /* PageRank */
    uint32_t pad;
    void *out;           /* V words, host or device memory; NULL = keep only the checksum */
    uint64_t tag;        /* returned in the completion */
} ptg_request;

typedef struct ptg_completion {
    uint64_t tag;
    int32_t status;          /* PTG_OK or a negative code */
    uint32_t batch_size;     /* requests served by the same launch */
    uint64_t checksum;       /* sum over v of out[v] * (2v + 1) mod 2^64 (result words, caller ids) */
    uint64_t submit_ns;      /* steady clock: submitted, launch started, result delivered */
    uint64_t start_ns;
    uint64_t done_ns;
} ptg_completion;

typedef struct ptg_server_summary {
    uint64_t completed, rejected, launches;
    uint64_t p50_ns, p99_ns, max_ns; /* latency submit -> done over completed requests */
    double mean_batch;               /* requests per launch */
    uint64_t first_submit_ns, last_done_ns;
    uint64_t state_bytes_peak;       /* most query-state bytes resident at once */
    uint64_t evictions;              /* graphs whose state was released to admit another */
} ptg_server_summary;

PTG_API int ptg_server_create(const ptg_server_config *cfg, ptg_server **out);
/* The graph must outlive the server and not be used elsewhere meanwhile. */
PTG_API int ptg_server_add_graph(ptg_server *s, ptg_graph *g, uint32_t *id);
PTG_API int ptg_server_submit(ptg_server *s, const ptg_request *r);
/* Submits r[i] at first_call + arrival_ns[i] (steady clock, open loop) from
 * a thread of its own and returns once all are submitted. */
PTG_API int ptg_server_replay(ptg_server *s, const ptg_request *r, const uint64_t *arrival_ns, uint64_t n);
/* Copies up to `max` completions; waits up to timeout_ms for the first.
 * Returns the number copied or a negative code. */
PTG_API int ptg_server_poll(ptg_server *s, ptg_completion *out, int max, int timeout_ms);
PTG_API int ptg_server_drain(ptg_server *s); /* waits until every submitted request completed */
/* The first failed launch's message ("" if none); valid until destroy. */
PTG_API const char *ptg_server_error(ptg_server *s);
/* paused != 0: no new launch starts (running ones finish; submits queue up). */
PTG_API int ptg_server_pause(ptg_server *s, int paused);
PTG_API int ptg_server_summary_get(ptg_server *s, ptg_server_summary *out);
PTG_API void ptg_server_destroy(ptg_server *s);

/* ---- C-ABI entry points named after the reference drivers ------------------ */
/* Same signature and meaning as ptGraph/main.cu:24-56: build the graph, run one
 * query, free everything, return elapsed milliseconds (or a negative code). */
PTG_API long bfsCaculateInShare(uint32_t testNumNodes, uint32_t testNumEdge, uint32_t *nodePointersI,
                        uint32_t *edgeList, uint32_t sourceNode);
PTG_API long bfsCaculateInOpt(uint32_t testNumNodes, uint32_t testNumEdge, uint32_t *nodePointersI,
                      uint32_t *edgeList, uint32_t sourceNode);
PTG_API long bfsCaculateInAsync(uint32_t testNumNodes, uint32_t testNumEdge, uint32_t *nodePointersI,
                        uint32_t *edgeList, uint32_t sourceNode);
PTG_API long bfsCaculateInAsyncSwap(uint32_t testNumNodes, uint32_t testNumEdge, uint32_t *nodePointersI,
                            uint32_t *edgeList, uint32_t sourceNode);
PTG_API long ssspCaculateInShare(uint32_t testNumNodes, uint32_t testNumEdge, uint32_t *nodePointersI,
                         EdgeWithWeight *edgeList, uint32_t sourceNode);
PTG_API long ssspCaculateInOpt(uint32_t testNumNodes, uint32_t testNumEdge, uint32_t *nodePointersI,
                       EdgeWithWeight *edgeList, uint32_t sourceNode);
PTG_API long ccCaculateInShare(uint32_t testNumNodes, uint32_t testNumEdge, uint32_t *nodePointersI,
                       uint32_t *edgeList);
PTG_API long ccCaculateInOpt(uint32_t testNumNodes, uint32_t testNumEdge, uint32_t *nodePointersI,
                     uint32_t *edgeList);
PTG_API long prCaculateInShare(uint32_t testNumNodes, uint32_t testNumEdge, uint32_t *nodePointersI,
                       uint32_t *edgeList);
PTG_API long prCaculateInOpt(uint32_t testNumNodes, uint32_t testNumEdge, uint32_t *nodePointersI,
                     uint32_t *edgeList);

/* ---- synthetic inputs and files (host only, no GPU) ------------------------ */
/* Undirected R-MAT: edge_factor * 2^scale samples from a counter-based
 * SplitMix64 stream, self-loops dropped, symmetrised, duplicates removed,
 * adjacency sorted ascending. No vertex permutation (vertex 0 is the hub). */
PTG_API int ptg_gen_rmat(uint32_t scale, uint32_t edge_factor, double a, double b, double c,
                 uint64_t seed, int threads, ptg_host_csr *out);
/* Erdos-Renyi with the same sample count, cleaned the same way. */
PTG_API int ptg_gen_er(uint32_t scale, uint32_t edge_factor, uint64_t seed, int threads,
               ptg_host_csr *out);
/* Symmetric weights w(u,v) = w(v,u) uniform in [lo, hi] from a hash of
 * (min(u,v), max(u,v), seed); allocates csr->weights. */
PTG_API int ptg_gen_weights(ptg_host_csr *csr, uint32_t lo, uint32_t hi, uint64_t seed, int threads);
PTG_API void ptg_host_csr_free(ptg_host_csr *csr);

/* .bcsr  = [u32 V][u32 E][u32 nodePointers x V][u32 col x E]      (main.cu:98-116, :2604-2614)
 * .bwcsr = [u32 V][u32 E][u32 nodePointers x V][EdgeWithWeight x E] (main.cu:452-471, :2575-2602)
 * Java/WebGraph export = big-endian [u32 V][u64 E][u64 ptr x V][u32 col x E] (main.cu:2479-2573). */
PTG_API int ptg_read_bcsr(const char *path, ptg_host_csr *out);
PTG_API int ptg_write_bcsr(const char *path, const ptg_host_csr *csr);
PTG_API int ptg_read_bwcsr(const char *path, ptg_host_csr *out);
PTG_API int ptg_write_bwcsr(const char *path, const ptg_host_csr *csr);
PTG_API int ptg_read_java_be(const char *path, ptg_host_csr *out);

#ifdef __cplusplus
}
#endif
#endif /* PTG_H */
