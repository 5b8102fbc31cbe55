/*
 * oracle.c -- CPU oracle for the graph hot path (BFS / SSSP / CC / PageRank).
 *
 * TEST INFRASTRUCTURE ONLY. Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load this library, and only to check
 * results against or as the timed CPU baseline; the product path never
 * calls it.
 *
 * Graphs use the reference layout: `ptr` has V entries and the degree of the
 * last vertex is E - ptr[V-1] (ptGraph/main.cu:257-264).
 *
 * Three independent families:
 *   1. orc_ref_loop: a restatement of the reference's frontier loop
 *      (label -> count -> scan -> compact -> clear -> relax), strict `<` relax
 *      with a CAS-based atomic min, its init conventions, T pthreads.
 *      Follows ptGraph/main.cu:1041-1123 (bfsCaculateInShare),
 *      :563-638 (ssspCaculateInShare), :961-1035 (ccCaculateInShare) and the
 *      kernels ptGraph/gpu_kernels.cu:3-13 (setLabeling), :58-66
 *      (setActiveNodeArray), :98-106 (setLabelDefault), :118-172
 *      (bfs_kernel / cc_kernel / sssp_kernel). It is also the timed CPU
 *      baseline.
 *   2. Textbook algorithms: queue BFS, binary-heap Dijkstra with uint64
 *      accumulation, reachability-min labels, union-find CC, FP64 PageRank.
 *   3. Brute force on tiny graphs: Floyd-Warshall and transitive closure.
 *   4. Full-size helpers, T pthreads: a shortest-path certificate checker
 *      (exact BFS levels / SSSP distances without recomputing them), FP64
 *      PageRank pulled over the out-lists of a symmetric graph, and the
 *      reference's host loops timed as written: the degree pass
 *      (ptGraph/main.cu:1703-1710) and the overload gather bfsDynamic with its
 *      thread fan-out (ptGraph/main.cu:1126-1156, :1915-1938).
 *
 * Pinning (tests/test_oracle.py): family 1 == family 2 == family 3 == scipy's
 * csgraph routines on random and hand-made graphs, closed forms on paths,
 * cycles, stars and complete graphs, and the committed golden fixtures.
 * PageRank (no arithmetic in the reference, ptGraph/main.cu:891-959) is pinned
 * against a dense numpy iteration and closed forms (uniform rank on regular
 * graphs, 1/V on edgeless graphs, total mass 1). Family 4: the certificate
 * accepts the textbook outputs and rejects planted mistakes (a level or
 * distance moved by one, a reached vertex dropped, an unreachable one
 * reached); the symmetric PageRank equals the push form to 1e-12; the degree
 * pass equals the closed form and the gather equals the concatenated lists.
 */
#define _GNU_SOURCE
#include <math.h>
#include <pthread.h>
#include <stdatomic.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#define ORC_BFS 0
#define ORC_SSSP 1
#define ORC_CC 2

static inline uint32_t deg_of(uint32_t V, uint64_t E, const uint32_t *ptr, uint32_t v) {
    /* ptGraph/main.cu:257-264: degree[i] = ptr[i+1]-ptr[i]; degree[V-1] = E-ptr[V-1] */
    return (v + 1 < V) ? ptr[v + 1] - ptr[v] : (uint32_t)(E - ptr[v]);
}

/* ------------------------------------------------------------------------- */
/* 1. Reference frontier loop                                                 */
/* ------------------------------------------------------------------------- */

typedef struct {
    int algo;
    uint32_t V;
    uint64_t E;
    const uint32_t *ptr, *col, *w;
    _Atomic uint32_t *value;
    _Atomic uint8_t *label;   /* bool label[V] (main.cu:1066) */
    uint32_t *labeling;       /* setLabeling output (uint 0/1) */
    uint32_t *prefix;         /* exclusive_scan(labeling) */
    uint32_t *active;         /* activeNodeList */
    uint64_t *part;           /* per-thread partial sums */
    uint32_t active_num;
    int nthreads;
    pthread_barrier_t bar;
    int stop;
    uint32_t iters;
    uint64_t node_sum;
} ref_ctx;

typedef struct {
    ref_ctx *c;
    int tid;
} ref_arg;

static inline void range_of(uint64_t n, int tid, int nt, uint64_t *b, uint64_t *e) {
    uint64_t chunk = (n + nt - 1) / nt;
    *b = chunk * tid < n ? chunk * tid : n;
    *e = *b + chunk < n ? *b + chunk : n;
}

static inline void atomic_min_u32(_Atomic uint32_t *p, uint32_t v) {
    /* atomicMin semantics (gpu_kernels.cu:128) as a CAS loop */
    uint32_t cur = atomic_load_explicit(p, memory_order_relaxed);
    while (v < cur &&
           !atomic_compare_exchange_weak_explicit(p, &cur, v, memory_order_relaxed,
                                                  memory_order_relaxed)) {
    }
}

/* setLabeling + thrust::reduce: labeling[v] = label[v] ? 1 : 0; returns the count */
static void phase_label(ref_ctx *c, int tid) {
    uint64_t b, e, s = 0;
    range_of(c->V, tid, c->nthreads, &b, &e);
    for (uint64_t v = b; v < e; v++) {
        uint32_t l = atomic_load_explicit(&c->label[v], memory_order_relaxed) ? 1u : 0u;
        c->labeling[v] = l;
        s += l;
    }
    c->part[tid] = s;
}

static void *ref_worker(void *argp) {
    ref_arg *a = (ref_arg *)argp;
    ref_ctx *c = a->c;
    int tid = a->tid, nt = c->nthreads;
    /* initial setLabeling + reduce (main.cu:1079-1083) */
    phase_label(c, tid);
    pthread_barrier_wait(&c->bar);
    if (tid == 0) {
        uint64_t s = 0;
        for (int t = 0; t < nt; t++) s += c->part[t];
        c->active_num = (uint32_t)s;
        c->node_sum = s;
        c->iters = 0;
        c->stop = (s == 0);
    }
    pthread_barrier_wait(&c->bar);
    while (!c->stop) {
        /* exclusive_scan over labeling (main.cu:1094): per-thread block sums,
         * then offsets (part[] still holds this iteration's counts) */
        uint64_t b, e;
        range_of(c->V, tid, nt, &b, &e);
        uint64_t off = 0;
        for (int t = 0; t < tid; t++) off += c->part[t];
        for (uint64_t v = b; v < e; v++) {
            c->prefix[v] = (uint32_t)off;
            off += c->labeling[v];
        }
        pthread_barrier_wait(&c->bar);
        /* setActiveNodeArray (gpu_kernels.cu:58-66) */
        for (uint64_t v = b; v < e; v++)
            if (atomic_load_explicit(&c->label[v], memory_order_relaxed)) c->active[c->prefix[v]] = (uint32_t)v;
        pthread_barrier_wait(&c->bar);
        /* setLabelDefault (gpu_kernels.cu:98-106) */
        uint64_t ab, ae;
        range_of(c->active_num, tid, nt, &ab, &ae);
        for (uint64_t i = ab; i < ae; i++)
            atomic_store_explicit(&c->label[c->active[i]], 0, memory_order_relaxed);
        pthread_barrier_wait(&c->bar);
        /* relax (gpu_kernels.cu:118-172): strict `<`, then atomicMin and label */
        for (uint64_t i = ab; i < ae; i++) {
            uint32_t id = c->active[i];
            uint32_t edgeIndex = c->ptr[id];
            uint32_t d = deg_of(c->V, c->E, c->ptr, id);
            uint32_t sourceValue = atomic_load_explicit(&c->value[id], memory_order_relaxed);
            for (uint32_t k = 0; k < d; k++) {
                uint32_t to = c->col[edgeIndex + k];
                uint32_t fin;
                if (c->algo == ORC_BFS)
                    fin = sourceValue + 1;
                else if (c->algo == ORC_SSSP)
                    fin = sourceValue + c->w[edgeIndex + k];
                else
                    fin = sourceValue;
                if (fin < atomic_load_explicit(&c->value[to], memory_order_relaxed)) {
                    atomic_min_u32(&c->value[to], fin);
                    atomic_store_explicit(&c->label[to], 1, memory_order_relaxed);
                }
            }
        }
        pthread_barrier_wait(&c->bar);
        /* setLabeling + reduce (main.cu:1102-1104) */
        phase_label(c, tid);
        pthread_barrier_wait(&c->bar);
        if (tid == 0) {
            uint64_t s = 0;
            for (int t = 0; t < nt; t++) s += c->part[t];
            c->iters++;
            c->active_num = (uint32_t)s;
            c->node_sum += s;
            c->stop = (s == 0);
        }
        pthread_barrier_wait(&c->bar);
    }
    return NULL;
}

/*
 * Runs the reference loop. Init conventions:
 *   BFS : value = 0xFFFFFFFF, value[src] = 1, label[src] (main.cu:1068-1076)
 *   SSSP: value = sentinel (0xFFFFFFFE in ssspCaculateInShare main.cu:584, or
 *         0xFFFFFFFF in ssspCaculateInOpt main.cu:742), value[src] = 1
 *   CC  : value[v] = v, every label set (main.cu:977-980)
 * Returns 0, or -1 on bad arguments / allocation failure.
 */
int orc_ref_loop(int algo, uint32_t V, uint64_t E, const uint32_t *ptr, const uint32_t *col,
                 const uint32_t *w, uint32_t src, uint32_t sentinel, int threads,
                 uint32_t *value_out, uint32_t *iters_out, uint64_t *node_sum_out) {
    if (V == 0 || !ptr || (E && !col) || !value_out) return -1;
    if (algo == ORC_SSSP && E && !w) return -1;
    if (algo != ORC_CC && src >= V) return -1;
    if (threads < 1) threads = 1;
    ref_ctx c;
    memset(&c, 0, sizeof c);
    c.algo = algo;
    c.V = V;
    c.E = E;
    c.ptr = ptr;
    c.col = col;
    c.w = w;
    c.nthreads = threads;
    c.value = (_Atomic uint32_t *)malloc(sizeof(uint32_t) * V);
    c.label = (_Atomic uint8_t *)malloc(V);
    c.labeling = (uint32_t *)malloc(sizeof(uint32_t) * V);
    c.prefix = (uint32_t *)malloc(sizeof(uint32_t) * V);
    c.active = (uint32_t *)malloc(sizeof(uint32_t) * V);
    c.part = (uint64_t *)calloc(threads, sizeof(uint64_t));
    if (!c.value || !c.label || !c.labeling || !c.prefix || !c.active || !c.part) goto fail;
    for (uint32_t v = 0; v < V; v++) {
        uint32_t init = algo == ORC_BFS ? 0xFFFFFFFFu : algo == ORC_SSSP ? sentinel : v;
        atomic_init(&c.value[v], init);
        atomic_init(&c.label[v], (uint8_t)(algo == ORC_CC));
    }
    if (algo != ORC_CC) {
        atomic_store(&c.value[src], 1u);
        atomic_store(&c.label[src], 1);
    }
    pthread_barrier_init(&c.bar, NULL, (unsigned)threads);
    pthread_t *th = (pthread_t *)malloc(sizeof(pthread_t) * threads);
    ref_arg *args = (ref_arg *)malloc(sizeof(ref_arg) * threads);
    for (int t = 0; t < threads; t++) {
        args[t].c = &c;
        args[t].tid = t;
        if (t) pthread_create(&th[t], NULL, ref_worker, &args[t]);
    }
    ref_worker(&args[0]);
    for (int t = 1; t < threads; t++) pthread_join(th[t], NULL);
    pthread_barrier_destroy(&c.bar);
    free(th);
    free(args);
    for (uint32_t v = 0; v < V; v++) value_out[v] = atomic_load(&c.value[v]);
    if (iters_out) *iters_out = c.iters;
    if (node_sum_out) *node_sum_out = c.node_sum;
    free((void *)c.value);
    free((void *)c.label);
    free(c.labeling);
    free(c.prefix);
    free(c.active);
    free(c.part);
    return 0;
fail:
    free((void *)c.value);
    free((void *)c.label);
    free(c.labeling);
    free(c.prefix);
    free(c.active);
    free(c.part);
    return -1;
}

/* Wall-clock of one orc_ref_loop call in ns (graph setup excluded, like the
 * reference's startProcessing placement main.cu:1084). */
int64_t orc_ref_loop_timed(int algo, uint32_t V, uint64_t E, const uint32_t *ptr,
                           const uint32_t *col, const uint32_t *w, uint32_t src,
                           uint32_t sentinel, int threads, uint32_t *value_out,
                           uint32_t *iters_out, uint64_t *node_sum_out) {
    struct timespec t0, t1;
    clock_gettime(CLOCK_MONOTONIC, &t0);
    int rc = orc_ref_loop(algo, V, E, ptr, col, w, src, sentinel, threads, value_out, iters_out,
                          node_sum_out);
    clock_gettime(CLOCK_MONOTONIC, &t1);
    if (rc) return -1;
    return (int64_t)(t1.tv_sec - t0.tv_sec) * 1000000000LL + (t1.tv_nsec - t0.tv_nsec);
}

/* ------------------------------------------------------------------------- */
/* 2. Textbook algorithms                                                     */
/* ------------------------------------------------------------------------- */

/* Queue BFS; output in the reference convention (src = 1, unreached 0xFFFFFFFF).
 * Returns eccentricity + 1 (number of levels) via *levels_out if non-null. */
int orc_bfs_queue(uint32_t V, uint64_t E, const uint32_t *ptr, const uint32_t *col, uint32_t src,
                  uint32_t *out, uint32_t *levels_out, uint32_t *reached_out,
                  uint64_t *edges_out) {
    if (V == 0 || src >= V || !out) return -1;
    uint32_t *q = (uint32_t *)malloc(sizeof(uint32_t) * V);
    if (!q) return -1;
    for (uint32_t v = 0; v < V; v++) out[v] = 0xFFFFFFFFu;
    uint64_t head = 0, tail = 0, edges = 0;
    uint32_t maxl = 1;
    out[src] = 1;
    q[tail++] = src;
    while (head < tail) {
        uint32_t u = q[head++];
        uint32_t d = deg_of(V, E, ptr, u);
        edges += d;
        if (out[u] > maxl) maxl = out[u];
        for (uint32_t k = 0; k < d; k++) {
            uint32_t x = col[ptr[u] + k];
            if (out[x] == 0xFFFFFFFFu) {
                out[x] = out[u] + 1;
                q[tail++] = x;
            }
        }
    }
    if (levels_out) *levels_out = maxl;
    if (reached_out) *reached_out = (uint32_t)tail;
    if (edges_out) *edges_out = edges;
    free(q);
    return 0;
}

/* Binary min-heap on (dist, vertex) with lazy deletion. */
typedef struct {
    uint64_t d;
    uint32_t v;
} hitem;

static void heap_push(hitem *h, uint64_t *n, hitem x) {
    uint64_t i = (*n)++;
    while (i) {
        uint64_t p = (i - 1) / 2;
        if (h[p].d <= x.d) break;
        h[i] = h[p];
        i = p;
    }
    h[i] = x;
}

static hitem heap_pop(hitem *h, uint64_t *n) {
    hitem top = h[0], last = h[--(*n)];
    uint64_t i = 0;
    for (;;) {
        uint64_t l = 2 * i + 1, r = l + 1, m = i;
        uint64_t md = last.d;
        if (l < *n && h[l].d < md) {
            m = l;
            md = h[l].d;
        }
        if (r < *n && h[r].d < md) m = r;
        if (m == i) break;
        h[i] = h[m];
        i = m;
    }
    h[i] = last;
    return top;
}

/* Dijkstra with uint64 distances; output = 1 + dist, unreached = sentinel.
 * Returns 1 if some finite 1 + dist does not fit below the sentinel. */
int orc_dijkstra(uint32_t V, uint64_t E, const uint32_t *ptr, const uint32_t *col,
                 const uint32_t *w, uint32_t src, uint32_t sentinel, uint32_t *out) {
    if (V == 0 || src >= V || !out || (E && !w)) return -1;
    uint64_t *dist = (uint64_t *)malloc(sizeof(uint64_t) * V);
    hitem *h = (hitem *)malloc(sizeof(hitem) * (E + 1));
    if (!dist || !h) {
        free(dist);
        free(h);
        return -1;
    }
    for (uint32_t v = 0; v < V; v++) dist[v] = UINT64_MAX;
    uint64_t n = 0;
    dist[src] = 0;
    heap_push(h, &n, (hitem){0, src});
    while (n) {
        hitem it = heap_pop(h, &n);
        if (it.d != dist[it.v]) continue;
        uint32_t d = deg_of(V, E, ptr, it.v);
        for (uint32_t k = 0; k < d; k++) {
            uint64_t e = (uint64_t)ptr[it.v] + k;
            uint64_t nd = it.d + w[e];
            if (nd < dist[col[e]]) {
                dist[col[e]] = nd;
                heap_push(h, &n, (hitem){nd, col[e]});
            }
        }
    }
    int overflow = 0;
    for (uint32_t v = 0; v < V; v++) {
        if (dist[v] == UINT64_MAX) {
            out[v] = sentinel;
        } else if (dist[v] + 1 >= (uint64_t)sentinel) {
            out[v] = sentinel;
            overflow = 1;
        } else {
            out[v] = (uint32_t)(dist[v] + 1);
        }
    }
    free(dist);
    free(h);
    return overflow;
}

/* label[v] = min u such that v is reachable from u (u = v included): the fixed
 * point of cc_kernel (gpu_kernels.cu:136-150) for any directed graph. Vertices
 * are visited in ascending id order; a BFS from u labels every unlabelled
 * vertex it reaches and stops at labelled ones (their whole reach already
 * carries a smaller label). */
int orc_cc_reach_min(uint32_t V, uint64_t E, const uint32_t *ptr, const uint32_t *col,
                     uint32_t *out) {
    if (V == 0 || !out) return -1;
    uint32_t *q = (uint32_t *)malloc(sizeof(uint32_t) * V);
    if (!q) return -1;
    for (uint32_t v = 0; v < V; v++) out[v] = 0xFFFFFFFFu;
    for (uint32_t s = 0; s < V; s++) {
        if (out[s] != 0xFFFFFFFFu) continue;
        uint64_t head = 0, tail = 0;
        out[s] = s;
        q[tail++] = s;
        while (head < tail) {
            uint32_t u = q[head++];
            uint32_t d = deg_of(V, E, ptr, u);
            for (uint32_t k = 0; k < d; k++) {
                uint32_t x = col[ptr[u] + k];
                if (out[x] == 0xFFFFFFFFu) {
                    out[x] = s;
                    q[tail++] = x;
                }
            }
        }
    }
    free(q);
    return 0;
}

static uint32_t uf_find(uint32_t *p, uint32_t x) {
    while (p[x] != x) {
        p[x] = p[p[x]];
        x = p[x];
    }
    return x;
}

/* Union-find over every edge taken as undirected; root = smallest id. */
int orc_cc_union_find(uint32_t V, uint64_t E, const uint32_t *ptr, const uint32_t *col,
                      uint32_t *out) {
    if (V == 0 || !out) return -1;
    for (uint32_t v = 0; v < V; v++) out[v] = v;
    for (uint32_t u = 0; u < V; u++) {
        uint32_t d = deg_of(V, E, ptr, u);
        for (uint32_t k = 0; k < d; k++) {
            uint32_t a = uf_find(out, u), b = uf_find(out, col[ptr[u] + k]);
            if (a == b) continue;
            if (a < b)
                out[b] = a;
            else
                out[a] = b;
        }
    }
    for (uint32_t v = 0; v < V; v++) out[v] = uf_find(out, v);
    return 0;
}

/* FP64 PageRank, push form over out-edges:
 * r0 = 1/V; r'(v) = (1-d)/V + d * (sum_{u->v} r(u)/outdeg(u) + D/V),
 * D = sum of r(u) over outdeg(u) == 0. */
int orc_pagerank_f64(uint32_t V, uint64_t E, const uint32_t *ptr, const uint32_t *col,
                     double damping, int iters, double *out) {
    if (V == 0 || !out || iters < 0) return -1;
    double *r = out;
    double *acc = (double *)malloc(sizeof(double) * V);
    if (!acc) return -1;
    for (uint32_t v = 0; v < V; v++) r[v] = 1.0 / V;
    for (int it = 0; it < iters; it++) {
        double dang = 0.0;
        memset(acc, 0, sizeof(double) * V);
        for (uint32_t u = 0; u < V; u++) {
            uint32_t d = deg_of(V, E, ptr, u);
            if (d == 0) {
                dang += r[u];
                continue;
            }
            double c = r[u] / d;
            for (uint32_t k = 0; k < d; k++) acc[col[ptr[u] + k]] += c;
        }
        double base = (1.0 - damping) / V + damping * dang / V;
        for (uint32_t v = 0; v < V; v++) r[v] = base + damping * acc[v];
    }
    free(acc);
    return 0;
}

/* ------------------------------------------------------------------------- */
/* 3. Brute force (tiny graphs, V <= 64)                                      */
/* ------------------------------------------------------------------------- */

/* All-pairs shortest paths by Floyd-Warshall (weights NULL = unit weights);
 * writes the row of `src` in the reference output convention. */
int orc_floyd_warshall(uint32_t V, uint64_t E, const uint32_t *ptr, const uint32_t *col,
                       const uint32_t *w, uint32_t src, uint32_t sentinel, uint32_t *out) {
    if (V == 0 || V > 64 || src >= V || !out) return -1;
    static const uint64_t INF = UINT64_MAX / 4;
    uint64_t D[64][64];
    for (uint32_t i = 0; i < V; i++)
        for (uint32_t j = 0; j < V; j++) D[i][j] = i == j ? 0 : INF;
    for (uint32_t u = 0; u < V; u++) {
        uint32_t d = deg_of(V, E, ptr, u);
        for (uint32_t k = 0; k < d; k++) {
            uint64_t e = (uint64_t)ptr[u] + k;
            uint64_t c = w ? w[e] : 1;
            if (c < D[u][col[e]]) D[u][col[e]] = c;
        }
    }
    for (uint32_t k = 0; k < V; k++)
        for (uint32_t i = 0; i < V; i++)
            for (uint32_t j = 0; j < V; j++)
                if (D[i][k] + D[k][j] < D[i][j]) D[i][j] = D[i][k] + D[k][j];
    for (uint32_t j = 0; j < V; j++) out[j] = D[src][j] >= INF ? sentinel : (uint32_t)(D[src][j] + 1);
    return 0;
}

/* Transitive closure; label[v] = min u with u ->* v (u = v included). */
int orc_closure_min(uint32_t V, uint64_t E, const uint32_t *ptr, const uint32_t *col,
                    uint32_t *out) {
    if (V == 0 || V > 64 || !out) return -1;
    uint64_t R[64];
    for (uint32_t i = 0; i < V; i++) R[i] = 1ull << i;
    for (uint32_t u = 0; u < V; u++) {
        uint32_t d = deg_of(V, E, ptr, u);
        for (uint32_t k = 0; k < d; k++) R[u] |= 1ull << col[ptr[u] + k];
    }
    for (uint32_t k = 0; k < V; k++)
        for (uint32_t i = 0; i < V; i++)
            if (R[i] >> k & 1) R[i] |= R[k];
    for (uint32_t v = 0; v < V; v++) {
        out[v] = v;
        for (uint32_t u = 0; u < v; u++)
            if (R[u] >> v & 1) {
                out[v] = u;
                break;
            }
    }
    return 0;
}

/* ------------------------------------------------------------------------- */
/* 4. Full-size checkers and the reference's host loops (T threads)           */
/* ------------------------------------------------------------------------- */

typedef struct {
    uint32_t V;
    uint64_t E;
    const uint32_t *ptr, *col, *w;
    const uint32_t *val;
    uint32_t unreached;
    _Atomic uint8_t *tight;
    _Atomic int bad;
    const double *contrib;
    double base, damping;
    double *next;
    int nthreads;
} mt_ctx;

typedef struct {
    mt_ctx *c;
    int tid;
} mt_arg;

static void run_threads(mt_ctx *c, void *(*fn)(void *)) {
    int T = c->nthreads < 1 ? 1 : c->nthreads;
    pthread_t *th = (pthread_t *)malloc(sizeof(pthread_t) * T);
    mt_arg *a = (mt_arg *)malloc(sizeof(mt_arg) * T);
    for (int t = 0; t < T; t++) {
        a[t].c = c;
        a[t].tid = t;
        if (t) pthread_create(&th[t], NULL, fn, &a[t]);
    }
    fn(&a[0]);
    for (int t = 1; t < T; t++) pthread_join(th[t], NULL);
    free(th);
    free(a);
}

/* Edge pass of the shortest-path certificate: for every edge (u, v) out of a
 * reached u, v is reached and val[v] <= val[u] + w; v gets a "tight" mark
 * when val[v] == val[u] + w. */
static void *cert_edges(void *p) {
    mt_arg *a = (mt_arg *)p;
    mt_ctx *c = a->c;
    uint64_t b, e;
    range_of(c->V, a->tid, c->nthreads, &b, &e);
    for (uint64_t u = b; u < e && !atomic_load_explicit(&c->bad, memory_order_relaxed); u++) {
        uint64_t vu = c->val[u];
        if (vu == c->unreached) continue;
        uint32_t d = deg_of(c->V, c->E, c->ptr, (uint32_t)u);
        for (uint32_t k = 0; k < d; k++) {
            uint64_t x = (uint64_t)c->ptr[u] + k;
            uint32_t v = c->col[x];
            uint64_t cand = vu + (c->w ? c->w[x] : 1u);
            uint64_t vv = c->val[v];
            if (vv == c->unreached || vv > cand) {
                atomic_store_explicit(&c->bad, 1, memory_order_relaxed);
                break;
            }
            if (vv == cand) atomic_store_explicit(&c->tight[v], 1, memory_order_relaxed);
        }
    }
    return NULL;
}

/*
 * Checks that `val` is exactly the BFS levels (w == NULL) or the SSSP
 * distances from `src` in the reference's output convention (val[src] = 1,
 * `unreached` for vertices src cannot reach), without computing them:
 *   (1) val[src] == 1;
 *   (2) every edge (u, v) out of a reached u has v reached and
 *       val[v] <= val[u] + w(u, v);
 *   (3) every reached v != src has an edge (u, v) with val[u] + w == val[v].
 * With every weight >= 1 these hold for the true distances and only for them:
 * (2) gives val <= true distance along every path from src, and from any
 * reached v the tight edges of (3) lead back, strictly decreasing, to src, so
 * val[v] is the length of a real path. Zero weights would let a tight cycle
 * stand in for a path, so they are refused.
 * Returns 0 if `val` passes, 1 if it fails, -1 on bad arguments (including a
 * zero weight).
 */
int orc_certify_paths(uint32_t V, uint64_t E, const uint32_t *ptr, const uint32_t *col,
                      const uint32_t *w, uint32_t src, uint32_t unreached, const uint32_t *val,
                      int threads) {
    if (V == 0 || src >= V || !val || !ptr || (E && !col)) return -1;
    if (w)
        for (uint64_t x = 0; x < E; x++)
            if (w[x] == 0) return -1;
    if (val[src] != 1) return 1;
    mt_ctx c;
    memset(&c, 0, sizeof c);
    c.V = V;
    c.E = E;
    c.ptr = ptr;
    c.col = col;
    c.w = w;
    c.val = val;
    c.unreached = unreached;
    c.nthreads = threads < 1 ? 1 : threads;
    c.tight = (_Atomic uint8_t *)calloc(V, 1);
    if (!c.tight) return -1;
    atomic_init(&c.bad, 0);
    run_threads(&c, cert_edges);
    int bad = atomic_load(&c.bad);
    for (uint32_t v = 0; v < V && !bad; v++)
        if (v != src && val[v] != unreached && !atomic_load_explicit(&c.tight[v], memory_order_relaxed))
            bad = 1;
    free((void *)c.tight);
    return bad;
}

static void *pr_pull(void *p) {
    mt_arg *a = (mt_arg *)p;
    mt_ctx *c = a->c;
    uint64_t b, e;
    range_of(c->V, a->tid, c->nthreads, &b, &e);
    for (uint64_t v = b; v < e; v++) {
        uint32_t d = deg_of(c->V, c->E, c->ptr, (uint32_t)v);
        double s = 0.0;
        for (uint32_t k = 0; k < d; k++) s += c->contrib[c->col[(uint64_t)c->ptr[v] + k]];
        c->next[v] = c->base + c->damping * s;
    }
    return NULL;
}

/* orc_pagerank_f64 for a graph that is symmetric as a multigraph (every pair
 * (u, v) occurs as often as (v, u)): the in-list of v is then its out-list, so
 * each vertex pulls r(u)/outdeg(u) over its own list, one thread per vertex
 * range. Same definition as orc_pagerank_f64; the caller vouches for symmetry. */
int orc_pagerank_f64_sym(uint32_t V, uint64_t E, const uint32_t *ptr, const uint32_t *col,
                         double damping, int iters, int threads, double *out) {
    if (V == 0 || !out || iters < 0) return -1;
    double *contrib = (double *)malloc(sizeof(double) * V);
    if (!contrib) return -1;
    mt_ctx c;
    memset(&c, 0, sizeof c);
    c.V = V;
    c.E = E;
    c.ptr = ptr;
    c.col = col;
    c.contrib = contrib;
    c.damping = damping;
    c.next = out;
    c.nthreads = threads < 1 ? 1 : threads;
    for (uint32_t v = 0; v < V; v++) out[v] = 1.0 / V;
    for (int it = 0; it < iters; it++) {
        double dang = 0.0;
        for (uint32_t u = 0; u < V; u++) {
            uint32_t d = deg_of(V, E, ptr, u);
            if (d == 0) {
                dang += out[u];
                contrib[u] = 0.0;
            } else {
                contrib[u] = out[u] / d;
            }
        }
        c.base = (1.0 - damping) / V + damping * dang / V;
        run_threads(&c, pr_pull);
    }
    free(contrib);
    return 0;
}

/* The reference's degree pass as written (ptGraph/main.cu:1703-1710):
 * degree[i] = ptr[i+1] - ptr[i] for i < V-1, stopping at an offset beyond E,
 * then degree[V-1] = E - ptr[V-1]. Returns the elapsed ns. */
int64_t orc_degree_pass(uint32_t V, uint64_t E, const uint32_t *ptr, uint32_t *degree) {
    struct timespec t0, t1;
    clock_gettime(CLOCK_MONOTONIC, &t0);
    for (uint32_t i = 0; i + 1 < V; i++) {
        if (ptr[i] > E) break;
        degree[i] = ptr[i + 1] - ptr[i];
    }
    degree[V - 1] = (uint32_t)(E - ptr[V - 1]);
    clock_gettime(CLOCK_MONOTONIC, &t1);
    return (int64_t)(t1.tv_sec - t0.tv_sec) * 1000000000LL + (t1.tv_nsec - t0.tv_nsec);
}

typedef struct {
    int tid, nthreads;
    uint32_t begin, n_active;
    const uint32_t *degree, *active_ptr, *ptr, *active, *col;
    uint32_t *gathered;
} gather_arg;

/* One worker of the overload gather, ptGraph/main.cu:1126-1156 (bfsDynamic):
 * chunk = ceil((n_active - begin) / T), worker t copies the adjacency of
 * active[t*chunk + begin .. min(+chunk, n_active)) to active_ptr[i]. */
static void *gather_worker(void *p) {
    gather_arg *a = (gather_arg *)p;
    /* float arithmetic as in the reference (exact below 2^24 vertices) */
    float todo = (float)(a->n_active - a->begin);
    uint32_t chunk = (uint32_t)ceilf(todo / (float)a->nthreads);
    uint64_t left = (uint64_t)a->tid * chunk + a->begin;
    uint64_t right = left + chunk < a->n_active ? left + chunk : a->n_active;
    for (uint64_t i = left; i < right; i++) {
        uint32_t v = a->active[i];
        uint32_t d = a->degree[v];
        uint64_t here = a->active_ptr[i], there = a->ptr[v];
        for (uint32_t j = 0; j < d; j++) a->gathered[here + j] = a->col[there + j];
    }
    return NULL;
}

/* The per-iteration host gather of the reference's swap path
 * (ptGraph/main.cu:1915-1938): T = 20 threads, or 1 below 50 overload
 * vertices, created and joined for this one gather; `active_ptr` holds the
 * exclusive prefix of the overload vertices' degrees. Returns elapsed ns
 * including thread creation and join, as the reference's cpu timer does. */
int64_t orc_overload_gather(int threads, uint32_t n_active, const uint32_t *degree,
                            const uint32_t *active_ptr, const uint32_t *ptr,
                            const uint32_t *active, const uint32_t *col, uint32_t *gathered) {
    struct timespec t0, t1;
    int T = threads < 1 ? 1 : threads;
    if (n_active < 50) T = 1;
    clock_gettime(CLOCK_MONOTONIC, &t0);
    pthread_t *th = (pthread_t *)malloc(sizeof(pthread_t) * T);
    gather_arg *a = (gather_arg *)malloc(sizeof(gather_arg) * T);
    for (int t = 0; t < T; t++) {
        a[t] = (gather_arg){t, T, 0, n_active, degree, active_ptr, ptr, active, col, gathered};
        pthread_create(&th[t], NULL, gather_worker, &a[t]);
    }
    for (int t = 0; t < T; t++) pthread_join(th[t], NULL);
    clock_gettime(CLOCK_MONOTONIC, &t1);
    free(th);
    free(a);
    return (int64_t)(t1.tv_sec - t0.tv_sec) * 1000000000LL + (t1.tv_nsec - t0.tv_nsec);
}
