// Explicit CSR-chunk residency: graphs whose edges do not fit the device
// budget (PTG_FLAG_STREAM).
//
// Replaces the reference's partition-swap path (bfsCaculateInAsyncSwap,
// ptGraph/main.cu:1681-2059: a static prefix of the edge list in device
// memory, the rest behind unified memory with per-iteration cudaMemAdvise
// toggling and a host-thread gather of the overloaded adjacency lists,
// main.cu:1126-1156). Here the columns stay in pinned host memory cut into
// vertex-aligned chunks of at most chunk_bytes; the device holds the vertex
// state and a cache of chunk slots sized by the budget. A search step lists
// the chunks its frontier touches, expands the resident ones first and
// streams the others with hipMemcpyAsync on a copy stream, each kernel
// waiting on its chunk's copy event and each copy on the event that frees
// its slot, so copies overlap with the expansion of earlier chunks.
//
// Vertices with more than kHubEdges edges are cut into pieces of kHubEdges
// (one wave each) so that no wave walks a hub's whole list; PageRank sums a
// hub's pieces in a fixed order afterwards. BFS, SSSP and directed CC run
// top-down frontier rounds (the reference's bfs_kernelDynamicSwap,
// sssp_kernelOpt and cc_kernelOpt relaxations, ptGraph/gpu_kernels.cu:325-418,
// with the same strict `<` then atomicMin rule); CC on a symmetric graph is
// one union-find pass over every chunk; PageRank pulls over the out-lists,
// which equal the in-lists on a symmetric graph. A weighted graph keeps its
// weights in a second pinned array; a chunk's weights are copied into the
// upper half of its slot only when a step needs them (SSSP).
#include <string.h>

#include <algorithm>
#include <thread>
#include <vector>

#include "device_util.h"
#include "fingerprint.h"
#include "relabel.h"
#include "union_find.h"

namespace ptg {

constexpr uint32_t kHubEdges = 4096;
constexpr uint32_t kStreamGrid = 1024;  // workgroups of the chunk kernels (fixed: partial slots)

struct ChunkDesc {
    uint64_t eb, ee;   // edge range
    uint32_t vb, vh;   // first vertex, first non-hub vertex
    uint32_t ve;       // one past the last vertex
    uint32_t pb, pe;   // hub pieces of the chunk
    uint32_t rb, re;   // hub runs (one per hub vertex) of the chunk
};

struct StreamState {
    std::vector<uint32_t> h_off; // V + 1 offsets (device order), kept to re-plan chunks
    uint32_t *h_col = nullptr;  // pinned host columns
    uint32_t *h_w = nullptr;    // pinned host weights (weighted graphs)
    uint64_t wofs = 0;          // words from a slot's start to its weights
    uint64_t slot_bytes = 0;
    std::vector<ChunkDesc> chunks;
    uint64_t max_chunk_edges = 0;
    // pieces (global edge sub-ranges of hub vertices) and runs (one per hub)
    uint32_t *d_pv = nullptr, *d_pe0 = nullptr, *d_pe1 = nullptr;
    uint32_t *d_rv = nullptr, *d_rpb = nullptr, *d_rpe = nullptr;
    uint32_t npieces = 0, nruns = 0;
    uint32_t split_rb = 0;      // runs [split_rb, nruns) belong to vertices larger than a chunk
    uint64_t zeros = 0;         // vertices without out-edges
    uint32_t *d_cvb = nullptr, *d_cve = nullptr, *d_need = nullptr, *h_need = nullptr;
    std::vector<uint32_t *> slot;
    std::vector<int> slot_chunk;
    std::vector<char> slot_w;   // the slot holds its chunk's weights too
    std::vector<uint64_t> slot_used;
    std::vector<int> chunk_slot;
    std::vector<hipEvent_t> slot_free, slot_ready;
    std::vector<hipEvent_t> tcopy;  // start/stop pairs of the copies of one query
    uint32_t tcopy_used = 0;
    hipStream_t copy = nullptr;
    uint32_t *d_bits[2] = {};
    unsigned long long *d_cnt = nullptr, *h_cnt = nullptr;  // 2 x kShards
    float *d_contrib[2] = {};
    unsigned long long *d_ppart = nullptr;  // per piece: Q1.63 sum
    unsigned long long *d_dpart = nullptr;  // 2 iterations x chunks x kStreamGrid: Q1.63 dangling mass
    float *d_base = nullptr;                // this iteration's base value
    uint64_t clock = 0;
    uint64_t loads = 0, hits = 0, bytes = 0;
};

// ---------------------------------------------------------------------------
// Kernels
// ---------------------------------------------------------------------------

// need[c] = 1 for every chunk holding a frontier vertex; cnt[0] += frontier size.
__global__ __launch_bounds__(kBlock) void k_stream_need(uint32_t nwords, const uint32_t *__restrict__ bits,
                                                        uint32_t nch, const uint32_t *__restrict__ cvb,
                                                        const uint32_t *__restrict__ cve, uint32_t *need,
                                                        unsigned long long *cnt) {
    uint64_t f = 0;
    for (uint32_t w = blockIdx.x * blockDim.x + threadIdx.x; w < nwords; w += gridDim.x * blockDim.x) {
        uint32_t m = bits[w];
        f += __popc(m);
        while (m) {
            const uint32_t v = w * 32 + __builtin_ctz(m);
            m &= m - 1;
            // chunks holding v (vb <= v < ve) form a run starting at the first
            // chunk with ve > v
            uint32_t lo = 0, hi = nch;
            while (lo < hi) {
                const uint32_t mid = (lo + hi) >> 1;
                if (cve[mid] <= v) lo = mid + 1; else hi = mid;
            }
            for (uint32_t c = lo; c < nch && cvb[c] <= v; c++) need[c] = 1u;
        }
    }
    f = wave_sum_u64(f);
    if (lane_id() == 0 && f) atomicAdd(&cnt[blockIdx.x % kShards], (unsigned long long)f);
}

struct ChunkArgs {
    uint64_t eb;
    uint32_t vh, ve;
    uint32_t pb, pe;
    const uint32_t *slot;   // columns eb.. of the chunk
    const uint32_t *wslot;  // weights eb.. of the chunk (steps that read them), else null
};

__device__ __forceinline__ void mark_next(uint32_t *nxt, uint32_t v, uint64_t &found) {
    const uint32_t bit = 1u << (v & 31);
    if (!(atomicOr(&nxt[v >> 5], bit) & bit)) found++;
}

// Frontier step ops: vertex(u) loads what u's edges need once; edge() applies
// the edge (u, v) of weight w and counts vertices newly put into the next
// frontier.
struct BfsStep {
    uint32_t *level, *nxt;
    uint32_t nl;  // level of the vertices this step reaches
    __device__ __forceinline__ uint32_t vertex(uint32_t) const { return 0u; }
    __device__ __forceinline__ void edge(uint32_t, uint32_t v, uint32_t, uint64_t &found) const {
        if (level[v] == 0xFFFFFFFFu) {
            level[v] = nl;
            mark_next(nxt, v, found);
        }
    }
};

// sssp_kernel / sssp_kernelOpt relaxation (gpu_kernels.cu:152-172, :392-418):
// new = value[u] + w in uint32; strict `<`, then atomicMin.
struct SsspStep {
    uint32_t *dist, *nxt;
    __device__ __forceinline__ uint32_t vertex(uint32_t u) const { return dist[u]; }
    __device__ __forceinline__ void edge(uint32_t du, uint32_t v, uint32_t w, uint64_t &found) const {
        const uint32_t nd = du + w;
        if (nd < dist[v] && nd < atomicMin(&dist[v], nd)) mark_next(nxt, v, found);
    }
};

// cc_kernel / cc_kernelOpt (gpu_kernels.cu:136-150, :370-390): the smaller
// label flows along the edge.
struct LabelStep {
    uint32_t *label, *nxt;
    __device__ __forceinline__ uint32_t vertex(uint32_t u) const { return label[u]; }
    __device__ __forceinline__ void edge(uint32_t lu, uint32_t v, uint32_t, uint64_t &found) const {
        if (lu < label[v] && lu < atomicMin(&label[v], lu)) mark_next(nxt, v, found);
    }
};

// Union-find hooking of every edge (symmetric graphs): one pass over all
// chunks gives the components.
struct HookStep {
    uint32_t *par;
    const uint32_t *key;
    __device__ __forceinline__ uint32_t vertex(uint32_t u) const { return u; }
    __device__ __forceinline__ void edge(uint32_t u, uint32_t v, uint32_t, uint64_t &) const { uf_unite(par, key, u, v); }
};

// One step over the frontier vertices of one resident chunk (`cur` bits;
// every vertex when cur is null): hub pieces a wave each, other vertices 64
// per wave step, lists above 32 edges walked by the whole wave.
template <class Op>
__global__ __launch_bounds__(kBlock) void k_stream_step(ChunkArgs c, const uint32_t *__restrict__ off,
                                                        const uint32_t *__restrict__ pv,
                                                        const uint32_t *__restrict__ pe0,
                                                        const uint32_t *__restrict__ pe1,
                                                        const uint32_t *__restrict__ cur, Op op,
                                                        unsigned long long *cnt) {
    const uint32_t lane = lane_id();
    const uint32_t gw = blockIdx.x * kWavesPerBlock + (threadIdx.x >> 6);
    const uint32_t nw = gridDim.x * kWavesPerBlock;
    uint64_t found = 0;
    auto weight = [&](uint64_t e) -> uint32_t { return c.wslot ? c.wslot[e - c.eb] : 1u; };
    for (uint32_t i = c.pb + gw; i < c.pe; i += nw) {
        const uint32_t v = pv[i];
        if (cur && !((cur[v >> 5] >> (v & 31)) & 1u)) continue;
        const uint32_t sv = op.vertex(v);
        const uint64_t e1 = pe1[i];
        for (uint64_t e = pe0[i] + lane; e < e1; e += 64) op.edge(sv, c.slot[e - c.eb], weight(e), found);
    }
    for (uint64_t base = (uint64_t)c.vh + (uint64_t)gw * 64; base < c.ve; base += (uint64_t)nw * 64) {
        const uint64_t v = base + lane;
        const bool act = v < c.ve && (!cur || ((cur[v >> 5] >> (v & 31)) & 1u));
        uint64_t b = 0, d = 0;
        uint32_t sv = 0;
        if (act) {
            b = off[v];
            d = off[v + 1] - b;
            sv = op.vertex((uint32_t)v);
        }
        uint64_t mm = ballot(act && d > 32);
        while (mm) {
            const uint32_t q = lowest_lane(mm);
            mm &= mm - 1;
            const uint64_t bq = read_lane64(b, q), dq = read_lane64(d, q);
            const uint32_t sq = read_lane(sv, q);
            for (uint64_t k = lane; k < dq; k += 64) op.edge(sq, c.slot[bq + k - c.eb], weight(bq + k), found);
        }
        if (act && d <= 32)
            for (uint64_t k = 0; k < d; k++) op.edge(sv, c.slot[b + k - c.eb], weight(b + k), found);
    }
    found = wave_sum_u64(found);
    if (lane == 0 && found) atomicAdd(&cnt[blockIdx.x % kShards], (unsigned long long)found);
}

// Every vertex in the frontier (bits of ids >= V stay clear).
__global__ __launch_bounds__(kBlock) void k_stream_all(uint32_t V, uint32_t *bits) {
    const uint32_t nwords = (V + 31) / 32;
    for (uint32_t w = blockIdx.x * blockDim.x + threadIdx.x; w < nwords; w += gridDim.x * blockDim.x) {
        const uint32_t left = V - w * 32;
        bits[w] = left >= 32 ? 0xFFFFFFFFu : ((1u << left) - 1u);
    }
}

struct PrStep {
    const uint32_t *off;
    const float *cin;
    float *cout;
    float *rank;       // last iteration only
    const float *base; // [0] = base of this iteration
    float damping;
};

// Sums are Q1.63 fixed point (fx63), added exactly, so chunking, piece
// boundaries and slot order change no bit of the result.
__device__ __forceinline__ void pr_finish(const PrStep &p, uint32_t v, unsigned long long sum, float base,
                                          unsigned long long &dang) {
    const float r = pr_rank(base, p.damping, sum);
    const uint32_t od = p.off[v + 1] - p.off[v];
    p.cout[v] = od ? r / (float)od : 0.0f;
    if (p.rank) p.rank[v] = r;
    if (!od) dang += fx63(r);
}

// PageRank over one resident chunk: hub pieces leave partial sums, every
// other vertex of the chunk is finished here; the workgroup's dangling mass
// goes to its fixed slot.
__global__ __launch_bounds__(kBlock) void k_stream_pr(ChunkArgs c, PrStep p, const uint32_t *__restrict__ pe0,
                                                      const uint32_t *__restrict__ pe1, unsigned long long *ppart,
                                                      unsigned long long *dpart) {
    __shared__ unsigned long long lds[kWavesPerBlock];
    const uint32_t lane = lane_id();
    const uint32_t gw = blockIdx.x * kWavesPerBlock + (threadIdx.x >> 6);
    const uint32_t nw = gridDim.x * kWavesPerBlock;
    const float base = p.base[0];
    for (uint32_t i = c.pb + gw; i < c.pe; i += nw) {
        unsigned long long a = 0;
        const uint64_t e1 = pe1[i];
        for (uint64_t e = pe0[i] + lane; e < e1; e += 64) a += fx63(p.cin[c.slot[e - c.eb]]);
        a = wave_sum_u64(a);
        if (lane == 0) ppart[i] = a;
    }
    unsigned long long dang = 0;
    for (uint64_t base_v = (uint64_t)c.vh + (uint64_t)gw * 64; base_v < c.ve; base_v += (uint64_t)nw * 64) {
        const uint64_t v = base_v + lane;
        const bool inb = v < c.ve;
        uint64_t b = 0, d = 0;
        if (inb) {
            b = p.off[v];
            d = p.off[v + 1] - b;
        }
        unsigned long long mine = 0;
        uint64_t mm = ballot(inb && d > 32);
        while (mm) {
            const uint32_t q = lowest_lane(mm);
            mm &= mm - 1;
            const uint64_t bq = read_lane64(b, q), dq = read_lane64(d, q);
            unsigned long long a = 0;
            for (uint64_t k = lane; k < dq; k += 64) a += fx63(p.cin[c.slot[bq + k - c.eb]]);
            a = wave_sum_u64(a);
            if (lane == q) mine = a;
        }
        if (inb && d <= 32)
            for (uint64_t k = 0; k < d; k++) mine += fx63(p.cin[c.slot[b + k - c.eb]]);
        if (inb) pr_finish(p, (uint32_t)v, mine, base, dang);
    }
    dang = wave_sum_u64(dang);
    if (lane == 0) lds[threadIdx.x >> 6] = dang;
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned long long t = 0;
        for (int w = 0; w < kWavesPerBlock; w++) t += lds[w];
        dpart[blockIdx.x] = t;
    }
}

// Finishes the hub vertices of one chunk: one wave per run of pieces.
__global__ __launch_bounds__(kBlock) void k_stream_pr_hubs(uint32_t rb, uint32_t re, const uint32_t *__restrict__ rv,
                                                           const uint32_t *__restrict__ rpb,
                                                           const uint32_t *__restrict__ rpe,
                                                           const unsigned long long *ppart, PrStep p) {
    const uint32_t lane = lane_id();
    const uint32_t gw = blockIdx.x * kWavesPerBlock + (threadIdx.x >> 6);
    const uint32_t nw = gridDim.x * kWavesPerBlock;
    const float base = p.base[0];
    for (uint32_t r = rb + gw; r < re; r += nw) {
        unsigned long long a = 0;
        for (uint32_t i = rpb[r] + lane; i < rpe[r]; i += 64) a += ppart[i];
        a = wave_sum_u64(a);
        if (lane == 0) {
            unsigned long long unused = 0;
            pr_finish(p, rv[r], a, base, unused);  // hubs have out-edges: never dangling
        }
    }
}

// base = (1-d)/V + d*D/V with D summed over every chunk's workgroup slots,
// or d0 (dangling vertices / V) for the first iteration. One workgroup.
__global__ __launch_bounds__(kBlock) void k_stream_pr_base(const unsigned long long *dpart, uint32_t nslots, bool first,
                                                           double d0, float damping, float inv_v, float *base) {
    __shared__ unsigned long long lds[kWavesPerBlock];
    unsigned long long t = 0;
    if (!first)
        for (uint32_t j = threadIdx.x; j < nslots; j += kBlock) t += dpart[j];
    t = wave_sum_u64(t);
    if (lane_id() == 0) lds[threadIdx.x >> 6] = t;
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned long long sum = 0;
        for (int w = 0; w < kWavesPerBlock; w++) sum += lds[w];
        base[0] = pr_base_value(damping, inv_v, first ? d0 : fx63_to_f64(sum));
    }
}

__global__ __launch_bounds__(kBlock) void k_stream_pr_init(uint32_t V, const uint32_t *__restrict__ off, float r0,
                                                           float *__restrict__ contrib) {
    for (uint32_t v = blockIdx.x * blockDim.x + threadIdx.x; v < V; v += gridDim.x * blockDim.x) {
        const uint32_t od = off[v + 1] - off[v];
        contrib[v] = od ? r0 / (float)od : 0.0f;
    }
}

// value[src] = 1 (the reference's source value) and src in the frontier.
__global__ void k_stream_seed(uint32_t *level, uint32_t *bits, uint32_t src) {
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        level[src] = 1u;
        bits[src >> 5] = 1u << (src & 31);
    }
}

// ---------------------------------------------------------------------------
// Host side
// ---------------------------------------------------------------------------

static uint64_t al256(uint64_t x) { return (x + 255) & ~255ull; }

int stream_plan(ptg_graph *g, uint64_t chunk_bytes);

// Cuts the CSR into vertex-aligned chunks of at most max_edges edges (a
// vertex with more edges than that gets chunks of its own) and the hub
// vertices (more than kHubEdges edges) into pieces.
static void plan_chunks(uint32_t V, const uint32_t *off, uint64_t max_edges, std::vector<ChunkDesc> *chunks,
                        std::vector<uint32_t> *pv, std::vector<uint32_t> *pe0, std::vector<uint32_t> *pe1,
                        std::vector<uint32_t> *rv, std::vector<uint32_t> *rpb, std::vector<uint32_t> *rpe,
                        std::vector<uint32_t> *sv, std::vector<uint32_t> *spb, std::vector<uint32_t> *spe) {
    auto deg = [&](uint32_t v) -> uint64_t { return (uint64_t)off[v + 1] - off[v]; };
    auto add_pieces = [&](uint32_t v, uint64_t e0, uint64_t e1) {
        for (uint64_t x = e0; x < e1; x += kHubEdges) {
            pv->push_back(v);
            pe0->push_back((uint32_t)x);
            pe1->push_back((uint32_t)std::min<uint64_t>(x + kHubEdges, e1));
        }
    };
    auto add_hub = [&](uint32_t v, uint64_t e0, uint64_t e1) {
        rv->push_back(v);
        rpb->push_back((uint32_t)pv->size());
        add_pieces(v, e0, e1);
        rpe->push_back((uint32_t)pv->size());
    };
    uint32_t v = 0;
    while (v < V) {
        ChunkDesc c{};
        if (deg(v) > max_edges) {
            // a vertex larger than a chunk: chunks of its own; its pieces
            // span them all and are summed once every chunk has run
            sv->push_back(v);
            spb->push_back((uint32_t)pv->size());
            for (uint64_t x = off[v]; x < off[v + 1]; x += max_edges) {
                ChunkDesc s{};
                s.eb = x;
                s.ee = std::min<uint64_t>(x + max_edges, off[v + 1]);
                s.vb = v;
                s.vh = v + 1;
                s.ve = v + 1;
                s.pb = (uint32_t)pv->size();
                add_pieces(v, s.eb, s.ee);
                s.pe = (uint32_t)pv->size();
                s.rb = s.re = (uint32_t)rv->size();
                chunks->push_back(s);
            }
            spe->push_back((uint32_t)pv->size());
            v++;
            continue;
        }
        c.eb = off[v];
        c.vb = v;
        c.pb = (uint32_t)pv->size();
        c.rb = (uint32_t)rv->size();
        while (v < V && deg(v) <= max_edges && (uint64_t)off[v + 1] - c.eb <= max_edges) v++;
        c.ve = v;
        c.ee = off[v];
        // hub vertices of the chunk: with degree ordering they lead the chunk;
        // otherwise the chunk keeps hubs as ordinary vertices up to the first non-hub
        uint32_t h = c.vb;
        while (h < c.ve && deg(h) > kHubEdges) {
            add_hub(h, off[h], off[h + 1]);
            h++;
        }
        c.vh = h;
        c.pe = (uint32_t)pv->size();
        c.re = (uint32_t)rv->size();
        chunks->push_back(c);
    }
}

// Symmetry of the host CSR as a multigraph: the multiset fingerprints of the
// pairs (u, v) and (v, u) agree (fingerprint.h), products taken per thread
// over a vertex range and then multiplied.
static bool host_symmetric(uint32_t V, const uint32_t *off, const uint32_t *col) {
    const int T = (int)std::max(1u, std::min(32u, std::thread::hardware_concurrency()));
    const FpKeys keys = fp_random_keys();
    std::vector<FpAcc> acc(T, FpAcc{{1, 1}, {1, 1}});
    std::vector<std::thread> th;
    for (int t = 0; t < T; t++)
        th.emplace_back([&, t] {
            const uint32_t b = (uint32_t)((uint64_t)V * t / T), e = (uint32_t)((uint64_t)V * (t + 1) / T);
            FpAcc a{{1, 1}, {1, 1}};
            for (uint32_t u = b; u < e; u++)
                for (uint32_t k = off[u]; k < off[u + 1]; k++) fp_edge(keys, a, u, col[k]);
            acc[t] = a;
        });
    for (auto &x : th) x.join();
    FpAcc all{{1, 1}, {1, 1}};
    for (const FpAcc &a : acc)
        for (int i = 0; i < 2; i++) {
            all.f[i] = fp_mul(all.f[i], a.f[i]);
            all.b[i] = fp_mul(all.b[i], a.b[i]);
        }
    return all.f[0] == all.b[0] && all.f[1] == all.b[1];
}

int create_streamed(ptg_graph *g, const uint32_t *ptr, const uint32_t *cols, const uint32_t *weights,
                    const EdgeWithWeight *wedges) {
    const uint32_t V = g->V;
    const uint64_t E = g->E;
    const ptg_config &c = g->cfg;
    int rc;
    StreamState *s = new StreamState();
    g->streamed = s;
    const bool weighted = g->weighted;
    // host CSR (relabeled by degree unless asked not to), columns and
    // weights pinned
    std::vector<uint32_t> off;
    Relabeled rl;
    const uint32_t *col = cols, *wt = weights;
    std::vector<uint32_t> tmp, tmpw;
    if (!(c.flags & PTG_FLAG_NO_RELABEL)) {
        relabel(V, E, ptr, cols, weighted ? weights : nullptr, cols ? nullptr : wedges, c.order_hub_degree, &rl);
        off.swap(rl.off);
        col = rl.col.data();
        wt = weighted ? rl.w.data() : nullptr;
        g->relabeled = true;
    } else {
        off.resize((size_t)V + 1);
        memcpy(off.data(), ptr, 4ull * V);
        off[V] = (uint32_t)E;
        if (!cols && wedges) {
            tmp.resize(E);
            if (weighted) tmpw.resize(E);
            for (uint64_t e = 0; e < E; e++) {
                tmp[e] = wedges[e].toNode;
                if (weighted) tmpw[e] = wedges[e].weight;
            }
            col = tmp.data();
            if (weighted) wt = tmpw.data();
        }
    }
    if (hipHostMalloc((void **)&s->h_col, 4 * std::max<uint64_t>(E, 1), hipHostMallocDefault) != hipSuccess)
        return fail(PTG_ERR_OUT_OF_MEMORY, "pinned host columns");
    if (E) memcpy(s->h_col, col, 4 * E);
    if (weighted) {
        if (hipHostMalloc((void **)&s->h_w, 4 * std::max<uint64_t>(E, 1), hipHostMallocDefault) != hipSuccess)
            return fail(PTG_ERR_OUT_OF_MEMORY, "pinned host weights");
        if (E) memcpy(s->h_w, wt, 4 * E);
    }
    tmpw.clear();
    tmpw.shrink_to_fit();
    rl.w.clear();
    rl.w.shrink_to_fit();
    g->symmetric = host_symmetric(V, off.data(), s->h_col) ? 1 : 0;
    tmp.clear();
    tmp.shrink_to_fit();
    rl.col.clear();
    rl.col.shrink_to_fit();
    s->h_off.swap(off);
    s->zeros = 0;
    for (uint32_t v = 0; v < V; v++) s->zeros += s->h_off[v + 1] == s->h_off[v];
    // device vertex state (independent of the chunk plan)
    const uint32_t nwords = (V + 31) / 32;
    if ((rc = galloc(g, (void **)&g->out.off, 4ull * (V + 1))) ||
        (rc = copy_sync(g, g->out.off, s->h_off.data(), 4ull * (V + 1), hipMemcpyHostToDevice)) ||
        (rc = galloc(g, (void **)&s->d_bits[0], 4ull * nwords)) ||
        (rc = galloc(g, (void **)&s->d_bits[1], 4ull * nwords)) ||
        (rc = galloc(g, (void **)&s->d_cnt, 16 * kShards)) || (rc = galloc(g, (void **)&g->d_val, 4ull * V)) ||
        (rc = galloc(g, (void **)&g->d_acc, 16 * kShards)) ||
        (rc = galloc(g, (void **)&s->d_contrib[0], 4ull * V)) ||
        (rc = galloc(g, (void **)&s->d_contrib[1], 4ull * V)) || (rc = galloc(g, (void **)&s->d_base, 16)))
        return rc;
    if (g->relabeled) {
        if ((rc = galloc(g, (void **)&g->d_perm, 4ull * V)) || (rc = galloc(g, (void **)&g->d_inv, 4ull * V)) ||
            (rc = galloc(g, (void **)&g->d_val2, 4ull * V)) ||
            (rc = copy_sync(g, g->d_perm, rl.perm.data(), 4ull * V, hipMemcpyHostToDevice)) ||
            (rc = copy_sync(g, g->d_inv, rl.inv.data(), 4ull * V, hipMemcpyHostToDevice)))
            return rc;
        g->h_perm.swap(rl.perm);
    }
    if (hipHostMalloc((void **)&s->h_cnt, 16 * kShards, hipHostMallocDefault) != hipSuccess)
        return fail(PTG_ERR_OUT_OF_MEMORY, "pinned counters");
    if (hipStreamCreateWithFlags(&s->copy, hipStreamNonBlocking) != hipSuccess)
        return fail(PTG_ERR_HIP, "copy stream");
    return stream_plan(g, c.chunk_bytes);
}

// Frees the chunk plan: its device tables, partials, slots and their events.
static void free_plan(ptg_graph *g) {
    StreamState *s = g->streamed;
    if (s->copy) (void)hipStreamSynchronize(s->copy);
    if (g->stream) (void)hipStreamSynchronize(g->stream);
    for (auto *&p : s->slot) {
        gfree(g, p);
        p = nullptr;
    }
    for (auto e : s->slot_free)
        if (e) (void)hipEventDestroy(e);
    for (auto e : s->slot_ready)
        if (e) (void)hipEventDestroy(e);
    s->slot.clear();
    s->slot_free.clear();
    s->slot_ready.clear();
    for (uint32_t **p : {&s->d_pv, &s->d_pe0, &s->d_pe1, &s->d_rv, &s->d_rpb, &s->d_rpe, &s->d_cvb, &s->d_cve,
                         &s->d_need}) {
        gfree(g, *p);
        *p = nullptr;
    }
    gfree(g, s->d_ppart);
    gfree(g, s->d_dpart);
    s->d_ppart = s->d_dpart = nullptr;
    if (s->h_need) (void)hipHostFree(s->h_need);
    s->h_need = nullptr;
    s->chunks.clear();
}

// Cuts the graph into chunks of at most chunk_bytes of columns and sizes the
// slot cache by the device budget (what the vertex state leaves, at least two
// slots, at most one per chunk; a weighted graph's slots hold the weights in
// their upper half).
int stream_plan(ptg_graph *g, uint64_t chunk_bytes) {
    StreamState *s = g->streamed;
    const uint32_t V = g->V;
    const ptg_config &c = g->cfg;
    int rc;
    free_plan(g);
    s->max_chunk_edges = std::max<uint64_t>(64, (chunk_bytes ? chunk_bytes : (256ull << 20)) / 4);
    std::vector<uint32_t> pv, pe0, pe1, rv, rpb, rpe, sv, spb, spe;
    plan_chunks(V, s->h_off.data(), s->max_chunk_edges, &s->chunks, &pv, &pe0, &pe1, &rv, &rpb, &rpe, &sv, &spb,
                &spe);
    // split-vertex runs go after the per-chunk runs
    s->split_rb = (uint32_t)rv.size();
    rv.insert(rv.end(), sv.begin(), sv.end());
    rpb.insert(rpb.end(), spb.begin(), spb.end());
    rpe.insert(rpe.end(), spe.begin(), spe.end());
    const uint32_t C = (uint32_t)s->chunks.size();
    s->npieces = (uint32_t)pv.size();
    s->nruns = (uint32_t)rv.size();
    uint64_t largest = 1;
    for (auto &ch : s->chunks) largest = std::max<uint64_t>(largest, ch.ee - ch.eb);
    auto upload = [&](uint32_t **d, const std::vector<uint32_t> &h) -> int {
        int r = galloc(g, (void **)d, 4 * std::max<size_t>(h.size(), 1));
        if (!r && !h.empty()) r = copy_sync(g, *d, h.data(), 4 * h.size(), hipMemcpyHostToDevice);
        return r;
    };
    std::vector<uint32_t> cvb(C), cve(C);
    for (uint32_t i = 0; i < C; i++) {
        cvb[i] = s->chunks[i].vb;
        cve[i] = s->chunks[i].ve;
    }
    if ((rc = upload(&s->d_pv, pv)) || (rc = upload(&s->d_pe0, pe0)) || (rc = upload(&s->d_pe1, pe1)) ||
        (rc = upload(&s->d_rv, rv)) || (rc = upload(&s->d_rpb, rpb)) || (rc = upload(&s->d_rpe, rpe)) ||
        (rc = upload(&s->d_cvb, cvb)) || (rc = upload(&s->d_cve, cve)) ||
        (rc = galloc(g, (void **)&s->d_need, 4 * std::max(C, 1u))) ||
        (rc = galloc(g, (void **)&s->d_ppart, 8ull * std::max(s->npieces, 1u))) ||
        (rc = galloc(g, (void **)&s->d_dpart, 8ull * 2 * std::max(C, 1u) * kStreamGrid)))
        return rc;
    if (hipHostMalloc((void **)&s->h_need, 4 * std::max(C, 1u), hipHostMallocDefault) != hipSuccess)
        return fail(PTG_ERR_OUT_OF_MEMORY, "pinned chunk marks");
    s->wofs = al256(4 * largest) / 4;
    const uint64_t slot_bytes = al256(4 * largest) * (s->h_w ? 2 : 1);
    s->slot_bytes = slot_bytes;
    uint64_t nslots = C;
    if (c.device_budget_bytes) {
        const uint64_t left = c.device_budget_bytes > g->device_bytes ? c.device_budget_bytes - g->device_bytes : 0;
        nslots = std::min<uint64_t>(C, left / slot_bytes);
    }
    if (nslots < std::min<uint64_t>(2, C))
        return fail(PTG_ERR_BUDGET, "device budget leaves room for fewer than two chunk slots");
    s->slot.assign(nslots, nullptr);
    s->slot_chunk.assign(nslots, -1);
    s->slot_w.assign(nslots, 0);
    s->slot_used.assign(nslots, 0);
    s->chunk_slot.assign(C, -1);
    s->slot_free.assign(nslots, nullptr);
    s->slot_ready.assign(nslots, nullptr);
    for (uint64_t i = 0; i < nslots; i++) {
        if ((rc = galloc(g, (void **)&s->slot[i], slot_bytes))) return rc;
        if (hipEventCreateWithFlags(&s->slot_free[i], hipEventDisableTiming) != hipSuccess ||
            hipEventCreateWithFlags(&s->slot_ready[i], hipEventDisableTiming) != hipSuccess)
            return fail(PTG_ERR_HIP, "hipEventCreate");
        PTG_HIP(hipEventRecord(s->slot_free[i], g->stream));
    }
    return PTG_OK;
}

void destroy_streamed(ptg_graph *g) {
    StreamState *s = g->streamed;
    if (!s) return;
    free_plan(g);
    if (s->copy) (void)hipStreamDestroy(s->copy);
    for (auto e : s->tcopy) (void)hipEventDestroy(e);
    for (auto *p : {s->d_bits[0], s->d_bits[1]}) gfree(g, p);
    gfree(g, s->d_cnt);
    gfree(g, s->d_contrib[0]);
    gfree(g, s->d_contrib[1]);
    gfree(g, s->d_base);
    if (s->h_col) (void)hipHostFree(s->h_col);
    if (s->h_w) (void)hipHostFree(s->h_w);
    if (s->h_cnt) (void)hipHostFree(s->h_cnt);
    delete s;
    g->streamed = nullptr;
}

// Copies `words` words from pinned host memory into a slot on the copy
// stream, timed, after the slot's previous readers are done; the graph's
// stream then waits for the copy.
static int slot_copy(ptg_graph *g, uint32_t sl, uint32_t *dst, const uint32_t *src, uint64_t words) {
    StreamState *s = g->streamed;
    const uint64_t bytes = 4 * words;
    if (!bytes) return PTG_OK;
    PTG_HIP(hipStreamWaitEvent(s->copy, s->slot_free[sl], 0));
    if (s->tcopy.size() < 2ull * (s->tcopy_used + 1)) {
        hipEvent_t a, b;
        PTG_HIP(hipEventCreate(&a));
        PTG_HIP(hipEventCreate(&b));
        s->tcopy.push_back(a);
        s->tcopy.push_back(b);
    }
    PTG_HIP(hipEventRecord(s->tcopy[2 * s->tcopy_used], s->copy));
    PTG_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, s->copy));
    PTG_HIP(hipEventRecord(s->tcopy[2 * s->tcopy_used + 1], s->copy));
    s->tcopy_used++;
    PTG_HIP(hipEventRecord(s->slot_ready[sl], s->copy));
    PTG_HIP(hipStreamWaitEvent(g->stream, s->slot_ready[sl], 0));
    s->bytes += bytes;
    return PTG_OK;
}

// Makes chunk `c` resident (with its weights when `need_w`) and returns its
// slot; the graph's stream waits for the copy. `pending[c']` marks chunks
// still to run in this step (kept if another slot can be freed).
static int acquire_chunk(ptg_graph *g, uint32_t c, const std::vector<char> &pending, bool need_w,
                         uint32_t *slot_out) {
    StreamState *s = g->streamed;
    const ChunkDesc &ch = s->chunks[c];
    int rc;
    s->clock++;
    if (s->chunk_slot[c] >= 0) {
        const uint32_t sl = (uint32_t)s->chunk_slot[c];
        s->slot_used[sl] = s->clock;
        if (need_w && !s->slot_w[sl]) {  // columns resident, weights not yet
            if ((rc = slot_copy(g, sl, s->slot[sl] + s->wofs, s->h_w + ch.eb, ch.ee - ch.eb))) return rc;
            s->slot_w[sl] = 1;
            s->loads++;
        } else {
            s->hits++;
        }
        *slot_out = sl;
        return PTG_OK;
    }
    // victim: an empty slot, else the least recently used one whose chunk
    // this step no longer needs, else any; chunks below `pinned` only when
    // nothing else is left
    const uint64_t nsl = s->slot.size();
    const uint32_t pinned = (uint32_t)std::min<uint64_t>(g->cfg.stream_pinned, nsl > 2 ? nsl - 2 : 0);
    int best = -1;
    uint64_t best_key = ~0ull;
    for (size_t i = 0; i < s->slot.size(); i++) {
        const int held = s->slot_chunk[i];
        uint64_t key = held < 0 ? 0 : (pending[held] ? (1ull << 62) + s->slot_used[i] : s->slot_used[i]);
        if (held >= 0 && (uint32_t)held < pinned) key |= 1ull << 63;
        if (key < best_key) {
            best_key = key;
            best = (int)i;
        }
    }
    const uint32_t sl = (uint32_t)best;
    if (s->slot_chunk[sl] >= 0) s->chunk_slot[s->slot_chunk[sl]] = -1;
    s->slot_chunk[sl] = (int)c;
    s->chunk_slot[c] = (int)sl;
    s->slot_used[sl] = s->clock;
    if ((rc = slot_copy(g, sl, s->slot[sl], s->h_col + ch.eb, ch.ee - ch.eb))) return rc;
    s->slot_w[sl] = 0;
    if (need_w) {
        if ((rc = slot_copy(g, sl, s->slot[sl] + s->wofs, s->h_w + ch.eb, ch.ee - ch.eb))) return rc;
        s->slot_w[sl] = 1;
    }
    s->loads++;
    *slot_out = sl;
    return PTG_OK;
}

static ChunkArgs chunk_args(const StreamState *s, uint32_t c, uint32_t sl, bool with_w = false) {
    const ChunkDesc &ch = s->chunks[c];
    ChunkArgs a;
    a.eb = ch.eb;
    a.vh = ch.vh;
    a.ve = ch.ve;
    a.pb = ch.pb;
    a.pe = ch.pe;
    a.slot = s->slot[sl];
    a.wslot = with_w ? s->slot[sl] + s->wofs : nullptr;
    return a;
}

// A graph whose last re-plan failed twice has no slots; its queries are refused.
static int check_plan(const ptg_graph *g) {
    const StreamState *s = g->streamed;
    if (s->slot.empty() || s->chunks.empty()) return fail(PTG_ERR_BUDGET, "streamed graph has no chunk plan");
    return PTG_OK;
}

static void begin_query(StreamState *s) {
    s->tcopy_used = 0;
    s->loads = s->hits = s->bytes = 0;
}

static int finish_query_stats(ptg_graph *g, ptg_stats *st) {
    StreamState *s = g->streamed;
    PTG_HIP(hipStreamSynchronize(s->copy));
    double ns = 0;
    for (uint32_t i = 0; i < s->tcopy_used; i++) {
        float ms = 0;
        PTG_HIP(hipEventElapsedTime(&ms, s->tcopy[2 * i], s->tcopy[2 * i + 1]));
        ns += (double)ms * 1e6;
    }
    st->host_link_bytes = s->bytes;
    st->host_link_ns = (uint64_t)ns;
    st->chunk_loads = (uint32_t)s->loads;
    st->chunk_hits = (uint32_t)s->hits;
    return PTG_OK;
}

static int sum_counter(ptg_graph *g, uint64_t *v) {
    StreamState *s = g->streamed;
    PTG_HIP(hipMemcpyAsync(s->h_cnt, s->d_cnt, 16 * kShards, hipMemcpyDeviceToHost, g->stream));
    PTG_HIP(hipStreamSynchronize(g->stream));
    uint64_t t = 0;
    for (uint32_t i = 0; i < 2 * kShards; i++) t += s->h_cnt[i];
    *v = t;
    return PTG_OK;
}

// Frontier rounds over the chunks, from the frontier in d_bits[0]: each round
// lists the chunks holding frontier vertices, expands the resident ones
// first, then streams and expands the others (`launch(args, cur, nxt, L)`
// enqueues one chunk's step on the graph's stream), and ends when a round
// puts no vertex into the next frontier.
template <class Launch>
static int stream_rounds(ptg_graph *g, bool need_w, Launch launch, uint32_t *iters_out, uint64_t *fsum_out) {
    StreamState *s = g->streamed;
    const uint32_t V = g->V;
    const uint32_t C = (uint32_t)s->chunks.size();
    const uint32_t nwords = (V + 31) / 32;
    int rc;
    uint32_t iters = 0;
    uint64_t fsum = 0;
    std::vector<char> pending(C, 0);
    for (uint32_t L = 0;; L++) {
        uint32_t *cur = s->d_bits[L & 1], *nxt = s->d_bits[(L + 1) & 1];
        // which chunks hold frontier vertices (and the frontier's size)
        if ((rc = fill_u32(g, s->d_need, C, 0u)) ||
            (rc = fill_u32(g, reinterpret_cast<uint32_t *>(s->d_cnt), 4 * kShards, 0u)))
            return rc;
        const uint32_t nb = std::min<uint32_t>((nwords + kBlock - 1) / kBlock, g->grid);
        k_stream_need<<<std::max(nb, 1u), kBlock, 0, g->stream>>>(nwords, cur, C, s->d_cvb, s->d_cve, s->d_need,
                                                                   s->d_cnt);
        PTG_HIP(hipGetLastError());
        PTG_HIP(hipMemcpyAsync(s->h_need, s->d_need, 4ull * C, hipMemcpyDeviceToHost, g->stream));
        uint64_t fsize = 0;
        if ((rc = sum_counter(g, &fsize))) return rc;  // also waits for h_need
        std::vector<uint32_t> order;
        for (uint32_t c = 0; c < C; c++)
            if (s->h_need[c] && s->chunk_slot[c] >= 0) order.push_back(c);
        for (uint32_t c = 0; c < C; c++)
            if (s->h_need[c] && s->chunk_slot[c] < 0) order.push_back(c);
        if (order.empty()) break;
        fsum += fsize;
        for (uint32_t c : order) pending[c] = 1;
        if ((rc = fill_u32(g, reinterpret_cast<uint32_t *>(s->d_cnt), 4 * kShards, 0u))) return rc;
        for (uint32_t c : order) {
            uint32_t sl = 0;
            if ((rc = acquire_chunk(g, c, pending, need_w, &sl))) return rc;
            pending[c] = 0;
            launch(chunk_args(s, c, sl, need_w), cur, nxt, L);
            PTG_HIP(hipGetLastError());
            PTG_HIP(hipEventRecord(s->slot_free[sl], g->stream));
        }
        uint64_t found = 0;
        if ((rc = sum_counter(g, &found))) return rc;
        iters++;
        if ((rc = fill_u32(g, cur, nwords, 0u))) return rc;
        if (!found) break;
    }
    *iters_out = iters;
    *fsum_out = fsum;
    return PTG_OK;
}

// Clears both frontier bitmaps and the query's statistics.
static int begin_rounds(ptg_graph *g) {
    StreamState *s = g->streamed;
    const uint32_t nwords = (g->V + 31) / 32;
    int rc;
    if ((rc = check_plan(g))) return rc;
    begin_query(s);
    if ((rc = fill_u32(g, s->d_bits[0], nwords, 0u)) || (rc = fill_u32(g, s->d_bits[1], nwords, 0u))) return rc;
    return PTG_OK;
}

static int end_query(ptg_graph *g, ptg_stats *st, uint32_t iters, uint64_t fsum) {
    PTG_HIP(hipEventRecord(g->ev1, g->stream));
    PTG_HIP(hipEventSynchronize(g->ev1));
    if (!st) return PTG_OK;
    memset(st, 0, sizeof *st);
    float ms = 0;
    PTG_HIP(hipEventElapsedTime(&ms, g->ev0, g->ev1));
    st->kernel_ns = (uint64_t)((double)ms * 1e6);
    st->iterations = iters;
    st->frontier_sum = fsum;
    return finish_query_stats(g, st);
}

int run_bfs_streamed(ptg_graph *g, uint32_t src, uint32_t *level, ptg_stats *st) {
    StreamState *s = g->streamed;
    int rc;
    PTG_HIP(hipEventRecord(g->ev0, g->stream));
    if ((rc = begin_rounds(g)) || (rc = fill_u32(g, level, g->V, 0xFFFFFFFFu))) return rc;
    k_stream_seed<<<1, 64, 0, g->stream>>>(level, s->d_bits[0], src);
    PTG_HIP(hipGetLastError());
    uint32_t iters = 0;
    uint64_t fsum = 0;
    rc = stream_rounds(g, false, [&](const ChunkArgs &a, const uint32_t *cur, uint32_t *nxt, uint32_t L) {
        k_stream_step<BfsStep><<<kStreamGrid, kBlock, 0, g->stream>>>(a, g->out.off, s->d_pv, s->d_pe0, s->d_pe1, cur,
                                                                     BfsStep{level, nxt, L + 2}, s->d_cnt);
    }, &iters, &fsum);
    if (rc || (rc = end_query(g, st, iters, fsum))) return rc;
    if (st) return reach_stats(g, level, 0xFFFFFFFFu, &st->reached, &st->edges_traversed);
    return PTG_OK;
}

int run_sssp_streamed(ptg_graph *g, uint32_t src, uint32_t *dist, ptg_stats *st) {
    StreamState *s = g->streamed;
    if (!s->h_w) return fail(PTG_ERR_UNSUPPORTED, "sssp: graph has no weights");
    const uint32_t sentinel = g->cfg.sssp_sentinel;
    int rc;
    PTG_HIP(hipEventRecord(g->ev0, g->stream));
    if ((rc = begin_rounds(g)) || (rc = fill_u32(g, dist, g->V, sentinel))) return rc;
    k_stream_seed<<<1, 64, 0, g->stream>>>(dist, s->d_bits[0], src);
    PTG_HIP(hipGetLastError());
    uint32_t iters = 0;
    uint64_t fsum = 0;
    rc = stream_rounds(g, true, [&](const ChunkArgs &a, const uint32_t *cur, uint32_t *nxt, uint32_t) {
        k_stream_step<SsspStep><<<kStreamGrid, kBlock, 0, g->stream>>>(a, g->out.off, s->d_pv, s->d_pe0, s->d_pe1, cur,
                                                                      SsspStep{dist, nxt}, s->d_cnt);
    }, &iters, &fsum);
    if (rc || (rc = end_query(g, st, iters, fsum))) return rc;
    if (st) return reach_stats(g, dist, sentinel, &st->reached, &st->edges_traversed);
    return PTG_OK;
}

int run_cc_streamed(ptg_graph *g, uint32_t *label, ptg_stats *st) {
    StreamState *s = g->streamed;
    const uint32_t C = (uint32_t)s->chunks.size();
    int rc;
    uint32_t iters = 0;
    uint64_t fsum = 0;
    PTG_HIP(hipEventRecord(g->ev0, g->stream));
    if ((rc = begin_rounds(g))) return rc;
    if (g->symmetric) {
        // union-find: one pass over every chunk, resident ones first
        if ((rc = uf_start(g, label))) return rc;
        std::vector<uint32_t> order;
        for (uint32_t c = 0; c < C; c++)
            if (s->chunk_slot[c] >= 0) order.push_back(c);
        for (uint32_t c = 0; c < C; c++)
            if (s->chunk_slot[c] < 0) order.push_back(c);
        std::vector<char> pending(C, 1);
        for (uint32_t c : order) {
            uint32_t sl = 0;
            if ((rc = acquire_chunk(g, c, pending, false, &sl))) return rc;
            pending[c] = 0;
            k_stream_step<HookStep><<<kStreamGrid, kBlock, 0, g->stream>>>(
                chunk_args(s, c, sl), g->out.off, s->d_pv, s->d_pe0, s->d_pe1, nullptr, HookStep{label, g->d_inv},
                s->d_cnt);
            PTG_HIP(hipGetLastError());
            PTG_HIP(hipEventRecord(s->slot_free[sl], g->stream));
        }
        if ((rc = uf_finish(g, label))) return rc;
        iters = 1;
        fsum = g->V;
    } else {
        // label propagation from the caller's ids, every vertex active first
        if ((rc = init_keys(g, label))) return rc;
        const uint32_t nwords = (g->V + 31) / 32;
        k_stream_all<<<std::max(1u, std::min<uint32_t>((nwords + kBlock - 1) / kBlock, g->grid)), kBlock, 0,
                       g->stream>>>(g->V, s->d_bits[0]);
        PTG_HIP(hipGetLastError());
        rc = stream_rounds(g, false, [&](const ChunkArgs &a, const uint32_t *cur, uint32_t *nxt, uint32_t) {
            k_stream_step<LabelStep><<<kStreamGrid, kBlock, 0, g->stream>>>(a, g->out.off, s->d_pv, s->d_pe0,
                                                                           s->d_pe1, cur, LabelStep{label, nxt},
                                                                           s->d_cnt);
        }, &iters, &fsum);
        if (rc) return rc;
    }
    if ((rc = end_query(g, st, iters, fsum))) return rc;
    if (st) {
        st->edges_traversed = g->E;
        return count_roots(g, label, &st->reached);
    }
    return PTG_OK;
}

int run_pagerank_streamed(ptg_graph *g, float damping, int iters, float *ranks, ptg_stats *st) {
    StreamState *s = g->streamed;
    if (!g->symmetric) return fail(PTG_ERR_UNSUPPORTED, "streamed PageRank needs a symmetric graph");
    if (int prc = check_plan(g)) return prc;
    const uint32_t V = g->V;
    const uint32_t C = (uint32_t)s->chunks.size();
    const float r0 = 1.0f / (float)V;
    int rc;
    begin_query(s);
    PTG_HIP(hipEventRecord(g->ev0, g->stream));
    if (iters == 0) {
        union {
            float f;
            uint32_t u;
        } cv;
        cv.f = r0;
        if ((rc = fill_u32(g, reinterpret_cast<uint32_t *>(ranks), V, cv.u))) return rc;
    } else {
        k_stream_pr_init<<<std::min<uint32_t>((V + kBlock - 1) / kBlock, g->grid), kBlock, 0, g->stream>>>(
            V, g->out.off, r0, s->d_contrib[0]);
        const double d0 = (double)s->zeros / (double)V;  // D0 = (vertices without out-edges) / V
        std::vector<char> pending(C, 0);
        const uint64_t slots_per_iter = (uint64_t)C * kStreamGrid;
        for (int t = 0; t < iters; t++) {
            unsigned long long *dprev = s->d_dpart + slots_per_iter * ((t + 1) & 1);
            unsigned long long *dcur = s->d_dpart + slots_per_iter * (t & 1);
            k_stream_pr_base<<<1, kBlock, 0, g->stream>>>(dprev, (uint32_t)slots_per_iter, t == 0, d0, damping, r0,
                                                         s->d_base);
            PrStep p{g->out.off, s->d_contrib[t & 1], s->d_contrib[(t + 1) & 1], t == iters - 1 ? ranks : nullptr,
                     s->d_base, damping};
            // alternate the order so the chunks left resident by one pass start the next
            for (uint32_t i = 0; i < C; i++) pending[i] = 1;
            for (uint32_t i = 0; i < C; i++) {
                const uint32_t c = (t & 1) ? C - 1 - i : i;
                uint32_t sl = 0;
                if ((rc = acquire_chunk(g, c, pending, false, &sl))) return rc;
                pending[c] = 0;
                k_stream_pr<<<kStreamGrid, kBlock, 0, g->stream>>>(chunk_args(s, c, sl), p, s->d_pe0, s->d_pe1,
                                                                   s->d_ppart, dcur + (uint64_t)c * kStreamGrid);
                const ChunkDesc &ch = s->chunks[c];
                if (ch.re > ch.rb)
                    k_stream_pr_hubs<<<std::min<uint32_t>((ch.re - ch.rb + 3) / 4, kStreamGrid), kBlock, 0,
                                       g->stream>>>(ch.rb, ch.re, s->d_rv, s->d_rpb, s->d_rpe, s->d_ppart, p);
                PTG_HIP(hipGetLastError());
                PTG_HIP(hipEventRecord(s->slot_free[sl], g->stream));
            }
            if (s->nruns > s->split_rb)
                k_stream_pr_hubs<<<std::min<uint32_t>((s->nruns - s->split_rb + 3) / 4, kStreamGrid), kBlock, 0,
                                   g->stream>>>(s->split_rb, s->nruns, s->d_rv, s->d_rpb, s->d_rpe, s->d_ppart, p);
            PTG_HIP(hipGetLastError());
        }
    }
    PTG_HIP(hipEventRecord(g->ev1, g->stream));
    PTG_HIP(hipEventSynchronize(g->ev1));
    if (st) {
        memset(st, 0, sizeof *st);
        float ms = 0;
        PTG_HIP(hipEventElapsedTime(&ms, g->ev0, g->ev1));
        st->kernel_ns = (uint64_t)((double)ms * 1e6);
        st->iterations = (uint32_t)iters;
        st->reached = V;
        st->edges_traversed = g->E * (uint64_t)iters;
        if ((rc = finish_query_stats(g, st))) return rc;
    }
    return PTG_OK;
}

int stream_info(const ptg_graph *g, uint32_t *chunks, uint32_t *slots, uint64_t *slot_bytes) {
    const StreamState *s = g->streamed;
    if (chunks) *chunks = (uint32_t)s->chunks.size();
    if (slots) *slots = (uint32_t)s->slot.size();
    if (slot_bytes) *slot_bytes = s->slot_bytes;
    return PTG_OK;
}

// Pinned host -> device copy rate: `reps` copies of `bytes` (timed with events).
int host_link_probe(uint64_t bytes, int reps, double *gbs) {
    void *h = nullptr, *d = nullptr;
    hipStream_t st = nullptr;
    hipEvent_t a = nullptr, b = nullptr;
    int rc = PTG_OK;
    if (hipHostMalloc(&h, bytes, hipHostMallocDefault) != hipSuccess) return fail(PTG_ERR_OUT_OF_MEMORY, "pinned probe");
    if (hipMalloc(&d, bytes) != hipSuccess) {
        (void)hipHostFree(h);
        return fail(PTG_ERR_OUT_OF_MEMORY, "device probe");
    }
    memset(h, 1, bytes);
    if (hipStreamCreateWithFlags(&st, hipStreamNonBlocking) != hipSuccess || hipEventCreate(&a) != hipSuccess ||
        hipEventCreate(&b) != hipSuccess) {
        rc = fail(PTG_ERR_HIP, "probe stream/events");
    } else {
        (void)hipMemcpyAsync(d, h, bytes, hipMemcpyHostToDevice, st);  // warm
        (void)hipEventRecord(a, st);
        for (int i = 0; i < reps; i++) (void)hipMemcpyAsync(d, h, bytes, hipMemcpyHostToDevice, st);
        (void)hipEventRecord(b, st);
        if (hipEventSynchronize(b) != hipSuccess) {
            rc = fail(PTG_ERR_HIP, "probe copy");
        } else {
            float ms = 0;
            (void)hipEventElapsedTime(&ms, a, b);
            *gbs = (double)bytes * reps / ((double)ms * 1e6);
        }
    }
    if (a) (void)hipEventDestroy(a);
    if (b) (void)hipEventDestroy(b);
    if (st) (void)hipStreamDestroy(st);
    (void)hipFree(d);
    (void)hipHostFree(h);
    return rc;
}

}  // namespace ptg
