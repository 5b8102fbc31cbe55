// Frontier kernels for gfx950: BFS (top-down + bottom-up), SSSP and CC.
//
// Replaces ptGraph's per-iteration pipeline (setLabeling -> thrust::reduce ->
// thrust::exclusive_scan -> setActiveNodeArray -> setLabelDefault -> relax,
// ptGraph/main.cu:1094-1106, kernels ptGraph/gpu_kernels.cu:3-172) by one
// persistent-grid kernel per round:
//   * the frontier lives in degree-binned queues (S: a lane per vertex,
//     M: a wave per vertex, L: hub adjacency cut into kChunkEdges chunks that
//     any workgroup may take), appended with one atomic per wave (ballot +
//     mbcnt), so no O(V) labelling, reduce or scan runs per round;
//   * relaxation keeps the reference's strict `<` test followed by atomicMin
//     (gpu_kernels.cu:127-131, :143-147, :165-169);
//   * every per-round counter block is kept in a device ring, so the host
//     can launch several rounds before it reads any of them back.
#include <string.h>

#include <algorithm>
#include <vector>

#include "device_util.h"

namespace ptg {

// ---------------------------------------------------------------------------
// Generic kernels
// ---------------------------------------------------------------------------

// Fills n words; `vec` (16-byte aligned p only) writes uint4 per lane.
__global__ __launch_bounds__(kBlock) void k_fill_u32(uint32_t *__restrict__ p, uint64_t n, uint32_t v, bool vec) {
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const uint64_t n4 = vec ? n / 4 : 0;
    uint4 *p4 = reinterpret_cast<uint4 *>(p);
    const uint4 vv = make_uint4(v, v, v, v);
    for (uint64_t j = i; j < n4; j += stride) p4[j] = vv;
    for (uint64_t j = n4 * 4 + i; j < n; j += stride) p[j] = v;
}

__global__ __launch_bounds__(kBlock) void k_iota_u32(uint32_t *__restrict__ p, uint32_t n) {
    const uint32_t stride = gridDim.x * blockDim.x;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) p[i] = i;
}

int fill_u32(ptg_graph *g, uint32_t *p, uint64_t n, uint32_t v) {
    if (n == 0) return PTG_OK;
    const bool vec = ((uintptr_t)p & 15u) == 0;
    uint32_t blocks = (uint32_t)std::min<uint64_t>(((vec ? n / 4 : n) + kBlock - 1) / kBlock + 1, g->grid);
    k_fill_u32<<<blocks, kBlock, 0, g->stream>>>(p, n, v, vec);
    PTG_HIP(hipGetLastError());
    return PTG_OK;
}

// ---------------------------------------------------------------------------
// Binned expansion: one kernel walks the L, M and S bins of a frontier and
// calls op.edge() for every out-edge, always from wave-uniform control flow.
// ---------------------------------------------------------------------------

template <class Op>
__global__ __launch_bounds__(kBlock) void k_expand(Frontier in, Frontier out, BinCounters *zero_ctr,
                                                   const uint32_t *__restrict__ off, Op op,
                                                   uint32_t small_deg, uint32_t large_deg) {
    if (zero_ctr && blockIdx.x == 0 && threadIdx.x < sizeof(BinCounters) / 4)
        reinterpret_cast<uint32_t *>(zero_ctr)[threadIdx.x] = 0u;
    const uint32_t lane = lane_id();
    const uint32_t wib = threadIdx.x >> 6;

    // L bin: chunks of hub adjacency lists, one workgroup per chunk.
    const unsigned long long lp = in.ctr->l_pack;
    const uint32_t nl = (uint32_t)(lp >> kLCountShift);
    const uint32_t nch = (uint32_t)(lp & kLChunkMask);
    for (uint32_t c = blockIdx.x; c < nch; c += gridDim.x) {
        uint32_t lo = 0, hi = nl;
        while (hi - lo > 1) {
            const uint32_t mid = (lo + hi) >> 1;
            if (in.lb[mid] <= c) lo = mid; else hi = mid;
        }
        const uint32_t v = in.ql[lo];
        const uint64_t b = (uint64_t)off[v] + (uint64_t)(c - in.lb[lo]) * kChunkEdges;
        const uint64_t e = min((uint64_t)off[v + 1], b + kChunkEdges);
        const uint32_t sv = op.vertex(v, threadIdx.x == 0);
        for (uint64_t base = b + wib * 64; base < e; base += kBlock)
            op.edge(base + lane < e, sv, base + lane, out, small_deg, large_deg);
    }

    // M bin: one wave per vertex, lanes stride the adjacency (coalesced).
    const uint32_t gw = blockIdx.x * kWavesPerBlock + wib;
    const uint32_t nw = gridDim.x * kWavesPerBlock;
    const uint32_t nm = in.ctr->n_m;
    for (uint32_t i = gw; i < nm; i += nw) {
        const uint32_t v = in.qm[i];
        const uint64_t b = off[v], e = off[v + 1];
        const uint32_t sv = op.vertex(v, lane == 0);
        for (uint64_t base = b; base < e; base += 64)
            op.edge(base + lane < e, sv, base + lane, out, small_deg, large_deg);
    }

    // S bin: one lane per vertex; the wave steps to its longest list.
    const uint32_t ns = in.ctr->n_s;
    for (uint64_t base = (uint64_t)gw * 64; base < ns; base += (uint64_t)nw * 64) {
        const uint64_t i = base + lane;
        const bool valid = i < ns;
        uint32_t b = 0, d = 0, sv = 0;
        if (valid) {
            const uint32_t v = in.qs[i];
            b = off[v];
            d = off[v + 1] - b;
            sv = op.vertex(v, true);
        }
        const uint32_t md = wave_max_u32(d);
        for (uint32_t j = 0; j < md; j++) op.edge(j < d, sv, (uint64_t)b + j, out, small_deg, large_deg);
    }
}

// ---------------------------------------------------------------------------
// BFS
// ---------------------------------------------------------------------------

// Top-down step of level-synchronous BFS. The visited bitmap decides the
// unique discoverer of a vertex (atomicOr), which writes level = current + 1;
// this equals the fixed point of bfs_kernel (gpu_kernels.cu:118-134).
struct BfsTopDown {
    const uint32_t *__restrict__ off;
    const uint32_t *__restrict__ col;
    uint32_t *vis;
    uint32_t *level;
    uint32_t next_level;

    __device__ __forceinline__ uint32_t vertex(uint32_t, bool) const { return 0; }

    __device__ __forceinline__ void edge(bool valid, uint32_t, uint64_t e, const Frontier &out,
                                         uint32_t sd, uint32_t ld) const {
        const uint32_t w = valid ? col[e] : 0u;
        const uint32_t bit = 1u << (w & 31);
        bool won = false;
        if (valid && !(vis[w >> 5] & bit)) won = !(atomicOr(&vis[w >> 5], bit) & bit);
        uint32_t d = 0;
        if (won) {
            level[w] = next_level;
            d = off[w + 1] - off[w];
        }
        count_found(out.ctr, won, d);
        push_binned(out, won && d > 0, w, d, sd, ld);
    }
};

// Bottom-up step (symmetric graphs only): every unvisited vertex looks for a
// neighbour in the current frontier bitmap and stops at the first one. A wave
// owns 64 consecutive vertices = two bitmap words, so no atomics are needed.
__global__ __launch_bounds__(kBlock) void k_bfs_bottom_up(uint32_t V, uint32_t npairs,
                                                          const uint32_t *__restrict__ off,
                                                          const uint32_t *__restrict__ col,
                                                          uint32_t *__restrict__ vis,
                                                          const uint32_t *__restrict__ front,
                                                          uint32_t *__restrict__ next_front,
                                                          uint32_t *__restrict__ level, uint32_t next_level,
                                                          BinCounters *ctr, BinCounters *zero_ctr) {
    if (zero_ctr && blockIdx.x == 0 && threadIdx.x < sizeof(BinCounters) / 4)
        reinterpret_cast<uint32_t *>(zero_ctr)[threadIdx.x] = 0u;
    const uint32_t lane = lane_id();
    const uint32_t gw = blockIdx.x * kWavesPerBlock + (threadIdx.x >> 6);
    const uint32_t nw = gridDim.x * kWavesPerBlock;
    for (uint32_t p = gw; p < npairs; p += nw) {
        const uint32_t v = p * 64 + lane;
        const uint32_t word = p * 2 + (lane >> 5);
        const uint32_t vw = vis[word];
        const bool todo = v < V && !((vw >> (lane & 31)) & 1u);
        bool hit = false;
        uint32_t d = 0;
        if (todo) {
            const uint32_t b = off[v], e = off[v + 1];
            d = e - b;
            for (uint32_t k = b; k < e; k++) {
                const uint32_t u = col[k];
                if ((front[u >> 5] >> (u & 31)) & 1u) {
                    hit = true;
                    break;
                }
            }
        }
        const uint64_t m = ballot(hit);
        if ((lane & 31) == 0) {
            const uint32_t bits = lane ? (uint32_t)(m >> 32) : (uint32_t)m;
            next_front[word] = bits;
            if (bits) vis[word] = vw | bits;
        }
        if (hit) level[v] = next_level;
        count_found(ctr, hit, d);
    }
}

// Queue -> bitmap (entering bottom-up); the bitmap must be cleared before.
__global__ __launch_bounds__(kBlock) void k_bins_to_bitmap(Frontier in, uint32_t *__restrict__ bits) {
    const uint32_t ns = in.ctr->n_s, nm = in.ctr->n_m;
    const uint32_t nl = (uint32_t)(in.ctr->l_pack >> kLCountShift);
    const uint64_t total = (uint64_t)ns + nm + nl;
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < total; i += stride) {
        const uint32_t v = i < ns ? in.qs[i] : (i < (uint64_t)ns + nm ? in.qm[i - ns] : in.ql[i - ns - nm]);
        atomicOr(&bits[v >> 5], 1u << (v & 31));
    }
}

// Bitmap -> queues (leaving bottom-up); found/edges are already counted.
__global__ __launch_bounds__(kBlock) void k_bitmap_to_bins(uint32_t V, uint32_t npairs,
                                                           const uint32_t *__restrict__ bits,
                                                           const uint32_t *__restrict__ off, Frontier out,
                                                           uint32_t sd, uint32_t ld) {
    const uint32_t lane = lane_id();
    const uint32_t gw = blockIdx.x * kWavesPerBlock + (threadIdx.x >> 6);
    const uint32_t nw = gridDim.x * kWavesPerBlock;
    for (uint32_t p = gw; p < npairs; p += nw) {
        const uint32_t v = p * 64 + lane;
        const bool in = v < V && ((bits[p * 2 + (lane >> 5)] >> (lane & 31)) & 1u);
        const uint32_t d = in ? off[v + 1] - off[v] : 0u;
        push_binned(out, in && d > 0, v, d, sd, ld);
    }
}

// Writes the source value and puts the source in the first frontier.
__global__ void k_seed(uint32_t src, uint32_t src_value, uint32_t *__restrict__ val, uint32_t *vis,
                       const uint32_t *__restrict__ off, Frontier out, uint32_t sd, uint32_t ld) {
    const bool leader = lane_id() == 0;
    if (leader) {
        val[src] = src_value;
        if (vis) vis[src >> 5] |= 1u << (src & 31);
    }
    const uint32_t d = off[src + 1] - off[src];
    count_found(out.ctr, leader, d);
    push_binned(out, leader && d > 0, src, d, sd, ld);
}

static inline Frontier ring_frontier(ptg_graph *g, int qi, uint32_t round) {
    return g->q[qi].with(g->d_ring + (round % kRing));
}

static int read_ring(ptg_graph *g, uint32_t from, uint32_t count) {
    // Copies ring entries [from, from + count) (mod kRing) into h_ring.
    uint32_t a = from % kRing;
    uint32_t first = std::min(count, kRing - a);
    PTG_HIP(hipMemcpyAsync(g->h_ring + a, g->d_ring + a, sizeof(BinCounters) * first, hipMemcpyDeviceToHost,
                           g->stream));
    if (count > first)
        PTG_HIP(hipMemcpyAsync(g->h_ring, g->d_ring, sizeof(BinCounters) * (count - first),
                               hipMemcpyDeviceToHost, g->stream));
    PTG_HIP(hipStreamSynchronize(g->stream));
    return PTG_OK;
}

static int clear_ring(ptg_graph *g) {
    return fill_u32(g, reinterpret_cast<uint32_t *>(g->d_ring), sizeof(BinCounters) / 4 * kRing, 0u);
}

int run_bfs(ptg_graph *g, uint32_t src, uint32_t *level, ptg_stats *st) {
    int sym = 0;
    int rc = check_symmetric(g, &sym);
    if (rc) return rc;
    const uint32_t sd = g->cfg.small_degree, ld = g->cfg.large_degree;
    uint32_t *vis = g->d_bits[2];
    uint32_t *fb[2] = {g->d_bits[0], g->d_bits[1]};
    LaunchTimer timer(g);
    PTG_HIP(hipEventRecord(g->ev0, g->stream));
    if ((rc = fill_u32(g, level, g->V, 0xFFFFFFFFu)) || (rc = fill_u32(g, vis, g->nwords, 0u)) ||
        (rc = clear_ring(g)))
        return rc;
    k_seed<<<1, 64, 0, g->stream>>>(src, 1u, level, vis, g->out.off, ring_frontier(g, 0, 0), sd, ld);
    PTG_HIP(hipGetLastError());

    bool bottom_up = false;
    uint64_t edges_seen = 0, reached = 0, frontier_sum = 0;
    uint32_t iters = 0;
    const double alpha = g->cfg.td_to_bu > 0 ? g->cfg.td_to_bu : 14.0;
    const double beta = g->cfg.bu_to_td > 0 ? g->cfg.bu_to_td : 24.0;
    for (uint32_t r = 0;; r++) {
        if ((rc = read_ring(g, r, 1))) return rc;
        const BinCounters c = g->h_ring[r % kRing];
        if (c.found == 0) break;
        iters++;
        reached += c.found;
        frontier_sum += c.found;
        edges_seen += c.edges;
        const int cur = r & 1, nxt = cur ^ 1;
        // Direction choice (Beamer, Asanovic, Patterson 2012).
        const double m_u = (double)(g->E > edges_seen ? g->E - edges_seen : 0);
        bool want_bu = bottom_up;
        if (!bottom_up && sym && (double)c.edges > m_u / alpha) want_bu = true;
        else if (bottom_up && (double)c.found < (double)g->V / beta) want_bu = false;
        Frontier fin = ring_frontier(g, cur, r), fout = ring_frontier(g, nxt, r + 1);
        BinCounters *zero = g->d_ring + ((r + 2) % kRing);
        if (want_bu && !bottom_up) {
            if ((rc = fill_u32(g, fb[cur], g->nwords, 0u))) return rc;
            k_bins_to_bitmap<<<g->grid, kBlock, 0, g->stream>>>(fin, fb[cur]);
        } else if (!want_bu && bottom_up) {
            k_bitmap_to_bins<<<g->grid, kBlock, 0, g->stream>>>(g->V, g->npairs, fb[cur], g->out.off, fin, sd, ld);
        }
        bottom_up = want_bu;
        timer.before();
        if (bottom_up) {
            k_bfs_bottom_up<<<g->grid, kBlock, 0, g->stream>>>(g->V, g->npairs, g->out.off, g->out.col, vis,
                                                               fb[cur], fb[nxt], level, r + 2, fout.ctr, zero);
        } else {
            BfsTopDown op{g->out.off, g->out.col, vis, level, r + 2};
            k_expand<BfsTopDown><<<g->grid, kBlock, 0, g->stream>>>(fin, fout, zero, g->out.off, op, sd, ld);
        }
        timer.after();
        PTG_HIP(hipGetLastError());
    }
    PTG_HIP(hipEventRecord(g->ev1, g->stream));
    PTG_HIP(hipEventSynchronize(g->ev1));
    if (st) {
        memset(st, 0, sizeof *st);
        float ms = 0;
        PTG_HIP(hipEventElapsedTime(&ms, g->ev0, g->ev1));
        st->kernel_ns = (uint64_t)((double)ms * 1e6);
        st->iterations = iters;
        st->reached = (uint32_t)reached;
        st->edges_traversed = edges_seen;
        st->frontier_sum = frontier_sum;
        if ((rc = timer.collect(st))) return rc;
    }
    return PTG_OK;
}

// ---------------------------------------------------------------------------
// SSSP and label-propagation CC: rounds of relaxation over a worklist.
// ---------------------------------------------------------------------------

// Relaxation of sssp_kernel (gpu_kernels.cu:152-172, new = val[u] + w) or of
// cc_kernel (gpu_kernels.cu:136-150, new = val[u]). An improved vertex joins
// the next round's queue once: its bit in `inq_next` dedupes the append, and
// it clears its bit in `inq_cur` when the round that holds it processes it.
template <bool kWeighted>
struct Relax {
    const uint32_t *__restrict__ off;
    const uint32_t *__restrict__ col;
    const uint32_t *__restrict__ wt;
    uint32_t *val;
    uint32_t *inq_cur;
    uint32_t *inq_next;

    __device__ __forceinline__ uint32_t vertex(uint32_t v, bool leader) const {
        if (leader) atomicAnd(&inq_cur[v >> 5], ~(1u << (v & 31)));
        return val[v];
    }

    __device__ __forceinline__ void edge(bool valid, uint32_t sv, uint64_t e, const Frontier &out,
                                         uint32_t sd, uint32_t ld) const {
        const uint32_t w = valid ? col[e] : 0u;
        const uint32_t nv = kWeighted ? sv + (valid ? wt[e] : 0u) : sv;
        bool enq = false;
        uint32_t d = 0;
        if (valid && nv < val[w]) {
            const uint32_t old = atomicMin(&val[w], nv);
            if (nv < old) {
                d = off[w + 1] - off[w];
                const uint32_t bit = 1u << (w & 31);
                if (d) enq = !(atomicOr(&inq_next[w >> 5], bit) & bit);
            }
        }
        count_found(out.ctr, enq, d);
        push_binned(out, enq, w, d, sd, ld);
    }
};

// Runs rounds until a round produces an empty queue. Round r reads queue
// r % 2 (round 0 may read `first`) and counters ring[r], appends to the other
// queue and ring[r + 1], and clears ring[r + 2]; so the host can launch a
// batch of rounds and read their counters afterwards.
template <bool kWeighted>
static int run_rounds(ptg_graph *g, uint32_t *val, const Frontier *first, LaunchTimer &timer,
                      uint32_t *iters_out, uint64_t *frontier_sum_out) {
    const uint32_t sd = g->cfg.small_degree, ld = g->cfg.large_degree;
    uint32_t *inq[2] = {g->d_bits[0], g->d_bits[1]};
    uint32_t r = 0, iters = 0, batch = 1;
    uint64_t fsum = 0;
    for (;;) {
        for (uint32_t k = 0; k < batch; k++, r++) {
            Frontier fin = (r == 0 && first) ? *first : ring_frontier(g, r & 1, r);
            Frontier fout = ring_frontier(g, (r + 1) & 1, r + 1);
            Relax<kWeighted> op{g->out.off, g->out.col, g->d_w, val, inq[(r + 1) & 1], inq[r & 1]};
            timer.before();
            k_expand<Relax<kWeighted>><<<g->grid, kBlock, 0, g->stream>>>(fin, fout, g->d_ring + ((r + 2) % kRing),
                                                                          g->out.off, op, sd, ld);
            timer.after();
            PTG_HIP(hipGetLastError());
        }
        int rc = read_ring(g, r - batch + 1, batch);
        if (rc) return rc;
        bool done = false;
        for (uint32_t k = r - batch + 1; k <= r; k++) {
            const BinCounters &c = g->h_ring[k % kRing];
            if (!done) iters++;
            if (bin_total(c) == 0) {
                done = true;
            } else {
                fsum += bin_total(c);
            }
        }
        if (done) break;
        if (batch < 16) batch *= 2;
    }
    *iters_out = iters;
    *frontier_sum_out = fsum;
    return PTG_OK;
}

int run_sssp(ptg_graph *g, uint32_t src, uint32_t *dist, ptg_stats *st) {
    const uint32_t sd = g->cfg.small_degree, ld = g->cfg.large_degree;
    int rc;
    LaunchTimer timer(g);
    PTG_HIP(hipEventRecord(g->ev0, g->stream));
    if ((rc = fill_u32(g, dist, g->V, g->cfg.sssp_sentinel)) || (rc = fill_u32(g, g->d_bits[0], g->nwords, 0u)) ||
        (rc = fill_u32(g, g->d_bits[1], g->nwords, 0u)) || (rc = clear_ring(g)))
        return rc;
    k_seed<<<1, 64, 0, g->stream>>>(src, 1u, dist, nullptr, g->out.off, ring_frontier(g, 0, 0), sd, ld);
    PTG_HIP(hipGetLastError());
    uint32_t iters = 0;
    uint64_t fsum = 0;
    if ((rc = run_rounds<true>(g, dist, nullptr, timer, &iters, &fsum))) return rc;
    PTG_HIP(hipEventRecord(g->ev1, g->stream));
    PTG_HIP(hipEventSynchronize(g->ev1));
    if (st) {
        memset(st, 0, sizeof *st);
        float ms = 0;
        PTG_HIP(hipEventElapsedTime(&ms, g->ev0, g->ev1));
        st->kernel_ns = (uint64_t)((double)ms * 1e6);
        st->iterations = iters;
        st->frontier_sum = fsum + 1;
        if ((rc = timer.collect(st))) return rc;
        if ((rc = reach_stats(g, dist, g->cfg.sssp_sentinel, &st->reached, &st->edges_traversed))) return rc;
    }
    return PTG_OK;
}

// ---------------------------------------------------------------------------
// CC: union-find hooking on symmetric graphs, label propagation otherwise.
// ---------------------------------------------------------------------------

// Root search with path halving. Parents only ever point to smaller ids, so
// any value read (even a stale one) is an ancestor and the walk terminates.
__device__ __forceinline__ uint32_t uf_root(uint32_t *par, uint32_t x) {
    for (;;) {
        const uint32_t p = par[x];
        if (p == x) return x;
        const uint32_t gp = par[p];
        if (gp == p) return p;
        par[x] = gp;
        x = gp;
    }
}

// Links the two trees, always hanging the larger root under the smaller one,
// so every root is the minimum vertex id of its tree.
__device__ __forceinline__ void uf_unite(uint32_t *par, uint32_t a, uint32_t b) {
    uint32_t ra = uf_root(par, a), rb = uf_root(par, b);
    while (ra != rb) {
        if (ra < rb) {
            const uint32_t t = ra;
            ra = rb;
            rb = t;
        }
        const uint32_t seen = atomicCAS(&par[ra], ra, rb);
        if (seen == ra) return;
        ra = uf_root(par, seen);
        rb = uf_root(par, rb);
    }
}

struct Hook {
    const uint32_t *__restrict__ col;
    uint32_t *par;
    __device__ __forceinline__ uint32_t vertex(uint32_t v, bool) const { return v; }
    __device__ __forceinline__ void edge(bool valid, uint32_t v, uint64_t e, const Frontier &, uint32_t,
                                         uint32_t) const {
        if (!valid) return;
        const uint32_t u = col[e];
        if (u < v) uf_unite(par, v, u);
    }
};

// par[v] = smallest of v and its first neighbour (a valid first hook).
__global__ __launch_bounds__(kBlock) void k_uf_init(uint32_t V, const uint32_t *__restrict__ off,
                                                    const uint32_t *__restrict__ col, uint32_t *__restrict__ par) {
    const uint32_t stride = gridDim.x * blockDim.x;
    for (uint32_t v = blockIdx.x * blockDim.x + threadIdx.x; v < V; v += stride) {
        const uint32_t b = off[v];
        uint32_t p = v;
        if (off[v + 1] > b) p = min(v, col[b]);
        par[v] = p;
    }
}

__global__ __launch_bounds__(kBlock) void k_uf_flatten(uint32_t V, uint32_t *par) {
    const uint32_t stride = gridDim.x * blockDim.x;
    for (uint32_t v = blockIdx.x * blockDim.x + threadIdx.x; v < V; v += stride) par[v] = uf_root(par, v);
}

// Number of v with label[v] == v.
__global__ __launch_bounds__(kBlock) void k_count_roots(uint32_t V, const uint32_t *__restrict__ label,
                                                        unsigned long long *out) {
    const uint32_t lane = lane_id();
    const uint32_t gw = blockIdx.x * kWavesPerBlock + (threadIdx.x >> 6);
    const uint32_t nw = gridDim.x * kWavesPerBlock;
    uint64_t cnt = 0;
    for (uint64_t base = (uint64_t)gw * 64; base < V; base += (uint64_t)nw * 64) {
        const uint64_t v = base + lane;
        cnt += (v < V && label[v] == v);
    }
    cnt = wave_sum_u64(cnt);
    if (lane == 0 && cnt) atomicAdd(out, (unsigned long long)cnt);
}

static int count_roots(ptg_graph *g, const uint32_t *label, uint32_t *roots) {
    unsigned long long *acc = g->d_acc;
    int rc = fill_u32(g, reinterpret_cast<uint32_t *>(acc), 2, 0u);
    if (rc) return rc;
    k_count_roots<<<g->grid, kBlock, 0, g->stream>>>(g->V, label, acc);
    PTG_HIP(hipGetLastError());
    unsigned long long h = 0;
    PTG_HIP(hipMemcpyAsync(&h, acc, sizeof h, hipMemcpyDeviceToHost, g->stream));
    PTG_HIP(hipStreamSynchronize(g->stream));
    *roots = (uint32_t)h;
    return PTG_OK;
}

int run_cc(ptg_graph *g, uint32_t *label, ptg_stats *st) {
    int sym = 0;
    int rc = check_symmetric(g, &sym);
    if (rc) return rc;
    const uint32_t sd = g->cfg.small_degree, ld = g->cfg.large_degree;
    uint32_t iters = 1;
    uint64_t fsum = g->V;
    LaunchTimer timer(g);
    PTG_HIP(hipEventRecord(g->ev0, g->stream));
    if (sym) {
        k_uf_init<<<g->grid, kBlock, 0, g->stream>>>(g->V, g->out.off, g->out.col, label);
        Hook op{g->out.col, label};
        Frontier fin = g->out.bins.q.with(g->out.bins.d_ctr);
        timer.before();
        k_expand<Hook><<<g->grid, kBlock, 0, g->stream>>>(fin, fin, nullptr, g->out.off, op, sd, ld);
        timer.after();
        k_uf_flatten<<<g->grid, kBlock, 0, g->stream>>>(g->V, label);
        PTG_HIP(hipGetLastError());
    } else {
        // Label propagation from value[v] = v with every vertex active
        // (ccCaculateInShare, ptGraph/main.cu:977-980).
        k_iota_u32<<<g->grid, kBlock, 0, g->stream>>>(label, g->V);
        if ((rc = fill_u32(g, g->d_bits[0], g->nwords, 0u)) || (rc = fill_u32(g, g->d_bits[1], g->nwords, 0u)) ||
            (rc = clear_ring(g)))
            return rc;
        Frontier first = g->out.bins.q.with(g->out.bins.d_ctr);
        uint64_t s = 0;
        if ((rc = run_rounds<false>(g, label, &first, timer, &iters, &s))) return rc;
        fsum = g->V + s;
    }
    PTG_HIP(hipEventRecord(g->ev1, g->stream));
    PTG_HIP(hipEventSynchronize(g->ev1));
    if (st) {
        memset(st, 0, sizeof *st);
        float ms = 0;
        PTG_HIP(hipEventElapsedTime(&ms, g->ev0, g->ev1));
        st->kernel_ns = (uint64_t)((double)ms * 1e6);
        st->iterations = iters;
        st->frontier_sum = fsum;
        st->edges_traversed = g->E;
        if ((rc = timer.collect(st))) return rc;
        // reached = number of distinct labels (components on a symmetric graph)
        if ((rc = count_roots(g, label, &st->reached))) return rc;
    }
    return PTG_OK;
}

// ---------------------------------------------------------------------------
// Symmetry check (done once per graph): every adjacency sorted ascending and
// every edge (v, u) matched by (u, v), found by binary search in u's list.
// ---------------------------------------------------------------------------

struct CheckSorted {
    const uint32_t *__restrict__ off;
    const uint32_t *__restrict__ col;
    uint32_t *flag;
    __device__ __forceinline__ uint32_t vertex(uint32_t v, bool) const { return off[v + 1]; }
    __device__ __forceinline__ void edge(bool valid, uint32_t end, uint64_t e, const Frontier &, uint32_t,
                                         uint32_t) const {
        if (valid && e + 1 < end && col[e] > col[e + 1]) *flag = 0u;
    }
};

struct CheckReverse {
    const uint32_t *__restrict__ off;
    const uint32_t *__restrict__ col;
    uint32_t *flag;
    __device__ __forceinline__ uint32_t vertex(uint32_t v, bool) const { return v; }
    __device__ __forceinline__ void edge(bool valid, uint32_t v, uint64_t e, const Frontier &, uint32_t,
                                         uint32_t) const {
        if (!valid) return;
        const uint32_t u = col[e];
        uint32_t lo = off[u], hi = off[u + 1];
        while (lo < hi) {
            const uint32_t mid = lo + ((hi - lo) >> 1);
            if (col[mid] < v) lo = mid + 1; else hi = mid;
        }
        if (lo == off[u + 1] || col[lo] != v) *flag = 0u;
    }
};

int check_symmetric(ptg_graph *g, int *sym) {
    if (g->symmetric >= 0) {
        *sym = g->symmetric;
        return PTG_OK;
    }
    uint32_t *flag = g->d_bits[0];
    int rc = fill_u32(g, flag, 2, 1u);
    if (rc) return rc;
    Frontier fin = g->out.bins.q.with(g->out.bins.d_ctr);
    const uint32_t sd = g->cfg.small_degree, ld = g->cfg.large_degree;
    k_expand<CheckSorted><<<g->grid, kBlock, 0, g->stream>>>(fin, fin, nullptr, g->out.off,
                                                             CheckSorted{g->out.off, g->out.col, flag}, sd, ld);
    k_expand<CheckReverse><<<g->grid, kBlock, 0, g->stream>>>(fin, fin, nullptr, g->out.off,
                                                              CheckReverse{g->out.off, g->out.col, flag + 1}, sd, ld);
    PTG_HIP(hipGetLastError());
    uint32_t h[2];
    PTG_HIP(hipMemcpyAsync(h, flag, sizeof h, hipMemcpyDeviceToHost, g->stream));
    PTG_HIP(hipStreamSynchronize(g->stream));
    // The reverse search assumes sorted lists; unsorted means "not proven".
    g->symmetric = (h[0] && h[1]) ? 1 : 0;
    *sym = g->symmetric;
    return PTG_OK;
}

// ---------------------------------------------------------------------------
// Reached count and m_r = sum of out-degrees of reached vertices.
// ---------------------------------------------------------------------------

__global__ __launch_bounds__(kBlock) void k_reach_stats(uint32_t V, const uint32_t *__restrict__ val,
                                                        uint32_t unreached, const uint32_t *__restrict__ off,
                                                        unsigned long long *out) {
    const uint32_t lane = lane_id();
    const uint32_t gw = blockIdx.x * kWavesPerBlock + (threadIdx.x >> 6);
    const uint32_t nw = gridDim.x * kWavesPerBlock;
    uint64_t cnt = 0, edges = 0;
    for (uint64_t base = (uint64_t)gw * 64; base < V; base += (uint64_t)nw * 64) {
        const uint64_t v = base + lane;
        if (v < V && val[v] != unreached) {
            cnt++;
            edges += off[v + 1] - off[v];
        }
    }
    cnt = wave_sum_u64(cnt);
    edges = wave_sum_u64(edges);
    if (lane == 0) {
        atomicAdd(&out[0], (unsigned long long)cnt);
        atomicAdd(&out[1], (unsigned long long)edges);
    }
}

int reach_stats(ptg_graph *g, const uint32_t *val, uint32_t unreached, uint32_t *reached, uint64_t *edges) {
    if (!reached && !edges) return PTG_OK;
    unsigned long long *acc = g->d_acc;
    int rc = fill_u32(g, reinterpret_cast<uint32_t *>(acc), 4, 0u);
    if (rc) return rc;
    k_reach_stats<<<g->grid, kBlock, 0, g->stream>>>(g->V, val, unreached, g->out.off, acc);
    PTG_HIP(hipGetLastError());
    unsigned long long h[2];
    PTG_HIP(hipMemcpyAsync(h, acc, sizeof h, hipMemcpyDeviceToHost, g->stream));
    PTG_HIP(hipStreamSynchronize(g->stream));
    if (reached) *reached = (uint32_t)h[0];
    if (edges) *edges = h[1];
    return PTG_OK;
}

}  // namespace ptg
