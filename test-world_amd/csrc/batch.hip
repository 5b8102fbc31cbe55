// Batched multi-source BFS and SSSP for gfx950: up to 64 sources per batch,
// each source's results identical to a single-source query.
//
// BFS (bit-parallel, "multi-source BFS": Then et al., VLDB 2014). Every vertex
// carries 64-bit masks: seen (sources that reached it) and a frontier mask
// per level, so one scan of an adjacency list serves all 64 sources. A level
// is top-down (frontier vertices OR their mask into their neighbours' next
// mask with a 64-bit atomicOr) or bottom-up (each vertex not yet seen by every
// source ORs its neighbours' frontier masks, stopping once every missing
// source is found; no atomics), chosen on the device from the previous
// level's counts. The level at which each source's bit reaches a vertex is
// kept as one mask per (level, vertex); the output kernel turns those masks
// into per-source levels.
//
// SSSP: one wave lane per source. Distances are stored vertex-major
// (dist[v * 64 + k]), so relaxing edge (v, w) for every source reads and
// updates one 256-byte row of w. The schedule is the near/far worklist of the
// single-source path with one threshold shared by all sources: a source's
// improvement below the threshold sets its bit in w's next-frontier mask,
// above it its bit in w's far mask; an empty near frontier raises the
// threshold and moves far bits below it into the frontier. Relaxation is the
// reference's rule (strict `<`, then atomicMin: sssp_kernel,
// ptGraph/gpu_kernels.cu:152-172), so every source reaches the same fixed
// point as ptg_sssp.
//
// Pull rounds (batch_pull). On a symmetric graph, a round in which most
// vertices were improved for some source can instead pull: every vertex w
// takes, for every source, the minimum of its distance and dist[u] + w(u, w)
// over its in-edges (u, w) from reached u, chunk by chunk, and applies it
// with one atomicMin per chunk instead of one per improving edge. A pull
// relaxes every edge a push round would (and edges out of unchanged vertices,
// whose candidates were applied before), so push and pull rounds in any mix
// reach the same fixed point. Which rounds pull is learned per graph: round r
// pulls when round r of the previous batch expanded at least E x batch_pull
// edges and round r - 1 of this one at least an eighth of that.
//
// Rounds run as hipGraph replays of a fixed number of levels/rounds; every
// kernel reads its round number from device memory and returns at once when
// the search has ended, so a replay past the end is harmless.
#include <string.h>

#include <algorithm>
#include <vector>

#include "static_expand.h"

namespace ptg {

constexpr uint32_t kBatchMax = 64;
constexpr uint32_t kBRing = 64;        // counter ring entries (> replay length + 2)
static_assert(sizeof(ptg_graph::batch_pull_rounds) == kBRing, "one learned pull flag per ring entry");
constexpr uint32_t kBfsReplay = 8;     // BFS levels per graph replay
constexpr uint32_t kSsspReplay = 16;   // SSSP rounds per graph replay
constexpr uint32_t kLevelCap = 32;     // BFS levels kept as masks; deeper -> per-source path
constexpr uint32_t kBatchChunk = 128;  // SSSP: default edges per static chunk
constexpr float kBatchPull = 0.25f;    // SSSP: default batch_pull

// Counters of one level / round, 16-way sharded.
struct BCtr {
    unsigned long long nf[kShards];    // vertices in the next frontier (SSSP expand: non-zero if any)
    unsigned long long mf[kShards];    // their out-degree sum
    unsigned long long munf[kShards];  // BFS: out-degree sum of vertices some source has not reached;
                                       // SSSP: edges visited by the round
    unsigned long long far[kShards];   // SSSP: far masks left non-empty; BFS: OR of the frontier
                                       // masks (sources that can still reach new vertices)
    uint32_t thr;                      // SSSP: threshold of the frontier this entry describes
    uint32_t overflow;                 // BFS: more than kLevelCap levels
    uint32_t pad[126];
};
static_assert(sizeof(BCtr) == 1024, "BCtr layout");

// Per-batch scalars in device memory (so captured graphs stay valid).
struct BParams {
    uint32_t n;                   // sources
    uint32_t base;                // first level / round of the current replay
    unsigned long long mask;      // low n bits
    unsigned long long visits;    // SSSP: edge visits of the rounds summed so far
};

struct BatchEnv {
    uint32_t V, nva;           // vertices; [0, nva) holds every vertex with out-degree > 0
    const uint32_t *off, *col, *wt;
    Frontier bins;             // static out-degree bins
    unsigned long long *seen;  // BFS
    unsigned long long *f0, *f1, *f2;  // rotating frontier masks
    unsigned long long *far;   // SSSP
    unsigned long long *lm;    // BFS: kLevelCap x V level masks
    uint32_t *dist;            // SSSP: V x 64
    const uint32_t *ch_v;      // SSSP: static chunk list, vertex of each chunk
    const uint32_t *ch_b;      //   and its first edge
    uint32_t nch, chunk;       // chunks; edges per chunk
    const uint2 *cwin;         // SSSP pull rounds: in-edge (source, weight) pairs; null = push only
    const uint8_t *pmode;      // per round (< kBRing): pull when the previous batch's round was dense
    unsigned long long pull_prev;  // and the previous round of this batch expanded >= this many edges
    unsigned long long pull_dense; // E x batch_pull: a dense round (the rule of a graph's first batch)
    uint32_t learned;          // pmode holds a previous batch's rounds
    uint32_t sentinel;         // SSSP: unreached value
    BCtr *ring;
    BParams *p;
    uint32_t sd;
    float alpha;
    uint32_t delta;
    int symmetric;
    uint32_t grid;             // launch width of the captured replays
};

struct BatchState {
    unsigned long long *seen = nullptr, *f[3] = {}, *far = nullptr, *lm = nullptr;
    uint32_t *dist = nullptr;
    uint32_t *ch_v = nullptr, *ch_b = nullptr;
    uint32_t nch = 0, chunk = 0;
    uint8_t *pmode = nullptr;                   // SSSP pull rounds (the graph's batch_pull_rounds)
    BCtr *h_ring = nullptr;                     // pinned copy of the counter ring (pull learning)
    uint32_t *out = nullptr;   // n x V staging for host outputs
    uint32_t *srcs = nullptr;  // device copy of the sources
    uint32_t *h_srcs = nullptr;
    BCtr *ring = nullptr;
    BParams *p = nullptr;
    BCtr *h_ctr = nullptr;     // pinned: one counter entry read back per replay
    BParams *h_p = nullptr;    // pinned copy of the batch scalars
    hipGraphExec_t bfs_exec = nullptr, sssp_exec = nullptr;
    BatchEnv env_bfs{}, env_sssp{};
};

__device__ __forceinline__ unsigned long long *rot(const BatchEnv &e, uint32_t i) {
    const uint32_t r = i % 3;
    return r == 0 ? e.f0 : (r == 1 ? e.f1 : e.f2);
}

__device__ __forceinline__ uint64_t wave_or_u64(uint64_t x) {
    for (int d = 32; d >= 1; d >>= 1) {
        const uint32_t lo = (uint32_t)__shfl_xor((int)(uint32_t)x, d, 64);
        const uint32_t hi = (uint32_t)__shfl_xor((int)(uint32_t)(x >> 32), d, 64);
        x |= (uint64_t)hi << 32 | lo;
    }
    return x;
}

struct BSum {
    unsigned long long nf, mf, munf, far;
};

__device__ __forceinline__ BSum bsum(const BCtr *c) {
    BSum s{0, 0, 0, 0};
    for (int i = 0; i < (int)kShards; i++) {
        s.nf += c->nf[i];
        s.mf += c->mf[i];
        s.munf += c->munf[i];
        s.far |= c->far[i];
    }
    return s;
}

// Zeroes ring entry `i` (every thread of workgroup 0 takes a share).
__device__ __forceinline__ void bctr_clear(BCtr *c) {
    uint32_t *w = reinterpret_cast<uint32_t *>(c);
    for (uint32_t i = threadIdx.x; i < sizeof(BCtr) / 4; i += blockDim.x) w[i] = 0;
}

// Adds per-thread tallies into shard blockIdx % kShards of `c`.
__device__ __forceinline__ void bctr_add(BCtr *c, uint64_t nf, uint64_t mf, uint64_t munf, uint64_t far) {
    __shared__ unsigned long long s[4][kWavesPerBlock];
    nf = wave_sum_u64(nf);
    mf = wave_sum_u64(mf);
    munf = wave_sum_u64(munf);
    far = wave_or_u64(far);
    const uint32_t wib = threadIdx.x >> 6;
    if (lane_id() == 0) {
        s[0][wib] = nf;
        s[1][wib] = mf;
        s[2][wib] = munf;
        s[3][wib] = far;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned long long a = 0, b = 0, d = 0, f = 0;
        for (int w = 0; w < kWavesPerBlock; w++) {
            a += s[0][w];
            b += s[1][w];
            d += s[2][w];
            f |= s[3][w];
        }
        const uint32_t sh = blockIdx.x % kShards;
        if (a) atomicAdd(&c->nf[sh], a);
        if (b) atomicAdd(&c->mf[sh], b);
        if (d) atomicAdd(&c->munf[sh], d);
        if (f) atomicOr(&c->far[sh], f);
    }
}

__global__ void k_batch_advance(BParams *p, uint32_t k) {
    if (threadIdx.x == 0 && blockIdx.x == 0) p->base += k;
}

// ---------------------------------------------------------------------------
// BFS
// ---------------------------------------------------------------------------

__device__ __forceinline__ bool bfs_bottom_up_dir(const BatchEnv &e, const BSum &s) {
    return e.symmetric && (double)s.mf * (double)e.alpha > (double)s.munf;
}
// Only sources with a non-empty frontier can reach anything new, so a
// vertex looks only for their bits: the bits of sources that cannot reach it
// would otherwise keep every bottom-up scan from stopping early.

// Top-down: OR the frontier mask into each neighbour's next mask, only with
// bits that neighbour has not seen.
struct MsPush {
    const uint32_t *__restrict__ col;
    const unsigned long long *seen, *cur;
    unsigned long long *nxt;
    __device__ __forceinline__ unsigned long long vertex(uint32_t v) const { return cur[v]; }
    __device__ __forceinline__ bool active(unsigned long long f) const { return f != 0; }
    __device__ __forceinline__ void edge(bool valid, unsigned long long f, uint64_t e) const {
        if (!valid) return;
        const uint32_t w = col[e];
        const unsigned long long b = f & ~seen[w];
        if (b && (nxt[w] & b) != b) atomicOr(&nxt[w], b);
    }
};

__global__ __launch_bounds__(kBlock) void k_msbfs_expand(BatchEnv e, uint32_t j) {
    const uint32_t L = e.p->base + j;
    const BSum s = bsum(e.ring + L % kBRing);
    if (s.nf == 0 || bfs_bottom_up_dir(e, s)) return;
    expand_static(e.bins, e.off, MsPush{e.col, e.seen, rot(e, L), rot(e, L + 1)});
}

// Finishes level L: in top-down mode the next masks are complete, in
// bottom-up mode each vertex computes its own. Updates seen, records the
// level mask, clears the buffer the next level accumulates into and counts.
__global__ __launch_bounds__(kBlock) void k_msbfs_finish(BatchEnv e, uint32_t j) {
    const uint32_t L = e.p->base + j;
    BCtr *in = e.ring + L % kBRing;
    BCtr *out = e.ring + (L + 1) % kBRing;
    const bool keep = L + 1 < kLevelCap;
    if (blockIdx.x == 0) {
        bctr_clear(e.ring + (L + 2) % kBRing);
        // sticky: every later entry reports that levels were not all kept
        if (threadIdx.x == 0 && (in->overflow || !keep)) out->overflow = 1;
    }
    const BSum s = bsum(in);
    if (s.nf == 0) return;
    const bool bu = bfs_bottom_up_dir(e, s);
    const unsigned long long mask = e.p->mask;
    const unsigned long long live = s.far & mask;  // sources whose frontier is non-empty
    const unsigned long long *cur = rot(e, L);
    unsigned long long *nxt = rot(e, L + 1);
    unsigned long long *clr = rot(e, L + 2);
    unsigned long long *lm = e.lm + (uint64_t)(keep ? L + 1 : 0) * e.V;
    const uint32_t lane = lane_id();
    const uint32_t gw = blockIdx.x * kWavesPerBlock + (threadIdx.x >> 6);
    const uint32_t nw = gridDim.x * kWavesPerBlock;
    const uint32_t np = (e.nva + 63) / 64;
    uint64_t t_nf = 0, t_mf = 0, t_munf = 0, t_live = 0;
    for (uint32_t p = gw; p < np; p += nw) {
        const uint32_t v = p * 64 + lane;
        const bool inb = v < e.nva;
        const unsigned long long sn = inb ? e.seen[v] : mask;
        const unsigned long long unseen = ~sn & live;
        uint32_t b = 0, d = 0;
        if (inb) {
            b = e.off[v];
            d = e.off[v + 1] - b;
        }
        unsigned long long nb = 0;
        if (!bu) {
            if (unseen) nb = nxt[v] & unseen;
        } else {
            // hubs: the whole wave scans 64 neighbours per step
            uint64_t mm = ballot(unseen && d > e.sd);
            while (mm) {
                const uint32_t q = lowest_lane(mm);
                mm &= mm - 1;
                const uint32_t bq = read_lane(b, q), dq = read_lane(d, q);
                const unsigned long long uq = read_lane64(unseen, q);
                unsigned long long acc = 0;
                for (uint32_t k = 0; k < dq; k += 64) {
                    unsigned long long a = 0;
                    if (k + lane < dq) a = cur[e.col[(uint64_t)bq + k + lane]];
                    acc |= wave_or_u64(a);
                    if ((acc & uq) == uq) break;
                }
                if (lane == q) nb = acc & uq;
            }
            if (unseen && d > 0 && d <= e.sd) {
                unsigned long long acc = 0;
                for (uint32_t k = 0; k < d; k += 4) {
                    uint32_t u[4];
#pragma unroll
                    for (int t = 0; t < 4; t++) u[t] = (k + t < d) ? e.col[(uint64_t)b + k + t] : 0u;
#pragma unroll
                    for (int t = 0; t < 4; t++)
                        if (k + t < d) acc |= cur[u[t]];
                    if ((acc & unseen) == unseen) break;
                }
                nb = acc & unseen;
            }
        }
        if (inb) {
            clr[v] = 0;
            if (keep) lm[v] = nb;
        }
        if (nb) {
            e.seen[v] = sn | nb;
            if (bu) nxt[v] = nb;
        }
        t_nf += nb != 0;
        t_mf += nb ? d : 0;
        t_munf += (inb && (~(sn | nb) & live)) ? d : 0;
        t_live |= nb;
    }
    bctr_add(out, t_nf, t_mf, t_munf, t_live);
}

// Seeds level 0 from the device copy of the sources (one workgroup).
__global__ __launch_bounds__(kBlock) void k_msbfs_seed(BatchEnv e, const uint32_t *srcs, uint32_t n,
                                                       unsigned long long E) {
    __shared__ uint32_t s_first[kBatchMax];
    bctr_clear(e.ring);
    bctr_clear(e.ring + 1);
    const uint32_t t = threadIdx.x;
    __syncthreads();
    if (t < n) {
        const uint32_t s = srcs[t];
        atomicOr(&e.seen[s], 1ull << t);
        atomicOr(&e.f0[s], 1ull << t);
        atomicOr(&e.lm[s], 1ull << t);
        bool first = true;
        for (uint32_t q = 0; q < t; q++) first = first && srcs[q] != s;
        s_first[t] = first;
    }
    __syncthreads();
    if (t == 0) {
        unsigned long long nf = 0, mf = 0;
        for (uint32_t q = 0; q < n; q++)
            if (s_first[q]) {
                nf++;
                mf += e.off[srcs[q] + 1] - e.off[srcs[q]];
            }
        e.ring[0].nf[0] = nf;
        e.ring[0].mf[0] = mf;
        e.ring[0].munf[0] = E;
        e.ring[0].far[0] = n == 64 ? ~0ull : ((1ull << n) - 1);
        e.p->n = n;
        e.p->base = 0;
        e.p->mask = n == 64 ? ~0ull : ((1ull << n) - 1);
    }
}

// out[k * V + c] = level of caller vertex c from source k (0xFFFFFFFF when
// unreached). A workgroup handles 256 caller ids: each lane turns its
// vertex's level masks into levels in LDS, then rows are written coalesced.
// Also counts reached (source, vertex) pairs and their out-degree sum.
__global__ __launch_bounds__(kBlock) void k_msbfs_output(BatchEnv e, uint32_t levels, const uint32_t *perm,
                                                         uint32_t *out, unsigned long long *acc) {
    __shared__ uint8_t tile[kBatchMax][kBlock];  // level, 0 = unreached (levels <= kLevelCap)
    const uint32_t n = e.p->n;
    uint64_t pairs = 0, edges = 0;
    for (uint64_t base = (uint64_t)blockIdx.x * kBlock; base < e.V; base += (uint64_t)gridDim.x * kBlock) {
        const uint64_t c = base + threadIdx.x;
        for (uint32_t k = 0; k < n; k++) tile[k][threadIdx.x] = 0;
        if (c < e.V) {
            const uint32_t dv = perm ? perm[c] : (uint32_t)c;
            const unsigned long long sn = e.seen[dv];
            const uint32_t lim = dv < e.nva ? levels : 1u;
            unsigned long long done = 0;
            for (uint32_t L = 0; L < lim && done != sn; L++) {
                unsigned long long m = e.lm[(uint64_t)L * e.V + dv] & sn & ~done;
                done |= m;
                while (m) {
                    const uint32_t k = (uint32_t)__builtin_ctzll(m);
                    m &= m - 1;
                    tile[k][threadIdx.x] = L + 1;
                }
            }
            const uint32_t np = (uint32_t)__popcll(sn);
            pairs += np;
            edges += (uint64_t)np * (e.off[dv + 1] - e.off[dv]);
        }
        __syncthreads();
        if (c < e.V)
            for (uint32_t k = 0; k < n; k++) {
                const uint32_t l = tile[k][threadIdx.x];
                out[(uint64_t)k * e.V + c] = l ? l : 0xFFFFFFFFu;
            }
        __syncthreads();
    }
    bctr_add(reinterpret_cast<BCtr *>(acc), pairs, edges, 0, 0);
}

// ---------------------------------------------------------------------------
// SSSP
// ---------------------------------------------------------------------------

// A chunk of the static chunk list as the round expands it.
// The wave-uniform fields are moved to scalar registers.
struct ChunkHead {
    unsigned long long a;  // sources active at the chunk's vertex
    uint32_t eb, ee;       // its edges
    uint32_t dv;           // this lane's distance of the vertex
    uint32_t mw, mwt;      // columns and weights of the first 64 edges
};

__device__ __forceinline__ uint32_t uniform(uint32_t x) { return (uint32_t)__builtin_amdgcn_readfirstlane((int)x); }

// Issues the loads a chunk needs before its rows (indexes clamped into the
// chunk, so every lane issues every load). v, eb, ee and a are wave-uniform.
__device__ __forceinline__ ChunkHead chunk_head(const BatchEnv &e, uint32_t v, uint32_t eb, uint32_t ee,
                                                unsigned long long a) {
    const uint32_t lane = lane_id();
    ChunkHead h;
    h.a = ((unsigned long long)uniform((uint32_t)(a >> 32)) << 32) | uniform((uint32_t)a);
    h.eb = uniform(eb);
    h.ee = uniform(ee);
    v = uniform(v);
    h.dv = e.dist[(uint64_t)v * 64 + lane];
    const uint32_t i = min(h.eb + lane, h.ee - 1);
    h.mw = e.col[i];
    h.mwt = e.wt[i];
    return h;
}

// Relaxes the chunk's edges [eb, ee) for every source whose bit is in `a`
// (lane k = source k; dv = this lane's distance of the chunk's vertex). Rows are read in
// groups of kP edges, the next group's loads issued before the current
// group's atomics. Every lane issues every load (lanes of sources not in `a`
// read the first active source's word of the same row, so no line is added),
// which keeps the loads out of divergent branches: the compiler can then
// count them and wait for one group while the next is in flight. The atomics
// return nothing the wave waits for: an improvement is taken from the row
// just read (values only fall, so every thread that lowers dist[w][k] sees
// nd < old and sets the bit; a bit set by a thread another one beat only
// re-expands w with its current value).
__device__ __forceinline__ void relax_chunk(const BatchEnv &e, const ChunkHead &h, uint32_t thr,
                                            unsigned long long *nxt, uint64_t &t_nf, uint64_t &t_far) {
    constexpr uint32_t kP = 4;  // 8 keeps more loads in flight but spills at 64 VGPRs
    const uint32_t lane = lane_id();
    const bool act = (h.a >> lane) & 1ull;
    const uint32_t sl = act ? lane : (uint32_t)__builtin_ctzll(h.a);
    const uint32_t dv = h.dv;
    uint32_t mw = h.mw, mwt = h.mwt;
    for (uint32_t base = h.eb; base < h.ee; base += 64) {
        const uint32_t cnt = min(64u, h.ee - base);
        if (base != h.eb) {
            const uint32_t i = min(base + lane, h.ee - 1);
            mw = e.col[i];
            mwt = e.wt[i];
        }
        // row words of edges q .. q + kP - 1 (indexes past cnt - 1 repeat the last edge)
        auto load = [&](uint32_t q, uint32_t (&old)[kP]) {
#pragma unroll
            for (uint32_t t = 0; t < kP; t++)
                old[t] = e.dist[(uint64_t)read_lane(mw, min(q + t, cnt - 1)) * 64 + sl];
        };
        auto relax = [&](uint32_t q, const uint32_t (&old)[kP]) {
#pragma unroll
            for (uint32_t t = 0; t < kP; t++) {
                const uint32_t i = min(q + t, cnt - 1);
                const uint32_t w = read_lane(mw, i);
                const uint32_t nd = dv + read_lane(mwt, i);
                const bool imp = act && q + t < cnt && nd < old[t];
                if (imp) atomicMin(&e.dist[(uint64_t)w * 64 + lane], nd);
                const unsigned long long mn = ballot(imp && nd < thr);
                const unsigned long long mf = ballot(imp && nd >= thr);
                if (lane == 0) {
                    if (mn) atomicOr(&nxt[w], mn);
                    if (mf) atomicOr(&e.far[w], mf);
                }
                t_nf |= mn != 0;
                t_far |= mf != 0;
            }
        };
        // two buffers in turn, so no loaded value is copied (a copy would
        // wait for its load)
        uint32_t ra[kP], rb[kP];
        load(0, ra);
        for (uint32_t q = 0;; q += 2 * kP) {
            if (q + kP >= cnt) {
                relax(q, ra);
                break;
            }
            load(q + kP, rb);
            relax(q, ra);
            if (q + 2 * kP >= cnt) {
                relax(q + kP, rb);
                break;
            }
            load(q + 2 * kP, ra);
            relax(q + kP, rb);
        }
    }
}

// Whether round R pulls (the push kernel then leaves the expansion to
// k_mssp_pull): an expand round of a graph with in-pairs, learned dense in
// the previous batch, whose previous round expanded enough edges (munf of
// the input entry). A graph's first batch has nothing learned: its round R
// pulls while round R - 1 was dense and expanded no fewer edges than round
// R - 2 (entry R - 1), so the tail after the dense rounds pushes again.
__device__ __forceinline__ bool pull_round(const BatchEnv &e, uint32_t R, const BSum &s) {
    if (!e.cwin || s.nf == 0 || R >= kBRing) return false;
    if (e.learned) return e.pmode[R] && s.munf >= e.pull_prev;
    if (s.munf < e.pull_dense || R < 2) return false;
    const BCtr *c = e.ring + (R - 1) % kBRing;
    unsigned long long before = 0;
    for (int i = 0; i < (int)kShards; i++) before += c->munf[i];
    return s.munf >= before;
}

// The loads a pull chunk needs before its rows: the vertex's distance in
// this lane and the first 64 in-edge pairs (indexes clamped into the chunk).
struct PullHead {
    uint32_t w, eb, ee;
    uint32_t dw;
    uint32_t mu, mwt;
};

__device__ __forceinline__ PullHead pull_head(const BatchEnv &e, uint32_t w, uint32_t eb, uint32_t ee) {
    const uint32_t lane = lane_id();
    PullHead h;
    h.w = uniform(w);
    h.eb = uniform(eb);
    h.ee = uniform(ee);
    h.dw = e.dist[(uint64_t)h.w * 64 + lane];
    const uint2 pr = e.cwin[min(h.eb + lane, h.ee - 1)];
    h.mu = pr.x;
    h.mwt = pr.y;
    return h;
}

// Pulls one chunk [eb, ee) of w's in-edges: lane k takes
// m = min(dist[w][k], dist[u][k] + weight) over the chunk's reached u and, if
// m is lower, applies it with one atomicMin and sets w's frontier bit (far
// bit at or above thr). Rows are read in groups of kP edges with the next
// group in flight, as in relax_chunk. Pulls run only where no path plus an
// edge can reach the sentinel, so no candidate wraps.
__device__ __forceinline__ void pull_chunk(const BatchEnv &e, const PullHead &h, uint32_t thr,
                                           unsigned long long *nxt, uint64_t &t_nf, uint64_t &t_far) {
    constexpr uint32_t kP = 4;
    const uint32_t lane = lane_id();
    uint32_t m = h.dw;
    uint32_t mu = h.mu, mwt = h.mwt;
    for (uint32_t base = h.eb; base < h.ee; base += 64) {
        const uint32_t cnt = min(64u, h.ee - base);
        if (base != h.eb) {
            const uint2 pr = e.cwin[min(base + lane, h.ee - 1)];
            mu = pr.x;
            mwt = pr.y;
        }
        auto load = [&](uint32_t q, uint32_t (&r)[kP]) {
#pragma unroll
            for (uint32_t t = 0; t < kP; t++)
                r[t] = e.dist[(uint64_t)read_lane(mu, min(q + t, cnt - 1)) * 64 + lane];
        };
        auto take = [&](uint32_t q, const uint32_t (&r)[kP]) {
#pragma unroll
            for (uint32_t t = 0; t < kP; t++) {
                const uint32_t c = r[t] + read_lane(mwt, min(q + t, cnt - 1));
                if (q + t < cnt && r[t] < e.sentinel && c < m) m = c;
            }
        };
        uint32_t ra[kP], rb[kP];
        load(0, ra);
        for (uint32_t q = 0;; q += 2 * kP) {
            if (q + kP >= cnt) {
                take(q, ra);
                break;
            }
            load(q + kP, rb);
            take(q, ra);
            if (q + 2 * kP >= cnt) {
                take(q + kP, rb);
                break;
            }
            load(q + 2 * kP, ra);
            take(q + kP, rb);
        }
    }
    const bool imp = m < h.dw;
    if (imp) atomicMin(&e.dist[(uint64_t)h.w * 64 + lane], m);
    const unsigned long long mn = ballot(imp && m < thr);
    const unsigned long long mf = ballot(imp && m >= thr);
    if (lane == 0) {
        if (mn) atomicOr(&nxt[h.w], mn);
        if (mf) atomicOr(&e.far[h.w], mf);
    }
    t_nf |= mn != 0;
    t_far |= mf != 0;
}

// One pull round (after k_mssp_round of the same round, which did the
// round's resets): every chunk of the static chunk list is pulled, the next
// chunk's head loads issued while the current one's rows are read. The
// round's edge count (munf of the output entry) is that of the chunks whose
// vertex is in the frontier, the edges a push round would have expanded.
__global__ __launch_bounds__(kBlock, 8) void k_mssp_pull(BatchEnv e, uint32_t j) {
    const uint32_t R = e.p->base + j;
    const BSum s = bsum(e.ring + R % kBRing);
    if (!pull_round(e, R, s)) return;
    BCtr *out = e.ring + (R + 1) % kBRing;
    const uint32_t thr = e.ring[R % kBRing].thr;
    const unsigned long long *cur = rot(e, R);
    unsigned long long *nxt = rot(e, R + 1);
    const uint32_t lane = lane_id();
    const uint64_t gw = (uint64_t)blockIdx.x * kWavesPerBlock + (threadIdx.x >> 6);
    const uint64_t nw = (uint64_t)gridDim.x * kWavesPerBlock;
    uint64_t t_nf = 0, t_far = s.far ? 1 : 0, t_vis = 0;
    for (uint64_t i0 = 0; gw + nw * i0 < e.nch; i0 += 64) {
        const uint64_t c = gw + nw * (i0 + lane);
        uint32_t v = 0, cb = 0, ce = 0;
        if (c < e.nch) {
            v = e.ch_v[c];
            cb = e.ch_b[c];
            ce = (uint32_t)min((uint64_t)e.off[v + 1], (uint64_t)cb + e.chunk);
            if (cur[v]) t_vis += ce - cb;
        }
        uint64_t mm = ballot(c < e.nch);
        auto head = [&]() {
            const uint32_t q = lowest_lane(mm);
            mm &= mm - 1;
            return pull_head(e, read_lane(v, q), read_lane(cb, q), read_lane(ce, q));
        };
        PullHead ha = head();
        for (;;) {
            const bool more = mm != 0;
            PullHead hb = ha;
            if (more) hb = head();
            pull_chunk(e, ha, thr, nxt, t_nf, t_far);
            if (!more) break;
            ha = hb;
        }
    }
    bctr_add(out, t_nf, 0, t_vis, t_far);
}

// One SSSP round: expand the frontier masks (near), or, when none is left,
// raise the threshold by delta and move far bits below it into the frontier.
__global__ __launch_bounds__(kBlock, 8) void k_mssp_round(BatchEnv e, uint32_t j) {
    const uint32_t R = e.p->base + j;
    BCtr *in = e.ring + R % kBRing;
    BCtr *out = e.ring + (R + 1) % kBRing;
    if (blockIdx.x == 0) bctr_clear(e.ring + (R + 2) % kBRing);
    const BSum s = bsum(in);
    // the previous round's edge visits (munf of its output entry)
    if (blockIdx.x == 0 && threadIdx.x == 0 && s.munf) atomicAdd(&e.p->visits, s.munf);
    const bool expand = s.nf > 0;
    if (!expand && !s.far) return;
    const uint32_t thr = in->thr;
    const uint32_t thr2 = thr > 0xFFFFFFFFu - e.delta ? 0xFFFFFFFFu : thr + e.delta;
    if (blockIdx.x == 0 && threadIdx.x == 0) out->thr = expand ? thr : thr2;
    const unsigned long long *cur = rot(e, R);
    unsigned long long *nxt = rot(e, R + 1);
    unsigned long long *clr = rot(e, R + 2);
    const uint32_t lane = lane_id();
    const uint32_t wib = threadIdx.x >> 6;
    const uint32_t gw = blockIdx.x * kWavesPerBlock + wib;
    const uint32_t nw = gridDim.x * kWavesPerBlock;
    const uint64_t gt = (uint64_t)blockIdx.x * kBlock + threadIdx.x;
    const uint64_t nt = (uint64_t)gridDim.x * kBlock;
    for (uint64_t v = gt; v < e.nva; v += nt) clr[v] = 0;
    uint64_t t_nf = 0, t_mf = 0, t_far = 0, t_vis = 0;
    if (expand && pull_round(e, R, s)) {
        // k_mssp_pull expands this round
    } else if (expand) {
        if (s.far) t_far = 1;  // the far set is untouched by an expand round
        // static chunks of at most e.chunk edges, dealt round-robin to the
        // waves (chunk gw + nw * i); lanes test 64 chunks' masks at once and
        // the wave relaxes each active chunk
        const uint64_t nch = e.nch;
        for (uint64_t i0 = 0; (uint64_t)gw + (uint64_t)nw * i0 < nch; i0 += 64) {
            const uint64_t c = (uint64_t)gw + (uint64_t)nw * (i0 + lane);
            uint32_t v = 0, cb = 0;
            unsigned long long a = 0;
            if (c < nch) {
                v = e.ch_v[c];
                cb = e.ch_b[c];
                a = cur[v];
            }
            // chunk ends for the 64 chunks at once; each active chunk's head
            // loads go out while the one before it relaxes
            uint32_t ce = 0;
            if (a) ce = (uint32_t)min((uint64_t)e.off[v + 1], (uint64_t)cb + e.chunk);
            uint64_t mm = ballot(a != 0);
            if (!mm) continue;
            auto head = [&]() {
                const uint32_t q = lowest_lane(mm);
                mm &= mm - 1;
                return chunk_head(e, read_lane(v, q), read_lane(cb, q), read_lane(ce, q), read_lane64(a, q));
            };
            ChunkHead ha = head();
            for (;;) {
                const bool more = mm != 0;
                ChunkHead hb = ha;
                if (more) hb = head();
                relax_chunk(e, ha, thr, nxt, t_nf, t_far);
                if (lane == 0) t_vis += ha.ee - ha.eb;
                if (!more) break;
                ha = hb;
            }
        }
    } else {
        // split: far bits with thr <= dist < thr2 move to the frontier; bits
        // below thr were improved again and expanded since, and are dropped
        for (uint64_t v = gt; v < e.nva; v += nt) {
            const unsigned long long fm = e.far[v];
            if (!fm) continue;
            unsigned long long move = 0, keep = 0;
            unsigned long long m = fm;
            while (m) {
                const uint32_t k = (uint32_t)__builtin_ctzll(m);
                m &= m - 1;
                const uint32_t d = e.dist[v * 64 + k];
                if (d >= thr2) keep |= 1ull << k;
                else if (d >= thr) move |= 1ull << k;
            }
            e.far[v] = keep;
            if (move) {
                nxt[v] = move;
                t_nf++;
                t_mf += e.off[v + 1] - e.off[v];
            }
            if (keep) t_far = 1;
        }
    }
    bctr_add(out, t_nf, t_mf, t_vis, t_far);
}

// Seeds round 0: dist[s][k] = 1 and bit k of the frontier mask of s.
__global__ __launch_bounds__(kBlock) void k_mssp_seed(BatchEnv e, const uint32_t *srcs, uint32_t n) {
    __shared__ uint32_t s_first[kBatchMax];
    bctr_clear(e.ring);
    bctr_clear(e.ring + 1);
    const uint32_t t = threadIdx.x;
    __syncthreads();
    if (t < n) {
        const uint32_t s = srcs[t];
        e.dist[(uint64_t)s * 64 + t] = 1u;
        atomicOr(&e.f0[s], 1ull << t);
        bool first = true;
        for (uint32_t q = 0; q < t; q++) first = first && srcs[q] != s;
        s_first[t] = first;
    }
    __syncthreads();
    if (t == 0) {
        unsigned long long nf = 0;
        for (uint32_t q = 0; q < n; q++) nf += s_first[q];
        e.ring[0].nf[0] = nf;
        e.ring[0].thr = 1u + e.delta;
        e.p->visits = 0;
        e.p->n = n;
        e.p->base = 0;
        e.p->mask = n == 64 ? ~0ull : ((1ull << n) - 1);
    }
}

// out[k * V + c] = dist[perm[c]][k]: 64 caller ids per wave, rows read 256
// bytes at a time and transposed through LDS. Counts reached pairs and their
// out-degree sum (the sentinel marks unreached).
__global__ __launch_bounds__(kBlock) void k_mssp_output(BatchEnv e, const uint32_t *perm, uint32_t sentinel,
                                                        uint32_t *out, unsigned long long *acc) {
    __shared__ uint32_t tile[kWavesPerBlock][kBatchMax][33];
    const uint32_t n = e.p->n;
    const uint32_t lane = lane_id(), wib = threadIdx.x >> 6;
    const uint32_t gw = blockIdx.x * kWavesPerBlock + wib;
    const uint32_t nw = gridDim.x * kWavesPerBlock;
    uint64_t pairs = 0, edges = 0;
    for (uint64_t base = (uint64_t)gw * 32; base < e.V; base += (uint64_t)nw * 32) {
        const uint64_t c = base + (lane & 31);
        const uint32_t dv = c < e.V ? (perm ? perm[c] : (uint32_t)c) : 0u;
        const uint32_t deg = c < e.V ? e.off[dv + 1] - e.off[dv] : 0u;
        const uint32_t cnt = (uint32_t)min((uint64_t)32, e.V - base);
        for (uint32_t q = 0; q < cnt; q++) {
            const uint32_t vq = read_lane(dv, q);
            const uint32_t dq = read_lane(deg, q);
            const uint32_t d = lane < n ? e.dist[(uint64_t)vq * 64 + lane] : sentinel;
            tile[wib][lane][q] = d;
            const uint32_t np = (uint32_t)__popcll(ballot(lane < n && d != sentinel));
            if (lane == 0) {
                pairs += np;
                edges += (uint64_t)np * dq;
            }
        }
        // two rows per store: lanes 0-31 row k, lanes 32-63 row k + 1
        for (uint32_t k = 0; k < n; k += 2) {
            const uint32_t kk = k + (lane >> 5);
            if (kk < n && c < e.V) out[(uint64_t)kk * e.V + c] = tile[wib][kk][lane & 31];
        }
    }
    bctr_add(reinterpret_cast<BCtr *>(acc), pairs, edges, 0, 0);
}

// ---------------------------------------------------------------------------
// Host side
// ---------------------------------------------------------------------------

static uint32_t active_vertices(const ptg_graph *g, int sym) {
    // relabeling orders vertices by out-degree, so the degree-0 ones are last;
    // on a symmetric graph they have no in-edges either and are never reached
    return (g->relabeled && sym) ? g->V - g->out.bins.n_zero : g->V;
}

static int ensure_batch(ptg_graph *g, bool sssp) {
    int rc;
    if (!g->batch) g->batch = new BatchState();
    BatchState *b = g->batch;
    const uint64_t V = g->V;
    if (!b->ring) {
        if ((rc = galloc(g, (void **)&b->seen, 8 * V)) || (rc = galloc(g, (void **)&b->f[0], 8 * V)) ||
            (rc = galloc(g, (void **)&b->f[1], 8 * V)) || (rc = galloc(g, (void **)&b->f[2], 8 * V)) ||
            (rc = galloc(g, (void **)&b->ring, sizeof(BCtr) * kBRing)) ||
            (rc = galloc(g, (void **)&b->p, sizeof(BParams))) || (rc = galloc(g, (void **)&b->srcs, 4 * kBatchMax)))
            return rc;
        if (hipHostMalloc((void **)&b->h_ctr, sizeof(BCtr), hipHostMallocDefault) != hipSuccess ||
            hipHostMalloc((void **)&b->h_p, sizeof(BParams), hipHostMallocDefault) != hipSuccess ||
            hipHostMalloc((void **)&b->h_srcs, 4 * kBatchMax, hipHostMallocDefault) != hipSuccess)
            return fail(PTG_ERR_OUT_OF_MEMORY, "pinned batch buffers");
    }
    if (!sssp && !b->lm && (rc = galloc(g, (void **)&b->lm, 8 * V * kLevelCap))) return rc;
    if (sssp) {
        if (!b->far && (rc = galloc(g, (void **)&b->far, 8 * V))) return rc;
        if (!b->pmode) {
            if ((rc = galloc(g, (void **)&b->pmode, kBRing))) return rc;
            if ((rc = copy_sync(g, b->pmode, g->batch_pull_rounds, kBRing, hipMemcpyHostToDevice))) return rc;
        }
        if (!b->h_ring && hipHostMalloc((void **)&b->h_ring, sizeof(BCtr) * kBRing, hipHostMallocDefault) != hipSuccess)
            return fail(PTG_ERR_OUT_OF_MEMORY, "pinned counter ring copy");
        if (!b->dist && (rc = galloc(g, (void **)&b->dist, 4 * V * kBatchMax))) return rc;
        const uint32_t chunk = g->cfg.batch_chunk ? g->cfg.batch_chunk : kBatchChunk;
        if (!b->ch_v || b->chunk != chunk) {
            // every adjacency list cut into chunks of `chunk` edges
            gfree(g, b->ch_v);
            gfree(g, b->ch_b);
            b->ch_v = b->ch_b = nullptr;
            std::vector<uint32_t> off(V + 1), cv, cb;
            if ((rc = copy_sync(g, off.data(), g->out.off, 4 * (V + 1), hipMemcpyDeviceToHost))) return rc;
            for (uint32_t v = 0; v < g->V; v++)
                for (uint32_t x = off[v]; x < off[v + 1]; x += std::min(chunk, off[v + 1] - x)) {
                    cv.push_back(v);
                    cb.push_back(x);
                }
            const uint64_t n = std::max<size_t>(cv.size(), 1);
            if ((rc = galloc(g, (void **)&b->ch_v, 4 * n)) || (rc = galloc(g, (void **)&b->ch_b, 4 * n))) return rc;
            b->chunk = chunk;
            if (!cv.empty()) {
                if ((rc = copy_sync(g, b->ch_v, cv.data(), 4 * cv.size(), hipMemcpyHostToDevice)) ||
                    (rc = copy_sync(g, b->ch_b, cb.data(), 4 * cb.size(), hipMemcpyHostToDevice)))
                    return rc;
            }
            b->nch = (uint32_t)cv.size();
        }
    }
    return PTG_OK;
}

// Learns which rounds of the next batch pull: round r did when round r of
// this batch expanded at least E x batch_pull edges (munf of entry r + 1).
// Rounds past the counter ring's length keep pushing.
static int learn_pull_rounds(ptg_graph *g, uint32_t rounds) {
    BatchState *b = g->batch;
    PTG_HIP(hipMemcpyAsync(b->h_ring, b->ring, sizeof(BCtr) * kBRing, hipMemcpyDeviceToHost, g->stream));
    PTG_HIP(hipStreamSynchronize(g->stream));
    const float f = g->cfg.batch_pull > 0 ? g->cfg.batch_pull : kBatchPull;
    const double dense = (double)g->E * f;
    uint8_t m[kBRing] = {};
    if (rounds + 2 < kBRing)
        for (uint32_t r = 0; r < rounds; r++) {
            uint64_t vis = 0;
            for (uint32_t i = 0; i < kShards; i++) vis += b->h_ring[r + 1].munf[i];
            m[r] = (double)vis >= dense ? 1 : 0;
        }
    g->batch_pull_learned = true;
    if (memcmp(m, g->batch_pull_rounds, kBRing) != 0) {
        memcpy(g->batch_pull_rounds, m, kBRing);
        return copy_sync(g, b->pmode, g->batch_pull_rounds, kBRing, hipMemcpyHostToDevice);
    }
    return PTG_OK;
}

static BatchEnv make_batch_env(ptg_graph *g, int sym, bool sssp) {
    BatchState *b = g->batch;
    BatchEnv e{};
    e.V = g->V;
    e.nva = active_vertices(g, sym);
    e.off = g->out.off;
    e.col = g->out.col;
    e.wt = g->d_w;
    e.bins = g->out.bins.q.with(g->out.bins.d_ctr);
    e.seen = b->seen;
    e.f0 = b->f[0];
    e.f1 = b->f[1];
    e.f2 = b->f[2];
    e.far = b->far;
    e.lm = b->lm;
    e.dist = b->dist;
    e.ch_v = b->ch_v;
    e.ch_b = b->ch_b;
    e.nch = b->nch;
    e.chunk = b->chunk;
    e.ring = b->ring;
    e.p = b->p;
    e.sd = g->cfg.small_degree ? g->cfg.small_degree : 32u;
    e.alpha = g->cfg.batch_td_to_bu > 0 ? g->cfg.batch_td_to_bu : (g->cfg.td_to_bu > 0 ? g->cfg.td_to_bu : 14.0f);
    e.delta = g->cfg.batch_delta ? g->cfg.batch_delta : 256u;
    e.symmetric = sym;
    e.grid = g->cfg.batch_grid ? g->cfg.batch_grid : 4 * g->grid;
    if (sssp && g->cfg.batch_sssp_grid) e.grid = g->cfg.batch_sssp_grid;
    if (sssp) {
        e.sentinel = g->cfg.sssp_sentinel;
        e.pmode = b->pmode;
        const float f = g->cfg.batch_pull > 0 ? g->cfg.batch_pull : kBatchPull;
        e.pull_prev = (unsigned long long)((double)g->E * f / 8.0);
        e.pull_dense = (unsigned long long)((double)g->E * f);
        e.learned = g->batch_pull_learned ? 1u : 0u;
    }
    return e;
}

// Captures one replay (kReplay levels/rounds + the base advance) into an
// executable graph on the graph's stream.
template <class Body>
static int capture(ptg_graph *g, hipGraphExec_t *exec, Body body) {
    hipGraph_t graph = nullptr;
    // relaxed: other threads (serving workers on other graphs) keep
    // allocating and copying while this stream records
    PTG_HIP(hipStreamBeginCapture(g->stream, hipStreamCaptureModeRelaxed));
    const int brc = body();
    hipError_t err = hipStreamEndCapture(g->stream, &graph);
    if (brc) {
        if (graph) (void)hipGraphDestroy(graph);
        return brc;
    }
    if (err != hipSuccess) return fail(PTG_ERR_HIP, "hipStreamEndCapture");
    err = hipGraphInstantiate(exec, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    if (err != hipSuccess) return fail(PTG_ERR_HIP, "hipGraphInstantiate");
    return PTG_OK;
}

static bool same_env(const BatchEnv &a, const BatchEnv &b) { return memcmp(&a, &b, sizeof a) == 0; }

// Replays `exec` (kReplay rounds each) until the counter entry after the last
// replayed round shows no work; returns the number of rounds run.
static int replay_until_idle(ptg_graph *g, hipGraphExec_t exec, uint32_t replay, bool sssp, uint32_t *rounds,
                             bool *overflow) {
    BatchState *b = g->batch;
    uint32_t done = 0;
    *overflow = false;
    for (;;) {
        PTG_HIP(hipGraphLaunch(exec, g->stream));
        done += replay;
        PTG_HIP(hipMemcpyAsync(b->h_ctr, b->ring + done % kBRing, sizeof(BCtr), hipMemcpyDeviceToHost, g->stream));
        PTG_HIP(hipStreamSynchronize(g->stream));
        uint64_t nf = 0, far = 0;
        for (uint32_t i = 0; i < kShards; i++) {
            nf += b->h_ctr->nf[i];
            far |= b->h_ctr->far[i];
        }
        if (b->h_ctr->overflow) *overflow = true;
        if (nf == 0 && (!sssp || far == 0)) break;
        if (*overflow) break;
        if (done > 4ull * g->V + 64) return fail(PTG_ERR_HIP, "batched search did not converge");
    }
    *rounds = done;
    return PTG_OK;
}

// Runs one batch; `out` is a device buffer of n x V words in caller ids.
static int run_batch(ptg_graph *g, bool sssp, const uint32_t *srcs, uint32_t n, uint32_t *out, ptg_stats *st) {
    int rc, sym = 0;
    if ((rc = check_symmetric(g, &sym))) return rc;
    if ((rc = ensure_batch(g, sssp))) return rc;
    BatchState *b = g->batch;
    const uint2 *cwin = nullptr;
    if (sssp && g->cfg.batch_pull >= 0 && (rc = sssp_in_pairs(g, &cwin))) return rc;
    BatchEnv env = make_batch_env(g, sym, sssp);
    env.cwin = cwin;
    hipGraphExec_t &exec = sssp ? b->sssp_exec : b->bfs_exec;
    BatchEnv &cap_env = sssp ? b->env_sssp : b->env_bfs;
    if (exec && !same_env(env, cap_env)) {
        (void)hipGraphExecDestroy(exec);
        exec = nullptr;
    }
    if (!exec) {
        const uint32_t grid = env.grid;
        rc = capture(g, &exec, [&]() -> int {
            if (sssp) {
                for (uint32_t j = 0; j < kSsspReplay; j++) {
                    k_mssp_round<<<grid, kBlock, 0, g->stream>>>(env, j);
                    if (env.cwin) k_mssp_pull<<<grid, kBlock, 0, g->stream>>>(env, j);
                }
                k_batch_advance<<<1, 64, 0, g->stream>>>(env.p, kSsspReplay);
            } else {
                for (uint32_t j = 0; j < kBfsReplay; j++) {
                    k_msbfs_expand<<<grid, kBlock, 0, g->stream>>>(env, j);
                    k_msbfs_finish<<<grid, kBlock, 0, g->stream>>>(env, j);
                }
                k_batch_advance<<<1, 64, 0, g->stream>>>(env.p, kBfsReplay);
            }
            PTG_HIP(hipGetLastError());
            return PTG_OK;
        });
        if (rc) {
            exec = nullptr;
            return rc;
        }
        cap_env = env;
    }
    for (uint32_t i = 0; i < n; i++) b->h_srcs[i] = g->relabeled ? g->h_perm[srcs[i]] : srcs[i];
    PTG_HIP(hipEventRecord(g->ev0, g->stream));
    PTG_HIP(hipMemcpyAsync(b->srcs, b->h_srcs, 4 * n, hipMemcpyHostToDevice, g->stream));
    const uint64_t V = g->V;
    if ((rc = fill_u32(g, reinterpret_cast<uint32_t *>(b->f[0]), 2 * V, 0u)) ||
        (rc = fill_u32(g, reinterpret_cast<uint32_t *>(b->f[1]), 2 * V, 0u)))
        return rc;
    if (sssp) {
        if ((rc = fill_u32(g, reinterpret_cast<uint32_t *>(b->far), 2 * V, 0u)) ||
            (rc = fill_u32(g, b->dist, V * kBatchMax, g->cfg.sssp_sentinel)))
            return rc;
        k_mssp_seed<<<1, kBlock, 0, g->stream>>>(env, b->srcs, n);
    } else {
        if ((rc = fill_u32(g, reinterpret_cast<uint32_t *>(b->seen), 2 * V, 0u)) ||
            (rc = fill_u32(g, reinterpret_cast<uint32_t *>(b->lm), 2 * V, 0u)))
            return rc;
        k_msbfs_seed<<<1, kBlock, 0, g->stream>>>(env, b->srcs, n, g->E);
    }
    PTG_HIP(hipGetLastError());
    uint32_t rounds = 0;
    bool overflow = false;
    if ((rc = replay_until_idle(g, exec, sssp ? kSsspReplay : kBfsReplay, sssp, &rounds, &overflow))) return rc;
    if (overflow) {
        // deeper than kLevelCap levels: answer each source on the single-source path
        memset(st, 0, sizeof *st);
        uint64_t edges = 0, reached = 0;
        for (uint32_t i = 0; i < n; i++) {
            ptg_stats s1;
            uint32_t *dst = out + (uint64_t)i * V;
            const uint32_t s = b->h_srcs[i];
            if ((rc = run_bfs(g, s, g->relabeled ? g->d_val : dst, &s1))) return rc;
            if (g->relabeled && (rc = unpermute(g, g->d_val, dst))) return rc;
            edges += s1.edges_traversed;
            reached += s1.reached;
            st->iterations = std::max(st->iterations, s1.iterations);
        }
        PTG_HIP(hipStreamSynchronize(g->stream));
        st->edges_traversed = edges;
        st->reached = (uint32_t)std::min<uint64_t>(reached, 0xFFFFFFFFu);
        return PTG_OK;
    }
    uint64_t visits = 0;
    if (sssp) {
        // rounds summed on the device + the last round's own count (entry `rounds`)
        for (uint32_t i = 0; i < kShards; i++) visits += b->h_ctr->munf[i];
        PTG_HIP(hipMemcpyAsync(b->h_p, b->p, sizeof(BParams), hipMemcpyDeviceToHost, g->stream));
        PTG_HIP(hipStreamSynchronize(g->stream));
        visits += b->h_p->visits;
        if (env.cwin && (rc = learn_pull_rounds(g, rounds))) return rc;
    }
    BCtr *accc = b->ring + (rounds + 3) % kBRing;  // an entry no live round uses
    const uint32_t *perm = g->relabeled ? g->d_perm : nullptr;
    if ((rc = fill_u32(g, reinterpret_cast<uint32_t *>(accc), sizeof(BCtr) / 4, 0u))) return rc;
    const uint32_t oblocks = (uint32_t)std::min<uint64_t>((V + kBlock - 1) / kBlock, g->grid);
    if (sssp)
        k_mssp_output<<<oblocks, kBlock, 0, g->stream>>>(env, perm, g->cfg.sssp_sentinel, out,
                                                         reinterpret_cast<unsigned long long *>(accc));
    else
        k_msbfs_output<<<oblocks, kBlock, 0, g->stream>>>(env, std::min(rounds + 1, kLevelCap), perm, out,
                                                          reinterpret_cast<unsigned long long *>(accc));
    PTG_HIP(hipGetLastError());
    PTG_HIP(hipEventRecord(g->ev1, g->stream));
    PTG_HIP(hipMemcpyAsync(b->h_ctr, accc, sizeof(BCtr), hipMemcpyDeviceToHost, g->stream));
    PTG_HIP(hipStreamSynchronize(g->stream));
    if (st) {
        memset(st, 0, sizeof *st);
        float ms = 0;
        PTG_HIP(hipEventElapsedTime(&ms, g->ev0, g->ev1));
        st->kernel_ns = (uint64_t)((double)ms * 1e6);
        uint64_t pairs = 0, edges = 0;
        for (uint32_t i = 0; i < kShards; i++) {
            pairs += b->h_ctr->nf[i];
            edges += b->h_ctr->mf[i];
        }
        st->reached = (uint32_t)std::min<uint64_t>(pairs, 0xFFFFFFFFu);
        st->edges_traversed = edges;
        st->iterations = rounds;
        st->frontier_sum = visits;  // SSSP: edges visited (each for every active source)
    }
    return PTG_OK;
}

int run_batch_query(ptg_graph *g, bool sssp, const uint32_t *srcs, uint32_t n, uint32_t *out, ptg_stats *st) {
    return run_batch(g, sssp, srcs, n, out, st);
}

int batch_staging(ptg_graph *g, uint32_t **out) {
    if (!g->batch) g->batch = new BatchState();
    BatchState *b = g->batch;
    if (!b->out) {
        int rc = galloc(g, (void **)&b->out, 4ull * g->V * kBatchMax);
        if (rc) return rc;
    }
    *out = b->out;
    return PTG_OK;
}

// Device bytes the next batch of this kind would allocate (state not yet
// resident), staging for n x V outputs included.
uint64_t batch_state_need(const ptg_graph *g, bool sssp) {
    const BatchState *b = g->batch;
    const uint64_t V = g->V;
    auto al = [](uint64_t x) { return (x + 255) & ~255ull; };
    uint64_t need = 0;
    if (!b || !b->ring) need += 4 * al(8 * V) + al(sizeof(BCtr) * kBRing) + 2 * al(256);
    if (!b || !b->out) need += al(4 * V * kBatchMax);
    if (!sssp && (!b || !b->lm)) need += al(8 * V * kLevelCap);
    if (sssp) {
        if (!b || !b->far) need += al(8 * V);
        if (!b || !b->dist) need += al(4 * V * kBatchMax);
        const uint32_t chunk = g->cfg.batch_chunk ? g->cfg.batch_chunk : kBatchChunk;
        if (!b || !b->ch_v || b->chunk != chunk) need += 2 * al(4 * (V + g->E / chunk + 1));
        if (!b || !b->pmode) need += al(kBRing);
        if (g->cfg.batch_pull >= 0) need += sssp_in_pairs_need(g);
    }
    return need;
}

void destroy_batch(ptg_graph *g) {
    BatchState *b = g->batch;
    if (!b) return;
    if (b->bfs_exec) (void)hipGraphExecDestroy(b->bfs_exec);
    if (b->sssp_exec) (void)hipGraphExecDestroy(b->sssp_exec);
    gfree(g, b->seen);
    for (auto *p : b->f) gfree(g, p);
    gfree(g, b->far);
    gfree(g, b->lm);
    gfree(g, b->dist);
    gfree(g, b->ch_v);
    gfree(g, b->ch_b);
    gfree(g, b->out);
    gfree(g, b->srcs);
    gfree(g, b->ring);
    gfree(g, b->p);
    if (b->h_ctr) (void)hipHostFree(b->h_ctr);
    if (b->h_srcs) (void)hipHostFree(b->h_srcs);
    if (b->h_p) (void)hipHostFree(b->h_p);
    if (b->h_ring) (void)hipHostFree(b->h_ring);
    gfree(g, b->pmode);
    delete b;
    g->batch = nullptr;
}

}  // namespace ptg
