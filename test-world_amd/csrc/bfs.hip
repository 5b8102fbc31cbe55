// Level-synchronous BFS for gfx950 with direction optimisation.
//
// Output convention of the reference (bfsCaculateInShare and siblings,
// ptGraph/main.cu:1068-1076): level[src] = 1, level[v] = 1 + hops,
// unreached 0xFFFFFFFF. Every vertex of frontier L gets level L + 1 however
// many frontier neighbours discover it, which is the fixed point of the
// reference's relax-and-activate loop (bfs_kernel, ptGraph/gpu_kernels.cu:
// 118-134), so duplicate discoveries are harmless and large frontiers are
// written with plain stores only.
#include <string.h>

#include "rounds.h"

namespace ptg {

constexpr uint32_t kUnreached = 0xFFFFFFFFu;

// Top-down step: a frontier vertex's unvisited neighbours join level + 1.
// Queue mode dedupes through the output bitmap (one atomicOr per candidate)
// so each vertex is appended once; flags mode just stores level and flag.
struct BfsTopDown {
    const uint32_t *__restrict__ off;
    const uint32_t *__restrict__ col;
    uint32_t *level;
    uint32_t *fout;
    uint8_t *flout;
    uint32_t next_level;

    __device__ __forceinline__ uint32_t vertex(uint32_t) const { return 0; }

    struct Pre {
        uint32_t w, lv;
        bool valid;
    };

    __device__ __forceinline__ Pre load(bool valid, uint32_t, uint64_t e) const {
        Pre p;
        p.valid = valid;
        p.w = valid ? col[e] : 0u;
        p.lv = valid ? level[p.w] : 0u;
        return p;
    }

    __device__ __forceinline__ void apply(const Pre &p, const Emit &em, Acc &acc) const {
        const uint32_t w = p.w;
        const bool fresh = p.valid && p.lv == kUnreached;
        if (em.mode == kModeQueue) {
            const uint32_t bit = 1u << (w & 31);
            bool won = false;
            uint32_t d = 0;
            if (fresh) won = !(atomicOr(&fout[w >> 5], bit) & bit);
            if (won) {
                level[w] = next_level;
                d = off[w + 1] - off[w];
            }
            emit(em, won, w, d, acc);
        } else {
            if (fresh) {
                level[w] = next_level;
                flout[w] = 1;
            }
            mark_any(em, fresh);
        }
    }
};

// Bottom-up step (symmetric graphs only): every unvisited vertex looks for a
// neighbour in the current frontier bitmap and stops at the first one. A wave
// owns 64 consecutive vertices = two bitmap words, so it writes its words
// without atomics. Lists longer than small_degree are scanned by the whole
// wave, 64 neighbours per step.
__device__ __forceinline__ void bfs_bottom_up(const RoundEnv &e, uint32_t *level, uint32_t next_level,
                                              const Emit &em, Acc &acc) {
    const uint32_t lane = lane_id();
    const uint32_t gw = blockIdx.x * kWavesPerBlock + (threadIdx.x >> 6);
    const uint32_t nw = gridDim.x * kWavesPerBlock;
    for (uint32_t p = gw; p < e.npairs; p += nw) {
        const uint32_t v = p * 64 + lane;
        const bool todo = v < e.V && level[v] == kUnreached;
        if (ballot(todo) == 0) continue;
        uint32_t b = 0, d = 0;
        if (todo) {
            b = e.off[v];
            d = e.off[v + 1] - b;
        }
        bool hit = false;
        uint64_t mm = ballot(todo && d > e.sd);
        while (mm) {
            const uint32_t j = lowest_lane(mm);
            mm &= mm - 1;
            const uint32_t bj = read_lane(b, j), dj = read_lane(d, j);
            for (uint32_t k = 0; k < dj; k += 64) {
                bool h = false;
                if (k + lane < dj) {
                    const uint32_t u = e.col[(uint64_t)bj + k + lane];
                    h = (e.fin[u >> 5] >> (u & 31)) & 1u;
                }
                if (ballot(h)) {
                    if (lane == j) hit = true;
                    break;
                }
            }
        }
        if (todo && d > 0 && d <= e.sd) {
            for (uint32_t k = 0; k < d; k += 4) {
                uint32_t u[4];
#pragma unroll
                for (int t = 0; t < 4; t++) u[t] = (k + t < d) ? e.col[(uint64_t)b + k + t] : kUnreached;
                bool h = false;
#pragma unroll
                for (int t = 0; t < 4; t++)
                    if (u[t] != kUnreached) h = h || ((e.fin[u[t] >> 5] >> (u[t] & 31)) & 1u);
                if (h) {
                    hit = true;
                    break;
                }
            }
        }
        const uint64_t m = ballot(hit);
        if ((lane & 31) == 0) {
            const uint32_t bits = lane ? (uint32_t)(m >> 32) : (uint32_t)m;
            if (bits) e.fout[p * 2 + (lane >> 5)] = bits;
        }
        if (hit) level[v] = next_level;
        emit(em, hit, v, d, acc);
    }
}

struct BfsParams {
    uint32_t *level;
    uint64_t E;
    float alpha, beta;
    int symmetric;
};

// One BFS level. Direction (Beamer, Asanovic, Patterson 2012): switch to
// bottom-up when the frontier's edges exceed the unexplored edges / alpha,
// back to top-down when the frontier has fewer than V / beta vertices.
__global__ __launch_bounds__(kBlock, 8) void k_bfs_round(RoundEnv e, BfsParams p) {
    RoundCtr *in = e.ring + e.r % kRing;
    RoundCtr *out = e.ring + (e.r + 1) % kRing;
    clear_ring_ahead(e);
    const InSummary s = summarize(in);
    if (s.found == 0) return;
    const uint64_t seen = in->seen + s.edges;
    const double m_u = (double)(p.E > seen ? p.E - seen : 0);
    uint32_t dir = in->dir;
    if (p.symmetric) {
        if (dir == kDirTopDown && (double)s.edges > m_u / p.alpha) dir = kDirBottomUp;
        else if (dir == kDirBottomUp && bfs_leaves_bottom_up(s.found, e.V, p.beta)) dir = kDirTopDown;
    }
    const uint32_t mode = dir == kDirBottomUp ? kModeBitmap : (s.edges <= e.qcap ? kModeQueue : kModeFlags);
    clear_bitmap_ahead(e);
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        out->mode = mode;
        out->dir = dir;
        out->seen = seen;
    }
    const Emit em = make_emit(e, out, mode);
    Acc acc;
    if (dir == kDirTopDown) {
        BfsTopDown op{e.off, e.col, p.level, e.fout, e.flout, e.r + 2};
        expand_frontier(e, in, op, em, acc);
    } else {
        bfs_bottom_up(e, p.level, e.r + 2, em, acc);
    }
    add_tallies(out, acc);
}

int run_bfs(ptg_graph *g, uint32_t src, uint32_t *level, ptg_stats *st) {
    int sym = 0;
    int rc = check_symmetric(g, &sym);
    if (rc) return rc;
    LaunchTimer timer(g);
    PTG_HIP(hipEventRecord(g->ev0, g->stream));
    if ((rc = reset_rounds(g, {FillSpan{level, g->V, kUnreached}}))) return rc;
    if ((rc = launch_seed(g, src, 1u, level, 0u))) return rc;
    const BfsParams p{level, g->E, g->cfg.td_to_bu > 0 ? g->cfg.td_to_bu : 14.0f,
                      g->cfg.bu_to_td > 0 ? g->cfg.bu_to_td : 24.0f, sym};
    uint32_t iters = 0;
    uint64_t reached = 0, edges = 0;
    auto finish = [&](bool speculative) -> int {
        PTG_HIP(hipEventRecord(g->ev1, g->stream));
        return deliver_output(g, speculative);
    };
    bool finished = false;
    rc = drive_rounds(
        g,
        [&](uint32_t r) -> int {
            RoundEnv env = make_env(g, r);
            env.bu_beta = p.beta;
            int rc2 = launch_prepare(g, env);
            if (rc2) return rc2;
            timer.launch(k_bfs_round, dim3(g->grid), dim3(kBlock), env, p);
            PTG_HIP(hipGetLastError());
            return PTG_OK;
        },
        [](const RoundCtr &c) { return c.any == 0; }, g->last_rounds[0], &iters, &reached, &edges, finish,
        &finished);
    if (rc) return rc;
    if (!finished && (rc = finish(false))) return rc;
    PTG_HIP(hipStreamSynchronize(g->stream));
    if (st) {
        memset(st, 0, sizeof *st);
        float ms = 0;
        PTG_HIP(hipEventElapsedTime(&ms, g->ev0, g->ev1));
        st->kernel_ns = (uint64_t)((double)ms * 1e6);
        st->iterations = iters;
        st->reached = (uint32_t)reached;
        st->edges_traversed = edges;
        st->frontier_sum = reached;
        if ((rc = timer.collect(st))) return rc;
    }
    return PTG_OK;
}

}  // namespace ptg
