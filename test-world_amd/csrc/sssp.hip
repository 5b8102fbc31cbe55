// SSSP (near/far worklists) and label propagation for gfx950.
//
// Relaxation keeps the reference's rule: new = val[u] + w (sssp_kernel,
// ptGraph/gpu_kernels.cu:152-172) or new = val[u] (cc_kernel,
// ptGraph/gpu_kernels.cu:136-150), a strict `<` test, then atomicMin. The
// fixed point does not depend on the order of relaxations, so the
// near/far schedule (Davidson, Baxter, Garland, Owens 2014) returns the same
// values as the reference's frontier loop: only vertices whose new value is
// below a threshold join the next frontier, the rest wait in a far set until
// the threshold has risen past them.
#include <string.h>

#include "rounds.h"

namespace ptg {

// One relaxation. Queue mode dedupes near vertices through the output
// bitmap so each is appended once; flags mode (large frontiers) stores a byte
// per improved vertex with no atomic beyond the atomicMin. Far vertices are
// marked in a byte array in both modes.
template <bool kWeighted>
struct Relax {
    const uint32_t *__restrict__ off;
    const uint32_t *__restrict__ col;
    const uint32_t *__restrict__ wt;
    uint32_t *val;
    uint32_t *fout;
    uint8_t *flout;
    uint8_t *far;
    uint32_t thr;

    __device__ __forceinline__ uint32_t vertex(uint32_t v) const { return val[v]; }

    __device__ __forceinline__ void edge(bool valid, uint32_t sv, uint64_t e, const Emit &em, Acc &acc) const {
        const uint32_t w = valid ? col[e] : 0u;
        const uint32_t nv = kWeighted ? sv + (valid ? wt[e] : 0u) : sv;
        bool improved = false;
        if (valid && nv < val[w]) improved = nv < atomicMin(&val[w], nv);
        const bool near = improved && nv < thr;
        if (improved && !near) {
            far[w] = 1;
            acc.far = 1;
        }
        if (em.mode == kModeQueue) {
            bool won = false;
            uint32_t d = 0;
            if (near) {
                const uint32_t bit = 1u << (w & 31);
                won = !(atomicOr(&fout[w >> 5], bit) & bit);
                if (won) d = off[w + 1] - off[w];
            }
            emit(em, won, w, d, acc);
        } else {
            if (near) flout[w] = 1;
            mark_any(em, near);
        }
    }
};

// Far-set split: raise the threshold to thr2 and move far vertices below it
// into the next frontier (bitmap mode, counted exactly); vertices already
// below the old threshold were processed when they got their value and are
// dropped. A wave owns 64 consecutive vertices.
__device__ __forceinline__ void sssp_split(const RoundEnv &e, const uint32_t *val, uint8_t *far, uint32_t thr,
                                           uint32_t thr2, const Emit &em, Acc &acc) {
    const uint32_t lane = lane_id();
    const uint32_t gw = blockIdx.x * kWavesPerBlock + (threadIdx.x >> 6);
    const uint32_t nw = gridDim.x * kWavesPerBlock;
    for (uint32_t p = gw; p < e.npairs; p += nw) {
        const uint32_t v = p * 64 + lane;
        const bool inb = far[v] != 0;  // the far array covers npairs * 64 bytes
        if (ballot(inb) == 0) continue;
        const uint32_t dv = inb ? val[v] : 0u;
        const bool move = inb && dv >= thr && dv < thr2;
        const bool keep = inb && dv >= thr2;
        if (inb && !keep) far[v] = 0;
        const uint32_t d = move ? e.off[v + 1] - e.off[v] : 0u;
        const uint64_t mm = ballot(move);
        if ((lane & 31) == 0) {
            const uint32_t bits = lane ? (uint32_t)(mm >> 32) : (uint32_t)mm;
            if (bits) e.fout[p * 2 + (lane >> 5)] = bits;
        }
        if (keep) acc.far = 1;
        emit(em, move, v, d, acc);
    }
}

struct RelaxParams {
    uint32_t *val;
    const uint32_t *wt;
    uint8_t *far;
    uint32_t delta;
};

// One round: expand the near frontier, or split the far set when the near
// frontier is empty. The far set is non-empty after the round when it was
// before an expand round or when anything was added or kept.
template <bool kWeighted>
__global__ __launch_bounds__(kBlock) void k_relax_round(RoundEnv e, RelaxParams p) {
    RoundCtr *in = e.ring + e.r % kRing;
    RoundCtr *out = e.ring + (e.r + 1) % kRing;
    clear_ring_ahead(e);
    const InSummary s = summarize(in);
    const bool expand = s.found > 0;
    if (!expand && !in->far_any) return;
    const uint32_t thr = in->thr;
    const uint32_t thr2 = thr > 0xFFFFFFFFu - p.delta ? 0xFFFFFFFFu : thr + p.delta;
    const uint32_t mode = !expand ? kModeBitmap : (s.edges <= e.qcap ? kModeQueue : kModeFlags);
    clear_bitmap_ahead(e);
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        out->mode = mode;
        out->thr = expand ? thr : thr2;
        if (expand && in->far_any) out->far_any = 1u;
    }
    const Emit em = make_emit(e, out, mode);
    Acc acc;
    if (expand) {
        Relax<kWeighted> op{e.off, e.col, p.wt, p.val, e.fout, e.flout, p.far, thr};
        expand_frontier(e, in, op, em, acc);
    } else {
        sssp_split(e, p.val, p.far, thr, thr2, em, acc);
    }
    if (ballot(acc.far != 0) && lane_id() == 0) out->far_any = 1u;
    acc.far = 0;
    add_tallies(out, acc);
}

template <bool kWeighted>
static int run_relax_rounds(ptg_graph *g, RelaxParams p, LaunchTimer &timer, uint32_t *iters, uint64_t *fsum) {
    return drive_rounds(
        g,
        [&](uint32_t r) -> int {
            const RoundEnv env = make_env(g, r);
            int rc = launch_prepare(g, env);
            if (rc) return rc;
            timer.launch(k_relax_round<kWeighted>, dim3(g->grid), dim3(kBlock), env, p);
            PTG_HIP(hipGetLastError());
            return PTG_OK;
        },
        [](const RoundCtr &c) { return c.any == 0 && c.far_any == 0; }, iters, fsum, nullptr);
}

int run_sssp(ptg_graph *g, uint32_t src, uint32_t *dist, ptg_stats *st) {
    int rc;
    LaunchTimer timer(g);
    const uint32_t delta = g->cfg.sssp_delta ? g->cfg.sssp_delta : 16u;
    PTG_HIP(hipEventRecord(g->ev0, g->stream));
    if ((rc = fill_u32(g, dist, g->V, g->cfg.sssp_sentinel)) ||
        (rc = fill_u32(g, reinterpret_cast<uint32_t *>(g->d_far), (uint64_t)g->npairs * 16, 0u)) ||
        (rc = reset_rounds(g)))
        return rc;
    if ((rc = launch_seed(g, src, 1u, dist, 1u + delta))) return rc;
    uint32_t iters = 0;
    uint64_t fsum = 0;
    if ((rc = run_relax_rounds<true>(g, RelaxParams{dist, g->d_w, g->d_far, delta}, timer, &iters, &fsum)))
        return rc;
    PTG_HIP(hipEventRecord(g->ev1, g->stream));
    PTG_HIP(hipEventSynchronize(g->ev1));
    if (st) {
        memset(st, 0, sizeof *st);
        float ms = 0;
        PTG_HIP(hipEventElapsedTime(&ms, g->ev0, g->ev1));
        st->kernel_ns = (uint64_t)((double)ms * 1e6);
        st->iterations = iters;
        st->frontier_sum = fsum;
        if ((rc = timer.collect(st))) return rc;
        if ((rc = reach_stats(g, dist, g->cfg.sssp_sentinel, &st->reached, &st->edges_traversed))) return rc;
    }
    return PTG_OK;
}

// Round 0 of label propagation: every vertex with out-degree > 0 is in the
// frontier (bitmap mode, counted here; the pre-pass queues heavy vertices).
__global__ __launch_bounds__(kBlock) void k_seed_all(RoundEnv e, uint32_t n_nonzero, unsigned long long E) {
    const uint32_t lane = lane_id();
    const uint32_t gw = blockIdx.x * kWavesPerBlock + (threadIdx.x >> 6);
    const uint32_t nw = gridDim.x * kWavesPerBlock;
    for (uint32_t p = gw; p < e.npairs; p += nw) {
        const uint32_t v = p * 64 + lane;
        const bool nz = v < e.V && e.off[v + 1] > e.off[v];
        const uint64_t m = ballot(nz);
        if ((lane & 31) == 0) e.fin[p * 2 + (lane >> 5)] = lane ? (uint32_t)(m >> 32) : (uint32_t)m;
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        RoundCtr *c = e.ring;
        c->mode = kModeBitmap;
        c->thr = 0xFFFFFFFFu;
        c->any = n_nonzero ? 1u : 0u;
        c->found_sh[0] = n_nonzero;
        c->edges_sh[0] = E;
    }
}

// Label propagation from label[v] (already initialised) with every vertex
// active: the fixed point of ccCaculateInShare (ptGraph/main.cu:961-1035).
int run_label_propagation(ptg_graph *g, uint32_t *label, LaunchTimer &timer, uint32_t *iters, uint64_t *fsum) {
    int rc;
    if ((rc = reset_rounds(g))) return rc;
    k_seed_all<<<g->grid, kBlock, 0, g->stream>>>(make_env(g, 0), g->V - g->out.bins.n_zero, g->E);
    PTG_HIP(hipGetLastError());
    return run_relax_rounds<false>(g, RelaxParams{label, nullptr, g->d_far, 0u}, timer, iters, fsum);
}

}  // namespace ptg
