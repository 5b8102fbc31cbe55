// SSSP (near/far worklists with light/heavy edges) and label propagation
// for gfx950.
//
// Relaxation keeps the reference's rule: new = val[u] + w (sssp_kernel,
// ptGraph/gpu_kernels.cu:152-172) or new = val[u] (cc_kernel,
// ptGraph/gpu_kernels.cu:136-150), a strict `<` test, then atomicMin. The
// fixed point does not depend on the order of relaxations, so the
// near/far schedule (Davidson, Baxter, Garland, Owens 2014) returns the same
// values as the reference's frontier loop: only vertices whose new value is
// below a threshold join the next frontier, the rest wait in a far set until
// the threshold has risen past them. As in delta-stepping (Meyer, Sanders
// 2003), the rounds of a bucket [lo, lo + delta) relax only light edges
// (weight < delta); when the bucket's frontier is empty its vertices' values
// are final, and one round relaxes their heavy edges once, every one of which
// lands at or above the bucket's end and so in the far set.
//
// By default there is a single unbounded bucket (every edge light, no far
// set): small vertices are relaxed as soon as they improve, and vertices of
// degree > defer_deg wait in the frontier until their value is within a
// window of the frontier's smallest value (rounds.h). On R-MAT graphs this
// scans each hub's list about once instead of once per improvement. Only the
// order of relaxations changes, so the values are the same.
#include <algorithm>
#include <vector>

#include <string.h>

#include "rounds.h"
#include "static_expand.h"

namespace ptg {

// sssp_delta 0: one unbounded bucket, no light/heavy split.
constexpr uint32_t kUnbounded = 0xFFFFFFFFu;
// Default deferral degree: vertices of degree > 64 wait for the window
// (measured best of 32..1024 on R-MAT scale 22 with weights in [1, 64]:
// 2.59 ms per query against 2.67 at 128 and 3.82 without deferral).
constexpr uint32_t kDefaultDeferDegree = 64;

// One relaxation. A candidate (new value below the target's value as read)
// issues an atomicMin whose result is not waited for, and counts as improved:
// on R-MAT graphs 99% of candidates do improve, and a vertex that joins the
// next frontier without having improved is only expanded again with its
// current value, which changes no result. Queue mode dedupes near vertices
// through the output bitmap so each is appended once; flags mode (large
// frontiers) stores a byte per improved vertex. Far vertices are marked in a
// byte array in both modes. kPacked reads each edge's column and weight as
// one 8-byte pair (cw) instead of two 4-byte loads.
template <bool kWeighted, bool kPacked = false>
struct Relax {
    const uint32_t *__restrict__ off;
    const uint32_t *__restrict__ col;
    const uint32_t *__restrict__ wt;
    const uint2 *__restrict__ cw;
    uint32_t *val;
    uint32_t *fout;
    uint8_t *flout;
    uint8_t *far;
    uint32_t thr;
    uint32_t defer_deg, lim;  // vertices of degree > defer_deg and value >= lim wait

    __device__ __forceinline__ uint32_t vertex(uint32_t v) const { return val[v]; }

    __device__ __forceinline__ bool defers(uint32_t d, uint32_t sv) const { return d > defer_deg && sv >= lim; }

    // Puts a deferred vertex (value sv, degree d) back into the output
    // frontier for lanes with pred. Wave-uniform.
    __device__ __forceinline__ void requeue(bool pred, uint32_t v, uint32_t d, uint32_t sv, const Emit &em,
                                            Acc &acc) const {
        if (pred) acc.fmin_inv = max(acc.fmin_inv, ~sv);
        if (em.mode == kModeQueue) {
            bool won = false;
            if (pred) {
                const uint32_t bit = 1u << (v & 31);
                won = !(atomicOr(&fout[v >> 5], bit) & bit);
            }
            emit(em, won, v, d, acc);
        } else {
            if (pred) flout[v] = 1;
            mark_any(em, pred);
        }
    }

    struct Pre {
        uint32_t w, nv, cur;
        bool valid;
    };

    // Loads only: the edge's target, its candidate value and the target's value.
    __device__ __forceinline__ Pre load(bool valid, uint32_t sv, uint64_t e) const {
        Pre p;
        p.valid = valid;
        if constexpr (kPacked) {
            const uint2 x = valid ? cw[e] : make_uint2(0u, 0u);
            p.w = x.x;
            p.nv = sv + x.y;
        } else {
            p.w = valid ? col[e] : 0u;
            p.nv = kWeighted ? sv + (valid ? wt[e] : 0u) : sv;
        }
        p.cur = valid ? val[p.w] : 0u;
        return p;
    }

    // The atomic of a candidate; no result is returned.
    __device__ __forceinline__ void issue(Pre &p) const {
        if (p.valid && p.nv < p.cur)
            (void)__hip_atomic_fetch_min(&val[p.w], p.nv, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }

    __device__ __forceinline__ void apply(const Pre &p, const Emit &em, Acc &acc) const {
        const uint32_t w = p.w, nv = p.nv;
        const bool improved = p.valid && nv < p.cur;
        const bool near = improved && nv < thr;
        if (near) acc.fmin_inv = max(acc.fmin_inv, ~nv);
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
        if (move) acc.fmin_inv = max(acc.fmin_inv, ~dv);
        emit(em, move, v, d, acc);
    }
}

struct RelaxParams {
    uint32_t *val;
    const uint32_t *wt;  // weights of the CSR the rounds expand (light edges)
    const uint2 *cw;     // (column, weight) pairs of the full CSR (one unbounded bucket)
    uint8_t *far;
    uint32_t delta;
    uint32_t split;      // 1: heavy edges live in the CSR below and get a round per bucket
    const uint32_t *hoff, *hcol, *hw;
    Frontier hbins;      // static degree bins of the heavy CSR
    uint32_t defer_deg;  // ~0u: no deferral
    uint32_t window;
};

// Heavy-edge round: every vertex whose final value lies in the bucket
// [lo, hi) relaxes its heavy edges once; each lands at or above hi, so an
// improvement only marks the far set.
struct HeavyRelax {
    const uint32_t *__restrict__ col;
    const uint32_t *__restrict__ wt;
    uint32_t *val;
    uint8_t *far;
    uint32_t lo, hi;
    __device__ __forceinline__ uint32_t vertex(uint32_t v) const { return val[v]; }
    __device__ __forceinline__ bool active(uint32_t d) const { return d >= lo && d < hi; }
    struct Pre {
        uint32_t w, nv, cur;
        bool valid;
    };
    __device__ __forceinline__ Pre load(bool valid, uint32_t d, uint64_t e) const {
        Pre p;
        p.valid = valid;
        p.w = valid ? col[e] : 0u;
        p.nv = valid ? d + wt[e] : 0u;
        p.cur = valid ? val[p.w] : 0u;
        return p;
    }
    __device__ __forceinline__ void apply(const Pre &p) const {
        if (p.valid && p.nv < p.cur && p.nv < atomicMin(&val[p.w], p.nv)) far[p.w] = 1;
    }
};

// One round: expand the near frontier; when it is empty, relax the heavy
// edges of the bucket just finished (once per bucket); then split the far set.
// The far set is non-empty after the round when it was before an expand
// round, when anything was added or kept, and (conservatively) after a heavy
// round. kBuckets = false (one unbounded bucket, no split: the far set stays
// empty) compiles the expand round only, with fewer registers, and reads
// packed (column, weight) pairs. Eight waves per SIMD keep the whole
// persistent grid (8 workgroups per CU) resident.
template <bool kWeighted, bool kBuckets>
__global__ __launch_bounds__(kBlock, 8) void k_relax_round(RoundEnv e, RelaxParams p) {
    RoundCtr *in = e.ring + e.r % kRing;
    RoundCtr *out = e.ring + (e.r + 1) % kRing;
    clear_ring_ahead(e);
    const InSummary s = summarize(in);
    const bool expand = s.found > 0;
    const bool heavy = kBuckets && !expand && p.split && in->need_heavy;
    if (!expand && !heavy && !(kBuckets && in->far_any)) return;
    const uint32_t thr = in->thr, lo = in->lo;
    const uint32_t thr2 = thr > 0xFFFFFFFFu - p.delta ? 0xFFFFFFFFu : thr + p.delta;
    const uint32_t mode = !expand ? kModeBitmap : (s.edges <= e.qcap ? kModeQueue : kModeFlags);
    clear_bitmap_ahead(e);
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        out->mode = mode;
        if (expand || heavy) {
            out->thr = thr;
            out->lo = lo;
            out->need_heavy = expand ? p.split : 0u;
            if (in->far_any || heavy) out->far_any = 1u;
        } else {
            out->thr = thr2;
            out->lo = thr;
            out->need_heavy = 0u;
        }
    }
    const Emit em = make_emit(e, out, mode);
    Acc acc;
    if (expand) {
        const uint32_t lim = p.defer_deg != 0xFFFFFFFFu ? defer_limit(in, p.window) : 0xFFFFFFFFu;
        Relax<kWeighted, kWeighted && !kBuckets> op{e.off,   e.col,   p.wt, p.cw,  p.val,       e.fout,
                                                     e.flout, p.far,   thr,  p.defer_deg, lim};
        expand_frontier(e, in, op, em, acc);
        if (lim != 0xFFFFFFFFu) requeue_deferred(e, in, op, em, acc);
    } else if constexpr (kBuckets) {
        if (heavy)
            expand_static(p.hbins, p.hoff, HeavyRelax{p.hcol, p.hw, p.val, p.far, lo, thr});
        else
            sssp_split(e, p.val, p.far, thr, thr2, em, acc);
    }
    if (ballot(acc.far != 0) && lane_id() == 0) out->far_any = 1u;
    acc.far = 0;
    add_tallies(out, acc);
}

template <bool kWeighted, class Finish>
static int run_relax_rounds(ptg_graph *g, RelaxParams p, const uint32_t *off, const uint32_t *col, LaunchTimer &timer,
                            uint32_t *iters, uint64_t *fsum, Finish finish, bool *finished) {
    return drive_rounds(
        g,
        [&](uint32_t r) -> int {
            RoundEnv env = make_env_csr(g, r, off, col);
            env.val = p.val;
            env.defer_deg = p.defer_deg;
            env.window = p.window;
            int rc = launch_prepare(g, env);
            if (rc) return rc;
            if (p.split)
                timer.launch(k_relax_round<kWeighted, true>, dim3(g->grid), dim3(kBlock), env, p);
            else
                timer.launch(k_relax_round<kWeighted, false>, dim3(g->grid), dim3(kBlock), env, p);
            PTG_HIP(hipGetLastError());
            return PTG_OK;
        },
        [](const RoundCtr &c) { return c.any == 0 && c.far_any == 0 && c.need_heavy == 0; },
        g->last_rounds[kWeighted ? 1 : 2], iters, fsum, nullptr, finish, finished);
}

// Counts the light edges (weight < delta) of every vertex: one wave each.
__global__ __launch_bounds__(kBlock) void k_count_light(uint32_t V, const uint32_t *__restrict__ off,
                                                        const uint32_t *__restrict__ w, uint32_t delta,
                                                        uint32_t *__restrict__ cnt) {
    const uint32_t lane = lane_id();
    const uint32_t gw = blockIdx.x * kWavesPerBlock + (threadIdx.x >> 6);
    const uint32_t nw = gridDim.x * kWavesPerBlock;
    for (uint32_t v = gw; v < V; v += nw) {
        uint32_t c = 0;
        for (uint64_t i = (uint64_t)off[v] + lane; i < off[v + 1]; i += 64) c += w[i] < delta;
        c = (uint32_t)wave_sum_u64(c);
        if (lane == 0) cnt[v] = c;
    }
}

// Splits every adjacency list into its light and heavy edges, each part in
// the list's order: one wave per vertex, ballot ranks for the positions.
__global__ __launch_bounds__(kBlock) void k_split_edges(uint32_t V, const uint32_t *__restrict__ off,
                                                        const uint32_t *__restrict__ col,
                                                        const uint32_t *__restrict__ w, uint32_t delta,
                                                        const uint32_t *__restrict__ loff,
                                                        const uint32_t *__restrict__ hoff, uint32_t *__restrict__ lcol,
                                                        uint32_t *__restrict__ lw, uint32_t *__restrict__ hcol,
                                                        uint32_t *__restrict__ hw) {
    const uint32_t lane = lane_id();
    const uint32_t gw = blockIdx.x * kWavesPerBlock + (threadIdx.x >> 6);
    const uint32_t nw = gridDim.x * kWavesPerBlock;
    for (uint32_t v = gw; v < V; v += nw) {
        uint64_t pl = loff[v], ph = hoff[v];
        const uint64_t end = off[v + 1];
        for (uint64_t base = off[v]; base < end; base += 64) {
            const uint64_t i = base + lane;
            const bool valid = i < end;
            const uint32_t c = valid ? col[i] : 0u, x = valid ? w[i] : 0u;
            const bool light = valid && x < delta, hv = valid && !light;
            const uint64_t ml = ballot(light), mh = ballot(hv);
            if (light) {
                lcol[pl + rank_below(ml)] = c;
                lw[pl + rank_below(ml)] = x;
            }
            if (hv) {
                hcol[ph + rank_below(mh)] = c;
                hw[ph + rank_below(mh)] = x;
            }
            pl += __popcll(ml);
            ph += __popcll(mh);
        }
    }
}

static void free_split(ptg_graph *g) {
    gfree(g, g->light.off);
    gfree(g, g->light.col);
    gfree(g, g->heavy.off);
    gfree(g, g->heavy.col);
    gfree(g, g->d_wl);
    gfree(g, g->d_wh);
    free_static_bins(g, &g->heavy.bins);
    g->light = Csr{};
    g->heavy = Csr{};
    g->d_wl = g->d_wh = nullptr;
    g->split_delta = 0;
}

// Builds (once per delta) the light and heavy CSRs of the weighted graph.
static int ensure_split(ptg_graph *g, uint32_t delta) {
    if (g->split_delta == delta && g->light.off) return PTG_OK;
    free_split(g);
    const uint32_t V = g->V;
    int rc;
    uint32_t *cnt = nullptr;
    if ((rc = galloc(g, (void **)&cnt, 4ull * V))) return rc;
    k_count_light<<<g->grid, kBlock, 0, g->stream>>>(V, g->out.off, g->d_w, delta, cnt);
    std::vector<uint32_t> c(V), off((size_t)V + 1), lo((size_t)V + 1), ho((size_t)V + 1);
    rc = copy_sync(g, c.data(), cnt, 4ull * V, hipMemcpyDeviceToHost);
    if (!rc) rc = copy_sync(g, off.data(), g->out.off, 4ull * (V + 1), hipMemcpyDeviceToHost);
    gfree(g, cnt);
    if (rc) return rc;
    lo[0] = ho[0] = 0;
    for (uint32_t v = 0; v < V; v++) {
        lo[v + 1] = lo[v] + c[v];
        ho[v + 1] = ho[v] + (off[v + 1] - off[v] - c[v]);
    }
    const uint64_t El = lo[V], Eh = ho[V];
    if ((rc = galloc(g, (void **)&g->light.off, 4ull * (V + 1))) ||
        (rc = galloc(g, (void **)&g->heavy.off, 4ull * (V + 1))) ||
        (rc = galloc(g, (void **)&g->light.col, 4 * std::max<uint64_t>(El, 1))) ||
        (rc = galloc(g, (void **)&g->d_wl, 4 * std::max<uint64_t>(El, 1))) ||
        (rc = galloc(g, (void **)&g->heavy.col, 4 * std::max<uint64_t>(Eh, 1))) ||
        (rc = galloc(g, (void **)&g->d_wh, 4 * std::max<uint64_t>(Eh, 1))) ||
        (rc = copy_sync(g, g->light.off, lo.data(), 4ull * (V + 1), hipMemcpyHostToDevice)) ||
        (rc = copy_sync(g, g->heavy.off, ho.data(), 4ull * (V + 1), hipMemcpyHostToDevice))) {
        free_split(g);
        return rc;
    }
    k_split_edges<<<g->grid, kBlock, 0, g->stream>>>(V, g->out.off, g->out.col, g->d_w, delta, g->light.off,
                                                     g->heavy.off, g->light.col, g->d_wl, g->heavy.col, g->d_wh);
    PTG_HIP(hipGetLastError());
    if ((rc = build_static_bins(g, ho.data(), V, &g->heavy.bins))) {
        free_split(g);
        return rc;
    }
    g->split_delta = delta;
    return PTG_OK;
}

void release_sssp_split(ptg_graph *g) {
    free_split(g);
    gfree(g, g->d_defer);
    g->d_defer = nullptr;
    gfree(g, g->d_cw);
    g->d_cw = nullptr;
}

uint64_t sssp_split_need(const ptg_graph *g) {
    if (!g->weighted) return 0;
    uint64_t need = g->d_defer ? 0 : 4ull * g->nwords;
    const uint32_t delta = g->cfg.sssp_delta ? g->cfg.sssp_delta : kUnbounded;
    if (delta == kUnbounded && !g->d_cw) need += 8ull * std::max<uint64_t>(g->E, 1);
    if (delta != kUnbounded && !(g->split_delta == delta && g->light.off))
        need += 8ull * g->E + 8ull * (g->V + 1) + 4ull * g->V + (64u << 10);
    return need;
}

// Sum of the weights, over 16 accumulators.
__global__ __launch_bounds__(kBlock) void k_weight_sum(const uint32_t *__restrict__ w, uint64_t E,
                                                       unsigned long long *acc) {
    uint64_t sum = 0;
    for (uint64_t i = (uint64_t)blockIdx.x * kBlock + threadIdx.x; i < E; i += (uint64_t)gridDim.x * kBlock) sum += w[i];
    sum = wave_sum_u64(sum);
    if (lane_id() == 0) atomicAdd(&acc[(blockIdx.x * kWavesPerBlock + (threadIdx.x >> 6)) % kShards], sum);
}

// (column, weight) of every edge side by side, for the default rounds.
__global__ __launch_bounds__(kBlock) void k_pack_edges(const uint32_t *__restrict__ col,
                                                       const uint32_t *__restrict__ w, uint64_t E,
                                                       uint2 *__restrict__ cw) {
    for (uint64_t i = (uint64_t)blockIdx.x * kBlock + threadIdx.x; i < E; i += (uint64_t)gridDim.x * kBlock)
        cw[i] = make_uint2(col[i], w[i]);
}

static int ensure_packed(ptg_graph *g) {
    if (g->d_cw) return PTG_OK;
    int rc = galloc(g, (void **)&g->d_cw, 8ull * std::max<uint64_t>(g->E, 1));
    if (rc) return rc;
    if (g->E) {
        k_pack_edges<<<g->grid, kBlock, 0, g->stream>>>(g->out.col, g->d_w, g->E, g->d_cw);
        PTG_HIP(hipGetLastError());
    }
    return PTG_OK;
}

// Allocates the (zeroed) deferred bitmap and, once per graph, sets the
// automatic deferral window: the mean weight / 16, at least 1 (2 for weights
// uniform in [1, 64]). A query that failed may have left bits set: clear them.
static int ensure_defer(ptg_graph *g) {
    int rc;
    if (!g->d_defer) {
        if ((rc = galloc(g, (void **)&g->d_defer, 4ull * g->nwords))) return rc;
        g->defer_dirty = true;
    }
    if (g->defer_dirty) {
        if ((rc = fill_u32(g, g->d_defer, g->nwords, 0u))) return rc;
        g->defer_dirty = false;
    }
    if (g->auto_window) return PTG_OK;
    unsigned long long sums[kShards];
    if ((rc = fill_u32(g, reinterpret_cast<uint32_t *>(g->d_acc), 2 * kShards, 0u))) return rc;
    if (g->E) {
        k_weight_sum<<<g->grid, kBlock, 0, g->stream>>>(g->d_w, g->E, g->d_acc);
        PTG_HIP(hipGetLastError());
    }
    if ((rc = copy_sync(g, sums, g->d_acc, sizeof sums, hipMemcpyDeviceToHost))) return rc;
    uint64_t total = 0;
    for (int i = 0; i < kShards; i++) total += sums[i];
    const uint64_t mean = g->E ? total / g->E : 0;
    g->auto_window = (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>((mean + 8) / 16, 1u << 30));
    return PTG_OK;
}

int run_sssp(ptg_graph *g, uint32_t src, uint32_t *dist, ptg_stats *st) {
    int rc;
    LaunchTimer timer(g);
    const uint32_t delta = g->cfg.sssp_delta ? g->cfg.sssp_delta : kUnbounded;
    const bool split = delta != kUnbounded;
    if (split ? (rc = ensure_split(g, delta)) : (rc = ensure_packed(g))) return rc;
    if ((rc = ensure_defer(g))) return rc;
    const uint32_t defer_deg = g->cfg.sssp_defer_degree ? g->cfg.sssp_defer_degree : kDefaultDeferDegree;
    const uint32_t window = g->cfg.sssp_defer_window ? g->cfg.sssp_defer_window : g->auto_window;
    const uint32_t *off = split ? g->light.off : g->out.off;
    const uint32_t *col = split ? g->light.col : g->out.col;
    PTG_HIP(hipEventRecord(g->ev0, g->stream));
    const FillSpan none{nullptr, 0, 0u};
    if ((rc = reset_rounds(g, {FillSpan{dist, g->V, g->cfg.sssp_sentinel},
                               split ? FillSpan{reinterpret_cast<uint32_t *>(g->d_far), (uint64_t)g->npairs * 16, 0u}
                                     : none,
                               st ? reach_reset_span(g) : none})))
        return rc;
    // closing work: end event, reach statistics (accumulators cleared again
    // if an earlier enqueue turned out early), output delivery
    bool acc_clean = true;
    auto finish = [&](bool speculative) -> int {
        PTG_HIP(hipEventRecord(g->ev1, g->stream));
        if (st) {
            const FillSpan z = reach_reset_span(g);
            int rc2;
            if ((!acc_clean && (rc2 = fill_spans(g, &z, 1))) ||
                (rc2 = reach_stats_launch(g, dist, g->cfg.sssp_sentinel)))
                return rc2;
            acc_clean = false;
        }
        return deliver_output(g, speculative);
    };
    bool finished = false;
    // round 0 expands the source (over the light CSR when split)
    const uint32_t thr0 = delta > 0xFFFFFFFEu ? 0xFFFFFFFFu : 1u + delta;
    if ((rc = launch_seed_csr(g, off, col, src, 1u, dist, thr0))) return rc;
    uint32_t iters = 0;
    uint64_t fsum = 0;
    RelaxParams p{dist, split ? g->d_wl : g->d_w, split ? nullptr : g->d_cw, g->d_far, delta, split ? 1u : 0u,
                  split ? g->heavy.off : nullptr, split ? g->heavy.col : nullptr, split ? g->d_wh : nullptr,
                  split ? g->heavy.bins.q.with(g->heavy.bins.d_ctr) : Frontier{}, defer_deg, window};
    if ((rc = run_relax_rounds<true>(g, p, off, col, timer, &iters, &fsum, finish, &finished))) {
        g->defer_dirty = true;
        return rc;
    }
    if (!finished && (rc = finish(false))) return rc;
    PTG_HIP(hipStreamSynchronize(g->stream));
    if (st) {
        memset(st, 0, sizeof *st);
        float ms = 0;
        PTG_HIP(hipEventElapsedTime(&ms, g->ev0, g->ev1));
        st->kernel_ns = (uint64_t)((double)ms * 1e6);
        st->iterations = iters;
        st->frontier_sum = fsum;
        if ((rc = timer.collect(st))) return rc;
        reach_stats_read(g, &st->reached, &st->edges_traversed);
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
    RelaxParams p{label, nullptr, nullptr, g->d_far, 0u, 0u, nullptr, nullptr, nullptr, Frontier{}, 0xFFFFFFFFu, 0u};
    bool finished = false;
    return run_relax_rounds<false>(g, p, g->out.off, g->out.col, timer, iters, fsum, [](bool) { return PTG_OK; },
                                   &finished);
}

}  // namespace ptg
