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
// Default deferral degree: vertices of degree > 128 wait for the window.
// With pull and sweep rounds, 128 against 64 (the best before sweeps): R-MAT
// scale 18 0.449 / 0.474 ms per query, scale 22 1.198 / 1.220, scale 24
// 3.474 / 3.666 (weights in [1, 64]; 96 and 192 within 1% of 128, 32 and
// 256 slower).
constexpr uint32_t kDefaultDeferDegree = 128;
// Default sweep threshold (sssp_sweep 0): a round sweeps when its push and
// its pull would each scan more than this share of the edges. A sweep reads
// every edge's 16-bit entry and weight byte and every pair's source and the
// source's value once, which costs about what pushing or pulling a quarter
// of the edges does (measured best of 0.1..0.4 at R-MAT scales 22 and 24).
constexpr double kDefaultSweepShare = 0.25;

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

struct PullParams {
    const uint2 *cwin;   // in-edge (source, weight) pairs
    uint32_t *val;
    float pull;          // as RelaxParams::pull
    uint32_t wmin;       // smallest edge weight
    uint32_t sent;       // value of unreached vertices
    const uint32_t *long_ids, *long_base;  // vertices of degree > tile_deg, first chunk of each
    const uint32_t *long_block;            // entry holding chunk b * kPullBlock, per block b
    uint32_t n_long, long_chunks;
    uint32_t tile_deg;                     // lists up to this long are walked in their tile
    uint64_t sweep;                        // as RelaxParams::sweep for this round
};

// Pull rounds: lists of up to tile_deg in-edges are walked by the wave that
// owns their tile, longer ones from a fixed list in blocks of kPullBlock
// 64-edge chunks. tile_deg is set per graph so that a tile full of such
// lists holds at most 4 x E / (waves of the grid) edges, within [sd, 256]:
// 256 at R-MAT scale 22, 58 at scale 18, where there are fewer tiles than
// waves and a heavy tile would hold up the round.
constexpr uint32_t kPullTileMax = 256;
constexpr uint32_t kPullBlock = 16;

// Pull round (symmetric graph, one unbounded bucket), chosen when the
// frontier's out-edges outnumber the in-edges of the vertices that can still
// improve. Only a vertex above `bound` (frontier minimum + smallest weight)
// can fall, since every edge out of the frontier carries at least that; such
// an open vertex takes the minimum of val[u] + w over its in-edges (u, w)
// from reached neighbours. A neighbour outside the frontier already satisfies
// val[v] <= val[u] + w, so reading it changes nothing, and a neighbour whose
// value falls during the round joins the next frontier, as an improved vertex
// does in a push round: the fixed point is the push rounds' one. Improved
// vertices go to the output flags.
//
// Vertices of degree > tile_deg are taken from a fixed list of them (built
// once per graph), the rest from the tiles of 64 vertices.

// The long lists: the fixed list's 64-edge chunks are cut into blocks of
// kPullBlock chunks, dealt to the waves round-robin (wave gw takes blocks gw,
// gw + nw, ...), so the open vertices, wherever they sit in the list, are
// spread over all waves. A wave skips the chunks of vertices that are not
// open and lowers every open vertex it touches to the minimum over its part
// of the list with one atomicMin.
__device__ __forceinline__ void pull_long(const RoundEnv &e, const PullParams &q, uint32_t bound, const Emit &em,
                                          Acc &acc, uint32_t gw, uint32_t nw) {
    const uint32_t lane = lane_id();
    const uint32_t C = q.long_chunks, n = q.n_long;
    const uint32_t nblocks = (C + kPullBlock - 1) / kPullBlock;
    for (uint32_t blk = gw; blk < nblocks; blk += nw) {
        const uint32_t c1 = min(C, (blk + 1) * kPullBlock);
        uint32_t c = blk * kPullBlock;
        uint32_t idx = q.long_block[blk];  // the entry holding chunk c
        uint32_t k = c - q.long_base[idx];
        while (c < c1) {
            // entries [idx, idx + nb), one per lane
            const uint32_t nb = min(64u, n - idx);
            uint32_t mv = 0, mb = 0, md = 0, mc = 0;
            if (lane < nb) {
                mv = q.long_ids[idx + lane];
                mb = e.off[mv];
                md = e.off[mv + 1] - mb;
                mc = q.val[mv];
            }
            for (uint32_t i = 0; i < nb && c < c1; i++) {
                const uint32_t v = __builtin_amdgcn_readlane(mv, i);
                const uint32_t vb = __builtin_amdgcn_readlane(mb, i);
                const uint32_t vd = __builtin_amdgcn_readlane(md, i);
                const uint32_t vc = __builtin_amdgcn_readlane(mc, i);
                const uint32_t nch = wave_chunks(vd);
                if (vc <= bound) {  // not open: skip its chunks in the block
                    c += min(nch - k, c1 - c);
                    k = 0;
                    continue;
                }
                uint32_t m = 0xFFFFFFFFu;
                while (k < nch && c < c1) {
                    const uint32_t gn = min(min(kGroup, nch - k), c1 - c);
                    const uint32_t b = k * kWaveChunk;
                    const uint32_t end = min(vd, b + kWaveChunk * gn);
                    uint2 x[kGroup];
#pragma unroll
                    for (uint32_t j = 0; j < kGroup; j++) {
                        const uint32_t t = b + 64 * j + lane;
                        x[j] = t < end ? q.cwin[(uint64_t)vb + t] : make_uint2(0u, 0u);
                    }
                    uint32_t cu[kGroup];
#pragma unroll
                    for (uint32_t j = 0; j < kGroup; j++) cu[j] = b + 64 * j + lane < end ? q.val[x[j].x] : q.sent;
#pragma unroll
                    for (uint32_t j = 0; j < kGroup; j++)
                        if (cu[j] < q.sent) m = min(m, cu[j] + x[j].y);
                    k += gn;
                    c += gn;
                }
                k = 0;  // the next entry starts at its first chunk
                m = wave_min_u32(m);
                // a candidate below the value read counts as improved without
                // waiting for the atomic: if another part of the list went
                // lower, the vertex is in the next frontier anyway
                const bool improved = lane == 0 && m < vc;
                if (improved) {
                    (void)__hip_atomic_fetch_min(&q.val[v], m, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                    e.flout[v] = 1;
                    acc.fmin_inv = max(acc.fmin_inv, ~m);
                }
                mark_any(em, improved);
            }
            idx += nb;
        }
    }
}

// A wave's LDS for the edge-parallel walk of its tile's short lists.
struct PullLds {
    uint32_t start[64], base[64], best[64];
};

// The tiles: a wave owns 64 consecutive vertices, and the open ones with
// 1..tile_deg in-edges are walked edge-parallel over the concatenation of
// their lists, a lane per edge (consecutive lanes on
// consecutive edges of a list); the owner is found by binary search over the
// lists' starts in LDS and its minimum lowered with an LDS atomic. Each value
// has one writer here and is stored without an atomic.
__device__ __forceinline__ void pull_tiles(const RoundEnv &e, const uint2 *__restrict__ cwin, uint32_t *val,
                                           uint32_t bound, uint32_t sent, uint32_t tile_deg, const Emit &em,
                                           Acc &acc, uint32_t gw, uint32_t nw) {
    __shared__ PullLds s_pull[kWavesPerBlock];
    PullLds &L = s_pull[threadIdx.x >> 6];
    const uint32_t lane = lane_id();
    for (uint32_t p = gw; p < e.npairs; p += nw) {
        const uint32_t v = p * 64 + lane;
        const uint32_t cur = v < e.V ? val[v] : 0u;
        const bool open = v < e.V && cur > bound;
        if (ballot(open) == 0) continue;
        uint32_t b = 0, d = 0;
        if (open) {
            b = e.off[v];
            d = e.off[v + 1] - b;
        }
        const bool sm = open && d > 0 && d <= tile_deg;
        const uint64_t msm = ballot(sm);
        if (msm == 0) continue;
        const uint32_t n = (uint32_t)__popcll(msm);
        const uint32_t k = rank_below(msm);
        const uint32_t ds = sm ? d : 0u;
        const uint32_t incl = wave_incl_scan_u32(ds);
        const uint32_t T = read_lane(incl, 63);
        if (sm) {
            L.start[k] = incl - ds;
            L.base[k] = b;
            L.best[k] = 0xFFFFFFFFu;
        }
        wave_lds_sync();
        for (uint32_t t0 = 0; t0 < T; t0 += kGroup * 64) {
            uint32_t own[kGroup];
            uint2 x[kGroup];
#pragma unroll
            for (uint32_t j = 0; j < kGroup; j++) {
                const uint32_t t = t0 + 64 * j + lane;
                uint32_t lo = 0, hi = n;
                while (hi - lo > 1) {
                    const uint32_t mid = (lo + hi) >> 1;
                    if (L.start[mid] <= t) lo = mid; else hi = mid;
                }
                own[j] = lo;
                x[j] = t < T ? cwin[(uint64_t)L.base[lo] + (t - L.start[lo])] : make_uint2(0u, 0u);
            }
            uint32_t cu[kGroup];
#pragma unroll
            for (uint32_t j = 0; j < kGroup; j++) cu[j] = t0 + 64 * j + lane < T ? val[x[j].x] : sent;
#pragma unroll
            for (uint32_t j = 0; j < kGroup; j++)
                if (cu[j] < sent) atomicMin(&L.best[own[j]], cu[j] + x[j].y);
        }
        wave_lds_sync();
        const uint32_t best = sm ? L.best[k] : 0xFFFFFFFFu;
        wave_lds_sync();  // the next tile rewrites the arrays
        const bool improved = sm && best < cur;
        if (improved) {
            val[v] = best;
            e.flout[v] = 1;
            acc.fmin_inv = max(acc.fmin_inv, ~best);
        }
        mark_any(em, improved);
    }
}

// The pre-pass of a round counts open vertices (and queues their long lists)
// only when the previous frontier had at least E / kPullGateShare edges: a
// round after a small frontier pushes few edges and would not pull.
constexpr uint64_t kPullGateShare = 64;

// Whether a round pulls: the pre-pass counted the in-edges of the vertices
// that can still improve (uedges), and they times `factor` are fewer than
// the out-edges a push round would scan (the frontier's, less those of its
// deferred vertices).
__device__ __forceinline__ bool pull_pays(const RoundCtr *in, uint64_t frontier_edges, float factor) {
    if (!in->ucount) return false;
    uint64_t u = 0, d = 0;
#pragma unroll
    for (int i = 0; i < kShards; i++) {
        u += in->uedges_sh[i];
        d += in->dedges_sh[i];
    }
    return (double)u * factor < (double)(frontier_edges - min(d, frontier_edges));
}

// Whether round e.r pulls (the push and pull kernels of the round decide
// alike): a frontier held as flags (the pre-pass counted the open in-edges)
// with a pull kernel behind the round (factor > 0) that pays.
__device__ __forceinline__ bool round_pulls(const RoundCtr *in, const InSummary &s, uint32_t qcap, float factor) {
    return factor > 0 && s.found > 0 && s.edges > qcap && in->mode == kModeFlags && pull_pays(in, s.edges, factor);
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
    float pull;          // > 0: a pull kernel follows each round, which pulls when
                         // open in-edges * pull < the push round's out-edges
    uint64_t sweep;      // > 0: sweep kernels follow the round, which sweep when
                         // round_sweeps(..., pull, sweep) (rounds.h)
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
        // a pulling round is expanded by k_pull_round, a sweeping one by the
        // sweep kernels, both launched behind this one
        const bool other = kWeighted && !kBuckets &&
                           (round_sweeps(in, s.found, s.edges, e.qcap, p.pull, p.sweep) ||
                            round_pulls(in, s, e.qcap, p.pull));
        if (!other) expand_frontier(e, in, op, em, acc);
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

// The pull expansion of round e.r, launched behind k_relax_round, which has
// done the rest of the round (output header, deferred vertices, clears);
// returns at once when the round does not pull. A kernel of its own leaves
// the push kernel's register allocation as it was.
__global__ __launch_bounds__(kBlock, 8) void k_pull_round(RoundEnv e, PullParams q) {
    RoundCtr *in = e.ring + e.r % kRing;
    const InSummary s = summarize(in);
    if (!round_pulls(in, s, e.qcap, q.pull) || round_sweeps(in, s.found, s.edges, e.qcap, q.pull, q.sweep)) return;
    RoundCtr *out = e.ring + (e.r + 1) % kRing;
    if (blockIdx.x == 0 && threadIdx.x == 0) in->pulled = 1u;
    const Emit em = make_emit(e, out, kModeFlags);
    Acc acc;
    const uint32_t gw = blockIdx.x * kWavesPerBlock + (threadIdx.x >> 6);
    const uint32_t nw = gridDim.x * kWavesPerBlock;
    const uint32_t bound = pull_bound(in, q.wmin);
    pull_long(e, q, bound, em, acc, gw, nw);
    pull_tiles(e, q.cwin, q.val, bound, q.sent, q.tile_deg, em, acc, gw, nw);
    add_tallies(out, acc);
}

// Rounds that launch the pull kernel: those within one of a round that pulled
// in either of the last two queries (rounds from 64 on always), and every
// round of each 16th query, so that a change of source or graph shape is
// seen again. A round without the kernel pushes, with the same result.
static uint64_t pull_window(ptg_graph *g) {
    if (g->pull_queries++ % 16 == 0) return ~0ull;
    const uint64_t h = g->pull_hist[0] | g->pull_hist[1];
    return h | (h << 1) | (h >> 1);
}

// Rounds that launch the sweep kernels: those within one of a round that
// swept in either of the last two queries, and every round of the queries
// whose pull window holds every round.
static uint64_t sweep_window(const ptg_graph *g, bool all) {
    if (all) return ~0ull;
    const uint64_t h = g->sweep_hist[0] | g->sweep_hist[1];
    return h | (h << 1) | (h >> 1);
}

template <bool kWeighted, class Finish>
static int run_relax_rounds(ptg_graph *g, RelaxParams p, const PullParams &q, const uint32_t *off,
                            const uint32_t *col, LaunchTimer &timer, uint32_t *iters, uint64_t *fsum, Finish finish,
                            bool *finished) {
    const bool look_all = g->pull_queries % 16 == 0;
    const uint64_t window = p.pull > 0 ? pull_window(g) : 0ull;
    const float factor = p.pull;
    const uint64_t sweep_edges = p.sweep;
    const uint64_t swindow = sweep_edges ? sweep_window(g, look_all) : 0ull;
    int rc = drive_rounds(
        g,
        [&](uint32_t r) -> int {
            RoundEnv env = make_env_csr(g, r, off, col);
            env.val = p.val;
            env.defer_deg = p.defer_deg;
            env.window = p.window;
            p.pull = factor > 0 && (r >= 64 || ((window >> r) & 1)) ? factor : 0.0f;
            p.sweep = sweep_edges && (r >= 64 || ((swindow >> r) & 1)) ? sweep_edges : 0ull;
            env.pull = p.pull > 0 || p.sweep ? 1u : 0u;
            env.pull_wmin = q.wmin;
            env.pull_gate = (uint32_t)std::min<uint64_t>(g->E / kPullGateShare, 0xFFFFFFFFu);
            int rc = launch_prepare(g, env);
            if (rc) return rc;
            if (p.split)
                timer.launch(k_relax_round<kWeighted, true>, dim3(g->grid), dim3(kBlock), env, p);
            else
                timer.launch(k_relax_round<kWeighted, false>, dim3(g->grid), dim3(kBlock), env, p);
            if (p.pull > 0) {
                PullParams qr = q;
                qr.sweep = p.sweep;
                timer.launch_rider(k_pull_round, dim3(g->grid), dim3(kBlock), env, qr);
            }
            PTG_HIP(hipGetLastError());
            if (p.sweep) return sssp_sweep_launch(g, env, p.val, q.sent, p.pull, p.sweep, timer);
            return PTG_OK;
        },
        [](const RoundCtr &c) { return c.any == 0 && c.far_any == 0 && c.need_heavy == 0; },
        g->last_rounds[kWeighted ? 1 : 2], iters, fsum, nullptr, finish, finished);
    if (!rc && factor > 0) {
        // a round that swept stays a candidate for pulls too
        g->pull_hist[1] = g->pull_hist[0];
        g->pull_hist[0] = g->pulled_mask | g->swept_mask;
    }
    if (!rc && sweep_edges) {
        g->sweep_hist[1] = g->sweep_hist[0];
        g->sweep_hist[0] = g->swept_mask;
    }
    return rc;
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

// Pull rounds run on a symmetric graph with one unbounded bucket, unless
// switched off, and only where no path value plus an edge can reach the
// unreached value (every value is a simple path's length, at most
// 1 + (V - 1) * wmax), so a candidate never wraps.
static bool pull_wanted(const ptg_graph *g, bool split) {
    return !split && g->cfg.sssp_pull >= 0 && g->symmetric != 0;
}

static bool pull_in_range(const ptg_graph *g) {
    return 1 + (uint64_t)g->V * g->wmax + g->wmax < (uint64_t)g->cfg.sssp_sentinel;
}

void release_sssp_split(ptg_graph *g) {
    free_split(g);
    release_sssp_pull(g);
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
    if (pull_wanted(g, delta != kUnbounded) && g->wsymmetric != 1 && !g->d_cwin)
        need += 8ull * std::max<uint64_t>(g->E, 1) + 4ull * g->V + 4;
    if (pull_wanted(g, delta != kUnbounded) && !g->long_built) need += 12ull * g->V + 12;
    if (pull_wanted(g, delta != kUnbounded)) need += sssp_sweep_need(g);
    if (delta != kUnbounded && !(g->split_delta == delta && g->light.off))
        need += 8ull * g->E + 8ull * (g->V + 1) + 4ull * g->V + (64u << 10);
    return need;
}

// Sum of the weights over 16 accumulators, smallest weight in acc[16],
// largest in acc[17].
__global__ __launch_bounds__(kBlock) void k_weight_sum(const uint32_t *__restrict__ w, uint64_t E,
                                                       unsigned long long *acc) {
    uint64_t sum = 0;
    uint32_t lo = 0xFFFFFFFFu, hi = 0;
    for (uint64_t i = (uint64_t)blockIdx.x * kBlock + threadIdx.x; i < E; i += (uint64_t)gridDim.x * kBlock) {
        const uint32_t x = w[i];
        sum += x;
        lo = min(lo, x);
        hi = max(hi, x);
    }
    sum = wave_sum_u64(sum);
    lo = wave_min_u32(lo);
    hi = wave_max_u32(hi);
    if (lane_id() == 0) {
        atomicAdd(&acc[(blockIdx.x * kWavesPerBlock + (threadIdx.x >> 6)) % kShards], sum);
        atomicMin(&acc[16], (unsigned long long)lo);
        atomicMax(&acc[17], (unsigned long long)hi);
    }
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
    unsigned long long sums[kShards + 2];
    uint32_t *acc = reinterpret_cast<uint32_t *>(g->d_acc);
    const FillSpan z[3] = {FillSpan{acc, 2 * kShards, 0u}, FillSpan{acc + 2 * kShards, 2, 0xFFFFFFFFu},
                           FillSpan{acc + 2 * kShards + 2, 2, 0u}};
    if ((rc = fill_spans(g, z, 3))) return rc;
    if (g->E) {
        k_weight_sum<<<g->grid, kBlock, 0, g->stream>>>(g->d_w, g->E, g->d_acc);
        PTG_HIP(hipGetLastError());
    }
    if ((rc = copy_sync(g, sums, g->d_acc, sizeof sums, hipMemcpyDeviceToHost))) return rc;
    uint64_t total = 0;
    for (int i = 0; i < kShards; i++) total += sums[i];
    const uint64_t mean = g->E ? total / g->E : 0;
    g->wmin = g->E ? (uint32_t)sums[kShards] : 0u;
    g->wmax = g->E ? (uint32_t)sums[kShards + 1] : 0u;
    g->auto_window = (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>((mean + 8) / 16, 1u << 30));
    return PTG_OK;
}

// In-edge (source, weight) pairs of a symmetric graph, in the out-CSR's
// offsets: one wave per source vertex u, every out-edge (u -> v, w) takes the
// next slot of v's list. A slot past the list means in- and out-degrees
// differ (the graph is not symmetric after all): *bad is set, nothing is
// written out of the list.
__global__ __launch_bounds__(kBlock) void k_in_pairs(uint32_t V, const uint32_t *__restrict__ off,
                                                     const uint2 *__restrict__ cw, uint32_t *__restrict__ cursor,
                                                     uint2 *__restrict__ cwin, uint32_t *__restrict__ bad) {
    const uint32_t lane = lane_id();
    const uint32_t gw = blockIdx.x * kWavesPerBlock + (threadIdx.x >> 6);
    const uint32_t nw = gridDim.x * kWavesPerBlock;
    for (uint32_t u = gw; u < V; u += nw) {
        const uint32_t end = off[u + 1];
        for (uint32_t i = off[u] + lane; i < end; i += 64) {
            const uint2 x = cw[i];
            const uint32_t slot = atomicAdd(&cursor[x.x], 1u);
            const uint32_t lo = off[x.x], hi = off[x.x + 1];
            if (slot < hi - lo)
                cwin[(uint64_t)lo + slot] = make_uint2(u, x.y);
            else
                *bad = 1u;
        }
    }
}

void release_sssp_pull(ptg_graph *g) {
    gfree(g, g->d_cwin);
    gfree(g, g->d_long_ids);
    gfree(g, g->d_long_base);
    gfree(g, g->d_long_block);
    g->d_cwin = nullptr;
    g->d_long_ids = g->d_long_base = g->d_long_block = nullptr;
    g->n_long = g->long_chunks = 0;
    g->long_built = false;
}

// Vertices of degree > tile_deg with the index of their first 64-edge
// chunk, in vertex order, and for every block of kPullBlock chunks the entry
// holding its first chunk (once per graph).
static int ensure_long_list(ptg_graph *g) {
    if (g->long_built) return PTG_OK;
    const uint32_t V = g->V;
    const uint64_t waves = (uint64_t)g->grid * kWavesPerBlock;
    const uint32_t tile_deg = (uint32_t)std::min<uint64_t>(
        kPullTileMax, std::max<uint64_t>(g->cfg.small_degree, 4 * g->E / (64 * waves)));
    std::vector<uint32_t> off((size_t)V + 1), ids, base, block;
    int rc = copy_sync(g, off.data(), g->out.off, 4ull * (V + 1), hipMemcpyDeviceToHost);
    if (rc) return rc;
    uint64_t chunks = 0;
    for (uint32_t v = 0; v < V; v++) {
        const uint32_t d = off[v + 1] - off[v];
        if (d <= tile_deg) continue;
        const uint64_t nch = (d + kWaveChunk - 1) / kWaveChunk;
        for (uint64_t c = (chunks + kPullBlock - 1) / kPullBlock * kPullBlock; c < chunks + nch; c += kPullBlock)
            block.push_back((uint32_t)ids.size());
        ids.push_back(v);
        base.push_back((uint32_t)chunks);
        chunks += nch;
    }
    const uint64_t n = ids.size(), nb = block.size();
    auto up = [&](uint32_t **dst, const std::vector<uint32_t> &src) -> int {
        int r = galloc(g, (void **)dst, 4 * std::max<uint64_t>(src.size(), 1));
        if (!r && !src.empty()) r = copy_sync(g, *dst, src.data(), 4 * src.size(), hipMemcpyHostToDevice);
        return r;
    };
    if ((rc = up(&g->d_long_ids, ids)) || (rc = up(&g->d_long_base, base)) || (rc = up(&g->d_long_block, block))) {
        gfree(g, g->d_long_ids);
        gfree(g, g->d_long_base);
        gfree(g, g->d_long_block);
        g->d_long_ids = g->d_long_base = g->d_long_block = nullptr;
        return rc;
    }
    (void)nb;
    g->n_long = (uint32_t)n;
    g->pull_tile_deg = tile_deg;
    g->long_chunks = (uint32_t)chunks;  // < E / 64 + V
    g->long_built = true;
    return PTG_OK;
}

// Builds the in-edge pairs once (needs the packed out-pairs). On a graph that
// turns out not to be symmetric the pairs are dropped and the graph is marked
// so.
static int ensure_pull(ptg_graph *g) {
    if (g->d_cwin) return PTG_OK;
    int rc;
    uint32_t *cursor = nullptr;
    if ((rc = galloc(g, (void **)&g->d_cwin, 8ull * std::max<uint64_t>(g->E, 1)))) return rc;
    if ((rc = galloc(g, (void **)&cursor, 4ull * g->V + 4))) {
        release_sssp_pull(g);
        return rc;
    }
    uint32_t bad = 0;
    if (!(rc = fill_u32(g, cursor, (uint64_t)g->V + 1, 0u))) {
        if (g->E) k_in_pairs<<<g->grid, kBlock, 0, g->stream>>>(g->V, g->out.off, g->d_cw, cursor, g->d_cwin,
                                                                 cursor + g->V);
        if (hipGetLastError() != hipSuccess) rc = fail(PTG_ERR_HIP, "k_in_pairs launch failed");
        if (!rc) rc = copy_sync(g, &bad, cursor + g->V, 4, hipMemcpyDeviceToHost);
    }
    gfree(g, cursor);
    if (rc || bad) release_sssp_pull(g);
    if (!rc && bad) g->symmetric = 0;
    return rc;
}

// The (source, weight) in-edge pairs of a symmetric weighted graph, in the
// out-CSR's offsets, for pull rounds of batched SSSP: the out-pairs when
// every edge's reverse carries the same weight, else the built in-pairs.
// *cwin stays null when a pull is not possible: the graph is not symmetric,
// or a path plus an edge could reach the sentinel (a candidate could wrap).
int sssp_in_pairs(ptg_graph *g, const uint2 **cwin) {
    *cwin = nullptr;
    if (!g->weighted) return PTG_OK;
    int rc, sym = 0, wsym = 0;
    if ((rc = check_symmetric(g, &sym))) return rc;
    if (!sym) return PTG_OK;
    if ((rc = ensure_packed(g)) || (rc = ensure_defer(g))) return rc;  // ensure_defer learns wmax
    if (!pull_in_range(g)) return PTG_OK;
    if ((rc = check_weight_symmetric(g, g->d_cw, &wsym))) return rc;
    if (!wsym && (rc = ensure_pull(g))) return rc;
    *cwin = wsym ? g->d_cw : g->d_cwin;
    return PTG_OK;
}

// Device bytes sssp_in_pairs would still allocate (an upper bound: the
// in-pairs are counted unless the weights are known to be symmetric).
uint64_t sssp_in_pairs_need(const ptg_graph *g) {
    if (!g->weighted || g->symmetric == 0) return 0;
    uint64_t need = 0;
    if (!g->d_cw) need += 8ull * std::max<uint64_t>(g->E, 1);
    if (!g->d_defer) need += 4ull * g->nwords;
    if (g->wsymmetric != 1 && !g->d_cwin) need += 8ull * std::max<uint64_t>(g->E, 1) + 4ull * g->V + 4;
    return need;
}

int run_sssp(ptg_graph *g, uint32_t src, uint32_t *dist, ptg_stats *st) {
    int rc;
    LaunchTimer timer(g);
    const uint32_t delta = g->cfg.sssp_delta ? g->cfg.sssp_delta : kUnbounded;
    const bool split = delta != kUnbounded;
    if (split ? (rc = ensure_split(g, delta)) : (rc = ensure_packed(g))) return rc;
    if ((rc = ensure_defer(g))) return rc;
    // in-edge pairs for pull rounds: the out-pairs themselves when every
    // edge's reverse carries the same weight, else the built in-pairs
    const uint2 *cwin = nullptr;
    if (pull_wanted(g, split)) {
        int sym = 0, wsym = 0;
        if ((rc = check_symmetric(g, &sym))) return rc;
        if (sym && pull_in_range(g)) {
            if ((rc = check_weight_symmetric(g, g->d_cw, &wsym))) return rc;
            if (!wsym && (rc = ensure_pull(g))) return rc;
            if ((rc = ensure_long_list(g))) return rc;
            cwin = wsym ? g->d_cw : g->d_cwin;
        }
    }
    // sweep rounds over the binned layout where pulls are possible
    uint64_t sweep_edges = 0;
    if (cwin && g->cfg.sssp_sweep >= 0) {
        bool ok = false;
        if ((rc = sssp_sweep_prepare(g, &ok))) return rc;
        const double share = g->cfg.sssp_sweep > 0 ? g->cfg.sssp_sweep : kDefaultSweepShare;
        if (ok) sweep_edges = std::max<uint64_t>(1, (uint64_t)((double)g->E * share));
    }
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
                  split ? g->heavy.bins.q.with(g->heavy.bins.d_ctr) : Frontier{}, defer_deg, window,
                  cwin ? (g->cfg.sssp_pull > 0 ? g->cfg.sssp_pull : 1.0f) : 0.0f, sweep_edges};
    const PullParams q{cwin,          dist,           p.pull,          g->wmin,  g->cfg.sssp_sentinel,
                       g->d_long_ids, g->d_long_base, g->d_long_block, g->n_long, g->long_chunks,
                       g->pull_tile_deg};
    if ((rc = run_relax_rounds<true>(g, p, q, off, col, timer, &iters, &fsum, finish, &finished))) {
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
        st->pull_rounds = g->pulled_rounds;
        st->sweep_rounds = g->swept_rounds;
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
    return run_relax_rounds<false>(g, p, PullParams{}, g->out.off, g->out.col, timer, iters, fsum,
                                   [](bool) { return PTG_OK; }, &finished);
}

}  // namespace ptg
