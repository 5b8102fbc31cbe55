// Round machinery shared by the BFS, SSSP and label-propagation kernels.
//
// Replaces ptGraph's per-iteration pipeline (setLabeling -> thrust::reduce ->
// thrust::exclusive_scan -> setActiveNodeArray -> setLabelDefault -> relax,
// ptGraph/main.cu:1094-1106, kernels ptGraph/gpu_kernels.cu:3-172) by two
// persistent-grid kernels per round: a pre-pass that turns the frontier into
// work lists, and the round kernel that expands it.
//
// A round's input frontier (ring entry r % kRing) comes in one of three forms:
//   * queue:  small frontiers. The producer appended every vertex once (its
//             bit in the frontier bitmap dedupes): degree <= sd to a sharded
//             S queue, larger ones to a sharded chunk queue.
//   * flags:  large frontiers. The producer only stored a byte per vertex
//             (plain stores, no atomics: duplicates are harmless); the
//             pre-pass packs the bytes into the bitmap, counts the frontier
//             and fills the chunk queue.
//   * bitmap: the producer wrote whole bitmap words itself (bottom-up BFS,
//             SSSP far-set splits, seeds) and counted exactly; the pre-pass
//             fills the chunk queue.
// The chunk queue holds vertices of degree > sd as ceil(d / kWaveChunk)
// chunks of 64 edges; wave w expands the contiguous chunk range
// [C*w/W, C*(w+1)/W), so a hub's list is spread over many waves and every
// wave gets about the same number of edges. Degree <= sd vertices are expanded one
// per lane from the S queue or straight from bitmap tiles of 64 vertices.
// Counters are summed per workgroup and added to 1 of 16 shards (a single hot
// address serialises atomics: measured 11 ns each on MI355X).
//
// SSSP defers high-degree vertices: one of degree > defer_deg whose value is
// at or above the frontier's smallest value plus `window` stays in the
// frontier without being expanded, so that a hub is expanded close to its
// final value rather than once per improvement. The pre-pass moves such
// vertices of flags and bitmap frontiers to a deferred bitmap (whole words,
// plain stores), which the round puts back into its output frontier and
// clears; queue frontiers (at most kQueueCap edges) skip them during the walk. The vertex with the smallest value is
// always expanded, so every round makes progress.
#pragma once

#include <initializer_list>

#include "device_util.h"

namespace ptg {

constexpr uint32_t kModeFlags = 2;

struct RoundEnv {
    RoundCtr *ring;
    uint32_t r;                   // round index
    const uint32_t *off, *col;
    uint32_t *fin, *fout, *fclr;  // frontier bitmaps: read, write, clear
    uint8_t *flin, *flout;        // frontier byte flags: read (pre-pass), write
    uint32_t *qin, *qout;         // sharded S queues read / written
    uint32_t *cqin, *cbin;        // sharded chunk queue read (ids, first chunks)
    uint32_t *cqout, *cbout;      // sharded chunk queue written
    uint32_t cq_cap;              // chunk-queue entries per shard
    uint32_t V, npairs, nwords;
    uint32_t sd;                  // small degree threshold
    uint32_t qcap;                // frontier edges up to which the output is queued
    const uint32_t *val;          // SSSP deferral: vertex values
    uint32_t *dbits;              // SSSP deferral: deferred vertices of the round (bitmap)
    uint32_t defer_deg;           // defer vertices of degree > defer_deg (~0u: never)
    uint32_t window;              // ... whose value >= frontier minimum + window
    float bu_beta;                // BFS: bottom-up -> top-down switch factor (0: not BFS)
    uint32_t pull;                // SSSP: flags pre-passes count the in-edges of the vertices
    uint32_t pull_wmin;           // ... whose value is above the frontier minimum + pull_wmin,
    uint32_t pull_gate;           // ... when the previous frontier had at least this many edges
};

// Saturating a + b.
__device__ __forceinline__ uint32_t add_sat(uint32_t a, uint32_t b) { return a > 0xFFFFFFFFu - b ? 0xFFFFFFFFu : a + b; }

// Values at or above this limit defer a vertex of degree > defer_deg in the
// round that reads `in` (no limit when the frontier minimum is unknown).
__device__ __forceinline__ uint32_t defer_limit(const RoundCtr *in, uint32_t window) {
    const uint32_t fi = in->fmin_inv;
    return fi ? add_sat(~fi, window) : 0xFFFFFFFFu;
}

// Values above this bound can still fall in the round that reads `in`:
// every edge out of the frontier carries at least the frontier minimum plus
// the smallest weight. Valid when in->fmin_inv != 0.
__device__ __forceinline__ uint32_t pull_bound(const RoundCtr *in, uint32_t wmin) { return add_sat(~in->fmin_inv, wmin); }

// Whether round `in` of an SSSP query sweeps: relaxes every edge through the
// binned layout (pagerank_bins.hip) instead of pushing or pulling. It does
// when its frontier is held as flags, the pre-pass counted the in-edges of
// the vertices that can still improve (ucount), and both the push (the
// frontier's out-edges less those of its deferred vertices) and, in a round
// whose pull kernel runs (factor > 0), the pull (those in-edges x factor)
// would scan more than min_edges edges (0: the round never sweeps). The push,
// pull and sweep kernels of a round all decide with this function.
__device__ __forceinline__ bool round_sweeps(const RoundCtr *in, uint32_t found, uint64_t edges, uint32_t qcap,
                                             float factor, uint64_t min_edges) {
    if (!min_edges || !found || edges <= qcap || in->mode != kModeFlags || !in->ucount) return false;
    uint64_t u = 0, d = 0;
#pragma unroll
    for (int i = 0; i < kShards; i++) {
        u += in->uedges_sh[i];
        d += in->dedges_sh[i];
    }
    const uint64_t push = edges - min(d, edges);
    if (push <= min_edges) return false;
    return !(factor > 0) || (double)u * factor > (double)min_edges;
}

// Per-thread tallies, summed per workgroup at the end of a round.
struct Acc {
    uint32_t found = 0;
    uint32_t far = 0;
    uint64_t edges = 0;
    uint32_t fmin_inv = 0;  // max of ~value over the vertices put into the output frontier
};

// Output side of a round.
struct Emit {
    uint32_t *q;                 // this workgroup's shard of the output S queue
    uint32_t *qn;                // its count
    uint32_t *cq, *cb;           // this workgroup's shard of the output chunk queue
    unsigned long long *cpack;   // its packed count
    uint32_t *any;               // set to 1 when anything is emitted
    uint32_t mode;               // kModeQueue, kModeFlags or kModeBitmap
    uint32_t sd;
};

__device__ __forceinline__ uint32_t wave_chunks(uint32_t d) { return (d + kWaveChunk - 1) / kWaveChunk; }

__device__ __forceinline__ Emit make_emit(const RoundEnv &e, RoundCtr *out, uint32_t mode) {
    const uint32_t shard = blockIdx.x % kShards;
    return Emit{e.qout + (uint64_t)shard * kQueueCap, &out->qn[shard],
                e.cqout + (uint64_t)shard * e.cq_cap, e.cbout + (uint64_t)shard * e.cq_cap,
                &out->cpack[shard], &out->any, mode, e.sd};
}

// Marks the output non-empty (one plain store per wave that has pred).
__device__ __forceinline__ void mark_any(const Emit &em, bool pred) {
    const uint64_t m = ballot(pred);
    if (m && lane_id() == lowest_lane(m)) *em.any = 1u;
}

// Appends chunk-queue entries (d > sd) with one packed atomic per wave.
__device__ __forceinline__ void append_chunks(uint32_t *cq, uint32_t *cb, unsigned long long *cpack, bool pc,
                                              uint32_t w, uint32_t d) {
    const uint64_t mc = ballot(pc);
    if (!mc) return;
    const uint32_t nch = pc ? wave_chunks(d) : 0u;
    const uint32_t incl = wave_incl_scan_u32(nch);
    const uint32_t total = read_lane(incl, 63);
    const uint32_t leader = lowest_lane(mc);
    unsigned long long old = 0;
    if (lane_id() == leader)
        old = atomicAdd(cpack, ((unsigned long long)__popcll(mc) << kLCountShift) | (unsigned long long)total);
    old = read_lane64(old, leader);
    if (pc) {
        const uint32_t slot = (uint32_t)(old >> kLCountShift) + rank_below(mc);
        cq[slot] = w;
        cb[slot] = (uint32_t)(old & kLChunkMask) + (incl - nch);
    }
}

// Records vertex w (out-degree d), newly in the next frontier, for every lane
// with pred, in queue or bitmap mode (the caller set its bitmap bit). Counts
// it, and in queue mode appends it to the S or chunk queue. Wave-uniform.
__device__ __forceinline__ void emit(const Emit &em, bool pred, uint32_t w, uint32_t d, Acc &acc) {
    if (pred) {
        acc.found++;
        acc.edges += d;
    }
    mark_any(em, pred);
    if (em.mode != kModeQueue) return;
    wave_append(em.q, em.qn, pred && d > 0 && d <= em.sd, w);
    append_chunks(em.cq, em.cb, em.cpack, pred && d > em.sd, w, d);
}

struct InSummary {
    uint32_t found;
    uint32_t far;
    uint64_t edges;
};

__device__ __forceinline__ InSummary summarize(const RoundCtr *in) {
    InSummary s{0, 0, 0};
#pragma unroll
    for (int i = 0; i < kShards; i++) {
        s.found += in->found_sh[i];
        s.far += in->far_sh[i];
        s.edges += in->edges_sh[i];
    }
    return s;
}

// BFS: a bottom-up round hands back to top-down once its output frontier
// has fewer than V / beta vertices (Beamer, Asanovic, Patterson 2012). Both
// the pre-pass and the round decide with this one function.
__device__ __forceinline__ bool bfs_leaves_bottom_up(uint32_t found, uint32_t V, float beta) {
    return (double)found < (double)V / beta;
}

// Zeroes the ring entry two rounds ahead (done by every round, active or
// not, so rounds launched past the end of a search always find zeros).
static_assert(sizeof(RoundCtr) / 4 <= kBlock, "one workgroup clears a ring entry");
__device__ __forceinline__ void clear_ring_ahead(const RoundEnv &e) {
    if (blockIdx.x == 0 && threadIdx.x < sizeof(RoundCtr) / 4)
        reinterpret_cast<uint32_t *>(e.ring + (e.r + 2) % kRing)[threadIdx.x] = 0u;
}

// Zeroes the frontier bitmap the round after next will write.
__device__ __forceinline__ void clear_bitmap_ahead(const RoundEnv &e) {
    const uint32_t stride = gridDim.x * blockDim.x;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < e.nwords; i += stride) e.fclr[i] = 0u;
}

// Adds the workgroup's tallies to shard blockIdx % kShards of `out`.
// Every thread of the workgroup must call it.
__device__ __forceinline__ void add_tallies(RoundCtr *out, const Acc &acc) {
    __shared__ uint32_t s_found[kWavesPerBlock], s_far[kWavesPerBlock], s_fmin[kWavesPerBlock];
    __shared__ unsigned long long s_edges[kWavesPerBlock];
    const uint32_t f = (uint32_t)wave_sum_u64(acc.found);
    const uint32_t fa = (uint32_t)wave_sum_u64(acc.far);
    const uint64_t ed = wave_sum_u64(acc.edges);
    const uint32_t fm = wave_max_u32(acc.fmin_inv);
    const uint32_t wib = threadIdx.x >> 6;
    if (lane_id() == 0) {
        s_found[wib] = f;
        s_far[wib] = fa;
        s_edges[wib] = ed;
        s_fmin[wib] = fm;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t tf = 0, tfa = 0, tm = 0;
        unsigned long long te = 0;
        for (int w = 0; w < kWavesPerBlock; w++) {
            tf += s_found[w];
            tfa += s_far[w];
            te += s_edges[w];
            tm = max(tm, s_fmin[w]);
        }
        const uint32_t shard = blockIdx.x % kShards;
        if (tf) atomicAdd(&out->found_sh[shard], tf);
        if (tfa) atomicAdd(&out->far_sh[shard], tfa);
        if (te) atomicAdd(&out->edges_sh[shard], te);
        if (tm && tm > out->fmin_inv) atomicMax(&out->fmin_inv, tm);  // no read when there is nothing to add
    }
}

// Ops that defer high-degree vertices (SSSP relaxation) provide
// defers(degree, value) and requeue(pred, v, degree, value, emit, acc).
template <class Op, class = void>
struct HasDefer {
    static constexpr bool value = false;
};
template <class Op>
struct HasDefer<Op, decltype((void)&Op::requeue)> {
    static constexpr bool value = true;
};

// Edges are taken kGroup at a time: every edge's loads (column, weight, the
// target's value) are issued before any is applied (compare, atomic, emit),
// so a lane keeps kGroup random reads in flight. A value read before another
// edge of the group is applied can only be larger than the current one
// (values only fall), so it lets at most more candidates reach the atomic,
// which decides. Ops with an issue() step send the group's atomics one after
// another before any result is used, so their round trips overlap too.
constexpr uint32_t kGroup = 4;

template <class Op, class = void>
struct HasIssue {
    static constexpr bool value = false;
};
template <class Op>
struct HasIssue<Op, decltype((void)&Op::issue)> {
    static constexpr bool value = true;
};

// load -> (issue) -> apply over a group of n edges.
template <class Op, uint32_t n, class Pre>
__device__ __forceinline__ void issue_group(const Op &op, Pre (&pre)[n]) {
    if constexpr (HasIssue<Op>::value) {
#pragma unroll
        for (uint32_t j = 0; j < n; j++) op.issue(pre[j]);
    }
}

// Lane-per-vertex expansion of the tile's vertices of degree 1..sd (`mem`
// says whether the lane holds a frontier vertex). Wave-uniform call.
template <class Op>
__device__ __forceinline__ void process_tile(const RoundEnv &e, bool mem, uint32_t v, Op &op, const Emit &em,
                                             Acc &acc) {
    uint32_t b = 0, d = 0;
    if (mem) {
        b = e.off[v];
        d = e.off[v + 1] - b;
    }
    const bool s = mem && d > 0 && d <= e.sd;
    const uint32_t ds = s ? d : 0u;
    const uint32_t sv = s ? op.vertex(v) : 0u;
    const uint32_t md = wave_max_u32(ds);
    for (uint32_t k = 0; k < md; k += kGroup) {
        typename Op::Pre pre[kGroup];
#pragma unroll
        for (uint32_t j = 0; j < kGroup; j++) pre[j] = op.load(k + j < ds, sv, (uint64_t)b + k + j);
        issue_group(op, pre);
#pragma unroll
        for (uint32_t j = 0; j < kGroup; j++)
            if (k + j < md) op.apply(pre[j], em, acc);
    }
}

// Sets the wave's issue priority (0..3; s_setprio takes an immediate).
__device__ __forceinline__ void set_wave_prio(uint32_t p) {
    switch (p) {
        case 0: __builtin_amdgcn_s_setprio(0); break;
        case 1: __builtin_amdgcn_s_setprio(1); break;
        case 2: __builtin_amdgcn_s_setprio(2); break;
        default: __builtin_amdgcn_s_setprio(3); break;
    }
}

// Walks the chunk queue: wave gw takes the contiguous global chunk range
// [C*gw/nw, C*(gw+1)/nw), finds its first entry by binary search and then
// steps through the chunks (kWaveChunk = 64 edges, a lane each), up to kGroup
// consecutive chunks of an entry at a time. The entries'
// ids, offsets and values are loaded 64 at a time, a lane each, and read
// back with readlane, so a short entry costs no load latency of its own.
// Equal shares do not finish together: the sequencer issues for the oldest
// wave first, and later-dispatched waves took up to 2.8x longer per edge. So
// a wave starts its range at priority 3 and steps down a level per quarter
// done: waves that are behind are issued for first.
template <class Op>
__device__ __forceinline__ void expand_chunks(const RoundEnv &e, const RoundCtr *in, Op &op, const Emit &em,
                                              Acc &acc, uint32_t gw, uint32_t nw) {
    const uint32_t lane = lane_id();
    uint32_t cpre[kShards + 1];
    cpre[0] = 0;
#pragma unroll
    for (int s = 0; s < kShards; s++) cpre[s + 1] = cpre[s] + (uint32_t)(in->cpack[s] & kLChunkMask);
    const uint32_t C = cpre[kShards];
    const uint32_t c0 = (uint32_t)((uint64_t)C * gw / nw), c1 = (uint32_t)((uint64_t)C * (gw + 1) / nw);
    if (c0 >= c1) return;
    int s = 0;
    while (s < kShards - 1 && c0 >= cpre[s + 1]) s++;
    const uint32_t *ids = e.cqin + (uint64_t)s * e.cq_cap;
    const uint32_t *bases = e.cbin + (uint64_t)s * e.cq_cap;
    uint32_t n = (uint32_t)(in->cpack[s] >> kLCountShift);
    const uint32_t local = c0 - cpre[s];
    uint32_t lo = 0, hi = n;
    while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) >> 1;
        if (bases[mid] <= local) lo = mid; else hi = mid;
    }
    uint32_t idx = lo;
    uint32_t k = local - bases[idx];
    uint32_t c = c0;
    const uint32_t len = c1 - c0;
    uint32_t quarter = 0, next_q = c0 + (len + 3) / 4;
    set_wave_prio(3);
    while (c < c1) {
        // entries [idx, idx + nb) of shard s, one per lane
        const uint32_t nb = min(64u, n - idx);
        uint32_t mv = 0, mb = 0, md = 0, ms = 0;
        if (lane < nb) {
            mv = ids[idx + lane];
            mb = e.off[mv];
            md = e.off[mv + 1] - mb;
            ms = op.vertex(mv);
        }
        for (uint32_t i = 0; i < nb && c < c1; i++) {
            const uint32_t v = __builtin_amdgcn_readlane(mv, i);
            const uint32_t vb = __builtin_amdgcn_readlane(mb, i);
            const uint32_t vd = __builtin_amdgcn_readlane(md, i);
            const uint32_t sv = __builtin_amdgcn_readlane(ms, i);
            const uint32_t nch = wave_chunks(vd);
            bool skip = false;
            if constexpr (HasDefer<Op>::value) {
                // A deferred entry (only queue frontiers reach here with one:
                // the pre-pass filters the others) goes back into the output
                // frontier once, from the wave that holds its first chunk. A
                // wave that read an older, larger value and skipped a later
                // chunk of an entry that another wave expanded left it to the
                // improvement in between, which put the vertex into the output.
                skip = op.defers(vd, sv);
                if (skip) {
                    if (k == 0) op.requeue(lane == 0, v, vd, sv, em, acc);
                    const uint32_t left = min(nch - k, c1 - c);
                    c += left;
                    k += left;
                }
            }
            // up to kGroup consecutive chunks of the entry at once
            while (k < nch && c < c1) {
                if (c >= next_q) {
                    quarter++;
                    next_q = c0 + (uint32_t)(((uint64_t)len * (quarter + 1) + 3) / 4);
                    set_wave_prio(3 - min(quarter, 3u));
                }
                const uint32_t gn = min(min(kGroup, nch - k), c1 - c);
                const uint32_t b = k * kWaveChunk;
                const uint32_t end = min(vd, b + kWaveChunk * gn);
                static_assert(kWaveChunk == 64, "a chunk is one 64-edge wave step");
                typename Op::Pre pre[kGroup];
#pragma unroll
                for (uint32_t j = 0; j < kGroup; j++) {
                    const uint32_t t = b + 64 * j + lane;
                    pre[j] = op.load(t < end, sv, (uint64_t)vb + t);
                }
                issue_group(op, pre);
#pragma unroll
                for (uint32_t j = 0; j < kGroup; j++)
                    if (b + 64 * j < end) op.apply(pre[j], em, acc);
                k += gn;
                c += gn;
            }
            if (k >= nch) k = 0;  // the next entry starts at its first chunk
        }
        idx += nb;
        if (idx >= n && c < c1) {  // on to the next non-empty shard
            do {
                s++;
            } while (s < kShards - 1 && cpre[s + 1] == cpre[s]);
            ids = e.cqin + (uint64_t)s * e.cq_cap;
            n = (uint32_t)(in->cpack[s] >> kLCountShift);
            idx = 0;
        }
    }
    set_wave_prio(0);
}

// Expands the frontier of round e.r: the chunk queue, then the degree <= sd
// vertices from the S queue or the bitmap.
template <class Op>
__device__ __forceinline__ void expand_frontier(const RoundEnv &e, const RoundCtr *in, Op &op, const Emit &em,
                                                Acc &acc) {
    const uint32_t lane = lane_id();
    const uint32_t gw = blockIdx.x * kWavesPerBlock + (threadIdx.x >> 6);
    const uint32_t nw = gridDim.x * kWavesPerBlock;
    expand_chunks(e, in, op, em, acc, gw, nw);
    if (in->mode == kModeQueue) {
        uint32_t pre[kShards + 1];
        pre[0] = 0;
#pragma unroll
        for (int s = 0; s < kShards; s++) pre[s + 1] = pre[s] + in->qn[s];
        const uint32_t total = pre[kShards];
        for (uint32_t t = gw * 64; t < total; t += nw * 64) {
            const uint32_t i = t + lane;
            const bool valid = i < total;
            uint32_t v = 0;
            if (valid) {
                int s = 0;
                while (s < kShards - 1 && i >= pre[s + 1]) s++;
                v = e.qin[(uint64_t)s * kQueueCap + (i - pre[s])];
            }
            process_tile(e, valid, v, op, em, acc);
        }
    } else {
        for (uint32_t p = gw; p < e.npairs; p += nw) {
            const uint32_t bits = e.fin[2 * p + (lane >> 5)];
            const bool mem = (bits >> (lane & 31)) & 1u;
            if (ballot(mem) == 0) continue;
            process_tile(e, mem, p * 64 + lane, op, em, acc);
        }
    }
}

// Puts the pre-pass's deferred vertices of round e.r back into the output
// frontier, a lane per vertex, and clears their bitmap words.
template <class Op>
__device__ __forceinline__ void requeue_deferred(const RoundEnv &e, const RoundCtr *in, Op &op, const Emit &em,
                                                 Acc &acc) {
    if (!in->ndefer) return;
    const uint32_t lane = lane_id();
    const uint32_t gw = blockIdx.x * kWavesPerBlock + (threadIdx.x >> 6);
    const uint32_t nw = gridDim.x * kWavesPerBlock;
    for (uint32_t p = gw; p < e.npairs; p += nw) {
        const uint32_t bits = e.dbits[2 * p + (lane >> 5)];
        const bool mem = (bits >> (lane & 31)) & 1u;
        if (ballot(mem) == 0) continue;
        if ((lane & 31) == 0 && bits) e.dbits[2 * p + (lane >> 5)] = 0u;
        const uint32_t v = p * 64 + lane;
        const uint32_t d = mem ? e.off[v + 1] - e.off[v] : 0u;
        op.requeue(mem, v, d, mem ? op.vertex(v) : 0u, em, acc);
    }
}

// ---- host side ---------------------------------------------------------------

RoundEnv make_env(ptg_graph *g, uint32_t r);
// Same over another CSR with the graph's vertex set (SSSP's light edges).
RoundEnv make_env_csr(ptg_graph *g, uint32_t r, const uint32_t *off, const uint32_t *col);
// Clears the frontier bitmaps, byte flags and counter ring before a search.
// Clears the frontier bitmaps, flags and counter ring, and the caller's
// `extra` spans (the query's value array and the like), in one launch.
int reset_rounds(ptg_graph *g, std::initializer_list<FillSpan> extra = {});
// Seeds round 0 with one source of value src_value (queue mode).
int launch_seed(ptg_graph *g, uint32_t src, uint32_t src_value, uint32_t *val, uint32_t thr);
int launch_seed_csr(ptg_graph *g, const uint32_t *off, const uint32_t *col, uint32_t src, uint32_t src_value,
                    uint32_t *val, uint32_t thr);
// Pre-pass of round r (a no-op for queue-mode frontiers).
int launch_prepare(ptg_graph *g, const RoundEnv &e);
int read_ring(ptg_graph *g, uint32_t from, uint32_t count);

struct HostSummary {
    uint32_t found, far;
    uint64_t edges;
};

inline HostSummary host_summary(const RoundCtr &c) {
    HostSummary s{0, 0, 0};
    for (int i = 0; i < kShards; i++) {
        s.found += c.found_sh[i];
        s.far += c.far_sh[i];
        s.edges += c.edges_sh[i];
    }
    return s;
}

// Launches rounds (pre-pass + round kernel, via launch(r)) in growing
// batches until a ring entry is idle. After a batch that ended with round r-1,
// entries 0..r-1 are complete and entry r holds only its header and `any`
// flag, so idleness is judged on every entry but counts on complete ones.
// Returns the number of rounds that did work and the sums of their counts.
// `finish(speculative)` enqueues the query's closing work (end event,
// statistics, output delivery). With a hint it goes out behind the first
// batch, before that batch's readback, so a query whose depth repeats waits
// once; *finished says whether that batch was the last (otherwise the caller
// enqueues it with speculative = false).
template <class Launch, class Idle, class Finish>
int drive_rounds(ptg_graph *g, Launch launch, Idle idle, uint32_t &hint, uint32_t *rounds_out, uint64_t *found_sum,
                 uint64_t *edges_sum, Finish finish, bool *finished) {
    // The first batch covers the previous query's rounds + 1 (one readback when
    // the depth repeats); later batches start at 4 and double up to 16. Rounds
    // launched past the end find an idle entry and return at once.
    uint32_t r = 0, batch = hint ? std::min(hint + 1, 64u) : 2u, rounds = 0;
    const bool hinted = hint != 0;
    bool first = true;
    *finished = false;
    uint64_t fs = 0, es = 0;
    uint32_t checked = 0;  // ring entries [0, checked) already summed
    g->frontiers.clear();
    g->pulled_rounds = 0;
    g->pulled_mask = 0;
    g->swept_rounds = 0;
    g->swept_mask = 0;
    for (;;) {
        for (uint32_t k = 0; k < batch; k++, r++) {
            int rc = launch(r);
            if (rc) return rc;
        }
        const bool spec = hinted && first;
        first = false;
        if (spec) {
            int rc = finish(true);
            if (rc) return rc;
        }
        int rc = read_ring(g, checked, r + 1 - checked);
        if (rc) return rc;
        bool done = false;
        for (; checked <= r; checked++) {
            const RoundCtr &c = g->h_ring[checked % kRing];
            if (idle(c)) {
                done = true;
                *finished = spec;
                break;
            }
            if (checked == r) break;  // not prepared yet: summed after the next batch
            const HostSummary s = host_summary(c);
            fs += s.found;
            es += s.edges;
            g->frontiers.push_back(s.found);
            g->pulled_rounds += c.pulled;
            if (c.pulled && checked < 64) g->pulled_mask |= 1ull << checked;
            g->swept_rounds += c.swept;
            if (c.swept && checked < 64) g->swept_mask |= 1ull << checked;
            rounds++;
        }
        if (done) break;
        batch = hinted && r <= hint + 1 ? 4u : std::min(batch * 2, 16u);
    }
    hint = rounds;
    *rounds_out = rounds;
    if (found_sum) *found_sum = fs;
    if (edges_sum) *edges_sum = es;
    return PTG_OK;
}

}  // namespace ptg
