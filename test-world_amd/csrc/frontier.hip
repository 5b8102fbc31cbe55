// Frontier kernels for gfx950: BFS (top-down + bottom-up), SSSP (near/far)
// and CC (union-find hooking, or label propagation on directed graphs).
//
// Replaces ptGraph's per-iteration pipeline (setLabeling -> thrust::reduce ->
// thrust::exclusive_scan -> setActiveNodeArray -> setLabelDefault -> relax,
// ptGraph/main.cu:1094-1106, kernels ptGraph/gpu_kernels.cu:3-172) by one
// persistent-grid kernel per round:
//   * a frontier is always a bitmap (three rotate: the round reads one, writes
//     the next and clears the third); vertices of degree <= small_degree are
//     expanded one lane each, straight from the bitmap or, while the frontier
//     is small, from a 16-way sharded queue; every other vertex sits in a
//     sharded chunk queue of kWaveChunk-edge chunks that waves take in equal
//     contiguous ranges, so hubs and heavy id ranges are spread over the GPU;
//   * counters are summed per workgroup and added to 1 of 16 shards, because a
//     single hot address serialises atomics (measured 11 ns each on MI355X);
//   * every decision (BFS direction, SSSP near/far split, queue vs bitmap) is
//     made on the device from the previous round's counters, so the host only
//     polls for termination every few rounds;
//   * relaxation keeps the reference's strict `<` test followed by atomicMin
//     (gpu_kernels.cu:127-131, :143-147, :165-169).
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
// Static-bin expansion (CC hooking, symmetry checks): walks the L, M and S
// bins of a fixed vertex set and calls op.edge() from wave-uniform control.
// ---------------------------------------------------------------------------

template <class Op>
__global__ __launch_bounds__(kBlock) void k_expand(Frontier in, const uint32_t *__restrict__ off, Op op) {
    const uint32_t lane = lane_id();
    const uint32_t wib = threadIdx.x >> 6;
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
        const uint32_t sv = op.vertex(v);
        for (uint64_t base = b + wib * 64; base < e; base += kBlock) op.edge(base + lane < e, sv, base + lane);
    }
    const uint32_t gw = blockIdx.x * kWavesPerBlock + wib;
    const uint32_t nw = gridDim.x * kWavesPerBlock;
    const uint32_t nm = in.ctr->n_m;
    for (uint32_t i = gw; i < nm; i += nw) {
        const uint32_t v = in.qm[i];
        const uint64_t b = off[v], e = off[v + 1];
        const uint32_t sv = op.vertex(v);
        for (uint64_t base = b; base < e; base += 64) op.edge(base + lane < e, sv, base + lane);
    }
    const uint32_t ns = in.ctr->n_s;
    for (uint64_t base = (uint64_t)gw * 64; base < ns; base += (uint64_t)nw * 64) {
        const uint64_t i = base + lane;
        const bool valid = i < ns;
        uint32_t b = 0, d = 0, sv = 0;
        if (valid) {
            const uint32_t v = in.qs[i];
            b = off[v];
            d = off[v + 1] - b;
            sv = op.vertex(v);
        }
        const uint32_t md = wave_max_u32(d);
        for (uint32_t j = 0; j < md; j++) op.edge(j < d, sv, (uint64_t)b + j);
    }
}

// ---------------------------------------------------------------------------
// Rounds: shared machinery
// ---------------------------------------------------------------------------

struct RoundEnv {
    RoundCtr *ring;
    uint32_t r;                   // round index
    const uint32_t *off, *col;
    uint32_t *fin, *fout, *fclr;  // frontier bitmaps: read, write, clear
    uint32_t *qin;                // sharded S queue read by this round
    uint32_t *qout;               // sharded S queue written by this round
    uint32_t *cqin, *cbin;        // sharded chunk queue read (ids, first chunks)
    uint32_t *cqout, *cbout;      // sharded chunk queue written
    uint32_t cq_cap;              // chunk-queue entries per shard
    uint32_t V, npairs, nwords;
    uint32_t sd;                  // small degree threshold
    uint32_t qcap;                // frontier edges up to which the output is queued
};

// Per-thread tallies, summed per workgroup at the end of a round.
struct Acc {
    uint32_t found = 0;
    uint32_t far = 0;
    uint64_t edges = 0;
};

// Output side of a round.
struct Emit {
    uint32_t *q;                 // this workgroup's shard of the output S queue
    uint32_t *qn;                // its count
    uint32_t *cq, *cb;           // this workgroup's shard of the output chunk queue
    unsigned long long *cpack;   // its packed count
    bool queue;                  // output frontier is queued (small) or bitmap-only
    uint32_t sd;
};

__device__ __forceinline__ uint32_t wave_chunks(uint32_t d) { return (d + kWaveChunk - 1) / kWaveChunk; }

__device__ __forceinline__ Emit make_emit(const RoundEnv &e, RoundCtr *out, bool queue) {
    const uint32_t shard = blockIdx.x % kShards;
    return Emit{e.qout + (uint64_t)shard * kQueueCap, &out->qn[shard],
                e.cqout + (uint64_t)shard * e.cq_cap, e.cbout + (uint64_t)shard * e.cq_cap,
                &out->cpack[shard], queue, e.sd};
}

// Records newly discovered vertex w (out-degree d) for every lane with pred.
// In queue mode it is appended to the S queue (d <= sd) or the chunk queue
// (d > sd); in bitmap mode the caller's bit in the output bitmap is all the
// next round needs. Wave-uniform call.
__device__ __forceinline__ void emit(const Emit &em, bool pred, uint32_t w, uint32_t d, Acc &acc) {
    if (pred) {
        acc.found++;
        acc.edges += d;
    }
    if (!em.queue) return;
    wave_append(em.q, em.qn, pred && d > 0 && d <= em.sd, w);
    const bool pc = pred && d > em.sd;
    const uint64_t mc = ballot(pc);
    if (mc) {
        const uint32_t nch = pc ? wave_chunks(d) : 0u;
        const uint32_t incl = wave_incl_scan_u32(nch);
        const uint32_t total = read_lane(incl, 63);
        const uint32_t leader = lowest_lane(mc);
        unsigned long long old = 0;
        if (lane_id() == leader)
            old = atomicAdd(em.cpack, ((unsigned long long)__popcll(mc) << kLCountShift) | (unsigned long long)total);
        old = read_lane64(old, leader);
        if (pc) {
            const uint32_t slot = (uint32_t)(old >> kLCountShift) + rank_below(mc);
            em.cq[slot] = w;
            em.cb[slot] = (uint32_t)(old & kLChunkMask) + (incl - nch);
        }
    }
}

struct InSummary {
    uint32_t found;
    uint32_t far;
    uint64_t edges;
};

__device__ __forceinline__ InSummary summarize(const RoundCtr *in) {
    InSummary s{0, in->far_base, 0};
#pragma unroll
    for (int i = 0; i < kShards; i++) {
        s.found += in->found_sh[i];
        s.far += in->far_sh[i];
        s.edges += in->edges_sh[i];
    }
    return s;
}

// Zeroes the ring entry two rounds ahead (done by every round, active or
// not, so rounds launched past the end of a search always find zeros).
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
__device__ __forceinline__ void finish_round(RoundCtr *out, Acc acc) {
    __shared__ uint32_t s_found[kWavesPerBlock], s_far[kWavesPerBlock];
    __shared__ unsigned long long s_edges[kWavesPerBlock];
    const uint32_t f = (uint32_t)wave_sum_u64(acc.found);
    const uint32_t fa = (uint32_t)wave_sum_u64(acc.far);
    const uint64_t ed = wave_sum_u64(acc.edges);
    const uint32_t wib = threadIdx.x >> 6;
    if (lane_id() == 0) {
        s_found[wib] = f;
        s_far[wib] = fa;
        s_edges[wib] = ed;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t tf = 0, tfa = 0;
        unsigned long long te = 0;
        for (int w = 0; w < kWavesPerBlock; w++) {
            tf += s_found[w];
            tfa += s_far[w];
            te += s_edges[w];
        }
        const uint32_t shard = blockIdx.x % kShards;
        if (tf) atomicAdd(&out->found_sh[shard], tf);
        if (tfa) atomicAdd(&out->far_sh[shard], tfa);
        if (te) atomicAdd(&out->edges_sh[shard], te);
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
    for (uint32_t k = 0; k < md; k++) op.edge(k < ds, sv, (uint64_t)b + k, em, acc);
}

// Walks the chunk queue: wave gw takes the contiguous global chunk range
// [C*gw/nw, C*(gw+1)/nw), finds its first entry by binary search and then
// steps through chunks (kWaveChunk edges, 64 lanes per step).
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
    uint32_t local = c0 - cpre[s];
    uint32_t lo = 0, hi = n;
    while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) >> 1;
        if (bases[mid] <= local) lo = mid; else hi = mid;
    }
    uint32_t idx = lo;
    uint32_t v = ids[idx];
    uint32_t vb = e.off[v], vd = e.off[v + 1] - vb;
    uint32_t k = local - bases[idx];
    uint32_t sv = op.vertex(v);
    for (uint32_t c = c0; c < c1; c++) {
        while (k >= wave_chunks(vd)) {  // next entry, possibly in the next shard
            k = 0;
            if (++idx >= n) {
                do {
                    s++;
                } while (s < kShards - 1 && cpre[s + 1] == cpre[s]);
                ids = e.cqin + (uint64_t)s * e.cq_cap;
                n = (uint32_t)(in->cpack[s] >> kLCountShift);
                idx = 0;
            }
            v = ids[idx];
            vb = e.off[v];
            vd = e.off[v + 1] - vb;
            sv = op.vertex(v);
        }
        const uint32_t b = k * kWaveChunk;
        const uint32_t end = min(vd, b + kWaveChunk);
        for (uint32_t t = b; t < end; t += 64) op.edge(t + lane < end, sv, (uint64_t)vb + t + lane, em, acc);
        k++;
    }
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

// Pre-pass of a bitmap frontier: puts every member of degree > sd into the
// chunk queue of round e.r. Workgroup b scans tiles [b*T, (b+1)*T), gathers
// members in LDS and appends them to shard b % 16 with one atomic per batch
// (plus an LDS scan for the chunk bases).
__global__ __launch_bounds__(kBlock) void k_chunkify(RoundEnv e) {
    constexpr uint32_t kBuf = 1024;
    __shared__ uint32_t s_v[kBuf], s_n[kBuf];
    __shared__ uint32_t s_cnt, s_wsum[kWavesPerBlock];
    __shared__ unsigned long long s_old;
    RoundCtr *in = e.ring + e.r % kRing;
    if (in->mode != kModeBitmap) return;
    const uint32_t lane = lane_id(), wib = threadIdx.x >> 6;
    const uint32_t T = (e.npairs + gridDim.x - 1) / gridDim.x;
    const uint32_t t_begin = blockIdx.x * T;
    const uint32_t t_end = min(e.npairs, t_begin + T);
    const uint32_t shard = blockIdx.x % kShards;
    uint32_t *ids = e.cqin + (uint64_t)shard * e.cq_cap;
    uint32_t *bases = e.cbin + (uint64_t)shard * e.cq_cap;
    if (threadIdx.x == 0) s_cnt = 0;
    __syncthreads();
    for (uint32_t t0 = t_begin; t0 < t_end; t0 += kWavesPerBlock) {
        const uint32_t p = t0 + wib;
        bool want = false;
        uint32_t v = p * 64 + lane, d = 0;
        if (p < t_end && ((e.fin[2 * p + (lane >> 5)] >> (lane & 31)) & 1u)) {
            d = e.off[v + 1] - e.off[v];
            want = d > e.sd;
        }
        const uint64_t m = ballot(want);
        if (m) {
            uint32_t base = 0;
            if (lane == lowest_lane(m)) base = atomicAdd(&s_cnt, (uint32_t)__popcll(m));
            base = read_lane(base, lowest_lane(m));
            if (want) {
                s_v[base + rank_below(m)] = v;
                s_n[base + rank_below(m)] = wave_chunks(d);
            }
        }
        __syncthreads();
        const uint32_t cnt = s_cnt;
        if (cnt > kBuf - kBlock || t0 + kWavesPerBlock >= t_end) {
            // exclusive scan of chunk counts: 4 entries per thread
            uint32_t loc[4], sum = 0;
#pragma unroll
            for (int j = 0; j < 4; j++) {
                const uint32_t i = threadIdx.x * 4 + j;
                loc[j] = i < cnt ? s_n[i] : 0u;
                sum += loc[j];
            }
            const uint32_t incl = wave_incl_scan_u32(sum);
            if (lane == 63) s_wsum[wib] = incl;
            __syncthreads();
            uint32_t woff = 0, total = 0;
            for (int w = 0; w < kWavesPerBlock; w++) {
                if (w < (int)wib) woff += s_wsum[w];
                total += s_wsum[w];
            }
            if (threadIdx.x == 0 && cnt)
                s_old = atomicAdd(&in->cpack[shard], ((unsigned long long)cnt << kLCountShift) | total);
            __syncthreads();
            const uint32_t slot0 = (uint32_t)(s_old >> kLCountShift);
            uint32_t run = (uint32_t)(s_old & kLChunkMask) + woff + (incl - sum);
#pragma unroll
            for (int j = 0; j < 4; j++) {
                const uint32_t i = threadIdx.x * 4 + j;
                if (i < cnt) {
                    ids[slot0 + i] = s_v[i];
                    bases[slot0 + i] = run;
                }
                run += loc[j];
            }
            __syncthreads();
            if (threadIdx.x == 0) s_cnt = 0;
            __syncthreads();
        }
    }
}

static inline RoundEnv make_env(ptg_graph *g, uint32_t r, uint32_t qcap) {
    RoundEnv e;
    e.ring = g->d_ring;
    e.r = r;
    e.off = g->out.off;
    e.col = g->out.col;
    e.fin = g->d_bits[r % 3];
    e.fout = g->d_bits[(r + 1) % 3];
    e.fclr = g->d_bits[(r + 2) % 3];
    e.qin = g->d_queue[r & 1];
    e.qout = g->d_queue[(r + 1) & 1];
    e.cqin = g->d_cq[r & 1];
    e.cbin = g->d_cb[r & 1];
    e.cqout = g->d_cq[(r + 1) & 1];
    e.cbout = g->d_cb[(r + 1) & 1];
    e.cq_cap = g->cq_cap;
    e.V = g->V;
    e.npairs = g->npairs;
    e.nwords = g->nwords;
    e.sd = g->cfg.small_degree;
    e.qcap = qcap;
    return e;
}

// Seeds round 0 with a single source (value src_value) in queue mode.
__global__ void k_seed(RoundEnv e, uint32_t src, uint32_t src_value, uint32_t *val, uint32_t *vis, uint32_t thr) {
    if (threadIdx.x != 0) return;
    RoundCtr *c = e.ring;
    const uint32_t d = e.off[src + 1] - e.off[src];
    val[src] = src_value;
    if (vis) vis[src >> 5] |= 1u << (src & 31);
    if (d > 0) e.fin[src >> 5] |= 1u << (src & 31);
    c->mode = kModeQueue;
    c->dir = kDirTopDown;
    c->thr = thr;
    c->found_sh[0] = 1;
    c->edges_sh[0] = d;
    if (d > e.sd) {
        e.cqin[0] = src;  // shard 0 of the chunk queue round 0 reads
        e.cbin[0] = 0;
        c->cpack[0] = (1ull << kLCountShift) | wave_chunks(d);
    } else if (d > 0) {
        e.qin[0] = src;  // shard 0 of the S queue round 0 reads
        c->qn[0] = 1;
    }
}

static int read_ring(ptg_graph *g, uint32_t from, uint32_t count) {
    // Copies ring entries [from, from + count) (mod kRing) into h_ring.
    uint32_t a = from % kRing;
    uint32_t first = std::min(count, kRing - a);
    PTG_HIP(hipMemcpyAsync(g->h_ring + a, g->d_ring + a, sizeof(RoundCtr) * first, hipMemcpyDeviceToHost,
                           g->stream));
    if (count > first)
        PTG_HIP(hipMemcpyAsync(g->h_ring, g->d_ring, sizeof(RoundCtr) * (count - first), hipMemcpyDeviceToHost,
                               g->stream));
    PTG_HIP(hipStreamSynchronize(g->stream));
    return PTG_OK;
}

static int clear_ring(ptg_graph *g) {
    return fill_u32(g, reinterpret_cast<uint32_t *>(g->d_ring), sizeof(RoundCtr) / 4 * kRing, 0u);
}

struct HostSummary {
    uint32_t found, far;
    uint64_t edges;
};

static HostSummary host_summary(const RoundCtr &c) {
    HostSummary s{0, c.far_base, 0};
    for (int i = 0; i < kShards; i++) {
        s.found += c.found_sh[i];
        s.far += c.far_sh[i];
        s.edges += c.edges_sh[i];
    }
    return s;
}

static int clear_frontier_bitmaps(ptg_graph *g) {
    for (int i = 0; i < 3; i++) {
        int rc = fill_u32(g, g->d_bits[i], g->nwords, 0u);
        if (rc) return rc;
    }
    return PTG_OK;
}

// Launches rounds in growing batches until one is idle. `launch(r)` enqueues
// round r; `idle(c)` says whether ring entry c describes a finished search.
// Returns the number of rounds that did work and the sums of their frontiers.
template <class Launch, class Idle>
static int drive_rounds(ptg_graph *g, Launch launch, Idle idle, uint32_t *rounds_out, uint64_t *found_sum,
                        uint64_t *edges_sum) {
    uint32_t r = 0, batch = 2, rounds = 0;
    uint64_t fs = 0, es = 0;
    uint32_t checked = 0;  // ring entries [0, checked) already summed
    for (;;) {
        for (uint32_t k = 0; k < batch; k++, r++) {
            int rc = launch(r);
            if (rc) return rc;
        }
        // entries 0..r are final (round r - 1 wrote entry r)
        int rc = read_ring(g, checked, r + 1 - checked);
        if (rc) return rc;
        bool done = false;
        for (; checked <= r; checked++) {
            const RoundCtr &c = g->h_ring[checked % kRing];
            if (idle(c)) {
                done = true;
                break;
            }
            const HostSummary s = host_summary(c);
            fs += s.found;
            es += s.edges;
            rounds++;
        }
        if (done) break;
        if (batch < 16) batch *= 2;
    }
    *rounds_out = rounds;
    if (found_sum) *found_sum = fs;
    if (edges_sum) *edges_sum = es;
    return PTG_OK;
}

// ---------------------------------------------------------------------------
// BFS
// ---------------------------------------------------------------------------

// Top-down step. The visited bitmap picks the unique discoverer of a vertex
// (atomicOr), which writes level = current + 1: the fixed point of
// bfs_kernel (gpu_kernels.cu:118-134).
struct BfsTopDown {
    const uint32_t *__restrict__ off;
    const uint32_t *__restrict__ col;
    uint32_t *vis;
    uint32_t *fout;
    uint32_t *level;
    uint32_t next_level;

    __device__ __forceinline__ uint32_t vertex(uint32_t) const { return 0; }

    __device__ __forceinline__ void edge(bool valid, uint32_t, uint64_t e, const Emit &em, Acc &acc) const {
        const uint32_t w = valid ? col[e] : 0u;
        const uint32_t bit = 1u << (w & 31);
        bool won = false;
        if (valid && !(vis[w >> 5] & bit)) won = !(atomicOr(&vis[w >> 5], bit) & bit);
        uint32_t d = 0;
        if (won) {
            level[w] = next_level;
            d = off[w + 1] - off[w];
            if (d) atomicOr(&fout[w >> 5], bit);
        }
        emit(em, won, w, d, acc);
    }
};

// Bottom-up step (symmetric graphs only): every unvisited vertex looks for a
// neighbour in the current frontier bitmap and stops at the first one. A wave
// owns 64 consecutive vertices = two bitmap words, so its words need no
// atomics. Lists longer than small_degree are scanned by the whole wave.
__device__ __forceinline__ void bfs_bottom_up(const RoundEnv &e, uint32_t *vis, uint32_t *level,
                                              uint32_t next_level, const Emit &em, Acc &acc) {
    const uint32_t lane = lane_id();
    const uint32_t gw = blockIdx.x * kWavesPerBlock + (threadIdx.x >> 6);
    const uint32_t nw = gridDim.x * kWavesPerBlock;
    for (uint32_t p = gw; p < e.npairs; p += nw) {
        const uint32_t v = p * 64 + lane;
        const uint32_t word = p * 2 + (lane >> 5);
        const uint32_t vw = vis[word];
        const bool todo = v < e.V && !((vw >> (lane & 31)) & 1u);
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
                for (int t = 0; t < 4; t++) u[t] = (k + t < d) ? e.col[(uint64_t)b + k + t] : 0xFFFFFFFFu;
                bool h = false;
#pragma unroll
                for (int t = 0; t < 4; t++)
                    if (u[t] != 0xFFFFFFFFu) h = h || ((e.fin[u[t] >> 5] >> (u[t] & 31)) & 1u);
                if (h) {
                    hit = true;
                    break;
                }
            }
        }
        const uint64_t m = ballot(hit);
        if ((lane & 31) == 0) {
            const uint32_t bits = lane ? (uint32_t)(m >> 32) : (uint32_t)m;
            if (bits) {
                e.fout[word] = bits;
                vis[word] = vw | bits;
            }
        }
        if (hit) level[v] = next_level;
        emit(em, hit, v, d, acc);
    }
}

struct BfsParams {
    uint32_t *vis;
    uint32_t *level;
    uint64_t E;
    float alpha, beta;
    int symmetric;
};

// One BFS level. Direction (Beamer, Asanovic, Patterson 2012): switch to
// bottom-up when the frontier's edges exceed the unexplored edges / alpha,
// back to top-down when the frontier has fewer than V / beta vertices.
__global__ __launch_bounds__(kBlock) void k_bfs_round(RoundEnv e, BfsParams p) {
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
        else if (dir == kDirBottomUp && (double)s.found < (double)e.V / p.beta) dir = kDirTopDown;
    }
    const bool queue = dir == kDirTopDown && s.edges <= e.qcap;
    clear_bitmap_ahead(e);
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        out->mode = queue ? kModeQueue : kModeBitmap;
        out->dir = dir;
        out->seen = seen;
    }
    const Emit em = make_emit(e, out, queue);
    Acc acc;
    if (dir == kDirTopDown) {
        BfsTopDown op{e.off, e.col, p.vis, e.fout, p.level, e.r + 2};
        expand_frontier(e, in, op, em, acc);
    } else {
        bfs_bottom_up(e, p.vis, p.level, e.r + 2, em, acc);
    }
    finish_round(out, acc);
}

int run_bfs(ptg_graph *g, uint32_t src, uint32_t *level, ptg_stats *st) {
    int sym = 0;
    int rc = check_symmetric(g, &sym);
    if (rc) return rc;
    uint32_t *vis = g->d_bits[3];
    LaunchTimer timer(g);
    PTG_HIP(hipEventRecord(g->ev0, g->stream));
    if ((rc = fill_u32(g, level, g->V, 0xFFFFFFFFu)) || (rc = fill_u32(g, vis, g->nwords, 0u)) ||
        (rc = clear_frontier_bitmaps(g)) || (rc = clear_ring(g)))
        return rc;
    const uint32_t qcap = kQueueCap;
    k_seed<<<1, 64, 0, g->stream>>>(make_env(g, 0, qcap), src, 1u, level, vis, 0u);
    PTG_HIP(hipGetLastError());
    BfsParams p{vis, level, g->E, g->cfg.td_to_bu > 0 ? g->cfg.td_to_bu : 14.0f,
                g->cfg.bu_to_td > 0 ? g->cfg.bu_to_td : 24.0f, sym};
    uint32_t iters = 0;
    uint64_t reached = 0, edges = 0;
    rc = drive_rounds(
        g,
        [&](uint32_t r) -> int {
            const RoundEnv env = make_env(g, r, qcap);
            timer.before();
            k_chunkify<<<g->grid, kBlock, 0, g->stream>>>(env);
            k_bfs_round<<<g->grid, kBlock, 0, g->stream>>>(env, p);
            timer.after();
            PTG_HIP(hipGetLastError());
            return PTG_OK;
        },
        [](const RoundCtr &c) { return host_summary(c).found == 0; }, &iters, &reached, &edges);
    if (rc) return rc;
    PTG_HIP(hipEventRecord(g->ev1, g->stream));
    PTG_HIP(hipEventSynchronize(g->ev1));
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

// ---------------------------------------------------------------------------
// SSSP (near/far worklists) and label-propagation CC
// ---------------------------------------------------------------------------

// Relaxation of sssp_kernel (gpu_kernels.cu:152-172, new = val[u] + w) or of
// cc_kernel (gpu_kernels.cu:136-150, new = val[u]). An improved vertex joins
// the next frontier once (its bit in fout dedupes) when its value is below
// the threshold, otherwise the far set.
template <bool kWeighted>
struct Relax {
    const uint32_t *__restrict__ off;
    const uint32_t *__restrict__ col;
    const uint32_t *__restrict__ wt;
    uint32_t *val;
    uint32_t *fout;
    uint32_t *far;
    uint32_t thr;

    __device__ __forceinline__ uint32_t vertex(uint32_t v) const { return val[v]; }

    __device__ __forceinline__ void edge(bool valid, uint32_t sv, uint64_t e, const Emit &em, Acc &acc) const {
        const uint32_t w = valid ? col[e] : 0u;
        const uint32_t nv = kWeighted ? sv + (valid ? wt[e] : 0u) : sv;
        bool near = false;
        uint32_t d = 0;
        if (valid && nv < val[w]) {
            const uint32_t old = atomicMin(&val[w], nv);
            if (nv < old) {
                d = off[w + 1] - off[w];
                const uint32_t bit = 1u << (w & 31);
                if (d) {
                    if (nv < thr)
                        near = !(atomicOr(&fout[w >> 5], bit) & bit);
                    else if (!(atomicOr(&far[w >> 5], bit) & bit))
                        acc.far++;
                }
            }
        }
        emit(em, near, w, d, acc);
    }
};

// Near/far split (Davidson, Baxter, Garland, Owens 2014): raise the threshold
// by delta and move far vertices below it into the next frontier; vertices
// whose value already fell below the old threshold were processed when they
// got it and are dropped.
__device__ __forceinline__ void sssp_split(const RoundEnv &e, const uint32_t *val, uint32_t *far, uint32_t thr,
                                           uint32_t thr2, const Emit &em, Acc &acc) {
    const uint32_t lane = lane_id();
    const uint32_t gw = blockIdx.x * kWavesPerBlock + (threadIdx.x >> 6);
    const uint32_t nw = gridDim.x * kWavesPerBlock;
    for (uint32_t p = gw; p < e.npairs; p += nw) {
        const uint32_t word = p * 2 + (lane >> 5);
        const uint32_t bits = far[word];
        const bool inb = (bits >> (lane & 31)) & 1u;
        if (ballot(inb) == 0) continue;
        const uint32_t v = p * 64 + lane;
        const uint32_t dv = inb ? val[v] : 0u;
        const bool move = inb && dv >= thr && dv < thr2;
        const bool keep = inb && dv >= thr2;
        const uint32_t d = move ? e.off[v + 1] - e.off[v] : 0u;
        const uint64_t mm = ballot(move), mk = ballot(keep);
        if ((lane & 31) == 0) {
            const int sh = lane ? 32 : 0;
            far[word] = (uint32_t)(mk >> sh);
            const uint32_t mv = (uint32_t)(mm >> sh);
            if (mv) e.fout[word] = mv;
        }
        acc.far += keep;
        emit(em, move, v, d, acc);
    }
}

struct SsspParams {
    uint32_t *val;
    const uint32_t *wt;
    uint32_t *far;
    uint32_t delta;
};

template <bool kWeighted>
__global__ __launch_bounds__(kBlock) void k_sssp_round(RoundEnv e, SsspParams p) {
    RoundCtr *in = e.ring + e.r % kRing;
    RoundCtr *out = e.ring + (e.r + 1) % kRing;
    clear_ring_ahead(e);
    const InSummary s = summarize(in);
    const bool expand = s.found > 0;
    if (!expand && s.far == 0) return;
    const uint32_t thr = in->thr;
    const uint32_t thr2 = thr > 0xFFFFFFFFu - p.delta ? 0xFFFFFFFFu : thr + p.delta;
    // the next frontier has at most (frontier edges) or (far set) vertices
    const bool queue = (expand ? s.edges : (uint64_t)s.far) <= e.qcap;
    clear_bitmap_ahead(e);
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        out->mode = queue ? kModeQueue : kModeBitmap;
        out->thr = expand ? thr : thr2;
        out->far_base = expand ? s.far : 0u;
    }
    const Emit em = make_emit(e, out, queue);
    Acc acc;
    if (expand) {
        Relax<kWeighted> op{e.off, e.col, p.wt, p.val, e.fout, p.far, thr};
        expand_frontier(e, in, op, em, acc);
    } else {
        sssp_split(e, p.val, p.far, thr, thr2, em, acc);
    }
    finish_round(out, acc);
}

template <bool kWeighted>
static int run_relax_rounds(ptg_graph *g, SsspParams p, LaunchTimer &timer, uint32_t *iters, uint64_t *fsum) {
    const uint32_t qcap = kQueueCap;
    return drive_rounds(
        g,
        [&](uint32_t r) -> int {
            const RoundEnv env = make_env(g, r, qcap);
            timer.before();
            k_chunkify<<<g->grid, kBlock, 0, g->stream>>>(env);
            k_sssp_round<kWeighted><<<g->grid, kBlock, 0, g->stream>>>(env, p);
            timer.after();
            PTG_HIP(hipGetLastError());
            return PTG_OK;
        },
        [](const RoundCtr &c) {
            const HostSummary s = host_summary(c);
            return s.found == 0 && s.far == 0;
        },
        iters, fsum, nullptr);
}

int run_sssp(ptg_graph *g, uint32_t src, uint32_t *dist, ptg_stats *st) {
    int rc;
    LaunchTimer timer(g);
    uint32_t *far = g->d_bits[3];
    const uint32_t delta = g->cfg.sssp_delta ? g->cfg.sssp_delta : 32u;
    PTG_HIP(hipEventRecord(g->ev0, g->stream));
    if ((rc = fill_u32(g, dist, g->V, g->cfg.sssp_sentinel)) || (rc = fill_u32(g, far, g->nwords, 0u)) ||
        (rc = clear_frontier_bitmaps(g)) || (rc = clear_ring(g)))
        return rc;
    const uint32_t thr0 = 1u + delta;
    k_seed<<<1, 64, 0, g->stream>>>(make_env(g, 0, kQueueCap), src, 1u, dist, nullptr, thr0);
    PTG_HIP(hipGetLastError());
    uint32_t iters = 0;
    uint64_t fsum = 0;
    if ((rc = run_relax_rounds<true>(g, SsspParams{dist, g->d_w, far, delta}, timer, &iters, &fsum))) return rc;
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
// frontier (bitmap mode; the pre-pass queues its heavier vertices).
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
        c->found_sh[0] = n_nonzero;
        c->edges_sh[0] = E;
    }
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
    __device__ __forceinline__ uint32_t vertex(uint32_t v) const { return v; }
    __device__ __forceinline__ void edge(bool valid, uint32_t v, uint64_t e) const {
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

// Number of v with label[v] == v (added to 1 of 4 accumulators).
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
    if (lane == 0 && cnt) atomicAdd(&out[blockIdx.x % 4], (unsigned long long)cnt);
}

static int count_roots(ptg_graph *g, const uint32_t *label, uint32_t *roots) {
    unsigned long long *acc = g->d_acc;
    int rc = fill_u32(g, reinterpret_cast<uint32_t *>(acc), 8, 0u);
    if (rc) return rc;
    k_count_roots<<<g->grid, kBlock, 0, g->stream>>>(g->V, label, acc);
    PTG_HIP(hipGetLastError());
    unsigned long long h[4] = {0, 0, 0, 0};
    PTG_HIP(hipMemcpyAsync(h, acc, sizeof h, hipMemcpyDeviceToHost, g->stream));
    PTG_HIP(hipStreamSynchronize(g->stream));
    *roots = (uint32_t)(h[0] + h[1] + h[2] + h[3]);
    return PTG_OK;
}

int run_cc(ptg_graph *g, uint32_t *label, ptg_stats *st) {
    int sym = 0;
    int rc = check_symmetric(g, &sym);
    if (rc) return rc;
    uint32_t iters = 1;
    uint64_t fsum = g->V;
    LaunchTimer timer(g);
    PTG_HIP(hipEventRecord(g->ev0, g->stream));
    if (sym) {
        k_uf_init<<<g->grid, kBlock, 0, g->stream>>>(g->V, g->out.off, g->out.col, label);
        Frontier fin = g->out.bins.q.with(g->out.bins.d_ctr);
        timer.before();
        k_expand<Hook><<<g->grid, kBlock, 0, g->stream>>>(fin, g->out.off, Hook{g->out.col, label});
        timer.after();
        k_uf_flatten<<<g->grid, kBlock, 0, g->stream>>>(g->V, label);
        PTG_HIP(hipGetLastError());
    } else {
        // Label propagation from value[v] = v with every vertex active
        // (ccCaculateInShare, ptGraph/main.cu:977-980).
        k_iota_u32<<<g->grid, kBlock, 0, g->stream>>>(label, g->V);
        if ((rc = clear_frontier_bitmaps(g)) || (rc = clear_ring(g))) return rc;
        k_seed_all<<<g->grid, kBlock, 0, g->stream>>>(make_env(g, 0, kQueueCap), g->V - g->out.bins.n_zero, g->E);
        PTG_HIP(hipGetLastError());
        uint64_t s = 0;
        if ((rc = run_relax_rounds<false>(g, SsspParams{label, nullptr, g->d_bits[3], 0u}, timer, &iters, &s)))
            return rc;
        fsum = s;
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
    __device__ __forceinline__ uint32_t vertex(uint32_t v) const { return off[v + 1]; }
    __device__ __forceinline__ void edge(bool valid, uint32_t end, uint64_t e) const {
        if (valid && e + 1 < end && col[e] > col[e + 1]) *flag = 0u;
    }
};

struct CheckReverse {
    const uint32_t *__restrict__ off;
    const uint32_t *__restrict__ col;
    uint32_t *flag;
    __device__ __forceinline__ uint32_t vertex(uint32_t v) const { return v; }
    __device__ __forceinline__ void edge(bool valid, uint32_t v, uint64_t e) const {
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
    uint32_t *flag = g->d_bits[4];
    int rc = fill_u32(g, flag, 2, 1u);
    if (rc) return rc;
    Frontier fin = g->out.bins.q.with(g->out.bins.d_ctr);
    k_expand<CheckSorted><<<g->grid, kBlock, 0, g->stream>>>(fin, g->out.off, CheckSorted{g->out.off, g->out.col, flag});
    k_expand<CheckReverse><<<g->grid, kBlock, 0, g->stream>>>(fin, g->out.off,
                                                              CheckReverse{g->out.off, g->out.col, flag + 1});
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
    if (lane == 0 && cnt) {
        const uint32_t s = 2 * (blockIdx.x % 2);
        atomicAdd(&out[s], (unsigned long long)cnt);
        atomicAdd(&out[s + 1], (unsigned long long)edges);
    }
}

int reach_stats(ptg_graph *g, const uint32_t *val, uint32_t unreached, uint32_t *reached, uint64_t *edges) {
    if (!reached && !edges) return PTG_OK;
    unsigned long long *acc = g->d_acc;
    int rc = fill_u32(g, reinterpret_cast<uint32_t *>(acc), 8, 0u);
    if (rc) return rc;
    k_reach_stats<<<g->grid, kBlock, 0, g->stream>>>(g->V, val, unreached, g->out.off, acc);
    PTG_HIP(hipGetLastError());
    unsigned long long h[4];
    PTG_HIP(hipMemcpyAsync(h, acc, sizeof h, hipMemcpyDeviceToHost, g->stream));
    PTG_HIP(hipStreamSynchronize(g->stream));
    if (reached) *reached = (uint32_t)(h[0] + h[2]);
    if (edges) *edges = h[1] + h[3];
    return PTG_OK;
}

}  // namespace ptg
