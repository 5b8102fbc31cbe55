// Round machinery: the per-round pre-pass, seeding, and ring readback.
#include <string.h>

#include <algorithm>

#include "rounds.h"

namespace ptg {

RoundEnv make_env(ptg_graph *g, uint32_t r) { return make_env_csr(g, r, g->out.off, g->out.col); }

RoundEnv make_env_csr(ptg_graph *g, uint32_t r, const uint32_t *off, const uint32_t *col) {
    RoundEnv e;
    e.ring = g->d_ring;
    e.r = r;
    e.off = off;
    e.col = col;
    e.fin = g->d_bits[r % 3];
    e.fout = g->d_bits[(r + 1) % 3];
    e.fclr = g->d_bits[(r + 2) % 3];
    e.flin = g->d_flags[r & 1];
    e.flout = g->d_flags[(r + 1) & 1];
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
    e.qcap = kQueueCap;
    e.val = nullptr;
    e.dbits = g->d_defer;
    e.defer_deg = 0xFFFFFFFFu;
    e.window = 0;
    e.bu_beta = 0.0f;
    e.pull = 0;
    e.pull_wmin = 0;
    e.pull_gate = 0;
    return e;
}

// LDS staging of the pre-pass: chunk-queue entries of the workgroup,
// flushed to its shard with one atomic per batch.
struct PrepLds {
    static constexpr uint32_t kU = 4;                       // tiles per wave per step
    static constexpr uint32_t kBuf = 2048;                  // entries; one step adds <= kBlock * kU
    static constexpr uint32_t kFlushAt = kBuf - kBlock * kU;
    static constexpr uint32_t kPer = kBuf / kBlock;         // entries per thread in the flush scan
    uint32_t v[kBuf], n[kBuf];
    uint32_t cnt, wsum[kWavesPerBlock];
    unsigned long long old;
    unsigned long long usum[kWavesPerBlock], dsum[kWavesPerBlock];
};

// Whether the pre-pass of round e.r counts the vertices that can still
// improve for an SSSP pull decision.
__device__ __forceinline__ bool counts_open(const RoundEnv &e, const RoundCtr *in) {
    if (!e.pull || in->fmin_inv == 0 || e.r == 0) return false;
    const RoundCtr *prev = e.ring + (e.r - 1) % kRing;
    uint64_t pe = 0;
#pragma unroll
    for (int i = 0; i < kShards; i++) pe += prev->edges_sh[i];
    return pe >= e.pull_gate;
}

// Tiles [t_begin, t_end) of the workgroup, kU per wave per step. Every load
// of a step (the frontier word or flag byte, offsets v and v + 1) is issued
// before any is used and none is skipped by a lane (indexes are clamped into
// range), so one memory latency covers kWavesPerBlock * kU tiles.
template <bool kFlags>
__device__ __forceinline__ void prepare_tiles(const RoundEnv &e, RoundCtr *in, uint32_t lim, PrepLds &L) {
    constexpr uint32_t kU = PrepLds::kU;
    constexpr uint32_t kStep = kWavesPerBlock * kU;
    const uint32_t lane = lane_id(), wib = threadIdx.x >> 6;
    const uint32_t T = (e.npairs + gridDim.x - 1) / gridDim.x;
    const uint32_t t_begin = blockIdx.x * T;
    const uint32_t t_end = min(e.npairs, t_begin + T);
    const uint32_t shard = blockIdx.x % kShards;
    uint32_t *ids = e.cqin + (uint64_t)shard * e.cq_cap;
    uint32_t *bases = e.cbin + (uint64_t)shard * e.cq_cap;
    Acc acc;
    // SSSP pull rounds: in-edges (= out-degrees, the graph is symmetric) of
    // the vertices whose value is above the bound
    const bool count_u = kFlags && counts_open(e, in);
    const uint32_t bound = count_u ? pull_bound(in, e.pull_wmin) : 0u;
    uint64_t uacc = 0, dacc = 0;
    if (threadIdx.x == 0) L.cnt = 0;
    __syncthreads();
    for (uint32_t t0 = t_begin; t0 < t_end; t0 += kStep) {
        uint32_t x[kU], o0[kU], o1[kU], vv[kU];
#pragma unroll
        for (uint32_t u = 0; u < kU; u++) {
            const uint32_t p = min(t0 + u * kWavesPerBlock + wib, t_end - 1);
            const uint32_t v = p * 64 + lane;
            if constexpr (kFlags)
                x[u] = e.flin[v];  // the flag array covers npairs * 64 bytes
            else
                x[u] = e.fin[2 * p + (lane >> 5)];
            o0[u] = e.off[min(v, e.V)];
            o1[u] = e.off[min(v + 1, e.V)];
            vv[u] = count_u ? e.val[min(v, e.V - 1)] : 0u;  // past V the degree is 0
        }
        uint32_t d[kU];
        bool want[kU];
#pragma unroll
        for (uint32_t u = 0; u < kU; u++) {
            const uint32_t p = t0 + u * kWavesPerBlock + wib;
            const bool ok = p < t_end;
            bool mem;
            if constexpr (kFlags) {
                mem = ok && x[u] != 0;
                if (mem) e.flin[p * 64 + lane] = 0;
                const uint64_t m = ballot(mem);
                if (ok && (lane & 31) == 0) e.fin[2 * p + (lane >> 5)] = lane ? (uint32_t)(m >> 32) : (uint32_t)m;
            } else {
                mem = ok && ((x[u] >> (lane & 31)) & 1u);
            }
            d[u] = mem ? o1[u] - o0[u] : 0u;
            if (count_u && ok && vv[u] > bound) uacc += o1[u] - o0[u];
            if (kFlags && mem) {
                acc.found++;
                acc.edges += d[u];
            }
            want[u] = mem && d[u] > e.sd;
        }
        if (lim != 0xFFFFFFFFu) {
            // SSSP deferral: heavy members at or above the limit wait in dbits
            uint32_t val[kU];
#pragma unroll
            for (uint32_t u = 0; u < kU; u++)
                val[u] = want[u] && d[u] > e.defer_deg
                             ? (count_u ? vv[u] : e.val[(t0 + u * kWavesPerBlock + wib) * 64 + lane])
                             : 0u;
#pragma unroll
            for (uint32_t u = 0; u < kU; u++) {
                const uint32_t p = t0 + u * kWavesPerBlock + wib;
                const bool defer = want[u] && d[u] > e.defer_deg && val[u] >= lim;
                want[u] = want[u] && !defer;
                if (defer) dacc += d[u];
                const uint64_t md = ballot(defer);
                if (md) {
                    if ((lane & 31) == 0) {
                        const uint32_t bits = lane ? (uint32_t)(md >> 32) : (uint32_t)md;
                        if (bits) e.dbits[2 * p + (lane >> 5)] = bits;
                    }
                    if (lane == 0) in->ndefer = 1u;
                }
            }
        }
#pragma unroll
        for (uint32_t u = 0; u < kU; u++) {
            const uint64_t m = ballot(want[u]);
            if (m) {
                uint32_t base = 0;
                if (lane == lowest_lane(m)) base = atomicAdd(&L.cnt, (uint32_t)__popcll(m));
                base = read_lane(base, lowest_lane(m));
                if (want[u]) {
                    L.v[base + rank_below(m)] = (t0 + u * kWavesPerBlock + wib) * 64 + lane;
                    L.n[base + rank_below(m)] = wave_chunks(d[u]);
                }
            }
        }
        __syncthreads();
        const uint32_t cnt = L.cnt;
        if (cnt > PrepLds::kFlushAt || t0 + kStep >= t_end) {
            // exclusive scan of chunk counts, kPer entries per thread
            uint32_t loc[PrepLds::kPer], sum = 0;
#pragma unroll
            for (uint32_t j = 0; j < PrepLds::kPer; j++) {
                const uint32_t i = threadIdx.x * PrepLds::kPer + j;
                loc[j] = i < cnt ? L.n[i] : 0u;
                sum += loc[j];
            }
            const uint32_t incl = wave_incl_scan_u32(sum);
            if (lane == 63) L.wsum[wib] = incl;
            __syncthreads();
            uint32_t woff = 0, total = 0;
            for (int w = 0; w < kWavesPerBlock; w++) {
                if (w < (int)wib) woff += L.wsum[w];
                total += L.wsum[w];
            }
            if (threadIdx.x == 0 && cnt)
                L.old = atomicAdd(&in->cpack[shard], ((unsigned long long)cnt << kLCountShift) | total);
            __syncthreads();
            const uint32_t slot0 = (uint32_t)(L.old >> kLCountShift);
            uint32_t run = (uint32_t)(L.old & kLChunkMask) + woff + (incl - sum);
#pragma unroll
            for (uint32_t j = 0; j < PrepLds::kPer; j++) {
                const uint32_t i = threadIdx.x * PrepLds::kPer + j;
                if (i < cnt) {
                    ids[slot0 + i] = L.v[i];
                    bases[slot0 + i] = run;
                }
                run += loc[j];
            }
            __syncthreads();
            if (threadIdx.x == 0) L.cnt = 0;
            __syncthreads();
        }
    }
    if (kFlags) add_tallies(in, acc);
    if (count_u) {
        const uint64_t us = wave_sum_u64(uacc), ds = wave_sum_u64(dacc);
        if (lane == 0) {
            L.usum[wib] = us;
            L.dsum[wib] = ds;
        }
        __syncthreads();
        if (threadIdx.x == 0) {
            unsigned long long tu = 0, td = 0;
            for (int w = 0; w < kWavesPerBlock; w++) {
                tu += L.usum[w];
                td += L.dsum[w];
            }
            if (tu) atomicAdd(&in->uedges_sh[shard], tu);
            if (td) atomicAdd(&in->dedges_sh[shard], td);
            if (blockIdx.x == 0) in->ucount = 1u;
        }
    }
}

// Pre-pass of round e.r for flags and bitmap frontiers. Workgroup b handles
// tiles [b*T, (b+1)*T): for a flags frontier it packs (and clears) the bytes
// into bitmap words and counts the frontier; for both it gathers members of
// degree > sd in LDS and appends them to chunk-queue shard b % 16 with one
// atomic per batch (an LDS scan gives the chunk bases). Members that the
// round defers (degree > defer_deg, value >= the frontier minimum + window)
// go to the deferred bitmap instead. For SSSP pull decisions a flags
// pre-pass also sums the in-degrees of the vertices whose value can still
// fall (above the frontier minimum + the smallest weight). A BFS round that
// stays bottom-up reads the bitmap alone, so its pre-pass has nothing to do.
__global__ __launch_bounds__(kBlock) void k_prepare(RoundEnv e) {
    __shared__ PrepLds L;
    RoundCtr *in = e.ring + e.r % kRing;
    const uint32_t mode = in->mode;
    if (mode == kModeQueue) return;
    if (e.bu_beta > 0.0f && mode == kModeBitmap && in->dir == kDirBottomUp &&
        !bfs_leaves_bottom_up(summarize(in).found, e.V, e.bu_beta))
        return;
    const uint32_t lim = e.defer_deg != 0xFFFFFFFFu ? defer_limit(in, e.window) : 0xFFFFFFFFu;
    if (mode == kModeFlags)
        prepare_tiles<true>(e, in, lim, L);
    else
        prepare_tiles<false>(e, in, lim, L);
}

// Seeds round 0 with a single source (value src_value) in queue mode.
__global__ void k_seed(RoundEnv e, uint32_t src, uint32_t src_value, uint32_t *val, uint32_t thr) {
    if (threadIdx.x != 0) return;
    RoundCtr *c = e.ring;
    const uint32_t d = e.off[src + 1] - e.off[src];
    val[src] = src_value;
    if (d > 0) e.fin[src >> 5] |= 1u << (src & 31);
    c->mode = kModeQueue;
    c->dir = kDirTopDown;
    c->thr = thr;
    c->lo = src_value;
    c->fmin_inv = ~src_value;
    c->any = 1u;
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

int reset_rounds(ptg_graph *g, std::initializer_list<FillSpan> extra) {
    FillSpan s[kMaxSpans];
    uint32_t k = 0;
    for (int i = 0; i < 3; i++) s[k++] = FillSpan{g->d_bits[i], g->nwords, 0u};
    for (int i = 0; i < 2; i++)
        s[k++] = FillSpan{reinterpret_cast<uint32_t *>(g->d_flags[i]), (uint64_t)g->npairs * 16, 0u};
    s[k++] = FillSpan{reinterpret_cast<uint32_t *>(g->d_ring), sizeof(RoundCtr) / 4 * kRing, 0u};
    if (k + extra.size() > kMaxSpans) return fail(PTG_ERR_INVALID_ARG, "reset_rounds: too many spans");
    for (const FillSpan &x : extra) s[k++] = x;
    return fill_spans(g, s, k);
}

int launch_seed_csr(ptg_graph *g, const uint32_t *off, const uint32_t *col, uint32_t src, uint32_t src_value,
                    uint32_t *val, uint32_t thr) {
    k_seed<<<1, 64, 0, g->stream>>>(make_env_csr(g, 0, off, col), src, src_value, val, thr);
    PTG_HIP(hipGetLastError());
    return PTG_OK;
}

int launch_seed(ptg_graph *g, uint32_t src, uint32_t src_value, uint32_t *val, uint32_t thr) {
    k_seed<<<1, 64, 0, g->stream>>>(make_env(g, 0), src, src_value, val, thr);
    PTG_HIP(hipGetLastError());
    return PTG_OK;
}

int launch_prepare(ptg_graph *g, const RoundEnv &e) {
    k_prepare<<<prepare_grid(g->grid), kBlock, 0, g->stream>>>(e);
    PTG_HIP(hipGetLastError());
    return PTG_OK;
}

int read_ring(ptg_graph *g, uint32_t from, uint32_t count) {
    // Copies ring entries [from, from + count) (mod kRing) into h_ring.
    const uint32_t a = from % kRing;
    const uint32_t first = std::min(count, kRing - a);
    PTG_HIP(hipMemcpyAsync(g->h_ring + a, g->d_ring + a, sizeof(RoundCtr) * first, hipMemcpyDeviceToHost,
                           g->stream));
    if (count > first)
        PTG_HIP(hipMemcpyAsync(g->h_ring, g->d_ring, sizeof(RoundCtr) * (count - first), hipMemcpyDeviceToHost,
                               g->stream));
    PTG_HIP(hipStreamSynchronize(g->stream));
    return PTG_OK;
}

}  // namespace ptg
