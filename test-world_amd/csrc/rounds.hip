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
    return e;
}

// Pre-pass of round e.r for flags and bitmap frontiers. Workgroup b handles
// tiles [b*T, (b+1)*T): for a flags frontier it packs (and clears) the bytes
// into bitmap words and counts the frontier; for both it gathers members of
// degree > sd in LDS and appends them to chunk-queue shard b % 16 with one
// atomic per batch (an LDS scan gives the chunk bases). Members that the
// round defers (degree > defer_deg, value >= the frontier minimum + window)
// go to the deferred bitmap instead. A BFS round that stays bottom-up reads
// the bitmap alone, so its pre-pass has nothing to do.
__global__ __launch_bounds__(kBlock) void k_prepare(RoundEnv e) {
    constexpr uint32_t kBuf = 1024;
    __shared__ uint32_t s_v[kBuf], s_n[kBuf];
    __shared__ uint32_t s_cnt, s_wsum[kWavesPerBlock];
    __shared__ unsigned long long s_old;
    RoundCtr *in = e.ring + e.r % kRing;
    const uint32_t mode = in->mode;
    if (mode == kModeQueue) return;
    if (e.bu_beta > 0.0f && mode == kModeBitmap && in->dir == kDirBottomUp &&
        !bfs_leaves_bottom_up(summarize(in).found, e.V, e.bu_beta))
        return;
    const bool from_flags = mode == kModeFlags;
    const uint32_t lane = lane_id(), wib = threadIdx.x >> 6;
    const uint32_t T = (e.npairs + gridDim.x - 1) / gridDim.x;
    const uint32_t t_begin = blockIdx.x * T;
    const uint32_t t_end = min(e.npairs, t_begin + T);
    const uint32_t shard = blockIdx.x % kShards;
    const uint32_t lim = e.defer_deg != 0xFFFFFFFFu ? defer_limit(in, e.window) : 0xFFFFFFFFu;
    uint32_t *ids = e.cqin + (uint64_t)shard * e.cq_cap;
    uint32_t *bases = e.cbin + (uint64_t)shard * e.cq_cap;
    Acc acc;
    if (threadIdx.x == 0) s_cnt = 0;
    __syncthreads();
    for (uint32_t t0 = t_begin; t0 < t_end; t0 += kWavesPerBlock) {
        const uint32_t p = t0 + wib;
        const uint32_t v = p * 64 + lane;
        bool mem = false;
        uint32_t d = 0;
        if (p < t_end) {
            if (from_flags) {
                mem = e.flin[v] != 0;  // the flag array covers npairs * 64 bytes
                if (mem) e.flin[v] = 0;
                const uint64_t m = ballot(mem);
                if ((lane & 31) == 0) e.fin[2 * p + (lane >> 5)] = lane ? (uint32_t)(m >> 32) : (uint32_t)m;
            } else {
                mem = (e.fin[2 * p + (lane >> 5)] >> (lane & 31)) & 1u;
            }
            if (mem) d = e.off[v + 1] - e.off[v];
        }
        if (from_flags && mem) {
            acc.found++;
            acc.edges += d;
        }
        bool want = mem && d > e.sd;
        if (lim != 0xFFFFFFFFu) {
            const bool defer = want && d > e.defer_deg && e.val[v] >= lim;
            want = want && !defer;
            const uint64_t md = ballot(defer);
            if (md) {
                if ((lane & 31) == 0) {
                    const uint32_t bits = lane ? (uint32_t)(md >> 32) : (uint32_t)md;
                    if (bits) e.dbits[2 * p + (lane >> 5)] = bits;
                }
                if (lane == 0) in->ndefer = 1u;
            }
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
            // exclusive scan of chunk counts, 4 entries per thread
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
    if (from_flags) add_tallies(in, acc);
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

int reset_rounds(ptg_graph *g) {
    int rc;
    for (int i = 0; i < 3; i++)
        if ((rc = fill_u32(g, g->d_bits[i], g->nwords, 0u))) return rc;
    for (int i = 0; i < 2; i++)
        if ((rc = fill_u32(g, reinterpret_cast<uint32_t *>(g->d_flags[i]), (uint64_t)g->npairs * 16, 0u))) return rc;
    return fill_u32(g, reinterpret_cast<uint32_t *>(g->d_ring), sizeof(RoundCtr) / 4 * kRing, 0u);
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
    k_prepare<<<g->grid, kBlock, 0, g->stream>>>(e);
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
