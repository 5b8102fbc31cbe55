// Connected components for gfx950.
//
// The reference's cc_kernel (ptGraph/gpu_kernels.cu:136-150) propagates
// value[v] = v along out-edges to a fixed point: label[v] = the smallest id u
// with a path u -> v. On a symmetric graph that is the smallest id of v's
// component, which union-find with min-id roots computes; on other graphs the
// frontier label propagation (sssp.hip) runs.
//
// The union-find follows Afforest (Sutton, Ben-Nun, Barak, IPDPS 2018): every
// vertex is first linked to its first kAfforestK neighbours; the component
// most of kAfforestSamples sampled vertices fall in (the giant component of a
// power-law graph) is found; then only vertices outside it link their
// remaining edges. An edge skipped from both ends joins two vertices that
// were both in that component when they were checked, and components only
// grow, so no link is lost.
// Ids are the caller's: with relabeling the union-find keys are the original
// ids, so every root is its component's minimum original id.
#include <string.h>

#include "rounds.h"
#include "static_expand.h"
#include "union_find.h"

namespace ptg {

int run_label_propagation(ptg_graph *g, uint32_t *label, LaunchTimer &timer, uint32_t *iters, uint64_t *fsum);

constexpr uint32_t kAfforestK = 2;
constexpr uint32_t kAfforestSamples = 1024;

// Links every vertex to its first kAfforestK neighbours.
__global__ __launch_bounds__(kBlock) void k_uf_link_first(uint32_t V, const uint32_t *__restrict__ off,
                                                          const uint32_t *__restrict__ col,
                                                          const uint32_t *__restrict__ key, uint32_t *par) {
    const uint32_t stride = gridDim.x * blockDim.x;
    for (uint32_t v = blockIdx.x * blockDim.x + threadIdx.x; v < V; v += stride) {
        const uint32_t b = off[v], d = min(off[v + 1] - b, kAfforestK);
        for (uint32_t i = 0; i < d; i++) uf_unite(par, key, v, col[b + i]);
    }
}

// One workgroup of kAfforestSamples threads: the root shared by the most
// sampled vertices (ties to the smaller root) goes to *giant. The samples are
// a fixed hash of the thread index, so the choice is repeatable.
__global__ __launch_bounds__(kAfforestSamples) void k_uf_sample(uint32_t V, uint32_t *par, uint32_t *giant) {
    constexpr uint32_t kSlots = 2 * kAfforestSamples;
    __shared__ uint32_t s_key[kSlots], s_cnt[kSlots];
    __shared__ unsigned long long s_best[kAfforestSamples / 64];
    for (uint32_t i = threadIdx.x; i < kSlots; i += blockDim.x) {
        s_key[i] = 0xFFFFFFFFu;
        s_cnt[i] = 0;
    }
    __syncthreads();
    uint32_t h = threadIdx.x * 0x9E3779B9u;
    h ^= h >> 16;
    h *= 0x85EBCA6Bu;
    h ^= h >> 13;
    const uint32_t v = (uint32_t)(((uint64_t)h * V) >> 32);
    const uint32_t r = uf_root(par, v);
    uint32_t slot = (r * 0x9E3779B1u) >> 21;  // 11 bits = kSlots
    for (;;) {
        const uint32_t old = atomicCAS(&s_key[slot], 0xFFFFFFFFu, r);
        if (old == 0xFFFFFFFFu || old == r) {
            atomicAdd(&s_cnt[slot], 1u);
            break;
        }
        slot = (slot + 1) & (kSlots - 1);
    }
    __syncthreads();
    // largest count, then smallest root: maximise (count << 32) | ~root
    unsigned long long best = 0;
    for (uint32_t i = threadIdx.x; i < kSlots; i += blockDim.x)
        if (s_cnt[i]) best = max(best, ((unsigned long long)s_cnt[i] << 32) | (uint32_t)~s_key[i]);
    best = wave_max_u64(best);
    if (lane_id() == 0) s_best[threadIdx.x >> 6] = best;
    __syncthreads();
    if (threadIdx.x == 0) {
        for (uint32_t w = 1; w < kAfforestSamples / 64; w++) best = max(best, s_best[w]);
        *giant = ~(uint32_t)best;
    }
}

// The remaining edges (from the (kAfforestK+1)-th on) of vertices outside the
// giant component; vertex() decides once per vertex (or hub chunk).
struct HookRest {
    const uint32_t *__restrict__ off;
    const uint32_t *__restrict__ col;
    const uint32_t *__restrict__ key;
    uint32_t *par;
    const uint32_t *giant;
    struct Src {
        uint32_t v, first;  // first = ~0u: nothing to link
    };
    __device__ __forceinline__ Src vertex(uint32_t v) const {
        const uint32_t b = off[v];
        if ((uint64_t)b + kAfforestK >= off[v + 1] || uf_root(par, v) == uf_root(par, *giant))
            return Src{v, 0xFFFFFFFFu};
        return Src{v, b + kAfforestK};
    }
    __device__ __forceinline__ bool active(Src s) const { return s.first != 0xFFFFFFFFu; }
    __device__ __forceinline__ void edge(bool valid, Src s, uint64_t e) const {
        if (valid && e >= s.first) uf_unite(par, key, s.v, col[e]);
    }
};

// par[v] = v, or its first neighbour when that one has a smaller key (a
// valid first hook).
__global__ __launch_bounds__(kBlock) void k_uf_init(uint32_t V, const uint32_t *__restrict__ off,
                                                    const uint32_t *__restrict__ col,
                                                    const uint32_t *__restrict__ key, uint32_t *__restrict__ par) {
    const uint32_t stride = gridDim.x * blockDim.x;
    for (uint32_t v = blockIdx.x * blockDim.x + threadIdx.x; v < V; v += stride) {
        const uint32_t b = off[v];
        uint32_t p = v;
        if (off[v + 1] > b && uf_key(key, col[b]) < uf_key(key, v)) p = col[b];
        par[v] = p;
    }
}

__global__ __launch_bounds__(kBlock) void k_uf_flatten(uint32_t V, uint32_t *par) {
    const uint32_t stride = gridDim.x * blockDim.x;
    for (uint32_t v = blockIdx.x * blockDim.x + threadIdx.x; v < V; v += stride) par[v] = uf_root(par, v);
}

// label[v] = key of its root (after flattening, par[v] is the root).
__global__ __launch_bounds__(kBlock) void k_map_keys(uint32_t V, const uint32_t *__restrict__ key,
                                                     uint32_t *__restrict__ par) {
    const uint32_t stride = gridDim.x * blockDim.x;
    for (uint32_t v = blockIdx.x * blockDim.x + threadIdx.x; v < V; v += stride) par[v] = key[par[v]];
}

// value[v] = key of v (label propagation starts from the caller's ids).
__global__ __launch_bounds__(kBlock) void k_init_keys(uint32_t V, const uint32_t *__restrict__ key,
                                                      uint32_t *__restrict__ val) {
    const uint32_t stride = gridDim.x * blockDim.x;
    for (uint32_t v = blockIdx.x * blockDim.x + threadIdx.x; v < V; v += stride) val[v] = key ? key[v] : v;
}

// Number of v whose label is its own key, one per component (summed per
// workgroup, then added to 1 of 16 shards).
__global__ __launch_bounds__(kBlock) void k_count_roots(uint32_t V, const uint32_t *__restrict__ label,
                                                        const uint32_t *__restrict__ key, unsigned long long *out) {
    __shared__ unsigned long long s_c[kWavesPerBlock];
    const uint32_t lane = lane_id(), wib = threadIdx.x >> 6;
    const uint32_t gw = blockIdx.x * kWavesPerBlock + wib;
    const uint32_t nw = gridDim.x * kWavesPerBlock;
    uint64_t cnt = 0;
    for (uint64_t base = (uint64_t)gw * 64; base < V; base += (uint64_t)nw * 64) {
        const uint64_t v = base + lane;
        cnt += (v < V && label[v] == (key ? key[v] : (uint32_t)v));
    }
    cnt = wave_sum_u64(cnt);
    if (lane == 0) s_c[wib] = cnt;
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned long long c = 0;
        for (int w = 0; w < kWavesPerBlock; w++) c += s_c[w];
        if (c) atomicAdd(&out[blockIdx.x % kShards], c);
    }
}

int count_roots(ptg_graph *g, const uint32_t *label, uint32_t *roots) {
    unsigned long long *acc = g->d_acc;
    int rc = fill_u32(g, reinterpret_cast<uint32_t *>(acc), 2 * kShards, 0u);
    if (rc) return rc;
    // a streaming scan: 512 workgroups, as for the reach statistics
    k_count_roots<<<std::min(g->grid, 512u), kBlock, 0, g->stream>>>(g->V, label, g->d_inv, acc);
    PTG_HIP(hipGetLastError());
    unsigned long long h[kShards];
    PTG_HIP(hipMemcpyAsync(h, acc, sizeof h, hipMemcpyDeviceToHost, g->stream));
    PTG_HIP(hipStreamSynchronize(g->stream));
    uint64_t c = 0;
    for (int s = 0; s < kShards; s++) c += h[s];
    *roots = (uint32_t)c;
    return PTG_OK;
}

// par[v] = v (a union-find with every vertex its own root).
__global__ __launch_bounds__(kBlock) void k_uf_iota(uint32_t V, uint32_t *__restrict__ par) {
    const uint32_t stride = gridDim.x * blockDim.x;
    for (uint32_t v = blockIdx.x * blockDim.x + threadIdx.x; v < V; v += stride) par[v] = v;
}

int uf_start(ptg_graph *g, uint32_t *par) {
    k_uf_iota<<<g->grid, kBlock, 0, g->stream>>>(g->V, par);
    PTG_HIP(hipGetLastError());
    return PTG_OK;
}

// Flattens the union-find in `par` and maps every root to its key (the
// caller's id when the graph is relabeled): par becomes the CC labels.
int uf_finish(ptg_graph *g, uint32_t *par) {
    k_uf_flatten<<<g->grid, kBlock, 0, g->stream>>>(g->V, par);
    if (g->d_inv) k_map_keys<<<g->grid, kBlock, 0, g->stream>>>(g->V, g->d_inv, par);
    PTG_HIP(hipGetLastError());
    return PTG_OK;
}

int init_keys(ptg_graph *g, uint32_t *val) {
    k_init_keys<<<g->grid, kBlock, 0, g->stream>>>(g->V, g->d_inv, val);
    PTG_HIP(hipGetLastError());
    return PTG_OK;
}

int run_cc(ptg_graph *g, uint32_t *label, ptg_stats *st) {
    int sym = 0;
    int rc = check_symmetric(g, &sym);
    if (rc) return rc;
    uint32_t iters = 1;
    uint64_t fsum = g->V;
    LaunchTimer timer(g);
    const uint32_t *key = g->d_inv;  // null when the device keeps the caller's ids
    PTG_HIP(hipEventRecord(g->ev0, g->stream));
    if (sym) {
        // every union-find launch is timed: together they are the computation
        uint32_t *giant = reinterpret_cast<uint32_t *>(g->d_acc);  // count_roots clears it afterwards
        const uint32_t *off = g->out.off, *col = g->out.col;
        timer.launch(k_uf_init, dim3(g->grid), dim3(kBlock), g->V, off, col, key, label);
        timer.launch(k_uf_link_first, dim3(g->grid), dim3(kBlock), g->V, off, col, key, label);
        if (g->V) timer.launch(k_uf_sample, dim3(1), dim3(kAfforestSamples), g->V, label, giant);
        const Frontier fin = g->out.bins.q.with(g->out.bins.d_ctr);
        timer.launch(k_expand<HookRest>, dim3(g->grid), dim3(kBlock), fin, off, HookRest{off, col, key, label, giant});
        timer.launch(k_uf_flatten, dim3(g->grid), dim3(kBlock), g->V, label);
        if (key) timer.launch(k_map_keys, dim3(g->grid), dim3(kBlock), g->V, key, label);
        PTG_HIP(hipGetLastError());
    } else {
        k_init_keys<<<g->grid, kBlock, 0, g->stream>>>(g->V, key, label);
        PTG_HIP(hipGetLastError());
        if ((rc = run_label_propagation(g, label, timer, &iters, &fsum))) return rc;
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

}  // namespace ptg
