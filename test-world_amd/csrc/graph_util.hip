// Small device utilities: fills, the result un-permutation, the one-time
// symmetry fingerprint and reached-set statistics.
#include <string.h>

#include <algorithm>
#include <random>
#include <vector>

#include "fingerprint.h"
#include "static_expand.h"

namespace ptg {

// Fills n words; `vec` (16-byte aligned p only) writes uint4 per lane.
__global__ __launch_bounds__(kBlock) void k_fill_u32(uint32_t *__restrict__ p, uint64_t n, uint32_t v, bool vec) {
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const uint64_t n4 = vec ? n / 4 : 0;
    uint4 *p4 = reinterpret_cast<uint4 *>(p);
    const uint4 vv = make_uint4(v, v, v, v);
    for (uint64_t j = i; j < n4; j += stride) p4[j] = vv;
    for (uint64_t j = n4 * 4 + i; j < n; j += stride) p[j] = v;
}

struct FillSet {
    FillSpan s[kMaxSpans];
    uint32_t k;
};

// Fills every span of the set: uint4 stores over each span's 16-byte-aligned
// body, words for its head and tail.
__global__ __launch_bounds__(kBlock) void k_fill_spans(FillSet f) {
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    for (uint32_t s = 0; s < f.k; s++) {
        uint32_t *p = f.s[s].p;
        const uint64_t n = f.s[s].n;
        const uint32_t v = f.s[s].v;
        const uint64_t head = min(n, (uint64_t)((16u - ((uint32_t)(uintptr_t)p & 15u)) & 15u) / 4);
        const uint64_t n4 = (n - head) / 4;
        uint4 *p4 = reinterpret_cast<uint4 *>(p + head);
        const uint4 vv = make_uint4(v, v, v, v);
        if (i < head) p[i] = v;
        for (uint64_t j = i; j < n4; j += stride) p4[j] = vv;
        for (uint64_t j = head + n4 * 4 + i; j < n; j += stride) p[j] = v;
    }
}

int fill_spans(ptg_graph *g, const FillSpan *spans, uint32_t k) {
    if (k > kMaxSpans) return fail(PTG_ERR_INVALID_ARG, "fill_spans: too many spans");
    FillSet f{};
    uint64_t total = 0;
    for (uint32_t i = 0; i < k; i++)
        if (spans[i].n) {
            f.s[f.k++] = spans[i];
            total += spans[i].n;
        }
    if (!f.k) return PTG_OK;
    const uint32_t blocks = (uint32_t)std::min<uint64_t>((total / 4 + kBlock - 1) / kBlock + 1, std::min(g->grid, 512u));
    k_fill_spans<<<blocks, kBlock, 0, g->stream>>>(f);
    PTG_HIP(hipGetLastError());
    return PTG_OK;
}

int fill_u32(ptg_graph *g, uint32_t *p, uint64_t n, uint32_t v) {
    if (n == 0) return PTG_OK;
    const bool vec = ((uintptr_t)p & 15u) == 0;
    const uint32_t blocks = (uint32_t)std::min<uint64_t>(((vec ? n / 4 : n) + kBlock - 1) / kBlock + 1, g->grid);
    k_fill_u32<<<blocks, kBlock, 0, g->stream>>>(p, n, v, vec);
    PTG_HIP(hipGetLastError());
    return PTG_OK;
}

// Streaming copy: uint4 per lane over the aligned body, bytes for the rest.
__global__ __launch_bounds__(kBlock) void k_copy_bytes(uint8_t *__restrict__ dst, const uint8_t *__restrict__ src,
                                                       uint64_t n, bool vec) {
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const uint64_t n16 = vec ? n / 16 : 0;
    const uint4 *s4 = reinterpret_cast<const uint4 *>(src);
    uint4 *d4 = reinterpret_cast<uint4 *>(dst);
    for (uint64_t j = i; j < n16; j += stride) d4[j] = s4[j];
    for (uint64_t j = n16 * 16 + i; j < n; j += stride) dst[j] = src[j];
}

int copy_bytes(void *dst, const void *src, uint64_t n, hipStream_t stream) {
    if (n == 0) return PTG_OK;
    const bool vec = (((uintptr_t)dst | (uintptr_t)src) & 15u) == 0;
    const uint64_t units = vec ? n / 16 : n;
    const uint32_t blocks = (uint32_t)std::min<uint64_t>((units + kBlock - 1) / kBlock + 1, 4096);
    k_copy_bytes<<<blocks, kBlock, 0, stream>>>(static_cast<uint8_t *>(dst), static_cast<const uint8_t *>(src), n, vec);
    PTG_HIP(hipGetLastError());
    return PTG_OK;
}

// out[v] = src[perm[v]]: a result in device ids back to the caller's ids.
__global__ __launch_bounds__(kBlock) void k_unpermute(uint32_t V, const uint32_t *__restrict__ perm,
                                                      const uint32_t *__restrict__ src, uint32_t *__restrict__ dst) {
    const uint32_t stride = gridDim.x * blockDim.x;
    for (uint32_t v = blockIdx.x * blockDim.x + threadIdx.x; v < V; v += stride) dst[v] = src[perm[v]];
}

int unpermute(ptg_graph *g, const uint32_t *src, uint32_t *dst) {
    k_unpermute<<<g->grid, kBlock, 0, g->stream>>>(g->V, g->d_perm, src, dst);
    PTG_HIP(hipGetLastError());
    return PTG_OK;
}

// ---------------------------------------------------------------------------
// Symmetry check (once per graph, at creation): the multiset fingerprints of
// the pairs (v, u) and (u, v) must agree (fingerprint.h). One pass over the
// static degree bins reads every column once; each workgroup leaves its four
// partial products, which the host multiplies.
// ---------------------------------------------------------------------------

struct FpOp {
    const uint32_t *__restrict__ col;
    FpKeys k;
    mutable FpAcc a;
    __device__ __forceinline__ uint32_t vertex(uint32_t v) const { return v; }
    __device__ __forceinline__ void edge(bool valid, uint32_t v, uint64_t e) const {
        if (valid) fp_edge(k, a, v, col[e]);
    }
};

// Triples (v, u, w) from (column, weight) pairs.
struct FpWOp {
    const uint2 *__restrict__ cw;
    FpKeys k;
    mutable FpAcc a;
    __device__ __forceinline__ uint32_t vertex(uint32_t v) const { return v; }
    __device__ __forceinline__ void edge(bool valid, uint32_t v, uint64_t e) const {
        if (valid) {
            const uint2 x = cw[e];
            fp_edge_w(k, a, v, x.x, x.y);
        }
    }
};

__device__ __forceinline__ uint64_t wave_fp_product(uint64_t x) {
    for (int d = 32; d >= 1; d >>= 1) {
        const uint32_t lo = (uint32_t)__shfl_xor((int)(uint32_t)x, d, 64);
        const uint32_t hi = (uint32_t)__shfl_xor((int)(uint32_t)(x >> 32), d, 64);
        x = fp_mul(x, (uint64_t)hi << 32 | lo);
    }
    return x;
}

template <class Op>
__global__ __launch_bounds__(kBlock) void k_fingerprint(Frontier in, const uint32_t *__restrict__ off, Op op,
                                                        unsigned long long *__restrict__ part) {
    __shared__ uint64_t s[4][kWavesPerBlock];
    expand_static(in, off, op);
    const uint64_t x[4] = {op.a.f[0], op.a.f[1], op.a.b[0], op.a.b[1]};
    const uint32_t wib = threadIdx.x >> 6;
#pragma unroll
    for (int i = 0; i < 4; i++) {
        const uint64_t y = wave_fp_product(x[i]);
        if (lane_id() == 0) s[i][wib] = y;
    }
    __syncthreads();
    if (threadIdx.x < 4) {
        uint64_t p = 1;
        for (int w = 0; w < kWavesPerBlock; w++) p = fp_mul(p, s[threadIdx.x][w]);
        part[4ull * blockIdx.x + threadIdx.x] = p;
    }
}

// Runs a fingerprint pass with `op` over the out-CSR; *equal = forward and
// backward products agree for both key sets.
template <class Op>
static int fingerprint_equal(ptg_graph *g, Op op, bool *equal) {
    const uint32_t blocks = g->grid;
    unsigned long long *part = nullptr;
    std::vector<unsigned long long> h(4ull * blocks);
    PTG_HIP(hipMalloc((void **)&part, h.size() * sizeof(unsigned long long)));
    const Frontier fin = g->out.bins.q.with(g->out.bins.d_ctr);
    k_fingerprint<<<blocks, kBlock, 0, g->stream>>>(fin, g->out.off, op, part);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) e = hipMemcpyAsync(h.data(), part, h.size() * sizeof(unsigned long long),
                                            hipMemcpyDeviceToHost, g->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(g->stream);
    (void)hipFree(part);
    if (e != hipSuccess) return hip_fail(e, "symmetry fingerprint");
    uint64_t p[4] = {1, 1, 1, 1};
    for (uint32_t b = 0; b < blocks; b++)
        for (int i = 0; i < 4; i++) p[i] = fp_mul(p[i], h[4ull * b + i]);
    *equal = p[0] == p[2] && p[1] == p[3];
    return PTG_OK;
}

int check_symmetric(ptg_graph *g, int *sym) {
    if (g->symmetric < 0) {
        bool eq = false;
        int rc = fingerprint_equal(g, FpOp{g->out.col, fp_random_keys(), FpAcc{{1, 1}, {1, 1}}}, &eq);
        if (rc) return rc;
        g->symmetric = eq ? 1 : 0;
    }
    *sym = g->symmetric;
    return PTG_OK;
}

int check_weight_symmetric(ptg_graph *g, const uint2 *cw, int *wsym) {
    if (g->wsymmetric < 0) {
        bool eq = false;
        int rc = fingerprint_equal(g, FpWOp{cw, fp_random_keys(), FpAcc{{1, 1}, {1, 1}}}, &eq);
        if (rc) return rc;
        g->wsymmetric = eq ? 1 : 0;
    }
    *wsym = g->wsymmetric;
    return PTG_OK;
}

FpKeys fp_random_keys() {
    std::random_device rd;
    FpKeys k{};
    auto draw = [&rd]() {
        uint64_t x = 0;
        while (x == 0) x = fp_reduce(((uint64_t)rd() << 32 | rd()) & ((1ull << 62) - 1));
        return x;
    };
    for (int i = 0; i < 2; i++) {
        k.r[i] = draw();
        k.s[i] = draw();
        k.t[i] = draw();
    }
    return k;
}

// ---------------------------------------------------------------------------
// Reached count and m_r = sum of out-degrees of reached vertices.
// ---------------------------------------------------------------------------

// Sums per workgroup, then adds to 1 of 16 shards of (count, edges) pairs.
__global__ __launch_bounds__(kBlock) void k_reach_stats(uint32_t V, const uint32_t *__restrict__ val,
                                                        uint32_t unreached, const uint32_t *__restrict__ off,
                                                        unsigned long long *out) {
    __shared__ unsigned long long s_c[kWavesPerBlock], s_e[kWavesPerBlock];
    const uint32_t lane = lane_id(), wib = threadIdx.x >> 6;
    const uint32_t gw = blockIdx.x * kWavesPerBlock + wib;
    const uint32_t nw = gridDim.x * kWavesPerBlock;
    uint64_t cnt = 0, edges = 0;
    // four 64-vertex tiles per step. Every lane issues every load (indexes
    // clamped into range, masked afterwards), so the compiler can keep all of
    // a step's loads in flight; off[v + 1] hits the lines off[v] brought in.
    constexpr uint32_t kU = 4;
    for (uint64_t base = (uint64_t)gw * 64; base < V; base += (uint64_t)nw * 64 * kU) {
        uint32_t x[kU], o[kU], o1[kU];
#pragma unroll
        for (uint32_t j = 0; j < kU; j++) {
            const uint64_t v = base + (uint64_t)j * nw * 64 + lane;
            const uint64_t vc = v < V ? v : V - 1;
            x[j] = val[vc];
            o[j] = off[vc];
            o1[j] = off[vc + 1];
        }
#pragma unroll
        for (uint32_t j = 0; j < kU; j++) {
            const uint64_t v = base + (uint64_t)j * nw * 64 + lane;
            if (v < V && x[j] != unreached) {
                cnt++;
                edges += o1[j] - o[j];
            }
        }
    }
    cnt = wave_sum_u64(cnt);
    edges = wave_sum_u64(edges);
    if (lane == 0) {
        s_c[wib] = cnt;
        s_e[wib] = edges;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned long long c = 0, e = 0;
        for (int w = 0; w < kWavesPerBlock; w++) {
            c += s_c[w];
            e += s_e[w];
        }
        const uint32_t s = 2 * (blockIdx.x % kShards);
        if (c) {
            atomicAdd(&out[s], c);
            atomicAdd(&out[s + 1], e);
        }
    }
}

FillSpan reach_reset_span(ptg_graph *g) { return FillSpan{reinterpret_cast<uint32_t *>(g->d_acc), 4 * kShards, 0u}; }

static int reach_count(ptg_graph *g, const uint32_t *val, uint32_t unreached, unsigned long long *h) {
    // 512 workgroups: twice as fast as the 2048 of a round grid for this scan
    k_reach_stats<<<std::min(g->grid, 512u), kBlock, 0, g->stream>>>(g->V, val, unreached, g->out.off, g->d_acc);
    PTG_HIP(hipGetLastError());
    PTG_HIP(hipMemcpyAsync(h, g->d_acc, sizeof(unsigned long long) * 2 * kShards, hipMemcpyDeviceToHost, g->stream));
    return PTG_OK;
}

static void reach_sum(const unsigned long long *h, uint32_t *reached, uint64_t *edges) {
    uint64_t c = 0, e = 0;
    for (int s = 0; s < kShards; s++) {
        c += h[2 * s];
        e += h[2 * s + 1];
    }
    if (reached) *reached = (uint32_t)c;
    if (edges) *edges = e;
}

// The in-query form reads back into the pinned h_acc (in-memory graphs).
int reach_stats_launch(ptg_graph *g, const uint32_t *val, uint32_t unreached) {
    if (!g->h_acc) return fail(PTG_ERR_UNSUPPORTED, "reach_stats_launch: no pinned accumulators");
    return reach_count(g, val, unreached, g->h_acc);
}

void reach_stats_read(const ptg_graph *g, uint32_t *reached, uint64_t *edges) { reach_sum(g->h_acc, reached, edges); }

int reach_stats(ptg_graph *g, const uint32_t *val, uint32_t unreached, uint32_t *reached, uint64_t *edges) {
    if (!reached && !edges) return PTG_OK;
    const FillSpan z = reach_reset_span(g);
    unsigned long long h[2 * kShards];
    int rc;
    if ((rc = fill_spans(g, &z, 1)) || (rc = reach_count(g, val, unreached, h))) return rc;
    PTG_HIP(hipStreamSynchronize(g->stream));
    reach_sum(h, reached, edges);
    return PTG_OK;
}

int deliver_output(ptg_graph *g, bool speculative) {
    if (g->deliver && (!speculative || g->deliver_on_device)) {
        int rc = g->deliver();
        if (rc) return rc;
        g->delivered = true;
    }
    return PTG_OK;
}


// ---------------------------------------------------------------------------
// Row checksums: sum over v of row[v] * (2v + 1) mod 2^64, per row.
// ---------------------------------------------------------------------------

__device__ __forceinline__ void flush_row(unsigned long long x, unsigned long long *dst, unsigned long long *s) {
    x = wave_sum_u64(x);
    if (lane_id() == 0) s[threadIdx.x >> 6] = x;
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned long long a = 0;
        for (int w = 0; w < kWavesPerBlock; w++) a += s[w];
        atomicAdd(dst, a);
    }
    __syncthreads();
}

// Workgroup b sums the contiguous tile range [b*T, (b+1)*T) of the n rows
// and adds once per row it touched.
__global__ __launch_bounds__(kBlock) void k_row_checksums(const uint32_t *__restrict__ rows, uint32_t n, uint32_t V,
                                                          unsigned long long *out) {
    __shared__ unsigned long long s[kWavesPerBlock];
    const uint64_t per_row = ((uint64_t)V + kBlock - 1) / kBlock;
    const uint64_t total = per_row * n;
    const uint64_t T = (total + gridDim.x - 1) / gridDim.x;
    const uint64_t t0 = (uint64_t)blockIdx.x * T, t1 = min(total, t0 + T);
    if (t0 >= t1) return;
    uint32_t cur = (uint32_t)(t0 / per_row);
    unsigned long long acc = 0;
    for (uint64_t t = t0; t < t1; t++) {
        const uint32_t r = (uint32_t)(t / per_row);
        if (r != cur) {
            flush_row(acc, &out[cur], s);
            acc = 0;
            cur = r;
        }
        const uint64_t v = (t % per_row) * kBlock + threadIdx.x;
        if (v < V) acc += (unsigned long long)rows[(uint64_t)r * V + v] * (2 * v + 1);
    }
    flush_row(acc, &out[cur], s);
}

int row_checksums(ptg_graph *g, hipStream_t stream, const uint32_t *rows, uint32_t n, unsigned long long *d_out) {
    if (n == 0) return PTG_OK;
    k_fill_u32<<<1, kBlock, 0, stream>>>(reinterpret_cast<uint32_t *>(d_out), 2ull * n, 0u, false);
    const uint64_t tiles = ((uint64_t)g->V + kBlock - 1) / kBlock * n;
    const uint32_t blocks = (uint32_t)std::min<uint64_t>(tiles, 1024);
    k_row_checksums<<<blocks, kBlock, 0, stream>>>(rows, n, g->V, d_out);
    PTG_HIP(hipGetLastError());
    return PTG_OK;
}

}  // namespace ptg
