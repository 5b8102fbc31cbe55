// PageRank for gfx950: pull over in-edges, FP32 with compensated sums.
//
// The reference has no PageRank arithmetic (prCaculateInShare,
// ptGraph/main.cu:891-959, launches nothing; testPagerank runs CC at
// ptGraph/main.cu:125). The definition used here:
//   r0(v) = 1/V
//   r'(v) = (1-d)/V + d * ( sum_{u->v} r(u)/outdeg(u) + D/V ),
//   D     = sum of r(u) over vertices with outdeg(u) = 0.
// Per iteration, a gather kernel over static in-degree bins (hub lists split
// into chunks) and a combine kernel that finishes the hubs; the next gather
// sums the dangling mass from both kernels' per-workgroup partials. Kernel
// boundaries order all of it, so no inter-workgroup fences or counters are
// needed. Every sum is a (hi, lo) pair updated with
// Knuth's TwoSum, and every reduction runs in a fixed order, so results are
// bit-identical run to run.
#include <string.h>

#include <algorithm>
#include <mutex>
#include <string>
#include <vector>

#include "device_util.h"

namespace ptg {

struct CF {
    float hi, lo;
};

__device__ __forceinline__ void cf_add(CF &a, float x) {
    const float s = a.hi + x;
    const float bb = s - a.hi;
    const float err = (a.hi - (s - bb)) + (x - bb);
    a.hi = s;
    a.lo += err;
}

__device__ __forceinline__ CF cf_merge(CF a, CF b) {
    const float s = a.hi + b.hi;
    const float bb = s - a.hi;
    const float err = (a.hi - (s - bb)) + (b.hi - bb);
    return CF{s, (a.lo + b.lo) + err};
}

__device__ __forceinline__ CF wave_cf_sum(CF a) {
    for (int d = 32; d >= 1; d >>= 1) {
        CF b;
        b.hi = __shfl_xor(a.hi, d, 64);
        b.lo = __shfl_xor(a.lo, d, 64);
        a = cf_merge(a, b);
    }
    return a;
}

// Sum over the workgroup in a fixed order; every thread gets the result.
__device__ __forceinline__ CF block_cf_sum(CF a, CF *lds) {
    a = wave_cf_sum(a);
    if (lane_id() == 0) lds[threadIdx.x >> 6] = a;
    __syncthreads();
    CF t = lds[0];
    for (int w = 1; w < kWavesPerBlock; w++) t = cf_merge(t, lds[w]);
    __syncthreads();
    return t;
}

struct PrArgs {
    Frontier bins;                  // in-degree bins
    const uint32_t *zero_list;      // vertices with in-degree 0
    uint32_t n_zero;
    const uint32_t *ioff, *icol;    // in-edges
    const uint32_t *ooff;           // out offsets (outdeg)
    const float *cin;               // r(u)/outdeg(u) of the previous iteration
    float *cout;                    // same for this iteration
    float *rank_out;                // ranks (last iteration only) or null
    const float *d0;                // (hi, lo) of the dangling mass D for iteration 0, or null
    const float *part_in;           // previous iteration's dangling partials: gather workgroups
    const float *cpart_in;          //   and combine workgroups, (hi, lo) each
    float *part_out, *cpart_out;    // this iteration's partials
    float *base_out;                // this iteration's base value, for the combine kernel
    float *lpart;                   // hub chunk partials (hi, lo)
    uint32_t gather_blocks, combine_blocks;
    uint32_t n_vert;                // V
    float damping, inv_v;
};

struct PrVertex {
    const PrArgs &a;
    float base;
    CF &dang;
    // r = base + d * sum; contrib = r / outdeg (0 for a dangling vertex,
    // whose r joins the dangling mass instead).
    __device__ __forceinline__ void finalize(uint32_t v, float sum) const {
        const float r = base + a.damping * sum;
        const uint32_t od = a.ooff[v + 1] - a.ooff[v];
        a.cout[v] = od ? r / (float)od : 0.0f;
        if (a.rank_out) a.rank_out[v] = r;
        if (!od) cf_add(dang, r);
    }
};

// base = (1-d)/V + d * D/V, with D summed from the previous iteration's
// partials in the same fixed order by every workgroup.
__device__ __forceinline__ float pr_base(const PrArgs &a, CF *lds) {
    float dang;
    if (a.d0) {
        dang = a.d0[0] + a.d0[1];
    } else {
        CF t{0.0f, 0.0f};
        for (uint32_t j = threadIdx.x; j < a.gather_blocks; j += kBlock)
            t = cf_merge(t, CF{a.part_in[2 * j], a.part_in[2 * j + 1]});
        for (uint32_t j = threadIdx.x; j < a.combine_blocks; j += kBlock)
            t = cf_merge(t, CF{a.cpart_in[2 * j], a.cpart_in[2 * j + 1]});
        t = block_cf_sum(t, lds);
        dang = t.hi + t.lo;
    }
    return (1.0f - a.damping) * a.inv_v + a.damping * (dang * a.inv_v);
}

constexpr uint32_t kPrHubLds = 5088;  // 20352 B: 8 workgroups of the grid per CU still fit

// Gathers every vertex's in-neighbour contributions. Hub lists are cut into
// chunks whose partial sums go to lpart; k_pr_combine finishes those
// vertices. Each workgroup leaves its share of the dangling mass in part_out.
__global__ __launch_bounds__(kBlock) void k_pr_gather(PrArgs a) {
    __shared__ CF lds[kWavesPerBlock];
    // The contributions of the first kPrHubLds vertices are staged in LDS:
    // in the default degree order these are the hubs, about a quarter of the
    // in-list entries at R-MAT scale 22. 20 KB leaves room for the 8
    // workgroups per CU of the grid; 32 KB (5 per CU) was 11% slower, and
    // 128 KB in 1024-thread workgroups (1 per CU) 20% slower.
    __shared__ float s_hub[kPrHubLds];
    const uint32_t nh = min(kPrHubLds, a.n_vert);
    for (uint32_t i = threadIdx.x; i < nh; i += kBlock) s_hub[i] = a.cin[i];
    __syncthreads();
    auto contrib = [&](uint32_t u) -> float { return u < nh ? s_hub[u] : a.cin[u]; };
    CF my_dang{0.0f, 0.0f};
    const float base = pr_base(a, lds);
    if (blockIdx.x == 0 && threadIdx.x == 0) a.base_out[0] = base;
    const PrVertex pv{a, base, my_dang};
    const uint32_t lane = lane_id();

    // Hubs: one workgroup per chunk of kChunkEdges in-edges.
    const unsigned long long lp = a.bins.ctr->l_pack;
    const uint32_t nl = (uint32_t)(lp >> kLCountShift);
    const uint32_t nch = (uint32_t)(lp & kLChunkMask);
    for (uint32_t c = blockIdx.x; c < nch; c += gridDim.x) {
        uint32_t lo = 0, hi = nl;
        while (hi - lo > 1) {
            const uint32_t mid = (lo + hi) >> 1;
            if (a.bins.lb[mid] <= c) lo = mid; else hi = mid;
        }
        const uint32_t v = a.bins.ql[lo];
        const uint64_t b = (uint64_t)a.ioff[v] + (uint64_t)(c - a.bins.lb[lo]) * kChunkEdges;
        const uint64_t e = min((uint64_t)a.ioff[v + 1], b + kChunkEdges);
        CF acc{0.0f, 0.0f};
        for (uint64_t i = b + threadIdx.x; i < e; i += kBlock) cf_add(acc, contrib(a.icol[i]));
        acc = block_cf_sum(acc, lds);
        if (threadIdx.x == 0) {
            a.lpart[2ull * c] = acc.hi;
            a.lpart[2ull * c + 1] = acc.lo;
        }
    }

    // Medium in-degree: one wave per vertex.
    const uint32_t gw = blockIdx.x * kWavesPerBlock + (threadIdx.x >> 6);
    const uint32_t nw = gridDim.x * kWavesPerBlock;
    const uint32_t nm = a.bins.ctr->n_m;
    for (uint32_t i = gw; i < nm; i += nw) {
        const uint32_t v = a.bins.qm[i];
        const uint64_t b = a.ioff[v], e = a.ioff[v + 1];
        CF acc{0.0f, 0.0f};
        for (uint64_t k = b + lane; k < e; k += 64) cf_add(acc, contrib(a.icol[k]));
        acc = wave_cf_sum(acc);
        if (lane == 0) pv.finalize(v, acc.hi + acc.lo);
    }

    // Small in-degree: one lane per vertex; then in-degree 0.
    const uint32_t gt = blockIdx.x * kBlock + threadIdx.x;
    const uint32_t nt = gridDim.x * kBlock;
    const uint32_t ns = a.bins.ctr->n_s;
    for (uint32_t i = gt; i < ns; i += nt) {
        const uint32_t v = a.bins.qs[i];
        const uint32_t b = a.ioff[v], e = a.ioff[v + 1];
        CF acc{0.0f, 0.0f};
        for (uint32_t k = b; k < e; k++) cf_add(acc, contrib(a.icol[k]));
        pv.finalize(v, acc.hi + acc.lo);
    }
    for (uint32_t i = gt; i < a.n_zero; i += nt) pv.finalize(a.zero_list[i], 0.0f);

    const CF bd = block_cf_sum(my_dang, lds);
    if (threadIdx.x == 0) {
        a.part_out[2 * blockIdx.x] = bd.hi;
        a.part_out[2 * blockIdx.x + 1] = bd.lo;
    }
}

// Finishes the hub vertices: one wave per hub sums its chunk partials (lane
// stride, then a fixed butterfly). Each workgroup leaves the dangling mass of
// its hubs in cpart_out.
__global__ __launch_bounds__(kBlock) void k_pr_combine(PrArgs a) {
    __shared__ CF lds[kWavesPerBlock];
    CF my_dang{0.0f, 0.0f};
    const PrVertex pv{a, a.base_out[0], my_dang};
    const uint32_t lane = lane_id();
    const uint32_t gw = blockIdx.x * kWavesPerBlock + (threadIdx.x >> 6);
    const uint32_t nw = gridDim.x * kWavesPerBlock;
    const uint32_t nl = (uint32_t)(a.bins.ctr->l_pack >> kLCountShift);
    for (uint32_t l = gw; l < nl; l += nw) {
        const uint32_t v = a.bins.ql[l];
        const uint32_t first = a.bins.lb[l];
        const uint32_t n = bin_chunks(a.ioff[v + 1] - a.ioff[v]);
        CF t{0.0f, 0.0f};
        for (uint32_t j = lane; j < n; j += 64)
            t = cf_merge(t, CF{a.lpart[2ull * (first + j)], a.lpart[2ull * (first + j) + 1]});
        t = wave_cf_sum(t);
        if (lane == 0) pv.finalize(v, t.hi + t.lo);
    }
    const CF bd = block_cf_sum(my_dang, lds);
    if (threadIdx.x == 0) {
        a.cpart_out[2 * blockIdx.x] = bd.hi;
        a.cpart_out[2 * blockIdx.x + 1] = bd.lo;
    }
}

// r0 = 1/V: contrib0(v) = (1/V)/outdeg(v), or 0 for a dangling vertex.
__global__ __launch_bounds__(kBlock) void k_pr_init(uint32_t V, const uint32_t *__restrict__ ooff, float r0,
                                                    float *__restrict__ contrib) {
    const uint32_t stride = gridDim.x * blockDim.x;
    for (uint32_t v = blockIdx.x * blockDim.x + threadIdx.x; v < V; v += stride) {
        const uint32_t od = ooff[v + 1] - ooff[v];
        contrib[v] = od ? r0 / (float)od : 0.0f;
    }
}

__global__ void k_set_f32x2(float *p, float x, float y) {
    if (threadIdx.x == 0) {
        p[0] = x;
        p[1] = y;
    }
}

// Transposed CSR on the host (counting sort by destination, sources visited in
// ascending order, so every in-list is sorted and the layout is deterministic).
int ensure_in_edges(ptg_graph *g) {
    if (g->have_in) return PTG_OK;
    int sym = 0;
    int rc = check_symmetric(g, &sym);
    if (rc) return rc;
    if (sym) {
        g->in = g->out;
        g->in_is_out = true;
        g->have_in = true;
        return PTG_OK;
    }
    const uint32_t V = g->V;
    const uint64_t E = g->E;
    std::vector<uint32_t> off((size_t)V + 1), col(E ? E : 1), toff((size_t)V + 1, 0), tcol(E ? E : 1);
    if ((rc = copy_sync(g, off.data(), g->out.off, sizeof(uint32_t) * ((uint64_t)V + 1), hipMemcpyDeviceToHost)) ||
        (rc = copy_sync(g, col.data(), g->out.col, sizeof(uint32_t) * E, hipMemcpyDeviceToHost)))
        return rc;
    for (uint64_t e = 0; e < E; e++) toff[col[e] + 1]++;
    for (uint32_t v = 0; v < V; v++) toff[v + 1] += toff[v];
    std::vector<uint32_t> cur(toff.begin(), toff.end() - 1);
    for (uint32_t u = 0; u < V; u++)
        for (uint32_t k = off[u]; k < off[u + 1]; k++) tcol[cur[col[k]]++] = u;
    if ((rc = galloc(g, (void **)&g->in.off, sizeof(uint32_t) * ((uint64_t)V + 1))) ||
        (rc = galloc(g, (void **)&g->in.col, sizeof(uint32_t) * E)))
        return rc;
    if ((rc = copy_sync(g, g->in.off, toff.data(), sizeof(uint32_t) * ((uint64_t)V + 1), hipMemcpyHostToDevice)) ||
        (rc = copy_sync(g, g->in.col, tcol.data(), sizeof(uint32_t) * E, hipMemcpyHostToDevice)))
        return rc;
    if ((rc = build_static_bins(g, toff.data(), V, &g->in.bins))) return rc;
    g->in_is_out = false;
    g->have_in = true;
    return PTG_OK;
}

constexpr uint32_t kMaxCombineBlocks = 256;

static uint32_t combine_blocks(const ptg_graph *g) {
    const uint64_t nl = g->in.bins.h_ctr.l_pack >> kLCountShift;
    return (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>(kMaxCombineBlocks, (nl + kWavesPerBlock - 1) / kWavesPerBlock));
}

// dangling and base partials: 2 iterations x (gather, combine) + base + D0
static uint64_t pr_part_floats(const ptg_graph *g) { return 2 * (2ull * g->grid + 2ull * kMaxCombineBlocks) + 4; }

static int ensure_pr_state(ptg_graph *g) {
    if (g->d_contrib[0]) return PTG_OK;
    int rc;
    const uint64_t chunks = g->in.bins.h_ctr.l_pack & kLChunkMask;
    // d_part: 2 iterations x (gather partials, combine partials) + base + D0
    const uint64_t part_floats = pr_part_floats(g);
    if ((rc = galloc(g, (void **)&g->d_contrib[0], sizeof(float) * g->V)) ||
        (rc = galloc(g, (void **)&g->d_contrib[1], sizeof(float) * g->V)) ||
        (rc = galloc(g, (void **)&g->d_part, sizeof(float) * part_floats)) ||
        (rc = galloc(g, (void **)&g->d_lpart, sizeof(float) * 2 * (chunks + 1))))
        return rc;
    g->lchunks_cap = (uint32_t)chunks;
    return PTG_OK;
}

uint64_t pr_state_need(const ptg_graph *g) {
    if (g->d_contrib[0]) return 0;
    auto al = [](uint64_t x) { return (x + 255) & ~255ull; };
    const uint64_t chunks = g->in.bins.h_ctr.l_pack & kLChunkMask;
    uint64_t need = 2 * al(4ull * g->V) + al(4 * (2 * (2ull * g->grid + 2ull * kMaxCombineBlocks) + 4)) +
                    al(8 * (chunks + 1));
    if (!g->have_in) need += al(4ull * (g->V + 1)) + al(4 * g->E) + al(8ull * g->V);  // upper bound: transpose
    return need;
}

void release_pr_state(ptg_graph *g) {
    for (auto *&c : g->d_contrib) {
        gfree(g, c);
        c = nullptr;
    }
    gfree(g, g->d_part);
    gfree(g, g->d_lpart);
    g->d_part = g->d_lpart = nullptr;
    g->lchunks_cap = 0;
}

// Per-iteration buffers of one PageRank run: contributions (2 x V floats),
// dangling and base partials, hub chunk partials.
struct PrBufs {
    float *contrib[2];
    float *part;
    float *lpart;
};


static int pagerank_core(ptg_graph *g, float damping, int iters, float *ranks, ptg_stats *st, const PrBufs &buf,
                         bool timed) {
    int rc;
    const uint32_t V = g->V;
    const float r0 = 1.0f / (float)V;
    LaunchTimer timer(g);
    if (!timed) timer.on = false;
    PTG_HIP(hipEventRecord(g->ev0, g->stream));
    if (iters == 0) {
        union {
            float f;
            uint32_t u;
        } cv;
        cv.f = r0;
        if ((rc = fill_u32(g, reinterpret_cast<uint32_t *>(ranks), V, cv.u))) return rc;
    } else {
        const uint32_t blocks = std::min<uint32_t>((V + kBlock - 1) / kBlock, g->grid);
        k_pr_init<<<blocks, kBlock, 0, g->stream>>>(V, g->out.off, r0, buf.contrib[0]);
        // partial buffers: [it parity][gather | combine], then base, then D0
        const uint32_t gb = g->grid, cb = combine_blocks(g);
        float *part[2] = {buf.part, buf.part + 2 * gb};
        float *cpart[2] = {buf.part + 4 * gb, buf.part + 4 * gb + 2 * kMaxCombineBlocks};
        float *base = buf.part + 4 * gb + 4 * kMaxCombineBlocks;
        float *d0 = base + 2;
        // D0 = (number of dangling vertices) / V, split into (hi, lo).
        const double d0v = (double)g->out.bins.n_zero / (double)V;
        const float d0h = (float)d0v;
        k_set_f32x2<<<1, 64, 0, g->stream>>>(d0, d0h, (float)(d0v - (double)d0h));
        PTG_HIP(hipGetLastError());
        for (int t = 0; t < iters; t++) {
            PrArgs a;
            a.bins = g->in.bins.q.with(g->in.bins.d_ctr);
            a.zero_list = g->in.bins.zero;
            a.n_zero = g->in.bins.n_zero;
            a.ioff = g->in.off;
            a.icol = g->in.col;
            a.ooff = g->out.off;
            a.cin = buf.contrib[t & 1];
            a.cout = buf.contrib[(t + 1) & 1];
            a.rank_out = (t == iters - 1) ? ranks : nullptr;
            a.d0 = t == 0 ? d0 : nullptr;
            a.part_in = part[(t + 1) & 1];
            a.cpart_in = cpart[(t + 1) & 1];
            a.part_out = part[t & 1];
            a.cpart_out = cpart[t & 1];
            a.base_out = base;
            a.lpart = buf.lpart;
            a.gather_blocks = gb;
            a.combine_blocks = cb;
            a.n_vert = V;
            a.damping = damping;
            a.inv_v = r0;
            timer.launch(k_pr_gather, dim3(g->grid), dim3(kBlock), a);
            k_pr_combine<<<cb, kBlock, 0, g->stream>>>(a);
            PTG_HIP(hipGetLastError());
        }
    }
    PTG_HIP(hipEventRecord(g->ev1, g->stream));
    PTG_HIP(hipEventSynchronize(g->ev1));
    if (st) {
        memset(st, 0, sizeof *st);
        float ms = 0;
        PTG_HIP(hipEventElapsedTime(&ms, g->ev0, g->ev1));
        st->kernel_ns = (uint64_t)((double)ms * 1e6);
        st->iterations = (uint32_t)iters;
        st->reached = V;
        st->edges_traversed = g->E * (uint64_t)iters;
        if ((rc = timer.collect(st))) return rc;
    }
    return PTG_OK;
}

int run_pagerank(ptg_graph *g, float damping, int iters, float *ranks, ptg_stats *st) {
    int rc;
    if ((rc = ensure_in_edges(g)) || (rc = ensure_pr_state(g))) return rc;
    const PrBufs buf{{g->d_contrib[0], g->d_contrib[1]}, g->d_part, g->d_lpart};
    return pagerank_core(g, damping, iters, ranks, st, buf, true);
}

// ---------------------------------------------------------------------------
// PageRank without allocation: the per-iteration buffers live in a fixed
// device array of kPrScratchBytes, one copy per GPU, shared by every call on
// that GPU (calls that overlap on one GPU race on it). Every word a run reads
// is written earlier in the same run (k_pr_init, the D0 pair, then each
// iteration's partials before the next reads them), so nothing is cleared.
// ---------------------------------------------------------------------------

static __device__ __align__(256) uint8_t scratch[kPrScratchBytes];

constexpr int kMaxDevices = 64;
static std::mutex g_scratch_mu;
static void *g_scratch_ptr[kMaxDevices];

// Looks up this GPU's copy of `scratch` once, with `device` current.
int pr_scratch_setup(int device) {
    if (device < 0 || device >= kMaxDevices) return fail(PTG_ERR_INVALID_ARG, "device index out of range");
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (g_scratch_ptr[device]) return PTG_OK;
    int cur = -1;
    PTG_HIP(hipGetDevice(&cur));
    if (cur != device) PTG_HIP(hipSetDevice(device));
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (cur != device) (void)hipSetDevice(cur);
    if (e != hipSuccess) return hip_fail(e, "hipGetSymbolAddress(scratch)");
    g_scratch_ptr[device] = p;
    return PTG_OK;
}

static uint64_t al256(uint64_t x) { return (x + 255) & ~255ull; }

uint64_t pr_scratch_need(const ptg_graph *g) {
    const uint64_t chunks = g->in.bins.h_ctr.l_pack & kLChunkMask;
    return 2 * al256(4ull * g->V) + al256(4 * pr_part_floats(g)) + al256(8 * (chunks + 1));
}

int run_pagerank_noalloc(ptg_graph *g, float damping, int iters, float *ranks, ptg_stats *st) {
    int rc;
    // in-edges without allocating: the graph's own lists when it is
    // symmetric, else a transpose an earlier ptg_pagerank built
    if (!g->have_in) {
        int sym = g->symmetric;
        if (sym < 0) return fail(PTG_ERR_UNSUPPORTED, "pagerank_noalloc: symmetry not yet known");
        if (!sym)
            return fail(PTG_ERR_UNSUPPORTED,
                        "pagerank_noalloc: graph is not symmetric and has no in-edges yet (run ptg_pagerank once)");
        if ((rc = ensure_in_edges(g))) return rc;  // aliases the out-lists, no allocation
    }
    const uint64_t need = pr_scratch_need(g);
    if (need > kPrScratchBytes)
        return fail(PTG_ERR_BUDGET, "pagerank_noalloc: needs " + std::to_string(need) + " scratch bytes, the maximum is " +
                                        std::to_string(kPrScratchBytes) + " (about 8M vertices)");
    if ((rc = pr_scratch_setup(g->device))) return rc;
    uint8_t *p = static_cast<uint8_t *>(g_scratch_ptr[g->device]);
    PrBufs buf;
    buf.contrib[0] = reinterpret_cast<float *>(p);
    p += al256(4ull * g->V);
    buf.contrib[1] = reinterpret_cast<float *>(p);
    p += al256(4ull * g->V);
    buf.part = reinterpret_cast<float *>(p);
    p += al256(4 * pr_part_floats(g));
    buf.lpart = reinterpret_cast<float *>(p);
    return pagerank_core(g, damping, iters, ranks, st, buf, false);
}

}  // namespace ptg
