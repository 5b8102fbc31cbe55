// PageRank for gfx950: pull over in-edges, FP32 ranks from exact sums.
//
// The reference has no PageRank arithmetic (prCaculateInShare,
// ptGraph/main.cu:891-959, launches nothing; testPagerank runs CC at
// ptGraph/main.cu:125). The definition used here:
//   r0(v) = 1/V
//   r'(v) = (1-d)/V + d * ( sum_{u->v} r(u)/outdeg(u) + D/V ),
//   D     = sum of r(u) over vertices with outdeg(u) = 0.
//
// Sums. Every sum of contributions, and the dangling mass, is an unsigned
// 64-bit fixed-point number with 63 fraction bits (Q1.63): each FP32 term in
// [0, 1] is truncated to a multiple of 2^-63 and added exactly. A sum is then
// independent of the order and grouping of its terms, and its error is below
// 2^-63 per term. Any split of the work (grid size, hub chunks, the
// allocation-free path's buffers) gives the same bits, run after run.
// The rank is rounded once: r = base + d * sum in FP64 from the exact sum,
// then to FP32.
//
// Work. Per iteration, a gather kernel over static in-degree bins (hub lists
// cut into chunks) and a combine kernel that finishes the hubs; the next
// gather adds up the dangling mass from both kernels' per-workgroup
// partials. Kernel boundaries order all of it, so no inter-workgroup fences
// or counters are needed.
#include <string.h>

#include <algorithm>
#include <mutex>
#include <string>
#include <vector>

#include "device_util.h"

namespace ptg {

// Hub contributions staged in LDS per 256 threads of a gather workgroup:
// 20352 B, so 8 workgroups of 256 threads (or 4 of 512, 2 of 1024) fit a CU.
constexpr uint32_t kPrHubLds = 5088;

// Sum over the workgroup (Waves waves); every thread gets the result.
template <uint32_t Waves>
__device__ __forceinline__ unsigned long long block_sum(unsigned long long x, unsigned long long *lds) {
    x = wave_sum_u64(x);
    if (lane_id() == 0) lds[threadIdx.x >> 6] = x;
    __syncthreads();
    unsigned long long t = 0;
    for (uint32_t w = 0; w < Waves; w++) t += lds[w];
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
    bool first;                     // iteration 0: the dangling mass is d0
    double d0;                      // D of iteration 0 (dangling vertices / V)
    const unsigned long long *part_in;   // previous iteration's dangling partials: gather workgroups
    const unsigned long long *cpart_in;  //   and combine workgroups
    unsigned long long *part_out, *cpart_out;  // this iteration's partials
    float *base_out;                // this iteration's base value, for the combine kernel
    unsigned long long *lpart;      // hub chunk sums
    uint32_t gather_blocks, combine_blocks;
    uint32_t n_vert;                // V
    float damping, inv_v;
};

struct PrVertex {
    const PrArgs &a;
    float base;
    unsigned long long &dang;
    // r = base + d * sum (rounded once); contrib = r / outdeg (0 for a
    // dangling vertex, whose r joins the dangling mass instead).
    __device__ __forceinline__ void finalize(uint32_t v, unsigned long long sum) const {
        const float r = pr_rank(base, a.damping, sum);
        const uint32_t od = a.ooff[v + 1] - a.ooff[v];
        a.cout[v] = od ? r / (float)od : 0.0f;
        if (a.rank_out) a.rank_out[v] = r;
        if (!od) dang += fx63(r);
    }
};

// base = (1-d)/V + d * D/V, D summed from the previous iteration's partials.
template <uint32_t TB>
__device__ __forceinline__ float pr_base(const PrArgs &a, unsigned long long *lds) {
    double dang;
    if (a.first) {
        dang = a.d0;
    } else {
        unsigned long long t = 0;
        for (uint32_t j = threadIdx.x; j < a.gather_blocks; j += TB) t += a.part_in[j];
        for (uint32_t j = threadIdx.x; j < a.combine_blocks; j += TB) t += a.cpart_in[j];
        dang = fx63_to_f64(block_sum<TB / 64>(t, lds));
    }
    return pr_base_value(a.damping, a.inv_v, dang);
}

// Gathers every vertex's in-neighbour contributions with TB threads per
// workgroup; the contributions of the first NH vertices (the hubs, in the
// default degree order: about a quarter of the in-list entries at R-MAT scale
// 22 for 5088 of them) come from LDS. Hub lists are cut into chunks whose sums
// go to lpart; k_pr_combine finishes those vertices. Each workgroup leaves its
// share of the dangling mass in part_out.
template <uint32_t TB, uint32_t NH>
__global__ __launch_bounds__(TB) void k_pr_gather(PrArgs a) {
    constexpr uint32_t kWaves = TB / 64;
    __shared__ unsigned long long lds[kWaves];
    __shared__ float s_hub[NH];
    const uint32_t nh = min(NH, a.n_vert);
    for (uint32_t i = threadIdx.x; i < nh; i += TB) s_hub[i] = a.cin[i];
    __syncthreads();
    auto term = [&](uint32_t u) -> unsigned long long { return fx63(u < nh ? s_hub[u] : a.cin[u]); };
    unsigned long long my_dang = 0;
    const float base = pr_base<TB>(a, lds);
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
        unsigned long long acc = 0;
        for (uint64_t i = b + threadIdx.x; i < e; i += TB) acc += term(a.icol[i]);
        acc = block_sum<kWaves>(acc, lds);
        if (threadIdx.x == 0) a.lpart[c] = acc;
    }

    // Medium in-degree: one wave per vertex.
    const uint32_t gw = blockIdx.x * kWaves + (threadIdx.x >> 6);
    const uint32_t nw = gridDim.x * kWaves;
    const uint32_t nm = a.bins.ctr->n_m;
    for (uint32_t i = gw; i < nm; i += nw) {
        const uint32_t v = a.bins.qm[i];
        unsigned long long acc = 0;
        for (uint64_t k = a.ioff[v] + lane; k < a.ioff[v + 1]; k += 64) acc += term(a.icol[k]);
        acc = wave_sum_u64(acc);
        if (lane == 0) pv.finalize(v, acc);
    }

    // Small in-degree: eight lanes per vertex (a wave takes 8 vertices), so a
    // column load touches 8 lines of 8 consecutive columns instead of 64
    // lines, one per lane; the 8 lanes' sums are added by shuffles.
    const uint32_t ns = a.bins.ctr->n_s;
    const uint32_t sub = lane & 7;
    for (uint32_t i0 = gw * 8; i0 < ns; i0 += nw * 8) {
        const uint32_t i = i0 + (lane >> 3);
        uint32_t v = 0;
        unsigned long long acc = 0;
        if (i < ns) {
            v = a.bins.qs[i];
            for (uint32_t k = a.ioff[v] + sub; k < a.ioff[v + 1]; k += 8) acc += term(a.icol[k]);
        }
        for (int d = 4; d >= 1; d >>= 1) {
            const uint32_t lo = (uint32_t)__shfl_xor((int)(uint32_t)acc, d, 64);
            const uint32_t hi = (uint32_t)__shfl_xor((int)(uint32_t)(acc >> 32), d, 64);
            acc += (unsigned long long)hi << 32 | lo;
        }
        if (sub == 0 && i < ns) pv.finalize(v, acc);
    }
    // then in-degree 0
    const uint32_t gt = blockIdx.x * TB + threadIdx.x;
    const uint32_t nt = gridDim.x * TB;
    for (uint32_t i = gt; i < a.n_zero; i += nt) pv.finalize(a.zero_list[i], 0ull);

    const unsigned long long bd = block_sum<kWaves>(my_dang, lds);
    if (threadIdx.x == 0) a.part_out[blockIdx.x] = bd;
}

// Finishes the hub vertices: one wave per hub adds its chunk sums. Each
// workgroup leaves the dangling mass of its hubs in cpart_out.
__global__ __launch_bounds__(kBlock) void k_pr_combine(PrArgs a) {
    __shared__ unsigned long long lds[kWavesPerBlock];
    unsigned long long my_dang = 0;
    const PrVertex pv{a, a.base_out[0], my_dang};
    const uint32_t lane = lane_id();
    const uint32_t gw = blockIdx.x * kWavesPerBlock + (threadIdx.x >> 6);
    const uint32_t nw = gridDim.x * kWavesPerBlock;
    const uint32_t nl = (uint32_t)(a.bins.ctr->l_pack >> kLCountShift);
    for (uint32_t l = gw; l < nl; l += nw) {
        const uint32_t v = a.bins.ql[l];
        const uint32_t first = a.bins.lb[l];
        const uint32_t n = bin_chunks(a.ioff[v + 1] - a.ioff[v]);
        unsigned long long t = 0;
        for (uint32_t j = lane; j < n; j += 64) t += a.lpart[first + j];
        t = wave_sum_u64(t);
        if (lane == 0) pv.finalize(v, t);
    }
    const unsigned long long bd = block_sum<kWavesPerBlock>(my_dang, lds);
    if (threadIdx.x == 0) a.cpart_out[blockIdx.x] = bd;
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

// Partial words: 2 iterations x (gather partials, combine partials; at most
// grid workgroups each), then the base value (a float in a 64-bit word).
static uint64_t pr_part_words(const ptg_graph *g) { return 4ull * g->grid + 1; }

static int ensure_pr_state(ptg_graph *g) {
    if (g->d_contrib[0]) return PTG_OK;
    int rc;
    const uint64_t chunks = g->in.bins.h_ctr.l_pack & kLChunkMask;
    if ((rc = galloc(g, (void **)&g->d_contrib[0], sizeof(float) * g->V)) ||
        (rc = galloc(g, (void **)&g->d_contrib[1], sizeof(float) * g->V)) ||
        (rc = galloc(g, (void **)&g->d_part, 8 * pr_part_words(g))) ||
        (rc = galloc(g, (void **)&g->d_lpart, 8 * (chunks + 1))))
        return rc;
    g->lchunks_cap = (uint32_t)chunks;
    return PTG_OK;
}

uint64_t pr_state_need(const ptg_graph *g) {
    auto al = [](uint64_t x) { return (x + 255) & ~255ull; };
    uint64_t need = 0;
    if (!g->d_contrib[0]) {
        const uint64_t chunks = g->in.bins.h_ctr.l_pack & kLChunkMask;
        need += 2 * al(4ull * g->V) + al(8 * pr_part_words(g)) + al(8 * (chunks + 1));
    }
    if (!g->have_in) need += al(4ull * (g->V + 1)) + al(4 * g->E) + al(8ull * g->V);  // upper bound: transpose
    if (pr_bins_wanted(g)) need += pr_bins_need(g);
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
    release_pr_bins(g);
}

// Per-iteration buffers of one PageRank run: contributions (2 x V floats),
// dangling partials and the base value, hub chunk sums.
struct PrBufs {
    float *contrib[2];
    unsigned long long *part;
    unsigned long long *lpart;
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
        // partial words: [it parity][gather | combine], then the base value;
        // the gather runs the grid's threads in workgroups of pr_block
        const uint32_t tb = g->cfg.pr_block ? g->cfg.pr_block : 256u;
        const uint32_t gb = std::max<uint32_t>(1, g->grid * 256u / tb);
        const uint32_t cb = std::min<uint32_t>(gb, combine_blocks(g));
        unsigned long long *part[2] = {buf.part, buf.part + gb};
        unsigned long long *cpart[2] = {buf.part + 2 * gb, buf.part + 3 * gb};
        float *base = reinterpret_cast<float *>(buf.part + 4 * gb);
        for (int t = 0; t < iters; t++) {
            PrArgs a;
            memset(&a, 0, sizeof a);
            a.bins = g->in.bins.q.with(g->in.bins.d_ctr);
            a.zero_list = g->in.bins.zero;
            a.n_zero = g->in.bins.n_zero;
            a.ioff = g->in.off;
            a.icol = g->in.col;
            a.ooff = g->out.off;
            a.cin = buf.contrib[t & 1];
            a.cout = buf.contrib[(t + 1) & 1];
            a.rank_out = (t == iters - 1) ? ranks : nullptr;
            a.first = t == 0;
            a.d0 = (double)g->out.bins.n_zero / (double)V;  // D0 = (dangling vertices) / V
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
            if (tb == 1024)
                timer.launch(k_pr_gather<1024, 4 * kPrHubLds>, dim3(gb), dim3(1024), a);
            else if (tb == 512)
                timer.launch(k_pr_gather<512, 2 * kPrHubLds>, dim3(gb), dim3(512), a);
            else
                timer.launch(k_pr_gather<256, kPrHubLds>, dim3(gb), dim3(256), a);
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
    if (pr_bins_wanted(g)) {
        bool done = false;
        if ((rc = run_pagerank_bins(g, damping, iters, ranks, st, &done)) || done) return rc;
    }
    if ((rc = ensure_in_edges(g)) || (rc = ensure_pr_state(g))) return rc;
    const PrBufs buf{{g->d_contrib[0], g->d_contrib[1]}, g->d_part, g->d_lpart};
    return pagerank_core(g, damping, iters, ranks, st, buf, true);
}

// ---------------------------------------------------------------------------
// PageRank without allocation: the per-iteration buffers live in a fixed
// device array of kPrScratchBytes, one copy per GPU, shared by every call on
// that GPU (calls that overlap on one GPU race on it). Every word a run reads
// is written earlier in the same run (k_pr_init's contributions, then each
// iteration's partials and base value before the next reads them; iteration
// 0 takes the dangling mass as a kernel argument), so nothing is cleared.
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
    return 2 * al256(4ull * g->V) + al256(8 * pr_part_words(g)) + al256(8 * (chunks + 1));
}

int run_pagerank_noalloc(ptg_graph *g, float damping, int iters, float *ranks, ptg_stats *st) {
    int rc;
    // in-edges without allocating: the graph's own lists when it is
    // symmetric, else a transpose an earlier in-list pull (ptg_pagerank
    // with pr_layout PTG_PR_PULL) built; the binned layout builds none
    if (!g->have_in) {
        int sym = g->symmetric;
        if (sym < 0) return fail(PTG_ERR_UNSUPPORTED, "pagerank_noalloc: symmetry not yet known");
        if (!sym)
            return fail(PTG_ERR_UNSUPPORTED,
                        "pagerank_noalloc: graph is not symmetric and has no in-edges yet (run ptg_pagerank "
                        "once with pr_layout PTG_PR_PULL)");
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
    buf.part = reinterpret_cast<unsigned long long *>(p);
    p += al256(8 * pr_part_words(g));
    buf.lpart = reinterpret_cast<unsigned long long *>(p);
    return pagerank_core(g, damping, iters, ranks, st, buf, false);
}

}  // namespace ptg
