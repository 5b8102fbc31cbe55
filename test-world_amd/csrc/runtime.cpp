// Host runtime: errors, device selection, buffers, and the graph handle.
// Replaces the per-call allocate/copy/free of every ptGraph driver
// (e.g. ptGraph/main.cu:1683-1830) with one resident handle per graph.
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <vector>

#include "ptg_internal.h"
#include "relabel.h"

namespace ptg {

static thread_local std::string g_last_error;

int fail(int code, const std::string &msg) {
    g_last_error = msg;
    return code;
}

int hip_fail(hipError_t e, const char *what) {
    g_last_error = std::string(what) + ": " + hipGetErrorString(e);
    if (e == hipErrorOutOfMemory || e == hipErrorMemoryAllocation) return PTG_ERR_OUT_OF_MEMORY;
    if (e == hipErrorNoDevice || e == hipErrorInvalidDevice) return PTG_ERR_NO_DEVICE;
    return PTG_ERR_HIP;
}

bool is_device_pointer(const void *p) {
    hipPointerAttribute_t a;
    hipError_t e = hipPointerGetAttributes(&a, p);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        return false;
    }
    return a.type == hipMemoryTypeDevice || a.type == hipMemoryTypeManaged;
}

int copy_sync(ptg_graph *g, void *dst, const void *src, uint64_t bytes, hipMemcpyKind kind) {
    if (bytes == 0) return PTG_OK;
    PTG_HIP(hipMemcpyAsync(dst, src, bytes, kind, g->stream));
    PTG_HIP(hipStreamSynchronize(g->stream));
    return PTG_OK;
}

// A non-blocking stream of this thread for the device-buffer utilities.
static int util_stream(hipStream_t *st) {
    thread_local hipStream_t s = nullptr;
    thread_local int dev = -1;
    int cur = 0;
    PTG_HIP(hipGetDevice(&cur));
    if (!s || dev != cur) {
        PTG_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
        dev = cur;
    }
    *st = s;
    return PTG_OK;
}

int galloc(ptg_graph *g, void **p, uint64_t bytes) {
    *p = nullptr;
    if (bytes == 0) bytes = 256;
    if (g->cfg.device_budget_bytes && g->device_bytes + bytes > g->cfg.device_budget_bytes)
        return fail(PTG_ERR_BUDGET, "device budget of " + std::to_string(g->cfg.device_budget_bytes) +
                                        " bytes exceeded (need " + std::to_string(g->device_bytes + bytes) + ")");
    hipError_t e = hipMalloc(p, bytes);
    if (e != hipSuccess) {
        *p = nullptr;
        return hip_fail(e, "hipMalloc");
    }
    g->device_bytes += bytes;
    g->alloc_bytes[*p] = bytes;
    return PTG_OK;
}

void gfree(ptg_graph *g, void *p) {
    if (!p) return;
    auto it = g->alloc_bytes.find(p);
    if (it != g->alloc_bytes.end()) {
        g->device_bytes -= it->second;
        g->alloc_bytes.erase(it);
    }
    (void)hipFree(p);
}

uint32_t chunk_queue_cap(uint32_t V, uint32_t grid) {
    // The bitmap pre-pass (prepare_grid(grid) workgroups) gives workgroup b
    // the tiles [b*T, (b+1)*T) and appends to shard b % kShards; queue-mode
    // rounds append at most min(V, kQueueCap) entries to any shard.
    const uint64_t npairs = (V + 63) / 64;
    const uint64_t pg = prepare_grid(grid);
    const uint64_t T = (npairs + pg - 1) / pg;
    const uint64_t blocks_per_shard = (pg + kShards - 1) / kShards;
    const uint64_t cap = std::max<uint64_t>(blocks_per_shard * 64 * T, std::min<uint64_t>(V, kQueueCap));
    return (uint32_t)std::min<uint64_t>(cap, 0xFFFFFFFFull);
}

int build_static_bins(ptg_graph *g, const uint32_t *h_off, uint32_t V, StaticBins *sb) {
    const uint32_t sd = g->cfg.small_degree, ld = g->cfg.large_degree;
    std::vector<uint32_t> s, m, l, lb, z;
    uint64_t chunks = 0;
    for (uint32_t v = 0; v < V; v++) {
        uint32_t d = h_off[v + 1] - h_off[v];
        if (d == 0)
            z.push_back(v);
        else if (d <= sd)
            s.push_back(v);
        else if (d <= ld)
            m.push_back(v);
        else {
            l.push_back(v);
            lb.push_back((uint32_t)chunks);
            chunks += (d + kChunkEdges - 1) / kChunkEdges;
        }
    }
    if (chunks > 0xFFFFFFFFull) return fail(PTG_ERR_UNSUPPORTED, "too many hub chunks");
    // five lists, each followed by one pad entry, then 8-byte aligned counters
    const uint64_t n = (uint64_t)s.size() + m.size() + 2 * l.size() + z.size() + 5;
    int rc = galloc(g, (void **)&sb->mem, n * sizeof(uint32_t) + 8 + sizeof(BinCounters));
    if (rc) return rc;
    uint32_t *p = sb->mem;
    auto put = [&](std::vector<uint32_t> &src, uint32_t *&dst) -> int {
        dst = p;
        if (!src.empty()) {
            const int rc = copy_sync(g, p, src.data(), src.size() * sizeof(uint32_t), hipMemcpyHostToDevice);
            if (rc) return rc;
        }
        p += src.size() + 1;
        return PTG_OK;
    };
    if ((rc = put(s, sb->q.qs)) || (rc = put(m, sb->q.qm)) || (rc = put(l, sb->q.ql)) ||
        (rc = put(lb, sb->q.lb)) || (rc = put(z, sb->zero)))
        return rc;
    // counters live after the lists (aligned to 8 bytes)
    uintptr_t cp = ((uintptr_t)p + 7) & ~(uintptr_t)7;
    sb->d_ctr = (BinCounters *)cp;
    memset(&sb->h_ctr, 0, sizeof sb->h_ctr);
    sb->h_ctr.n_s = (uint32_t)s.size();
    sb->h_ctr.n_m = (uint32_t)m.size();
    sb->h_ctr.l_pack = ((unsigned long long)l.size() << kLCountShift) | chunks;
    sb->h_ctr.found = (uint32_t)(s.size() + m.size() + l.size());
    sb->n_zero = (uint32_t)z.size();
    if (int rc = copy_sync(g, sb->d_ctr, &sb->h_ctr, sizeof(BinCounters), hipMemcpyHostToDevice)) return rc;
    return PTG_OK;
}

void free_static_bins(ptg_graph *g, StaticBins *sb) {
    gfree(g, sb->mem);
    *sb = StaticBins{};
}

LaunchTimer::LaunchTimer(ptg_graph *gr) : g(gr), on((gr->cfg.flags & PTG_FLAG_TIMING) != 0) {
    g->tev_used = 0;
}

bool LaunchTimer::next(hipEvent_t *start, hipEvent_t *stop) {
    if (!on) return false;
    const size_t need = 2ull * (g->tev_used + 1);
    if (g->tev.size() < need) {
        if (g->tev.size() >= 2ull * 65536) {  // cap: stop timing very long queries
            on = false;
            return false;
        }
        while (g->tev.size() < need) {
            hipEvent_t e;
            if (hipEventCreate(&e) != hipSuccess) {
                on = false;
                return false;
            }
            g->tev.push_back(e);
        }
    }
    *start = g->tev[2 * g->tev_used];
    *stop = g->tev[2 * g->tev_used + 1];
    g->tev_used++;
    return true;
}

int LaunchTimer::collect(ptg_stats *st) {
    if (!st) return PTG_OK;
    double ns = 0;
    for (uint32_t i = 0; i < g->tev_used; i++) {
        float ms = 0;
        PTG_HIP(hipEventElapsedTime(&ms, g->tev[2 * i], g->tev[2 * i + 1]));
        ns += (double)ms * 1e6;
    }
    st->main_kernel_ns = (uint64_t)ns;
    st->main_launches = g->tev_used - std::min(riders, g->tev_used);
    return PTG_OK;
}

static void destroy(ptg_graph *g) {
    if (!g) return;
    if (g->stream) (void)hipStreamSynchronize(g->stream);
    gfree(g, g->out.off);
    gfree(g, g->out.col);
    gfree(g, g->d_w);
    free_static_bins(g, &g->out.bins);
    if (g->have_in && !g->in_is_out) {
        gfree(g, g->in.off);
        gfree(g, g->in.col);
        free_static_bins(g, &g->in.bins);
    }
    gfree(g, g->d_val);
    gfree(g, g->d_val2);
    gfree(g, g->d_perm);
    gfree(g, g->d_inv);
    for (auto *b : g->d_bits) gfree(g, b);
    for (auto *f : g->d_flags) gfree(g, f);
    gfree(g, g->d_far);
    gfree(g, g->d_qmem);
    gfree(g, g->d_ring);
    gfree(g, g->d_acc);
    for (auto *c : g->d_contrib) gfree(g, c);
    destroy_batch(g);
    destroy_streamed(g);
    release_sssp_split(g);
    destroy_pr_bins(g);
    gfree(g, g->d_part);
    gfree(g, g->d_lpart);
    if (g->h_ring) (void)hipHostFree(g->h_ring);
    if (g->h_acc) (void)hipHostFree(g->h_acc);
    if (g->ev0) (void)hipEventDestroy(g->ev0);
    if (g->ev1) (void)hipEventDestroy(g->ev1);
    for (auto e : g->tev) (void)hipEventDestroy(e);
    if (g->stream) (void)hipStreamDestroy(g->stream);
    delete g;
}

static int create_impl(uint32_t V, uint64_t E, const uint32_t *ptr, const uint32_t *cols,
                       const uint32_t *weights, const EdgeWithWeight *wedges, const ptg_config *cfg,
                       ptg_graph **out) {
    if (!out) return fail(PTG_ERR_INVALID_ARG, "graph_create: out is null");
    *out = nullptr;
    ptg_config c;
    if (cfg)
        c = *cfg;
    else
        ptg_config_default(&c);
    if (c.small_degree == 0 || c.large_degree < c.small_degree || c.large_degree < 64)
        return fail(PTG_ERR_INVALID_ARG, "graph_create: need 0 < small_degree <= large_degree, large_degree >= 64");
    if (c.sssp_sentinel == 0) c.sssp_sentinel = 0xFFFFFFFEu;
    if (c.sssp_pull != c.sssp_pull || c.batch_pull != c.batch_pull || c.sssp_sweep != c.sssp_sweep)
        return fail(PTG_ERR_INVALID_ARG, "graph_create: NaN sssp_pull, batch_pull or sssp_sweep");
    if (c.pr_block && c.pr_block != 256 && c.pr_block != 512 && c.pr_block != 1024)
        return fail(PTG_ERR_INVALID_ARG, "graph_create: pr_block is 0, 256, 512 or 1024");
    if (c.pr_layout > PTG_PR_BINNED) return fail(PTG_ERR_INVALID_ARG, "graph_create: pr_layout is 0, 1 or 2");
    if (!(c.flags & PTG_FLAG_NO_VALIDATE)) {
        int rc = ptg_validate_csr(V, E, ptr, cols, wedges);
        if (rc) return rc;
    } else if (V == 0 || !ptr || E > 0xFFFFFFFFull || (E && !cols && !wedges)) {
        return fail(PTG_ERR_INVALID_ARG, "graph_create: bad sizes or null arrays");
    }
    const bool weighted = (weights != nullptr) || (wedges != nullptr);
    const bool streamed = (c.flags & PTG_FLAG_STREAM) != 0;
    const uint64_t need = ptg_graph_estimate_bytes(V, E, weighted, &c);
    if (!streamed && c.device_budget_bytes && need > c.device_budget_bytes)
        return fail(PTG_ERR_BUDGET, "graph needs " + std::to_string(need) + " device bytes, budget is " +
                                        std::to_string(c.device_budget_bytes));
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) {
        (void)hipGetLastError();
        return fail(PTG_ERR_NO_DEVICE, "no HIP device available");
    }
    ptg_graph *g = new ptg_graph();
    g->V = V;
    g->E = E;
    g->cfg = c;
    g->weighted = weighted;
    int rc = PTG_OK;
#define TRY(x)                 \
    do {                       \
        rc = (x);              \
        if (rc) {              \
            destroy(g);        \
            return rc;         \
        }                      \
    } while (0)
    TRY(hipGetDevice(&g->device) == hipSuccess ? PTG_OK : fail(PTG_ERR_NO_DEVICE, "hipGetDevice failed"));
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, g->device) != hipSuccess) TRY(fail(PTG_ERR_HIP, "hipGetDeviceProperties"));
    g->num_cus = prop.multiProcessorCount;
    g->grid = c.grid_blocks ? c.grid_blocks : (uint32_t)g->num_cus * 8u;
    if (hipStreamCreateWithFlags(&g->stream, hipStreamNonBlocking) != hipSuccess)
        TRY(fail(PTG_ERR_HIP, "hipStreamCreateWithFlags"));
    if (hipEventCreate(&g->ev0) != hipSuccess || hipEventCreate(&g->ev1) != hipSuccess)
        TRY(fail(PTG_ERR_HIP, "hipEventCreate"));
    if (streamed) {
        TRY(create_streamed(g, ptr, cols, weights, wedges));
        g->create_bytes = g->device_bytes;
        *out = g;
        return PTG_OK;
    }

    // CSR in HBM: V+1 offsets, relabeled by degree unless asked not to.
    std::vector<uint32_t> h_off;
    Relabeled rl;
    const bool do_relabel = !(c.flags & PTG_FLAG_NO_RELABEL);
    const uint32_t *up_col = cols, *up_w = weights;
    if (do_relabel) {
        relabel(V, E, ptr, cols, weights, wedges, c.order_hub_degree, &rl);
        h_off.swap(rl.off);
        up_col = rl.col.data();
        up_w = g->weighted ? rl.w.data() : nullptr;
        g->relabeled = true;
    } else {
        h_off.resize((size_t)V + 1);
        memcpy(h_off.data(), ptr, sizeof(uint32_t) * V);
        h_off[V] = (uint32_t)E;
    }
    TRY(galloc(g, (void **)&g->out.off, sizeof(uint32_t) * ((uint64_t)V + 1)));
    TRY(galloc(g, (void **)&g->out.col, sizeof(uint32_t) * E));
    if (g->weighted) TRY(galloc(g, (void **)&g->d_w, sizeof(uint32_t) * E));
    TRY(copy_sync(g, g->out.off, h_off.data(), sizeof(uint32_t) * ((uint64_t)V + 1), hipMemcpyHostToDevice));
    if (up_col) {
        if (E) TRY(copy_sync(g, g->out.col, up_col, sizeof(uint32_t) * E, hipMemcpyHostToDevice));
        if (up_w && E) TRY(copy_sync(g, g->d_w, up_w, sizeof(uint32_t) * E, hipMemcpyHostToDevice));
    } else if (E) {
        // AoS EdgeWithWeight -> SoA columns + weights, in host blocks.
        const uint64_t blk = 1ull << 22;
        std::vector<uint32_t> tc(std::min<uint64_t>(blk, E)), tw(std::min<uint64_t>(blk, E));
        for (uint64_t e0 = 0; e0 < E; e0 += blk) {
            uint64_t n = std::min(blk, E - e0);
            for (uint64_t i = 0; i < n; i++) {
                tc[i] = wedges[e0 + i].toNode;
                tw[i] = wedges[e0 + i].weight;
            }
            TRY(copy_sync(g, g->out.col + e0, tc.data(), n * 4, hipMemcpyHostToDevice));
            TRY(copy_sync(g, g->d_w + e0, tw.data(), n * 4, hipMemcpyHostToDevice));
        }
    }
    if (do_relabel) {
        TRY(galloc(g, (void **)&g->d_perm, sizeof(uint32_t) * V));
        TRY(galloc(g, (void **)&g->d_inv, sizeof(uint32_t) * V));
        TRY(galloc(g, (void **)&g->d_val2, sizeof(uint32_t) * V));
        TRY(copy_sync(g, g->d_perm, rl.perm.data(), sizeof(uint32_t) * V, hipMemcpyHostToDevice));
        TRY(copy_sync(g, g->d_inv, rl.inv.data(), sizeof(uint32_t) * V, hipMemcpyHostToDevice));
        g->h_perm.swap(rl.perm);
    }
    TRY(build_static_bins(g, h_off.data(), V, &g->out.bins));

    // Query state.
    g->npairs = (V + 63) / 64;
    g->nwords = 2 * g->npairs;
    TRY(galloc(g, (void **)&g->d_val, sizeof(uint32_t) * V));
    for (auto &b : g->d_bits) TRY(galloc(g, (void **)&b, sizeof(uint32_t) * g->nwords));
    for (auto &f : g->d_flags) TRY(galloc(g, (void **)&f, 64ull * g->npairs));
    TRY(galloc(g, (void **)&g->d_far, 64ull * g->npairs));
    // per buffer: sharded S queue and sharded chunk queue (ids + chunk bases)
    g->cq_cap = chunk_queue_cap(V, g->grid);
    const uint64_t qwords = (uint64_t)kShards * kQueueCap;
    const uint64_t cwords = (uint64_t)kShards * g->cq_cap;
    const uint64_t per_set = qwords + 2 * cwords;
    TRY(galloc(g, (void **)&g->d_qmem, sizeof(uint32_t) * 2 * per_set));
    for (int i = 0; i < 2; i++) {
        uint32_t *p = g->d_qmem + i * per_set;
        g->d_queue[i] = p;
        g->d_cq[i] = p + qwords;
        g->d_cb[i] = p + qwords + cwords;
    }
    TRY(galloc(g, (void **)&g->d_ring, sizeof(RoundCtr) * kRing));
    TRY(galloc(g, (void **)&g->d_acc, sizeof(unsigned long long) * 2 * kShards));
    if (hipHostMalloc((void **)&g->h_ring, sizeof(RoundCtr) * kRing, hipHostMallocDefault) != hipSuccess ||
        hipHostMalloc((void **)&g->h_acc, sizeof(unsigned long long) * 2 * kShards, hipHostMallocDefault) !=
            hipSuccess)
        TRY(fail(PTG_ERR_OUT_OF_MEMORY, "hipHostMalloc(counters)"));
    {
        int sym = 0;
        TRY(check_symmetric(g, &sym));
    }
#undef TRY
    g->create_bytes = g->device_bytes;
    *out = g;
    return PTG_OK;
}

}  // namespace ptg

using namespace ptg;

// Runs `fn` (which writes a V-entry result in device ids into the buffer it
// is given) and delivers the result in the caller's ids to `user`, which may
// be device or host memory.
template <class T, class F>
static int with_output(ptg_graph *g, T *user, F fn) {
    if (!user) return fail(PTG_ERR_INVALID_ARG, "null output pointer");
    const bool dev = is_device_pointer(user);
    uint32_t *target = reinterpret_cast<uint32_t *>(user);
    if (!g->relabeled && dev) return fn(reinterpret_cast<T *>(target));
    // computed into d_val, then unpermuted into the caller's ids and/or copied
    // to the caller's host buffer
    auto deliver = [g, user, dev, target]() -> int {
        const uint32_t *src = g->d_val;
        if (g->relabeled) {
            uint32_t *dst = dev ? target : g->d_val2;
            int rc = unpermute(g, g->d_val, dst);
            if (rc || dev) return rc;
            src = dst;
        }
        PTG_HIP(hipMemcpyAsync(user, src, (uint64_t)g->V * sizeof(T), hipMemcpyDeviceToHost, g->stream));
        return PTG_OK;
    };
    g->deliver = deliver;
    g->deliver_on_device = dev;
    g->delivered = false;
    int rc = fn(reinterpret_cast<T *>(g->d_val));
    const bool done = g->delivered;
    g->deliver = nullptr;
    g->delivered = false;
    if (rc) return rc;
    if (done) return PTG_OK;  // the query enqueued it and waited
    if ((rc = deliver())) return rc;
    PTG_HIP(hipStreamSynchronize(g->stream));
    return PTG_OK;
}

static uint32_t map_source(const ptg_graph *g, uint32_t src) { return g->relabeled ? g->h_perm[src] : src; }

extern "C" {

const char *ptg_strerror(int code) {
    switch (code) {
        case PTG_OK: return "success";
        case PTG_ERR_NO_DEVICE: return "no usable HIP device";
        case PTG_ERR_INVALID_ARG: return "invalid argument";
        case PTG_ERR_INVALID_GRAPH: return "invalid CSR graph";
        case PTG_ERR_OUT_OF_MEMORY: return "out of memory";
        case PTG_ERR_BUDGET: return "device memory budget exceeded";
        case PTG_ERR_HIP: return "HIP runtime error";
        case PTG_ERR_UNSUPPORTED: return "unsupported operation";
        case PTG_ERR_REJECTED: return "request rejected by admission control";
        case PTG_ERR_IO: return "file I/O error";
        default: return "unknown error";
    }
}

const char *ptg_last_error(void) { return g_last_error.c_str(); }

void ptg_config_default(ptg_config *c) {
    if (!c) return;
    memset(c, 0, sizeof *c);
    c->device_budget_bytes = 0;
    c->chunk_bytes = 256ull << 20;
    c->sssp_sentinel = 0xFFFFFFFEu;  // ssspCaculateInShare, ptGraph/main.cu:584
    c->grid_blocks = 0;
    c->small_degree = 32;
    c->large_degree = 4096;
    c->td_to_bu = 14.0f;
    c->bu_to_td = 24.0f;
    c->sssp_delta = 0;
    c->flags = 0;
    // hubs first, then breadth-first order: measured +13% BFS and +2%
    // PageRank against pure degree order on R-MAT scale 22
    c->order_hub_degree = 16;
}

int ptg_init(int device) {
    int n = 0;
    hipError_t e = hipGetDeviceCount(&n);
    if (e != hipSuccess || n == 0) {
        (void)hipGetLastError();
        return fail(PTG_ERR_NO_DEVICE, "no HIP device available");
    }
    if (device < 0 || device >= n) return fail(PTG_ERR_INVALID_ARG, "device index out of range");
    PTG_HIP(hipSetDevice(device));
    return pr_scratch_setup(device);
}

int ptg_device_info(int *num_cus, uint64_t *total_mem, uint64_t *free_mem) {
    int dev = 0;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n == 0) {
        (void)hipGetLastError();
        return fail(PTG_ERR_NO_DEVICE, "no HIP device available");
    }
    PTG_HIP(hipGetDevice(&dev));
    hipDeviceProp_t p;
    PTG_HIP(hipGetDeviceProperties(&p, dev));
    size_t fr = 0, tot = 0;
    PTG_HIP(hipMemGetInfo(&fr, &tot));
    if (num_cus) *num_cus = p.multiProcessorCount;
    if (total_mem) *total_mem = tot;
    if (free_mem) *free_mem = fr;
    return PTG_OK;
}

int ptg_device_alloc(uint64_t bytes, void **ptr) {
    if (!ptr) return fail(PTG_ERR_INVALID_ARG, "device_alloc: null ptr");
    PTG_HIP(hipMalloc(ptr, bytes ? bytes : 1));
    return PTG_OK;
}

int ptg_device_free(void *ptr) {
    if (ptr) PTG_HIP(hipFree(ptr));
    return PTG_OK;
}

int ptg_memcpy_d2h(void *dst, const void *src, uint64_t bytes) {
    hipStream_t st;
    if (int rc = util_stream(&st)) return rc;
    PTG_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, st));
    PTG_HIP(hipStreamSynchronize(st));
    return PTG_OK;
}

int ptg_memcpy_h2d(void *dst, const void *src, uint64_t bytes) {
    hipStream_t st;
    if (int rc = util_stream(&st)) return rc;
    PTG_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, st));
    PTG_HIP(hipStreamSynchronize(st));
    return PTG_OK;
}

int ptg_graph_stream_info(const ptg_graph *g, uint32_t *chunks, uint32_t *slots, uint64_t *slot_bytes) {
    if (!g) return fail(PTG_ERR_INVALID_ARG, "null graph");
    if (!g->streamed) return fail(PTG_ERR_UNSUPPORTED, "not a streamed graph");
    return stream_info(g, chunks, slots, slot_bytes);
}

int ptg_host_link_probe(uint64_t bytes, int reps, double *gbs) {
    if (!gbs || bytes == 0 || reps < 1) return fail(PTG_ERR_INVALID_ARG, "host_link_probe: bad arguments");
    return host_link_probe(bytes, reps, gbs);
}

int ptg_memcpy_d2d(void *dst, const void *src, uint64_t bytes) {
    if (bytes == 0) return PTG_OK;
    if (!dst || !src) return fail(PTG_ERR_INVALID_ARG, "null pointer");
    if (!is_device_pointer(dst) || !is_device_pointer(src))
        return fail(PTG_ERR_INVALID_ARG, "ptg_memcpy_d2d needs two device pointers");
    hipStream_t st;
    int rc = util_stream(&st);
    if (rc || (rc = copy_bytes(dst, src, bytes, st))) return rc;
    PTG_HIP(hipStreamSynchronize(st));
    return PTG_OK;
}

int ptg_graph_create(uint32_t V, uint64_t E, const uint32_t *nodePointers, const uint32_t *edgeList,
                     const EdgeWithWeight *wedges, const ptg_config *cfg, ptg_graph **out) {
    if (edgeList && wedges)
        return fail(PTG_ERR_INVALID_ARG, "graph_create: pass edgeList or wedges, not both");
    return create_impl(V, E, nodePointers, edgeList, nullptr, wedges, cfg, out);
}

int ptg_graph_create_soa(uint32_t V, uint64_t E, const uint32_t *nodePointers, const uint32_t *cols,
                         const uint32_t *weights, const ptg_config *cfg, ptg_graph **out) {
    if (E && !cols) return fail(PTG_ERR_INVALID_ARG, "graph_create_soa: null cols");
    static const uint32_t dummy = 0;
    return create_impl(V, E, nodePointers, cols ? cols : &dummy, weights, nullptr, cfg, out);
}

void ptg_graph_destroy(ptg_graph *g) { destroy(g); }

uint64_t ptg_graph_estimate_bytes(uint32_t V, uint64_t E, int weighted, const ptg_config *cfg) {
    ptg_config c;
    if (cfg)
        c = *cfg;
    else
        ptg_config_default(&c);
    const uint64_t v1 = (uint64_t)V + 1;
    const uint64_t large = c.large_degree ? c.large_degree : 4096;
    const uint64_t lcap = std::min<uint64_t>(V, E / (large + 1) + 1);
    const uint64_t nwords = 2ull * ((V + 63) / 64);
    const uint32_t grid = c.grid_blocks ? c.grid_blocks : 2048u;
    auto al = [](uint64_t b) { return (b + 255) & ~255ull; };
    uint64_t b = al(4 * v1) + al(4 * E) + (weighted ? al(4 * E) : 0);  // CSR
    b += al(4 * (uint64_t)V + 4 * (2 * lcap) + 64);                    // static bins
    b += al(4 * (uint64_t)V) + 5 * al(4 * nwords) + 3 * al(32 * nwords); // values, bitmaps, byte flags
    if (!(c.flags & PTG_FLAG_NO_RELABEL)) b += 3 * al(4 * (uint64_t)V); // permutation, inverse, results
    b += al(4 * 2 * ((uint64_t)kShards * kQueueCap + 2ull * kShards * chunk_queue_cap(V, grid)));  // queues
    b += al(sizeof(RoundCtr) * kRing) + al(16 * kShards);              // counter ring, accumulators
    return b;
}

uint64_t ptg_graph_device_bytes(const ptg_graph *g) { return g ? g->device_bytes : 0; }

int ptg_graph_set_config(ptg_graph *g, const ptg_config *c) {
    if (!g || !c) return fail(PTG_ERR_INVALID_ARG, "null graph or config");
    const ptg_config &o = g->cfg;
    const uint64_t o_chunk = o.chunk_bytes;
    if (c->device_budget_bytes != o.device_budget_bytes || (c->chunk_bytes != o.chunk_bytes && !g->streamed) ||
        c->sssp_sentinel != o.sssp_sentinel || c->grid_blocks != o.grid_blocks || c->small_degree != o.small_degree ||
        c->large_degree != o.large_degree || c->order_hub_degree != o.order_hub_degree ||
        ((c->flags ^ o.flags) & ~PTG_FLAG_TIMING))
        return fail(PTG_ERR_INVALID_ARG, "only tuning fields can change after graph creation");
    if (!(c->td_to_bu >= 0) || !(c->bu_to_td >= 0) || !(c->batch_td_to_bu >= 0))
        return fail(PTG_ERR_INVALID_ARG, "negative or NaN switch threshold");
    if (c->sssp_pull != c->sssp_pull || c->batch_pull != c->batch_pull || c->sssp_sweep != c->sssp_sweep)
        return fail(PTG_ERR_INVALID_ARG, "NaN sssp_pull, batch_pull or sssp_sweep");
    if (c->pr_block && c->pr_block != 256 && c->pr_block != 512 && c->pr_block != 1024)
        return fail(PTG_ERR_INVALID_ARG, "pr_block is 0, 256, 512 or 1024");
    if (c->pr_layout > PTG_PR_BINNED) return fail(PTG_ERR_INVALID_ARG, "pr_layout is 0, 1 or 2");
    const bool replan = g->streamed && c->chunk_bytes != o.chunk_bytes;
    g->cfg = *c;
    g->pull_hist[0] = g->pull_hist[1] = ~0ull;  // SSSP pull and sweep rounds are looked for again
    g->sweep_hist[0] = g->sweep_hist[1] = ~0ull;
    g->pull_queries = 0;
    if (replan) {
        // a streamed graph is cut into chunks of the new size; on failure
        // (budget) the previous size is planned again
        const uint64_t want = c->chunk_bytes;
        int rc = stream_plan(g, want);
        if (rc) {
            const std::string why = ptg_last_error();
            g->cfg.chunk_bytes = o_chunk;
            if (stream_plan(g, o_chunk) == PTG_OK) return fail(rc, why);
            return rc;
        }
    }
    return PTG_OK;
}

int ptg_graph_get_config(const ptg_graph *g, ptg_config *c) {
    if (!g || !c) return fail(PTG_ERR_INVALID_ARG, "null graph or config");
    *c = g->cfg;
    return PTG_OK;
}

int ptg_graph_frontier_sizes(const ptg_graph *g, uint64_t *sizes, uint32_t cap, uint32_t *count) {
    if (!g || !count || (cap && !sizes)) return fail(PTG_ERR_INVALID_ARG, "null argument");
    const uint32_t n = (uint32_t)g->frontiers.size();
    for (uint32_t i = 0; i < n && i < cap; i++) sizes[i] = g->frontiers[i];
    *count = n;
    return PTG_OK;
}

int ptg_graph_set_flags(ptg_graph *g, uint32_t flags) {
    if (!g) return fail(PTG_ERR_INVALID_ARG, "null graph");
    if ((flags ^ g->cfg.flags) & ~PTG_FLAG_TIMING)
        return fail(PTG_ERR_INVALID_ARG, "only PTG_FLAG_TIMING can change after graph creation");
    g->cfg.flags = flags;
    return PTG_OK;
}

int ptg_graph_is_symmetric(ptg_graph *g) {
    if (!g) return fail(PTG_ERR_INVALID_ARG, "null graph");
    int s = 0;
    int rc = check_symmetric(g, &s);
    return rc ? rc : s;
}

int ptg_bfs(ptg_graph *g, uint32_t src, uint32_t *levels_out, ptg_stats *st) {
    if (!g) return fail(PTG_ERR_INVALID_ARG, "null graph");
    if (src >= g->V) return fail(PTG_ERR_INVALID_ARG, "bfs: source >= V");
    const uint32_t s = map_source(g, src);
    if (g->streamed) return with_output(g, levels_out, [&](uint32_t *o) { return run_bfs_streamed(g, s, o, st); });
    return with_output(g, levels_out, [&](uint32_t *o) { return run_bfs(g, s, o, st); });
}

int ptg_sssp(ptg_graph *g, uint32_t src, uint32_t *dist_out, ptg_stats *st) {
    if (!g) return fail(PTG_ERR_INVALID_ARG, "null graph");
    if (src >= g->V) return fail(PTG_ERR_INVALID_ARG, "sssp: source >= V");
    if (!g->weighted) return fail(PTG_ERR_UNSUPPORTED, "sssp: graph has no weights");
    const uint32_t s = map_source(g, src);
    if (g->streamed) return with_output(g, dist_out, [&](uint32_t *o) { return run_sssp_streamed(g, s, o, st); });
    return with_output(g, dist_out, [&](uint32_t *o) { return run_sssp(g, s, o, st); });
}

int ptg_cc(ptg_graph *g, uint32_t *labels_out, ptg_stats *st) {
    if (!g) return fail(PTG_ERR_INVALID_ARG, "null graph");
    if (g->streamed) return with_output(g, labels_out, [&](uint32_t *o) { return run_cc_streamed(g, o, st); });
    return with_output(g, labels_out, [&](uint32_t *o) { return run_cc(g, o, st); });
}

int ptg_pagerank(ptg_graph *g, float damping, int iterations, float *ranks_out, ptg_stats *st) {
    if (!g) return fail(PTG_ERR_INVALID_ARG, "null graph");
    if (iterations < 0 || !(damping >= 0.0f && damping <= 1.0f))
        return fail(PTG_ERR_INVALID_ARG, "pagerank: need iterations >= 0 and 0 <= damping <= 1");
    if (g->streamed)
        return with_output(g, ranks_out, [&](float *o) { return run_pagerank_streamed(g, damping, iterations, o, st); });
    return with_output(g, ranks_out, [&](float *o) { return run_pagerank(g, damping, iterations, o, st); });
}

int ptg_pagerank_noalloc(ptg_graph *g, float damping, int iterations, float *ranks_out, ptg_stats *st) {
    if (!g) return fail(PTG_ERR_INVALID_ARG, "null graph");
    if (iterations < 0 || !(damping >= 0.0f && damping <= 1.0f))
        return fail(PTG_ERR_INVALID_ARG, "pagerank_noalloc: need iterations >= 0 and 0 <= damping <= 1");
    if (g->streamed) return fail(PTG_ERR_UNSUPPORTED, "pagerank_noalloc: not available on a streamed graph");
    if (!ranks_out || !is_device_pointer(ranks_out))
        return fail(PTG_ERR_INVALID_ARG, "pagerank_noalloc: the output must be device memory");
    return with_output(g, ranks_out, [&](float *o) { return run_pagerank_noalloc(g, damping, iterations, o, st); });
}

// Shared checks and output delivery of the two batched queries.
static int batch_query(ptg_graph *g, bool sssp, const uint32_t *srcs, int n, uint32_t *out, ptg_stats *st) {
    if (!g) return fail(PTG_ERR_INVALID_ARG, "null graph");
    if (!srcs || !out) return fail(PTG_ERR_INVALID_ARG, "null sources or output");
    if (n < 1 || n > 64) return fail(PTG_ERR_INVALID_ARG, "a batch holds 1 to 64 sources");
    for (int i = 0; i < n; i++)
        if (srcs[i] >= g->V) return fail(PTG_ERR_INVALID_ARG, "source out of range");
    if (g->streamed) return fail(PTG_ERR_UNSUPPORTED, "batched queries are not available on a streamed graph");
    if (sssp && !g->weighted) return fail(PTG_ERR_UNSUPPORTED, "SSSP needs a weighted graph");
    ptg_stats local;
    if (!st) st = &local;
    if (is_device_pointer(out)) return run_batch_query(g, sssp, srcs, (uint32_t)n, out, st);
    uint32_t *stage = nullptr;
    int rc = batch_staging(g, &stage);
    if (rc) return rc;
    if ((rc = run_batch_query(g, sssp, srcs, (uint32_t)n, stage, st))) return rc;
    PTG_HIP(hipMemcpyAsync(out, stage, 4ull * g->V * (uint64_t)n, hipMemcpyDeviceToHost, g->stream));
    PTG_HIP(hipStreamSynchronize(g->stream));
    return PTG_OK;
}

int ptg_bfs_batch(ptg_graph *g, const uint32_t *srcs, int n, uint32_t *levels_out, ptg_stats *st) {
    return batch_query(g, false, srcs, n, levels_out, st);
}

int ptg_sssp_batch(ptg_graph *g, const uint32_t *srcs, int n, uint32_t *dist_out, ptg_stats *st) {
    return batch_query(g, true, srcs, n, dist_out, st);
}

}  // extern "C"
