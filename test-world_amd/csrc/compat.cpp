// Entry points with the names, signatures and return value of ptGraph's
// drivers (ptGraph/main.cu:24-56): each builds the graph on the device, runs
// one query, frees everything and returns the elapsed milliseconds of the
// query (the reference's `startProcessing` .. end placement, main.cu:1084 and
// :1107). Like the reference they discard the result array; ptg_bfs & co.
// return it. On failure they return the negative ptg error code instead of
// calling exit() (main.cu:15-22).
#include <chrono>
#include <vector>

#include "ptg_internal.h"

namespace {

// When the resident graph would not fit 90% of the device's free memory, a
// driver streams the edges from pinned host memory like the reference's
// partition drivers (bfsCaculateInOpt / bfsCaculateInAsyncSwap,
// main.cu:1180-2059; ssspCaculateInOpt :678-889; ccCaculateInOpt :222-440).
// Streamed PageRank needs a graph that is symmetric as a multigraph; on any
// other graph that does not fit, the PageRank drivers return
// PTG_ERR_UNSUPPORTED.
template <class F>
long run_one(uint32_t V, uint32_t E, const uint32_t *ptr, const uint32_t *cols, const EdgeWithWeight *wedges,
             uint32_t sentinel, bool may_stream, F query) {
    ptg_config cfg;
    ptg_config_default(&cfg);
    cfg.sssp_sentinel = sentinel;
    if (may_stream) {
        int cus = 0;
        uint64_t total = 0, free_b = 0;
        const uint64_t need = ptg_graph_estimate_bytes(V, E, wedges ? 1 : 0, &cfg);
        if (ptg_device_info(&cus, &total, &free_b) == PTG_OK && need > free_b / 10 * 9) {
            cfg.flags |= PTG_FLAG_STREAM;
            cfg.device_budget_bytes = free_b / 10 * 9;
        }
    }
    ptg_graph *g = nullptr;
    int rc = ptg_graph_create(V, E, ptr, cols, wedges, &cfg, &g);
    if (rc) return rc;
    std::vector<uint32_t> out(V);
    auto t0 = std::chrono::steady_clock::now();
    rc = query(g, out.data());
    auto t1 = std::chrono::steady_clock::now();
    ptg_graph_destroy(g);
    if (rc) return rc;
    return (long)std::chrono::duration_cast<std::chrono::milliseconds>(t1 - t0).count();
}

}  // namespace

extern "C" {

long bfsCaculateInShare(uint32_t V, uint32_t E, uint32_t *ptr, uint32_t *edgeList, uint32_t src) {
    return run_one(V, E, ptr, edgeList, nullptr, 0xFFFFFFFEu, true,
                   [&](ptg_graph *g, uint32_t *o) { return ptg_bfs(g, src, o, nullptr); });
}

long bfsCaculateInOpt(uint32_t V, uint32_t E, uint32_t *ptr, uint32_t *edgeList, uint32_t src) {
    return bfsCaculateInShare(V, E, ptr, edgeList, src);
}

long bfsCaculateInAsync(uint32_t V, uint32_t E, uint32_t *ptr, uint32_t *edgeList, uint32_t src) {
    return bfsCaculateInShare(V, E, ptr, edgeList, src);
}

long bfsCaculateInAsyncSwap(uint32_t V, uint32_t E, uint32_t *ptr, uint32_t *edgeList, uint32_t src) {
    return bfsCaculateInShare(V, E, ptr, edgeList, src);
}

long ssspCaculateInShare(uint32_t V, uint32_t E, uint32_t *ptr, EdgeWithWeight *edgeList, uint32_t src) {
    // unreached = UINT_MAX - 1 (main.cu:584)
    return run_one(V, E, ptr, nullptr, edgeList, 0xFFFFFFFEu, true,
                   [&](ptg_graph *g, uint32_t *o) { return ptg_sssp(g, src, o, nullptr); });
}

long ssspCaculateInOpt(uint32_t V, uint32_t E, uint32_t *ptr, EdgeWithWeight *edgeList, uint32_t src) {
    // unreached = UINT_MAX (main.cu:742)
    return run_one(V, E, ptr, nullptr, edgeList, 0xFFFFFFFFu, true,
                   [&](ptg_graph *g, uint32_t *o) { return ptg_sssp(g, src, o, nullptr); });
}

long ccCaculateInShare(uint32_t V, uint32_t E, uint32_t *ptr, uint32_t *edgeList) {
    return run_one(V, E, ptr, edgeList, nullptr, 0xFFFFFFFEu, true,
                   [&](ptg_graph *g, uint32_t *o) { return ptg_cc(g, o, nullptr); });
}

long ccCaculateInOpt(uint32_t V, uint32_t E, uint32_t *ptr, uint32_t *edgeList) {
    return ccCaculateInShare(V, E, ptr, edgeList);
}

// The reference's PageRank drivers compute nothing; these run the defined
// PageRank (damping 0.85, 20 iterations).
long prCaculateInShare(uint32_t V, uint32_t E, uint32_t *ptr, uint32_t *edgeList) {
    return run_one(V, E, ptr, edgeList, nullptr, 0xFFFFFFFEu, true, [&](ptg_graph *g, uint32_t *o) {
        return ptg_pagerank(g, 0.85f, 20, reinterpret_cast<float *>(o), nullptr);
    });
}

long prCaculateInOpt(uint32_t V, uint32_t E, uint32_t *ptr, uint32_t *edgeList) {
    return prCaculateInShare(V, E, ptr, edgeList);
}

}  // extern "C"
