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

template <class F>
long run_one(uint32_t V, uint32_t E, const uint32_t *ptr, const uint32_t *cols, const EdgeWithWeight *wedges,
             uint32_t sentinel, F query) {
    ptg_config cfg;
    ptg_config_default(&cfg);
    cfg.sssp_sentinel = sentinel;
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
    return run_one(V, E, ptr, edgeList, nullptr, 0xFFFFFFFEu,
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
    return run_one(V, E, ptr, nullptr, edgeList, 0xFFFFFFFEu,
                   [&](ptg_graph *g, uint32_t *o) { return ptg_sssp(g, src, o, nullptr); });
}

long ssspCaculateInOpt(uint32_t V, uint32_t E, uint32_t *ptr, EdgeWithWeight *edgeList, uint32_t src) {
    // unreached = UINT_MAX (main.cu:742)
    return run_one(V, E, ptr, nullptr, edgeList, 0xFFFFFFFFu,
                   [&](ptg_graph *g, uint32_t *o) { return ptg_sssp(g, src, o, nullptr); });
}

long ccCaculateInShare(uint32_t V, uint32_t E, uint32_t *ptr, uint32_t *edgeList) {
    return run_one(V, E, ptr, edgeList, nullptr, 0xFFFFFFFEu,
                   [&](ptg_graph *g, uint32_t *o) { return ptg_cc(g, o, nullptr); });
}

long ccCaculateInOpt(uint32_t V, uint32_t E, uint32_t *ptr, uint32_t *edgeList) {
    return ccCaculateInShare(V, E, ptr, edgeList);
}

// The reference's PageRank drivers compute nothing; these run the defined
// PageRank (damping 0.85, 20 iterations).
long prCaculateInShare(uint32_t V, uint32_t E, uint32_t *ptr, uint32_t *edgeList) {
    return run_one(V, E, ptr, edgeList, nullptr, 0xFFFFFFFEu, [&](ptg_graph *g, uint32_t *o) {
        return ptg_pagerank(g, 0.85f, 20, reinterpret_cast<float *>(o), nullptr);
    });
}

long prCaculateInOpt(uint32_t V, uint32_t E, uint32_t *ptr, uint32_t *edgeList) {
    return prCaculateInShare(V, E, ptr, edgeList);
}

}  // extern "C"
