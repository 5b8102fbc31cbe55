// CPU driver for the sanitizer builds (tests/test_sanitizers.py): runs the
// threaded host code of the library (graph generators, weights, file readers
// and writers, CSR validation, vertex relabeling) and the threaded oracle
// routines on small graphs, checking their results against each other, so
// AddressSanitizer / UndefinedBehaviorSanitizer and ThreadSanitizer see every
// code path with real data. Prints "sanitize ok" and exits 0 on success.
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <string>
#include <vector>

#include "ptg.h"
#include "relabel.h"

namespace ptg {
// The library's error recorder lives in runtime.cpp (HIP); this build has
// only the host files, so failures are recorded here.
static std::string g_msg;
int fail(int code, const std::string &msg) {
    g_msg = msg;
    return code;
}
}  // namespace ptg

extern "C" {
int orc_ref_loop(int algo, uint32_t V, uint64_t E, const uint32_t *ptr, const uint32_t *col, const uint32_t *w,
                 uint32_t src, uint32_t sentinel, int threads, uint32_t *value_out, uint32_t *iters_out,
                 uint64_t *node_sum_out);
int orc_bfs_queue(uint32_t V, uint64_t E, const uint32_t *ptr, const uint32_t *col, uint32_t src, uint32_t *out,
                  uint32_t *levels_out, uint32_t *reached_out, uint64_t *edges_out);
int orc_dijkstra(uint32_t V, uint64_t E, const uint32_t *ptr, const uint32_t *col, const uint32_t *w, uint32_t src,
                 uint32_t sentinel, uint32_t *out);
int orc_cc_reach_min(uint32_t V, uint64_t E, const uint32_t *ptr, const uint32_t *col, uint32_t *out);
int orc_cc_union_find(uint32_t V, uint64_t E, const uint32_t *ptr, const uint32_t *col, uint32_t *out);
int orc_pagerank_f64(uint32_t V, uint64_t E, const uint32_t *ptr, const uint32_t *col, double damping, int iters,
                     double *out);
int orc_pagerank_f64_sym(uint32_t V, uint64_t E, const uint32_t *ptr, const uint32_t *col, double damping, int iters,
                         int threads, double *out);
int orc_certify_paths(uint32_t V, uint64_t E, const uint32_t *ptr, const uint32_t *col, const uint32_t *w,
                      uint32_t src, uint32_t unreached, const uint32_t *val, int threads);
int64_t orc_degree_pass(uint32_t V, uint64_t E, const uint32_t *ptr, uint32_t *degree);
int64_t orc_overload_gather(int threads, uint32_t n_active, const uint32_t *degree, const uint32_t *active_ptr,
                            const uint32_t *ptr, const uint32_t *active, const uint32_t *col, uint32_t *gathered);
}

static int failures = 0;
#define CHECK(c)                                                        \
    do {                                                                \
        if (!(c)) {                                                     \
            fprintf(stderr, "check failed at line %d: %s\n", __LINE__, #c); \
            failures++;                                                 \
        }                                                               \
    } while (0)

static bool same_csr(const ptg_host_csr &a, const ptg_host_csr &b, bool weights) {
    if (a.V != b.V || a.E != b.E) return false;
    if (memcmp(a.offsets, b.offsets, 4ull * (a.V + 1))) return false;
    if (a.E && memcmp(a.cols, b.cols, 4 * a.E)) return false;
    if (weights && a.E && memcmp(a.weights, b.weights, 4 * a.E)) return false;
    return true;
}

static void check_oracle(const ptg_host_csr &g) {
    const uint32_t V = g.V;
    const uint64_t E = g.E;
    const uint32_t *ptr = g.offsets, *col = g.cols, *w = g.weights;
    std::vector<uint32_t> a(V), b(V);
    uint32_t it = 0, lev = 0, reached = 0;
    uint64_t ns = 0, edges = 0;
    for (uint32_t src : {0u, V / 3, V - 1}) {
        CHECK(orc_ref_loop(0, V, E, ptr, col, nullptr, src, 0xFFFFFFFEu, 4, a.data(), &it, &ns) == 0);
        CHECK(orc_bfs_queue(V, E, ptr, col, src, b.data(), &lev, &reached, &edges) == 0);
        CHECK(a == b);
        CHECK(orc_certify_paths(V, E, ptr, col, nullptr, src, 0xFFFFFFFFu, a.data(), 4) == 0);
        CHECK(orc_ref_loop(1, V, E, ptr, col, w, src, 0xFFFFFFFEu, 3, a.data(), &it, &ns) == 0);
        CHECK(orc_dijkstra(V, E, ptr, col, w, src, 0xFFFFFFFEu, b.data()) == 0);
        CHECK(a == b);
        CHECK(orc_certify_paths(V, E, ptr, col, w, src, 0xFFFFFFFEu, a.data(), 4) == 0);
    }
    CHECK(orc_ref_loop(2, V, E, ptr, col, nullptr, 0, 0, 4, a.data(), &it, &ns) == 0);
    CHECK(orc_cc_union_find(V, E, ptr, col, b.data()) == 0);
    CHECK(a == b);
    CHECK(orc_cc_reach_min(V, E, ptr, col, b.data()) == 0);
    CHECK(a == b);
    std::vector<double> p(V), q(V);
    CHECK(orc_pagerank_f64(V, E, ptr, col, 0.85, 10, p.data()) == 0);
    CHECK(orc_pagerank_f64_sym(V, E, ptr, col, 0.85, 10, 4, q.data()) == 0);
    double l1 = 0;
    for (uint32_t v = 0; v < V; v++) l1 += p[v] > q[v] ? p[v] - q[v] : q[v] - p[v];
    CHECK(l1 < 1e-12);
    std::vector<uint32_t> deg(V);
    CHECK(orc_degree_pass(V, E, ptr, deg.data()) >= 0);
    std::vector<uint32_t> active, aptr;
    uint64_t total = 0;
    for (uint32_t v = 0; v < V; v += 3) {
        active.push_back(v);
        aptr.push_back((uint32_t)total);
        total += deg[v];
    }
    std::vector<uint32_t> out(total + 1);
    CHECK(orc_overload_gather(20, (uint32_t)active.size(), deg.data(), aptr.data(), ptr, active.data(), col,
                              out.data()) >= 0);
    for (size_t i = 0; i < active.size(); i++)
        CHECK(memcmp(out.data() + aptr[i], col + ptr[active[i]], 4ull * deg[active[i]]) == 0);
}

int main(int argc, char **argv) {
    const char *dir = argc > 1 ? argv[1] : ".";
    ptg_host_csr g1{}, g2{}, er{}, r{};
    CHECK(ptg_gen_rmat(11, 16, 0.57, 0.19, 0.19, 1, 8, &g1) == PTG_OK);
    CHECK(ptg_gen_rmat(11, 16, 0.57, 0.19, 0.19, 1, 3, &g2) == PTG_OK);
    CHECK(same_csr(g1, g2, false));  // the same graph whatever the thread count
    CHECK(ptg_gen_weights(&g1, 1, 64, 2, 8) == PTG_OK);
    CHECK(ptg_gen_weights(&g2, 1, 64, 2, 2) == PTG_OK);
    CHECK(same_csr(g1, g2, true));
    CHECK(ptg_gen_er(10, 8, 9, 8, &er) == PTG_OK);
    CHECK(ptg_gen_weights(&er, 1, 9, 3, 4) == PTG_OK);
    CHECK(ptg_validate_csr(g1.V, g1.E, g1.offsets, g1.cols, nullptr) == PTG_OK);
    // file formats round trips
    const std::string b = std::string(dir) + "/g.bcsr", bw = std::string(dir) + "/g.bwcsr";
    CHECK(ptg_write_bcsr(b.c_str(), &g1) == PTG_OK);
    CHECK(ptg_read_bcsr(b.c_str(), &r) == PTG_OK);
    CHECK(same_csr(g1, r, false));
    ptg_host_csr_free(&r);
    CHECK(ptg_write_bwcsr(bw.c_str(), &g1) == PTG_OK);
    CHECK(ptg_read_bwcsr(bw.c_str(), &r) == PTG_OK);
    CHECK(same_csr(g1, r, true));
    ptg_host_csr_free(&r);
    CHECK(ptg_read_bcsr((std::string(dir) + "/missing.bcsr").c_str(), &r) == PTG_ERR_IO);
    // relabeling (threaded list rebuild) keeps every degree and edge count
    for (uint32_t hub : {0u, 3u}) {
        ptg::Relabeled rl;
        ptg::relabel(g1.V, g1.E, g1.offsets, g1.cols, g1.weights, nullptr, hub, &rl);
        std::vector<char> seen(g1.V, 0);
        for (uint32_t v = 0; v < g1.V; v++) {
            CHECK(rl.perm[rl.inv[v]] == v);
            seen[rl.perm[v]] = 1;
            CHECK(rl.off[rl.perm[v] + 1] - rl.off[rl.perm[v]] == g1.offsets[v + 1] - g1.offsets[v]);
        }
        for (char c : seen) CHECK(c);
        CHECK(rl.off[g1.V] == g1.E && rl.col.size() == g1.E && rl.w.size() == g1.E);
    }
    check_oracle(g1);
    check_oracle(er);
    ptg_host_csr_free(&g1);
    ptg_host_csr_free(&g2);
    ptg_host_csr_free(&er);
    if (failures) {
        fprintf(stderr, "%d checks failed\n", failures);
        return 1;
    }
    printf("sanitize ok\n");
    return 0;
}
