// Autotuner: picks the graph's runtime tuning (BFS push/pull switch
// thresholds, SSSP bucket step and deferral, launch width and static chunk
// size of the batched kernels; for a streamed graph the chunk size and the
// number of pinned leading chunks) by timing the graph's own queries.
//
// Coordinate search: each knob in turn is set to every value of its grid with
// the others held, the query that knob affects is timed twice per value (HIP
// events around the whole query, the faster run kept, so a one-time capture or
// chunk-list rebuild is not charged), and the fastest value stays. Results do
// not depend on any knob (every schedule reaches the same fixed point), so
// only device time is compared.
#include <string.h>

#include <algorithm>
#include <vector>

#include "ptg_internal.h"

namespace ptg {

namespace {

enum Query { kBfs, kSssp, kBfsBatch, kSsspBatch, kStream, kPagerank };

struct Tuner {
    ptg_graph *g;
    std::vector<uint32_t> plain;  // device ids, up to 4
    std::vector<uint32_t> batch;  // caller ids, up to 64

    int time_once(Query q, double *ns) {
        ptg_stats st;
        int rc = PTG_OK;
        double t = 0;
        if (q == kStream) {
            // streamed graph: BFS from the sources, then 2 PageRank
            // iterations on a symmetric graph (copies included: the query's
            // stream waits for them)
            for (uint32_t s : plain) {
                if ((rc = run_bfs_streamed(g, s, g->d_val, &st))) return rc;
                t += (double)st.kernel_ns;
            }
            if (g->symmetric == 1) {
                if ((rc = run_pagerank_streamed(g, 0.85f, 2, reinterpret_cast<float *>(g->d_val), &st))) return rc;
                t += (double)st.kernel_ns;
            }
        } else if (q == kPagerank) {
            if ((rc = run_pagerank(g, 0.85f, 3, reinterpret_cast<float *>(g->d_val), &st))) return rc;
            t = (double)st.kernel_ns;
        } else if (q == kBfs || q == kSssp) {
            for (uint32_t s : plain) {
                rc = q == kBfs ? run_bfs(g, s, g->d_val, &st) : run_sssp(g, s, g->d_val, &st);
                if (rc) return rc;
                t += (double)st.kernel_ns;
            }
        } else {
            uint32_t *out = nullptr;
            if ((rc = batch_staging(g, &out))) return rc;
            rc = run_batch_query(g, q == kSsspBatch, batch.data(), (uint32_t)batch.size(), out, &st);
            if (rc) return rc;
            t = (double)st.kernel_ns;
        }
        *ns = t;
        return PTG_OK;
    }

    int time_best(Query q, double *ns) {
        double a = 0, b = 0;
        int rc = time_once(q, &a);
        if (!rc) rc = time_once(q, &b);
        *ns = std::min(a, b);
        return rc;
    }

    // Tries every value of one knob; leaves the fastest one set.
    template <class T>
    int knob(Query q, T *field, const std::vector<T> &values) {
        T best = *field;
        double best_ns = 0;
        int rc = time_best(q, &best_ns);
        if (rc) return rc;
        for (T v : values) {
            if (v == best) continue;
            const T keep = *field;
            *field = v;
            double ns = 0;
            if ((rc = time_best(q, &ns))) {
                *field = keep;
                return rc;
            }
            if (ns < best_ns) {
                best_ns = ns;
                best = v;
            }
            *field = keep;
        }
        *field = best;
        return PTG_OK;
    }
};

}  // namespace

}  // namespace ptg

using namespace ptg;

extern "C" int ptg_autotune(ptg_graph *g, const uint32_t *srcs, int n, uint32_t what, ptg_config *best) {
    if (!g || !srcs || n < 1) return fail(PTG_ERR_INVALID_ARG, "autotune needs a graph and at least one source");
    for (int i = 0; i < n; i++)
        if (srcs[i] >= g->V) return fail(PTG_ERR_INVALID_ARG, "source out of range");
    if (g->streamed ? (what & ~PTG_TUNE_STREAM) != 0 : (what & PTG_TUNE_STREAM) != 0)
        return fail(PTG_ERR_UNSUPPORTED, g->streamed ? "autotune: a streamed graph tunes PTG_TUNE_STREAM only"
                                                     : "autotune: PTG_TUNE_STREAM needs a streamed graph");
    ptg_config &c = g->cfg;
    if (g->streamed) {
        Tuner t{g, {}, {}};
        for (int i = 0; i < n && i < 2; i++) t.plain.push_back(g->relabeled ? g->h_perm[srcs[i]] : srcs[i]);
        int rc = PTG_OK;
        // chunk size: each value re-cuts the chunks (sizes the budget cannot
        // hold two slots of are skipped)
        double best_ns = 0;
        uint64_t best_chunk = c.chunk_bytes;
        if ((rc = t.time_best(kStream, &best_ns))) return rc;
        for (uint64_t mb : {64ull, 128ull, 256ull, 512ull, 1024ull}) {
            const uint64_t cb = mb << 20;
            if (cb == best_chunk) continue;
            c.chunk_bytes = cb;
            if (stream_plan(g, cb) != PTG_OK) continue;
            double ns = 0;
            if ((rc = t.time_best(kStream, &ns))) break;
            if (ns < best_ns) {
                best_ns = ns;
                best_chunk = cb;
            }
        }
        c.chunk_bytes = best_chunk;
        const int prc = stream_plan(g, best_chunk);
        if (!rc) rc = prc;
        uint32_t nsl = 0;
        if (!rc) stream_info(g, nullptr, &nsl, nullptr);
        const uint32_t top = nsl > 2 ? nsl - 2 : 0;
        if (!rc) rc = t.knob(kStream, &c.stream_pinned, std::vector<uint32_t>{0, top / 2, top});
        if (best) *best = c;
        return rc;
    }
    // explicit values for every knob the search moves
    if (c.td_to_bu <= 0) c.td_to_bu = 14.0f;
    if (c.bu_to_td <= 0) c.bu_to_td = 24.0f;
    if (c.batch_td_to_bu <= 0) c.batch_td_to_bu = c.td_to_bu;
    if (!c.batch_grid) c.batch_grid = 4 * g->grid;
    if (!c.batch_sssp_grid) c.batch_sssp_grid = c.batch_grid;
    if (!c.batch_delta) c.batch_delta = 256;
    if (!c.batch_chunk) c.batch_chunk = 128;
    if (c.batch_pull == 0) c.batch_pull = 0.25f;  // the batched SSSP default
    Tuner t{g, {}, {}};
    for (int i = 0; i < n && i < 4; i++) t.plain.push_back(g->relabeled ? g->h_perm[srcs[i]] : srcs[i]);
    for (int i = 0; i < n && i < 64; i++) t.batch.push_back(srcs[i]);
    const uint32_t G = g->grid;
    int rc = PTG_OK;
    if (what & PTG_TUNE_BFS) {
        if (!rc) rc = t.knob(kBfs, &c.td_to_bu, std::vector<float>{4, 8, 14, 24, 40});
        if (!rc) rc = t.knob(kBfs, &c.bu_to_td, std::vector<float>{8, 16, 24, 48, 96, 200});
    }
    if ((what & PTG_TUNE_SSSP) && g->weighted) {
        // 0 = automatic (one unbounded bucket; degree 128; mean weight / 16)
        if (!rc) rc = t.knob(kSssp, &c.sssp_delta, std::vector<uint32_t>{0, 24, 64, 128});
        if (!rc) rc = t.knob(kSssp, &c.sssp_defer_window, std::vector<uint32_t>{0, 1, 3, 4, 6});
        if (!rc) rc = t.knob(kSssp, &c.sssp_defer_degree, std::vector<uint32_t>{0, 32, 64, 256, 0xFFFFFFFFu});
        // 0 = pull when the open in-edges are fewer than the frontier's edges; -1 = never
        if (!rc) rc = t.knob(kSssp, &c.sssp_pull, std::vector<float>{0, 0.5f, 2, -1});
        // 0 = sweep a round whose push and pull would each scan > E / 4; -1 =
        // never (a graph set to never keeps it: its caller keeps the binned
        // layout's memory free)
        if (!rc && c.sssp_pull >= 0 && c.sssp_sweep >= 0)
            rc = t.knob(kSssp, &c.sssp_sweep, std::vector<float>{0, 0.15f, 0.4f, -1});
        if (!rc && !c.sssp_delta) release_sssp_split(g);  // the light/heavy CSRs tried above
        if (!rc && c.sssp_pull < 0) release_sssp_pull(g);
    }
    if (what & PTG_TUNE_BFS_BATCH) {
        if (!rc) rc = t.knob(kBfsBatch, &c.batch_td_to_bu, std::vector<float>{2, 6, 14, 30});
        if (!rc) rc = t.knob(kBfsBatch, &c.batch_grid, std::vector<uint32_t>{G, 2 * G, 4 * G, 8 * G});
    }
    if (what & PTG_TUNE_PAGERANK) {
        // the layout (in-list pull or binned; the search starts from an
        // explicit one so an automatic layout never stays set), then, for the
        // pull, the threads per gather workgroup (its LDS hub stage grows
        // with it)
        if (c.pr_layout == PTG_PR_AUTO) c.pr_layout = PTG_PR_BINNED;
        if (!rc) rc = t.knob(kPagerank, &c.pr_layout, std::vector<uint32_t>{PTG_PR_PULL, PTG_PR_BINNED});
        if (!c.pr_block) c.pr_block = 256;
        if (!rc && c.pr_layout == PTG_PR_PULL)
            rc = t.knob(kPagerank, &c.pr_block, std::vector<uint32_t>{256, 512, 1024});
    }
    if ((what & PTG_TUNE_SSSP_BATCH) && g->weighted) {
        if (!rc) rc = t.knob(kSsspBatch, &c.batch_delta, std::vector<uint32_t>{32, 64, 128, 256, 1024, 4096});
        if (!rc) rc = t.knob(kSsspBatch, &c.batch_sssp_grid, std::vector<uint32_t>{G, 2 * G, 4 * G, 8 * G, 16 * G, 32 * G, 64 * G});
        if (!rc) rc = t.knob(kSsspBatch, &c.batch_chunk, std::vector<uint32_t>{64, 128, 256, 512, 1024});
        // a value's first batch learns its pull rounds, the second (timed
        // too; the faster counts) uses them; -1 = push only
        if (!rc) rc = t.knob(kSsspBatch, &c.batch_pull, std::vector<float>{0.25f, 0.1f, 0.5f, -1});
    }
    if (best) *best = c;
    return rc;
}
