// Serving loop over resident graphs: continuous batching, memory admission,
// latency accounting.
//
// One worker thread per graph issues that graph's launches on the graph's own
// stream, so launches on different graphs overlap on the device. A worker
// whose slot is free forms its next launch from the oldest pending request:
// BFS / SSSP requests of that kind are gathered (in arrival order, up to
// max_batch) into one batched launch; identical CC / PageRank requests share
// one run when coalescing is on. Before launching, the worker reserves the
// query-state bytes the launch will allocate against the state budget; if
// they do not fit, idle graphs' state is released least recently used first,
// and failing that the worker waits for a running launch to finish. A launch
// that cannot fit even on an otherwise empty budget completes its requests
// with PTG_ERR_REJECTED.
#include <string.h>

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "ptg_internal.h"

namespace ptg {
namespace {

uint64_t now_ns() {
    return (uint64_t)std::chrono::duration_cast<std::chrono::nanoseconds>(
               std::chrono::steady_clock::now().time_since_epoch())
        .count();
}

struct Pending {
    ptg_request r;
    uint64_t submit_ns;
};

struct Slot {
    ptg_graph *g = nullptr;
    uint32_t id = 0;
    std::deque<Pending> q;
    std::thread worker;
    bool running = false;
    uint64_t held = 0;       // query-state bytes resident on this graph
    uint64_t reserved = 0;   // bytes reserved by the running launch
    uint64_t last_used = 0;  // for least-recently-used release
    uint64_t base_bytes = 0; // graph bytes when created (CSR and fixed state)
    bool drained = false;    // state released since the last launch
    bool releasing = false;  // another thread is freeing this graph's state (lock not held)
    uint32_t *row = nullptr;              // V words: single-query results (caller ids)
    unsigned long long *sums = nullptr;   // 64 row checksums (device)
    unsigned long long *h_sums = nullptr; // pinned copy
};

}  // namespace
}  // namespace ptg

using namespace ptg;

struct ptg_server {
    ptg_server_config cfg{};
    std::mutex mu;
    std::condition_variable cv_work;
    std::condition_variable cv_done;
    std::vector<std::unique_ptr<Slot>> slots;
    std::deque<ptg_completion> done;
    std::vector<uint64_t> lat;
    uint64_t submitted = 0, finished = 0, rejected = 0, launches = 0, served = 0;
    uint64_t first_submit = 0, last_done = 0, state_peak = 0, evictions = 0;
    uint32_t inflight = 0;
    bool stop = false;
    bool paused = false;
    std::string first_error;  // first failed launch: graph, kind and the library's message
};

namespace ptg {
namespace {

uint64_t resident(const ptg_server *s) {
    uint64_t t = 0;
    for (auto &p : s->slots) t += p->held + p->reserved;
    return t;
}

// Frees the graph's query state; returns the bytes it still holds. Runs
// without the server lock (hipFree waits for the device); the caller marks
// the slot `releasing` first, so its worker does not launch meanwhile, and
// stores the result in `held` under the lock.
uint64_t release_state(Slot *sl) {
    destroy_batch(sl->g);
    release_pr_state(sl->g);
    release_sssp_split(sl->g);
    return sl->g->device_bytes - sl->base_bytes;
}

// Releases `victim`'s state with the lock dropped around the frees.
void release_unlocked(ptg_server *s, Slot *victim, std::unique_lock<std::mutex> &lk) {
    const uint64_t before = victim->held;
    victim->releasing = true;
    lk.unlock();
    const uint64_t after = release_state(victim);
    lk.lock();
    victim->held = after;
    victim->releasing = false;
    victim->drained = true;
    if (after < before) s->evictions++;
    s->cv_work.notify_all();
}

// Records completions of `reqs` (lock held).
void complete(ptg_server *s, const std::vector<Pending> &reqs, int status, uint32_t batch, const uint64_t *sums,
              uint64_t start, uint64_t end) {
    for (size_t i = 0; i < reqs.size(); i++) {
        ptg_completion c{};
        c.tag = reqs[i].r.tag;
        c.status = status;
        c.batch_size = batch;
        c.checksum = sums ? sums[i] : 0;
        c.submit_ns = reqs[i].submit_ns;
        c.start_ns = start;
        c.done_ns = end;
        s->done.push_back(c);
        if (status == PTG_OK) s->lat.push_back(end - reqs[i].submit_ns);
        if (status == PTG_ERR_REJECTED) s->rejected++;
        s->finished++;
        s->last_done = std::max(s->last_done, end);
    }
    s->cv_done.notify_all();
}

// Takes the next launch's requests from the slot's queue (lock held).
std::vector<Pending> form_launch(ptg_server *s, Slot *sl) {
    std::vector<Pending> job;
    const ptg_request head = sl->q.front().r;
    const bool batchable = head.kind == PTG_Q_BFS || head.kind == PTG_Q_SSSP;
    auto same = [&](const ptg_request &r) {
        if (r.kind != head.kind) return false;
        if (head.kind == PTG_Q_PAGERANK) return r.iterations == head.iterations && r.damping == head.damping;
        return true;
    };
    uint32_t limit = batchable ? std::max<uint32_t>(1, s->cfg.max_batch) : (s->cfg.coalesce ? ~0u : 1u);
    if (sl->g->streamed) limit = batchable ? 1u : limit;  // a streamed graph has no batched path
    if (batchable && limit > 1) {
        // a batched launch costs about the same for any number of sources:
        // below min_batch pending requests, serve the oldest one alone
        size_t pending = 0;
        for (const Pending &p : sl->q) pending += same(p.r);
        if (pending < s->cfg.min_batch) limit = 1;
    }
    std::deque<Pending> rest;
    while (!sl->q.empty()) {
        Pending p = sl->q.front();
        sl->q.pop_front();
        if (job.size() < limit && same(p.r)) job.push_back(p);
        else rest.push_back(p);
    }
    sl->q.swap(rest);
    return job;
}

// Query-state bytes the launch will allocate (worker's own graph).
uint64_t launch_need(const Slot *sl, uint32_t kind, size_t n) {
    if ((kind == PTG_Q_BFS || kind == PTG_Q_SSSP) && n > 1) return batch_state_need(sl->g, kind == PTG_Q_SSSP);
    if (kind == PTG_Q_PAGERANK) return pr_state_need(sl->g);
    if (kind == PTG_Q_SSSP) return sssp_split_need(sl->g);
    return 0;
}

// Runs one launch and delivers its results; fills per-request checksums.
int run_launch(Slot *sl, const std::vector<Pending> &job, std::vector<uint64_t> *sums) {
    ptg_graph *g = sl->g;
    const uint64_t V = g->V;
    const ptg_request &h = job.front().r;
    const size_t n = job.size();
    ptg_stats st;
    int rc;
    const uint32_t *rows;
    uint32_t nrows;
    if ((h.kind == PTG_Q_BFS || h.kind == PTG_Q_SSSP) && n > 1) {
        uint32_t *stage = nullptr;
        if ((rc = batch_staging(g, &stage))) return rc;
        std::vector<uint32_t> srcs(n);
        for (size_t i = 0; i < n; i++) srcs[i] = job[i].r.source;
        rc = h.kind == PTG_Q_BFS ? ptg_bfs_batch(g, srcs.data(), (int)n, stage, &st)
                                 : ptg_sssp_batch(g, srcs.data(), (int)n, stage, &st);
        if (rc) return rc;
        rows = stage;
        nrows = (uint32_t)n;
    } else {
        switch (h.kind) {
            case PTG_Q_BFS: rc = ptg_bfs(g, h.source, sl->row, &st); break;
            case PTG_Q_SSSP: rc = ptg_sssp(g, h.source, sl->row, &st); break;
            case PTG_Q_CC: rc = ptg_cc(g, sl->row, &st); break;
            default: rc = ptg_pagerank(g, h.damping, (int)h.iterations, reinterpret_cast<float *>(sl->row), &st);
        }
        if (rc) return rc;
        rows = sl->row;
        nrows = 1;
    }
    if ((rc = row_checksums(g, g->stream, rows, nrows, sl->sums))) return rc;
    PTG_HIP(hipMemcpyAsync(sl->h_sums, sl->sums, 8ull * nrows, hipMemcpyDeviceToHost, g->stream));
    for (size_t i = 0; i < n; i++) {
        void *out = job[i].r.out;
        if (!out) continue;
        const uint32_t *src = rows + (nrows > 1 ? i * V : 0);
        if (is_device_pointer(out)) {
            if ((rc = copy_bytes(out, src, 4 * V, g->stream))) return rc;
        } else {
            PTG_HIP(hipMemcpyAsync(out, src, 4 * V, hipMemcpyDeviceToHost, g->stream));
        }
    }
    PTG_HIP(hipStreamSynchronize(g->stream));
    sums->resize(n);
    for (size_t i = 0; i < n; i++) (*sums)[i] = sl->h_sums[nrows > 1 ? i : 0];
    return PTG_OK;
}

void worker_loop(ptg_server *s, Slot *sl) {
    (void)hipSetDevice(sl->g->device);
    std::unique_lock<std::mutex> lk(s->mu);
    for (;;) {
        s->cv_work.wait(lk, [&] {
            return s->stop || (!s->paused && !sl->q.empty() && !sl->releasing &&
                               (!s->cfg.max_inflight || s->inflight < s->cfg.max_inflight));
        });
        if (s->stop) return;
        std::vector<Pending> job = form_launch(s, sl);
        const uint32_t kind = job.front().r.kind;
        uint64_t need = 0;
        // memory admission; `need` is recomputed every time round, since
        // another worker may have released this graph's state meanwhile
        const uint64_t budget = s->cfg.state_budget_bytes;
        bool rejected = false;
        for (;;) {
            if (sl->releasing) {  // another worker is freeing this graph's state
                s->cv_work.wait(lk);
                if (s->stop) return;
                continue;
            }
            need = launch_need(sl, kind, job.size());
            if (!budget || resident(s) + need <= budget) break;
            // idle graphs that still hold releasable state, least recently used first
            Slot *victim = nullptr;
            for (auto &p : s->slots)
                if (p.get() != sl && !p->running && !p->drained && !p->releasing &&
                    (!victim || p->last_used < victim->last_used))
                    victim = p.get();
            if (!victim && !sl->drained) victim = sl;  // this graph's own state of other kinds
            if (victim) {
                release_unlocked(s, victim, lk);
                if (s->stop) return;
                continue;
            }
            bool others_busy = false;
            for (auto &p : s->slots) others_busy = others_busy || p->running || p->releasing;
            if (!others_busy) {
                rejected = true;
                break;
            }
            s->cv_work.wait(lk);
            if (s->stop) return;
        }
        if (rejected) {
            const uint64_t t = now_ns();
            complete(s, job, PTG_ERR_REJECTED, (uint32_t)job.size(), nullptr, t, t);
            continue;
        }
        sl->running = true;
        sl->drained = false;
        sl->reserved = need;
        s->inflight++;
        s->launches++;
        s->served += job.size();
        s->state_peak = std::max(s->state_peak, resident(s));
        lk.unlock();
        const uint64_t start = now_ns();
        std::vector<uint64_t> sums;
        const int rc = run_launch(sl, job, &sums);
        const uint64_t end = now_ns();
        lk.lock();
        if (rc && s->first_error.empty())
            s->first_error = "graph " + std::to_string(sl->id) + " kind " + std::to_string(job.front().r.kind) +
                             " batch " + std::to_string(job.size()) + ": " + ptg_last_error();
        sl->running = false;
        sl->reserved = 0;
        sl->held = sl->g->device_bytes - sl->base_bytes;
        sl->last_used = end;
        s->inflight--;
        s->state_peak = std::max(s->state_peak, resident(s));
        complete(s, job, rc, (uint32_t)job.size(), rc ? nullptr : sums.data(), start, end);
        s->cv_work.notify_all();
    }
}

}  // namespace
}  // namespace ptg

extern "C" {

int ptg_server_create(const ptg_server_config *cfg, ptg_server **out) {
    if (!out) return fail(PTG_ERR_INVALID_ARG, "null output");
    *out = nullptr;
    ptg_server_config c{};
    if (cfg) c = *cfg;
    if (c.max_batch == 0) c.max_batch = 64;
    if (c.max_batch > 64) return fail(PTG_ERR_INVALID_ARG, "max_batch above 64");
    if (c.queue_cap == 0) c.queue_cap = 1u << 20;
    if (c.min_batch == 0) c.min_batch = 8;
    ptg_server *s = new ptg_server();
    s->cfg = c;
    *out = s;
    return PTG_OK;
}

int ptg_server_add_graph(ptg_server *s, ptg_graph *g, uint32_t *id) {
    if (!s || !g) return fail(PTG_ERR_INVALID_ARG, "null server or graph");
    auto sl = std::make_unique<Slot>();
    sl->g = g;
    PTG_HIP(hipSetDevice(g->device));
    // bytes allocated after creation (batch / PageRank state of earlier
    // queries) stay resident and count as held query state
    sl->base_bytes = g->create_bytes;
    if (hipMalloc((void **)&sl->row, 4ull * std::max<uint32_t>(g->V, 1)) != hipSuccess ||
        hipMalloc((void **)&sl->sums, 8 * 64) != hipSuccess ||
        hipHostMalloc((void **)&sl->h_sums, 8 * 64, hipHostMallocDefault) != hipSuccess) {
        if (sl->row) (void)hipFree(sl->row);
        if (sl->sums) (void)hipFree(sl->sums);
        return fail(PTG_ERR_OUT_OF_MEMORY, "server buffers");
    }
    std::lock_guard<std::mutex> lk(s->mu);
    sl->held = g->device_bytes - sl->base_bytes;
    sl->id = (uint32_t)s->slots.size();
    if (id) *id = sl->id;
    Slot *raw = sl.get();
    s->slots.push_back(std::move(sl));
    raw->worker = std::thread(worker_loop, s, raw);
    return PTG_OK;
}

int ptg_server_submit(ptg_server *s, const ptg_request *r) {
    if (!s || !r) return fail(PTG_ERR_INVALID_ARG, "null server or request");
    std::lock_guard<std::mutex> lk(s->mu);
    if (r->graph >= s->slots.size()) return fail(PTG_ERR_INVALID_ARG, "unknown graph id");
    Slot *sl = s->slots[r->graph].get();
    if (r->kind > PTG_Q_PAGERANK) return fail(PTG_ERR_INVALID_ARG, "unknown query kind");
    if ((r->kind == PTG_Q_BFS || r->kind == PTG_Q_SSSP) && r->source >= sl->g->V)
        return fail(PTG_ERR_INVALID_ARG, "source out of range");
    if (r->kind == PTG_Q_SSSP && !sl->g->weighted) return fail(PTG_ERR_UNSUPPORTED, "SSSP needs a weighted graph");
    if (r->kind == PTG_Q_PAGERANK && !(r->damping >= 0.0f && r->damping <= 1.0f))
        return fail(PTG_ERR_INVALID_ARG, "damping outside [0, 1]");
    const uint64_t t = now_ns();
    if (sl->q.size() >= s->cfg.queue_cap) {
        s->rejected++;
        return fail(PTG_ERR_REJECTED, "queue full");
    }
    if (!s->submitted) s->first_submit = t;
    s->submitted++;
    sl->q.push_back(Pending{*r, t});
    s->cv_work.notify_all();
    return PTG_OK;
}

int ptg_server_replay(ptg_server *s, const ptg_request *r, const uint64_t *arrival_ns, uint64_t n) {
    if (!s || (n && (!r || !arrival_ns))) return fail(PTG_ERR_INVALID_ARG, "null argument");
    int first_err = PTG_OK;
    std::thread t([&] {
        const auto t0 = std::chrono::steady_clock::now();
        for (uint64_t i = 0; i < n; i++) {
            std::this_thread::sleep_until(t0 + std::chrono::nanoseconds(arrival_ns[i]));
            const int rc = ptg_server_submit(s, &r[i]);
            if (rc && rc != PTG_ERR_REJECTED && !first_err) first_err = rc;
        }
    });
    t.join();
    return first_err;
}

int ptg_server_poll(ptg_server *s, ptg_completion *out, int max, int timeout_ms) {
    if (!s || (max > 0 && !out)) return fail(PTG_ERR_INVALID_ARG, "null argument");
    std::unique_lock<std::mutex> lk(s->mu);
    if (s->done.empty() && timeout_ms > 0)
        s->cv_done.wait_for(lk, std::chrono::milliseconds(timeout_ms), [&] { return !s->done.empty(); });
    int k = 0;
    while (k < max && !s->done.empty()) {
        out[k++] = s->done.front();
        s->done.pop_front();
    }
    return k;
}

const char *ptg_server_error(ptg_server *s) {
    if (!s) return "";
    std::lock_guard<std::mutex> lk(s->mu);
    return s->first_error.c_str();
}

int ptg_server_pause(ptg_server *s, int paused) {
    if (!s) return fail(PTG_ERR_INVALID_ARG, "null server");
    {
        std::lock_guard<std::mutex> lk(s->mu);
        s->paused = paused != 0;
    }
    s->cv_work.notify_all();
    return PTG_OK;
}

int ptg_server_drain(ptg_server *s) {
    if (!s) return fail(PTG_ERR_INVALID_ARG, "null server");
    std::unique_lock<std::mutex> lk(s->mu);
    if (s->paused) return fail(PTG_ERR_INVALID_ARG, "drain on a paused server would not return");
    s->cv_done.wait(lk, [&] { return s->finished >= s->submitted; });
    return PTG_OK;
}

int ptg_server_summary_get(ptg_server *s, ptg_server_summary *o) {
    if (!s || !o) return fail(PTG_ERR_INVALID_ARG, "null argument");
    std::lock_guard<std::mutex> lk(s->mu);
    memset(o, 0, sizeof *o);
    std::vector<uint64_t> l = s->lat;
    std::sort(l.begin(), l.end());
    auto pct = [&](double p) -> uint64_t {
        if (l.empty()) return 0;
        const size_t i = (size_t)std::min<double>((double)l.size() - 1, p * (double)(l.size() - 1) + 0.5);
        return l[i];
    };
    o->completed = l.size();
    o->rejected = s->rejected;
    o->launches = s->launches;
    o->p50_ns = pct(0.50);
    o->p99_ns = pct(0.99);
    o->max_ns = l.empty() ? 0 : l.back();
    o->mean_batch = s->launches ? (double)s->served / (double)s->launches : 0.0;
    o->first_submit_ns = s->first_submit;
    o->last_done_ns = s->last_done;
    o->state_bytes_peak = s->state_peak;
    o->evictions = s->evictions;
    return PTG_OK;
}

void ptg_server_destroy(ptg_server *s) {
    if (!s) return;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        s->stop = true;
    }
    s->cv_work.notify_all();
    for (auto &p : s->slots)
        if (p->worker.joinable()) p->worker.join();
    for (auto &p : s->slots) {
        if (p->row) (void)hipFree(p->row);
        if (p->sums) (void)hipFree(p->sums);
        if (p->h_sums) (void)hipHostFree(p->h_sums);
    }
    delete s;
}

}  // extern "C"
