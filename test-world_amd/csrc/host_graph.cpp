// Host-side graph utilities: CSR validation, synthetic R-MAT / Erdos-Renyi
// generation, symmetric weights, and the reference's on-disk formats.
// None of this touches the GPU.
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <atomic>
#include <string>
#include <thread>
#include <vector>

#include "ptg.h"

namespace ptg {
int fail(int code, const std::string &msg);
}

namespace {

inline uint64_t splitmix64(uint64_t x) {
    // SplitMix64 (Steele, Lea, Flood 2014), used as a counter-based stream.
    x += 0x9E3779B97F4A7C15ull;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
    return x ^ (x >> 31);
}

int clamp_threads(int t) {
    if (t <= 0) {
        unsigned h = std::thread::hardware_concurrency();
        t = h ? (int)h : 1;
    }
    return std::min(t, 64);
}

template <class F>
void parallel_for(uint64_t n, int threads, F f) {
    // f(begin, end, tid) over contiguous ranges.
    if (threads <= 1 || n < 4096) {
        f(0, n, 0);
        return;
    }
    std::vector<std::thread> th;
    uint64_t chunk = (n + threads - 1) / threads;
    for (int t = 0; t < threads; t++) {
        uint64_t b = std::min(n, chunk * t), e = std::min(n, b + chunk);
        th.emplace_back([=] { f(b, e, t); });
    }
    for (auto &x : th) x.join();
}

// One sample of the edge stream: R-MAT quadrant descent or a uniform pair.
struct Sampler {
    uint32_t scale;
    uint64_t base;
    bool rmat;
    uint32_t ta, tab, tabc;  // cumulative thresholds in 2^-32 units
    uint32_t words;          // 64-bit hashes per sample
    uint32_t V;

    inline void edge(uint64_t i, uint32_t &u, uint32_t &v) const {
        if (!rmat) {
            uint64_t h = splitmix64(base + i);
            u = (uint32_t)(((uint64_t)(uint32_t)h * V) >> 32);
            v = (uint32_t)(((h >> 32) * V) >> 32);
            return;
        }
        uint32_t uu = 0, vv = 0;
        for (uint32_t j = 0; j < words; j++) {
            uint64_t h = splitmix64(base + i * words + j);
            for (int half = 0; half < 2; half++) {
                uint32_t k = 2 * j + half;
                if (k >= scale) break;
                uint32_t r = half ? (uint32_t)(h >> 32) : (uint32_t)h;
                uint32_t qu, qv;
                if (r < ta) {
                    qu = 0; qv = 0;
                } else if (r < tab) {
                    qu = 0; qv = 1;
                } else if (r < tabc) {
                    qu = 1; qv = 0;
                } else {
                    qu = 1; qv = 1;
                }
                uu |= qu << (scale - 1 - k);
                vv |= qv << (scale - 1 - k);
            }
        }
        u = uu;
        v = vv;
    }
};

uint32_t thresh(double p) {
    double x = p * 4294967296.0;
    if (x <= 0) return 0;
    if (x >= 4294967295.0) return 0xFFFFFFFFu;
    return (uint32_t)x;
}

// Builds a clean undirected CSR from the sample stream: drop self-loops, add
// both directions, sort each adjacency ascending, drop duplicates.
int build_csr(const Sampler &s, uint64_t nsamples, int threads, ptg_host_csr *out) {
    const uint32_t V = s.V;
    std::vector<std::atomic<uint32_t>> deg(V);
    for (auto &d : deg) d.store(0, std::memory_order_relaxed);
    parallel_for(nsamples, threads, [&](uint64_t b, uint64_t e, int) {
        for (uint64_t i = b; i < e; i++) {
            uint32_t u, v;
            s.edge(i, u, v);
            if (u == v) continue;
            deg[u].fetch_add(1, std::memory_order_relaxed);
            deg[v].fetch_add(1, std::memory_order_relaxed);
        }
    });
    std::vector<uint64_t> off(V + 1);
    off[0] = 0;
    for (uint32_t v = 0; v < V; v++) off[v + 1] = off[v] + deg[v].load(std::memory_order_relaxed);
    const uint64_t raw = off[V];
    uint32_t *tmp = (uint32_t *)malloc(sizeof(uint32_t) * (raw ? raw : 1));
    if (!tmp) return ptg::fail(PTG_ERR_OUT_OF_MEMORY, "generator: host allocation failed");
    std::vector<std::atomic<uint64_t>> cur(V);
    for (uint32_t v = 0; v < V; v++) cur[v].store(off[v], std::memory_order_relaxed);
    parallel_for(nsamples, threads, [&](uint64_t b, uint64_t e, int) {
        for (uint64_t i = b; i < e; i++) {
            uint32_t u, v;
            s.edge(i, u, v);
            if (u == v) continue;
            tmp[cur[u].fetch_add(1, std::memory_order_relaxed)] = v;
            tmp[cur[v].fetch_add(1, std::memory_order_relaxed)] = u;
        }
    });
    // Sort + dedupe each adjacency in place; blocks of vertices handed out dynamically.
    std::vector<uint32_t> nd(V);
    std::atomic<uint32_t> next{0};
    const uint32_t blk = 4096;
    auto worker = [&] {
        for (;;) {
            uint32_t b = next.fetch_add(blk);
            if (b >= V) break;
            uint32_t e = std::min<uint64_t>((uint64_t)b + blk, V);
            for (uint32_t v = b; v < e; v++) {
                uint32_t *p = tmp + off[v], *q = tmp + off[v + 1];
                std::sort(p, q);
                nd[v] = (uint32_t)(std::unique(p, q) - p);
            }
        }
    };
    {
        std::vector<std::thread> th;
        for (int t = 1; t < threads; t++) th.emplace_back(worker);
        worker();
        for (auto &x : th) x.join();
    }
    uint32_t *offsets = (uint32_t *)malloc(sizeof(uint32_t) * ((uint64_t)V + 1));
    if (!offsets) {
        free(tmp);
        return ptg::fail(PTG_ERR_OUT_OF_MEMORY, "generator: host allocation failed");
    }
    uint64_t acc = 0;
    for (uint32_t v = 0; v < V; v++) {
        if (acc > 0xFFFFFFFFull) break;
        offsets[v] = (uint32_t)acc;
        acc += nd[v];
    }
    if (acc > 0xFFFFFFFFull) {
        free(tmp);
        free(offsets);
        return ptg::fail(PTG_ERR_UNSUPPORTED, "generator: more than 2^32-1 edges");
    }
    offsets[V] = (uint32_t)acc;
    uint32_t *cols = (uint32_t *)malloc(sizeof(uint32_t) * (acc ? acc : 1));
    if (!cols) {
        free(tmp);
        free(offsets);
        return ptg::fail(PTG_ERR_OUT_OF_MEMORY, "generator: host allocation failed");
    }
    parallel_for(V, threads, [&](uint64_t b, uint64_t e, int) {
        for (uint64_t v = b; v < e; v++)
            memcpy(cols + offsets[v], tmp + off[v], sizeof(uint32_t) * nd[v]);
    });
    free(tmp);
    out->V = V;
    out->E = acc;
    out->offsets = offsets;
    out->cols = cols;
    out->weights = nullptr;
    return PTG_OK;
}

bool read_all(FILE *f, void *p, size_t bytes) {
    char *c = (char *)p;
    while (bytes) {
        size_t n = fread(c, 1, bytes, f);
        if (n == 0) return false;
        c += n;
        bytes -= n;
    }
    return true;
}

uint32_t be32(const unsigned char *b) {
    return (uint32_t)b[0] << 24 | (uint32_t)b[1] << 16 | (uint32_t)b[2] << 8 | (uint32_t)b[3];
}

uint64_t be64(const unsigned char *b) { return (uint64_t)be32(b) << 32 | be32(b + 4); }

// Reads [u32 V][u32 E][u32 ptr x V] and checks the pointer array.
int read_header_ptr(FILE *f, const char *path, ptg_host_csr *out) {
    uint32_t hdr[2];
    if (!read_all(f, hdr, sizeof hdr)) return ptg::fail(PTG_ERR_IO, std::string("short header: ") + path);
    out->V = hdr[0];
    out->E = hdr[1];
    if (out->V == 0) return ptg::fail(PTG_ERR_IO, std::string("zero vertices: ") + path);
    out->offsets = (uint32_t *)malloc(sizeof(uint32_t) * ((uint64_t)out->V + 1));
    if (!out->offsets) return ptg::fail(PTG_ERR_OUT_OF_MEMORY, "reader: host allocation failed");
    if (!read_all(f, out->offsets, sizeof(uint32_t) * out->V))
        return ptg::fail(PTG_ERR_IO, std::string("short pointer array: ") + path);
    out->offsets[out->V] = (uint32_t)out->E;
    return PTG_OK;
}

}  // namespace

extern "C" {

int ptg_validate_csr(uint32_t V, uint64_t E, const uint32_t *ptr, const uint32_t *col,
                     const EdgeWithWeight *wedges) {
    if (V == 0 || !ptr) return ptg::fail(PTG_ERR_INVALID_ARG, "validate: V == 0 or null nodePointers");
    if (E > 0xFFFFFFFFull) return ptg::fail(PTG_ERR_INVALID_GRAPH, "validate: E exceeds uint32 (reference `uint E`)");
    if (E && !col && !wedges) return ptg::fail(PTG_ERR_INVALID_ARG, "validate: null edge array");
    for (uint32_t v = 0; v + 1 < V; v++)
        if (ptr[v] > ptr[v + 1])
            return ptg::fail(PTG_ERR_INVALID_GRAPH,
                             "validate: nodePointers decrease at vertex " + std::to_string(v));
    if (ptr[V - 1] > E)
        return ptg::fail(PTG_ERR_INVALID_GRAPH, "validate: nodePointers[V-1] > E");
    for (uint64_t e = ptr[0]; e < E; e++) {
        uint32_t c = col ? col[e] : wedges[e].toNode;
        if (c >= V)
            return ptg::fail(PTG_ERR_INVALID_GRAPH,
                             "validate: column " + std::to_string(c) + " >= V at edge " + std::to_string(e));
    }
    return PTG_OK;
}

int ptg_gen_rmat(uint32_t scale, uint32_t edge_factor, double a, double b, double c, uint64_t seed,
                 int threads, ptg_host_csr *out) {
    if (!out || scale == 0 || scale > 31 || edge_factor == 0)
        return ptg::fail(PTG_ERR_INVALID_ARG, "gen_rmat: bad scale/edge_factor/out");
    if (a < 0 || b < 0 || c < 0 || a + b + c > 1.0)
        return ptg::fail(PTG_ERR_INVALID_ARG, "gen_rmat: probabilities must be >= 0 with a+b+c <= 1");
    memset(out, 0, sizeof *out);
    Sampler s;
    s.scale = scale;
    s.base = splitmix64(seed);
    s.rmat = true;
    s.ta = thresh(a);
    s.tab = thresh(a + b);
    s.tabc = thresh(a + b + c);
    s.words = (scale + 1) / 2;
    s.V = 1u << scale;
    return build_csr(s, (uint64_t)edge_factor << scale, clamp_threads(threads), out);
}

int ptg_gen_er(uint32_t scale, uint32_t edge_factor, uint64_t seed, int threads, ptg_host_csr *out) {
    if (!out || scale == 0 || scale > 31 || edge_factor == 0)
        return ptg::fail(PTG_ERR_INVALID_ARG, "gen_er: bad scale/edge_factor/out");
    memset(out, 0, sizeof *out);
    Sampler s{};
    s.scale = scale;
    s.base = splitmix64(seed ^ 0x5851F42D4C957F2Dull);
    s.rmat = false;
    s.V = 1u << scale;
    return build_csr(s, (uint64_t)edge_factor << scale, clamp_threads(threads), out);
}

int ptg_gen_weights(ptg_host_csr *csr, uint32_t lo, uint32_t hi, uint64_t seed, int threads) {
    if (!csr || !csr->offsets || (csr->E && !csr->cols) || lo > hi)
        return ptg::fail(PTG_ERR_INVALID_ARG, "gen_weights: bad csr or range");
    free(csr->weights);
    csr->weights = (uint32_t *)malloc(sizeof(uint32_t) * (csr->E ? csr->E : 1));
    if (!csr->weights) return ptg::fail(PTG_ERR_OUT_OF_MEMORY, "gen_weights: host allocation failed");
    const uint64_t base = splitmix64(seed);
    const uint64_t span = (uint64_t)hi - lo + 1;
    parallel_for(csr->V, clamp_threads(threads), [&](uint64_t b, uint64_t e, int) {
        for (uint64_t u = b; u < e; u++)
            for (uint64_t k = csr->offsets[u]; k < csr->offsets[u + 1]; k++) {
                uint64_t v = csr->cols[k];
                uint64_t mn = std::min<uint64_t>(u, v), mx = std::max<uint64_t>(u, v);
                uint64_t h = splitmix64(base + (mn << 32 | mx));
                csr->weights[k] = (uint32_t)(lo + h % span);
            }
    });
    return PTG_OK;
}

void ptg_host_csr_free(ptg_host_csr *csr) {
    if (!csr) return;
    free(csr->offsets);
    free(csr->cols);
    free(csr->weights);
    memset(csr, 0, sizeof *csr);
}

int ptg_read_bcsr(const char *path, ptg_host_csr *out) {
    if (!path || !out) return ptg::fail(PTG_ERR_INVALID_ARG, "read_bcsr: null argument");
    memset(out, 0, sizeof *out);
    FILE *f = fopen(path, "rb");
    if (!f) return ptg::fail(PTG_ERR_IO, std::string("cannot open ") + path);
    int rc = read_header_ptr(f, path, out);
    if (rc == PTG_OK) {
        out->cols = (uint32_t *)malloc(sizeof(uint32_t) * (out->E ? out->E : 1));
        if (!out->cols)
            rc = ptg::fail(PTG_ERR_OUT_OF_MEMORY, "read_bcsr: host allocation failed");
        else if (!read_all(f, out->cols, sizeof(uint32_t) * out->E))
            rc = ptg::fail(PTG_ERR_IO, std::string("short edge array: ") + path);
    }
    fclose(f);
    if (rc != PTG_OK) ptg_host_csr_free(out);
    return rc;
}

int ptg_read_bwcsr(const char *path, ptg_host_csr *out) {
    if (!path || !out) return ptg::fail(PTG_ERR_INVALID_ARG, "read_bwcsr: null argument");
    memset(out, 0, sizeof *out);
    FILE *f = fopen(path, "rb");
    if (!f) return ptg::fail(PTG_ERR_IO, std::string("cannot open ") + path);
    int rc = read_header_ptr(f, path, out);
    if (rc == PTG_OK) {
        out->cols = (uint32_t *)malloc(sizeof(uint32_t) * (out->E ? out->E : 1));
        out->weights = (uint32_t *)malloc(sizeof(uint32_t) * (out->E ? out->E : 1));
        if (!out->cols || !out->weights) rc = ptg::fail(PTG_ERR_OUT_OF_MEMORY, "read_bwcsr: host allocation failed");
        const uint64_t blk = 1 << 16;
        std::vector<EdgeWithWeight> buf(blk);
        for (uint64_t e = 0; rc == PTG_OK && e < out->E; e += blk) {
            uint64_t n = std::min(blk, out->E - e);
            if (!read_all(f, buf.data(), sizeof(EdgeWithWeight) * n)) {
                rc = ptg::fail(PTG_ERR_IO, std::string("short edge array: ") + path);
                break;
            }
            for (uint64_t i = 0; i < n; i++) {
                out->cols[e + i] = buf[i].toNode;
                out->weights[e + i] = buf[i].weight;
            }
        }
    }
    fclose(f);
    if (rc != PTG_OK) ptg_host_csr_free(out);
    return rc;
}

int ptg_write_bcsr(const char *path, const ptg_host_csr *csr) {
    if (!path || !csr || !csr->offsets || csr->E > 0xFFFFFFFFull)
        return ptg::fail(PTG_ERR_INVALID_ARG, "write_bcsr: bad arguments");
    FILE *f = fopen(path, "wb");
    if (!f) return ptg::fail(PTG_ERR_IO, std::string("cannot create ") + path);
    uint32_t hdr[2] = {csr->V, (uint32_t)csr->E};
    bool ok = fwrite(hdr, sizeof hdr, 1, f) == 1 &&
              fwrite(csr->offsets, sizeof(uint32_t), csr->V, f) == csr->V &&
              (csr->E == 0 || fwrite(csr->cols, sizeof(uint32_t), csr->E, f) == csr->E);
    ok = (fclose(f) == 0) && ok;
    return ok ? PTG_OK : ptg::fail(PTG_ERR_IO, std::string("write failed: ") + path);
}

int ptg_write_bwcsr(const char *path, const ptg_host_csr *csr) {
    if (!path || !csr || !csr->offsets || !csr->weights || csr->E > 0xFFFFFFFFull)
        return ptg::fail(PTG_ERR_INVALID_ARG, "write_bwcsr: bad arguments (weights required)");
    FILE *f = fopen(path, "wb");
    if (!f) return ptg::fail(PTG_ERR_IO, std::string("cannot create ") + path);
    uint32_t hdr[2] = {csr->V, (uint32_t)csr->E};
    bool ok = fwrite(hdr, sizeof hdr, 1, f) == 1 &&
              fwrite(csr->offsets, sizeof(uint32_t), csr->V, f) == csr->V;
    const uint64_t blk = 1 << 16;
    std::vector<EdgeWithWeight> buf(blk);
    for (uint64_t e = 0; ok && e < csr->E; e += blk) {
        uint64_t n = std::min(blk, csr->E - e);
        for (uint64_t i = 0; i < n; i++) buf[i] = {csr->cols[e + i], csr->weights[e + i]};
        ok = fwrite(buf.data(), sizeof(EdgeWithWeight), n, f) == n;
    }
    ok = (fclose(f) == 0) && ok;
    return ok ? PTG_OK : ptg::fail(PTG_ERR_IO, std::string("write failed: ") + path);
}

int ptg_read_java_be(const char *path, ptg_host_csr *out) {
    if (!path || !out) return ptg::fail(PTG_ERR_INVALID_ARG, "read_java_be: null argument");
    memset(out, 0, sizeof *out);
    FILE *f = fopen(path, "rb");
    if (!f) return ptg::fail(PTG_ERR_IO, std::string("cannot open ") + path);
    unsigned char h[12];
    int rc = PTG_OK;
    if (!read_all(f, h, 12)) {
        fclose(f);
        return ptg::fail(PTG_ERR_IO, std::string("short header: ") + path);
    }
    out->V = be32(h);
    out->E = be64(h + 4);
    if (out->V == 0) rc = ptg::fail(PTG_ERR_IO, "read_java_be: zero vertices");
    else if (out->E > 0xFFFFFFFFull)
        rc = ptg::fail(PTG_ERR_UNSUPPORTED, "read_java_be: more than 2^32-1 edges (uint32 offsets)");
    if (rc == PTG_OK) {
        out->offsets = (uint32_t *)malloc(sizeof(uint32_t) * ((uint64_t)out->V + 1));
        out->cols = (uint32_t *)malloc(sizeof(uint32_t) * (out->E ? out->E : 1));
        if (!out->offsets || !out->cols) rc = ptg::fail(PTG_ERR_OUT_OF_MEMORY, "read_java_be: allocation failed");
    }
    std::vector<unsigned char> buf(1 << 16);
    for (uint64_t v = 0; rc == PTG_OK && v < out->V;) {
        uint64_t n = std::min<uint64_t>(buf.size() / 8, out->V - v);
        if (!read_all(f, buf.data(), n * 8)) {
            rc = ptg::fail(PTG_ERR_IO, "read_java_be: short pointer array");
            break;
        }
        for (uint64_t i = 0; i < n; i++) {
            uint64_t p = be64(buf.data() + 8 * i);
            if (p > out->E) {
                rc = ptg::fail(PTG_ERR_IO, "read_java_be: pointer beyond E");
                break;
            }
            out->offsets[v + i] = (uint32_t)p;
        }
        v += n;
    }
    if (rc == PTG_OK) out->offsets[out->V] = (uint32_t)out->E;
    for (uint64_t e = 0; rc == PTG_OK && e < out->E;) {
        uint64_t n = std::min<uint64_t>(buf.size() / 4, out->E - e);
        if (!read_all(f, buf.data(), n * 4)) {
            rc = ptg::fail(PTG_ERR_IO, "read_java_be: short edge array");
            break;
        }
        for (uint64_t i = 0; i < n; i++) out->cols[e + i] = be32(buf.data() + 4 * i);
        e += n;
    }
    fclose(f);
    if (rc != PTG_OK) ptg_host_csr_free(out);
    return rc;
}

}  // extern "C"
