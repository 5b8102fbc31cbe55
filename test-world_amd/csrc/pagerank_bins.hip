// PageRank over destination partitions ("binned" PageRank) for gfx950.
//
// The same definition and the same exact Q1.63 sums as the in-list pull of
// pagerank.hip (r' = base + d * sum, base = (1-d)/V + d*D/V), so the two give
// the same bits. What changes is how a sum is gathered. The pull reads
// r(u)/outdeg(u) at a random address for every edge, and a random 4-byte
// read costs the same address-unit time whether it hits L2 or not. Here:
//
//   partitions  destinations in ranges of kPW = 16384 vertices; a workgroup
//               keeps the 16384 sums of one range in LDS (128 KB of Q1.63);
//   pairs       one per (source u, partition p) with an edge u -> v, v in p,
//               ordered by partition, then source; bsrc[j] is pair j's
//               source. In the default degree order most of a vertex's edges
//               go to the hubs, which share partition 0, so there are far
//               fewer pairs than edges (0.16 E at R-MAT scale 22);
//   entries     one 16-bit word per edge, in pair order and, within a pair,
//               by destination: the destination's offset in its partition
//               (14 bits) and, in bit 15, whether the edge starts a new pair.
//               0x4000 is a null entry (padding to a multiple of 4);
//   segments    every kSeg entries of a partition: the index of the pair that
//               holds the entry before the segment, so a wave can start there;
//   chunks      one workgroup's share of k_prb_acc: a partition, or a run
//               of segments of a large one (the hub partitions). Each chunk
//               leaves its LDS sums as a row of kPW partial sums;
//   units       one workgroup's share of k_prb_fin: 1024 vertices of a
//               partition of at most kRowsDirect rows (a thread per vertex
//               adds its rows), or 64 vertices of a partition cut into more
//               (the 16 waves split the rows).
//
// Per iteration, k_prb_acc streams the entries (2 bytes per edge) and the
// pairs' sources (4 bytes per pair, read in order), reads each pair's
// contribution r(u)/outdeg(u) (within a partition the sources ascend, so a
// wave's reads share lines where the sources are dense) and adds each
// edge's term into LDS with a 64-bit LDS atomic; k_prb_fin adds up every
// vertex's rows and finishes it: its rank, its share of the dangling mass
// and its contribution for the next iteration (one coalesced store per
// vertex). Workgroup 0 of k_prb_acc sums the previous iteration's dangling
// partials into the base value k_prb_fin uses. Kernel boundaries order
// everything. (Writing each pair's contribution to a bin in pair order
// instead, so that k_prb_acc reads the bins as a stream, took 20 M scattered
// stores per iteration at R-MAT scale 22 and was slower.)
//
// The layout is built once per graph on the host from the device out-lists
// (the edges u -> v themselves, so a directed graph needs no transpose) and
// kept in host memory, so releasing the device copy and uploading it again
// is a plain copy.
//
// SSSP sweep rounds use the same layout with one more array, a weight byte
// per entry (graphs whose weights are at most 255). A sweep relaxes every
// edge u -> v with the value u had when the round began: k_sw_acc reads each
// entry's source value through its pair's source, as k_prb_acc reads
// contributions, takes per chunk the minimum of value + weight over the
// entries of each destination in LDS (a 32-bit LDS atomicMin) and writes the
// chunk's row of minima, and k_sw_fin
// takes each vertex's minimum over its partition's rows; a vertex whose
// minimum is below its value takes it and joins the next frontier, as a
// pushed vertex does. Relaxing an edge out of a vertex outside the frontier
// changes nothing (its edges were relaxed at its current value), so a sweep
// round reaches what a push round would, by the same strict-< rule, and the
// distances are the same.
#include <string.h>

#include <algorithm>
#include <string>
#include <thread>
#include <vector>

#include "rounds.h"

namespace ptg {

constexpr uint32_t kPWBits = 14;
constexpr uint32_t kPW = 1u << kPWBits;   // destinations per partition
constexpr uint32_t kSeg = 4096;           // entries per segment
constexpr uint32_t kAccBlock = 1024;      // threads of k_prb_acc / k_prb_fin
constexpr uint32_t kAccWaves = kAccBlock / 64;
constexpr uint32_t kTile = 64;            // vertices of a k_prb_fin unit whose rows the waves split
constexpr uint32_t kRowsDirect = 4;       // at most this many rows: a thread per vertex adds them
constexpr uint16_t kEntFirst = 0x8000, kEntNull = 0x4000, kEntDest = kPW - 1;
constexpr uint64_t kNull4 = 0x4000400040004000ull;  // four null entries
constexpr int kGroups = 4;                           // 256-entry groups per step of a k_prb_acc wave
// Chunk target: the entries / (2 x CUs); a partition of more entries is cut
// into chunks of about that many. Finer chunks balance the launch better but
// each writes a row of kPW sums (R-MAT-22: 1, 2, 3, 4 and 6 chunks per CU,
// cutting above the target or twice it, all measured; 2 and the target were
// the fastest).
constexpr uint64_t kChunksPerCu = 2;
constexpr uint64_t kChunkSplit = 1;

struct PrbChunk {
    uint64_t e0, e1;  // entries [e0, e1), within one partition
    uint32_t seg0;    // segment of e0
    uint32_t row;     // row of partial sums it writes
};

struct PrbUnit {
    uint32_t v0;     // first vertex (within one partition)
    uint32_t nv;     // vertices: up to kAccBlock, or kTile when nrows > kRowsDirect
    uint32_t row0;   // the partition's rows
    uint32_t nrows;
};

// Host copy of the layout.
struct PrBinsHost {
    std::vector<uint16_t> ent;
    std::vector<uint8_t> went;  // SSSP sweeps: the weight of every entry (0 for nulls); empty when not built
    std::vector<uint32_t> bsrc; // the source of every pair, in pair order
    bool has_w = false;
    std::vector<uint32_t> seg_cur;
    std::vector<PrbChunk> chunks;
    std::vector<PrbUnit> units;
    uint32_t n_rows = 0;
    uint64_t n_pairs = 0;
};

struct PrBins {
    PrBinsHost h;
    bool built = false;
    bool usable = false;  // auto layout: false when pairs would exceed half the edges
    bool sweep_refused = false;  // SSSP sweeps: the layout would not pay (pairs above half the edges)
    // device copy
    uint16_t *ent = nullptr;
    uint8_t *went = nullptr;
    uint32_t *bsrc = nullptr;
    uint32_t *seg_cur = nullptr;
    PrbChunk *chunks = nullptr;
    PrbUnit *units = nullptr;
    float *con[2] = {};  // contribution r(v)/outdeg(v) of every vertex, two iterations in turn
    unsigned long long *rows = nullptr;
    unsigned long long *part = nullptr;  // 2 iterations x units, then the base value
    bool on_device = false;
};

struct PrbArgs {
    const uint16_t *ent;
    const uint32_t *seg_cur;
    const PrbChunk *chunks;
    const PrbUnit *units;
    const uint32_t *ooff;  // out offsets (outdeg)
    const uint32_t *bsrc;  // the source of every pair
    const float *con_in;   // contributions of this iteration
    float *con_out;        // ... of the next
    float *rank_out;       // last iteration only
    unsigned long long *rows;
    const unsigned long long *part_in;  // previous iteration's dangling partials (one per unit)
    unsigned long long *part_out;
    float *base;           // this iteration's base value
    uint32_t n_units;
    uint32_t n_vert;
    uint32_t last_pair;    // pairs - 1
    bool first;            // iteration 0: the dangling mass is d0
    bool init;             // k_prb_fin writes the contributions of r0 = 1/V only
    double d0;
    float damping, inv_v;
};

// Workgroup 0 of k_prb_acc: base = (1-d)/V + d * D/V from the previous
// iteration's dangling partials.
__device__ __forceinline__ void prb_base(const PrbArgs &a, unsigned long long *lds) {
    double dang = a.d0;
    if (!a.first) {
        unsigned long long t = 0;
        for (uint32_t j = threadIdx.x; j < a.n_units; j += kAccBlock) t += a.part_in[j];
        t = wave_sum_u64(t);
        if (lane_id() == 0) lds[threadIdx.x >> 6] = t;
        __syncthreads();
        unsigned long long s = 0;
        for (uint32_t w = 0; w < kAccWaves; w++) s += lds[w];
        dang = fx63_to_f64(s);
    }
    if (threadIdx.x == 0) *a.base = pr_base_value(a.damping, a.inv_v, dang);
}

// One workgroup per chunk: adds every entry's term into LDS sums of its
// partition and writes them as the chunk's row.
__global__ __launch_bounds__(kAccBlock) void k_prb_acc(PrbArgs a) {
    __shared__ unsigned long long s_sum[kPW];
    __shared__ unsigned long long s_red[kAccWaves];
    const PrbChunk ch = a.chunks[blockIdx.x];
    if (blockIdx.x == 0) prb_base(a, s_red);
    for (uint32_t i = threadIdx.x; i < kPW; i += kAccBlock) s_sum[i] = 0;
    __syncthreads();
    const uint32_t wave = threadIdx.x >> 6, lane = lane_id();
    uint32_t j = ch.seg0 + wave;
    for (uint64_t s = ch.e0 + (uint64_t)wave * kSeg; s < ch.e1; s += (uint64_t)kAccWaves * kSeg, j += kAccWaves) {
        const uint64_t se = s + kSeg < ch.e1 ? s + kSeg : ch.e1;
        uint32_t cur = a.seg_cur[j];  // pair holding the entry before s
        // kGroups groups of 256 entries per step, four per lane; lanes past
        // the end read nulls (no pair start). The next step's entries are
        // loaded before this step's bins are used.
        auto load4 = [&](uint64_t e) -> uint64_t {
            return e < se ? *reinterpret_cast<const uint64_t *>(a.ent + e) : kNull4;
        };
        uint64_t w[kGroups];
#pragma unroll
        for (int q = 0; q < kGroups; q++) w[q] = load4(s + 256ull * q + 4ull * lane);
        for (uint64_t g = s; g < se; g += 256ull * kGroups) {
            uint64_t nxt[kGroups];
#pragma unroll
            for (int q = 0; q < kGroups; q++) nxt[q] = load4(g + 256ull * (kGroups + q) + 4ull * lane);
            uint32_t src[kGroups][4];
            float x[kGroups][4];
#pragma unroll
            for (int q = 0; q < kGroups; q++) {
                const uint64_t w4 = w[q];
                const uint32_t cnt = (uint32_t)((w4 >> 15) & 1) + (uint32_t)((w4 >> 31) & 1) +
                                     (uint32_t)((w4 >> 47) & 1) + (uint32_t)((w4 >> 63) & 1);
                const uint64_t b0 = ballot(cnt & 1), b1 = ballot(cnt & 2), b2 = ballot(cnt & 4);
                uint32_t pr = cur + rank_below(b0) + 2 * rank_below(b1) + 4 * rank_below(b2);
                cur += (uint32_t)(__popcll(b0) + 2 * __popcll(b1) + 4 * __popcll(b2));
                // every lane's pair index is a valid bin (a null entry, or a
                // lane past the end, repeats the pair before it); clamped
                // regardless
#pragma unroll
                for (int i = 0; i < 4; i++) {
                    pr += (uint32_t)(w4 >> (16 * i + 15)) & 1u;
                    src[q][i] = a.bsrc[min(pr, a.last_pair)];
                }
            }
            // every pair's contribution read before any is used
#pragma unroll
            for (int q = 0; q < kGroups; q++)
#pragma unroll
                for (int i = 0; i < 4; i++) x[q][i] = a.con_in[src[q][i]];
#pragma unroll
            for (int q = 0; q < kGroups; q++)
#pragma unroll
                for (int i = 0; i < 4; i++) {
                    const uint32_t d = (uint32_t)(w[q] >> (16 * i)) & 0xFFFFu;
                    if (!(d & kEntNull))
                        __hip_atomic_fetch_add(&s_sum[d & kEntDest], fx63(x[q][i]), __ATOMIC_RELAXED,
                                               __HIP_MEMORY_SCOPE_WORKGROUP);
                }
#pragma unroll
            for (int q = 0; q < kGroups; q++) w[q] = nxt[q];
        }
    }
    __syncthreads();
    unsigned long long *row = a.rows + (uint64_t)ch.row * kPW;
    for (uint32_t i = threadIdx.x; i < kPW; i += kAccBlock) row[i] = s_sum[i];
}

// One workgroup per unit: adds up each vertex's rows and finishes it (rank,
// dangling share, contribution for the next iteration). With `init`, writes
// the contributions of r0 = 1/V instead.
__global__ __launch_bounds__(kAccBlock) void k_prb_fin(PrbArgs a) {
    __shared__ unsigned long long s_part[kAccWaves][kTile];
    __shared__ unsigned long long s_red[kAccWaves];
    const PrbUnit u = a.units[blockIdx.x];
    const uint32_t wave = threadIdx.x >> 6, lane = lane_id();
    const uint32_t off0 = u.v0 & kEntDest;
    unsigned long long sum = 0;
    if (!a.init) {
        if (u.nrows > kRowsDirect) {
            for (uint32_t r = wave; r < u.nrows; r += kAccWaves) sum += a.rows[(uint64_t)(u.row0 + r) * kPW + off0 + lane];
            s_part[wave][lane] = sum;
            __syncthreads();
            sum = 0;
            if (threadIdx.x < kTile)
                for (uint32_t w = 0; w < kAccWaves; w++) sum += s_part[w][threadIdx.x];
        } else if (threadIdx.x < u.nv) {
            for (uint32_t r = 0; r < u.nrows; r++) sum += a.rows[(uint64_t)(u.row0 + r) * kPW + off0 + threadIdx.x];
        }
    }
    unsigned long long dang = 0;
    if (threadIdx.x < u.nv) {
        const uint32_t v = u.v0 + threadIdx.x;
        const uint32_t od = a.ooff[v + 1] - a.ooff[v];
        float c;
        if (a.init) {
            c = od ? a.inv_v / (float)od : 0.0f;
        } else {
            const float r = pr_rank(*a.base, a.damping, sum);
            if (a.rank_out) a.rank_out[v] = r;
            if (!od) dang = fx63(r);
            c = od ? r / (float)od : 0.0f;
        }
        a.con_out[v] = c;
    }
    if (a.init) return;
    dang = wave_sum_u64(dang);
    if (lane == 0) s_red[wave] = dang;
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned long long t = 0;
        for (uint32_t w = 0; w < kAccWaves; w++) t += s_red[w];
        a.part_out[blockIdx.x] = t;
    }
}

// ---------------------------------------------------------------------------
// Host build of the layout.
// ---------------------------------------------------------------------------

static int build_threads() {
    const unsigned h = std::thread::hardware_concurrency();
    return (int)std::max(1u, std::min(h ? h : 1u, 16u));
}

template <class F>
static void run_ranges(const std::vector<uint32_t> &vb, F f) {
    std::vector<std::thread> th;
    for (size_t t = 0; t + 1 < vb.size(); t++) th.emplace_back([=] { f((int)t, vb[t], vb[t + 1]); });
    for (auto &x : th) x.join();
}

// Builds the layout from CSR out-lists (each list ascending). Returns false,
// with nothing built, when `require_gain` and the pairs exceed half the edges.
static bool build_layout(uint32_t V, uint64_t E, const uint32_t *off, const uint32_t *col, const uint32_t *w,
                         uint32_t num_cus, bool require_gain, PrBinsHost *out) {
    const uint32_t P = (V + kPW - 1) >> kPWBits;
    // thread ranges of about equal edges
    const int T = (int)std::max<uint64_t>(1, std::min<uint64_t>(build_threads(), (uint64_t)V / 4096 + 1));
    std::vector<uint32_t> vb(T + 1, V);
    vb[0] = 0;
    for (int t = 1; t < T; t++) {
        const uint64_t target = E * (uint64_t)t / (uint64_t)T;
        vb[t] = (uint32_t)(std::upper_bound(off, off + V, (uint32_t)target) - off);
        vb[t] = std::max(vb[t], vb[t - 1]);
    }
    // pass 1: pairs and entries per (thread, partition)
    std::vector<std::vector<uint64_t>> cp(T, std::vector<uint64_t>(P, 0)), ce(T, std::vector<uint64_t>(P, 0));
    run_ranges(vb, [&](int t, uint32_t b, uint32_t e) {
        auto &mp = cp[t];
        auto &me = ce[t];
        for (uint32_t u = b; u < e; u++) {
            uint32_t prev = 0xFFFFFFFFu;
            for (uint32_t k = off[u]; k < off[u + 1]; k++) {
                const uint32_t p = col[k] >> kPWBits;
                me[p]++;
                if (p != prev) {
                    mp[p]++;
                    prev = p;
                }
            }
        }
    });
    uint64_t pairs = 0;
    for (int t = 0; t < T; t++)
        for (uint32_t p = 0; p < P; p++) pairs += cp[t][p];
    if (require_gain && pairs * 2 > E) return false;
    // partition offsets (entries padded to a multiple of 4), thread cursors
    std::vector<uint64_t> ent_off(P + 1, 0), pair_off(P + 1, 0);
    std::vector<uint32_t> seg_off(P + 1, 0);
    for (uint32_t p = 0; p < P; p++) {
        uint64_t ne = 0, npp = 0;
        for (int t = 0; t < T; t++) {
            ne += ce[t][p];
            npp += cp[t][p];
        }
        ent_off[p + 1] = ent_off[p] + ((ne + 3) & ~3ull);
        pair_off[p + 1] = pair_off[p] + npp;
        seg_off[p + 1] = seg_off[p] + (uint32_t)((ent_off[p + 1] - ent_off[p] + kSeg - 1) / kSeg);
    }
    std::vector<std::vector<uint64_t>> ec(T, std::vector<uint64_t>(P)), pc(T, std::vector<uint64_t>(P));
    for (uint32_t p = 0; p < P; p++) {
        uint64_t e = ent_off[p], q = pair_off[p];
        for (int t = 0; t < T; t++) {
            ec[t][p] = e;
            pc[t][p] = q;
            e += ce[t][p];
            q += cp[t][p];
        }
    }
    out->n_pairs = pairs;
    out->bsrc.assign(pairs ? pairs : 1, 0);
    out->ent.assign(ent_off[P] ? ent_off[P] : 4, kEntNull);
    out->has_w = w != nullptr;
    if (w)
        out->went.assign(out->ent.size(), 0);
    else
        std::vector<uint8_t>().swap(out->went);
    out->seg_cur.assign(seg_off[P] ? seg_off[P] : 1, 0);
    // pass 2: entries, the pairs' sources, segment starts
    run_ranges(vb, [&](int t, uint32_t b, uint32_t e) {
        auto &ecur = ec[t];
        auto &pcur = pc[t];
        for (uint32_t u = b; u < e; u++) {
            uint32_t prev = 0xFFFFFFFFu;
            uint32_t pp = 0;
            for (uint32_t k = off[u]; k < off[u + 1]; k++) {
                const uint32_t v = col[k];
                const uint32_t p = v >> kPWBits;
                const bool first = p != prev;
                if (first) {
                    pp = (uint32_t)pcur[p]++;
                    out->bsrc[pp] = u;
                    prev = p;
                }
                const uint64_t x = ecur[p]++;
                out->ent[x] = (uint16_t)((v & kEntDest) | (first ? kEntFirst : 0));
                if (w) out->went[x] = (uint8_t)w[k];
                const uint64_t rel = x - ent_off[p];
                if (rel % kSeg == 0) out->seg_cur[seg_off[p] + rel / kSeg] = pp - (first ? 1u : 0u);
            }
        }
    });
    // chunks: a partition of at most kChunkSplit x target entries is one
    // chunk; larger ones are cut into runs of whole segments of about target
    // entries. Each chunk has a row; the units of a partition add up its rows.
    const uint64_t target = std::max<uint64_t>(16ull * kSeg, ent_off[P] / (kChunksPerCu * std::max<uint32_t>(1, num_cus)));
    out->chunks.clear();
    out->units.clear();
    out->n_rows = 0;
    for (uint32_t p = 0; p < P; p++) {
        const uint64_t n = ent_off[p + 1] - ent_off[p];
        const uint32_t row0 = out->n_rows;
        if (n <= kChunkSplit * target) {
            out->chunks.push_back(PrbChunk{ent_off[p], ent_off[p + 1], seg_off[p], out->n_rows++});
        } else {
            const uint64_t nseg = seg_off[p + 1] - seg_off[p];
            const uint64_t k = (n + target - 1) / target;
            const uint64_t per = (nseg + k - 1) / k;  // segments per chunk
            for (uint64_t s = 0; s < nseg; s += per) {
                const uint64_t e0 = ent_off[p] + s * kSeg;
                const uint64_t e1 = std::min<uint64_t>(ent_off[p + 1], e0 + per * kSeg);
                out->chunks.push_back(PrbChunk{e0, e1, seg_off[p] + (uint32_t)s, out->n_rows++});
            }
        }
        const uint32_t nrows = out->n_rows - row0;
        const uint32_t step = nrows > kRowsDirect ? kTile : kAccBlock;
        const uint32_t v0 = p << kPWBits, v1 = std::min<uint32_t>(V, v0 + kPW);
        for (uint32_t v = v0; v < v1; v += step) out->units.push_back(PrbUnit{v, std::min(step, v1 - v), row0, nrows});
    }
    // largest chunks first, so the tail of the launch is short ones
    std::stable_sort(out->chunks.begin(), out->chunks.end(),
                     [](const PrbChunk &x, const PrbChunk &y) { return x.e1 - x.e0 > y.e1 - y.e0; });
    return true;
}

static uint64_t al256(uint64_t x) { return (x + 255) & ~255ull; }

static uint64_t device_bytes_of(const PrBinsHost &h, uint32_t V) {
    return al256(2 * h.ent.size()) + al256(h.went.size()) + al256(4 * h.bsrc.size()) + al256(4 * h.seg_cur.size()) +
           al256(sizeof(PrbChunk) * h.chunks.size()) + al256(sizeof(PrbUnit) * h.units.size()) +
           2 * al256(4ull * V) + al256(8ull * kPW * h.n_rows) + al256(8 * (2 * h.units.size() + 1));
}

static void free_device(ptg_graph *g, PrBins *b) {
    gfree(g, b->ent);
    gfree(g, b->went);
    gfree(g, b->bsrc);
    gfree(g, b->seg_cur);
    gfree(g, b->chunks);
    gfree(g, b->units);
    gfree(g, b->con[0]);
    gfree(g, b->con[1]);
    gfree(g, b->rows);
    gfree(g, b->part);
    b->ent = nullptr;
    b->went = nullptr;
    b->bsrc = b->seg_cur = nullptr;
    b->chunks = nullptr;
    b->units = nullptr;
    b->con[0] = b->con[1] = nullptr;
    b->rows = b->part = nullptr;
    b->on_device = false;
}

static int upload(ptg_graph *g, PrBins *b) {
    const PrBinsHost &h = b->h;
    int rc;
    if ((rc = galloc(g, (void **)&b->ent, 2 * h.ent.size())) ||
        (rc = galloc(g, (void **)&b->bsrc, 4 * h.bsrc.size())) ||
        (rc = galloc(g, (void **)&b->seg_cur, 4 * h.seg_cur.size())) ||
        (rc = galloc(g, (void **)&b->chunks, sizeof(PrbChunk) * h.chunks.size())) ||
        (rc = galloc(g, (void **)&b->units, sizeof(PrbUnit) * h.units.size())) ||
        (rc = galloc(g, (void **)&b->con[0], 4ull * g->V)) || (rc = galloc(g, (void **)&b->con[1], 4ull * g->V)) ||
        (rc = galloc(g, (void **)&b->rows, 8ull * kPW * h.n_rows)) ||
        (rc = galloc(g, (void **)&b->part, 8 * (2 * h.units.size() + 1))) ||
        (h.has_w && (rc = galloc(g, (void **)&b->went, h.went.size())))) {
        free_device(g, b);
        return rc;
    }
    if ((rc = copy_sync(g, b->ent, h.ent.data(), 2 * h.ent.size(), hipMemcpyHostToDevice)) ||
        (rc = copy_sync(g, b->bsrc, h.bsrc.data(), 4 * h.bsrc.size(), hipMemcpyHostToDevice)) ||
        (rc = copy_sync(g, b->seg_cur, h.seg_cur.data(), 4 * h.seg_cur.size(), hipMemcpyHostToDevice)) ||
        (rc = copy_sync(g, b->chunks, h.chunks.data(), sizeof(PrbChunk) * h.chunks.size(), hipMemcpyHostToDevice)) ||
        (rc = copy_sync(g, b->units, h.units.data(), sizeof(PrbUnit) * h.units.size(), hipMemcpyHostToDevice)) ||
        (h.has_w && (rc = copy_sync(g, b->went, h.went.data(), h.went.size(), hipMemcpyHostToDevice)))) {
        free_device(g, b);
        return rc;
    }
    b->on_device = true;
    return PTG_OK;
}

// Whether the layout's device copy fits the graph's device budget (galloc
// refuses allocations past it).
static bool fits_budget(const ptg_graph *g, const PrBins *b) {
    return !g->cfg.device_budget_bytes || g->device_bytes + device_bytes_of(b->h, g->V) <= g->cfg.device_budget_bytes;
}

// Builds the layout from the device out-lists (with a weight byte per entry
// when with_w; the weights must be at most 255).
static int build_from_device(ptg_graph *g, bool require_gain, bool *usable, bool with_w = false) {
    const uint32_t V = g->V;
    const uint64_t E = g->E;
    std::vector<uint32_t> off((size_t)V + 1), col(E ? E : 1), w;
    int rc;
    if ((rc = copy_sync(g, off.data(), g->out.off, 4ull * ((uint64_t)V + 1), hipMemcpyDeviceToHost)) ||
        (E && (rc = copy_sync(g, col.data(), g->out.col, 4 * E, hipMemcpyDeviceToHost))))
        return rc;
    if (with_w) {
        w.resize(E ? E : 1);
        if (E && (rc = copy_sync(g, w.data(), g->d_w, 4 * E, hipMemcpyDeviceToHost))) return rc;
    }
    *usable = build_layout(V, E, off.data(), col.data(), with_w ? w.data() : nullptr,
                           (uint32_t)std::max(1, g->num_cus), require_gain, &g->prb->h);
    return PTG_OK;
}

// Builds (once) and uploads the layout; *use = false when the graph takes
// the in-list pull instead (automatic layout and more pairs than E / 2).
static int ensure_bins(ptg_graph *g, bool *use) {
    *use = false;
    if (!g->prb) g->prb = new PrBins();
    PrBins *b = g->prb;
    const bool forced = g->cfg.pr_layout == PTG_PR_BINNED;
    int rc;
    if (!b->built) {
        if ((rc = build_from_device(g, !forced, &b->usable))) return rc;
        b->built = true;
    }
    if (!b->usable) {
        if (!forced) return PTG_OK;
        if ((rc = build_from_device(g, false, &b->usable))) return rc;  // refused earlier by the automatic layout
    }
    // the automatic layout gives way to the in-list pull when the layout does
    // not fit the graph's device budget
    if (!b->on_device && !forced && !fits_budget(g, b)) return PTG_OK;
    if (!b->on_device && (rc = upload(g, b))) return rc;
    *use = true;
    return PTG_OK;
}

bool pr_bins_wanted(const ptg_graph *g) { return g->cfg.pr_layout != PTG_PR_PULL; }

uint64_t pr_bins_need(const ptg_graph *g) {
    const PrBins *b = g->prb;
    if (b && b->built) return b->usable && !b->on_device ? device_bytes_of(b->h, g->V) : 0;
    // not built yet: entries, pairs at most E / 2 (a source each), two
    // contribution arrays, and a row per partition and per extra chunk
    const uint64_t P = (g->V + kPW - 1) / kPW;
    const uint64_t rows = 2 * P + kChunksPerCu * (uint64_t)std::max(1, g->num_cus);
    return al256(2 * (g->E + 4 * P)) + al256(4 * (g->E / 2 + 1)) + 2 * al256(4ull * g->V) +
           al256(8ull * kPW * rows) + al256(4 * (g->E / kSeg + P + 1)) + (1u << 20);
}

void release_pr_bins(ptg_graph *g) {
    if (g->prb) free_device(g, g->prb);
}

void destroy_pr_bins(ptg_graph *g) {
    if (!g->prb) return;
    free_device(g, g->prb);
    delete g->prb;
    g->prb = nullptr;
}

int run_pagerank_bins(ptg_graph *g, float damping, int iters, float *ranks, ptg_stats *st, bool *done) {
    *done = false;
    bool use = false;
    int rc = ensure_bins(g, &use);
    if (rc || !use) return rc;
    *done = true;
    PrBins *b = g->prb;
    const uint32_t V = g->V;
    const float r0 = 1.0f / (float)V;
    LaunchTimer timer(g);
    PTG_HIP(hipEventRecord(g->ev0, g->stream));
    if (iters == 0) {
        union {
            float f;
            uint32_t u;
        } cv;
        cv.f = r0;
        if ((rc = fill_u32(g, reinterpret_cast<uint32_t *>(ranks), V, cv.u))) return rc;
    } else {
        const uint32_t nc = (uint32_t)b->h.chunks.size(), nu = (uint32_t)b->h.units.size();
        PrbArgs a;
        memset(&a, 0, sizeof a);
        a.ent = b->ent;
        a.seg_cur = b->seg_cur;
        a.chunks = b->chunks;
        a.units = b->units;
        a.ooff = g->out.off;
        a.bsrc = b->bsrc;
        a.rows = b->rows;
        a.base = reinterpret_cast<float *>(b->part + 2ull * nu);
        a.n_units = nu;
        a.n_vert = V;
        a.last_pair = (uint32_t)(b->h.n_pairs ? b->h.n_pairs - 1 : 0);
        a.d0 = (double)g->out.bins.n_zero / (double)V;  // D0 = (vertices without out-edges) / V
        a.damping = damping;
        a.inv_v = r0;
        a.init = true;
        a.con_out = b->con[0];
        k_prb_fin<<<nu, kAccBlock, 0, g->stream>>>(a);
        PTG_HIP(hipGetLastError());
        a.init = false;
        for (int t = 0; t < iters; t++) {
            a.con_in = b->con[t & 1];
            a.con_out = b->con[(t + 1) & 1];
            a.rank_out = t == iters - 1 ? ranks : nullptr;
            a.part_in = b->part + (uint64_t)nu * ((t + 1) & 1);
            a.part_out = b->part + (uint64_t)nu * (t & 1);
            a.first = t == 0;
            timer.launch(k_prb_acc, dim3(nc), dim3(kAccBlock), a);
            timer.launch_rider(k_prb_fin, dim3(nu), dim3(kAccBlock), a);
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

// ---------------------------------------------------------------------------
// SSSP sweep rounds.
// ---------------------------------------------------------------------------

struct SwArgs {
    RoundEnv e;               // the round: ring entry e.r in, e.r + 1 out, output flags
    const uint16_t *ent;
    const uint8_t *went;
    const uint32_t *seg_cur;
    const PrbChunk *chunks;
    const PrbUnit *units;
    const uint32_t *bsrc;     // the source of each pair, in bin order
    uint32_t *rows;           // per chunk, the minima of its partition's 16384 destinations
    uint32_t *val;
    uint32_t sent;            // value of unreached vertices (their bins relax nothing)
    uint32_t last_pair;
    uint32_t n_vert;
    uint32_t n_units;
    float factor;             // the round's pull factor (0: no pull kernel this round)
    uint64_t min_edges;       // the round's sweep threshold
};

// Every kernel of a sweep checks first whether round e.r sweeps, with the
// decision its push and pull kernels make.
__device__ __forceinline__ bool sw_active(const SwArgs &a) {
    const RoundCtr *in = a.e.ring + a.e.r % kRing;
    uint32_t found = 0;
    uint64_t edges = 0;
#pragma unroll
    for (int i = 0; i < kShards; i++) {
        found += in->found_sh[i];
        edges += in->edges_sh[i];
    }
    return round_sweeps(in, found, edges, a.e.qcap, a.factor, a.min_edges);
}

// One workgroup per chunk: the minimum of bin + weight over the chunk's
// entries of every destination of its partition, in LDS, written as the
// chunk's row (0xFFFFFFFF where no entry reaches a destination). The walk is
// k_prb_acc's: four 256-entry groups in flight per wave, each entry's pair
// found from the segment's first pair and the entries' first-of-pair bits.
__global__ __launch_bounds__(kAccBlock) void k_sw_acc(SwArgs a) {
    __shared__ uint32_t s_min[kPW];
    if (!sw_active(a)) return;
    const PrbChunk ch = a.chunks[blockIdx.x];
    for (uint32_t i = threadIdx.x; i < kPW; i += kAccBlock) s_min[i] = 0xFFFFFFFFu;
    __syncthreads();
    const uint32_t wave = threadIdx.x >> 6, lane = lane_id();
    uint32_t j = ch.seg0 + wave;
    for (uint64_t s = ch.e0 + (uint64_t)wave * kSeg; s < ch.e1; s += (uint64_t)kAccWaves * kSeg, j += kAccWaves) {
        const uint64_t se = s + kSeg < ch.e1 ? s + kSeg : ch.e1;
        uint32_t cur = a.seg_cur[j];  // pair holding the entry before s
        auto load4 = [&](uint64_t e) -> uint64_t {
            return e < se ? *reinterpret_cast<const uint64_t *>(a.ent + e) : kNull4;
        };
        auto loadw = [&](uint64_t e) -> uint32_t {
            return e < se ? *reinterpret_cast<const uint32_t *>(a.went + e) : 0u;
        };
        uint64_t w[kGroups];
        uint32_t wt[kGroups];
#pragma unroll
        for (int q = 0; q < kGroups; q++) {
            w[q] = load4(s + 256ull * q + 4ull * lane);
            wt[q] = loadw(s + 256ull * q + 4ull * lane);
        }
        for (uint64_t g = s; g < se; g += 256ull * kGroups) {
            uint64_t nxt[kGroups];
            uint32_t nwt[kGroups];
#pragma unroll
            for (int q = 0; q < kGroups; q++) {
                nxt[q] = load4(g + 256ull * (kGroups + q) + 4ull * lane);
                nwt[q] = loadw(g + 256ull * (kGroups + q) + 4ull * lane);
            }
            uint32_t src[kGroups][4], x[kGroups][4];
#pragma unroll
            for (int q = 0; q < kGroups; q++) {
                const uint64_t w4 = w[q];
                const uint32_t cnt = (uint32_t)((w4 >> 15) & 1) + (uint32_t)((w4 >> 31) & 1) +
                                     (uint32_t)((w4 >> 47) & 1) + (uint32_t)((w4 >> 63) & 1);
                const uint64_t b0 = ballot(cnt & 1), b1 = ballot(cnt & 2), b2 = ballot(cnt & 4);
                uint32_t pr = cur + rank_below(b0) + 2 * rank_below(b1) + 4 * rank_below(b2);
                cur += (uint32_t)(__popcll(b0) + 2 * __popcll(b1) + 4 * __popcll(b2));
#pragma unroll
                for (int i = 0; i < 4; i++) {
                    pr += (uint32_t)(w4 >> (16 * i + 15)) & 1u;
                    src[q][i] = a.bsrc[min(pr, a.last_pair)];
                }
            }
            // every source's value read before any is used
#pragma unroll
            for (int q = 0; q < kGroups; q++)
#pragma unroll
                for (int i = 0; i < 4; i++) x[q][i] = a.val[src[q][i]];
#pragma unroll
            for (int q = 0; q < kGroups; q++)
#pragma unroll
                for (int i = 0; i < 4; i++) {
                    const uint32_t d = (uint32_t)(w[q] >> (16 * i)) & 0xFFFFu;
                    const uint32_t xv = x[q][i];
                    if (!(d & kEntNull) && xv < a.sent)
                        __hip_atomic_fetch_min(&s_min[d & kEntDest], xv + ((wt[q] >> (8 * i)) & 0xFFu),
                                               __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
                }
#pragma unroll
            for (int q = 0; q < kGroups; q++) {
                w[q] = nxt[q];
                wt[q] = nwt[q];
            }
        }
    }
    __syncthreads();
    uint32_t *row = a.rows + (uint64_t)ch.row * kPW;
    for (uint32_t i = threadIdx.x; i < kPW; i += kAccBlock) row[i] = s_min[i];
}

// One workgroup per unit: each vertex's minimum over its partition's rows;
// a vertex whose minimum is below its value takes it and goes to the output
// flags. As in a pull round, the output entry gets only `any` and the
// smallest value put into it; the next pre-pass counts the flags.
__global__ __launch_bounds__(kAccBlock) void k_sw_fin(SwArgs a) {
    __shared__ uint32_t s_part[kAccWaves][kTile];
    __shared__ uint32_t s_any[kAccWaves], s_fmin[kAccWaves];
    if (!sw_active(a)) return;
    const uint32_t wave = threadIdx.x >> 6, lane = lane_id();
    bool improved = false;
    uint32_t fmin_inv = 0;
    // a fixed grid walks the units (an idle launch, in a round that does not
    // sweep, then costs a few workgroups, not one per unit)
    for (uint32_t ui = blockIdx.x; ui < a.n_units; ui += gridDim.x) {
        const PrbUnit u = a.units[ui];
        const uint32_t off0 = u.v0 & kEntDest;
        uint32_t m = 0xFFFFFFFFu;
        if (u.nrows > kRowsDirect) {
            for (uint32_t r = wave; r < u.nrows; r += kAccWaves)
                m = min(m, a.rows[(uint64_t)(u.row0 + r) * kPW + off0 + lane]);
            s_part[wave][lane] = m;
            __syncthreads();
            m = 0xFFFFFFFFu;
            if (threadIdx.x < kTile)
                for (uint32_t w = 0; w < kAccWaves; w++) m = min(m, s_part[w][threadIdx.x]);
            __syncthreads();  // s_part is rewritten by the next unit
        } else if (threadIdx.x < u.nv) {
            for (uint32_t r = 0; r < u.nrows; r++) m = min(m, a.rows[(uint64_t)(u.row0 + r) * kPW + off0 + threadIdx.x]);
        }
        if (threadIdx.x < u.nv) {
            const uint32_t v = u.v0 + threadIdx.x;
            if (m < a.val[v]) {
                a.val[v] = m;
                a.e.flout[v] = 1;
                improved = true;
                fmin_inv = max(fmin_inv, ~m);
            }
        }
    }
    const uint32_t any = ballot(improved) ? 1u : 0u;
    const uint32_t fm = wave_max_u32(fmin_inv);
    if (lane == 0) {
        s_any[wave] = any;
        s_fmin[wave] = fm;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t ta = 0, tm = 0;
        for (uint32_t w = 0; w < kAccWaves; w++) {
            ta |= s_any[w];
            tm = max(tm, s_fmin[w]);
        }
        RoundCtr *out = a.e.ring + (a.e.r + 1) % kRing;
        if (ta) out->any = 1u;
        if (tm && tm > out->fmin_inv) atomicMax(&out->fmin_inv, tm);
        if (blockIdx.x == 0) (a.e.ring + a.e.r % kRing)->swept = 1u;
    }
}

int sssp_sweep_prepare(ptg_graph *g, bool *ok) {
    *ok = false;
    if (!g->weighted || g->cfg.sssp_sweep < 0 || g->wmax > 255 || g->E == 0) return PTG_OK;
    if (!g->prb) g->prb = new PrBins();
    PrBins *b = g->prb;
    // refused once, refused for good: the automatic layout's rule (pairs at
    // most half the edges), whether a forced PageRank layout exists or not
    if (b->sweep_refused) return PTG_OK;
    if (b->built && (!b->usable || 2 * b->h.n_pairs > g->E)) {
        b->sweep_refused = true;
        return PTG_OK;
    }
    int rc;
    if (!b->built || !b->h.has_w) {
        // (re)built with the weight bytes; the entries and pairs come out the
        // same. A refusal builds nothing and leaves the automatic layout's
        // answer (not usable), which PageRank reads the same way.
        free_device(g, b);
        bool usable = false;
        if ((rc = build_from_device(g, true, &usable, true))) return rc;
        if (!usable) {
            b->sweep_refused = true;
            if (!b->built) {
                b->built = true;
                b->usable = false;
            }
            return PTG_OK;
        }
        b->built = true;
        b->usable = true;
    }
    // sweeps are optional: without room in the device budget the rounds push
    // and pull
    if (!b->on_device && !fits_budget(g, b)) return PTG_OK;
    if (!b->on_device && (rc = upload(g, b))) return rc;
    *ok = true;
    return PTG_OK;
}

uint64_t sssp_sweep_need(const ptg_graph *g) {
    if (!g->weighted || g->cfg.sssp_sweep < 0 || (g->auto_window && g->wmax > 255)) return 0;
    const PrBins *b = g->prb;
    if (b && b->sweep_refused) return 0;
    if (b && b->built && b->h.has_w) return b->usable && !b->on_device ? device_bytes_of(b->h, g->V) : 0;
    // not built with weights yet: the layout's bound plus a byte per entry
    return pr_bins_need(g) + ((g->E + 4ull * ((g->V + kPW - 1) / kPW) + 255) & ~255ull);
}

int sssp_sweep_launch(ptg_graph *g, const RoundEnv &env, uint32_t *val, uint32_t sent, float pull_factor,
                      uint64_t min_edges, LaunchTimer &timer) {
    const PrBins *b = g->prb;
    if (!b || !b->on_device || !b->h.has_w) return fail(PTG_ERR_UNSUPPORTED, "sweep round without its layout");
    SwArgs a;
    memset(&a, 0, sizeof a);
    a.e = env;
    a.ent = b->ent;
    a.went = b->went;
    a.seg_cur = b->seg_cur;
    a.chunks = b->chunks;
    a.units = b->units;
    a.bsrc = b->bsrc;
    a.rows = reinterpret_cast<uint32_t *>(b->rows);
    a.val = val;
    a.sent = sent;
    a.last_pair = (uint32_t)(b->h.n_pairs ? b->h.n_pairs - 1 : 0);
    a.n_vert = g->V;
    a.n_units = (uint32_t)b->h.units.size();
    a.factor = pull_factor;
    a.min_edges = min_edges;
    const uint32_t nc = (uint32_t)b->h.chunks.size(), nu = (uint32_t)b->h.units.size();
    timer.launch_rider(k_sw_acc, dim3(nc), dim3(kAccBlock), a);
    timer.launch_rider(k_sw_fin, dim3(std::min(nu, 2u * (uint32_t)std::max(1, g->num_cus))), dim3(kAccBlock), a);
    PTG_HIP(hipGetLastError());
    return PTG_OK;
}

int pr_bins_info(const ptg_graph *g, uint64_t *pairs, uint32_t *chunks, uint32_t *rows) {
    const PrBins *b = g->prb;
    const bool ok = b && b->built && b->usable;
    if (pairs) *pairs = ok ? b->h.n_pairs : 0;
    if (chunks) *chunks = ok ? (uint32_t)b->h.chunks.size() : 0;
    if (rows) *rows = ok ? b->h.n_rows : 0;
    return PTG_OK;
}

}  // namespace ptg
