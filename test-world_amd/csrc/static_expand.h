// Expansion over a fixed vertex set binned by degree (static bins built when
// the graph is created): CC hooking and the symmetry checks use it.
#pragma once

#include <type_traits>

#include "device_util.h"

namespace ptg {

// Ops with a Pre type split an edge into load() (reads only) and apply();
// the walks below then issue kStaticGroup edges' loads before applying any
// (a value read before another edge of the group is applied can only be
// larger than the current one, so at most more candidates reach the op's
// atomic, which decides).
constexpr uint32_t kStaticGroup = 4;
template <class T, class = void>
struct HasPre : std::false_type {};
template <class T>
struct HasPre<T, std::void_t<typename T::Pre>> : std::true_type {};

// op.active(sv) when the op defines it (vertices it returns false for are
// skipped), else true.
template <class Op, class T>
__device__ __forceinline__ auto op_active(const Op &op, T sv, int) -> decltype(op.active(sv)) {
    return op.active(sv);
}
template <class Op, class T>
__device__ __forceinline__ bool op_active(const Op &, T, long) {
    return true;
}

// Walks the L (hub chunks of kChunkEdges, one workgroup each), M (one wave
// per vertex) and S (one lane per vertex) bins and calls op.edge() from
// wave-uniform control flow. op.vertex(v) gives the value passed to edge().
template <class Op>
__device__ __forceinline__ void expand_static(const Frontier &in, const uint32_t *__restrict__ off, const Op &op) {
    const uint32_t lane = lane_id();
    const uint32_t wib = threadIdx.x >> 6;
    const unsigned long long lp = in.ctr->l_pack;
    const uint32_t nl = (uint32_t)(lp >> kLCountShift);
    const uint32_t nch = (uint32_t)(lp & kLChunkMask);
    for (uint32_t c = blockIdx.x; c < nch; c += gridDim.x) {
        uint32_t lo = 0, hi = nl;
        while (hi - lo > 1) {
            const uint32_t mid = (lo + hi) >> 1;
            if (in.lb[mid] <= c) lo = mid; else hi = mid;
        }
        const uint32_t v = in.ql[lo];
        const uint64_t b = (uint64_t)off[v] + (uint64_t)(c - in.lb[lo]) * kChunkEdges;
        const uint64_t e = min((uint64_t)off[v + 1], b + kChunkEdges);
        const auto sv = op.vertex(v);
        if (!op_active(op, sv, 0)) continue;
        if constexpr (HasPre<Op>::value) {
            for (uint64_t base = b + wib * 64; base < e; base += (uint64_t)kBlock * kStaticGroup) {
                typename Op::Pre pre[kStaticGroup];
#pragma unroll
                for (uint32_t j = 0; j < kStaticGroup; j++) {
                    const uint64_t i = base + (uint64_t)j * kBlock + lane;
                    pre[j] = op.load(i < e, sv, i);
                }
#pragma unroll
                for (uint32_t j = 0; j < kStaticGroup; j++) op.apply(pre[j]);
            }
        } else {
            for (uint64_t base = b + wib * 64; base < e; base += kBlock) op.edge(base + lane < e, sv, base + lane);
        }
    }
    const uint32_t gw = blockIdx.x * kWavesPerBlock + wib;
    const uint32_t nw = gridDim.x * kWavesPerBlock;
    const uint32_t nm = in.ctr->n_m;
    for (uint32_t i = gw; i < nm; i += nw) {
        const uint32_t v = in.qm[i];
        const auto sv = op.vertex(v);
        if (!op_active(op, sv, 0)) continue;
        const uint64_t b = off[v], e = off[v + 1];
        if constexpr (HasPre<Op>::value) {
            for (uint64_t base = b; base < e; base += 64 * kStaticGroup) {
                typename Op::Pre pre[kStaticGroup];
#pragma unroll
                for (uint32_t j = 0; j < kStaticGroup; j++) {
                    const uint64_t i = base + 64 * j + lane;
                    pre[j] = op.load(i < e, sv, i);
                }
#pragma unroll
                for (uint32_t j = 0; j < kStaticGroup; j++) op.apply(pre[j]);
            }
        } else {
            for (uint64_t base = b; base < e; base += 64) op.edge(base + lane < e, sv, base + lane);
        }
    }
    const uint32_t ns = in.ctr->n_s;
    for (uint64_t base = (uint64_t)gw * 64; base < ns; base += (uint64_t)nw * 64) {
        const uint64_t i = base + lane;
        const bool valid = i < ns;
        uint32_t b = 0, d = 0;
        decltype(op.vertex(0u)) sv{};
        if (valid) {
            const uint32_t v = in.qs[i];
            sv = op.vertex(v);
            if (op_active(op, sv, 0)) {
                b = off[v];
                d = off[v + 1] - b;
            }
        }
        const uint32_t md = wave_max_u32(d);
        if constexpr (HasPre<Op>::value) {
            for (uint32_t j = 0; j < md; j += kStaticGroup) {
                typename Op::Pre pre[kStaticGroup];
#pragma unroll
                for (uint32_t t = 0; t < kStaticGroup; t++) pre[t] = op.load(j + t < d, sv, (uint64_t)b + j + t);
#pragma unroll
                for (uint32_t t = 0; t < kStaticGroup; t++) op.apply(pre[t]);
            }
        } else {
            for (uint32_t j = 0; j < md; j++) op.edge(j < d, sv, (uint64_t)b + j);
        }
    }
}

template <class Op>
__global__ __launch_bounds__(kBlock) void k_expand(Frontier in, const uint32_t *__restrict__ off, Op op) {
    expand_static(in, off, op);
}

}  // namespace ptg
