// Wave64 primitives and the binned-frontier append used by every HIP kernel.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "ptg_internal.h"

namespace ptg {

__device__ __forceinline__ uint32_t lane_id() {
    return __builtin_amdgcn_mbcnt_hi(~0u, __builtin_amdgcn_mbcnt_lo(~0u, 0u));
}

// Number of set bits of `m` below this lane.
__device__ __forceinline__ uint32_t rank_below(uint64_t m) {
    return __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
}

__device__ __forceinline__ uint64_t ballot(bool p) { return __ballot(p ? 1 : 0); }

__device__ __forceinline__ uint32_t lowest_lane(uint64_t m) { return (uint32_t)__builtin_ctzll(m); }

__device__ __forceinline__ uint32_t read_lane(uint32_t x, uint32_t lane) {
    return (uint32_t)__shfl((int)x, (int)lane, 64);
}

__device__ __forceinline__ uint64_t read_lane64(uint64_t x, uint32_t lane) {
    uint32_t lo = read_lane((uint32_t)x, lane), hi = read_lane((uint32_t)(x >> 32), lane);
    return (uint64_t)hi << 32 | lo;
}

// Orders the wave's LDS accesses: writes before the call are visible to
// reads after it by any lane of the same wave.
__device__ __forceinline__ void wave_lds_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Wave-wide sums / max; every lane of the wave must be active.
__device__ __forceinline__ uint32_t wave_max_u32(uint32_t x) {
    for (int d = 32; d >= 1; d >>= 1) {
        uint32_t y = (uint32_t)__shfl_xor((int)x, d, 64);
        x = x > y ? x : y;
    }
    return x;
}

__device__ __forceinline__ uint32_t wave_min_u32(uint32_t x) { return ~wave_max_u32(~x); }

__device__ __forceinline__ uint64_t wave_max_u64(uint64_t x) {
    for (int d = 32; d >= 1; d >>= 1) {
        const uint32_t lo = (uint32_t)__shfl_xor((int)(uint32_t)x, d, 64);
        const uint32_t hi = (uint32_t)__shfl_xor((int)(uint32_t)(x >> 32), d, 64);
        const uint64_t y = (uint64_t)hi << 32 | lo;
        x = x > y ? x : y;
    }
    return x;
}

__device__ __forceinline__ uint64_t wave_sum_u64(uint64_t x) {
    for (int d = 32; d >= 1; d >>= 1) {
        uint32_t lo = (uint32_t)__shfl_xor((int)(uint32_t)x, d, 64);
        uint32_t hi = (uint32_t)__shfl_xor((int)(uint32_t)(x >> 32), d, 64);
        x += (uint64_t)hi << 32 | lo;
    }
    return x;
}

// Inclusive prefix sum across the wave.
__device__ __forceinline__ uint32_t wave_incl_scan_u32(uint32_t x) {
    const uint32_t lane = lane_id();
    for (int d = 1; d < 64; d <<= 1) {
        uint32_t y = (uint32_t)__shfl_up((int)x, d, 64);
        if (lane >= (uint32_t)d) x += y;
    }
    return x;
}

// Appends `item` for every lane with `pred`, one atomic per wave.
__device__ __forceinline__ void wave_append(uint32_t *q, uint32_t *cnt, bool pred, uint32_t item) {
    const uint64_t m = ballot(pred);
    if (m == 0) return;
    const uint32_t leader = lowest_lane(m);
    uint32_t base = 0;
    if (lane_id() == leader) base = atomicAdd(cnt, (uint32_t)__popcll(m));
    base = read_lane(base, leader);
    if (pred) q[base + rank_below(m)] = item;
}

__device__ __forceinline__ uint32_t bin_chunks(uint32_t deg) {
    return (deg + kChunkEdges - 1) / kChunkEdges;
}

// Appends vertex `v` (out-degree `d`, d > 0 when pred) to the S/M/L bin that
// matches its degree. The L bin stores (v, first chunk index) with one packed
// 64-bit atomic per wave so that chunk bases grow with the slot index.
// Must be called by every active lane of the wave (wave-uniform call site).
__device__ __forceinline__ void push_binned(const Frontier &out, bool pred, uint32_t v, uint32_t d,
                                            uint32_t small_deg, uint32_t large_deg) {
    const bool ps = pred && d <= small_deg;
    const bool pm = pred && d > small_deg && d <= large_deg;
    const bool pl = pred && d > large_deg;
    wave_append(out.qs, &out.ctr->n_s, ps, v);
    wave_append(out.qm, &out.ctr->n_m, pm, v);
    const uint64_t ml = ballot(pl);
    if (ml) {
        const uint32_t nch = pl ? bin_chunks(d) : 0u;
        const uint32_t incl = wave_incl_scan_u32(nch);
        const uint32_t total = read_lane(incl, 63);
        const uint32_t leader = lowest_lane(ml);
        unsigned long long old = 0;
        if (lane_id() == leader)
            old = atomicAdd(&out.ctr->l_pack,
                            ((unsigned long long)__popcll(ml) << kLCountShift) | (unsigned long long)total);
        old = read_lane64(old, leader);
        if (pl) {
            const uint32_t slot = (uint32_t)(old >> kLCountShift) + rank_below(ml);
            out.ql[slot] = v;
            out.lb[slot] = (uint32_t)(old & kLChunkMask) + (incl - nch);
        }
    }
}

// Adds `pred` lanes to found/edges counters (one atomic pair per wave).
__device__ __forceinline__ void count_found(BinCounters *ctr, bool pred, uint32_t d) {
    const uint64_t m = ballot(pred);
    if (m == 0) return;
    const uint64_t s = wave_sum_u64(pred ? (uint64_t)d : 0ull);
    if (lane_id() == lowest_lane(m)) {
        atomicAdd(&ctr->found, (uint32_t)__popcll(m));
        atomicAdd(&ctr->edges, (unsigned long long)s);
    }
}


// PageRank sums as unsigned 64-bit fixed point with 63 fraction bits
// (Q1.63). Each FP32 term in [0, 1] is truncated to a multiple of 2^-63 and
// added exactly, so a sum does not depend on the order or grouping of its
// terms. The resident and streamed PageRank kernels both round through
// pr_base_value and pr_rank, so they give the same bits.
__device__ __forceinline__ unsigned long long fx63(float x) {
    const uint32_t b = __float_as_uint(x);
    const int e = (int)((b >> 23) & 0xFFu);
    if (e == 0) return 0;  // zero and subnormals
    const unsigned long long m = (b & 0x7FFFFFu) | 0x800000u;
    const int sh = e - 87;  // x = m * 2^(e - 150), so x * 2^63 = m * 2^(e - 87); e <= 127
    return sh >= 0 ? m << sh : (sh > -64 ? m >> -sh : 0ull);
}

__device__ __forceinline__ double fx63_to_f64(unsigned long long s) { return (double)s * 0x1p-63; }

// base = (1-d)/V + d * D/V in FP64, rounded once to FP32.
__device__ __forceinline__ float pr_base_value(float damping, float inv_v, double dang) {
    return (float)((1.0 - (double)damping) * (double)inv_v + (double)damping * dang * (double)inv_v);
}

// r = base + d * sum in FP64 from the exact sum, rounded once to FP32.
__device__ __forceinline__ float pr_rank(float base, float damping, unsigned long long sum) {
    return (float)((double)base + (double)damping * fx63_to_f64(sum));
}

}  // namespace ptg
