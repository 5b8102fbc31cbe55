// Multiset fingerprint of a graph's edge pairs, for the symmetry test.
//
// A CSR is symmetric as a multigraph when the multiset of its pairs (v, u)
// equals the multiset of the reversed pairs (u, v), repeats counted. Both
// multisets are mapped to products over GF(p), p = 2^61 - 1:
//     F = prod over edges (v, u) of (r - v - s*u),   B = the same over (u, v).
// Distinct pairs with v, u < 2^32 < p give distinct linear factors in (r, s),
// so two different multisets of E pairs give different polynomials of degree
// E, and for keys r, s drawn uniformly at random they agree with probability
// at most E / p (Schwartz-Zippel): 2^-30 at E = 2^31. Two independent key
// pairs are used, so a non-symmetric graph passes with probability at most
// (E / p)^2. A symmetric graph always passes, with any keys.
//
// Weights are symmetric when the multiset of triples (v, u, w) equals that of
// (u, v, w); the factors are then (r - v - s*u - t*w), linear in (r, s, t)
// and distinct for distinct triples, so the same bound holds.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace ptg {

constexpr uint64_t kFpP = (1ull << 61) - 1;

__host__ __device__ __forceinline__ uint64_t fp_reduce(uint64_t x) {
    x = (x & kFpP) + (x >> 61);
    return x >= kFpP ? x - kFpP : x;
}

// a * b mod p for a, b < p.
__host__ __device__ __forceinline__ uint64_t fp_mul(uint64_t a, uint64_t b) {
#ifdef __HIP_DEVICE_COMPILE__
    const uint64_t lo = a * b, hi = __umul64hi(a, b);
#else
    const unsigned __int128 t = (unsigned __int128)a * b;
    const uint64_t lo = (uint64_t)t, hi = (uint64_t)(t >> 64);
#endif
    // a * b < 2^122: low 61 bits plus the rest, which is < 2^61
    return fp_reduce((lo & kFpP) + ((lo >> 61) | (hi << 3)));
}

// r - v - s*u mod p.
__host__ __device__ __forceinline__ uint64_t fp_term(uint64_t r, uint64_t s, uint32_t v, uint32_t u) {
    const uint64_t su = fp_mul(s, u);
    const uint64_t t = fp_reduce(su + v);  // < 2^61 + 2^32: one reduction suffices
    return r >= t ? r - t : r + kFpP - t;
}

// r - v - s*u - t*w mod p.
__host__ __device__ __forceinline__ uint64_t fp_term_w(uint64_t r, uint64_t s, uint64_t t, uint32_t v, uint32_t u,
                                                       uint32_t w) {
    const uint64_t x = fp_reduce(fp_reduce(fp_mul(s, u) + fp_mul(t, w)) + v);  // each sum < 2^62
    return r >= x ? r - x : r + kFpP - x;
}

// Two key sets; accumulated products forward (v, u) and backward (u, v).
struct FpKeys {
    uint64_t r[2], s[2], t[2];
};

struct FpAcc {
    uint64_t f[2], b[2];
};

__host__ __device__ __forceinline__ void fp_edge(const FpKeys &k, FpAcc &a, uint32_t v, uint32_t u) {
#pragma unroll
    for (int i = 0; i < 2; i++) {
        a.f[i] = fp_mul(a.f[i], fp_term(k.r[i], k.s[i], v, u));
        a.b[i] = fp_mul(a.b[i], fp_term(k.r[i], k.s[i], u, v));
    }
}

__host__ __device__ __forceinline__ void fp_edge_w(const FpKeys &k, FpAcc &a, uint32_t v, uint32_t u, uint32_t w) {
#pragma unroll
    for (int i = 0; i < 2; i++) {
        a.f[i] = fp_mul(a.f[i], fp_term_w(k.r[i], k.s[i], k.t[i], v, u, w));
        a.b[i] = fp_mul(a.b[i], fp_term_w(k.r[i], k.s[i], k.t[i], u, v, w));
    }
}

// Fresh random keys in [1, p) (host only).
FpKeys fp_random_keys();

}  // namespace ptg
