// Lock-free union-find with minimum-key roots, shared by the resident CC
// (cc.hip) and the streamed CC (stream.hip).
#pragma once

#include "device_util.h"

namespace ptg {

// Root search with path halving. A parent always has a smaller key than its
// child, so any value read, even a stale one, is an ancestor and the walk
// terminates.
__device__ __forceinline__ uint32_t uf_root(uint32_t *par, uint32_t x) {
    for (;;) {
        const uint32_t p = par[x];
        if (p == x) return x;
        const uint32_t gp = par[p];
        if (gp == p) return p;
        par[x] = gp;
        x = gp;
    }
}

__device__ __forceinline__ uint32_t uf_key(const uint32_t *key, uint32_t x) { return key ? key[x] : x; }

// Links the two trees, always hanging the root with the larger key under the
// other one, so every root is the minimum-key vertex of its tree.
__device__ __forceinline__ void uf_unite(uint32_t *par, const uint32_t *key, uint32_t a, uint32_t b) {
    uint32_t ra = uf_root(par, a), rb = uf_root(par, b);
    while (ra != rb) {
        if (uf_key(key, ra) < uf_key(key, rb)) {
            const uint32_t t = ra;
            ra = rb;
            rb = t;
        }
        const uint32_t seen = atomicCAS(&par[ra], ra, rb);
        if (seen == ra) return;
        ra = uf_root(par, seen);
        rb = uf_root(par, rb);
    }
}

}  // namespace ptg
