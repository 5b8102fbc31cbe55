// Degree-ordered vertex relabeling (relabel.cpp).
#pragma once

#include <stdint.h>

#include <vector>

#include "ptg.h"

namespace ptg {

struct Relabeled {
    std::vector<uint32_t> perm;  // old id -> new id
    std::vector<uint32_t> inv;   // new id -> old id
    std::vector<uint32_t> off;   // V + 1 offsets in new ids
    std::vector<uint32_t> col;   // E columns in new ids, each list ascending
    std::vector<uint32_t> w;     // E weights (empty for an unweighted graph)
};

// Relabels a reference-layout CSR (V-entry `ptr`) by descending out-degree.
// Exactly one of cols / wedges is used; weights may be null.
void relabel_by_degree(uint32_t V, uint64_t E, const uint32_t *ptr, const uint32_t *cols, const uint32_t *weights,
                       const EdgeWithWeight *wedges, Relabeled *out);
// hub_degree == 0: by degree as above; otherwise vertices of out-degree >
// hub_degree first by degree, then the others in breadth-first order from
// them, vertices without out-edges last.
void relabel(uint32_t V, uint64_t E, const uint32_t *ptr, const uint32_t *cols, const uint32_t *weights,
             const EdgeWithWeight *wedges, uint32_t hub_degree, Relabeled *out);

}  // namespace ptg
