// spk_treelet.h -- the optimal topology of a small treelet (Karras and Aila, "Fast parallel
// construction of high-quality bounding volume hierarchies", HPG 2013, section 4).
//
// A treelet is a node and up to kTreeletLeaves of its descendants taken as leaves (each
// leaf a whole subtree with a known SAH cost).  Over the subsets S of those leaves
//     cost(S) = c_node * area(S) + min over splits {P, S \ P} of cost(P) + cost(S \ P)
// with cost({leaf}) = that subtree's own cost and area(S) the half surface of the union
// of the leaves' boxes.  Subsets are visited in increasing numeric order, so every proper
// subset is done before its superset; each unordered split is tried once (P holds the
// lowest leaf of S).  The host compiles the same code for the exhaustive check in tests/.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace spk {

constexpr int kTreeletLeaves = 7;
constexpr int kTreeletSubsets = 1 << kTreeletLeaves;

struct TreeletBox {
    float lo[3], hi[3];
};

__host__ __device__ inline float treelet_half_area(const TreeletBox& b) {
    const float dx = b.hi[0] - b.lo[0], dy = b.hi[1] - b.lo[1], dz = b.hi[2] - b.lo[2];
    return dx * dy + dy * dz + dz * dx;
}

__host__ __device__ inline void treelet_union(TreeletBox& a, const TreeletBox& b) {
    for (int k = 0; k < 3; k++) {
        a.lo[k] = fminf(a.lo[k], b.lo[k]);
        a.hi[k] = fmaxf(a.hi[k], b.hi[k]);
    }
}

__host__ __device__ inline TreeletBox treelet_subset_box(const TreeletBox* box, uint32_t s) {
    TreeletBox u = {{3.0e38f, 3.0e38f, 3.0e38f}, {-3.0e38f, -3.0e38f, -3.0e38f}};
    for (int i = 0; i < kTreeletLeaves; i++)
        if (s >> i & 1u) treelet_union(u, box[i]);
    return u;
}

// Fills cost(S) for every non-empty subset of the nl leaves (box[i], leaf_cost[i]);
// `cost(s)` is a reference accessor into caller storage of kTreeletSubsets floats.
template <class Cost>
__host__ __device__ inline void treelet_costs(const TreeletBox* box, const float* leaf_cost, int nl, float c_node,
                                              Cost cost) {
    const uint32_t full = (1u << nl) - 1u;
    for (uint32_t s = 1; s <= full; s++) {
        if ((s & (s - 1u)) == 0u) {
            int i = 0;
            while (!(s >> i & 1u)) i++;
            cost(s) = leaf_cost[i];
            continue;
        }
        const uint32_t low = s & (0u - s);
        float best = 3.0e38f;
        // proper subsets p of s holding its lowest leaf: p = low | q, q a subset of s \ low
        const uint32_t rest = s ^ low;
        for (uint32_t q = (rest - 1u) & rest;; q = (q - 1u) & rest) {
            const uint32_t p = low | q;
            best = fminf(best, cost(p) + cost(s ^ p));
            if (q == 0u) break;
        }
        cost(s) = c_node * treelet_half_area(treelet_subset_box(box, s)) + best;
    }
}

// The split of subset s (at least two leaves) that treelet_costs chose: the part holding
// the lowest leaf.
template <class Cost>
__host__ __device__ inline uint32_t treelet_split(uint32_t s, Cost cost) {
    const uint32_t low = s & (0u - s), rest = s ^ low;
    float best = 3.0e38f;
    uint32_t arg = low;
    for (uint32_t q = (rest - 1u) & rest;; q = (q - 1u) & rest) {
        const uint32_t p = low | q;
        const float c = cost(p) + cost(s ^ p);
        if (c < best) {
            best = c;
            arg = p;
        }
        if (q == 0u) break;
    }
    return arg;
}

}  // namespace spk
