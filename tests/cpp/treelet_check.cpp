// Host check of the treelet topology search (spica_amd/csrc/spk_treelet.h): on random
// treelets of 2..7 leaves (random boxes, random subtree costs) the dynamic programme's
// cost of the whole set equals the cheapest of all binary trees over the leaves, found
// by enumerating every tree ((2n - 3)!! of them: 10395 for 7 leaves), and the tree rebuilt
// from treelet_split has that cost.
//
//   treelet_check <trials>    prints one summary line; exits non-zero on any mismatch
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <random>
#include <vector>

#include "../../spica_amd/csrc/spk_treelet.h"

using namespace spk;

// SAH cost of a tree given as nested splits: every binary tree over the set s is reached by
// choosing the part holding the lowest leaf and recursing on both parts.
static void all_costs(uint32_t s, const TreeletBox* box, const float* lc, float cn, std::vector<float>& out) {
    if ((s & (s - 1u)) == 0u) {
        int i = 0;
        while (!(s >> i & 1u)) i++;
        out.assign(1, lc[i]);
        return;
    }
    const float a = cn * treelet_half_area(treelet_subset_box(box, s));
    const uint32_t low = s & (0u - s), rest = s ^ low;
    out.clear();
    for (uint32_t q = (rest - 1u) & rest;; q = (q - 1u) & rest) {
        const uint32_t p = low | q;
        std::vector<float> l, r;
        all_costs(p, box, lc, cn, l);
        all_costs(s ^ p, box, lc, cn, r);
        for (float x : l)
            for (float y : r) out.push_back(a + x + y);
        if (q == 0u) break;
    }
}

// Cost of the tree rebuilt from the recorded splits, as the builder rebuilds it.
static double rebuilt_cost(uint32_t s, const TreeletBox* box, const float* lc, float cn, const std::vector<float>& c,
                           int& internal) {
    if ((s & (s - 1u)) == 0u) {
        int i = 0;
        while (!(s >> i & 1u)) i++;
        return lc[i];
    }
    internal++;
    const uint32_t p = treelet_split(s, [&](uint32_t t) -> float { return c[t]; });
    return cn * treelet_half_area(treelet_subset_box(box, s)) + rebuilt_cost(p, box, lc, cn, c, internal) +
           rebuilt_cost(s ^ p, box, lc, cn, c, internal);
}

int main(int argc, char** argv) {
    const int trials = argc > 1 ? atoi(argv[1]) : 200;
    std::mt19937 rng(12345);
    std::uniform_real_distribution<float> u(0.0f, 1.0f);
    long bad = 0, trees = 0;
    double worst = 0.0;
    for (int t = 0; t < trials; t++) {
        const int nl = 2 + t % (kTreeletLeaves - 1);
        TreeletBox box[kTreeletLeaves];
        float lc[kTreeletLeaves];
        for (int i = 0; i < nl; i++) {
            for (int k = 0; k < 3; k++) {
                const float c = 10.0f * u(rng), e = (t % 3 == 0 ? 0.01f : 1.0f) * u(rng);
                box[i].lo[k] = c - e;
                box[i].hi[k] = c + e;
            }
            lc[i] = 4.0f * u(rng) * treelet_half_area(box[i]);
        }
        const float cn = 0.5f + u(rng);
        std::vector<float> cost(kTreeletSubsets, 0.0f);
        treelet_costs(box, lc, nl, cn, [&](uint32_t s) -> float& { return cost[s]; });
        const uint32_t full = (1u << nl) - 1u;
        std::vector<float> all;
        all_costs(full, box, lc, cn, all);
        trees += (long)all.size();
        float best = 3.0e38f;
        for (float x : all) best = std::fmin(best, x);
        int internal = 0;
        const double rb = rebuilt_cost(full, box, lc, cn, cost, internal);
        const double e1 = std::fabs(cost[full] - best) / best, e2 = std::fabs(rb - best) / best;
        worst = std::fmax(worst, std::fmax(e1, e2));
        if (e1 > 1e-5 || e2 > 1e-5 || internal != nl - 1) bad++;
    }
    printf("treelets %d trees %ld mismatches %ld worst rel %.3g\n", trials, trees, bad, worst);
    return bad ? 1 : 0;
}
