// lbvh.hip -- linear BVH built on the GPU (Karras, "Maximizing parallelism in the
// construction of BVHs, octrees, and k-d trees", HPG 2012).
//
//   1. 63-bit Morton code (21 bits per axis) of each triangle centroid in the scene box,
//      sorted with its index (hipCUB radix sort); 30-bit codes gave many equal keys on
//      scenes of small, dense meshes, and equal keys split by index, not by space;
//   2. one thread per internal node finds the node's key range and split from the
//      common-prefix lengths of neighbouring codes (ties broken by index), giving a
//      binary radix tree whose node i covers a contiguous range of sorted triangles;
//   3. boxes bottom-up: each leaf walks to the root, the second thread to reach a node
//      (an atomic counter per node) merges its children's boxes and computes the node's
//      SAH cost; a subtree of at most max_leaf triangles whose cost as one leaf is no
//      higher than as a subtree is marked to collapse (the rule of Karras and Aila, "Fast
//      parallel construction of high-quality BVHs", HPG 2013);
//   4. treelet restructuring (Karras and Aila 2013): bottom-up sweeps in which every node
//      of at least kTreeletMinTris triangles takes the treelet of its up to 7 largest
//      descendants and rebuilds the treelet's inner topology with the cheapest SAH cost
//      (spk_treelet.h), followed by a top-down pass, one level per launch, that gives every
//      subtree its contiguous range of the final triangle order (depth first);
//   5. boxes, SAH costs and the leaf collapse of step 3 again over the restructured tree;
//   6. emit binary nodes (collapsed subtrees as leaves over their contiguous ranges), then
//      renumber the reachable nodes depth first on the host; the scene then collapses
//      them into 4-wide nodes like the SAH tree's.
// Node 0 is the root, as the traversal expects.  The tree answers the same closest hits
// as the SAH tree (only exact-tie winners may differ); it is cheaper to build and somewhat
// slower to traverse.
#include <hip/hip_runtime.h>
#include <rocprim/device/device_radix_sort.hpp>

#include <algorithm>
#include <cstring>
#include <vector>

#include "spk_internal.h"
#include "spk_treelet.h"

namespace spk {

namespace {

// Spreads the low 21 bits of v to every third bit of a 64-bit word.
__device__ __forceinline__ uint64_t expand21(uint32_t v) {
    uint64_t x = v & 0x1fffffu;
    x = (x | (x << 32)) & 0x1f00000000ffffull;
    x = (x | (x << 16)) & 0x1f0000ff0000ffull;
    x = (x | (x << 8)) & 0x100f00f00f00f00full;
    x = (x | (x << 4)) & 0x10c30c30c30c30c3ull;
    x = (x | (x << 2)) & 0x1249249249249249ull;
    return x;
}

__global__ void k_morton(const float* __restrict__ v, uint32_t n, float3 lo, float3 scale, uint64_t* codes,
                         uint32_t* idx) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const float* t = v + 9 * (size_t)i;
        const float cx = (t[0] + t[3] + t[6]) * (1.0f / 3.0f);
        const float cy = (t[1] + t[4] + t[7]) * (1.0f / 3.0f);
        const float cz = (t[2] + t[5] + t[8]) * (1.0f / 3.0f);
        const uint32_t x = (uint32_t)fminf(fmaxf((cx - lo.x) * scale.x, 0.0f), 2097151.0f);
        const uint32_t y = (uint32_t)fminf(fmaxf((cy - lo.y) * scale.y, 0.0f), 2097151.0f);
        const uint32_t z = (uint32_t)fminf(fmaxf((cz - lo.z) * scale.z, 0.0f), 2097151.0f);
        codes[i] = (expand21(x) << 2) | (expand21(y) << 1) | expand21(z);
        idx[i] = i;
    }
}

// Common-prefix length of the keys at i and j (the index breaks ties), -1 out of range.
__device__ __forceinline__ int delta(const uint64_t* __restrict__ c, int n, int i, int j) {
    if (j < 0 || j >= n) return -1;
    const uint64_t a = c[i], b = c[j];
    if (a == b) return 64 + __clz((uint32_t)i ^ (uint32_t)j);
    return __clzll((long long)(a ^ b));
}

// Internal nodes 0 .. n-2; child ids < n-1 are internal nodes, id (n-1) + k is leaf k.
__global__ void k_hierarchy(const uint64_t* __restrict__ c, int n, int2* child, int2* range, int* parent) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n - 1; i += gridDim.x * blockDim.x) {
        const int d = (delta(c, n, i, i + 1) - delta(c, n, i, i - 1)) >= 0 ? 1 : -1;
        const int dmin = delta(c, n, i, i - d);
        int lmax = 2;
        while (delta(c, n, i, i + lmax * d) > dmin) lmax *= 2;
        int l = 0;
        for (int t = lmax / 2; t >= 1; t /= 2)
            if (delta(c, n, i, i + (l + t) * d) > dmin) l += t;
        const int j = i + l * d;
        const int dnode = delta(c, n, i, j);
        int s = 0, t = l;
        do {
            t = (t + 1) / 2;
            if (delta(c, n, i, i + (s + t) * d) > dnode) s += t;
        } while (t > 1);
        const int gamma = i + s * d + (d < 0 ? -1 : 0);
        const int lo = min(i, j), hi = max(i, j);
        const int left = lo == gamma ? (n - 1) + gamma : gamma;
        const int right = hi == gamma + 1 ? (n - 1) + gamma + 1 : gamma + 1;
        child[i] = make_int2(left, right);
        range[i] = make_int2(lo, hi);
        parent[left] = i;
        parent[right] = i;
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) parent[0] = -1;
}

struct BoxF {
    float lo[3], hi[3];
};

__device__ __forceinline__ float half_area(const BoxF& b) {
    const float dx = b.hi[0] - b.lo[0], dy = b.hi[1] - b.lo[1], dz = b.hi[2] - b.lo[2];
    return dx * dy + dy * dz + dz * dx;
}

// Boxes and SAH costs bottom-up.  cost = node_cost * A + cost(left) + cost(right) for an
// interior node, tri_cost * count * A for a leaf; a subtree of at most max_leaf triangles
// whose leaf cost is no higher becomes one leaf (collapse[p] = 1).
__global__ void k_refit(const float* __restrict__ v, const uint32_t* __restrict__ order, int n,
                        const int2* __restrict__ child, const int2* __restrict__ range, const int* __restrict__ parent,
                        BoxF* box, float* cost, uint32_t* collapse, uint32_t* flag, int max_leaf, float cn, float ct) {
    for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < n; k += gridDim.x * blockDim.x) {
        const float* t = v + 9 * (size_t)order[k];
        BoxF b;
        for (int a = 0; a < 3; a++) {
            b.lo[a] = fminf(fminf(t[a], t[3 + a]), t[6 + a]);
            b.hi[a] = fmaxf(fmaxf(t[a], t[3 + a]), t[6 + a]);
        }
        int node = (n - 1) + k;
        box[node] = b;
        cost[node] = ct * half_area(b);
        __threadfence();
        int p = parent[node];
        while (p >= 0) {
            // the first thread to arrive leaves; the second has both children's results
            if (atomicAdd(&flag[p], 1u) == 0u) break;
            __threadfence();
            const int2 ch = child[p];
            const BoxF l = box[ch.x], r = box[ch.y];
            BoxF m;
            for (int a = 0; a < 3; a++) {
                m.lo[a] = fminf(l.lo[a], r.lo[a]);
                m.hi[a] = fmaxf(l.hi[a], r.hi[a]);
            }
            const float area = half_area(m);
            const float c_int = cn * area + cost[ch.x] + cost[ch.y];
            const int2 rg = range[p];
            const int cnt = rg.y - rg.x + 1;
            const float c_leaf = cnt <= max_leaf ? ct * (float)cnt * area : 3.0e38f;
            box[p] = m;
            cost[p] = fminf(c_int, c_leaf);
            collapse[p] = c_leaf <= c_int ? 1u : 0u;
            __threadfence();
            p = parent[p];
        }
    }
}

// Leaf k (node (n - 1) + k, triangle order[k] of the Morton order): box, SAH cost, one
// triangle, and its position in the triangle order (the Morton position until restructured).
__global__ void k_leaves(const float* __restrict__ v, const uint32_t* __restrict__ order, int n, BoxF* box, float* cost,
                         int* size, int2* range, float ct) {
    for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < n; k += gridDim.x * blockDim.x) {
        const float* t = v + 9 * (size_t)order[k];
        BoxF b;
        for (int a = 0; a < 3; a++) {
            b.lo[a] = fminf(fminf(t[a], t[3 + a]), t[6 + a]);
            b.hi[a] = fmaxf(fmaxf(t[a], t[3 + a]), t[6 + a]);
        }
        const int node = (n - 1) + k;
        box[node] = b;
        cost[node] = ct * half_area(b);
        size[node] = 1;
        range[node] = make_int2(k, k);
    }
}

constexpr int kTreeletBlock = 64;
constexpr int kTreeletMinTris = 8;

__device__ __forceinline__ TreeletBox to_tb(const BoxF& b) {
    TreeletBox t;
    for (int a = 0; a < 3; a++) {
        t.lo[a] = b.lo[a];
        t.hi[a] = b.hi[a];
    }
    return t;
}

// Re-optimises the treelet under p (p's box, size and cost are current, as are those of every
// node below it).  Its leaves: p's children, then repeatedly the largest-area interior leaf
// replaced by its two children, up to kTreeletLeaves; its interior nodes are p and the nodes
// so replaced, and they are reused for the new topology.
template <class Cost>
__device__ void treelet_optimise(int p, int n, int2* child, int* parent, BoxF* box, float* cost, int* size, float cn,
                                 Cost C) {
    int leaf[kTreeletLeaves], inner[kTreeletLeaves - 1];
    int nl = 2, ni = 1;
    inner[0] = p;
    const int2 c0 = child[p];
    leaf[0] = c0.x;
    leaf[1] = c0.y;
    while (nl < kTreeletLeaves) {
        int best = -1;
        float ba = -1.0f;
        for (int i = 0; i < nl; i++)
            if (leaf[i] < n - 1) {
                const float a = half_area(box[leaf[i]]);
                if (a > ba) {
                    ba = a;
                    best = i;
                }
            }
        if (best < 0) break;
        const int x = leaf[best];
        inner[ni++] = x;
        const int2 cx = child[x];
        leaf[best] = cx.x;
        leaf[nl++] = cx.y;
    }
    if (nl < 3) return;  // two leaves have one topology
    TreeletBox tb[kTreeletLeaves];
    float lc[kTreeletLeaves];
    for (int i = 0; i < nl; i++) {
        tb[i] = to_tb(box[leaf[i]]);
        lc[i] = cost[leaf[i]];
    }
    treelet_costs(tb, lc, nl, cn, C);
    const uint32_t full = (1u << nl) - 1u;
    if (!(C(full) < cost[p] * (1.0f - 1e-6f))) return;
    // new topology top-down (each split's parts go to a leaf or the next spare interior node),
    // then boxes, sizes and costs of the interior nodes, children before parents
    uint32_t set[kTreeletLeaves - 1];
    int node[kTreeletLeaves - 1];
    int head = 0, tail = 1, next = 1;
    set[0] = full;
    node[0] = p;
    while (head < tail) {
        const uint32_t s = set[head];
        const int x = node[head];
        head++;
        const uint32_t a = treelet_split(s, C);
        const uint32_t parts[2] = {a, s ^ a};
        int ids[2];
        for (int k = 0; k < 2; k++) {
            const uint32_t q = parts[k];
            if ((q & (q - 1u)) == 0u) {
                ids[k] = leaf[__ffs(q) - 1];
            } else {
                ids[k] = inner[next++];
                set[tail] = q;
                node[tail] = ids[k];
                tail++;
            }
            parent[ids[k]] = x;
        }
        child[x] = make_int2(ids[0], ids[1]);
    }
    for (int j = tail - 1; j >= 0; j--) {
        const int x = node[j];
        const int2 ch = child[x];
        const BoxF l = box[ch.x], r = box[ch.y];
        BoxF m;
        for (int a = 0; a < 3; a++) {
            m.lo[a] = fminf(l.lo[a], r.lo[a]);
            m.hi[a] = fmaxf(l.hi[a], r.hi[a]);
        }
        box[x] = m;
        size[x] = size[ch.x] + size[ch.y];
        cost[x] = cn * half_area(m) + cost[ch.x] + cost[ch.y];
    }
}

// One bottom-up sweep: the second thread to reach a node (one atomic counter per node, as in
// k_refit) has both subtrees final, refits the node and re-optimises its treelet.  Every
// node a treelet touches lies under its root, whose subtrees are finished; the fence after
// the counter makes the other subtree's writes visible.  Costs are the plain binary SAH
// (the leaf collapse is decided after restructuring).
__global__ __launch_bounds__(kTreeletBlock) void k_treelet(int n, int2* child, int* parent, BoxF* box, float* cost,
                                                          int* size, uint32_t* flag, float cn) {
    __shared__ float dp[kTreeletSubsets * kTreeletBlock];  // per thread one column of subset costs
    float* col = dp + threadIdx.x;
    auto C = [col](uint32_t s) -> float& { return col[s * kTreeletBlock]; };
    for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < n; k += gridDim.x * blockDim.x) {
        int p = parent[(n - 1) + k];
        while (p >= 0) {
            if (atomicAdd(&flag[p], 1u) == 0u) break;
            __threadfence();
            const int2 ch = child[p];
            const BoxF l = box[ch.x], r = box[ch.y];
            BoxF m;
            for (int a = 0; a < 3; a++) {
                m.lo[a] = fminf(l.lo[a], r.lo[a]);
                m.hi[a] = fmaxf(l.hi[a], r.hi[a]);
            }
            box[p] = m;
            size[p] = size[ch.x] + size[ch.y];
            cost[p] = cn * half_area(m) + cost[ch.x] + cost[ch.y];
            if (size[p] >= kTreeletMinTris) treelet_optimise(p, n, child, parent, box, cost, size, cn, C);
            __threadfence();
            p = parent[p];
        }
    }
}

// One level of the top-down pass: each frontier node hands its range start to its children
// (left first) and appends its interior children to the next frontier (one atomic per wave).
__global__ void k_ranges_level(const int* __restrict__ frontier, uint32_t nf, const int2* __restrict__ child,
                               const int* __restrict__ size, int2* range, int* next, uint32_t* next_count, int n) {
    for (uint32_t base = blockIdx.x * blockDim.x; base < nf; base += gridDim.x * blockDim.x) {
        const uint32_t i = base + threadIdx.x;
        int c[2] = {-1, -1};
        if (i < nf) {
            const int x = frontier[i];
            const int2 ch = child[x];
            const int s = range[x].x;
            const int sl = size[ch.x], sr = size[ch.y];
            range[ch.x] = make_int2(s, s + sl - 1);
            range[ch.y] = make_int2(s + sl, s + sl + sr - 1);
            c[0] = ch.x < n - 1 ? ch.x : -1;
            c[1] = ch.y < n - 1 ? ch.y : -1;
        }
        for (int k = 0; k < 2; k++) {
            const bool push = c[k] >= 0;
            const uint64_t m = __ballot(push);
            if (m == 0) continue;
            const int lane = __lane_id();
            const int leader = __ffsll((unsigned long long)m) - 1;
            uint32_t b = 0;
            if (lane == leader) b = atomicAdd(next_count, (uint32_t)__popcll(m));
            b = __shfl(b, leader);
            if (push) next[b + __popcll(m & ((1ull << lane) - 1ull))] = c[k];
        }
    }
}

// Leaf k's triangle goes to its range start in the final order.
__global__ void k_leaf_order(const uint32_t* __restrict__ morton_order, const int2* __restrict__ range, int n,
                             uint32_t* order) {
    for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < n; k += gridDim.x * blockDim.x)
        order[range[(n - 1) + k].x] = morton_order[k];
}

__device__ __forceinline__ int32_t leaf_link_dev(uint32_t first, uint32_t count) {
    return ~(int32_t)((first << kLeafCountBits) | (count - 1));
}

__global__ void k_emit(const int2* __restrict__ child, const int2* __restrict__ range, const BoxF* __restrict__ box, int n,
                       const uint32_t* __restrict__ collapse, BvhNodeHost* out) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n - 1; i += gridDim.x * blockDim.x) {
        const int2 ch = child[i];
        const BoxF l = box[ch.x], r = box[ch.y];
        BvhNodeHost h;
        h.a[0] = l.lo[0]; h.a[1] = l.hi[0]; h.a[2] = l.lo[1]; h.a[3] = l.hi[1];
        h.b[0] = r.lo[0]; h.b[1] = r.hi[0]; h.b[2] = r.lo[1]; h.b[3] = r.hi[1];
        h.c[0] = l.lo[2]; h.c[1] = l.hi[2]; h.c[2] = r.lo[2]; h.c[3] = r.hi[2];
        const int ids[2] = {ch.x, ch.y};
        for (int side = 0; side < 2; side++) {
            const int c = ids[side];
            if (c >= n - 1) {
                h.d[side] = leaf_link_dev((uint32_t)range[c].x, 1u);
            } else {
                const int2 rg = range[c];
                h.d[side] = collapse[c] ? leaf_link_dev((uint32_t)rg.x, (uint32_t)(rg.y - rg.x + 1)) : c;
            }
        }
        h.d[2] = h.d[3] = 0;
        out[i] = h;
    }
}

struct Bufs {
    DevBuf dv, codes, idx, codes_s, idx_s, tmp, child, range, parent, box, cost, collapse, flag, nodes, size, front[2],
        front_n, order;
    ~Bufs() {
        for (DevBuf* b : {&dv, &codes, &idx, &codes_s, &idx_s, &tmp, &child, &range, &parent, &box, &cost, &collapse,
                          &flag, &nodes, &size, &front[0], &front[1], &front_n, &order})
            b->release();
    }
};

// Depth-first renumbering of the reachable nodes: the left child follows its parent,
// and interior nodes inside collapsed subtrees are dropped.
void compact(const std::vector<BvhNodeHost>& in, BvhResult& res) {
    res.nodes.clear();
    res.max_depth = 0;
    res.num_leaves = 0;
    struct Item {
        int32_t id, parent, side, depth;
    };
    std::vector<Item> st;
    st.push_back({0, -1, 0, 1});
    while (!st.empty()) {
        const Item it = st.back();
        st.pop_back();
        const int32_t nid = (int32_t)res.nodes.size();
        res.nodes.push_back(in[(size_t)it.id]);
        if (it.parent >= 0) res.nodes[(size_t)it.parent].d[it.side] = nid;
        res.max_depth = std::max(res.max_depth, it.depth);
        const BvhNodeHost& h = in[(size_t)it.id];
        for (int side = 1; side >= 0; side--) {   // right pushed first: left is visited next
            if (h.d[side] < 0) res.num_leaves++;
            else st.push_back({h.d[side], nid, side, it.depth + 1});
        }
    }
}

}  // namespace

int build_bvh_gpu(const float* verts, int64_t n, const BvhBuildOpts& opts, BvhResult& res) {
    if (opts.max_leaf < 1 || opts.max_leaf > kMaxLeafTris) {
        set_error("build_bvh_gpu: max_leaf %d outside [1, %d]", opts.max_leaf, kMaxLeafTris);
        return SPK_ERR_INVALID;
    }
    if (n > (int64_t)1 << 28) {
        set_error("build_bvh_gpu: %lld triangles (at most 2^28)", (long long)n);
        return SPK_ERR_INVALID;
    }
    if (n < 2) {   // a one-triangle scene is one leaf: nothing to sort
        res = build_bvh(verts, n, opts);
        return SPK_OK;
    }
    float lo[3] = {3.0e38f, 3.0e38f, 3.0e38f}, hi[3] = {-3.0e38f, -3.0e38f, -3.0e38f};
    for (int64_t i = 0; i < n; i++)
        for (int a = 0; a < 3; a++) {
            const float c = (verts[9 * i + a] + verts[9 * i + 3 + a] + verts[9 * i + 6 + a]) * (1.0f / 3.0f);
            lo[a] = std::min(lo[a], c);
            hi[a] = std::max(hi[a], c);
        }
    float scale[3];
    for (int a = 0; a < 3; a++) scale[a] = hi[a] > lo[a] ? 2097152.0f / (hi[a] - lo[a]) : 0.0f;
    const int N = (int)n;
    Bufs B;
    int rc;
    if ((rc = B.dv.upload(verts, (size_t)n * 9 * 4)) || (rc = B.codes.alloc((size_t)n * 8)) ||
        (rc = B.idx.alloc((size_t)n * 4)) || (rc = B.codes_s.alloc((size_t)n * 8)) || (rc = B.idx_s.alloc((size_t)n * 4)) ||
        (rc = B.child.alloc((size_t)n * 8)) || (rc = B.range.alloc((size_t)(2 * n) * 8)) ||
        (rc = B.size.alloc((size_t)(2 * n) * 4)) ||
        (rc = B.parent.alloc((size_t)(2 * n) * 4)) || (rc = B.box.alloc((size_t)(2 * n) * sizeof(BoxF))) ||
        (rc = B.cost.alloc((size_t)(2 * n) * 4)) || (rc = B.collapse.alloc((size_t)n * 4)) ||
        (rc = B.flag.alloc((size_t)n * 4)) || (rc = B.nodes.alloc((size_t)(n - 1) * sizeof(BvhNodeHost))))
        return rc;
    const int blocks = (int)std::min<int64_t>((n + 255) / 256, 65536);
    hipLaunchKernelGGL(k_morton, dim3(blocks), dim3(256), 0, 0, B.dv.as<float>(), (uint32_t)n,
                       make_float3(lo[0], lo[1], lo[2]), make_float3(scale[0], scale[1], scale[2]), B.codes.as<uint64_t>(),
                       B.idx.as<uint32_t>());
    SPK_CHECK_HIP(hipGetLastError());
    size_t tbytes = 0;
    SPK_CHECK_HIP(rocprim::radix_sort_pairs(nullptr, tbytes, B.codes.as<uint64_t>(), B.codes_s.as<uint64_t>(),
                                            B.idx.as<uint32_t>(), B.idx_s.as<uint32_t>(), (uint32_t)N, 0u, 63u,
                                            (hipStream_t)0));
    if ((rc = B.tmp.alloc(std::max<size_t>(tbytes, 16)))) return rc;
    SPK_CHECK_HIP(rocprim::radix_sort_pairs(B.tmp.p, tbytes, B.codes.as<uint64_t>(), B.codes_s.as<uint64_t>(),
                                            B.idx.as<uint32_t>(), B.idx_s.as<uint32_t>(), (uint32_t)N, 0u, 63u,
                                            (hipStream_t)0));
    hipLaunchKernelGGL(k_hierarchy, dim3(blocks), dim3(256), 0, 0, B.codes_s.as<uint64_t>(), N, B.child.as<int2>(),
                       B.range.as<int2>(), B.parent.as<int>());
    SPK_CHECK_HIP(hipGetLastError());
    hipLaunchKernelGGL(k_leaves, dim3(blocks), dim3(256), 0, 0, B.dv.as<float>(), B.idx_s.as<uint32_t>(), N,
                       B.box.as<BoxF>(), B.cost.as<float>(), B.size.as<int>(), B.range.as<int2>(), opts.tri_cost);
    SPK_CHECK_HIP(hipGetLastError());
    const uint32_t* leaf_order = B.idx_s.as<uint32_t>();
    if (opts.treelet_passes > 0 && n >= kTreeletMinTris) {
        const int tblocks = (int)std::min<int64_t>((n + kTreeletBlock - 1) / kTreeletBlock, 65536);
        for (int pass = 0; pass < opts.treelet_passes; pass++) {
            SPK_CHECK_HIP(hipMemset(B.flag.p, 0, (size_t)n * 4));
            hipLaunchKernelGGL(k_treelet, dim3(tblocks), dim3(kTreeletBlock), 0, 0, N, B.child.as<int2>(),
                               B.parent.as<int>(), B.box.as<BoxF>(), B.cost.as<float>(), B.size.as<int>(),
                               B.flag.as<uint32_t>(), opts.node_cost);
            SPK_CHECK_HIP(hipGetLastError());
        }
        // ranges top-down, one tree level per launch, from the root's [0, n - 1]
        if ((rc = B.front[0].alloc((size_t)n * 4)) || (rc = B.front[1].alloc((size_t)n * 4)) ||
            (rc = B.front_n.alloc(4)) || (rc = B.order.alloc((size_t)n * 4)))
            return rc;
        const int root = 0;
        const int2 root_range = make_int2(0, N - 1);
        SPK_CHECK_HIP(hipMemcpy(B.front[0].p, &root, 4, hipMemcpyHostToDevice));
        SPK_CHECK_HIP(hipMemcpy(B.range.p, &root_range, 8, hipMemcpyHostToDevice));
        uint32_t nf = 1;
        for (int level = 0; nf > 0; level++) {
            if (level > N) {
                set_error("build_bvh_gpu: restructured tree deeper than its triangle count");
                return SPK_ERR_STATE;
            }
            SPK_CHECK_HIP(hipMemset(B.front_n.p, 0, 4));
            const int lb = (int)std::min<int64_t>((nf + 255) / 256, 65536);
            hipLaunchKernelGGL(k_ranges_level, dim3(lb), dim3(256), 0, 0, B.front[level & 1].as<int>(), nf,
                               B.child.as<int2>(), B.size.as<int>(), B.range.as<int2>(), B.front[(level + 1) & 1].as<int>(),
                               B.front_n.as<uint32_t>(), N);
            SPK_CHECK_HIP(hipGetLastError());
            SPK_CHECK_HIP(hipMemcpy(&nf, B.front_n.p, 4, hipMemcpyDeviceToHost));
        }
        hipLaunchKernelGGL(k_leaf_order, dim3(blocks), dim3(256), 0, 0, B.idx_s.as<uint32_t>(), B.range.as<int2>(), N,
                           B.order.as<uint32_t>());
        SPK_CHECK_HIP(hipGetLastError());
        leaf_order = B.order.as<uint32_t>();
    }
    SPK_CHECK_HIP(hipMemset(B.flag.p, 0, (size_t)n * 4));
    hipLaunchKernelGGL(k_refit, dim3(blocks), dim3(256), 0, 0, B.dv.as<float>(), B.idx_s.as<uint32_t>(), N,
                       B.child.as<int2>(), B.range.as<int2>(), B.parent.as<int>(), B.box.as<BoxF>(), B.cost.as<float>(),
                       B.collapse.as<uint32_t>(), B.flag.as<uint32_t>(), opts.max_leaf, opts.node_cost, opts.tri_cost);
    SPK_CHECK_HIP(hipGetLastError());
    hipLaunchKernelGGL(k_emit, dim3(blocks), dim3(256), 0, 0, B.child.as<int2>(), B.range.as<int2>(), B.box.as<BoxF>(), N,
                       B.collapse.as<uint32_t>(), B.nodes.as<BvhNodeHost>());
    SPK_CHECK_HIP(hipGetLastError());
    std::vector<BvhNodeHost> raw((size_t)(n - 1));
    res.order.resize((size_t)n);
    SPK_CHECK_HIP(hipMemcpy(raw.data(), B.nodes.p, (size_t)(n - 1) * sizeof(BvhNodeHost), hipMemcpyDeviceToHost));
    SPK_CHECK_HIP(hipMemcpy(res.order.data(), leaf_order, (size_t)n * 4, hipMemcpyDeviceToHost));
    compact(raw, res);
    return SPK_OK;
}

}  // namespace spk
