// bvh_build.cpp -- host build of the traversal BVH.
//
// spica builds a BVH2 with one primitive per leaf, 16 SAH buckets on the longest
// centroid axis and a median split below 9 primitives, one heap node at a time
// (sources/accelerators/bvh.cc:139-237).  The GPU traversal wants something
// different: few, fat, cache-line-sized nodes and short leaves.  This builder
// bins all three axes (32 bins), stops at up to kMaxLeafTris triangles per leaf
// when the SAH says a leaf is cheaper, builds large subtrees on parallel threads
// and then flattens depth-first into binary nodes, which collapse_bvh4 turns into 4-wide
// nodes and quantize_bvh4 into the 64-byte nodes the traversal reads (one node fetch =
// four box tests).  The
// closest hit does not depend on the tree (only exact-tie winners do), so the tree is
// free to differ from spica's.
#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstring>
#include <memory>
#include <mutex>
#include <thread>
#include <vector>

#include "spk_internal.h"

namespace spk {

namespace {

struct Box {
    float lo[3], hi[3];
    void reset() {
        for (int i = 0; i < 3; i++) {
            lo[i] = FLT_MAX;
            hi[i] = -FLT_MAX;
        }
    }
    void grow(const Box& b) {
        for (int i = 0; i < 3; i++) {
            lo[i] = std::min(lo[i], b.lo[i]);
            hi[i] = std::max(hi[i], b.hi[i]);
        }
    }
    void grow(const float* p) {
        for (int i = 0; i < 3; i++) {
            lo[i] = std::min(lo[i], p[i]);
            hi[i] = std::max(hi[i], p[i]);
        }
    }
    float half_area() const {
        float dx = hi[0] - lo[0], dy = hi[1] - lo[1], dz = hi[2] - lo[2];
        if (dx < 0 || dy < 0 || dz < 0) return 0.0f;
        return dx * dy + dy * dz + dz * dx;
    }
};

struct PrimRef {
    Box b;
    float c[3];
    uint32_t id;
};

// Child reference during the build: pool/index of an interior node, or a leaf.
struct Ref {
    int32_t pool;
    int32_t idx;      // interior node index in the pool, or -1 for a leaf
    int64_t first;    // leaf: first PrimRef
    int32_t count;    // leaf: triangle count
};

struct TmpNode {
    Box box[2];
    Ref child[2];
};

struct Pool {
    std::vector<TmpNode> nodes;
};

constexpr int kBins = 32;
constexpr int kMaxBuildDepth = 48;

class Builder {
public:
    Builder(std::vector<PrimRef>& refs, const BvhBuildOpts& opts) : refs_(refs), opts_(opts) {}

    Ref build_root(Box& root_box) {
        pools_.emplace_back(new Pool());
        return build(0, 0, (int64_t)refs_.size(), 0, root_box);
    }

    std::vector<std::unique_ptr<Pool>>& pools() { return pools_; }
    int max_depth() const { return max_depth_; }

private:
    int new_pool() {
        std::lock_guard<std::mutex> g(mu_);
        pools_.emplace_back(new Pool());
        return (int)pools_.size() - 1;
    }
    Pool* pool(int p) {
        std::lock_guard<std::mutex> g(mu_);
        return pools_[p].get();
    }
    void note_depth(int d) {
        std::lock_guard<std::mutex> g(mu_);
        max_depth_ = std::max(max_depth_, d);
    }

    Ref make_leaf(int64_t b, int64_t e) {
        Ref r;
        r.pool = -1;
        r.idx = -1;
        r.first = b;
        r.count = (int32_t)(e - b);
        return r;
    }

    // Splits [b, e) at the best binned-SAH plane over all three axes.  Returns
    // the split position, or -1 when a leaf is cheaper (and allowed).
    int64_t find_split(int64_t b, int64_t e, const Box& box, int depth) {
        int64_t n = e - b;
        Box cb;
        cb.reset();
        for (int64_t i = b; i < e; i++) cb.grow(refs_[i].c);
        float best_cost = FLT_MAX;
        int best_axis = -1, best_bin = -1;
        for (int ax = 0; ax < 3; ax++) {
            float ext = cb.hi[ax] - cb.lo[ax];
            if (!(ext > 0.0f)) continue;
            float scale = (float)kBins * (1.0f - 1e-6f) / ext;
            Box bb[kBins];
            int64_t cnt[kBins];
            for (int k = 0; k < kBins; k++) {
                bb[k].reset();
                cnt[k] = 0;
            }
            for (int64_t i = b; i < e; i++) {
                int k = std::min(kBins - 1, std::max(0, (int)((refs_[i].c[ax] - cb.lo[ax]) * scale)));
                bb[k].grow(refs_[i].b);
                cnt[k]++;
            }
            float right_area[kBins];
            int64_t right_cnt[kBins];
            Box acc;
            acc.reset();
            int64_t c = 0;
            for (int k = kBins - 1; k > 0; k--) {
                acc.grow(bb[k]);
                c += cnt[k];
                right_area[k] = acc.half_area();
                right_cnt[k] = c;
            }
            acc.reset();
            c = 0;
            for (int k = 0; k < kBins - 1; k++) {
                acc.grow(bb[k]);
                c += cnt[k];
                if (c == 0 || right_cnt[k + 1] == 0) continue;
                float cost = acc.half_area() * (float)c + right_area[k + 1] * (float)right_cnt[k + 1];
                if (cost < best_cost) {
                    best_cost = cost;
                    best_axis = ax;
                    best_bin = k;
                }
            }
        }
        float parent = box.half_area();
        bool must_split = n > opts_.max_leaf;
        if (!must_split && depth >= kMaxBuildDepth) return -1;
        if (best_axis >= 0 && parent > 0.0f) {
            float split_cost = opts_.node_cost + opts_.tri_cost * best_cost / parent;
            float leaf_cost = opts_.tri_cost * (float)n;
            if (!must_split && leaf_cost <= split_cost) return -1;
            float ext = cb.hi[best_axis] - cb.lo[best_axis];
            float scale = (float)kBins * (1.0f - 1e-6f) / ext;
            float lo = cb.lo[best_axis];
            int ax = best_axis, bin = best_bin;
            PrimRef* mid = std::partition(refs_.data() + b, refs_.data() + e, [&](const PrimRef& p) {
                int k = std::min(kBins - 1, std::max(0, (int)((p.c[ax] - lo) * scale)));
                return k <= bin;
            });
            int64_t m = mid - refs_.data();
            if (m > b && m < e) return m;
        }
        if (!must_split) return -1;
        // Degenerate centroids (or depth cap): split by count along the longest axis.
        int ax = 0;
        float ext0 = cb.hi[0] - cb.lo[0], ext1 = cb.hi[1] - cb.lo[1], ext2 = cb.hi[2] - cb.lo[2];
        if (ext1 > ext0 && ext1 >= ext2) ax = 1;
        else if (ext2 > ext0 && ext2 > ext1) ax = 2;
        int64_t m = b + n / 2;
        std::nth_element(refs_.data() + b, refs_.data() + m, refs_.data() + e, [ax](const PrimRef& x, const PrimRef& y) {
            return x.c[ax] < y.c[ax] || (x.c[ax] == y.c[ax] && x.id < y.id);
        });
        return m;
    }

    Ref build(int pool_id, int64_t b, int64_t e, int depth, Box& out_box) {
        out_box.reset();
        for (int64_t i = b; i < e; i++) out_box.grow(refs_[i].b);
        note_depth(depth);
        int64_t n = e - b;
        if (n <= 1) return make_leaf(b, e);
        int64_t m = find_split(b, e, out_box, depth);
        if (m < 0) return make_leaf(b, e);
        TmpNode node;
        if (n >= opts_.parallel_threshold && depth < opts_.parallel_depth) {
            int left_pool = new_pool();
            Ref left;
            std::thread t([&]() { left = build(left_pool, b, m, depth + 1, node.box[0]); });
            node.child[1] = build(pool_id, m, e, depth + 1, node.box[1]);
            t.join();
            node.child[0] = left;
        } else {
            node.child[0] = build(pool_id, b, m, depth + 1, node.box[0]);
            node.child[1] = build(pool_id, m, e, depth + 1, node.box[1]);
        }
        Pool* p = pool(pool_id);
        p->nodes.push_back(node);
        Ref r;
        r.pool = pool_id;
        r.idx = (int32_t)p->nodes.size() - 1;
        r.first = 0;
        r.count = 0;
        return r;
    }

    std::vector<PrimRef>& refs_;
    BvhBuildOpts opts_;
    std::vector<std::unique_ptr<Pool>> pools_;
    std::mutex mu_;
    int max_depth_ = 0;
};

int32_t leaf_link(int64_t first, int32_t count) { return ~(int32_t)((first << kLeafCountBits) | (count - 1)); }

}  // namespace

BvhResult build_bvh(const float* verts, int64_t n, const BvhBuildOpts& opts) {
    BvhResult res;
    res.max_depth = 0;
    res.num_leaves = 0;
    if (n <= 0) return res;
    std::vector<PrimRef> refs((size_t)n);
    for (int64_t i = 0; i < n; i++) {
        PrimRef& r = refs[(size_t)i];
        r.b.reset();
        for (int v = 0; v < 3; v++) r.b.grow(verts + 9 * i + 3 * v);
        for (int a = 0; a < 3; a++) r.c[a] = 0.5f * (r.b.lo[a] + r.b.hi[a]);
        r.id = (uint32_t)i;
    }
    Builder bld(refs, opts);
    Box root_box;
    Ref root = bld.build_root(root_box);
    res.max_depth = bld.max_depth() + 1;

    // Flatten depth-first (left subtree first), interior nodes only.
    auto& pools = bld.pools();
    std::vector<std::pair<Ref, int32_t>> stack;  // (ref, slot to patch: node*2 + side, or -1)
    auto emit_node = [&](const BvhNodeHost& h) {
        res.nodes.push_back(h);
        return (int32_t)res.nodes.size() - 1;
    };
    auto box_to_node = [](BvhNodeHost& h, const Box& l, const Box& r) {
        h.a[0] = l.lo[0]; h.a[1] = l.hi[0]; h.a[2] = l.lo[1]; h.a[3] = l.hi[1];
        h.b[0] = r.lo[0]; h.b[1] = r.hi[0]; h.b[2] = r.lo[1]; h.b[3] = r.hi[1];
        h.c[0] = l.lo[2]; h.c[1] = l.hi[2]; h.c[2] = r.lo[2]; h.c[3] = r.hi[2];
    };
    if (root.idx < 0) {
        // A single leaf: wrap it in a root whose two links name the same leaf.
        BvhNodeHost h;
        std::memset(&h, 0, sizeof(h));
        box_to_node(h, root_box, root_box);
        h.d[0] = h.d[1] = leaf_link(root.first, root.count);
        emit_node(h);
        res.num_leaves = 1;
    } else {
        stack.push_back({root, -1});
        while (!stack.empty()) {
            auto item = stack.back();
            stack.pop_back();
            const TmpNode& tn = pools[item.first.pool]->nodes[item.first.idx];
            BvhNodeHost h;
            std::memset(&h, 0, sizeof(h));
            box_to_node(h, tn.box[0], tn.box[1]);
            int32_t id = emit_node(h);
            if (item.second >= 0) res.nodes[item.second >> 1].d[item.second & 1] = id;
            for (int side = 1; side >= 0; side--) {
                const Ref& c = tn.child[side];
                if (c.idx < 0) {
                    res.nodes[id].d[side] = leaf_link(c.first, c.count);
                    res.num_leaves++;
                } else {
                    stack.push_back({c, id * 2 + side});
                }
            }
        }
    }
    res.order.resize((size_t)n);
    for (int64_t i = 0; i < n; i++) res.order[(size_t)i] = refs[(size_t)i].id;
    return res;
}

// ---- 4-wide collapse ------------------------------------------------------------------
// Each 4-wide node takes the two children of a binary node and keeps opening the interior
// child of largest surface area into its two children until it holds four (or only
// leaves): the 4-wide tree visits about half the nodes with the same leaves and the same
// triangle order.  Nodes are numbered depth first from the root (0).
int collapse_bvh4(const std::vector<BvhNodeHost>& n2, std::vector<Bvh4NodeHost>& out) {
    out.clear();
    if (n2.empty()) return 0;
    struct Item {
        float lo[3], hi[3];
        int32_t link;
    };
    auto child = [&](int32_t node, int side) {
        const BvhNodeHost& h = n2[(size_t)node];
        Item it;
        const float* xy = side == 0 ? h.a : h.b;
        it.lo[0] = xy[0];
        it.hi[0] = xy[1];
        it.lo[1] = xy[2];
        it.hi[1] = xy[3];
        it.lo[2] = h.c[2 * side];
        it.hi[2] = h.c[2 * side + 1];
        it.link = h.d[side];
        return it;
    };
    auto area = [](const Item& it) {
        const float dx = it.hi[0] - it.lo[0], dy = it.hi[1] - it.lo[1], dz = it.hi[2] - it.lo[2];
        return dx * dy + dy * dz + dz * dx;
    };
    struct Work {
        int32_t node2, node4, depth;
    };
    std::vector<Work> st;
    out.push_back(Bvh4NodeHost{});
    st.push_back({0, 0, 1});
    int max_depth = 0;
    while (!st.empty()) {
        const Work w = st.back();
        st.pop_back();
        max_depth = std::max(max_depth, w.depth);
        Item items[4];
        int n = 0;
        items[n++] = child(w.node2, 0);
        items[n++] = child(w.node2, 1);
        while (n < 4) {
            int best = -1;
            float best_area = -1.0f;
            for (int j = 0; j < n; j++)
                if (items[j].link >= 0 && area(items[j]) > best_area) {
                    best = j;
                    best_area = area(items[j]);
                }
            if (best < 0) break;
            const int32_t opened = items[best].link;
            items[best] = child(opened, 0);
            items[n++] = child(opened, 1);
        }
        Bvh4NodeHost h;
        memset(&h, 0, sizeof(h));
        for (int j = 0; j < 4; j++) {
            if (j >= n) {
                h.link[j] = kBvh4Empty;
                continue;
            }
            for (int a = 0; a < 3; a++) {
                h.lo[a][j] = items[j].lo[a];
                h.hi[a][j] = items[j].hi[a];
            }
            if (items[j].link >= 0) {
                const int32_t idx = (int32_t)out.size();
                out.push_back(Bvh4NodeHost{});
                st.push_back({items[j].link, idx, w.depth + 1});
                h.link[j] = idx;
            } else {
                h.link[j] = items[j].link;
            }
        }
        out[(size_t)w.node4] = h;
    }
    return max_depth;
}

// ---- 8-bit child boxes ----------------------------------------------------------------
// Per axis the node's box [lo, hi] (the union of its children) is split into 255 steps:
// scale = (hi - lo) / 255, raised (by doubling increments) until code 255 reaches hi.  A child's lower plane takes the
// largest code whose decoding fmaf(q, scale, lo) is <= the plane, its upper plane the
// smallest code whose decoding is >= it.  Device and host fmaf are both correctly rounded,
// and decode_q flushes subnormals as the traversal kernels do, so the traversal decodes
// exactly the boxes checked here (a negative subnormal lower bound is stored as -FLT_MIN,
// which the flush keeps below the plane).  A flat axis has scale 0 and decodes every code
// to the plane itself.
void quantize_bvh4(const std::vector<Bvh4NodeHost>& in, std::vector<Bvh4QNodeHost>& out) {
    out.resize(in.size());
    for (size_t i = 0; i < in.size(); i++) {
        const Bvh4NodeHost& h = in[i];
        Bvh4QNodeHost& q = out[i];
        memset(&q, 0, sizeof(q));
        for (int c = 0; c < 4; c++) q.link[c] = h.link[c];
        for (int a = 0; a < 3; a++) {
            float lo = FLT_MAX, hi = -FLT_MAX;
            for (int c = 0; c < 4; c++)
                if (h.link[c] != kBvh4Empty) {
                    lo = std::min(lo, h.lo[a][c]);
                    hi = std::max(hi, h.hi[a][c]);
                }
            if (lo > hi) lo = hi = 0.0f;  // no children (never for a built tree)
            if (flush_subnormal(lo) > lo) lo = -FLT_MIN;
            float sc = (hi - lo) / 255.0f;
            // raise the step until code 255 reaches hi; the increment doubles, so a nearly flat
            // box far from the origin needs a few rounds, not one per ulp of the step
            for (float add = std::nextafter(sc, FLT_MAX) - sc; decode_q(255.0f, sc, lo) < hi; add *= 2.0f) sc += add;
            q.origin[a] = lo;
            q.scale[a] = sc;
            uint32_t wlo = 0, whi = 0;
            for (int c = 0; c < 4; c++) {
                int ql = 0, qh = 0;
                if (h.link[c] != kBvh4Empty) {
                    const float cl = h.lo[a][c], ch = h.hi[a][c];
                    if (sc > 0.0f) {
                        ql = std::min(255, std::max(0, (int)std::floor((cl - lo) / sc)));
                        qh = std::min(255, std::max(0, (int)std::ceil((ch - lo) / sc)));
                        while (ql > 0 && decode_q((float)ql, sc, lo) > cl) ql--;
                        while (ql < 255 && decode_q((float)(ql + 1), sc, lo) <= cl) ql++;
                        while (qh < 255 && decode_q((float)qh, sc, lo) < ch) qh++;
                        while (qh > 0 && decode_q((float)(qh - 1), sc, lo) >= ch) qh--;
                    }
                }
                wlo |= (uint32_t)ql << (8 * c);
                whi |= (uint32_t)qh << (8 * c);
            }
            q.qlo[a] = wlo;
            q.qhi[a] = whi;
        }
    }
}

}  // namespace spk
