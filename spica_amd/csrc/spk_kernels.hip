// spk_kernels.hip -- gfx950 kernels of the wavefront path tracer.
//
// spica traces one path per thread-loop iteration, depth first, on the CPU
// (sources/integrators/path/path.cc:42-125, sources/core/integrator.cc:77-97).
// Here the path's bounces are split into stages over compacted SoA queues:
//
//   pass_init -> camera -> [extend -> shade -> shadow -> mis] x (maxDepth+1) -> film
//
// * extend / shadow / mis: persistent BVH traversal (closest hit / any hit /
//   closest hit), rays pulled 64 at a time per wave from an atomic work counter,
//   short per-lane stack in LDS (spilling to HBM only when a ray goes deeper than
//   the LDS slots), interior-node loop with postponed leaves that the wave leaves
//   only when every lane holds a leaf (wave vote).
// * shade: one entry per hit; emission, spica's NEE (light sample + BSDF sample
//   with the power heuristic, mis.cc:21-136), BSDF continuation and Russian
//   roulette; it appends to the next extend queue, the shadow queue and the MIS
//   queue with one atomic per wave per queue (ballot + mbcnt prefix).
// * film: per pixel, the samples of the pass are added in sample order into
//   Kahan-compensated float sums of L and L^2: deterministic and reproducible.
#include "spk_internal.h"
#include "spk_guide.h"
#include "spk_sampler.h"

namespace spk {

namespace {

constexpr float kEPS = 1.0e-12f;      // core/common.h:55
constexpr float kINFTY = 1.0e32f;     // core/common.h:54
constexpr float kPI = 3.14159265358979323846f;
constexpr float kINV_PI = 0.31830988618379067154f;
constexpr float kRobust = 1.0000005f; // widens the slab interval by ~4 ulp
constexpr int32_t kSent = 0x7fffffff; // "no node"

// ---- small vector helpers ---------------------------------------------------------
__device__ __forceinline__ float3 mk(float x, float y, float z) { return make_float3(x, y, z); }
__device__ __forceinline__ float3 xyz(float4 v) { return make_float3(v.x, v.y, v.z); }
__device__ __forceinline__ float3 add(float3 a, float3 b) { return mk(a.x + b.x, a.y + b.y, a.z + b.z); }
__device__ __forceinline__ float3 sub(float3 a, float3 b) { return mk(a.x - b.x, a.y - b.y, a.z - b.z); }
__device__ __forceinline__ float3 mul(float3 a, float s) { return mk(a.x * s, a.y * s, a.z * s); }
__device__ __forceinline__ float3 cmul(float3 a, float3 b) { return mk(a.x * b.x, a.y * b.y, a.z * b.z); }
__device__ __forceinline__ float3 neg(float3 a) { return mk(-a.x, -a.y, -a.z); }
__device__ __forceinline__ float dot(float3 a, float3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ __forceinline__ float3 cross(float3 a, float3 b) {
    return mk(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
__device__ __forceinline__ float3 normalize(float3 a) {
    float n = sqrtf(a.x * a.x + a.y * a.y + a.z * a.z);
    return mk(a.x / n, a.y / n, a.z / n);
}
__device__ __forceinline__ bool black(float3 a) { return a.x == 0.0f && a.y == 0.0f && a.z == 0.0f; }
__device__ __forceinline__ float gray(float3 c) { return 0.2126f * c.x + 0.7152f * c.y + 0.0722f * c.z; }
__device__ __forceinline__ float clampf(float v, float lo, float hi) { return fminf(fmaxf(v, lo), hi); }
__device__ __forceinline__ float comp(float3 a, int i) { return i == 0 ? a.x : (i == 1 ? a.y : a.z); }

// float.h:23-45 in single precision (zero goes down for Up, up for Down).
__device__ __forceinline__ float next_up(float f) {
    if (__builtin_isinf(f) && f > 0.0f) return f;
    if (f == -0.0f) f = 0.0f;
    uint32_t u = __float_as_uint(f);
    u = f > 0.0f ? u + 1u : u - 1u;
    return __uint_as_float(u);
}
__device__ __forceinline__ float next_down(float f) {
    if (__builtin_isinf(f) && f > 0.0f) return f;
    if (f == -0.0f) f = 0.0f;
    uint32_t u = __float_as_uint(f);
    u = f > 0.0f ? u - 1u : u + 1u;
    return __uint_as_float(u);
}
// interaction.cc:16-30
__device__ __forceinline__ float3 offset_origin(float3 p, float3 n, float3 w) {
    float3 off = mul(n, 1.0e-3f);
    if (dot(w, n) < 0.0f) off = neg(off);
    float3 po = add(p, off);
    return mk(off.x > 0.0f ? next_up(po.x) : next_down(po.x), off.y > 0.0f ? next_up(po.y) : next_down(po.y),
              off.z > 0.0f ? next_up(po.z) : next_down(po.z));
}
// vect_math.h:121-128
__device__ __forceinline__ void coordinate_system(float3 w, float3& u, float3& v) {
    if (fabsf(w.x) > fabsf(w.y)) u = normalize(mk(-w.z, 0.0f, w.x));
    else u = normalize(mk(0.0f, w.z, -w.y));
    v = normalize(cross(w, u));
}

// Closest-hit traversal fits 64 VGPRs without spilling when asked to: 8 waves per SIMD
// instead of the 7 the compiler picks by itself (C1 closest-hit launches 2% shorter).  The
// batched and visit-counting variants hold more per lane and would spill at 8; they run at 7
// (72 VGPRs; the batched closest hit needs 75 unbounded, and at 7 waves the C2 batch is 1%
// faster than at 6).
template <bool BATCH, bool COUNT>
constexpr int traversal_waves() { return BATCH ? 7 : COUNT ? 7 : 8; }

// ---- block-level queue append -----------------------------------------------------------
// One counter atomic per block and queue instead of one per wave.  The queue counters are
// the hottest addresses of a pass and the atomics on one address are performed one after
// another, so at a wave per atomic they took about 60% of the shading kernel's time on a
// full first bounce.  Every thread of the block calls this together (block-uniform loops).
// lds holds 2 * NQ * (NW + 1) words, used by alternate calls (parity), so a wave that has
// moved on to the next call never overwrites totals a slower wave still reads: it cannot
// pass that next call's first barrier before every wave has finished reading.
template <int NW, int NQ>
__device__ __forceinline__ void block_append(const bool (&pred)[NQ], uint32_t* const (&counter)[NQ],
                                             uint32_t (&slot)[NQ], uint32_t* lds, uint32_t parity) {
    constexpr int R = NW + 1;  // per queue: the wave totals, then the block's base
    uint32_t* buf = lds + parity * (NQ * R);
    const int w = (int)(threadIdx.x >> 6);
    uint32_t prefix[NQ];
#pragma unroll
    for (int q = 0; q < NQ; q++) {
        const unsigned long long m = __ballot(pred[q] ? 1 : 0);
        prefix[q] = __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
        if (__lane_id() == 0) buf[q * R + w] = (uint32_t)__popcll(m);
    }
    __syncthreads();
    if (threadIdx.x < (unsigned)NQ) {
        const int q = (int)threadIdx.x;
        uint32_t sum = 0;
#pragma unroll
        for (int v = 0; v < NW; v++) sum += buf[q * R + v];
        uint32_t* c = counter[0];
#pragma unroll
        for (int k = 1; k < NQ; k++)
            if (q == k) c = counter[k];
        buf[q * R + NW] = sum != 0 ? atomicAdd(c, sum) : 0u;
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < NQ; q++) {
        uint32_t b = buf[q * R + NW];
        for (int v = 0; v < w; v++) b += buf[q * R + v];
        slot[q] = b + prefix[q];
    }
}

// This file is compiled with fast single-precision division (reciprocal and multiply) and
// flushes single-precision denormals; the two reciprocals of the traversal stay correctly
// rounded: the double quotient of two floats rounded to float is the correctly rounded float
// quotient (53 >= 2 * 24 + 2 bits make the double rounding innocuous).
__device__ __forceinline__ float rcp_rn(float x) { return (float)(1.0 / (double)x); }

// ---- ray/triangle (triangle.cc:98-178 in float) -----------------------------------------
__device__ __forceinline__ bool tri_hit(const float4* __restrict__ tris, uint32_t k, float3 o, float3 d, float tmax,
                                        float& t, float& u, float& v) {
    float4 T0 = tris[3 * k], T1 = tris[3 * k + 1], T2 = tris[3 * k + 2];
    float3 e1 = xyz(T1), e2 = xyz(T2);
    float3 pv = cross(d, e2);
    float det = dot(e1, pv);
    if (det > -kEPS && det < kEPS) return false;
    float inv = rcp_rn(det);
    float3 tv = sub(o, xyz(T0));
    float uu = dot(tv, pv) * inv;
    if (uu < 0.0f || uu > 1.0f) return false;
    float3 qv = cross(tv, e1);
    float vv = dot(d, qv) * inv;
    if (vv < 0.0f || uu + vv > 1.0f) return false;
    float tt = dot(e2, qv) * inv;
    if (tt <= kEPS || tt > tmax) return false;
    t = tt;
    u = uu;
    v = vv;
    return true;
}

__device__ __forceinline__ float3 safe_inv(float3 d) {
    // ray.cc:43-47: zero components become INFTY; tiny ones are clamped to it
    auto one = [](float x) {
        if (x == 0.0f) return kINFTY;
        float r = rcp_rn(x);
        return fabsf(r) > kINFTY ? copysignf(kINFTY, r) : r;
    };
    return mk(one(d.x), one(d.y), one(d.z));
}

struct Counters {
    uint32_t nodes, tris;
};


// One 4-wide node (8 float4: lo.x, lo.y, lo.z, hi.x, hi.y, hi.z of the four children,
// their links, pad): slab-tests the four boxes against [0, tbest] in FMA form (plane t =
// p * (1/d) - o * (1/d), one fused multiply-add per plane), continues with the
// nearest hit child and pushes the other hit children farthest first, so they pop in
// near-to-far order.  Returns the next node or leaf link (kSent when the stack is empty).
template <bool COUNT, class Push, class Pop>
__device__ __forceinline__ int32_t visit4(const float4* __restrict__ nodes, int32_t node, float3 inv, float3 oinv,
                                          float tbest, Push& push, Pop& pop, Counters& cnt) {
    // 64-byte node: origin, scale and 8-bit child planes (decoded as fmaf(q, scale, origin),
    // boxes rounded outwards at build time), then the four links
    const float4* nd = nodes + 4 * node;
    const float4 A = nd[0], B = nd[1], C = nd[2], lk = nd[3];
    if (COUNT) cnt.nodes++;
    const uint32_t qlx = __float_as_uint(B.z), qly = __float_as_uint(B.w), qlz = __float_as_uint(C.x);
    const uint32_t qhx = __float_as_uint(C.y), qhy = __float_as_uint(C.z), qhz = __float_as_uint(C.w);
    auto dq = [](uint32_t w, int c, float sc, float o) { return fmaf((float)((w >> (8 * c)) & 255u), sc, o); };
    auto plane4 = [&](uint32_t w, float sc, float o) {
        return make_float4(dq(w, 0, sc, o), dq(w, 1, sc, o), dq(w, 2, sc, o), dq(w, 3, sc, o));
    };
    const float4 lx = plane4(qlx, A.w, A.x), ly = plane4(qly, B.x, A.y), lz = plane4(qlz, B.y, A.z);
    const float4 hx = plane4(qhx, A.w, A.x), hy = plane4(qhy, B.x, A.y), hz = plane4(qhz, B.y, A.z);
    const float kNone = __int_as_float(0x7f800000);  // +inf: sorts after every hit
    auto slab = [&](float ax, float ay, float az, float bx, float by, float bz, int32_t link) {
        const float l_x = fmaf(ax, inv.x, -oinv.x), h_x = fmaf(bx, inv.x, -oinv.x);
        const float l_y = fmaf(ay, inv.y, -oinv.y), h_y = fmaf(by, inv.y, -oinv.y);
        const float l_z = fmaf(az, inv.z, -oinv.z), h_z = fmaf(bz, inv.z, -oinv.z);
        const float tn = fmaxf(fmaxf(fminf(l_x, h_x), fminf(l_y, h_y)), fmaxf(fminf(l_z, h_z), 0.0f));
        const float tf = fminf(fminf(fmaxf(l_x, h_x), fmaxf(l_y, h_y)), fminf(fmaxf(l_z, h_z), tbest));
        return (tn <= tf * kRobust && link != kSent) ? tn : kNone;
    };
    int32_t c0 = __float_as_int(lk.x), c1 = __float_as_int(lk.y), c2 = __float_as_int(lk.z), c3 = __float_as_int(lk.w);
    float k0 = slab(lx.x, ly.x, lz.x, hx.x, hy.x, hz.x, c0);
    float k1 = slab(lx.y, ly.y, lz.y, hx.y, hy.y, hz.y, c1);
    float k2 = slab(lx.z, ly.z, lz.z, hx.z, hy.z, hz.z, c2);
    float k3 = slab(lx.w, ly.w, lz.w, hx.w, hy.w, hz.w, c3);
    const int n = (k0 != kNone) + (k1 != kNone) + (k2 != kNone) + (k3 != kNone);
    if (n == 0) return pop();
    auto cx = [](float& ka, int32_t& ca, float& kb, int32_t& cb) {
        const bool sw = kb < ka;
        const float kt = sw ? kb : ka;
        const int32_t ct = sw ? cb : ca;
        kb = sw ? ka : kb;
        cb = sw ? ca : cb;
        ka = kt;
        ca = ct;
    };
    // ascending by entry distance (misses last): a 5-comparator network
    cx(k0, c0, k1, c1);
    cx(k2, c2, k3, c3);
    cx(k0, c0, k2, c2);
    cx(k1, c1, k3, c3);
    cx(k1, c1, k2, c2);
    if (n > 3) push(c3);
    if (n > 2) push(c2);
    if (n > 1) push(c1);
    return c0;
}

// Closest hit (ANY = false) or first hit (ANY = true).  Returns the leaf-order
// triangle index or -1.  `col` is this lane's LDS stack column (slot s at
// col[s * BLOCK]); `spill` holds entries below the S most recent ones.
template <int BLOCK, int S, bool ANY, bool COUNT>
__device__ __forceinline__ int32_t traverse(const SceneDev& sc, float3 o, float3 d, float tmax, float& t_out, float& u_out,
                                            float& v_out, uint32_t* col, uint32_t* spill, bool active, Counters& cnt) {
    const float4* __restrict__ nodes = sc.nodes;
    const float4* __restrict__ tris = sc.tris;
    float3 inv = safe_inv(d);
    const float3 oinv = mk(o.x * inv.x, o.y * inv.y, o.z * inv.z);
    int32_t node = (active && sc.has_bvh) ? 0 : kSent;
    int32_t leaf = 0;  // pending leaf link (negative) or 0
    int sp = 0, spilled = 0;
    int32_t hit = -1;
    float tbest = tmax, hu = 0.0f, hv = 0.0f;

    auto push = [&](int32_t x) {
        if (sp - spilled == S) {
            spill[spilled] = col[(spilled & (S - 1)) * BLOCK];
            spilled++;
        }
        col[(sp & (S - 1)) * BLOCK] = (uint32_t)x;
        sp++;
    };
    auto pop = [&]() -> int32_t {
        if (sp == 0) return kSent;
        sp--;
        if (sp < spilled) {
            spilled = sp;
            return (int32_t)spill[sp];
        }
        return (int32_t)col[(sp & (S - 1)) * BLOCK];
    };

    for (;;) {
        // Interior nodes until every lane of the wave holds a leaf (or is done).
        while (node >= 0 && node != kSent) {
            node = visit4<COUNT>(nodes, node, inv, oinv, tbest, push, pop, cnt);
            if (node < 0 && leaf == 0) {
                leaf = node;
                node = pop();
            }
            if (__all(leaf != 0)) break;
        }
        // Leaves.
        while (leaf < 0) {
            uint32_t x = ~(uint32_t)leaf;
            uint32_t first = x >> kLeafCountBits, n = (x & (kMaxLeafTris - 1)) + 1;
            for (uint32_t k = first; k < first + n; k++) {
                float t, u, v;
                if (COUNT) cnt.tris++;
                if (tri_hit(tris, k, o, d, tbest, t, u, v)) {
                    hit = (int32_t)k;
                    if (ANY) break;
                    tbest = t;
                    hu = u;
                    hv = v;
                }
            }
            if (ANY && hit >= 0) {
                node = kSent;
                leaf = 0;
                sp = 0;
                spilled = 0;
                break;
            }
            leaf = 0;
            if (node < 0) {
                leaf = node;
                node = pop();
            }
        }
        if (node == kSent && leaf == 0) break;
    }
    t_out = tbest;
    u_out = hu;
    v_out = hv;
    return hit;
}

// Static work split for the persistent traversal grids.  The hardware deals
// workgroups round-robin over the 8 XCDs; renumbering so that XCD x owns the
// logical blocks [x * G/8, (x+1) * G/8) gives each XCD (and its L2) a contiguous
// slice of the queue, i.e. neighbouring, coherent rays.  Each wave then takes
// 64-item chunks grid-stride: no shared counter, no atomics.
constexpr uint32_t kXcds = 8;
struct WaveSplit {
    uint32_t first;   // first item of this wave
    uint32_t stride;  // items between successive chunks of this wave
};
template <int BLOCK>
__device__ __forceinline__ WaveSplit wave_split() {
    uint32_t g = gridDim.x, b = blockIdx.x;
    uint32_t logical = b;
    if (g % kXcds == 0) logical = (b % kXcds) * (g / kXcds) + b / kXcds;
    uint32_t wave = logical * (BLOCK / 64) + threadIdx.x / 64;
    WaveSplit s;
    s.first = wave * 64u;
    s.stride = g * (uint32_t)BLOCK;
    return s;
}

// Streaming traversal for the persistent queue kernels: every lane traces its own ray
// and, when at least kRefill lanes of the wave are idle, the wave hands the next rays of
// its slice (the same 64-ray chunks wave_split assigns, in order) to the idle lanes, so
// a wave is not held by its slowest ray (dynamic ray fetch, Aila and Laine 2009).
// load(i, o, d, tmax) reads ray i and returns its scene (batch: an entry of the scene
// table; single: null, the kernel's scene argument is used); done(i, hit, t, u, v, scene)
// consumes its result.
#ifndef SPK_REFILL
#define SPK_REFILL 8
#endif
constexpr int kRefill = SPK_REFILL;

template <bool BATCH>
__device__ __forceinline__ const SceneDev& pick_scene(const SceneDev* p, const SceneDev& s) {
    if constexpr (BATCH) return *p;
    else return s;
}

template <int BLOCK, int S, bool ANY, bool COUNT, bool BATCH, class Load, class Done>
__device__ __forceinline__ void traverse_stream(const SceneDev& sc0, const SceneDev* scenes, uint32_t count, Load load,
                                                Done done, uint32_t* col, uint32_t* spill, Counters& cnt) {
    const WaveSplit ws = wave_split<BLOCK>();
    bool active = false;
    uint32_t idx = 0;
    const SceneDev* sc = scenes;  // batch only: this lane's scene
    float3 o = mk(0, 0, 0), d = mk(0, 0, 1), inv = mk(1, 1, 1), oinv = mk(0, 0, 0);
    int32_t node = kSent, leaf = 0;
    int sp = 0, spilled = 0;
    int32_t hit = -1;
    float tbest = 0.0f, hu = 0.0f, hv = 0.0f;
    uint32_t cursor = 0;  // wave-uniform: positions of the wave's slice handed out so far
    auto index_of = [&](uint32_t pos) { return ws.first + (pos >> 6) * ws.stride + (pos & 63u); };
    bool exhausted = index_of(0) >= count;

    auto push = [&](int32_t x) {
        if (sp - spilled == S) {
            spill[spilled] = col[(spilled & (S - 1)) * BLOCK];
            spilled++;
        }
        col[(sp & (S - 1)) * BLOCK] = (uint32_t)x;
        sp++;
    };
    auto pop = [&]() -> int32_t {
        if (sp == 0) return kSent;
        sp--;
        if (sp < spilled) {
            spilled = sp;
            return (int32_t)spill[sp];
        }
        return (int32_t)col[(sp & (S - 1)) * BLOCK];
    };

    for (;;) {
        if (!exhausted) {
            const uint64_t idle = __ballot(!active);
            const uint32_t n_idle = (uint32_t)__popcll(idle);
            if (n_idle >= (uint32_t)kRefill || n_idle == 64u) {
                if (!active) {
                    const uint32_t i = index_of(cursor + (uint32_t)__popcll(idle & ((1ull << __lane_id()) - 1ull)));
                    if (i < count) {
                        float tmax;
                        const SceneDev* ls = load(i, o, d, tmax);
                        if constexpr (BATCH) sc = ls;
                        idx = i;
                        active = true;
                        inv = safe_inv(d);
                        oinv = mk(o.x * inv.x, o.y * inv.y, o.z * inv.z);
                        node = pick_scene<BATCH>(sc, sc0).has_bvh ? 0 : kSent;
                        leaf = 0;
                        sp = 0;
                        spilled = 0;
                        hit = -1;
                        tbest = tmax;
                        hu = hv = 0.0f;
                    }
                }
                cursor += n_idle;
                exhausted = index_of(cursor) >= count;
            }
        }
        if (!__any(active)) break;
        const float4* __restrict__ nodes = pick_scene<BATCH>(sc, sc0).nodes;
        const float4* __restrict__ tris = pick_scene<BATCH>(sc, sc0).tris;
        // Interior nodes until every busy lane holds a leaf (or is done).
        while (node >= 0 && node != kSent) {
            node = visit4<COUNT>(nodes, node, inv, oinv, tbest, push, pop, cnt);
            if (node < 0 && leaf == 0) {
                leaf = node;
                node = pop();
            }
            if (__all(leaf != 0)) break;
        }
        // Leaves.
        while (leaf < 0) {
            uint32_t x = ~(uint32_t)leaf;
            uint32_t first = x >> kLeafCountBits, n = (x & (kMaxLeafTris - 1)) + 1;
            for (uint32_t k = first; k < first + n; k++) {
                float t, u, v;
                if (COUNT) cnt.tris++;
                if (tri_hit(tris, k, o, d, tbest, t, u, v)) {
                    hit = (int32_t)k;
                    if (ANY) break;
                    tbest = t;
                    hu = u;
                    hv = v;
                }
            }
            if (ANY && hit >= 0) {
                node = kSent;
                leaf = 0;
                sp = 0;
                spilled = 0;
                break;
            }
            leaf = 0;
            if (node < 0) {
                leaf = node;
                node = pop();
            }
        }
        if (active && node == kSent && leaf == 0) {
            done(idx, hit, tbest, hu, hv, pick_scene<BATCH>(sc, sc0));
            active = false;
        }
    }
}

// ---- standalone trace (spk_trace) --------------------------------------------------------
template <int BLOCK, int S, bool ANY>
__global__ __launch_bounds__(BLOCK) __attribute__((amdgpu_waves_per_eu(8))) void k_trace(SceneDev sc, TraceBufs b, uint32_t* work, uint32_t* spill_all,
                                                 int spill_cap) {
    __shared__ uint32_t lds[S * BLOCK];
    uint32_t* col = lds + threadIdx.x;
    uint32_t* spill = spill_all + (size_t)(blockIdx.x * BLOCK + threadIdx.x) * spill_cap;
    const uint64_t n = (uint64_t)b.n;
    const WaveSplit ws = wave_split<BLOCK>();
    for (uint64_t base = ws.first; base < n; base += ws.stride) {
        uint64_t i = base + __lane_id();
        bool active = i < n;
        float3 o = mk(0, 0, 1), d = mk(0, 0, 1);
        float tmax = kINFTY;
        if (active) {
            o = mk(b.org[3 * i], b.org[3 * i + 1], b.org[3 * i + 2]);
            d = normalize(mk(b.dir[3 * i], b.dir[3 * i + 1], b.dir[3 * i + 2]));
            if (b.tmax) tmax = b.tmax[i];
        }
        float t, u, v;
        Counters cnt = {0, 0};
        int32_t k = traverse<BLOCK, S, ANY, false>(sc, o, d, tmax, t, u, v, col, spill, active, cnt);
        if (active) {
            if (ANY) {
                b.prim_out[i] = k >= 0 ? 1 : 0;
            } else {
                b.prim_out[i] = k >= 0 ? sc.tri_ids[k].x : -1;
                if (b.t_out) b.t_out[i] = k >= 0 ? t : -1.0f;
            }
        }
    }
}

// ---- pass bookkeeping -----------------------------------------------------------------------
__global__ void k_set_pass(PassState* ps, uint32_t index) {
    if (threadIdx.x == 0 && blockIdx.x == 0) ps->pass_index = index;
}

template <bool BATCH>
__global__ void k_pass_init(const ConfigDev* __restrict__ cfg, PassState* ps, const PassDev* __restrict__ passes,
                            uint32_t num_passes) {
    // One block: thread 0 computes the pass, all threads clear the counters.
    __shared__ uint32_t n_sh;
    if (threadIdx.x == 0) {
        uint32_t j = ps->pass_index;
        if (BATCH) {
            const bool done = j >= num_passes;
            PassDev P = {0u, 0u, 0u, 0u};
            if (!done) P = passes[j];
            ps->done = done ? 1u : 0u;
            ps->chunk_begin = P.chunk_begin;
            ps->chunk_end = P.chunk_end;
            ps->n = P.n_paths;
            ps->n_pixels = P.n_pixels;
            ps->sample_base = 0;
            ps->pixel_base = 0;
            ps->n_pix = 0;
            ps->k_samples = 0;
        } else {
            uint32_t total = cfg->end_sample - cfg->first_sample;
            uint32_t sample_base, pixel_base, n_pix, k;
            bool done;
            if (cfg->chunks > 1) {
                uint32_t s = j / cfg->chunks, c = j % cfg->chunks;
                done = s >= total;
                sample_base = cfg->first_sample + s;
                pixel_base = c * cfg->queue;
                n_pix = min(cfg->queue, cfg->npix - pixel_base);
                k = 1;
            } else {
                uint32_t s0 = j * cfg->kmax;
                done = s0 >= total;
                sample_base = cfg->first_sample + s0;
                pixel_base = 0;
                n_pix = cfg->npix;
                k = done ? 0 : min(cfg->kmax, total - s0);
            }
            ps->done = done ? 1u : 0u;
            ps->sample_base = sample_base;
            ps->pixel_base = pixel_base;
            ps->n_pix = done ? 0u : n_pix;
            ps->k_samples = done ? 0u : k;
            ps->n = done ? 0u : n_pix * k;
        }
        ps->pass_index = j + 1;
        n_sh = ps->n;
    }
    __syncthreads();
    for (int i = threadIdx.x; i < kMaxDepthCap + 2; i += blockDim.x) {
        ps->countA[i] = 0;
        ps->countS[i] = 0;
        ps->countM[i] = 0;
        ps->workA[i] = 0;
        ps->workS[i] = 0;
        ps->workM[i] = 0;
        ps->countG[i] = 0;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        // classic passes start with all their camera paths in the queue; regenerating passes
        // with an empty one that k_plan fills
        ps->countA[0] = cfg->regen ? 0u : n_sh;
        ps->rg_next = 0;
        ps->rg_nin = 0;
        ps->rg_ncont = 0;
        ps->rg_mnew = 0;
        ps->rg_first = 0;
        ps->rg_done = 0;
    }
}

// One iteration of a regenerating pass, before its kernels (one thread).  The queue in `slot`
// holds the continuations the previous iteration produced; the pass's next unstarted camera
// paths fill it up to the in-flight capacity.  The previous iteration's queue counts go to the
// totals (the last iteration's are added by k_film).
__global__ void k_plan(const ConfigDev* __restrict__ cfg, PassState* ps, int slot) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    const int o = slot ^ 1;
    ps->total_extend += ps->rg_nin;
    ps->total_shadow += ps->countS[o];
    ps->total_mis += ps->countM[o];
    ps->total_guided += ps->countG[o];
    ps->countS[o] = 0;
    ps->countM[o] = 0;
    ps->countG[o] = 0;
    const uint32_t ncont = ps->countA[slot];
    const uint32_t next = ps->rg_next, total = ps->done ? 0u : ps->n;
    const uint32_t room = cfg->inflight > ncont ? cfg->inflight - ncont : 0u;
    const uint32_t m = min(room, total - min(next, total));
    ps->rg_ncont = ncont;
    ps->rg_mnew = m;
    ps->rg_first = next;
    ps->rg_next = next + m;
    ps->rg_nin = ncont + m;
    ps->countA[slot] = ncont + m;
    ps->countA[o] = 0;
    ps->total_paths += m;
    ps->rg_done = (ncont + m == 0) ? 1u : 0u;
}

// Last chunk c in [b, e) whose offset field (path_off or pix_off) is <= i.
template <bool PIX>
__device__ __forceinline__ uint32_t find_chunk(const ChunkDev* __restrict__ ch, uint32_t b, uint32_t e, uint32_t i) {
    while (e - b > 1) {
        const uint32_t m = (b + e) >> 1;
        const uint32_t off = PIX ? ch[m].pix_off : ch[m].path_off;
        if (off <= i) b = m;
        else e = m;
    }
    return b;
}

__device__ __forceinline__ uint32_t batch_pixel(uint32_t field) { return field & ((1u << kBatchPixelBits) - 1u); }
__device__ __forceinline__ uint32_t batch_image(uint32_t field) { return field >> kBatchPixelBits; }

// Pixel-order index -> raster (x, y): bands of 8 rows, 8-column blocks inside a
// band, so one wave of 64 consecutive paths covers an 8x8 tile.
__device__ __forceinline__ void swizzle(uint32_t q, int W, int H, int& x, int& y) {
    uint32_t band = q / (8u * (uint32_t)W);
    uint32_t rows = min(8u, (uint32_t)H - 8u * band);
    uint32_t r = q - band * 8u * (uint32_t)W;
    uint32_t blk = r / (8u * rows);
    uint32_t cols = min(8u, (uint32_t)W - 8u * blk);
    uint32_t within = r - blk * 8u * rows;
    x = (int)(8u * blk + within % cols);
    y = (int)(8u * band + within / cols);
}

// ---- camera (cameras/perspective.cc:53-74, integrator.cc:84-86) ---------------------------
template <bool BATCH>
__global__ __launch_bounds__(256) void k_camera(const ConfigDev* __restrict__ cfg, RenderBufs rb, int slot) {
    const PassState* ps = rb.ps;
    // classic: the pass's paths 0 .. n-1 at queue positions 0 .. n-1; regen: paths first ..
    // first + n - 1 appended after the n_cont continuations of queue `slot`
    const bool rg = cfg->regen != 0;
    const uint32_t n = rg ? ps->rg_mnew : ps->n;
    const uint32_t first = rg ? ps->rg_first : 0u, dst0 = rg ? ps->rg_ncont : 0u;
    const int q = rg ? (slot & 1) : 0;
    const uint32_t n_pix = ps->n_pix, pixel_base = ps->pixel_base, sample_base = ps->sample_base;
    const uint32_t cb = ps->chunk_begin, ce = ps->chunk_end;
    const int kind = cfg->sampler;
    for (uint32_t t = blockIdx.x * blockDim.x + threadIdx.x; t < n; t += gridDim.x * blockDim.x) {
        const uint32_t i = first + t;  // path id within the pass
        int x, y;
        uint32_t pixel, sample, flags, field;
        CameraDev cam;
        uint32_t seed;
        if (BATCH) {
            const ChunkDev ch = rb.chunks[find_chunk<false>(rb.chunks, cb, ce, i)];
            const uint32_t local = i - ch.path_off;
            const uint32_t s_local = local / ch.n_pix, j = local - s_local * ch.n_pix;
            const ImageDev& im = rb.images[ch.img];
            cam = im.cam;
            seed = im.seed;
            swizzle(ch.pixel_base + j, cam.W, cam.H, x, y);
            pixel = (uint32_t)(y * cam.W + x);
            sample = ch.sample_base + s_local;
            flags = sample << 7;
            field = pixel | (ch.img << kBatchPixelBits);
        } else {
            cam = cfg->cam;
            seed = cfg->seed;
            const uint32_t s_local = i / n_pix, j = i - s_local * n_pix;
            swizzle(pixel_base + j, cfg->W, cfg->H, x, y);
            pixel = (uint32_t)(y * cfg->W + x);
            sample = sample_base + s_local;
            flags = s_local << 7;
            field = pixel;
        }
        const int W = cam.W, H = cam.H;
        float2 uf = sample_2d(kind, seed, pixel, sample, 0);
        float sx = 2.0f * ((float)x + uf.x) / (float)W - 1.0f;
        float sy = 1.0f - 2.0f * ((float)y + uf.y) / (float)H;
        float3 dir = normalize(mk(sx * cam.sx_scale, sy * cam.sy_scale, 1.0f));
        float3 org = mk(0, 0, 0);
        if (cam.lens_r > 0.0f) {
            float2 ul = sample_2d(kind, seed, pixel, sample, 2);
            // sampling.cc:151-164 concentric disk
            float ox = 2.0f * ul.x - 1.0f, oy = 2.0f * ul.y - 1.0f, dx = 0.0f, dy = 0.0f;
            if (ox != 0.0f || oy != 0.0f) {
                float r, th;
                if (fabsf(ox) > fabsf(oy)) {
                    r = ox;
                    th = kPI * (oy / ox) / 4.0f;
                } else {
                    r = oy;
                    th = (kPI / 2.0f) - kPI * (ox / oy) / 4.0f;
                }
                dx = r * cosf(th);
                dy = r * sinf(th);
            }
            float ft = cam.focal / dir.z;
            float3 pf = mul(dir, ft);
            org = mk(dx * cam.lens_r, dy * cam.lens_r, 0.0f);
            dir = normalize(sub(pf, org));
        }
        const float* m = cam.m;
        float3 ow = mk(m[0] * org.x + m[1] * org.y + m[2] * org.z + m[3], m[4] * org.x + m[5] * org.y + m[6] * org.z + m[7],
                       m[8] * org.x + m[9] * org.y + m[10] * org.z + m[11]);
        float3 dw = normalize(mk(m[0] * dir.x + m[1] * dir.y + m[2] * dir.z, m[4] * dir.x + m[5] * dir.y + m[6] * dir.z,
                                 m[8] * dir.x + m[9] * dir.y + m[10] * dir.z));
        rb.qa_o[q][dst0 + t] = make_float4(ow.x, ow.y, ow.z, __uint_as_float(i));
        rb.qa_d[q][dst0 + t] = make_float4(dw.x, dw.y, dw.z, __uint_as_float(flags));
        rb.qa_b[q][dst0 + t] = make_float4(1.0f, 1.0f, 1.0f, __uint_as_float(field));
        rb.path_L[i] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    }
}

// ---- extend: closest hit over the extend queue -------------------------------------------------
template <int BLOCK, int S, bool COUNT, bool BATCH>
__global__ __launch_bounds__(BLOCK) __attribute__((amdgpu_waves_per_eu(traversal_waves<BATCH, COUNT>()))) void k_extend(SceneDev sc, RenderBufs rb, int slot) {
    __shared__ uint32_t lds[S * BLOCK];
    PassState* ps = rb.ps;
    const uint32_t count = ps->countA[slot];
    if (count == 0) return;
    uint32_t* col = lds + threadIdx.x;
    uint32_t* spill = rb.spill + (size_t)(blockIdx.x * BLOCK + threadIdx.x) * rb.spill_cap;
    const float4* __restrict__ qo = rb.qa_o[slot & 1];
    const float4* __restrict__ qd = rb.qa_d[slot & 1];
    const float4* __restrict__ qb = rb.qa_b[slot & 1];
    Counters cnt = {0, 0};
    uint32_t rays = 0;
    auto load = [&](uint32_t i, float3& o, float3& d, float& tmax) -> const SceneDev* {
        o = xyz(qo[i]);
        d = xyz(qd[i]);
        tmax = kINFTY;
        return BATCH ? rb.scenes + batch_image(__float_as_uint(qb[i].w)) : nullptr;
    };
    auto done = [&](uint32_t i, int32_t k, float t, float u, float v, const SceneDev&) {
        rb.hit[i] = make_float4(t, u, v, __int_as_float(k));
        if (COUNT) rays++;
    };
    traverse_stream<BLOCK, S, false, COUNT, BATCH>(sc, rb.scenes, count, load, done, col, spill, cnt);
    if (COUNT) {
        atomicAdd(&ps->count_nodes, (unsigned long long)cnt.nodes);
        atomicAdd(&ps->count_tris, (unsigned long long)cnt.tris);
        atomicAdd(&ps->count_rays, (unsigned long long)rays);
    }
}

// ---- BSDF (core/bsdf.cc, core/bxdf.cc, core/microfacet.cc, core/fresnel.cc) -------------------
enum { BX_REFLECTION = 1, BX_TRANSMISSION = 2, BX_DIFFUSE = 4, BX_GLOSSY = 8, BX_SPECULAR = 16, BX_ALL = 31 };

struct Bsdf {
    int lobe, type, dist;
    float3 R, T, eta, k;  // T is Ks for the plastics
    float etaB, alpha, specW;
    float3 t, b, n;  // frame
};

template <bool W>
__device__ __forceinline__ int lobe_type(int lobe) {
    if (W) {
        if (lobe == LOBE_PLASTIC) return BX_REFLECTION | BX_SPECULAR | BX_DIFFUSE;
        if (lobe == LOBE_MICROTRANS) return BX_REFLECTION | BX_TRANSMISSION | BX_GLOSSY;
        if (lobe == LOBE_ROUGHPLASTIC) return BX_REFLECTION | BX_DIFFUSE | BX_GLOSSY;
    }
    switch (lobe) {
    case LOBE_LAMBERT: return BX_REFLECTION | BX_DIFFUSE;
    case LOBE_SPECREFL: return BX_REFLECTION | BX_SPECULAR;
    case LOBE_FRESNELSPEC: return BX_REFLECTION | BX_TRANSMISSION | BX_SPECULAR;
    case LOBE_MICROREFL: return BX_REFLECTION | BX_GLOSSY;
    default: return 0;
    }
}

__device__ __forceinline__ float3 fr_conductor(float cosI, float3 eta, float3 kk) {
    cosI = clampf(cosI, -1.0f, 1.0f);
    float r[3];
#pragma unroll
    for (int c = 0; c < 3; c++) {
        float e = comp(eta, c), ek = comp(kk, c);
        float cos2 = cosI * cosI, sin2 = 1.0f - cos2;
        float e2 = e * e, ek2 = ek * ek;
        float t0 = e2 - ek2 - sin2;
        float a2b2 = sqrtf(t0 * t0 + 4.0f * e2 * ek2);
        float t1 = a2b2 + cos2;
        float a = sqrtf(0.5f * (a2b2 + t0));
        float t2 = 2.0f * cosI * a;
        float Rs = (t1 - t2) / (t1 + t2);
        float t3 = cos2 * a2b2 + sin2 * sin2;
        float t4 = t2 * sin2;
        float Rp = Rs * (t3 - t4) / (t3 + t4 + kEPS);
        r[c] = 0.5f * (Rp + Rs);
    }
    return mk(r[0], r[1], r[2]);
}

__device__ __forceinline__ float fr_dielectric(float cosI, float etaI, float etaT) {
    cosI = clampf(cosI, -1.0f, 1.0f);
    if (!(cosI > 0.0f)) {
        float tmp = etaI;
        etaI = etaT;
        etaT = tmp;
        cosI = fabsf(cosI);
    }
    float sinI = sqrtf(fmaxf(0.0f, 1.0f - cosI * cosI));
    float sinT = etaI / etaT * sinI;
    if (sinT >= 1.0f) return 1.0f;
    float cosT = sqrtf(fmaxf(0.0f, 1.0f - sinT * sinT));
    float rpar = ((etaT * cosI) - (etaI * cosT)) / ((etaT * cosI) + (etaI * cosT));
    float rper = ((etaI * cosI) - (etaT * cosT)) / ((etaI * cosI) + (etaT * cosT));
    return 0.5f * (rpar * rpar + rper * rper);
}

__device__ __forceinline__ float sin2t(float3 w) { return fmaxf(0.0f, 1.0f - w.z * w.z); }
__device__ __forceinline__ float cos_phi(float3 w) {
    float s = sqrtf(sin2t(w));
    return s == 0.0f ? 1.0f : clampf(w.x / s, -1.0f, 1.0f);
}
__device__ __forceinline__ float sin_phi(float3 w) {
    float s = sqrtf(sin2t(w));
    return s == 0.0f ? 0.0f : clampf(w.y / s, -1.0f, 1.0f);
}
__device__ __forceinline__ float ggx_D(float alpha, float3 wh) {
    float c2 = wh.z * wh.z;
    float t2 = sin2t(wh) / (c2 + kEPS);
    if (__builtin_isinf(t2)) return 0.0f;
    float cp = cos_phi(wh), sp = sin_phi(wh);
    float e = (cp * cp / (alpha * alpha) + sp * sp / (alpha * alpha)) * t2;
    return 1.0f / (kPI * alpha * alpha * c2 * c2 * (1.0f + e) * (1.0f + e));
}
__device__ __forceinline__ float ggx_lambda(float alpha, float3 w) {
    float at = fabsf(sqrtf(sin2t(w)) / (w.z + kEPS));
    if (__builtin_isinf(at)) return 0.0f;
    float cp = cos_phi(w), sp = sin_phi(w);
    float al = sqrtf(cp * cp * alpha * alpha + sp * sp * alpha * alpha);
    float a = al * at;
    return (-1.0f + sqrtf(1.0f + a * a)) / 2.0f;
}
// microfacet.cc:15-56: visible-slope sampling, rational fit for the y slope.
__device__ __forceinline__ void ggx_sample11(float cosT, float u0, float u1, float& sx, float& sy) {
    if (cosT > 0.9999f) {
        float r = sqrtf(u0 / (1.0f - u0));
        float phi = 2.0f * kPI * u1;
        sx = r * cosf(phi);
        sy = r * sinf(phi);
        return;
    }
    float sinT = sqrtf(fmaxf(0.0f, 1.0f - cosT * cosT));
    float tanT = sinT / cosT;
    float a = 1.0f / tanT;
    float G1 = 2.0f / (1.0f + sqrtf(1.0f + 1.0f / (a * a)));
    float A = 2.0f * u0 / G1 - 1.0f;
    float tmp = 1.0f / (A * A - 1.0f);
    if (tmp > 1.0e10f) tmp = 1.0e10f;
    float B = tanT;
    float D = sqrtf(fmaxf(0.0f, B * B * tmp * tmp - (A * A - B * B) * tmp));
    float x1 = B * tmp - D, x2 = B * tmp + D;
    sx = (A < 0.0f || x2 > 1.0f / tanT) ? x1 : x2;
    float S, U;
    if (u1 > 0.5f) {
        S = 1.0f;
        U = 2.0f * (u1 - 0.5f);
    } else {
        S = -1.0f;
        U = 2.0f * (0.5f - u1);
    }
    float z = (U * (U * (U * 0.27385f - 0.73369f) + 0.46341f)) / (U * (U * (U * 0.093073f + 0.309420f) - 1.0f) + 0.597999f);
    sy = S * z * sqrtf(1.0f + sx * sx);
}
__device__ __forceinline__ float3 ggx_sample_wh(float alpha, float3 wo, float u0, float u1) {
    bool flip = wo.z < 0.0f;
    float3 wi = flip ? neg(wo) : wo;
    float3 ws = normalize(mk(alpha * wi.x, alpha * wi.y, wi.z));
    float sx, sy;
    ggx_sample11(ws.z, u0, u1, sx, sy);
    float cp = cos_phi(ws), sp = sin_phi(ws);
    float tmp = cp * sx - sp * sy;
    sy = sp * sx + cp * sy;
    sx = tmp * alpha;
    sy = sy * alpha;
    float3 wh = normalize(mk(-sx, -sy, 1.0f));
    return flip ? neg(wh) : wh;
}
// math.h:14-65: erfinv (Giles 2011, single-precision coefficients) and erf
// (Abramowitz & Stegun 7.1.26), as spica's Beckmann sampling uses them.
__device__ __forceinline__ float erfinv_s(float x) {
    x = clampf(x, -0.99999f, 0.99999f);
    float w = -logf((1.0f - x) * (1.0f + x)), p;
    if (w < 5.0f) {
        w = w - 2.5f;
        p = 2.81022636e-08f;
        p = 3.43273939e-07f + p * w;
        p = -3.5233877e-06f + p * w;
        p = -4.39150654e-06f + p * w;
        p = 0.00021858087f + p * w;
        p = -0.00125372503f + p * w;
        p = -0.00417768164f + p * w;
        p = 0.246640727f + p * w;
        p = 1.50140941f + p * w;
    } else {
        w = sqrtf(w) - 3.0f;
        p = -0.000200214257f;
        p = 0.000100950558f + p * w;
        p = 0.00134934322f + p * w;
        p = -0.00367342844f + p * w;
        p = 0.00573950773f + p * w;
        p = -0.0076224613f + p * w;
        p = 0.00943887047f + p * w;
        p = 1.00167406f + p * w;
        p = 2.83297682f + p * w;
    }
    return p * x;
}
__device__ __forceinline__ float erf_s(float x) {
    const float sign = x < 0.0f ? -1.0f : 1.0f;
    x = fabsf(x);
    const float t = 1.0f / (1.0f + 0.3275911f * x);
    const float y = 1.0f - (((((1.061405429f * t - 1.453152027f) * t) + 1.421413741f) * t - 0.284496736f) * t +
                            0.254829592f) * t * expf(-x * x);
    return sign * y;
}
// microfacet.cc:277-288
__device__ __forceinline__ float bk_D(float alpha, float3 wh) {
    float c2 = wh.z * wh.z;
    float t2 = sin2t(wh) / (c2 + kEPS);
    if (__builtin_isinf(t2)) return 0.0f;
    float cp = cos_phi(wh), sp = sin_phi(wh);
    float ab2 = cp * cp / (alpha * alpha) + sp * sp / (alpha * alpha);
    return expf(-t2 * ab2) / (kPI * alpha * alpha * c2 * c2);
}
// microfacet.cc:342-355 (Walter et al. 2007's rational fit of the Smith lambda)
__device__ __forceinline__ float bk_lambda(float alpha, float3 w) {
    float at = fabsf(sqrtf(sin2t(w)) / (w.z + kEPS));
    if (__builtin_isinf(at)) return 0.0f;
    float cp = cos_phi(w), sp = sin_phi(w);
    float al = sqrtf(cp * cp * alpha * alpha + sp * sp * alpha * alpha);
    float a = 1.0f / (al * at);
    if (a >= 1.6f) return 0.0f;
    return (1.0f - 1.259f * a + 0.396f * a * a) / (3.535f * a + 2.181f * a * a);
}
// microfacet.cc:80-131: visible slopes by Jakob 2014's bisection-Newton solve (16 steps at most)
__device__ void bk_sample11(float cosI, float U1, float U2, float& sx, float& sy) {
    if (cosI > 0.9999f) {
        float r = sqrtf(-logf(1.0f - U1));
        sx = r * cosf(2.0f * kPI * U2);
        sy = r * sinf(2.0f * kPI * U2);
        return;
    }
    const float sinI = sqrtf(fmaxf(0.0f, 1.0f - cosI * cosI));
    const float tanI = sinI / cosI, cotI = 1.0f / tanI;
    float a = -1.0f, c = erf_s(cotI);
    const float x0 = fmaxf(U1, 1.0e-6f);
    const float th = acosf(cosI);
    const float fit = 1.0f + th * (-0.876f + th * (0.4265f - 0.0594f * th));
    float bb = c - (1.0f + c) * powf(1.0f - x0, fit);
    const float kSqrtPiInv = 0.5641895835477563f;  // 1 / sqrt(pi)
    const float norm = 1.0f / (1.0f + c + kSqrtPiInv * tanI * expf(-cotI * cotI));
    for (int it = 0; it < 16; it++) {
        if (bb < a || c < bb) bb = 0.5f * (a + c);
        const float xm = erfinv_s(bb);
        const float value = norm * (1.0f + bb + kSqrtPiInv * tanI * expf(-xm * xm)) - x0;
        const float deriv = norm * (1.0f - xm * tanI);
        if (fabsf(value) < 1.0e-6f) break;
        if (value > 0.0f) c = bb;
        else a = bb;
        bb -= value / deriv;
    }
    sx = erfinv_s(bb);
    sy = erfinv_s(2.0f * fmaxf(U2, 1.0e-6f) - 1.0f);
}
__device__ __forceinline__ float3 bk_sample_wh(float alpha, float3 wo, float u0, float u1) {
    bool flip = wo.z < 0.0f;
    float3 wi = flip ? neg(wo) : wo;
    float3 ws = normalize(mk(alpha * wi.x, alpha * wi.y, wi.z));
    float sx, sy;
    bk_sample11(ws.z, u0, u1, sx, sy);
    float cp = cos_phi(ws), sp = sin_phi(ws);
    float tmp = cp * sx - sp * sy;
    sy = (sp * sx + cp * sy) * alpha;
    sx = tmp * alpha;
    float3 wh = normalize(mk(-sx, -sy, 1.0f));
    return flip ? neg(wh) : wh;
}
// MicrofacetDistribution of the material's "distribution" (microfacet.cc:166-193)
// W (wide): the kernel may meet Beckmann and the plastic / glossy-transmission lobes.
// Scenes without them run kernels built with W = false, which leave that code out and
// keep the registers of the four basic lobes.
template <bool W>
__device__ __forceinline__ float mf_D(const Bsdf& b, float3 wh) {
    return (W && b.dist) ? bk_D(b.alpha, wh) : ggx_D(b.alpha, wh);
}
template <bool W>
__device__ __forceinline__ float mf_lambda(const Bsdf& b, float3 w) {
    return (W && b.dist) ? bk_lambda(b.alpha, w) : ggx_lambda(b.alpha, w);
}
template <bool W>
__device__ __forceinline__ float mf_G(const Bsdf& b, float3 wo, float3 wi, float3 wh) {
    if (wi.z * dot(wi, wh) <= 0.0f) return 0.0f;
    if (wo.z * dot(wo, wh) <= 0.0f) return 0.0f;
    return 1.0f / (1.0f + mf_lambda<W>(b, wo) + mf_lambda<W>(b, wi));
}
template <bool W>
__device__ __forceinline__ float mf_pdf(const Bsdf& b, float3 wo, float3 wh) {
    float g1 = wo.z * dot(wo, wh) <= 0.0f ? 0.0f : 1.0f / (1.0f + mf_lambda<W>(b, wo));
    return mf_D<W>(b, wh) * g1 * fabsf(dot(wo, wh)) / fabsf(wo.z);
}
template <bool W>
__device__ __forceinline__ float3 mf_sample_wh(const Bsdf& b, float3 wo, float u0, float u1) {
    return (W && b.dist) ? bk_sample_wh(b.alpha, wo, u0, u1) : ggx_sample_wh(b.alpha, wo, u0, u1);
}
// fresnel.cc:87-95
__device__ __forceinline__ float fr_diffuse_reflectance(float eta) {
    if (eta >= 1.0f) return -1.4399f / (eta * eta) + 0.7099f / eta + 0.6681f + 0.0636f * eta;
    return -0.4399f + 0.7099f / eta - 0.3319f / (eta * eta) + 0.0636f / (eta * eta * eta);
}
template <bool W>
__device__ __forceinline__ float3 micro_f(const Bsdf& b, float3 wo, float3 wi) {
    float cO = fabsf(wo.z), cI = fabsf(wi.z);
    float3 wh = add(wi, wo);
    if (cI == 0.0f || cO == 0.0f) return mk(0, 0, 0);
    if (wh.x == 0.0f && wh.y == 0.0f && wh.z == 0.0f) return mk(0, 0, 0);
    wh = normalize(wh);
    float3 F = fr_conductor(fabsf(dot(wi, wh)), b.eta, b.k);
    float s = mf_D<W>(b, wh) * mf_G<W>(b, wo, wi, wh) / (4.0f * cI * cO);
    return mul(cmul(b.R, F), s);
}
__device__ __forceinline__ bool refract_v(float3 wi, float3 n, float eta, float3& wt) {
    float cosI = dot(n, wi);
    float sin2I = fmaxf(0.0f, 1.0f - cosI * cosI);
    float sin2T = eta * eta * sin2I;
    if (sin2T >= 1.0f) return false;
    float cosT = sqrtf(1.0f - sin2T);
    wt = add(mul(neg(wi), eta), mul(n, eta * cosI - cosT));
    return true;
}

// bxdf.cc:333-354: MicrofacetTransmission::f(wo, wi, wh)
__device__ float3 mtrans_f_wh(const Bsdf& b, float3 wo, float3 wi, float3 wh) {
    const float cO = wo.z, cI = wi.z;
    if (cO == 0.0f || cI == 0.0f) return mk(0, 0, 0);
    if (wo.z * dot(wo, wh) <= 0.0f || wi.z * dot(wi, wh) <= 0.0f) return mk(0, 0, 0);
    const float F = fr_dielectric(dot(wo, wh), 1.0f, b.etaB);
    if (wo.z * wi.z > 0.0f) return mul(b.R, F * mf_D<true>(b, wh) * mf_G<true>(b, wo, wi, wh) / (4.0f * cO * cI));
    const float eta = dot(wo, wh) > 0.0f ? b.etaB : 1.0f / b.etaB;
    const float factor = 1.0f / eta;
    const float sd = dot(wo, wh) + eta * dot(wi, wh);
    return mul(b.T, (1.0f - F) * fabsf(mf_D<true>(b, wh) * mf_G<true>(b, wo, wi, wh) * eta * eta * fabsf(dot(wi, wh)) *
                                      fabsf(dot(wo, wh)) * factor * factor / (cI * cO * sd * sd)));
}
// bxdf.cc:408-423
__device__ float mtrans_pdf_wh(const Bsdf& b, float3 wo, float3 wi, float3 wh) {
    if (wo.z * dot(wo, wh) <= 0.0f || wi.z * dot(wi, wh) <= 0.0f) return 0.0f;
    const float F = fr_dielectric(dot(wo, wh), 1.0f, b.etaB);
    if (wo.z * wi.z > 0.0f) return F * mf_pdf<true>(b, wo, wh) / (4.0f * fabsf(dot(wo, wh)));
    const float eta = dot(wo, wh) > 0.0f ? b.etaB : 1.0f / b.etaB;
    const float sd = dot(wo, wh) + eta * dot(wi, wh);
    return (1.0f - F) * mf_pdf<true>(b, wo, wh) * fabsf((eta * eta * dot(wi, wh)) / (sd * sd));
}
// bxdf.cc:315-331, 394-406: the half vector of a (wo, wi) pair on the +z side
__device__ __forceinline__ float3 mtrans_wh(const Bsdf& b, float3 wo, float3 wi) {
    float3 wh = wo.z * wi.z > 0.0f ? normalize(add(wo, wi))
                                   : normalize(add(wo, mul(wi, wo.z > 0.0f ? b.etaB : 1.0f / b.etaB)));
    return wh.z >= 0.0f ? wh : neg(wh);
}
// plastic.cc:28, roughplastic.cc:28: probability of the specular branch
__device__ __forceinline__ float plastic_prob_spec(const Bsdf& b, float Fo) {
    return (Fo * b.specW) / (Fo * b.specW + (1.0f - Fo) * (1.0f - b.specW));
}
// The plastics' diffuse term with the internal-scattering normalisation (plastic.cc:62-67)
__device__ __forceinline__ float3 plastic_diffuse(const Bsdf& b, float Fo, float cosI) {
    const float invEta = 1.0f / b.etaB;
    const float Fi = fr_dielectric(cosI, 1.0f, b.etaB);
    return mul(b.R, (1.0f - Fo) * (1.0f - Fi) * invEta * invEta * kINV_PI / (1.0f - fr_diffuse_reflectance(b.etaB)));
}
__device__ __forceinline__ float3 cosine_hemisphere(float u0, float u1) {
    float ox = 2.0f * u0 - 1.0f, oy = 2.0f * u1 - 1.0f, dx = 0.0f, dy = 0.0f;
    if (ox != 0.0f || oy != 0.0f) {
        float r, th;
        if (fabsf(ox) > fabsf(oy)) {
            r = ox;
            th = kPI * (oy / ox) / 4.0f;
        } else {
            r = oy;
            th = (kPI / 2.0f) - kPI * (ox / oy) / 4.0f;
        }
        dx = r * cosf(th);
        dy = r * sinf(th);
    }
    return mk(dx, dy, sqrtf(fmaxf(0.0f, 1.0f - dx * dx - dy * dy)));
}

// The plastics and the glossy transmission (wide kernels only).
__device__ __forceinline__ float3 lobe_f_ext(const Bsdf& b, float3 wo, float3 wi) {
    switch (b.lobe) {
    case LOBE_PLASTIC: {  // plastic.cc:32-46 (the "specular" test compares wi with its own mirror image)
        const float Fo = fr_dielectric(wo.z, 1.0f, b.etaB);
        if (dot(wi, mk(-wi.x, -wi.y, wi.z)) > 1.0f - kEPS) return mul(b.T, Fo / fabsf(wi.z));
        return plastic_diffuse(b, Fo, wi.z);
    }
    case LOBE_MICROTRANS:  // bxdf.cc:315-331
        if (wo.z == 0.0f || wi.z == 0.0f) return mk(0, 0, 0);
        return mtrans_f_wh(b, wo, wi, mtrans_wh(b, wo, wi));
    case LOBE_ROUGHPLASTIC: {  // roughplastic.cc:35-47
        if (wo.z == 0.0f || wi.z == 0.0f || !(wo.z * wi.z > 0.0f)) return mk(0, 0, 0);
        const float3 wh = normalize(add(wi, wo));
        const float F = fr_dielectric(dot(wo, wh), 1.0f, b.etaB);
        return add(mul(b.T, F * mf_D<true>(b, wh) * mf_G<true>(b, wo, wi, wh) / (4.0f * wo.z * wi.z)),
                   mul(b.R, (1.0f - F) * kINV_PI));
    }
    default:
        return mk(0, 0, 0);
    }
}
__device__ __forceinline__ float lobe_pdf_ext(const Bsdf& b, float3 wo, float3 wi) {
    switch (b.lobe) {
    case LOBE_PLASTIC: {  // plastic.cc:71-82
        const float ps = plastic_prob_spec(b, fr_dielectric(wo.z, 1.0f, b.etaB));
        if (dot(wi, mk(-wi.x, -wi.y, wi.z)) > 1.0f - kEPS) return ps;
        return (1.0f - ps) * fabsf(wi.z) * kINV_PI;
    }
    case LOBE_MICROTRANS:  // bxdf.cc:394-406
        return mtrans_pdf_wh(b, wo, wi, mtrans_wh(b, wo, wi));
    case LOBE_ROUGHPLASTIC: {  // roughplastic.cc:88-97 (F-weighted, not the sampling weights)
        if (!(wo.z * wi.z > 0.0f)) return 0.0f;
        const float3 wh = normalize(add(wi, wo));
        const float F = fr_dielectric(dot(wo, wh), 1.0f, b.etaB);
        return F * mf_pdf<true>(b, wo, wh) / (4.0f * fabsf(dot(wo, wh))) + (1.0f - F) * fabsf(wi.z) * kINV_PI;
    }
    default:
        return 0.0f;
    }
}
__device__ __forceinline__ float3 lobe_sample_ext(const Bsdf& b, float3 wo, float3& wi, float u0, float u1, float& pdf,
                                               int& stype, float uh) {
    switch (b.lobe) {
    case LOBE_PLASTIC: {  // plastic.cc:48-69 (no hemisphere flip of the diffuse sample)
        const float Fo = fr_dielectric(wo.z, 1.0f, b.etaB);
        const float ps = plastic_prob_spec(b, Fo);
        if (uh < ps) {
            wi = mk(-wo.x, -wo.y, wo.z);
            pdf = ps;
            return mul(b.T, Fo / fabsf(wi.z));
        }
        wi = cosine_hemisphere(u0, u1);
        pdf = (1.0f - ps) * fabsf(wi.z) * kINV_PI;
        return plastic_diffuse(b, Fo, wi.z);
    }
    case LOBE_MICROTRANS: {  // bxdf.cc:356-392
        if (wo.z == 0.0f) return mk(0, 0, 0);
        const float3 wh = mf_sample_wh<true>(b, wo, u0, u1);
        const float cO = dot(wo, wh.z > 0.0f ? wh : neg(wh));
        const float F = fr_dielectric(cO, 1.0f, b.etaB);
        if (uh < F) {
            wi = sub(mul(wh, 2.0f * dot(wh, wo)), wo);
            if (!(wo.z * wi.z > 0.0f)) return mk(0, 0, 0);
            stype = BX_REFLECTION | BX_GLOSSY;
        } else {
            if (!refract_v(wo, wh, cO > 0.0f ? 1.0f / b.etaB : b.etaB, wi)) return mk(0, 0, 0);
            if (wo.z * wi.z > 0.0f) return mk(0, 0, 0);
            stype = BX_TRANSMISSION | BX_GLOSSY;
        }
        const float3 whp = wh.z >= 0.0f ? wh : neg(wh);
        pdf = mtrans_pdf_wh(b, wo, wi, whp);
        return mtrans_f_wh(b, wo, wi, whp);
    }
    case LOBE_ROUGHPLASTIC: {  // roughplastic.cc:49-80 (the diffuse branch reuses u0, u1)
        if (wo.z == 0.0f) return mk(0, 0, 0);
        const float3 wh = mf_sample_wh<true>(b, wo, u0, u1);
        const float Fo = fr_dielectric(dot(wo, wh), 1.0f, b.etaB);
        const float ps = plastic_prob_spec(b, Fo);
        if (uh < ps) {
            wi = sub(mul(wh, 2.0f * dot(wh, wo)), wo);
            if (!(wo.z * wi.z > 0.0f)) return mk(0, 0, 0);
            pdf = ps * mf_pdf<true>(b, wo, wh) / (4.0f * fabsf(dot(wo, wh)));
            return mul(b.T, Fo * mf_D<true>(b, wh) * mf_G<true>(b, wo, wi, wh) / (4.0f * wo.z * wi.z));
        }
        wi = cosine_hemisphere(u0, u1);
        pdf = (1.0f - ps) * fabsf(wi.z) * kINV_PI;
        return plastic_diffuse(b, Fo, wi.z);
    }
    default:
        return mk(0, 0, 0);
    }
}

template <bool W>
__device__ float3 lobe_f(const Bsdf& b, float3 wo, float3 wi) {
    const float deltaEps = 1.0e-4f;
    switch (b.lobe) {
    case LOBE_LAMBERT:
        return mul(b.R, kINV_PI);
    case LOBE_SPECREFL:
        if (dot(mk(-wo.x, -wo.y, wo.z), wi) > 1.0f - deltaEps) return mul(b.R, 1.0f / fabsf(wi.z));
        return mk(0, 0, 0);
    case LOBE_FRESNELSPEC: {
        float F = fr_dielectric(wo.z, 1.0f, b.etaB);
        if (wo.z * wi.z > 0.0f) {
            if (dot(mk(-wo.x, -wo.y, wo.z), wi) > 1.0f - deltaEps) return mul(b.R, F / fabsf(wi.z));
        } else {
            bool entering = wo.z > 0.0f;
            float etaI = entering ? 1.0f : b.etaB, etaT = entering ? b.etaB : 1.0f;
            float3 wt;
            if (!refract_v(wo, wo.z < 0.0f ? mk(0, 0, -1) : mk(0, 0, 1), etaI / etaT, wt)) return mk(0, 0, 0);
            if (dot(wi, wt) > 1.0f - deltaEps) return mul(b.T, (1.0f - F) / fabsf(wi.z));
        }
        return mk(0, 0, 0);
    }
    case LOBE_MICROREFL:
        return micro_f<W>(b, wo, wi);
    case LOBE_PLASTIC:
    case LOBE_MICROTRANS:
    case LOBE_ROUGHPLASTIC:
        if (W) return lobe_f_ext(b, wo, wi);
        return mk(0, 0, 0);
    default:
        return mk(0, 0, 0);
    }
}
template <bool W>
__device__ float lobe_pdf(const Bsdf& b, float3 wo, float3 wi) {
    const float deltaEps = 1.0e-4f;
    switch (b.lobe) {
    case LOBE_LAMBERT:
        return wo.z * wi.z > 0.0f ? fabsf(wi.z) * kINV_PI : 0.0f;
    case LOBE_SPECREFL:
        return dot(mk(-wo.x, -wo.y, wo.z), wi) > 1.0f - deltaEps ? 1.0f : 0.0f;
    case LOBE_FRESNELSPEC: {
        float F = fr_dielectric(wo.z, 1.0f, b.etaB);
        if (wo.z * wi.z > 0.0f) {
            if (dot(mk(-wo.x, -wo.y, wo.z), wi) > 1.0f - deltaEps) return F;
        } else {
            bool entering = wo.z > 0.0f;
            float etaI = entering ? 1.0f : b.etaB, etaT = entering ? b.etaB : 1.0f;
            float3 wt;
            if (!refract_v(wo, wo.z < 0.0f ? mk(0, 0, -1) : mk(0, 0, 1), etaI / etaT, wt)) return 0.0f;
            if (dot(wi, wt) > 1.0f - deltaEps) return 1.0f - F;
        }
        return 0.0f;
    }
    case LOBE_MICROREFL: {
        if (!(wo.z * wi.z > 0.0f)) return 0.0f;
        float3 wh = normalize(add(wo, wi));
        return mf_pdf<W>(b, wo, wh) / (4.0f * dot(wo, wh));
    }
    case LOBE_PLASTIC:
    case LOBE_MICROTRANS:
    case LOBE_ROUGHPLASTIC:
        if (W) return lobe_pdf_ext(b, wo, wi);
        return 0.0f;
    default:
        return 0.0f;
    }
}
// Returns f; pdf stays 0 on failure; stype receives the sampled component type.  uh is
// the BxDF's hidden uniform: spica's MicrofacetTransmission, PlasticBRDF and
// RoughPlasticBRDF draw it from a thread-local Random; here it is a sampler dimension.
template <bool W>
__device__ float3 lobe_sample(const Bsdf& b, float3 wo, float3& wi, float u0, float u1, float& pdf, int& stype, float uh) {
    switch (b.lobe) {
    case LOBE_LAMBERT: {
        wi = cosine_hemisphere(u0, u1);
        if (wo.z < 0.0f) wi.z *= -1.0f;
        pdf = lobe_pdf<W>(b, wo, wi);
        return lobe_f<W>(b, wo, wi);
    }
    case LOBE_SPECREFL: {
        wi = mk(-wo.x, -wo.y, wo.z);
        pdf = 1.0f;
        float3 F = fr_conductor(fabsf(wi.z), b.eta, b.k);
        return mul(cmul(F, b.R), 1.0f / fabsf(wi.z));
    }
    case LOBE_FRESNELSPEC: {
        float F = fr_dielectric(wo.z, 1.0f, b.etaB);
        if (u0 < F) {
            wi = mk(-wo.x, -wo.y, wo.z);
            stype = BX_SPECULAR | BX_REFLECTION;
            pdf = F;
            return mul(b.R, F / fabsf(wi.z));
        }
        bool entering = wo.z > 0.0f;
        float etaI = entering ? 1.0f : b.etaB, etaT = entering ? b.etaB : 1.0f;
        if (!refract_v(wo, wo.z < 0.0f ? mk(0, 0, -1) : mk(0, 0, 1), etaI / etaT, wi)) return mk(0, 0, 0);
        float3 ft = mul(b.T, (1.0f - F) * ((etaI * etaI) / (etaT * etaT)));
        stype = BX_SPECULAR | BX_TRANSMISSION;
        pdf = 1.0f - F;
        return mul(ft, 1.0f / fabsf(wi.z));
    }
    case LOBE_MICROREFL: {
        if (wo.z == 0.0f) return mk(0, 0, 0);
        float3 wh = mf_sample_wh<W>(b, wo, u0, u1);
        wi = sub(mul(wh, 2.0f * dot(wh, wo)), wo);
        if (!(wo.z * wi.z > 0.0f)) return mk(0, 0, 0);
        pdf = mf_pdf<W>(b, wo, wh) / (4.0f * dot(wo, wh));
        return micro_f<W>(b, wo, wi);
    }
    case LOBE_PLASTIC:
    case LOBE_MICROTRANS:
    case LOBE_ROUGHPLASTIC:
        if (W) return lobe_sample_ext(b, wo, wi, u0, u1, pdf, stype, uh);
        return mk(0, 0, 0);
    default:
        return mk(0, 0, 0);
    }
}
__device__ __forceinline__ float3 to_local(const Bsdf& b, float3 v) { return mk(dot(b.t, v), dot(b.b, v), dot(b.n, v)); }
__device__ __forceinline__ float3 to_world(const Bsdf& b, float3 v) {
    return add(add(mul(b.t, v.x), mul(b.b, v.y)), mul(b.n, v.z));
}
// numComponents(All & ~Specular) > 0 (bsdf.cc:28-36, subset rule): the path's NEE gate
__device__ __forceinline__ bool bsdf_has_nonspecular(const Bsdf& b) {
    return b.lobe != LOBE_NONE && (b.type & BX_SPECULAR) == 0;
}
// BxDFs whose sample() draws a hidden uniform (only in wide kernels)
template <bool W>
__device__ __forceinline__ bool lobe_has_hidden(const Bsdf& b) { return W && b.lobe >= LOBE_PLASTIC; }
// bsdf.cc:49-63
template <bool W>
__device__ float3 bsdf_f(const Bsdf& b, float3 woW, float3 wiW, int mask) {
    if (b.lobe == LOBE_NONE) return mk(0, 0, 0);
    bool reflect = dot(wiW, b.n) * dot(woW, b.n) > 0.0f;
    if ((b.type & mask) != 0 && ((reflect && (b.type & BX_REFLECTION)) || (!reflect && (b.type & BX_TRANSMISSION))))
        return lobe_f<W>(b, to_local(b, woW), to_local(b, wiW));
    return mk(0, 0, 0);
}
// bsdf.cc:124-136
template <bool W>
__device__ float bsdf_pdf(const Bsdf& b, float3 woW, float3 wiW, int mask) {
    if (b.lobe == LOBE_NONE || (b.type & mask) == 0) return 0.0f;
    return lobe_pdf<W>(b, to_local(b, woW), to_local(b, wiW));
}
// bsdf.cc:65-122 (one BxDF)
template <bool W>
__device__ float3 bsdf_sample(const Bsdf& b, float3 woW, float3& wiW, float u0, float u1, float& pdf, int mask, int& stype,
                              float uh) {
    pdf = 0.0f;
    stype = 0;
    if (b.lobe == LOBE_NONE || (b.type & mask) != b.type) return mk(0, 0, 0);
    float3 wo = normalize(to_local(b, woW));
    if (wo.z == 0.0f) return mk(0, 0, 0);
    stype = b.type;
    float3 wi = mk(0, 0, 1);
    float3 f = lobe_sample<W>(b, wo, wi, u0, u1, pdf, stype, uh);
    if (pdf == 0.0f) {
        stype = 0;
        return mk(0, 0, 0);
    }
    wiW = normalize(to_world(b, wi));
    return f;
}

// ---- environment map (lights/envmap.cc, core/mipmap.cc, core/sampling.cc) ----------------------
__device__ __forceinline__ float3 env_texel(const EnvDev& e, int s, int t) {
    s = (s % e.w + e.w) % e.w;
    t = (t % e.h + e.h) % e.h;
    return xyz(e.tex[t * e.w + s]);
}
// mipmap.cc:71-79 at level 0 (truncation toward zero, as spica does)
__device__ float3 env_lookup(const EnvDev& e, float u, float v) {
    float s = u * (float)e.w - 0.5f, t = v * (float)e.h - 0.5f;
    int si = (int)s, ti = (int)t;
    float ds = s - (float)si, dt = t - (float)ti;
    float3 r = mul(env_texel(e, si, ti), (1.0f - ds) * (1.0f - dt));
    r = add(r, mul(env_texel(e, si + 1, ti), ds * (1.0f - dt)));
    r = add(r, mul(env_texel(e, si, ti + 1), (1.0f - ds) * dt));
    r = add(r, mul(env_texel(e, si + 1, ti + 1), ds * dt));
    return r;
}
__device__ __forceinline__ float3 m3(const float* m, float3 v) {
    return mk(m[0] * v.x + m[1] * v.y + m[2] * v.z, m[3] * v.x + m[4] * v.y + m[5] * v.z, m[6] * v.x + m[7] * v.y + m[8] * v.z);
}
__device__ __forceinline__ float sph_phi(float3 d) {
    float p = atan2f(d.y, d.x);
    return p < 0.0f ? p + 2.0f * kPI : p;
}
// envmap.cc:130-134
__device__ float3 env_Le(const EnvDev& e, float3 dir) {
    float3 d = normalize(m3(e.w2l, dir));
    return env_lookup(e, sph_phi(d) * (0.5f * kINV_PI), acosf(clampf(d.z, -1.0f, 1.0f)) * kINV_PI);
}
// envmap.cc:81-88
__device__ float env_pdf(const EnvDev& e, float3 dir) {
    float3 d = m3(e.w2l, dir);
    float theta = acosf(clampf(d.z, -1.0f, 1.0f));
    float phi = sph_phi(d);
    float st = sinf(theta);
    if (st == 0.0f) return 0.0f;
    int iu = min(max((int)(phi * (0.5f * kINV_PI) * (float)e.w), 0), e.w - 1);
    int iv = min(max((int)(theta * kINV_PI * (float)e.h), 0), e.h - 1);
    return e.pdf2d[iv * e.w + iu] / (2.0f * kPI * kPI * st);
}
// sampling.cc:94-97: upper_bound - 1 over cdf[0..size)
__device__ __forceinline__ int find_interval(const float* cdf, int size, float v) {
    int lo = 0, hi = size;
    while (lo < hi) {
        int mid = (lo + hi) >> 1;
        if (v < cdf[mid]) hi = mid;
        else lo = mid + 1;
    }
    return min(lo - 1, size - 1);
}
// envmap.cc:60-79
__device__ float3 env_sample(const EnvDev& e, float u0, float u1, float3& dir, float& pdf, float& uu, float& vv) {
    int vo = min(max(find_interval(e.marg_cdf, e.h + 1, u1), 0), e.h - 1);
    float c0 = e.marg_cdf[vo], c1 = e.marg_cdf[vo + 1];
    float du = u1 - c0;
    if (c1 - c0 > 0.0f) du /= (c1 - c0);
    float p1 = e.marg_pdf[vo];
    vv = ((float)vo + du) / (float)e.h;
    const float* rc = e.cond_cdf + (size_t)vo * (e.w + 1);
    int uo = min(max(find_interval(rc, e.w + 1, u0), 0), e.w - 1);
    float d0 = rc[uo], d1 = rc[uo + 1];
    float dv = u0 - d0;
    if (d1 - d0 > 0.0f) dv /= (d1 - d0);
    float p0 = e.cond_pdf[(size_t)vo * e.w + uo];
    uu = ((float)uo + dv) / (float)e.w;
    float mp = p0 * p1;
    if (mp == 0.0f) {
        pdf = 0.0f;
        return mk(0, 0, 0);
    }
    float theta = vv * kPI, phi = uu * (2.0f * kPI);
    float st = sinf(theta), ct = cosf(theta);
    dir = m3(e.l2w, mk(st * cosf(phi), st * sinf(phi), ct));
    pdf = st == 0.0f ? 0.0f : mp / (2.0f * kPI * kPI * st);
    return env_lookup(e, uu, vv);
}

// Shape::pdf(ref, wi) (core/shape.cc:35-48): re-intersect the light triangle
// from the offset origin.
__device__ float area_pdf(const LightDev& L, float3 p, float3 ng, float3 wi) {
    float3 o = offset_origin(p, ng, wi);
    float3 e1 = xyz(L.e1), e2 = xyz(L.e2);
    float3 pv = cross(wi, e2);
    float det = dot(e1, pv);
    if (det > -kEPS && det < kEPS) return 0.0f;
    float inv = 1.0f / det;
    float3 tv = sub(o, xyz(L.p0));
    float u = dot(tv, pv) * inv;
    if (u < 0.0f || u > 1.0f) return 0.0f;
    float3 qv = cross(tv, e1);
    float v = dot(wi, qv) * inv;
    if (v < 0.0f || u + v > 1.0f) return 0.0f;
    float t = dot(e2, qv) * inv;
    if (t <= kEPS || t > kINFTY) return 0.0f;
    float3 hp = add(o, mul(wi, t));
    float3 dd = sub(p, hp);
    float r = dot(dd, dd) / (fabsf(dot(xyz(L.ng), neg(wi))) * L.p0.w);
    return __builtin_isinf(r) ? 0.0f : r;
}

__device__ __forceinline__ float power_heuristic(float f, float g) { return (f * f) / (f * f + g * g); }

// Material lobe and parameters of a surface hit (bsdf.cc:14-21).
// Scenes that reach kernels built with W = false hold no Beckmann material, so the lobe
// word there is the lobe itself.
template <bool W>
__device__ __forceinline__ void bsdf_material(const SceneDev& sc, uint32_t mbits, Bsdf& b) {
    const MatDev M = sc.mats[mbits & 0x7fffffffu];
    const int lb = __float_as_int(M.a.w);
    b.lobe = W ? (lb & 0xff) : lb;
    b.dist = W ? (lb >> 8) : 0;
    b.specW = M.d.w;
    b.type = lobe_type<W>(b.lobe);
    b.R = xyz(M.a);
    b.T = xyz(M.b);
    b.etaB = M.b.w;
    b.eta = xyz(M.c);
    b.alpha = M.c.w;
    b.k = xyz(M.d);
}

// Shading frame: the geometric frame, or the interpolated normal's frame when the
// triangle is smooth and the normals differ (triangle.cc:127-155).
__device__ __forceinline__ void bsdf_frame(float3 ng, bool smooth, float3 ns, Bsdf& b) {
    float3 dpdu, dpdv;
    coordinate_system(ng, dpdu, dpdv);
    b.n = ng;
    b.t = dpdu;
    b.b = dpdv;
    if (smooth && fabsf(dot(ns, ng)) < 1.0f - 1.0e-6f) {
        float3 ss = normalize(cross(ns, ng));
        float3 ts = normalize(cross(ns, ss));
        b.n = normalize(cross(ss, ts));
        b.t = ss;
        b.b = ts;
    }
}

// The hidden uniform of a path's BSDF sample at bounce `depth`: which = 0 for the NEE
// BSDF sample, 1 for the continuation; past the path's 8 dimensions per bounce and the
// learned proposal's block.
__device__ __forceinline__ uint32_t hidden_dim_path(int depth, int which) {
    return 4u + 8u * 128u + 2u * (uint32_t)depth + (uint32_t)which;
}

// Vertices whose continuation the learned proposal samples: Lambert and GGX reflection
// (the guide kernel is built for the four basic lobes; the plastics' sample() does not
// even agree with their f() and pdf(), roughplastic.cc:49-97).
// Vertices whose continuation the proposal may draw: diffuse ones, and microfacet reflections
// (GGX) at least `min_alpha` rough.  A narrow glossy lobe turns with the outgoing direction
// faster than the network's inputs resolve, so the mixture wastes its share of the samples
// there and BSDF sampling alone does better.
__device__ __forceinline__ bool guide_eligible(const Bsdf& b, float min_alpha) {
    return (b.lobe == LOBE_LAMBERT && !black(b.R)) || (b.lobe == LOBE_MICROREFL && b.dist == 0 && b.alpha >= min_alpha);
}

// The direction bsdf_sample draws at a guided vertex (Lambert or GGX reflection) from the same
// uniforms, without its f and pdf: the guide kernel evaluates both once for whichever strategy
// drew the direction.  False where bsdf_sample finds no direction.
__device__ bool guide_bsdf_dir(const Bsdf& b, float3 woW, float u0, float u1, float3& wiW) {
    const float3 wo = normalize(to_local(b, woW));
    if (wo.z == 0.0f) return false;
    float3 wi;
    if (b.lobe == LOBE_LAMBERT) {
        wi = cosine_hemisphere(u0, u1);
        if (wo.z < 0.0f) wi.z *= -1.0f;
    } else {
        const float3 wh = mf_sample_wh<false>(b, wo, u0, u1);
        wi = sub(mul(wh, 2.0f * dot(wh, wo)), wo);
        if (!(wo.z * wi.z > 0.0f)) return false;
    }
    wiW = normalize(to_world(b, wi));
    return true;
}

// ---- shade ---------------------------------------------------------------------------------------
// Shading runs in blocks of its own size, independent of the traversal kernels' block:
// the queue appends are aggregated per block (block_append).
constexpr int kShadeBlock = 256;

// The basic kernel is held to 128 VGPRs (4 waves per SIMD); the wide one holds more lobes.
template <bool WIDE>
constexpr int shade_waves() { return WIDE ? 1 : 4; }

template <int BLOCK, bool GUIDED, bool BATCH, bool WIDE>
__global__ __launch_bounds__(BLOCK) __attribute__((amdgpu_waves_per_eu(shade_waves<WIDE>()))) void k_shade(SceneDev sc_arg, const ConfigDev* __restrict__ cfgp, RenderBufs rb,
                                                 int slot, int oslot) {
    PassState* ps = rb.ps;
    const uint32_t count = ps->countA[slot];
    if (count == 0) return;
    const int kind = cfgp->sampler;
    const uint32_t cfg_seed = cfgp->seed;
    const int max_depth = cfgp->max_depth, rr_depth = cfgp->rr_depth;
    // directlighting.cc:21-57: emission at the first hit only, NEE at every hit, and the
    // recursion only through specular reflection (see the continuation below)
    const bool direct = cfgp->integrator == SPK_INTEGRATOR_DIRECT;
    const float guide_min_alpha = GUIDED ? cfgp->guide_min_alpha : 0.0f;
    const int guide_depth = GUIDED ? cfgp->guide_depth : 0;
    const uint32_t sample_base = ps->sample_base;
    const int cur = slot & 1, nxt = oslot & 1;
    __shared__ uint32_t app_lds[2 * 4 * (BLOCK / 64 + 1)];
    uint32_t parity = 0;
    for (uint32_t base = blockIdx.x * BLOCK; base < count; base += gridDim.x * BLOCK) {
        const uint32_t i = base + threadIdx.x;
        const bool active = i < count;
        bool pushA = false, pushS = false, pushM = false, pushG = false;
        float4 gq0 = make_float4(0, 0, 0, 0), gq1 = gq0, gq3 = gq0;
        float3 nA_o = mk(0, 0, 0), nA_d = mk(0, 0, 1), nA_beta = mk(0, 0, 0);
        uint32_t nA_flags = 0;
        float3 s_o = mk(0, 0, 0), s_d = mk(0, 0, 1), s_c = mk(0, 0, 0);
        float s_tmax = 0.0f;
        float3 m_o = mk(0, 0, 0), m_d = mk(0, 0, 1), m_c = mk(0, 0, 0);
        int m_light = 0;
        uint32_t path = 0, pixel = 0, pfield = 0, img = 0;
        float3 Ladd = mk(0, 0, 0);
        if (active) {
            float4 O = rb.qa_o[cur][i], D = rb.qa_d[cur][i], B = rb.qa_b[cur][i], H = rb.hit[i];
            path = __float_as_uint(O.w);
            const uint32_t flags = __float_as_uint(D.w);
            const bool specular = (flags >> 6) & 1u;
            // the vertex's bounce count travels in the entry (bits 0..5): one queue may hold
            // paths of every depth when spent paths are replaced by new camera paths
            const int depth = (int)(flags & 0x3fu);
            const uint32_t dim_base = direct ? 4u + 10u * (uint32_t)depth : 4u + 8u * (uint32_t)depth;
            // batch entries carry the image in the pixel field and the absolute sample
            pfield = __float_as_uint(B.w);
            img = BATCH ? batch_image(pfield) : 0u;
            pixel = BATCH ? batch_pixel(pfield) : pfield;
            const SceneDev& sc = BATCH ? rb.scenes[img] : sc_arg;
            const uint32_t seed = BATCH ? rb.images[img].seed : cfg_seed;
            const float nl = (float)sc.num_lights;
            const uint32_t sample = BATCH ? (flags >> 7) : sample_base + (flags >> 7);
            float3 ro = xyz(O), rd = xyz(D), beta = xyz(B);
            const int32_t k = __float_as_int(H.w);
            const bool emit = direct ? depth == 0 : (depth == 0 || specular);
            if (k < 0) {
                // directlighting adds the lights' Le(ray) at every miss (directlighting.cc:30-34)
                if ((emit || direct) && sc.env.present) Ladd = cmul(beta, env_Le(sc.env, rd));
            } else {
                float4 NG = sc.tri_ng[k];
                int2 ids = sc.tri_ids[k];
                float3 ng = xyz(NG);
                uint32_t mbits = __float_as_uint(NG.w);
                float3 p = add(ro, mul(rd, H.x));
                if (emit && ids.y >= 0 && dot(ng, neg(rd)) > 0.0f) Ladd = cmul(beta, xyz(sc.lights[ids.y].L));
                if (direct || depth < max_depth) {
                    // BSDF frame (bsdf.cc:14-21 from triangle.cc:127-155)
                    Bsdf b;
                    bsdf_material<WIDE>(sc, mbits, b);
                    float3 ns = ng;
                    if (mbits >> 31) {
                        const float4* tn = sc.tri_n + 3 * k;
                        ns = normalize(add(add(mul(xyz(tn[0]), 1.0f - H.y - H.z), mul(xyz(tn[1]), H.y)),
                                           mul(xyz(tn[2]), H.z)));
                    }
                    bsdf_frame(ng, (mbits >> 31) != 0u, ns, b);
                    const float3 wo = neg(rd);
                    // NEE: uniformSampleOneLight + estimateDirectLight (mis.cc:21-136)
                    // the path skips NEE at BSDFs without a non-specular component (path.cc:77);
                    // directlighting always calls it (directlighting.cc:47-49)
                    if ((direct ? b.lobe != LOBE_NONE : bsdf_has_nonspecular(b)) && sc.num_lights > 0) {
                        const int mask = BX_ALL & ~BX_SPECULAR;
                        float upick = sample_1d(kind, seed, pixel, sample, dim_base + 0);
                        int li = min((int)(upick * nl), sc.num_lights - 1);
                        float2 ul = sample_2d(kind, seed, pixel, sample, dim_base + 1);
                        float2 us = sample_2d(kind, seed, pixel, sample, dim_base + 3);
                        const bool is_env = li >= sc.num_areas;
                        float3 wi = mk(0, 0, 1), Li = mk(0, 0, 0);
                        float lightPdf = 0.0f;
                        float3 lp = mk(0, 0, 0), ln = mk(0, 0, 0);
                        if (!is_env) {
                            const LightDev& L = sc.lights[li];
                            float a0 = ul.x, a1 = ul.y;
                            if (ul.x + ul.y >= 1.0f) {
                                a0 = 1.0f - ul.x;
                                a1 = 1.0f - ul.y;
                            }
                            lp = add(add(xyz(L.p0), mul(xyz(L.e1), a0)), mul(xyz(L.e2), a1));
                            ln = add(add(mul(xyz(L.n0), 1.0f - a0 - a1), mul(xyz(L.n1), a0)), mul(xyz(L.n2), a1));
                            wi = normalize(sub(lp, p));
                            lightPdf = area_pdf(L, p, ng, wi);
                            if (dot(ln, neg(wi)) > 0.0f) Li = xyz(L.L);
                        } else {
                            float uu, vv;
                            Li = env_sample(sc.env, ul.x, ul.y, wi, lightPdf, uu, vv);
                            lp = add(p, mul(wi, 2.0f * sc.env.radius));
                        }
                        if (lightPdf > 0.0f && !black(Li)) {
                            float3 f = mul(bsdf_f<WIDE>(b, wo, wi, mask), fabsf(dot(wi, b.n)));
                            float bsdfPdf = bsdf_pdf<WIDE>(b, wo, wi, mask);
                            if (!black(f)) {
                                // spawnRayTo(Interaction) (interaction.cc:88-93)
                                float3 so = offset_origin(p, ng, sub(lp, p));
                                float3 tg = offset_origin(lp, ln, sub(so, lp));
                                float3 dv = sub(tg, so);
                                float dist = sqrtf(dot(dv, dv));
                                float w = power_heuristic(lightPdf, bsdfPdf);
                                pushS = true;
                                s_o = so;
                                s_d = mk(dv.x / dist, dv.y / dist, dv.z / dist);
                                s_tmax = dist;
                                s_c = mul(cmul(beta, cmul(f, Li)), nl * w / lightPdf);
                            }
                        }
                        float bp = 0.0f;
                        int st = 0;
                        float3 wi2 = mk(0, 0, 1);
                        const float uh2 = lobe_has_hidden<WIDE>(b) ? sample_1d(kind, seed, pixel, sample,
                                                                         direct ? dim_base + 9 : hidden_dim_path(depth, 0))
                                                             : 0.5f;
                        float3 f2 = bsdf_sample<WIDE>(b, wo, wi2, us.x, us.y, bp, mask, st, uh2);
                        f2 = mul(f2, fabsf(dot(wi2, b.n)));
                        if (!black(f2) && bp > 0.0f) {
                            float lp2 = is_env ? env_pdf(sc.env, wi2) : area_pdf(sc.lights[li], p, ng, wi2);
                            if (lp2 != 0.0f) {
                                float w = power_heuristic(bp, lp2);
                                pushM = true;
                                m_o = offset_origin(p, ng, wi2);
                                m_d = wi2;
                                m_light = li;
                                m_c = mul(cmul(beta, f2), nl * w / bp);
                            }
                        }
                    }
                    // Continuation (path.cc:83-94) and Russian roulette (path.cc:116-121); with a
                    // guide the vertices it may draw for (guide_eligible, within guide_depth
                    // bounces) go to the guide queue instead.
                    if (direct) {
                        // specularReflect (integrator.cc:112-130) while depth + 1 < maxDepth, where
                        // spica's depth grows by 2 per specular bounce (Li -> specularReflect(depth + 1)
                        // -> Li(depth + 2)); weighted by |wi . n_geometric| / pdf, no roulette.
                        // Only the mirror's type (R|S) is a subset of R|S; specularTransmit never finds a
                        // component: no lobe's type is a subset of T|S (FresnelSpecular is R|T|S,
                        // MicrofacetTransmission R|T|G).
                        if (2 * depth + 1 < max_depth) {
                            float2 uc = sample_2d(kind, seed, pixel, sample, dim_base + 5);
                            float pdf = 0.0f;
                            int st = 0;
                            float3 wi = mk(0, 0, 1);
                            float3 f = bsdf_sample<WIDE>(b, wo, wi, uc.x, uc.y, pdf, BX_REFLECTION | BX_SPECULAR, st, 0.5f);
                            const float cg = fabsf(dot(wi, ng));
                            if (pdf > 0.0f && !black(f) && cg != 0.0f) {
                                pushA = true;
                                nA_o = offset_origin(p, ng, wi);
                                nA_d = wi;
                                nA_beta = cmul(beta, mul(f, cg / pdf));
                                nA_flags = (flags & ~0x7fu) | (uint32_t)(depth + 1) | (1u << 6);
                            }
                        }
                    } else if (GUIDED && depth < guide_depth && guide_eligible(b, guide_min_alpha)) {
                        pushG = true;
                        gq0 = make_float4(p.x, p.y, p.z, __uint_as_float(path));
                        gq1 = make_float4(wo.x, wo.y, wo.z, __uint_as_float(flags));
                        nA_beta = beta;
                        gq3 = make_float4(ns.x, ns.y, ns.z, __int_as_float(k));
                    } else {
                    float2 uc = sample_2d(kind, seed, pixel, sample, dim_base + 5);
                    float pdf = 0.0f;
                    int st = 0;
                    float3 wi = mk(0, 0, 1);
                    const float uh = lobe_has_hidden<WIDE>(b) ? sample_1d(kind, seed, pixel, sample, hidden_dim_path(depth, 1)) : 0.5f;
                    float3 f = bsdf_sample<WIDE>(b, wo, wi, uc.x, uc.y, pdf, BX_ALL, st, uh);
                    if (!black(f) && pdf != 0.0f) {
                        float3 nb = cmul(beta, mul(f, fabsf(dot(wi, b.n)) / pdf));
                        bool cont = true;
                        if (depth > rr_depth) {
                            float q = fminf(0.95f, gray(nb));
                            if (sample_1d(kind, seed, pixel, sample, dim_base + 7) > q) cont = false;
                            else nb = mul(nb, 1.0f / q);
                        }
                        if (cont) {
                            pushA = true;
                            nA_o = offset_origin(p, ng, wi);
                            nA_d = wi;
                            nA_beta = nb;
                            nA_flags = (flags & ~0x7fu) | (uint32_t)(depth + 1) | ((st & BX_SPECULAR) ? (1u << 6) : 0u);
                        }
                    }
                    }
                }
            }
            if (!black(Ladd)) {
                float4 Lv = rb.path_L[path];
                rb.path_L[path] = make_float4(Lv.x + Ladd.x, Lv.y + Ladd.y, Lv.z + Ladd.z, Lv.w);
            }
        }
        constexpr int NQ = GUIDED ? 4 : 3;
        bool pr[NQ];
        uint32_t* ctr[NQ];
        uint32_t sl[NQ];
        pr[0] = pushA;
        ctr[0] = &ps->countA[oslot];
        pr[1] = pushS;
        ctr[1] = &ps->countS[slot];
        pr[2] = pushM;
        ctr[2] = &ps->countM[slot];
        if constexpr (GUIDED) {
            pr[NQ - 1] = pushG;
            ctr[NQ - 1] = &ps->countG[slot];
        }
        block_append<BLOCK / 64, NQ>(pr, ctr, sl, app_lds, parity);
        parity ^= 1u;
        const uint32_t sa = sl[0], ss = sl[1], sm = sl[2];
        if constexpr (GUIDED) {
            const uint32_t sg = sl[NQ - 1];
            if (pushG) {
                rb.qg[0][sg] = gq0;
                rb.qg[1][sg] = gq1;
                rb.qg[2][sg] = make_float4(nA_beta.x, nA_beta.y, nA_beta.z, __uint_as_float(pfield));
                rb.qg[3][sg] = gq3;
            }
        }
        if (pushA) {
            rb.qa_o[nxt][sa] = make_float4(nA_o.x, nA_o.y, nA_o.z, __uint_as_float(path));
            rb.qa_d[nxt][sa] = make_float4(nA_d.x, nA_d.y, nA_d.z, __uint_as_float(nA_flags));
            rb.qa_b[nxt][sa] = make_float4(nA_beta.x, nA_beta.y, nA_beta.z, __uint_as_float(pfield));
        }
        if (pushS) {
            rb.qs_o[ss] = make_float4(s_o.x, s_o.y, s_o.z, __uint_as_float(path));
            rb.qs_d[ss] = make_float4(s_d.x, s_d.y, s_d.z, s_tmax);
            rb.qs_c[ss] = make_float4(s_c.x, s_c.y, s_c.z, __uint_as_float(img));
        }
        if (pushM) {
            rb.qm_o[sm] = make_float4(m_o.x, m_o.y, m_o.z, __uint_as_float(path));
            rb.qm_d[sm] = make_float4(m_d.x, m_d.y, m_d.z, __int_as_float(m_light));
            rb.qm_c[sm] = make_float4(m_c.x, m_c.y, m_c.z, __uint_as_float(img));
        }
    }
}

// ---- shadow: NEE light-sample occlusion (visibility_tester.cc:21-24) ------------------------
template <int BLOCK, int S, bool BATCH, bool COUNT>
__global__ __launch_bounds__(BLOCK) void k_shadow(SceneDev sc, RenderBufs rb, int slot) {
    __shared__ uint32_t lds[S * BLOCK];
    PassState* ps = rb.ps;
    const uint32_t count = ps->countS[slot];
    if (count == 0) return;
    uint32_t* col = lds + threadIdx.x;
    uint32_t* spill = rb.spill + (size_t)(blockIdx.x * BLOCK + threadIdx.x) * rb.spill_cap;
    Counters cnt = {0, 0};
    auto load = [&](uint32_t i, float3& o, float3& d, float& tmax) -> const SceneDev* {
        const float4 O = rb.qs_o[i], D = rb.qs_d[i];
        o = xyz(O);
        d = xyz(D);
        tmax = D.w;
        return BATCH ? rb.scenes + __float_as_uint(rb.qs_c[i].w) : nullptr;
    };
    uint32_t rays = 0;
    auto done = [&](uint32_t i, int32_t k, float, float, float, const SceneDev&) {
        if (k < 0) {
            const uint32_t path = __float_as_uint(rb.qs_o[i].w);
            const float4 c = rb.qs_c[i];
            float4 Lv = rb.path_L[path];
            rb.path_L[path] = make_float4(Lv.x + c.x, Lv.y + c.y, Lv.z + c.z, Lv.w);
        }
        if (COUNT) rays++;
    };
    traverse_stream<BLOCK, S, true, COUNT, BATCH>(sc, rb.scenes, count, load, done, col, spill, cnt);
    if (COUNT) {
        atomicAdd(&ps->count_nodes_s, (unsigned long long)cnt.nodes);
        atomicAdd(&ps->count_tris_s, (unsigned long long)cnt.tris);
        atomicAdd(&ps->count_rays_s, (unsigned long long)rays);
    }
}

// ---- MIS: NEE BSDF-sample ray; counts if it reaches the chosen light (mis.cc:114-133) ----------
template <int BLOCK, int S, bool BATCH>
__global__ __launch_bounds__(BLOCK) __attribute__((amdgpu_waves_per_eu(BATCH ? 7 : 8))) void k_mis(SceneDev sc, RenderBufs rb, int slot) {
    __shared__ uint32_t lds[S * BLOCK];
    PassState* ps = rb.ps;
    const uint32_t count = ps->countM[slot];
    if (count == 0) return;
    uint32_t* col = lds + threadIdx.x;
    uint32_t* spill = rb.spill + (size_t)(blockIdx.x * BLOCK + threadIdx.x) * rb.spill_cap;
    Counters cnt = {0, 0};
    auto load = [&](uint32_t i, float3& o, float3& d, float& tmax) -> const SceneDev* {
        o = xyz(rb.qm_o[i]);
        d = xyz(rb.qm_d[i]);
        tmax = kINFTY;
        return BATCH ? rb.scenes + __float_as_uint(rb.qm_c[i].w) : nullptr;
    };
    auto done = [&](uint32_t i, int32_t k, float, float, float, const SceneDev& s) {
        const float4 D = rb.qm_d[i];
        const int li = __float_as_int(D.w);
        float3 Le = mk(0, 0, 0);
        if (k >= 0) {
            if (li < s.num_areas && s.tri_ids[k].y == li && dot(xyz(s.tri_ng[k]), neg(xyz(D))) > 0.0f)
                Le = xyz(s.lights[li].L);
        } else if (li >= s.num_areas) {
            Le = env_Le(s.env, xyz(D));
        }
        if (!black(Le)) {
            const uint32_t path = __float_as_uint(rb.qm_o[i].w);
            const float4 c = rb.qm_c[i];
            float4 Lv = rb.path_L[path];
            rb.path_L[path] = make_float4(Lv.x + c.x * Le.x, Lv.y + c.y * Le.y, Lv.z + c.z * Le.z, Lv.w);
        }
    };
    traverse_stream<BLOCK, S, false, false, BATCH>(sc, rb.scenes, count, load, done, col, spill, cnt);
}

// ---- guide: learned continuation proposal ---------------------------------------------------------
// One persistent block of 512 threads per CU (the bf16 weights fill 58 KB of LDS once per
// block).  Each wave takes 64 queued vertices: their inputs go through the MFMA network
// together, then every lane samples its vertex's continuation from the one-sample MIS
// mixture frac * guide + (1 - frac) * BSDF (path.cc:83-94 with the proposal added), applies
// Russian roulette (path.cc:116-121) and, in training passes, records (inputs, w_i,
// radiance so far, throughput) for the network.
// One 512-thread block per CU: the weights (58 KB) and, per wave, the network outputs staged in
// fp32 (64 x 41 floats) take 144 KB of LDS.  The mixture reads its parameters lobe by lobe from
// the staged outputs.  (Staging in fp16 with 768-thread blocks, 3 waves per SIMD at 168 VGPRs,
// measured the same: 7.40 against 7.32 G vertices/s after training on C3; the kernel is bound
// by VALU issue, not latency.)
constexpr int kGuideBlock = 512;
constexpr int kGuideObStride = 41;  // floats per lane: 40 outputs, padded to an odd count

__device__ __forceinline__ uint32_t guide_dim(int depth) { return 4u + 8u * (uint32_t)(kMaxDepthCap + depth) + 1u; }

__device__ __forceinline__ void guide_inputs(const GuideNetDev& net, float3 p, float3 n, float3 wo, float (&x)[7]) {
    x[0] = clampf((p.x - net.lo[0]) * net.inv_ext[0], 0.0f, 1.0f);
    x[1] = clampf((p.y - net.lo[1]) * net.inv_ext[1], 0.0f, 1.0f);
    x[2] = clampf((p.z - net.lo[2]) * net.inv_ext[2], 0.0f, 1.0f);
    float2 fn = dir_features(n.x, n.y, n.z), fw = dir_features(wo.x, wo.y, wo.z);
    x[3] = fn.x;
    x[4] = fn.y;
    x[5] = fw.x;
    x[6] = fw.y;
}

// Batches (BATCH): every image renders one scene (the guide's), queue entries carry the image in
// the pixel field and the absolute sample index, and each image has its own sampler seed.
template <bool BATCH>
__global__ __launch_bounds__(kGuideBlock) void k_guide(SceneDev sc, const ConfigDev* __restrict__ cfgp, RenderBufs rb,
                                                      GuideNetDev net, int slot, int oslot) {
    __shared__ __attribute__((aligned(16))) uint16_t w_lds[kGFwdSize];
    __shared__ float b_lds[kGBiasSize];
    __shared__ __attribute__((aligned(16))) float obuf[kGuideBlock / 64][64 * kGuideObStride];
    PassState* ps = rb.ps;
    const uint32_t count = ps->countG[slot];
    if (count == 0) return;
    for (int i = threadIdx.x; i < kGFwdSize / 8; i += kGuideBlock)
        reinterpret_cast<u32x4_t*>(w_lds)[i] = reinterpret_cast<const u32x4_t*>(net.wf)[i];
    for (int i = threadIdx.x; i < kGBiasSize; i += kGuideBlock) b_lds[i] = net.bias[i];
    __syncthreads();
    const int kind = cfgp->sampler;
    const uint32_t cfg_seed = cfgp->seed;
    const int rr_depth = cfgp->rr_depth;
    const uint32_t sample_base = ps->sample_base;
    const int nxt = oslot & 1;
    const int lane = __lane_id(), wv = threadIdx.x >> 6, r = lane & 15, g = lane >> 4;
    float* ob = obuf[wv];  // the inputs (8 floats per lane), then the network's outputs
    __shared__ uint32_t app_lds[2 * 2 * (kGuideBlock / 64 + 1)];
    uint32_t parity = 0;
    // block-uniform trip count: the appends below synchronise the block
    for (uint32_t bb = blockIdx.x * kGuideBlock; bb < count; bb += gridDim.x * kGuideBlock) {
        const uint32_t base = bb + (uint32_t)wv * 64u;
        const uint32_t i = base + lane;
        const bool active = i < count;
        float4 G0 = make_float4(0, 0, 0, 0), G1 = make_float4(0, 0, 1, 0), G3 = make_float4(0, 0, 1, 0);
        float4 G2 = make_float4(0, 0, 0, 0), NG = make_float4(0, 0, 1, 0);
        if (active) {
            G0 = rb.qg[0][i];
            G1 = rb.qg[1][i];
            G3 = rb.qg[3][i];
            // issued before the network so that their latency overlaps the MFMA chain
            G2 = rb.qg[2][i];
            NG = sc.tri_ng[__float_as_int(G3.w)];
        }
        float x[7];
        guide_inputs(net, xyz(G0), xyz(G3), xyz(G1), x);
#pragma unroll
        for (int d = 0; d < 7; d++) ob[lane * 8 + d] = x[d];
        wave_lds_sync();
        f32x4_t out[3][4];
        NoHook hook;
        mlp_forward<4>(opaque(w_lds), opaque(b_lds), [&](int ct, int d) { return ob[(16 * ct + r) * 8 + d]; }, out,
                       hook);
        wave_lds_sync();
#pragma unroll
        for (int mt = 0; mt < 3; mt++)
#pragma unroll
            for (int ct = 0; ct < 4; ct++)
#pragma unroll
                for (int k = 0; k < 4; k++) {
                    const int u = 16 * mt + 4 * g + k;
                    if (u < kGOut)
                        ob[(16 * ct + r) * kGuideObStride + u] = out[mt][ct][k];
                }
        wave_lds_sync();
        bool pushA = false, pushR = false;
        float3 nA_o = mk(0, 0, 0), nA_d = mk(0, 0, 1), nA_beta = mk(0, 0, 0), beta = mk(0, 0, 0);
        uint32_t nA_flags = 0, path = 0, pfield = 0, rsample = 0;
        int depth = 0;
        if (active) {
            const float* orow = ob + lane * kGuideObStride;
            const auto o = [orow](int u) { return orow[u]; };
            const float ls = mix_log_norm(o);
            path = __float_as_uint(G0.w);
            pfield = __float_as_uint(G2.w);
            const uint32_t pixel = BATCH ? batch_pixel(pfield) : pfield;
            const uint32_t seed = BATCH ? rb.images[batch_image(pfield)].seed : cfg_seed;
            beta = xyz(G2);
            const uint32_t flags = __float_as_uint(G1.w);
            depth = (int)(flags & 0x3fu);
            const uint32_t dim_base = 4u + 8u * (uint32_t)depth;
            const uint32_t sample = BATCH ? (flags >> 7) : sample_base + (flags >> 7);
            rsample = sample;
            const uint32_t mbits = __float_as_uint(NG.w);
            const float3 ng = xyz(NG), p = xyz(G0), wo = xyz(G1);
            Bsdf b;
            bsdf_material<false>(sc, mbits, b);
            bsdf_frame(ng, (mbits >> 31) != 0u, xyz(G3), b);
            const float2 uc = sample_2d(kind, seed, pixel, sample, dim_base + 5);
            const float us = sample_1d(kind, seed, pixel, sample, guide_dim(depth));
            float3 wi = mk(0, 0, 1);
            bool ok = true;
            if (us < net.frac) wi = mix_sample_o(o, ls, us / net.frac, uc.x, uc.y);
            else ok = guide_bsdf_dir(b, wo, uc.x, uc.y, wi);
            if (ok) {
                const float3 f = bsdf_f<false>(b, wo, wi, BX_ALL);
                const float pb = bsdf_pdf<false>(b, wo, wi, BX_ALL);
                const float pdf = (1.0f - net.frac) * pb + net.frac * __expf(mix_log_pdf_o(o, ls, wi.x, wi.y, wi.z));
                if (!black(f) && pdf > 0.0f) {
                    const float cw = fabsf(dot(wi, b.n));
                    float3 nb = cmul(beta, mul(f, cw / pdf));
                    bool cont = true;
                    if (depth > rr_depth) {
                        // Survival from the throughput BSDF sampling would give this direction,
                        // beta f cos / p_bsdf: the mixture's pdf is highest where the network expects
                        // the most light, so beta f cos / p_mix is lowest there, and a roulette on it
                        // would end the very paths the guide sends towards the light.  Any survival
                        // probability in (0, 1] with the 1/q reweighting keeps the estimate unbiased.
                        float q = fminf(0.95f, gray(pb > 0.0f ? cmul(beta, mul(f, cw / pb)) : nb));
                        if (sample_1d(kind, seed, pixel, sample, dim_base + 7) > q) cont = false;
                        else nb = mul(nb, 1.0f / q);
                    }
                    if (cont) {
                        pushA = true;
                        if (net.record) {
                            // keyed by (pixel field, depth, sample), not the pass-local path slot, so
                            // the kept records do not depend on how the samples are cut into passes
                            uint32_t h0 = pfield, h1 = (uint32_t)depth, h2 = sample, h3 = 0x7f4a7c15u;
                            pcg4d(h0, h1, h2, h3);
                            pushR = u32_to_unit(h0) < net.rec_keep;
                        }
                        nA_o = offset_origin(p, ng, wi);
                        nA_d = wi;
                        nA_beta = nb;
                        nA_flags = (flags & ~0x7fu) | (uint32_t)(depth + 1);
                    }
                }
            }
        }
        const bool pr[2] = {pushA, pushR};
        uint32_t* const ctr[2] = {&ps->countA[oslot], net.rec_count};
        uint32_t sl[2];
        block_append<kGuideBlock / 64, 2>(pr, ctr, sl, app_lds, parity);
        parity ^= 1u;
        const uint32_t sa = sl[0];
        if (pushA) {
            rb.qa_o[nxt][sa] = make_float4(nA_o.x, nA_o.y, nA_o.z, __uint_as_float(path));
            rb.qa_d[nxt][sa] = make_float4(nA_d.x, nA_d.y, nA_d.z, __uint_as_float(nA_flags));
            rb.qa_b[nxt][sa] = make_float4(nA_beta.x, nA_beta.y, nA_beta.z, __uint_as_float(pfield));
        }
        if (net.record) {
            const uint32_t sr = sl[1];
            if (pushR && sr < net.rec_cap) {
                const float4 L = rb.path_L[path];
                net.rec[0][sr] = make_float4(x[0], x[1], x[2], x[3]);
                net.rec[1][sr] = make_float4(x[4], x[5], x[6], __uint_as_float(path));
                net.rec[2][sr] = make_float4(nA_d.x, nA_d.y, nA_d.z, __uint_as_float((uint32_t)depth));
                net.rec[3][sr] = make_float4(L.x, L.y, L.z, __uint_as_float(rsample));
                net.rec[4][sr] = make_float4(beta.x, beta.y, beta.z, __uint_as_float(pfield));
            }
        }
        wave_lds_sync();
    }
}

// Training records of the pass that just ended: rec[3] held the path's radiance when the
// vertex was recorded; once the path has ended it becomes the radiance gathered after the
// vertex, L_end - L_before.  This runs before the next pass's camera kernel reuses the path
// slots, so a training round that spans several passes trains on its own paths' radiance.
__global__ __launch_bounds__(256) void k_guide_finalize(GuideNetDev net, const float4* __restrict__ path_L) {
    const uint32_t lo = *net.rec_fin, hi = min(*net.rec_count, net.rec_cap);
    for (uint32_t i = lo + blockIdx.x * blockDim.x + threadIdx.x; i < hi; i += gridDim.x * blockDim.x) {
        const float4 r3 = net.rec[3][i];
        const float4 L = path_L[__float_as_uint(net.rec[1][i].w)];
        net.rec[3][i] = make_float4(L.x - r3.x, L.y - r3.y, L.z - r3.z, r3.w);
    }
}
__global__ void k_guide_finalize_mark(GuideNetDev net) {
    if (threadIdx.x == 0 && blockIdx.x == 0) *net.rec_fin = min(*net.rec_count, net.rec_cap);
}

// ---- film: Kahan-compensated accumulation (film.cc:65-74) -------------------------------------------
__device__ __forceinline__ void kahan(float& s, float& c, float x) {
    float y = x - c;
    float t = s + y;
    c = (t - s) - y;
    s = t;
}
template <bool BATCH>
__global__ __launch_bounds__(256) void k_film(const ConfigDev* __restrict__ cfg, RenderBufs rb, int max_depth,
                                             int last_slot) {
    PassState* ps = rb.ps;
    const uint32_t n_items = BATCH ? ps->n_pixels : ps->n_pix;
    const uint32_t cb = ps->chunk_begin, ce = ps->chunk_end;
    for (uint32_t t = blockIdx.x * blockDim.x + threadIdx.x; t < n_items; t += gridDim.x * blockDim.x) {
        // one film pixel: its k samples of this pass in sample order
        uint32_t j, k, n_pix, pixel_base, path_base, film_off = 0;
        int W, H;
        if (BATCH) {
            const ChunkDev ch = rb.chunks[find_chunk<true>(rb.chunks, cb, ce, t)];
            const ImageDev& im = rb.images[ch.img];
            j = t - ch.pix_off;
            k = ch.k;
            n_pix = ch.n_pix;
            pixel_base = ch.pixel_base;
            path_base = ch.path_off;
            film_off = im.film_off;
            W = im.cam.W;
            H = im.cam.H;
        } else {
            j = t;
            k = ps->k_samples;
            n_pix = ps->n_pix;
            pixel_base = ps->pixel_base;
            path_base = 0;
            W = cfg->W;
            H = cfg->H;
        }
        int x, y;
        swizzle(pixel_base + j, W, H, x, y);
        uint32_t f = film_off + (uint32_t)(y * W + (W - 1 - x));
        float4 s = rb.film_sum[f], c = rb.film_c[f], q = rb.film_sq[f], qc = rb.film_sqc[f];
        for (uint32_t si = 0; si < k; si++) {
            float4 L = rb.path_L[path_base + si * n_pix + j];
            kahan(s.x, c.x, L.x);
            kahan(s.y, c.y, L.y);
            kahan(s.z, c.z, L.z);
            kahan(q.x, qc.x, L.x * L.x);
            kahan(q.y, qc.y, L.y * L.y);
            kahan(q.z, qc.z, L.z * L.z);
        }
        s.w += (float)k;
        rb.film_sum[f] = s;
        rb.film_c[f] = c;
        rb.film_sq[f] = q;
        rb.film_sqc[f] = qc;
    }
    if (blockIdx.x == 0 && threadIdx.x == 0 && !ps->done && cfg->regen) {
        // the last iteration's queue counts (k_plan added the earlier ones and the paths)
        ps->total_extend += ps->rg_nin;
        ps->total_shadow += ps->countS[last_slot];
        ps->total_mis += ps->countM[last_slot];
        ps->total_guided += ps->countG[last_slot];
        ps->total_samples += ps->k_samples;
    } else if (blockIdx.x == 0 && threadIdx.x == 0 && !ps->done) {
        unsigned long long e = 0, sh = 0, mi = 0, gu = 0;
        for (int d = 0; d <= max_depth; d++) {
            e += ps->countA[d];
            sh += ps->countS[d];
            mi += ps->countM[d];
            gu += ps->countG[d];
        }
        ps->total_guided += gu;
        ps->total_paths += ps->n;
        ps->total_extend += e;
        ps->total_shadow += sh;
        ps->total_mis += mi;
        ps->total_samples += ps->k_samples;
    }
}

__global__ void k_clear_film(RenderBufs rb, uint32_t npix) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < npix; i += gridDim.x * blockDim.x) {
        float4 z = make_float4(0, 0, 0, 0);
        rb.film_sum[i] = z;
        rb.film_c[i] = z;
        rb.film_sq[i] = z;
        rb.film_sqc[i] = z;
    }
}

// relMSE in float64: mean over pixels and channels of (I - R)^2 / (R^2 + eps).
__global__ __launch_bounds__(256) void k_relmse(RenderBufs rb, uint32_t npix, const double* ref, double eps, double* partial) {
    __shared__ double red[256];
    double acc = 0.0;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < npix; i += gridDim.x * blockDim.x) {
        float4 s = rb.film_sum[i];
        double inv = 1.0 / ((double)s.w + 1.0e-12);
        double I[3] = {(double)s.x * inv, (double)s.y * inv, (double)s.z * inv};
        for (int c = 0; c < 3; c++) {
            double R = ref[3 * (size_t)i + c];
            double dd = I[c] - R;
            acc += dd * dd / (R * R + eps);
        }
    }
    red[threadIdx.x] = acc;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) partial[blockIdx.x] = red[0];
}
__global__ __launch_bounds__(256) void k_relmse_pixels(RenderBufs rb, int64_t n, const int32_t* __restrict__ pixels,
                                                       const double* ref, double eps, double* partial) {
    __shared__ double red[256];
    double acc = 0.0;
    for (int64_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        float4 s = rb.film_sum[pixels[i]];
        double inv = 1.0 / ((double)s.w + 1.0e-12);
        double I[3] = {(double)s.x * inv, (double)s.y * inv, (double)s.z * inv};
        for (int c = 0; c < 3; c++) {
            double R = ref[3 * i + c];
            double dd = I[c] - R;
            acc += dd * dd / (R * R + eps);
        }
    }
    red[threadIdx.x] = acc;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) partial[blockIdx.x] = red[0];
}
__global__ __launch_bounds__(256) void k_reduce_sum(const double* partial, int n, double* out) {
    __shared__ double red[256];
    double acc = 0.0;
    for (int i = threadIdx.x; i < n; i += 256) acc += partial[i];
    red[threadIdx.x] = acc;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[0] = red[0];
}

// LDS stack slots per lane.  16 (4 KiB per 64-lane block) lets the traversal kernels run
// 7-8 waves per SIMD; a 4-wide tree rarely needs more (entries beyond spill to memory).
// 32 slots held them at 5 waves: C1 2963 -> 3569 Mrays/s going from 32 to 16.
constexpr int kStackSlots = 16;

__global__ void k_sampler(int kind, uint32_t seed, int64_t n, const uint32_t* px, const uint32_t* sm, const uint32_t* dm,
                          uint32_t* out) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        out[i] = sample_u32(kind, seed, px[i], sm[i], dm[i]);
}

__global__ void k_zero_u32(uint32_t* p, int n) {
    for (int i = threadIdx.x; i < n; i += blockDim.x) p[i] = 0u;
}

}  // namespace

int zero_u32(uint32_t* p, int n, hipStream_t st) {
    hipLaunchKernelGGL(k_zero_u32, dim3(1), dim3(64), 0, st, p, n);
    SPK_CHECK_HIP(hipGetLastError());
    return SPK_OK;
}

int launch_sampler(int kind, uint32_t seed, int64_t n, const uint32_t* px, const uint32_t* sm, const uint32_t* dm,
                   uint32_t* out, hipStream_t st) {
    hipLaunchKernelGGL(k_sampler, dim3(256), dim3(256), 0, st, kind, seed, n, px, sm, dm, out);
    SPK_CHECK_HIP(hipGetLastError());
    return SPK_OK;
}

int trace_blocks_per_cu(int block) {
    int nb = 0;
    hipError_t e;
    switch (block) {
    case 64: e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_extend<64, kStackSlots, false, false>, 64, 0); break;
    case 256: e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_extend<256, kStackSlots, false, false>, 256, 0); break;
    default: e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_extend<128, kStackSlots, false, false>, 128, 0); break;
    }
    if (e != hipSuccess || nb < 1) return 2;
    return nb;
}

// ---- launchers ---------------------------------------------------------------------------------------
int launch_trace(const SceneDev& s, const TraceBufs& b, uint32_t* work_counter, uint32_t* spill, int spill_cap, int grid,
                 hipStream_t st) {
    constexpr int B = 128;
    if (b.any_hit)
        hipLaunchKernelGGL((k_trace<B, kStackSlots, true>), dim3(grid), dim3(B), 0, st, s, b, work_counter, spill, spill_cap);
    else
        hipLaunchKernelGGL((k_trace<B, kStackSlots, false>), dim3(grid), dim3(B), 0, st, s, b, work_counter, spill, spill_cap);
    SPK_CHECK_HIP(hipGetLastError());
    return SPK_OK;
}

int launch_set_pass(PassState* ps, uint32_t index, hipStream_t st) {
    hipLaunchKernelGGL(k_set_pass, dim3(1), dim3(64), 0, st, ps, index);
    SPK_CHECK_HIP(hipGetLastError());
    return SPK_OK;
}

int launch_clear_film(const RenderBufs& b, uint32_t npix, hipStream_t st) {
    hipLaunchKernelGGL(k_clear_film, dim3(1024), dim3(256), 0, st, b, npix);
    SPK_CHECK_HIP(hipGetLastError());
    return SPK_OK;
}

int launch_relmse(const RenderBufs& b, const ConfigDev& cfg, const double* ref, double eps, double* out_partial, int nblocks,
                  hipStream_t st) {
    hipLaunchKernelGGL(k_relmse, dim3(nblocks), dim3(256), 0, st, b, cfg.npix, ref, eps, out_partial);
    hipLaunchKernelGGL(k_reduce_sum, dim3(1), dim3(256), 0, st, out_partial, nblocks, out_partial + nblocks);
    SPK_CHECK_HIP(hipGetLastError());
    return SPK_OK;
}

int launch_relmse_pixels(const RenderBufs& b, int64_t n, const int32_t* pixels, const double* ref, double eps,
                         double* out_partial, int nblocks, hipStream_t st) {
    hipLaunchKernelGGL(k_relmse_pixels, dim3(nblocks), dim3(256), 0, st, b, n, pixels, ref, eps, out_partial);
    hipLaunchKernelGGL(k_reduce_sum, dim3(1), dim3(256), 0, st, out_partial, nblocks, out_partial + nblocks);
    SPK_CHECK_HIP(hipGetLastError());
    return SPK_OK;
}

// The stage kernels of one bounce (classic: depth d, queue slot d, output slot d + 1) or of one
// iteration of a regenerating pass (queue slot it & 1, output slot the other one, preceded by
// k_plan and the camera kernel that refills the queue).  Profile events per step (e0 = the
// step's first event; nd = maxDepth + 1 for a classic pass, 1 for an iteration): closest hit
// e0 + 2d, + 1; guide e0 + 2nd + 2d, + 1; shade e0 + 4nd + 2d, + 1; shadow e0 + 6nd + 2d, + 1.
template <int B, bool BATCH>
static int launch_stages(const SceneDev& s, const RenderBufs& b, const ConfigDev& cfg, const LaunchShape& ls,
                         hipStream_t st, int slot, int oslot, int d, int nd, bool last, hipEvent_t* ev, int n_ev,
                         bool count, const GuideNetDev* guide) {
    auto mark = [&](int e) -> int {
        if (ev && e < n_ev) SPK_CHECK_HIP(hipEventRecord(ev[e], st));
        return SPK_OK;
    };
    if (mark(2 * d)) return SPK_ERR_HIP;
    if (count)
        hipLaunchKernelGGL((k_extend<B, kStackSlots, true, BATCH>), dim3(ls.grid_trace), dim3(B), 0, st, s, b, slot);
    else
        hipLaunchKernelGGL((k_extend<B, kStackSlots, false, BATCH>), dim3(ls.grid_trace), dim3(B), 0, st, s, b, slot);
    if (mark(2 * d + 1)) return SPK_ERR_HIP;
    constexpr int SB = kShadeBlock;
    const dim3 gs(ls.grid_stream * 2 * 64 / SB), bs(SB);
    if (mark(4 * nd + 2 * d)) return SPK_ERR_HIP;
    if (guide) {
        if (cfg.wide) hipLaunchKernelGGL((k_shade<SB, true, BATCH, true>), gs, bs, 0, st, s, b.cfg, b, slot, oslot);
        else hipLaunchKernelGGL((k_shade<SB, true, BATCH, false>), gs, bs, 0, st, s, b.cfg, b, slot, oslot);
    } else {
        if (cfg.wide) hipLaunchKernelGGL((k_shade<SB, false, BATCH, true>), gs, bs, 0, st, s, b.cfg, b, slot, oslot);
        else hipLaunchKernelGGL((k_shade<SB, false, BATCH, false>), gs, bs, 0, st, s, b.cfg, b, slot, oslot);
    }
    if (mark(4 * nd + 2 * d + 1)) return SPK_ERR_HIP;
    if (!last) {
        if (mark(6 * nd + 2 * d)) return SPK_ERR_HIP;
        if (count)
            hipLaunchKernelGGL((k_shadow<B, kStackSlots, BATCH, true>), dim3(ls.grid_trace), dim3(B), 0, st, s, b, slot);
        else
            hipLaunchKernelGGL((k_shadow<B, kStackSlots, BATCH, false>), dim3(ls.grid_trace), dim3(B), 0, st, s, b, slot);
        if (mark(6 * nd + 2 * d + 1)) return SPK_ERR_HIP;
        hipLaunchKernelGGL((k_mis<B, kStackSlots, BATCH>), dim3(ls.grid_trace), dim3(B), 0, st, s, b, slot);
        // after the NEE kernels, so a training record sees this vertex's direct light
        if (guide) {
            if (mark(2 * nd + 2 * d)) return SPK_ERR_HIP;
            hipLaunchKernelGGL(k_guide<BATCH>, dim3(ls.grid_guide), dim3(kGuideBlock), 0, st, s, b.cfg, b, *guide, slot,
                               oslot);
            if (mark(2 * nd + 2 * d + 1)) return SPK_ERR_HIP;
        }
    }
    SPK_CHECK_HIP(hipGetLastError());
    return SPK_OK;
}

template <bool BATCH>
static int launch_film(const RenderBufs& b, const ConfigDev& cfg, const LaunchShape& ls, hipStream_t st, int last_slot,
                       const GuideNetDev* guide) {
    hipLaunchKernelGGL(k_film<BATCH>, dim3(ls.grid_stream), dim3(256), 0, st, b.cfg, b, cfg.max_depth, last_slot);
    if (guide && guide->record) {
        hipLaunchKernelGGL(k_guide_finalize, dim3(ls.grid_stream), dim3(256), 0, st, *guide, (const float4*)b.path_L);
        hipLaunchKernelGGL(k_guide_finalize_mark, dim3(1), dim3(64), 0, st, *guide);
    }
    SPK_CHECK_HIP(hipGetLastError());
    return SPK_OK;
}

// A classic pass: all of its camera paths at once, then maxDepth + 1 bounces over shrinking queues.
template <int B, bool BATCH>
static int launch_pass_b(const SceneDev& s, const RenderBufs& b, const ConfigDev& cfg, const LaunchShape& ls,
                         hipStream_t st, hipEvent_t* ev_ext, int n_ev_ext, bool count, const GuideNetDev* guide) {
    hipLaunchKernelGGL(k_pass_init<BATCH>, dim3(1), dim3(256), 0, st, b.cfg, b.ps, b.passes, b.num_passes);
    hipLaunchKernelGGL(k_camera<BATCH>, dim3(ls.grid_stream), dim3(256), 0, st, b.cfg, b, 0);
    const int nd = cfg.max_depth + 1;
    for (int d = 0; d <= cfg.max_depth; d++) {
        int rc = launch_stages<B, BATCH>(s, b, cfg, ls, st, d, d + 1, d, nd, d == cfg.max_depth, ev_ext, n_ev_ext, count,
                                         guide);
        if (rc) return rc;
    }
    return launch_film<BATCH>(b, cfg, ls, st, 0, guide);
}

template <int B, bool BATCH>
// n_ev: profile events recorded per iteration (8 = every stage, 2 = the closest hit), eight
// slots apart.
static int launch_iters_b(const SceneDev& s, const RenderBufs& b, const ConfigDev& cfg, const LaunchShape& ls,
                          hipStream_t st, int it0, int n_it, hipEvent_t* ev, int n_ev, bool count,
                          const GuideNetDev* guide) {
    for (int k = 0; k < n_it; k++) {
        const int slot = (it0 + k) & 1;
        hipLaunchKernelGGL(k_plan, dim3(1), dim3(64), 0, st, b.cfg, b.ps, slot);
        hipLaunchKernelGGL(k_camera<BATCH>, dim3(ls.grid_stream), dim3(256), 0, st, b.cfg, b, slot);
        int rc = launch_stages<B, BATCH>(s, b, cfg, ls, st, slot, slot ^ 1, 0, 1, false, ev ? ev + 8 * k : nullptr,
                                         n_ev, count, guide);
        if (rc) return rc;
    }
    return SPK_OK;
}

int launch_pass(const SceneDev& s, const RenderBufs& b, const ConfigDev& cfg, const LaunchShape& ls, hipStream_t st,
                hipEvent_t* ev_ext, int n_ev_ext, bool count, const GuideNetDev* guide) {
    switch (ls.block) {
    case 64: return launch_pass_b<64, false>(s, b, cfg, ls, st, ev_ext, n_ev_ext, count, guide);
    case 256: return launch_pass_b<256, false>(s, b, cfg, ls, st, ev_ext, n_ev_ext, count, guide);
    default: return launch_pass_b<128, false>(s, b, cfg, ls, st, ev_ext, n_ev_ext, count, guide);
    }
}

int launch_regen_begin(const RenderBufs& b, bool batch, hipStream_t st) {
    if (batch) hipLaunchKernelGGL(k_pass_init<true>, dim3(1), dim3(256), 0, st, b.cfg, b.ps, b.passes, b.num_passes);
    else hipLaunchKernelGGL(k_pass_init<false>, dim3(1), dim3(256), 0, st, b.cfg, b.ps, b.passes, b.num_passes);
    SPK_CHECK_HIP(hipGetLastError());
    return SPK_OK;
}

int launch_regen_iters(const SceneDev& s, const RenderBufs& b, const ConfigDev& cfg, const LaunchShape& ls,
                       hipStream_t st, bool batch, int it0, int n_it, hipEvent_t* ev, int ev_per_it, bool count,
                       const GuideNetDev* guide) {
    const int n_ev = ev ? ev_per_it : 0;
    if (batch) {
        switch (ls.block) {
        case 64: return launch_iters_b<64, true>(s, b, cfg, ls, st, it0, n_it, ev, n_ev, count, guide);
        case 256: return launch_iters_b<256, true>(s, b, cfg, ls, st, it0, n_it, ev, n_ev, count, guide);
        default: return launch_iters_b<128, true>(s, b, cfg, ls, st, it0, n_it, ev, n_ev, count, guide);
        }
    }
    switch (ls.block) {
    case 64: return launch_iters_b<64, false>(s, b, cfg, ls, st, it0, n_it, ev, n_ev, count, guide);
    case 256: return launch_iters_b<256, false>(s, b, cfg, ls, st, it0, n_it, ev, n_ev, count, guide);
    default: return launch_iters_b<128, false>(s, b, cfg, ls, st, it0, n_it, ev, n_ev, count, guide);
    }
}

int launch_regen_end(const RenderBufs& b, const ConfigDev& cfg, const LaunchShape& ls, hipStream_t st, bool batch,
                     int last_slot, const GuideNetDev* guide) {
    return batch ? launch_film<true>(b, cfg, ls, st, last_slot, guide)
                 : launch_film<false>(b, cfg, ls, st, last_slot, guide);
}

int launch_batch_pass(const RenderBufs& b, const ConfigDev& cfg, const LaunchShape& ls, hipStream_t st,
                      const SceneDev* guide_scene, const GuideNetDev* guide) {
    SceneDev none;
    memset(&none, 0, sizeof(none));
    const SceneDev& s = guide_scene ? *guide_scene : none;
    switch (ls.block) {
    case 64: return launch_pass_b<64, true>(s, b, cfg, ls, st, nullptr, 0, false, guide);
    case 256: return launch_pass_b<256, true>(s, b, cfg, ls, st, nullptr, 0, false, guide);
    default: return launch_pass_b<128, true>(s, b, cfg, ls, st, nullptr, 0, false, guide);
    }
}

}  // namespace spk
