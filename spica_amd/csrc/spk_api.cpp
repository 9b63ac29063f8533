// spk_api.cpp -- C ABI of libspk.so: scene upload, batched traces, wavefront renders.
#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <chrono>
#include <memory>
#include <vector>

#include "spk_internal.h"
#include "spk_sampler.h"

namespace spk {

static thread_local char g_err[1024] = "";
static thread_local int g_device = 0;

void set_error(const char* fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof(g_err), fmt, ap);
    va_end(ap);
}
void clear_error() { g_err[0] = 0; }

int DevBuf::alloc(size_t nbytes) {
    release();
    if (nbytes == 0) nbytes = 16;
    hipError_t e = hipMalloc(&p, nbytes);
    if (e != hipSuccess) {
        p = nullptr;
        set_error("hipMalloc(%zu) failed: %s", nbytes, hipGetErrorString(e));
        return SPK_ERR_NOMEM;
    }
    bytes = nbytes;
    return SPK_OK;
}
int DevBuf::upload(const void* src, size_t nbytes) {
    int rc = alloc(nbytes);
    if (rc) return rc;
    if (src && nbytes) SPK_CHECK_HIP(hipMemcpy(p, src, nbytes, hipMemcpyHostToDevice));
    return SPK_OK;
}
void DevBuf::release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    bytes = 0;
}

int zero_u32(uint32_t* p, int n, hipStream_t st);

// Paths per regenerating pass (one path radiance slot each, 16 bytes): the in-flight queues are
// the render's queue size, refilled from the pass's unstarted paths.
#ifndef SPK_REGEN_PASS_LOG2
#define SPK_REGEN_PASS_LOG2 28
#endif
constexpr int64_t kRegenPassPaths = int64_t(1) << SPK_REGEN_PASS_LOG2;
// Iterations of a regenerating pass launched between polls of its done flag.
constexpr int kRegenGroup = 4;
int launch_sampler(int kind, uint32_t seed, int64_t n, const uint32_t* px, const uint32_t* sm, const uint32_t* dm,
                   uint32_t* out, hipStream_t st);

}  // namespace spk

using namespace spk;

// ---------------------------------------------------------------------------------
// double-precision helpers for the host-side set-up (spica computes in FP64)
// ---------------------------------------------------------------------------------
namespace {

struct D3 {
    double x, y, z;
};
inline D3 d3(double x, double y, double z) { return D3{x, y, z}; }
inline D3 dsub(D3 a, D3 b) { return d3(a.x - b.x, a.y - b.y, a.z - b.z); }
inline D3 dadd(D3 a, D3 b) { return d3(a.x + b.x, a.y + b.y, a.z + b.z); }
inline D3 dmul(D3 a, double s) { return d3(a.x * s, a.y * s, a.z * s); }
inline double ddot(D3 a, D3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline D3 dcross(D3 a, D3 b) { return d3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x); }
inline double dnorm(D3 a) { return sqrt(ddot(a, a)); }
inline D3 dnormalize(D3 a) {
    double n = dnorm(a);
    return d3(a.x / n, a.y / n, a.z / n);
}
// vect_math.h:121-128
inline void dcoord(D3 w, D3* u, D3* v) {
    if (fabs(w.x) > fabs(w.y)) *u = dnormalize(d3(-w.z, 0.0, w.x));
    else *u = dnormalize(d3(0.0, w.z, -w.y));
    *v = dnormalize(dcross(w, *u));
}
inline float4 f4(D3 a, float w) { return make_float4((float)a.x, (float)a.y, (float)a.z, w); }
inline double dgray(double r, double g, double b) { return 0.2126 * r + 0.7152 * g + 0.0722 * b; }

// Triangle frame as spica builds it (triangle.cc:17-52, 119-138).
struct TriFrame {
    D3 p[3], n[3], fn, ng;
    double area;
    bool smooth;
};
TriFrame tri_frame(const float* v, const float* nrm) {
    TriFrame t;
    for (int i = 0; i < 3; i++) t.p[i] = d3(v[3 * i], v[3 * i + 1], v[3 * i + 2]);
    D3 e1 = dsub(t.p[1], t.p[0]), e2 = dsub(t.p[2], t.p[0]);
    D3 c = dcross(e1, e2);
    t.smooth = false;
    // a triangle whose three normals are all zero is flat (a flat mesh among smooth ones)
    bool zero = nrm != nullptr;
    for (int i = 0; nrm && i < 9; i++) zero = zero && nrm[i] == 0.0f;
    if (!nrm || zero) {
        t.fn = dnormalize(c);
        t.n[0] = t.n[1] = t.n[2] = t.fn;
    } else {
        for (int i = 0; i < 3; i++) t.n[i] = dnormalize(d3(nrm[3 * i], nrm[3 * i + 1], nrm[3 * i + 2]));
        D3 fn = c;
        if (dnorm(fn) < 1e-12) fn = dmul(dadd(dadd(t.n[0], t.n[1]), t.n[2]), 1.0 / 3.0);
        t.fn = dnormalize(fn);
        for (int i = 0; i < 3; i++)
            if (fabs(ddot(t.n[i], t.fn)) < 1.0 - 1e-12) t.smooth = true;
    }
    D3 du, dv;
    dcoord(t.fn, &du, &dv);
    t.ng = dnormalize(dcross(du, dv));
    t.area = 0.5 * dnorm(c);
    return t;
}

// ---- environment map distribution (envmap.cc:17-47, mipmap.cc, sampling.cc) ----
struct Mip {
    int levels;
    std::vector<int> w, h;
    std::vector<std::vector<double>> px;
    D3 texel(int l, int s, int t) const {
        int W = w[l], H = h[l];
        s = (s % W + W) % W;
        t = (t % H + H) % H;
        const double* p = &px[l][((size_t)t * W + s) * 3];
        return d3(p[0], p[1], p[2]);
    }
    D3 bilinear(int l, double s0, double t0) const {
        l = std::max(0, std::min(l, levels));
        double s = s0 * w[l] - 0.5, t = t0 * h[l] - 0.5;
        int si = (int)s, ti = (int)t;
        double ds = s - si, dt = t - ti;
        D3 r = dmul(texel(l, si, ti), (1 - ds) * (1 - dt));
        r = dadd(r, dmul(texel(l, si + 1, ti), ds * (1 - dt)));
        r = dadd(r, dmul(texel(l, si, ti + 1), (1 - ds) * dt));
        r = dadd(r, dmul(texel(l, si + 1, ti + 1), ds * dt));
        return r;
    }
    D3 lookup(double s, double t, double width) const {
        double level = levels - 1 + log2(std::max(width, 1.0e-8));
        if (level < 0) return bilinear(0, s, t);
        if (level >= levels - 1) return texel(levels - 1, 0, 0);
        int l = (int)level;
        double delta = level - l;
        return dadd(dmul(bilinear(l, s, t), 1 - delta), dmul(bilinear(l + 1, s, t), delta));
    }
};
int round_up_pow2(int v) {
    v--;
    v |= v >> 1;
    v |= v >> 2;
    v |= v >> 4;
    v |= v >> 8;
    v |= v >> 16;
    return v + 1;
}
int log2_int(int v) {
    int r = 0;
    while (v) {
        r++;
        v >>= 1;
    }
    return r;
}
void build_mip(Mip& m, int w, int h, std::vector<double>&& base) {
    int rx = w, ry = h;
    bool p2 = (w && !(w & (w - 1))) && (h && !(h & (h - 1)));
    if (!p2) {
        rx = round_up_pow2(w);
        ry = round_up_pow2(h);
    }
    m.levels = 1 + log2_int(std::max(rx, ry));
    m.w.assign(m.levels, 0);
    m.h.assign(m.levels, 0);
    m.px.assign(m.levels, {});
    m.w[0] = w;
    m.h[0] = h;
    m.px[0] = std::move(base);
    for (int i = 1; i < m.levels; i++) {
        int sr = std::max(1, m.w[i - 1] / 2), tr = std::max(1, m.h[i - 1] / 2);
        m.w[i] = sr;
        m.h[i] = tr;
        m.px[i].assign((size_t)sr * tr * 3, 0.0);
        for (int y = 0; y < tr; y++)
            for (int x = 0; x < sr; x++) {
                D3 s = dadd(dadd(m.texel(i - 1, 2 * x, 2 * y), m.texel(i - 1, 2 * x + 1, 2 * y)),
                            dadd(m.texel(i - 1, 2 * x, 2 * y + 1), m.texel(i - 1, 2 * x + 1, 2 * y + 1)));
                s = dmul(s, 0.25);
                double* p = &m.px[i][((size_t)y * sr + x) * 3];
                p[0] = s.x;
                p[1] = s.y;
                p[2] = s.z;
            }
    }
}
struct Dist1 {
    std::vector<double> func, cdf;
    double integral;
    void init(const double* f, int n) {
        func.assign(f, f + n);
        cdf.assign(n + 1, 0.0);
        for (int i = 1; i <= n; i++) cdf[i] = cdf[i - 1] + func[i - 1] / n;
        integral = cdf[n];
        if (integral == 0.0) {
            for (int i = 1; i <= n; i++) cdf[i] = (double)i / n;
        } else {
            for (int i = 1; i <= n; i++) cdf[i] /= integral;
        }
    }
};

}  // namespace

// ---------------------------------------------------------------------------------
// handles
// ---------------------------------------------------------------------------------
struct spk_scene {
    int device = 0;
    SceneDev dev;
    DevBuf nodes, tris, tri_ng, tri_ids, tri_n, mats, lights;
    DevBuf env_tex, env_marg_cdf, env_marg_pdf, env_cond_cdf, env_cond_pdf, env_pdf2d;
    DevBuf trace_work, trace_spill;
    int trace_grid = 0;
    int spill_cap = 0;
    float lo[3] = {0.0f, 0.0f, 0.0f}, hi[3] = {1.0f, 1.0f, 1.0f};  // bounding box of the triangles
    spk_scene_stats stats;
};

// PerspectiveCamera (cameras/perspective.cc:20-51) as the kernels use it.
static void camera_dev(const spk_camera_desc* cam, CameraDev& d) {
    memset(&d, 0, sizeof(d));
    const double PI = 3.14159265358979323846;
    double tanHalf = tan(cam->fov_deg * PI / 180.0 / 2.0);
    double aspect = (double)cam->width / cam->height;
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 4; j++) d.m[i * 4 + j] = (float)cam->to_world[i * 4 + j];
    d.sx_scale = (float)(aspect * tanHalf);
    d.sy_scale = (float)tanHalf;
    d.lens_r = (float)cam->aperture_radius;
    d.focal = (float)cam->focus_distance;
    d.W = cam->width;
    d.H = cam->height;
}

struct spk_render {
    spk_scene* scene = nullptr;
    int device = 0;
    hipStream_t stream = nullptr;
    ConfigDev cfg;
    spk_render_desc desc;
    spk_render_opts opts;
    RenderBufs bufs;
    LaunchShape shape;
    DevBuf qa_o[2], qa_d[2], qa_b[2], hit, qs_o, qs_d, qs_c, qm_o, qm_d, qm_c, path_L;
    DevBuf film_sum, film_c, film_sq, film_sqc, spill, cfg_dev, ps_dev, scratch;
    uint32_t capacity = 0;          // paths per pass (path radiance slots)
    uint32_t inflight = 0;          // queue capacity: = capacity for classic passes
    uint32_t* rg_flag = nullptr;    // regen: pinned host words the done flag is copied to
    hipEvent_t rg_ev[2] = {nullptr, nullptr};
    hipGraphExec_t rg_exec = nullptr;  // use_graph with regen: one captured group of iterations
    int64_t rg_iterations = 0;      // regen iterations launched
    int64_t samples_done = 0;
    double extend_ms = 0.0;
    int64_t extend_launches = 0;
    double guide_ms = 0.0;          // profile = 1: summed guide-kernel time
    int64_t guide_launches = 0;
    double shade_ms = 0.0, shadow_ms = 0.0;
    int64_t shade_launches = 0, shadow_launches = 0, shadow_rays_timed = 0;
    int64_t extend_rays_timed = 0;
    std::vector<hipEvent_t> events;
    hipGraphExec_t pass_exec = nullptr;  // use_graph: one captured plain pass, replayed
    // learned proposal
    spk_guide* guide = nullptr;
    DevBuf qg[4], rec[5], rec_s[5], rec_keys, rec_idx, rec_tmp;  // rec_s: the records sorted
    uint32_t rec_cap = 0;
    int64_t guide_trained = 0;   // training samples per pixel done since attaching
    int64_t train_passes = 0;
    // Device buffers free themselves after this body; the stream, events and graph are
    // released here, so an early error return from spk_render_create leaks nothing.
    ~spk_render() {
        (void)hipSetDevice(device);
        if (stream) {
            (void)hipStreamSynchronize(stream);
            (void)hipStreamDestroy(stream);
        }
        for (hipEvent_t e : events) (void)hipEventDestroy(e);
        if (pass_exec) (void)hipGraphExecDestroy(pass_exec);
        if (rg_exec) (void)hipGraphExecDestroy(rg_exec);
        for (hipEvent_t e : rg_ev)
            if (e) (void)hipEventDestroy(e);
        if (rg_flag) (void)hipHostFree(rg_flag);
    }
};

extern "C" {

const char* spk_last_error(void) { return g_err; }

const char* spk_version(void) { return "spica-amd 0.1 (gfx950 wavefront path tracer)"; }

int spk_init(int device) {
    clear_error();
    int n = 0;
    SPK_CHECK_HIP(hipGetDeviceCount(&n));
    if (device < 0 || device >= n) {
        set_error("spk_init: device %d out of range (%d devices)", device, n);
        return SPK_ERR_INVALID;
    }
    SPK_CHECK_HIP(hipSetDevice(device));
    g_device = device;
    return SPK_OK;
}

int spk_scene_create(const spk_scene_desc* d, spk_scene** out) { return spk_scene_create_ex(d, nullptr, out); }

int spk_scene_build_opts_default(spk_scene_build_opts* o) {
    clear_error();
    if (!o) {
        set_error("spk_scene_build_opts_default: null argument");
        return SPK_ERR_INVALID;
    }
    o->builder = SPK_BVH_SAH;
    o->max_leaf = 4;
    o->treelet_passes = 6;
    return SPK_OK;
}

int spk_scene_create_ex(const spk_scene_desc* d, const spk_scene_build_opts* bopts, spk_scene** out) {
    clear_error();
    if (!d || !out) {
        set_error("spk_scene_create: null argument");
        return SPK_ERR_INVALID;
    }
    spk_scene_build_opts bop;
    spk_scene_build_opts_default(&bop);
    if (bopts) bop = *bopts;
    if ((bop.builder != SPK_BVH_SAH && bop.builder != SPK_BVH_LBVH) || bop.max_leaf < 1 || bop.max_leaf > kMaxLeafTris ||
        bop.treelet_passes < 0 || bop.treelet_passes > 16) {
        set_error("spk_scene_create: bad build options (builder %d, max_leaf %d, treelet_passes %d; max_leaf in [1, %d], "
                  "treelet_passes in [0, 16])",
                  bop.builder, bop.max_leaf, bop.treelet_passes, kMaxLeafTris);
        return SPK_ERR_INVALID;
    }
    *out = nullptr;
    const int64_t n = d->num_tris;
    if (n < 0 || (n > 0 && (!d->verts || !d->mat_ids)) || n >= (int64_t(1) << 28)) {
        set_error("spk_scene_create: bad triangle arrays (num_tris=%lld)", (long long)n);
        return SPK_ERR_INVALID;
    }
    if (d->num_materials <= 0 && n > 0) {
        set_error("spk_scene_create: no materials");
        return SPK_ERR_INVALID;
    }
    if (d->num_area_lights < 0 || (d->num_area_lights > 0 && (!d->area_tri || !d->area_radiance || !d->light_ids))) {
        set_error("spk_scene_create: bad area light arrays");
        return SPK_ERR_INVALID;
    }
    for (int64_t i = 0; i < n; i++) {
        if (d->mat_ids[i] < 0 || d->mat_ids[i] >= d->num_materials) {
            set_error("spk_scene_create: triangle %lld has material %d outside [0,%d)", (long long)i, d->mat_ids[i],
                      d->num_materials);
            return SPK_ERR_INVALID;
        }
        int li = d->light_ids ? d->light_ids[i] : -1;
        if (li < -1 || li >= d->num_area_lights || (li >= 0 && d->area_tri[li] != i)) {
            set_error("spk_scene_create: triangle %lld has inconsistent light id %d", (long long)i, li);
            return SPK_ERR_INVALID;
        }
    }
    for (int i = 0; i < d->num_area_lights; i++) {
        int t = d->area_tri[i];
        if (t < 0 || t >= n || d->light_ids[t] != i) {
            set_error("spk_scene_create: area light %d names triangle %d inconsistently", i, t);
            return SPK_ERR_INVALID;
        }
    }
    for (int m = 0; m < d->num_materials; m++) {
        int ty = d->materials[m].type;
        if (ty < SPK_MAT_DIFFUSE || ty > SPK_MAT_ROUGHPLASTIC) {
            set_error("spk_scene_create: material %d has unknown type %d", m, ty);
            return SPK_ERR_INVALID;
        }
        const int dist = d->materials[m].distribution;
        if ((ty == SPK_MAT_ROUGHCONDUCTOR || ty == SPK_MAT_ROUGHDIELECTRIC || ty == SPK_MAT_ROUGHPLASTIC) &&
            dist != SPK_MICROFACET_GGX && dist != SPK_MICROFACET_BECKMANN) {
            set_error("spk_scene_create: material %d has unknown microfacet distribution %d", m, dist);
            return SPK_ERR_INVALID;
        }
    }
    if (d->env_width > 0 && (d->env_height <= 0 || !d->env_rgb)) {
        set_error("spk_scene_create: bad environment map");
        return SPK_ERR_INVALID;
    }
    // spica's light keeps the transpose of toWorld and inverts that whole 4x4 matrix
    // (envmap.cc:19, light.cc:6-10); for an affine toWorld the transpose's last column is
    // (0, 0, 0, 1), so the direction part of that inverse is the inverse of the transposed 3x3,
    // which is what the kernels apply.  A projective toWorld would break that identity.
    if (d->env_width > 0 && (d->env_to_world[12] != 0.0 || d->env_to_world[13] != 0.0 ||
                             d->env_to_world[14] != 0.0 || d->env_to_world[15] != 1.0)) {
        set_error("spk_scene_create: environment toWorld must be affine (last row 0 0 0 1)");
        return SPK_ERR_INVALID;
    }

    std::unique_ptr<spk_scene> s(new spk_scene());
    s->device = g_device;
    SPK_CHECK_HIP(hipSetDevice(s->device));
    memset(&s->dev, 0, sizeof(s->dev));
    memset(&s->stats, 0, sizeof(s->stats));
    int rc;

    // BVH
    auto t0 = std::chrono::steady_clock::now();
    BvhBuildOpts bo;
    bo.max_leaf = bop.max_leaf;
    bo.treelet_passes = bop.treelet_passes;
    BvhResult bvh;
    if (bop.builder == SPK_BVH_LBVH) {
        if ((rc = build_bvh_gpu(d->verts, n, bo, bvh))) return rc;
    } else {
        bvh = build_bvh(d->verts, n, bo);
    }
    std::vector<Bvh4QNodeHost> nodes4;
    int depth4 = 0;
    {
        std::vector<Bvh4NodeHost> wide;
        depth4 = collapse_bvh4(bvh.nodes, wide);
        quantize_bvh4(wide, nodes4);
    }
    auto t1 = std::chrono::steady_clock::now();

    // Leaf-ordered triangle data (FP64 set-up, FP32 storage).
    std::vector<float4> tris((size_t)n * 3), tri_ng((size_t)n), tri_n;
    std::vector<int2> tri_ids((size_t)n);
    bool any_smooth = false;
    std::vector<TriFrame> frames((size_t)n);
    for (int64_t i = 0; i < n; i++) {
        frames[(size_t)i] = tri_frame(d->verts + 9 * i, d->normals ? d->normals + 9 * i : nullptr);
        any_smooth = any_smooth || frames[(size_t)i].smooth;
    }
    if (any_smooth) tri_n.resize((size_t)n * 3);
    for (int64_t k = 0; k < n; k++) {
        uint32_t prim = bvh.order[(size_t)k];
        const TriFrame& f = frames[prim];
        const float* v = d->verts + 9 * (size_t)prim;
        float4 T0 = make_float4(v[0], v[1], v[2], 0.0f);
        memcpy(&T0.w, &prim, 4);
        tris[3 * k] = T0;
        tris[3 * k + 1] = f4(dsub(f.p[1], f.p[0]), 0.0f);
        tris[3 * k + 2] = f4(dsub(f.p[2], f.p[0]), 0.0f);
        uint32_t mbits = (uint32_t)d->mat_ids[prim] | (f.smooth ? 0x80000000u : 0u);
        float w;
        memcpy(&w, &mbits, 4);
        tri_ng[k] = f4(f.ng, w);
        tri_ids[k] = make_int2((int)prim, d->light_ids ? d->light_ids[prim] : -1);
        if (any_smooth)
            for (int j = 0; j < 3; j++) tri_n[3 * k + j] = f4(f.n[j], 0.0f);
    }

    // Materials (bsdfs/*.cc setScatterFuncs).
    std::vector<MatDev> mats((size_t)std::max(1, d->num_materials));
    for (int m = 0; m < d->num_materials; m++) {
        const spk_material& M = d->materials[m];
        MatDev md;
        memset(&md, 0, sizeof(md));
        auto cl = [](double x) { return std::max(0.0, std::min(x, 1.0e32)); };
        int lobe = LOBE_NONE;
        D3 R = d3(1, 1, 1), T = d3(0, 0, 0);
        double etaB = 1.0, alpha = 0.0, specW = 0.0;
        const int dist = M.distribution == SPK_MICROFACET_BECKMANN ? 1 : 0;
        switch (M.type) {
        case SPK_MAT_DIFFUSE:
            R = d3(cl(M.reflectance[0]), cl(M.reflectance[1]), cl(M.reflectance[2]));
            lobe = (R.x == 0 && R.y == 0 && R.z == 0) ? LOBE_NONE : LOBE_LAMBERT;
            break;
        case SPK_MAT_CONDUCTOR:
            lobe = LOBE_SPECREFL;
            break;
        case SPK_MAT_DIELECTRIC:
            R = d3(cl(M.reflectance[0]), cl(M.reflectance[1]), cl(M.reflectance[2]));
            T = d3(cl(M.transmittance[0]), cl(M.transmittance[1]), cl(M.transmittance[2]));
            etaB = dgray(M.int_ior, M.int_ior, M.int_ior);
            lobe = (R.x == 0 && R.y == 0 && R.z == 0 && T.x == 0 && T.y == 0 && T.z == 0) ? LOBE_NONE : LOBE_FRESNELSPEC;
            break;
        case SPK_MAT_ROUGHCONDUCTOR:
            alpha = dgray(M.alpha, M.alpha, M.alpha);
            lobe = alpha == 0.0 ? LOBE_SPECREFL : LOBE_MICROREFL;
            break;
        case SPK_MAT_PLASTIC:  // PlasticBRDF(Kd, Ks, 1, gray(intIOR)), no clamping
            R = d3(M.reflectance[0], M.reflectance[1], M.reflectance[2]);
            T = d3(M.specular[0], M.specular[1], M.specular[2]);
            etaB = dgray(M.int_ior, M.int_ior, M.int_ior);
            specW = dgray(T.x, T.y, T.z) / (dgray(R.x, R.y, R.z) + dgray(T.x, T.y, T.z));
            lobe = LOBE_PLASTIC;
            break;
        case SPK_MAT_ROUGHDIELECTRIC:
            R = d3(cl(M.reflectance[0]), cl(M.reflectance[1]), cl(M.reflectance[2]));
            T = d3(cl(M.transmittance[0]), cl(M.transmittance[1]), cl(M.transmittance[2]));
            etaB = dgray(M.int_ior, M.int_ior, M.int_ior);
            alpha = dgray(M.alpha, M.alpha, M.alpha);
            if (R.x == 0 && R.y == 0 && R.z == 0 && T.x == 0 && T.y == 0 && T.z == 0) lobe = LOBE_NONE;
            else lobe = alpha == 0.0 ? LOBE_FRESNELSPEC : LOBE_MICROTRANS;
            break;
        case SPK_MAT_ROUGHPLASTIC:
            R = d3(M.reflectance[0], M.reflectance[1], M.reflectance[2]);
            T = d3(M.specular[0], M.specular[1], M.specular[2]);
            if (T.x == 0 && T.y == 0 && T.z == 0) {  // LambertianReflection(kd), unclamped
                lobe = LOBE_LAMBERT;
                break;
            }
            etaB = dgray(M.int_ior, M.int_ior, M.int_ior);
            alpha = std::max(0.001, dgray(M.alpha, M.alpha, M.alpha));
            specW = dgray(T.x, T.y, T.z) / (dgray(R.x, R.y, R.z) + dgray(T.x, T.y, T.z));
            lobe = LOBE_ROUGHPLASTIC;
            break;
        }
        const int lobe_bits = lobe | (dist << 8);
        float lw;
        memcpy(&lw, &lobe_bits, 4);
        md.a = f4(R, lw);
        md.b = f4(T, (float)etaB);
        md.c = f4(d3(M.eta[0], M.eta[1], M.eta[2]), (float)alpha);
        md.d = f4(d3(M.k[0], M.k[1], M.k[2]), (float)specW);
        mats[(size_t)m] = md;
    }

    // Area lights.
    std::vector<LightDev> lights((size_t)std::max(1, d->num_area_lights));
    for (int i = 0; i < d->num_area_lights; i++) {
        const TriFrame& f = frames[(size_t)d->area_tri[i]];
        LightDev L;
        L.p0 = f4(f.p[0], (float)f.area);
        L.e1 = f4(dsub(f.p[1], f.p[0]), 0.0f);
        L.e2 = f4(dsub(f.p[2], f.p[0]), 0.0f);
        L.n0 = f4(f.n[0], 0.0f);
        L.n1 = f4(f.n[1], 0.0f);
        L.n2 = f4(f.n[2], 0.0f);
        L.L = f4(d3(d->area_radiance[3 * i], d->area_radiance[3 * i + 1], d->area_radiance[3 * i + 2]), 0.0f);
        L.ng = f4(f.ng, 0.0f);
        lights[(size_t)i] = L;
    }

    // Environment map.
    EnvDev env;
    memset(&env, 0, sizeof(env));
    if (d->env_width > 0) {
        int w = d->env_width, h = d->env_height;
        std::vector<double> base((size_t)w * h * 3);
        for (size_t i = 0; i < base.size(); i++) base[i] = (double)d->env_rgb[i] * d->env_scale;
        Mip mip;
        build_mip(mip, w, h, std::move(base));
        double filter = 1.0 / std::max(w, h);
        std::vector<double> g((size_t)w * h);
        const double PI = 3.14159265358979323846;
        for (int v = 0; v < h; v++) {
            double vp = (double)(v + 0.5) / h, st = sin(PI * (v + 0.5) / h);
            for (int u = 0; u < w; u++) {
                D3 c = mip.lookup((double)(u + 0.5) / w, vp, filter);
                g[(size_t)v * w + u] = dgray(c.x, c.y, c.z) * st;
            }
        }
        std::vector<Dist1> cond((size_t)h);
        std::vector<double> mf((size_t)h);
        for (int v = 0; v < h; v++) {
            cond[(size_t)v].init(&g[(size_t)v * w], w);
            mf[(size_t)v] = cond[(size_t)v].integral;
        }
        Dist1 marg;
        marg.init(mf.data(), h);
        std::vector<float4> tex((size_t)w * h);
        for (size_t i = 0; i < tex.size(); i++)
            tex[i] = make_float4((float)mip.px[0][3 * i], (float)mip.px[0][3 * i + 1], (float)mip.px[0][3 * i + 2], 0.0f);
        std::vector<float> mcdf((size_t)h + 1), mpdf((size_t)h), ccdf((size_t)h * (w + 1)), cpdf((size_t)h * w),
            pdf2((size_t)h * w);
        for (int v = 0; v <= h; v++) mcdf[(size_t)v] = (float)marg.cdf[(size_t)v];
        for (int v = 0; v < h; v++) {
            mpdf[(size_t)v] = (float)(marg.func[(size_t)v] / marg.integral);
            for (int u = 0; u <= w; u++) ccdf[(size_t)v * (w + 1) + u] = (float)cond[(size_t)v].cdf[(size_t)u];
            for (int u = 0; u < w; u++) {
                cpdf[(size_t)v * w + u] = (float)(cond[(size_t)v].func[(size_t)u] / cond[(size_t)v].integral);
                pdf2[(size_t)v * w + u] = (float)(cond[(size_t)v].func[(size_t)u] / marg.integral);
            }
        }
        if ((rc = s->env_tex.upload(tex.data(), tex.size() * sizeof(float4)))) return rc;
        if ((rc = s->env_marg_cdf.upload(mcdf.data(), mcdf.size() * 4))) return rc;
        if ((rc = s->env_marg_pdf.upload(mpdf.data(), mpdf.size() * 4))) return rc;
        if ((rc = s->env_cond_cdf.upload(ccdf.data(), ccdf.size() * 4))) return rc;
        if ((rc = s->env_cond_pdf.upload(cpdf.data(), cpdf.size() * 4))) return rc;
        if ((rc = s->env_pdf2d.upload(pdf2.data(), pdf2.size() * 4))) return rc;
        env.present = 1;
        env.w = w;
        env.h = h;
        env.radius = (float)d->env_radius;
        env.tex = s->env_tex.as<float4>();
        env.marg_cdf = s->env_marg_cdf.as<float>();
        env.marg_pdf = s->env_marg_pdf.as<float>();
        env.cond_cdf = s->env_cond_cdf.as<float>();
        env.cond_pdf = s->env_cond_pdf.as<float>();
        env.pdf2d = s->env_pdf2d.as<float>();
        // envmap.cc:19 keeps the transpose of toWorld; Light keeps its inverse.
        double l2w[3][3], w2l[3][3];
        for (int i = 0; i < 3; i++)
            for (int j = 0; j < 3; j++) l2w[i][j] = d->env_to_world[j * 4 + i];
        double det = l2w[0][0] * (l2w[1][1] * l2w[2][2] - l2w[1][2] * l2w[2][1]) -
                     l2w[0][1] * (l2w[1][0] * l2w[2][2] - l2w[1][2] * l2w[2][0]) +
                     l2w[0][2] * (l2w[1][0] * l2w[2][1] - l2w[1][1] * l2w[2][0]);
        if (det == 0.0) {
            set_error("spk_scene_create: singular environment transform");
            return SPK_ERR_INVALID;
        }
        double id = 1.0 / det;
        w2l[0][0] = (l2w[1][1] * l2w[2][2] - l2w[1][2] * l2w[2][1]) * id;
        w2l[0][1] = (l2w[0][2] * l2w[2][1] - l2w[0][1] * l2w[2][2]) * id;
        w2l[0][2] = (l2w[0][1] * l2w[1][2] - l2w[0][2] * l2w[1][1]) * id;
        w2l[1][0] = (l2w[1][2] * l2w[2][0] - l2w[1][0] * l2w[2][2]) * id;
        w2l[1][1] = (l2w[0][0] * l2w[2][2] - l2w[0][2] * l2w[2][0]) * id;
        w2l[1][2] = (l2w[0][2] * l2w[1][0] - l2w[0][0] * l2w[1][2]) * id;
        w2l[2][0] = (l2w[1][0] * l2w[2][1] - l2w[1][1] * l2w[2][0]) * id;
        w2l[2][1] = (l2w[0][1] * l2w[2][0] - l2w[0][0] * l2w[2][1]) * id;
        w2l[2][2] = (l2w[0][0] * l2w[1][1] - l2w[0][1] * l2w[1][0]) * id;
        for (int i = 0; i < 3; i++)
            for (int j = 0; j < 3; j++) {
                env.l2w[i * 3 + j] = (float)l2w[i][j];
                env.w2l[i * 3 + j] = (float)w2l[i][j];
            }
    }

    // Upload.
    if ((rc = s->nodes.upload(nodes4.data(), nodes4.size() * sizeof(Bvh4QNodeHost)))) return rc;
    if ((rc = s->tris.upload(tris.data(), tris.size() * sizeof(float4)))) return rc;
    if ((rc = s->tri_ng.upload(tri_ng.data(), tri_ng.size() * sizeof(float4)))) return rc;
    if ((rc = s->tri_ids.upload(tri_ids.data(), tri_ids.size() * sizeof(int2)))) return rc;
    if (any_smooth && (rc = s->tri_n.upload(tri_n.data(), tri_n.size() * sizeof(float4)))) return rc;
    if ((rc = s->mats.upload(mats.data(), mats.size() * sizeof(MatDev)))) return rc;
    if ((rc = s->lights.upload(lights.data(), lights.size() * sizeof(LightDev)))) return rc;
    auto t2 = std::chrono::steady_clock::now();

    SceneDev& sd = s->dev;
    sd.nodes = s->nodes.as<float4>();
    sd.tris = s->tris.as<float4>();
    sd.tri_ng = s->tri_ng.as<float4>();
    sd.tri_ids = s->tri_ids.as<int2>();
    sd.tri_n = any_smooth ? s->tri_n.as<float4>() : nullptr;
    sd.mats = s->mats.as<MatDev>();
    sd.lights = s->lights.as<LightDev>();
    sd.num_areas = d->num_area_lights;
    sd.num_lights = d->num_area_lights + (env.present ? 1 : 0);
    sd.has_bvh = n > 0 ? 1 : 0;
    sd.wide = 0;
    for (int m = 0; m < d->num_materials; m++) {
        const spk_material& M = d->materials[m];
        if (M.type >= SPK_MAT_PLASTIC || (M.type == SPK_MAT_ROUGHCONDUCTOR && M.distribution == SPK_MICROFACET_BECKMANN))
            sd.wide = 1;
    }
    sd.env = env;

    // a 4-wide visit pushes at most three children, so a ray's stack never holds more
    // than 3 per level below the root
    s->spill_cap = std::max(8, 3 * depth4 + 2);
    for (int k = 0; k < 3; k++) {
        s->lo[k] = 3.0e38f;
        s->hi[k] = -3.0e38f;
    }
    for (int64_t i = 0; i < 9 * n; i++) {
        s->lo[i % 3] = std::min(s->lo[i % 3], d->verts[i]);
        s->hi[i % 3] = std::max(s->hi[i % 3], d->verts[i]);
    }
    if (n == 0)
        for (int k = 0; k < 3; k++) {
            s->lo[k] = 0.0f;
            s->hi[k] = 1.0f;
        }
    s->stats.num_tris = n;
    s->stats.num_nodes = (int64_t)nodes4.size();
    s->stats.num_leaves = bvh.num_leaves;
    s->stats.max_depth = depth4;
    s->stats.num_lights = sd.num_lights;
    s->stats.build_ms = std::chrono::duration<double, std::milli>(t1 - t0).count();
    s->stats.upload_ms = std::chrono::duration<double, std::milli>(t2 - t1).count();
    s->stats.device_bytes = (int64_t)(s->nodes.bytes + s->tris.bytes + s->tri_ng.bytes + s->tri_ids.bytes +
                                      s->tri_n.bytes + s->mats.bytes + s->lights.bytes + s->env_tex.bytes +
                                      s->env_cond_cdf.bytes + s->env_cond_pdf.bytes + s->env_pdf2d.bytes);
    *out = s.release();
    return SPK_OK;
}

int spk_scene_destroy(spk_scene* s) {
    clear_error();
    if (!s) return SPK_OK;
    (void)hipSetDevice(s->device);
    (void)hipDeviceSynchronize();
    DevBuf* bufs[] = {&s->nodes, &s->tris, &s->tri_ng, &s->tri_ids, &s->tri_n, &s->mats, &s->lights,
                      &s->env_tex, &s->env_marg_cdf, &s->env_marg_pdf, &s->env_cond_cdf, &s->env_cond_pdf,
                      &s->env_pdf2d, &s->trace_work, &s->trace_spill};
    for (DevBuf* b : bufs) b->release();
    delete s;
    return SPK_OK;
}

}  // extern "C"

void spk::scene_trace_info(const spk_scene* s, SceneTraceInfo* out) {
    out->dev = &s->dev;
    out->device = s->device;
    out->spill_cap = s->spill_cap;
}

extern "C" {

int spk_scene_stats_get(const spk_scene* s, spk_scene_stats* out) {
    clear_error();
    if (!s || !out) {
        set_error("spk_scene_stats_get: null argument");
        return SPK_ERR_INVALID;
    }
    *out = s->stats;
    return SPK_OK;
}

int spk_trace(spk_scene* s, const float* org, const float* dir, const float* tmax, int64_t n, int32_t any_hit,
              int32_t* prim_out, float* t_out, void* stream) {
    clear_error();
    if (!s || n < 0 || (n > 0 && (!org || !dir || !prim_out))) {
        set_error("spk_trace: bad arguments");
        return SPK_ERR_INVALID;
    }
    if (n >= (int64_t(1) << 32) - 64) {
        set_error("spk_trace: n too large");
        return SPK_ERR_INVALID;
    }
    if (n == 0) return SPK_OK;
    SPK_CHECK_HIP(hipSetDevice(s->device));
    hipStream_t st = (hipStream_t)stream;
    int rc;
    if (!s->trace_work.p) {
        hipDeviceProp_t prop;
        SPK_CHECK_HIP(hipGetDeviceProperties(&prop, s->device));
        s->trace_grid = prop.multiProcessorCount * std::max(1, trace_blocks_per_cu(128));
        if ((rc = s->trace_work.alloc(64))) return rc;
        if ((rc = s->trace_spill.alloc((size_t)s->trace_grid * 128 * s->spill_cap * 4))) return rc;
    }
    if ((rc = zero_u32(s->trace_work.as<uint32_t>(), 16, st))) return rc;
    TraceBufs b;
    b.org = org;
    b.dir = dir;
    b.tmax = tmax;
    b.prim_out = prim_out;
    b.t_out = t_out;
    b.n = n;
    b.any_hit = any_hit ? 1 : 0;
    return launch_trace(s->dev, b, s->trace_work.as<uint32_t>(), s->trace_spill.as<uint32_t>(), s->spill_cap,
                        s->trace_grid, st);
}

int spk_render_create(spk_scene* s, const spk_camera_desc* cam, const spk_render_desc* desc, const spk_render_opts* opts,
                      spk_render** out) {
    clear_error();
    if (!s || !cam || !desc || !out) {
        set_error("spk_render_create: null argument");
        return SPK_ERR_INVALID;
    }
    *out = nullptr;
    if (cam->width <= 0 || cam->height <= 0 || (int64_t)cam->width * cam->height >= (int64_t(1) << 31)) {
        set_error("spk_render_create: bad film size %dx%d", cam->width, cam->height);
        return SPK_ERR_INVALID;
    }
    if (desc->max_depth < 0 || desc->max_depth > 60) {
        set_error("spk_render_create: maxDepth %d outside [0, 60]", desc->max_depth);
        return SPK_ERR_INVALID;
    }
    if (desc->sampler != SPK_SAMPLER_INDEPENDENT && desc->sampler != SPK_SAMPLER_SOBOL) {
        set_error("spk_render_create: sampler %d not available on the GPU", desc->sampler);
        return SPK_ERR_INVALID;
    }
    if (desc->integrator != SPK_INTEGRATOR_PATH && desc->integrator != SPK_INTEGRATOR_DIRECT) {
        set_error("spk_render_create: unknown integrator %d", desc->integrator);
        return SPK_ERR_INVALID;
    }
    if (cam->to_world[12] != 0.0 || cam->to_world[13] != 0.0 || cam->to_world[14] != 0.0 || cam->to_world[15] != 1.0) {
        set_error("spk_render_create: camera toWorld must be affine");
        return SPK_ERR_INVALID;
    }
    std::unique_ptr<spk_render> r(new spk_render());
    r->scene = s;
    r->device = s->device;
    SPK_CHECK_HIP(hipSetDevice(r->device));
    r->desc = *desc;
    spk_render_opts o;
    memset(&o, 0, sizeof(o));
    if (opts) o = *opts;
    if (o.queue_size <= 0) o.queue_size = int64_t(1) << 21;
    if (o.queue_size < 64) o.queue_size = 64;
    if (o.queue_size > (int64_t(1) << 27)) o.queue_size = int64_t(1) << 27;
    if (o.block_size != 0 && o.block_size != 64 && o.block_size != 128 && o.block_size != 256) {
        set_error("spk_render_create: block_size %d not one of 64, 128, 256", o.block_size);
        return SPK_ERR_INVALID;
    }
    if (o.block_size == 0) o.block_size = 128;
    r->opts = o;

    ConfigDev& c = r->cfg;
    memset(&c, 0, sizeof(c));
    camera_dev(cam, c.cam);
    c.W = cam->width;
    c.H = cam->height;
    c.npix = (uint32_t)cam->width * (uint32_t)cam->height;
    // Regenerating passes (the default) cover up to kRegenPassPaths paths each with queue_size
    // paths in flight; classic passes hold all of their paths in the queues at once.
    const bool regen = o.classic_passes == 0;
    uint32_t qcap = regen ? (uint32_t)std::max<int64_t>(o.queue_size, kRegenPassPaths) : (uint32_t)o.queue_size;
    if (c.npix <= qcap) {
        // samples per pass; bounded so the per-path sample slot fits the queue flags
        // (bits 7..31) and the film kernel's per-pixel loop stays short.  A regenerating pass
        // holds at most the render's planned samples (later runs just take more passes; how
        // samples are cut into passes does not change the film), so small renders do not
        // hold path slots for 2^28 paths.
        c.kmax = std::min<uint32_t>(qcap / c.npix, 1u << 14);
        if (regen) c.kmax = std::min<uint32_t>(c.kmax, (uint32_t)std::max(1, desc->spp));
        c.chunks = 1;
        r->capacity = c.kmax * c.npix;
    } else {
        c.kmax = 1;
        c.chunks = (c.npix + qcap - 1) / qcap;
        r->capacity = qcap;
    }
    c.queue = c.chunks > 1 ? qcap : r->capacity;
    r->inflight = regen ? (uint32_t)std::min<int64_t>(o.queue_size, r->capacity) : r->capacity;
    c.regen = regen ? 1 : 0;
    c.inflight = r->inflight;
    c.max_depth = desc->max_depth;
    c.rr_depth = desc->rr_depth;
    c.sampler = desc->sampler;
    c.seed = desc->seed;
    c.integrator = desc->integrator;
    c.wide = s->dev.wide;

    // Buffers: queues hold the paths in flight, path radiance every path of a pass.
    int rc;
    size_t Q = r->inflight;
    for (int k = 0; k < 2; k++) {
        if ((rc = r->qa_o[k].alloc(Q * 16)) || (rc = r->qa_d[k].alloc(Q * 16)) || (rc = r->qa_b[k].alloc(Q * 16))) return rc;
    }
    if ((rc = r->hit.alloc(Q * 16)) || (rc = r->qs_o.alloc(Q * 16)) || (rc = r->qs_d.alloc(Q * 16)) ||
        (rc = r->qs_c.alloc(Q * 16)) || (rc = r->qm_o.alloc(Q * 16)) || (rc = r->qm_d.alloc(Q * 16)) ||
        (rc = r->qm_c.alloc(Q * 16)) || (rc = r->path_L.alloc((size_t)r->capacity * 16)))
        return rc;
    if (regen) {
        SPK_CHECK_HIP(hipHostMalloc((void**)&r->rg_flag, 2 * sizeof(uint32_t), hipHostMallocDefault));
        r->rg_flag[0] = r->rg_flag[1] = 0;
        for (hipEvent_t& e : r->rg_ev) SPK_CHECK_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    }
    size_t P = c.npix;
    if ((rc = r->film_sum.alloc(P * 16)) || (rc = r->film_c.alloc(P * 16)) || (rc = r->film_sq.alloc(P * 16)) ||
        (rc = r->film_sqc.alloc(P * 16)))
        return rc;
    hipDeviceProp_t prop;
    SPK_CHECK_HIP(hipGetDeviceProperties(&prop, r->device));
    r->shape.block = o.block_size;
    r->shape.grid_trace = prop.multiProcessorCount * std::max(1, trace_blocks_per_cu(o.block_size));
    r->shape.grid_stream = prop.multiProcessorCount * 8;
    r->shape.grid_guide = prop.multiProcessorCount;
    if ((rc = r->spill.alloc((size_t)r->shape.grid_trace * o.block_size * s->spill_cap * 4))) return rc;
    if ((rc = r->cfg_dev.alloc(sizeof(ConfigDev))) || (rc = r->ps_dev.alloc(sizeof(PassState)))) return rc;
    if ((rc = r->scratch.alloc(4096 * sizeof(double)))) return rc;
    SPK_CHECK_HIP(hipMemset(r->ps_dev.p, 0, sizeof(PassState)));
    SPK_CHECK_HIP(hipStreamCreateWithFlags(&r->stream, hipStreamNonBlocking));

    RenderBufs& b = r->bufs;
    for (int k = 0; k < 2; k++) {
        b.qa_o[k] = r->qa_o[k].as<float4>();
        b.qa_d[k] = r->qa_d[k].as<float4>();
        b.qa_b[k] = r->qa_b[k].as<float4>();
    }
    b.hit = r->hit.as<float4>();
    b.qs_o = r->qs_o.as<float4>();
    b.qs_d = r->qs_d.as<float4>();
    b.qs_c = r->qs_c.as<float4>();
    b.qm_o = r->qm_o.as<float4>();
    b.qm_d = r->qm_d.as<float4>();
    b.qm_c = r->qm_c.as<float4>();
    b.path_L = r->path_L.as<float4>();
    b.film_sum = r->film_sum.as<float4>();
    b.film_c = r->film_c.as<float4>();
    b.film_sq = r->film_sq.as<float4>();
    b.film_sqc = r->film_sqc.as<float4>();
    b.spill = r->spill.as<uint32_t>();
    b.spill_cap = s->spill_cap;
    b.cfg = r->cfg_dev.as<ConfigDev>();
    b.ps = r->ps_dev.as<PassState>();
    if ((rc = launch_clear_film(b, c.npix, r->stream))) return rc;
    SPK_CHECK_HIP(hipStreamSynchronize(r->stream));
    *out = r.release();
    return SPK_OK;
}

int spk_render_destroy(spk_render* r) {
    clear_error();
    delete r;
    return SPK_OK;
}

}  // extern "C"

namespace {

GuideNetDev guide_net(spk_render* r, bool record) {
    GuideNetDev n;
    memset(&n, 0, sizeof(n));
    spk_guide* g = r->guide;
    n.wf = g->wf.as<uint16_t>();
    n.bias = g->bias.as<float>();
    for (int k = 0; k < 3; k++) {
        float ext = r->scene->hi[k] - r->scene->lo[k];
        n.lo[k] = r->scene->lo[k];
        n.inv_ext[k] = ext > 0.0f ? 1.0f / ext : 0.0f;
    }
    n.frac = g->opts.guide_fraction;
    n.record = record ? 1 : 0;
    for (int k = 0; k < 5; k++) n.rec[k] = r->rec[k].as<float4>();
    n.rec_cap = r->rec_cap;
    n.rec_count = g->state.as<uint32_t>();      // GuideState::rec_count, its first word
    n.rec_fin = g->state.as<uint32_t>() + 1;    // GuideState::rec_fin, the second
    return n;
}

// The Adam steps after a training round: the round's records sorted into a fixed order, then
// steps_per_pass steps on batches drawn from them.
int guide_train_round(spk_render* r, GuideNetDev net) {
    spk_guide* g = r->guide;
    float4* in[5];
    float4* out[5];
    for (int q = 0; q < 5; q++) {
        in[q] = r->rec[q].as<float4>();
        out[q] = r->rec_s[q].as<float4>();
    }
    int rc;
    if ((rc = sort_guide_records(in, out, r->rec_cap, g->state.as<uint32_t>(), r->rec_keys.as<uint64_t>(),
                                 r->rec_idx.as<uint32_t>(), r->rec_tmp.p, r->rec_tmp.bytes, r->stream)))
        return rc;
    for (int q = 0; q < 5; q++) net.rec[q] = out[q];
    const TrainBufs tb = g->train_bufs(g->bcap);
    for (int st = 0; st < g->opts.steps_per_pass; st++) {
        uint32_t h0 = g->opts.seed, h1 = (uint32_t)g->steps, h2 = 0x6a09e667u, h3 = 0xbb67ae85u;
        pcg4d(h0, h1, h2, h3);
        if ((rc = launch_guide_step(g->bufs(), tb, &net, g->adam(), h0, 0u, true, nullptr, r->stream))) return rc;
        g->steps++;
    }
    return SPK_OK;
}

// Profile events of regenerating passes: 8 per iteration (closest hit, guide, shade, shadow).
int ensure_events(spk_render* r, size_t need) {
    while (r->events.size() < need) {
        hipEvent_t e;
        SPK_CHECK_HIP(hipEventCreate(&e));
        r->events.push_back(e);
    }
    return SPK_OK;
}

// One regenerating pass: set-up, iterations in groups of kRegenGroup until one finds no work,
// then the film.  The done flag is copied to pinned host memory after every group and read one
// group later, so the host never stalls the stream (at most one group of empty iterations runs
// at the end).  With use_graph one captured group (an even number of iterations, so every group
// has the same queue slots) is replayed.  Profile events are appended from index *ev_next.
int run_pass_regen(spk_render* r, const SceneDev& sd, bool batch, bool use_graph, const GuideNetDev* net,
                   size_t* ev_next) {
    int rc;
    if ((rc = launch_regen_begin(r->bufs, batch, r->stream))) return rc;
    const bool prof = r->opts.profile != 0 && ev_next;
    const bool count = r->opts.count_visits != 0;
    static_assert(kRegenGroup % 2 == 0, "a group keeps the slot pattern");
    // Every queued path advances one bounce per iteration and ends after maxDepth + 1, and an
    // iteration starts new paths whenever the queue has room, so a pass needs at most this many
    // iterations; beyond it something is wrong and the pass is abandoned with an error.
    const int64_t cap_paths = batch ? (int64_t)r->capacity : (int64_t)std::max<uint32_t>(r->capacity, 1);
    const int64_t max_it = (cap_paths / std::max<uint32_t>(r->inflight, 1) + 2) * (r->cfg.max_depth + 2) +
                           4 * kRegenGroup;
    int64_t it = 0;
    for (int g = 0;; g++) {
        if (it > max_it) {
            SPK_CHECK_HIP(hipStreamSynchronize(r->stream));
            set_error("regenerating pass did not finish after %lld iterations", (long long)it);
            return SPK_ERR_STATE;
        }
        if (use_graph && !prof && !net) {
            if (!r->rg_exec) {
                hipGraph_t gr;
                SPK_CHECK_HIP(hipStreamBeginCapture(r->stream, hipStreamCaptureModeThreadLocal));
                int crc = launch_regen_iters(sd, r->bufs, r->cfg, r->shape, r->stream, batch, 0, kRegenGroup, nullptr, 0,
                                             count, nullptr);
                hipError_t e = hipStreamEndCapture(r->stream, &gr);
                if (crc) return crc;
                SPK_CHECK_HIP(e);
                e = hipGraphInstantiate(&r->rg_exec, gr, nullptr, nullptr, 0);
                (void)hipGraphDestroy(gr);
                SPK_CHECK_HIP(e);
            }
            SPK_CHECK_HIP(hipGraphLaunch(r->rg_exec, r->stream));
        } else {
            hipEvent_t* ev = nullptr;
            if (prof) {
                if ((rc = ensure_events(r, *ev_next + 8 * kRegenGroup))) return rc;
                ev = &r->events[*ev_next];
                *ev_next += 8 * kRegenGroup;
            }
            if ((rc = launch_regen_iters(sd, r->bufs, r->cfg, r->shape, r->stream, batch, (int)it, kRegenGroup, ev,
                                         r->opts.profile >= 2 ? 8 : 2, count, net)))
                return rc;
        }
        it += kRegenGroup;
        SPK_CHECK_HIP(hipMemcpyAsync(&r->rg_flag[g & 1], &r->bufs.ps->rg_done, sizeof(uint32_t), hipMemcpyDeviceToHost,
                                     r->stream));
        SPK_CHECK_HIP(hipEventRecord(r->rg_ev[g & 1], r->stream));
        if (g >= 1) {
            SPK_CHECK_HIP(hipEventSynchronize(r->rg_ev[(g - 1) & 1]));
            if (r->rg_flag[(g - 1) & 1]) break;
        }
    }
    r->rg_iterations += it;
    return launch_regen_end(r->bufs, r->cfg, r->shape, r->stream, batch, (int)((it - 1) & 1), net);
}

// Adds the profile events of iterations [0, n_it) stored from index 0 (8 per iteration; the
// closest-hit pair only at profile level 1).
int regen_profile(spk_render* r, int64_t n_it, bool guided) {
    for (int64_t k = 0; k < n_it; k++) {
        float ms = 0.0f;
        hipEvent_t* e = &r->events[(size_t)k * 8];
        SPK_CHECK_HIP(hipEventElapsedTime(&ms, e[0], e[1]));
        r->extend_ms += ms;
        r->extend_launches++;
        if (r->opts.profile < 2) continue;
        if (guided) {
            SPK_CHECK_HIP(hipEventElapsedTime(&ms, e[2], e[3]));
            r->guide_ms += ms;
            r->guide_launches++;
        }
        SPK_CHECK_HIP(hipEventElapsedTime(&ms, e[4], e[5]));
        r->shade_ms += ms;
        r->shade_launches++;
        SPK_CHECK_HIP(hipEventElapsedTime(&ms, e[6], e[7]));
        r->shadow_ms += ms;
        r->shadow_launches++;
    }
    return SPK_OK;
}

// Samples [first, first + n_spp) as passes of the compiled pipeline.
int run_passes(spk_render* r, int64_t first, int32_t n_spp, const GuideNetDev* net) {
    ConfigDev& c = r->cfg;
    c.first_sample = (uint32_t)first;
    c.end_sample = (uint32_t)(first + n_spp);
    SPK_CHECK_HIP(hipMemcpyAsync(r->cfg_dev.p, &c, sizeof(ConfigDev), hipMemcpyHostToDevice, r->stream));
    int rc;
    if ((rc = launch_set_pass(r->bufs.ps, 0, r->stream))) return rc;
    int64_t passes = c.chunks > 1 ? (int64_t)n_spp * c.chunks : ((int64_t)n_spp + c.kmax - 1) / c.kmax;
    // per depth: start/end of the closest-hit, guide (guided renders), shade and shadow launches
    const int nd = c.max_depth + 1;
    int nev = 8 * nd;
    int64_t ext_before = 0, shadow_before = 0;
    if (r->opts.profile) {
        PassState ps;
        SPK_CHECK_HIP(hipStreamSynchronize(r->stream));
        SPK_CHECK_HIP(hipMemcpy(&ps, r->ps_dev.p, sizeof(PassState), hipMemcpyDeviceToHost));
        ext_before = (int64_t)ps.total_extend;
        shadow_before = (int64_t)ps.total_shadow;
        if (!c.regen && (rc = ensure_events(r, (size_t)passes * nev))) return rc;
    }
    if (c.regen) {
        size_t ev_next = 0;
        for (int64_t p = 0; p < passes; p++)
            if ((rc = run_pass_regen(r, r->scene->dev, false, r->opts.use_graph != 0, net, &ev_next))) return rc;
        if (r->opts.profile) {
            SPK_CHECK_HIP(hipStreamSynchronize(r->stream));
            PassState ps;
            SPK_CHECK_HIP(hipMemcpy(&ps, r->ps_dev.p, sizeof(PassState), hipMemcpyDeviceToHost));
            r->extend_rays_timed += (int64_t)ps.total_extend - ext_before;
            r->shadow_rays_timed += (int64_t)ps.total_shadow - shadow_before;
            if ((rc = regen_profile(r, (int64_t)(ev_next / 8), net != nullptr))) return rc;
        }
        return SPK_OK;
    }
    if (r->opts.use_graph && !r->opts.profile && !net) {
        // The pass reads its sample range from device memory, so one capture serves every pass.
        if (!r->pass_exec) {
            hipGraph_t g;
            SPK_CHECK_HIP(hipStreamBeginCapture(r->stream, hipStreamCaptureModeThreadLocal));
            int crc = launch_pass(r->scene->dev, r->bufs, c, r->shape, r->stream, nullptr, 0, r->opts.count_visits != 0,
                                  nullptr);
            hipError_t e = hipStreamEndCapture(r->stream, &g);
            if (crc) return crc;
            SPK_CHECK_HIP(e);
            e = hipGraphInstantiate(&r->pass_exec, g, nullptr, nullptr, 0);
            (void)hipGraphDestroy(g);
            SPK_CHECK_HIP(e);
        }
        for (int64_t p = 0; p < passes; p++) SPK_CHECK_HIP(hipGraphLaunch(r->pass_exec, r->stream));
        return SPK_OK;
    }
    for (int64_t p = 0; p < passes; p++) {
        hipEvent_t* ev = r->opts.profile ? &r->events[(size_t)p * nev] : nullptr;
        // level 1 records the closest-hit pairs only (indices below 2 nd)
        const int n_ev = !ev ? 0 : r->opts.profile >= 2 ? nev : 2 * nd;
        if ((rc = launch_pass(r->scene->dev, r->bufs, c, r->shape, r->stream, ev, n_ev, r->opts.count_visits != 0,
                              net)))
            return rc;
    }
    if (r->opts.profile) {
        SPK_CHECK_HIP(hipStreamSynchronize(r->stream));
        PassState ps;
        SPK_CHECK_HIP(hipMemcpy(&ps, r->ps_dev.p, sizeof(PassState), hipMemcpyDeviceToHost));
        r->extend_rays_timed += (int64_t)ps.total_extend - ext_before;
        r->shadow_rays_timed += (int64_t)ps.total_shadow - shadow_before;
        for (int64_t p = 0; p < passes; p++)
            for (int d = 0; d <= c.max_depth; d++) {
                float ms = 0.0f;
                SPK_CHECK_HIP(hipEventElapsedTime(&ms, r->events[(size_t)p * nev + 2 * d], r->events[(size_t)p * nev + 2 * d + 1]));
                r->extend_ms += ms;
                r->extend_launches++;
                if (r->opts.profile < 2) continue;
                if (net && d < c.max_depth) {
                    const size_t g = (size_t)p * nev + 2 * (nd + d);
                    SPK_CHECK_HIP(hipEventElapsedTime(&ms, r->events[g], r->events[g + 1]));
                    r->guide_ms += ms;
                    r->guide_launches++;
                }
                const size_t sh = (size_t)p * nev + 4 * nd + 2 * d;
                SPK_CHECK_HIP(hipEventElapsedTime(&ms, r->events[sh], r->events[sh + 1]));
                r->shade_ms += ms;
                r->shade_launches++;
                if (d < c.max_depth) {
                    const size_t so = (size_t)p * nev + 6 * nd + 2 * d;
                    SPK_CHECK_HIP(hipEventElapsedTime(&ms, r->events[so], r->events[so + 1]));
                    r->shadow_ms += ms;
                    r->shadow_launches++;
                }
            }
    }
    return SPK_OK;
}

}  // namespace

extern "C" {

int spk_render_run(spk_render* r, int32_t n_spp, int32_t sync) {
    clear_error();
    if (!r || n_spp < 0) {
        set_error("spk_render_run: bad arguments");
        return SPK_ERR_INVALID;
    }
    if (n_spp == 0) return SPK_OK;
    SPK_CHECK_HIP(hipSetDevice(r->device));
    int64_t first = (int64_t)r->desc.first_sample + r->samples_done;
    if (first + n_spp >= (int64_t(1) << 31)) {
        set_error("spk_render_run: sample index overflow");
        return SPK_ERR_INVALID;
    }
    int rc;
    if (!r->guide) {
        if ((rc = run_passes(r, first, n_spp, nullptr))) return rc;
        r->samples_done += n_spp;
    } else {
        // Online training: rounds of train_pass_spp samples per pixel, each keeping a
        // 1/train_pass_spp subsample of its vertices as records and followed by Adam steps
        // on them; then the remaining samples with the network fixed.
        spk_guide* g = r->guide;
        int32_t left = n_spp;
        while (left > 0) {
            const bool train = r->guide_trained < g->opts.train_spp && g->opts.steps_per_pass > 0;
            const int32_t k = train ? (int32_t)std::min<int64_t>({(int64_t)left, (int64_t)g->opts.train_pass_spp,
                                                                  g->opts.train_spp - r->guide_trained})
                                    : left;
            // rec_count and rec_fin are the first two words of the guide state
            if (train && (rc = zero_u32(g->state.as<uint32_t>(), 2, r->stream))) return rc;
            GuideNetDev net = guide_net(r, train);
            net.rec_keep = 1.0f / (float)k;
            if ((rc = run_passes(r, first, k, &net))) return rc;
            if (train) {
                if ((rc = guide_train_round(r, net))) return rc;
                r->guide_trained += k;
                r->train_passes++;
            }
            first += k;
            left -= k;
            r->samples_done += k;
        }
    }
    if (sync) SPK_CHECK_HIP(hipStreamSynchronize(r->stream));
    return SPK_OK;
}

int spk_render_set_guide(spk_render* r, spk_guide* g) {
    clear_error();
    if (!r) return SPK_ERR_INVALID;
    SPK_CHECK_HIP(hipSetDevice(r->device));
    SPK_CHECK_HIP(hipStreamSynchronize(r->stream));
    if (g && g->device != r->device) {
        set_error("spk_render_set_guide: guide and render on different devices");
        return SPK_ERR_INVALID;
    }
    if (g && r->cfg.integrator != SPK_INTEGRATOR_PATH) {
        set_error("spk_render_set_guide: the learned proposal samples path continuations; directlighting has none");
        return SPK_ERR_INVALID;
    }
    if (g) SPK_CHECK_HIP(hipStreamSynchronize(g->stream));
    r->guide = g;
    r->guide_trained = 0;
    // read by the shading kernels of guided runs (batches launch without re-uploading the config)
    r->cfg.guide_min_alpha = g ? g->opts.glossy_min_alpha : 0.0f;
    r->cfg.guide_depth = g ? g->opts.guided_depth : 0;
    SPK_CHECK_HIP(hipMemcpy(r->cfg_dev.p, &r->cfg, sizeof(ConfigDev), hipMemcpyHostToDevice));
    if (!g || r->qg[0].p) return SPK_OK;
    int rc;
    const size_t Q = r->inflight;
    for (int k = 0; k < 4; k++)
        if ((rc = r->qg[k].alloc(Q * 16))) return rc;
    // training records of one round: about one kept vertex per pixel, room for four, up to 2^25
    // (160 B each with the sorted copy: 5.4 GB at the cap, which an 8-view 1080p batch reaches)
    r->rec_cap = (uint32_t)std::min<uint64_t>((uint64_t)r->cfg.npix * 4u, uint64_t(1) << 25);
    for (int k = 0; k < 5; k++)
        if ((rc = r->rec[k].alloc((size_t)r->rec_cap * 16)) || (rc = r->rec_s[k].alloc((size_t)r->rec_cap * 16))) return rc;
    size_t sort_bytes = 0;
    if ((rc = guide_record_sort_bytes(r->rec_cap, &sort_bytes))) return rc;
    if ((rc = r->rec_keys.alloc((size_t)r->rec_cap * 16)) || (rc = r->rec_idx.alloc((size_t)r->rec_cap * 8)) ||
        (rc = r->rec_tmp.alloc(std::max<size_t>(sort_bytes, 16))))
        return rc;
    r->bufs.qg[0] = r->qg[0].as<float4>();
    r->bufs.qg[1] = r->qg[1].as<float4>();
    r->bufs.qg[2] = r->qg[2].as<float4>();
    r->bufs.qg[3] = r->qg[3].as<float4>();
    return SPK_OK;
}

int spk_render_sync(spk_render* r) {
    clear_error();
    if (!r) return SPK_ERR_INVALID;
    SPK_CHECK_HIP(hipSetDevice(r->device));
    SPK_CHECK_HIP(hipStreamSynchronize(r->stream));
    return SPK_OK;
}

int spk_render_film(spk_render* r, double* sum_rgb, double* sumsq_rgb, double* count) {
    clear_error();
    if (!r) return SPK_ERR_INVALID;
    SPK_CHECK_HIP(hipSetDevice(r->device));
    SPK_CHECK_HIP(hipStreamSynchronize(r->stream));
    size_t P = r->cfg.npix;
    std::vector<float4> s(P), c(P), q(P), qc(P);
    SPK_CHECK_HIP(hipMemcpy(s.data(), r->film_sum.p, P * 16, hipMemcpyDeviceToHost));
    SPK_CHECK_HIP(hipMemcpy(c.data(), r->film_c.p, P * 16, hipMemcpyDeviceToHost));
    SPK_CHECK_HIP(hipMemcpy(q.data(), r->film_sq.p, P * 16, hipMemcpyDeviceToHost));
    SPK_CHECK_HIP(hipMemcpy(qc.data(), r->film_sqc.p, P * 16, hipMemcpyDeviceToHost));
    for (size_t i = 0; i < P; i++) {
        // The Kahan compensation holds the negated lost low-order part.
        if (sum_rgb) {
            sum_rgb[3 * i] = (double)s[i].x - (double)c[i].x;
            sum_rgb[3 * i + 1] = (double)s[i].y - (double)c[i].y;
            sum_rgb[3 * i + 2] = (double)s[i].z - (double)c[i].z;
        }
        if (sumsq_rgb) {
            sumsq_rgb[3 * i] = (double)q[i].x - (double)qc[i].x;
            sumsq_rgb[3 * i + 1] = (double)q[i].y - (double)qc[i].y;
            sumsq_rgb[3 * i + 2] = (double)q[i].z - (double)qc[i].z;
        }
        if (count) count[i] = (double)s[i].w;
    }
    return SPK_OK;
}

int spk_render_image(spk_render* r, float* rgb) {
    clear_error();
    if (!r || !rgb) return SPK_ERR_INVALID;
    size_t P = r->cfg.npix;
    std::vector<double> sum(P * 3), cnt(P);
    int rc = spk_render_film(r, sum.data(), nullptr, cnt.data());
    if (rc) return rc;
    for (size_t i = 0; i < P; i++)
        for (int k = 0; k < 3; k++) rgb[3 * i + k] = (float)(sum[3 * i + k] / (cnt[i] + 1e-12));
    return SPK_OK;
}

int spk_render_relmse(spk_render* r, const double* ref_rgb, double eps, double* out) {
    clear_error();
    if (!r || !ref_rgb || !out) return SPK_ERR_INVALID;
    SPK_CHECK_HIP(hipSetDevice(r->device));
    size_t P = r->cfg.npix;
    DevBuf ref;
    int rc;
    if ((rc = ref.upload(ref_rgb, P * 3 * sizeof(double)))) return rc;
    const int nblocks = 1024;
    if ((rc = launch_relmse(r->bufs, r->cfg, ref.as<double>(), eps, r->scratch.as<double>(), nblocks, r->stream))) {
        ref.release();
        return rc;
    }
    double total = 0.0;
    SPK_CHECK_HIP(hipMemcpyAsync(&total, r->scratch.as<double>() + nblocks, sizeof(double), hipMemcpyDeviceToHost, r->stream));
    SPK_CHECK_HIP(hipStreamSynchronize(r->stream));
    ref.release();
    *out = total / (double)(P * 3);
    return SPK_OK;
}

int spk_render_relmse_pixels(spk_render* r, int64_t n, const int32_t* pixels, const double* ref_rgb, double eps,
                             double* out) {
    clear_error();
    if (!r || n <= 0 || !pixels || !ref_rgb || !out) return SPK_ERR_INVALID;
    for (int64_t i = 0; i < n; i++)
        if (pixels[i] < 0 || (uint32_t)pixels[i] >= r->cfg.npix) {
            set_error("spk_render_relmse_pixels: pixel %d outside the film", pixels[i]);
            return SPK_ERR_INVALID;
        }
    SPK_CHECK_HIP(hipSetDevice(r->device));
    DevBuf px, ref;
    int rc;
    if ((rc = px.upload(pixels, (size_t)n * 4)) || (rc = ref.upload(ref_rgb, (size_t)n * 3 * sizeof(double)))) return rc;
    const int nblocks = 1024;
    if ((rc = launch_relmse_pixels(r->bufs, n, px.as<int32_t>(), ref.as<double>(), eps, r->scratch.as<double>(), nblocks,
                                   r->stream)))
        return rc;
    double total = 0.0;
    SPK_CHECK_HIP(hipMemcpyAsync(&total, r->scratch.as<double>() + nblocks, sizeof(double), hipMemcpyDeviceToHost, r->stream));
    SPK_CHECK_HIP(hipStreamSynchronize(r->stream));
    *out = total / (double)(n * 3);
    return SPK_OK;
}

int spk_render_stats_get(spk_render* r, spk_render_stats* out) {
    clear_error();
    if (!r || !out) return SPK_ERR_INVALID;
    SPK_CHECK_HIP(hipSetDevice(r->device));
    SPK_CHECK_HIP(hipStreamSynchronize(r->stream));
    PassState ps;
    SPK_CHECK_HIP(hipMemcpy(&ps, r->ps_dev.p, sizeof(PassState), hipMemcpyDeviceToHost));
    memset(out, 0, sizeof(*out));
    out->samples_done = r->samples_done;
    out->paths = (int64_t)ps.total_paths;
    out->rays_extend = (int64_t)ps.total_extend;
    out->rays_shadow = (int64_t)ps.total_shadow;
    out->rays_mis = (int64_t)ps.total_mis;
    out->passes = (int64_t)ps.total_samples;
    out->extend_ms = r->extend_ms;
    out->extend_launches = r->extend_launches;
    out->extend_rays_timed = r->extend_rays_timed;
    out->count_nodes = (int64_t)ps.count_nodes;
    out->count_tris = (int64_t)ps.count_tris;
    out->count_rays = (int64_t)ps.count_rays;
    out->guided_vertices = (int64_t)ps.total_guided;
    out->guide_ms = r->guide_ms;
    out->guide_launches = r->guide_launches;
    out->shade_ms = r->shade_ms;
    out->shade_launches = r->shade_launches;
    out->shadow_ms = r->shadow_ms;
    out->shadow_launches = r->shadow_launches;
    out->shadow_rays_timed = r->shadow_rays_timed;
    out->count_nodes_shadow = (int64_t)ps.count_nodes_s;
    out->count_tris_shadow = (int64_t)ps.count_tris_s;
    out->count_rays_shadow = (int64_t)ps.count_rays_s;
    out->train_passes = r->train_passes;
    return SPK_OK;
}

int spk_sampler_u32(int32_t kind, uint32_t seed, int64_t n, const uint32_t* pixel, const uint32_t* sample,
                    const uint32_t* dim, uint32_t* out, int32_t on_device) {
    clear_error();
    if (n < 0 || (n > 0 && (!pixel || !sample || !dim || !out)) || (kind != SAMPLER_INDEPENDENT && kind != SAMPLER_SOBOL)) {
        set_error("spk_sampler_u32: bad arguments");
        return SPK_ERR_INVALID;
    }
    if (!on_device) {
        for (int64_t i = 0; i < n; i++) out[i] = sample_u32(kind, seed, pixel[i], sample[i], dim[i]);
        return SPK_OK;
    }
    if (n == 0) return SPK_OK;
    SPK_CHECK_HIP(hipSetDevice(g_device));
    DevBuf a, b, c, o;
    int rc;
    if ((rc = a.upload(pixel, n * 4)) || (rc = b.upload(sample, n * 4)) || (rc = c.upload(dim, n * 4)) || (rc = o.alloc(n * 4)))
        return rc;
    if ((rc = launch_sampler(kind, seed, n, a.as<uint32_t>(), b.as<uint32_t>(), c.as<uint32_t>(), o.as<uint32_t>(), nullptr)))
        return rc;
    SPK_CHECK_HIP(hipDeviceSynchronize());
    SPK_CHECK_HIP(hipMemcpy(out, o.p, n * 4, hipMemcpyDeviceToHost));
    a.release();
    b.release();
    c.release();
    o.release();
    return SPK_OK;
}

// ---- batched rendering ---------------------------------------------------------------------

}  // extern "C"

struct spk_batch {
    spk_render* r = nullptr;          // queues, pass state, stream; its film holds every image
    std::vector<spk_scene*> scenes;
    std::vector<ImageDev> images;
    std::vector<ChunkDev> chunks;
    std::vector<PassDev> passes;
    DevBuf scenes_dev, images_dev, chunks_dev, passes_dev;
    int64_t paths = 0;
    int64_t film_pixels = 0;
    int graph = 0;
    hipGraphExec_t exec = nullptr;    // one pass (graph = 1) or the whole batch (graph = 2)
    // learned proposal (one network for every image of one scene)
    std::vector<int32_t> pass_round;  // training round of each pass, -1 after training
    std::vector<uint32_t> spp;        // samples per image
    ~spk_batch() {
        if (r) (void)hipSetDevice(r->device);
        if (exec) (void)hipGraphExecDestroy(exec);
        if (r) spk_render_destroy(r);
    }
};

namespace {

// Work of image `img` as chunks of whole-frame sample groups, or pixel ranges at one
// sample each when the frame exceeds the queue.
void batch_chunks(uint32_t img, uint32_t npix, uint32_t spp, uint32_t cap, std::vector<ChunkDev>& out) {
    if (npix <= cap) {
        const uint32_t k = std::min<uint32_t>({spp, cap / npix, 1u << 14});
        for (uint32_t s0 = 0; s0 < spp; s0 += k) {
            ChunkDev c = {img, 0u, npix, s0, std::min(k, spp - s0), 0u, 0u, 0u};
            out.push_back(c);
        }
    } else {
        for (uint32_t s0 = 0; s0 < spp; s0++)
            for (uint32_t p0 = 0; p0 < npix; p0 += cap) {
                ChunkDev c = {img, p0, std::min(cap, npix - p0), s0, 1u, 0u, 0u, 0u};
                out.push_back(c);
            }
    }
}

// Guided batches: the first `train_spp` samples of every image as training rounds of
// `round_spp` samples, round-major (round r holds samples [r round_spp, (r + 1) round_spp) of
// every image, so the network learns from every view), then the remaining samples; rounds[c]
// is chunk c's round or -1.
void batch_chunks_guided(const std::vector<uint32_t>& npix, const std::vector<uint32_t>& spp, uint32_t cap,
                         uint32_t train_spp, uint32_t round_spp, std::vector<ChunkDev>& out,
                         std::vector<int32_t>& rounds) {
    uint32_t tmax = 0;
    for (size_t i = 0; i < spp.size(); i++) tmax = std::max(tmax, std::min(train_spp, spp[i]));
    const uint32_t nr = (tmax + round_spp - 1) / round_spp;
    for (uint32_t rd = 0; rd < nr; rd++)
        for (size_t i = 0; i < spp.size(); i++) {
            const uint32_t s0 = rd * round_spp, s1 = std::min({(rd + 1) * round_spp, train_spp, spp[i]});
            if (s0 >= s1) continue;
            std::vector<ChunkDev> part;
            batch_chunks((uint32_t)i, npix[i], s1 - s0, cap, part);
            for (ChunkDev c : part) {
                c.sample_base += s0;
                out.push_back(c);
                rounds.push_back((int32_t)rd);
            }
        }
    for (size_t i = 0; i < spp.size(); i++) {
        const uint32_t s0 = std::min(train_spp, spp[i]);
        if (s0 >= spp[i]) continue;
        std::vector<ChunkDev> part;
        batch_chunks((uint32_t)i, npix[i], spp[i] - s0, cap, part);
        for (ChunkDev c : part) {
            c.sample_base += s0;
            out.push_back(c);
            rounds.push_back(-1);
        }
    }
}

// Packs chunks in order into passes of at most `cap` paths, at most one chunk per image
// per pass (two chunks of one image in a pass could touch the same film pixel), and, when
// `rounds` is given, chunks of one training round only.
void batch_passes(std::vector<ChunkDev>& ch, uint32_t cap, std::vector<PassDev>& out,
                  const std::vector<int32_t>* rounds = nullptr, std::vector<int32_t>* pass_round = nullptr) {
    size_t i = 0;
    while (i < ch.size()) {
        PassDev p = {(uint32_t)i, (uint32_t)i, 0u, 0u};
        std::vector<uint32_t> imgs;
        const int32_t rd = rounds ? (*rounds)[i] : -1;
        while (i < ch.size()) {
            ChunkDev& c = ch[i];
            const uint32_t paths = c.n_pix * c.k;
            if (p.n_paths + paths > cap) break;
            if (std::find(imgs.begin(), imgs.end(), c.img) != imgs.end()) break;
            if (rounds && (*rounds)[i] != rd) break;
            c.path_off = p.n_paths;
            c.pix_off = p.n_pixels;
            p.n_paths += paths;
            p.n_pixels += c.n_pix;
            imgs.push_back(c.img);
            i++;
        }
        p.chunk_end = (uint32_t)i;
        out.push_back(p);
        if (pass_round) pass_round->push_back(rd);
    }
}

int batch_launch_all(spk_batch* b) {
    spk_render* r = b->r;
    int rc;
    if ((rc = launch_clear_film(r->bufs, (uint32_t)b->film_pixels, r->stream))) return rc;
    if ((rc = launch_set_pass(r->bufs.ps, 0, r->stream))) return rc;
    for (size_t p = 0; p < b->passes.size(); p++)
        if ((rc = launch_batch_pass(r->bufs, r->cfg, r->shape, r->stream))) return rc;
    return SPK_OK;
}

}  // namespace

extern "C" {

int spk_batch_create(const spk_batch_desc* d, const spk_render_opts* opts, spk_batch** out) {
    clear_error();
    if (!d || !out || d->num_images <= 0 || !d->images) {
        set_error("spk_batch_create: no images");
        return SPK_ERR_INVALID;
    }
    *out = nullptr;
    if ((uint32_t)d->num_images > kBatchMaxImages) {
        set_error("spk_batch_create: %d images, at most %u", d->num_images, kBatchMaxImages);
        return SPK_ERR_INVALID;
    }
    if (d->graph < 0 || d->graph > 2) {
        set_error("spk_batch_create: graph mode %d not 0, 1 or 2", d->graph);
        return SPK_ERR_INVALID;
    }
    std::unique_ptr<spk_batch> b(new spk_batch());
    b->graph = d->graph;
    int64_t film = 0, spill_cap = 0;
    const int device = d->images[0].scene ? d->images[0].scene->device : 0;
    for (int i = 0; i < d->num_images; i++) {
        const spk_batch_image& im = d->images[i];
        const spk_camera_desc& c = im.camera;
        if (!im.scene || im.scene->device != device) {
            set_error("spk_batch_create: image %d has no scene or a scene on another device", i);
            return SPK_ERR_INVALID;
        }
        if (c.width <= 0 || c.height <= 0 || (int64_t)c.width * c.height > (int64_t(1) << kBatchPixelBits)) {
            set_error("spk_batch_create: image %d film %dx%d outside 1 .. 2^%u pixels", i, c.width, c.height,
                      kBatchPixelBits);
            return SPK_ERR_INVALID;
        }
        if (c.to_world[12] != 0.0 || c.to_world[13] != 0.0 || c.to_world[14] != 0.0 || c.to_world[15] != 1.0) {
            set_error("spk_batch_create: image %d camera toWorld must be affine", i);
            return SPK_ERR_INVALID;
        }
        if (im.spp <= 0 || im.spp >= (1 << 25)) {
            set_error("spk_batch_create: image %d spp %d outside [1, 2^25)", i, im.spp);
            return SPK_ERR_INVALID;
        }
        ImageDev id;
        memset(&id, 0, sizeof(id));
        camera_dev(&c, id.cam);
        id.seed = im.seed;
        id.film_off = (uint32_t)film;
        film += (int64_t)c.width * c.height;
        if (film >= (int64_t(1) << 31)) {
            set_error("spk_batch_create: films exceed 2^31 pixels");
            return SPK_ERR_INVALID;
        }
        b->images.push_back(id);
        b->scenes.push_back(im.scene);
        b->spp.push_back((uint32_t)im.spp);
        spill_cap = std::max<int64_t>(spill_cap, im.scene->spill_cap);
    }
    b->film_pixels = film;
    // The queues, pass state and film come from a render of a 1-row "film" holding every
    // image's pixels; its camera is never used by the batch kernels.
    spk_camera_desc fc;
    memset(&fc, 0, sizeof(fc));
    fc.width = (int32_t)film;
    fc.height = 1;
    fc.fov_deg = 45.0;
    fc.to_world[0] = fc.to_world[5] = fc.to_world[10] = fc.to_world[15] = 1.0;
    spk_render_desc rd;
    memset(&rd, 0, sizeof(rd));
    // the planned samples bound a regenerating pass: the largest image's, so passes hold as
    // many whole samples of the film as fit
    rd.spp = 1;
    for (int i = 0; i < d->num_images; i++) rd.spp = std::max(rd.spp, d->images[i].spp);
    rd.max_depth = d->max_depth;
    rd.rr_depth = d->rr_depth;
    rd.sampler = d->sampler;
    int rc;
    spk_scene* s0 = b->scenes[0];
    const int s0_cap = s0->spill_cap;
    s0->spill_cap = (int)spill_cap;  // size the traversal spill for the deepest BVH
    rc = spk_render_create(s0, &fc, &rd, opts, &b->r);
    s0->spill_cap = s0_cap;
    if (rc) return rc;
    spk_render* r = b->r;
    for (int i = 0; i < d->num_images; i++) r->cfg.wide |= b->scenes[i]->dev.wide;
    const uint32_t cap = r->capacity;
    for (int i = 0; i < d->num_images; i++) {
        const spk_camera_desc& c = d->images[i].camera;
        batch_chunks((uint32_t)i, (uint32_t)(c.width * c.height), (uint32_t)d->images[i].spp, cap, b->chunks);
        b->paths += (int64_t)c.width * c.height * d->images[i].spp;
    }
    batch_passes(b->chunks, cap, b->passes);
    std::vector<SceneDev> sd;
    for (spk_scene* sc : b->scenes) sd.push_back(sc->dev);
    if ((rc = b->scenes_dev.upload(sd.data(), sd.size() * sizeof(SceneDev))) ||
        (rc = b->images_dev.upload(b->images.data(), b->images.size() * sizeof(ImageDev))) ||
        (rc = b->chunks_dev.upload(b->chunks.data(), b->chunks.size() * sizeof(ChunkDev))) ||
        (rc = b->passes_dev.upload(b->passes.data(), b->passes.size() * sizeof(PassDev))))
        return rc;
    r->bufs.scenes = b->scenes_dev.as<SceneDev>();
    r->bufs.images = b->images_dev.as<ImageDev>();
    r->bufs.chunks = b->chunks_dev.as<ChunkDev>();
    r->bufs.passes = b->passes_dev.as<PassDev>();
    r->bufs.num_passes = (uint32_t)b->passes.size();
    r->bufs.spill_cap = (int)spill_cap;
    SPK_CHECK_HIP(hipMemcpy(r->cfg_dev.p, &r->cfg, sizeof(ConfigDev), hipMemcpyHostToDevice));
    if (b->graph && !r->cfg.regen) {
        // Captured work: kernel launches only (no memset nodes).
        hipGraph_t g;
        SPK_CHECK_HIP(hipStreamBeginCapture(r->stream, hipStreamCaptureModeThreadLocal));
        int crc = b->graph == 2 ? batch_launch_all(b.get()) : launch_batch_pass(r->bufs, r->cfg, r->shape, r->stream);
        hipError_t e = hipStreamEndCapture(r->stream, &g);
        if (crc) return crc;
        SPK_CHECK_HIP(e);
        e = hipGraphInstantiate(&b->exec, g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        SPK_CHECK_HIP(e);
    }
    *out = b.release();
    return SPK_OK;
}

int spk_batch_destroy(spk_batch* b) {
    clear_error();
    delete b;
    return SPK_OK;
}

int spk_batch_set_guide(spk_batch* b, spk_guide* g) {
    clear_error();
    if (!b) return SPK_ERR_INVALID;
    spk_render* r = b->r;
    SPK_CHECK_HIP(hipSetDevice(r->device));
    if (g) {
        for (spk_scene* sc : b->scenes)
            if (sc != b->scenes[0]) {
                set_error("spk_batch_set_guide: a guide serves the images of one scene (image scenes differ)");
                return SPK_ERR_INVALID;
            }
    }
    int rc;
    if ((rc = spk_render_set_guide(r, g))) return rc;
    // re-cut the work: training rounds first (or back to image order without a guide)
    std::vector<uint32_t> npix;
    for (const ImageDev& im : b->images) npix.push_back((uint32_t)im.cam.W * (uint32_t)im.cam.H);
    b->chunks.clear();
    b->passes.clear();
    b->pass_round.clear();
    if (g) {
        std::vector<int32_t> rounds;
        batch_chunks_guided(npix, b->spp, r->capacity, (uint32_t)std::max(0, g->opts.train_spp),
                            (uint32_t)g->opts.train_pass_spp, b->chunks, rounds);
        batch_passes(b->chunks, r->capacity, b->passes, &rounds, &b->pass_round);
    } else {
        for (size_t i = 0; i < b->images.size(); i++) batch_chunks((uint32_t)i, npix[i], b->spp[i], r->capacity, b->chunks);
        batch_passes(b->chunks, r->capacity, b->passes);
    }
    if ((rc = b->chunks_dev.upload(b->chunks.data(), b->chunks.size() * sizeof(ChunkDev))) ||
        (rc = b->passes_dev.upload(b->passes.data(), b->passes.size() * sizeof(PassDev))))
        return rc;
    r->bufs.chunks = b->chunks_dev.as<ChunkDev>();
    r->bufs.passes = b->passes_dev.as<PassDev>();
    r->bufs.num_passes = (uint32_t)b->passes.size();
    if (b->exec) {  // captured with the old tables
        (void)hipGraphExecDestroy(b->exec);
        b->exec = nullptr;
    }
    return SPK_OK;
}

int spk_batch_run(spk_batch* b, int32_t sync) {
    clear_error();
    if (!b) return SPK_ERR_INVALID;
    spk_render* r = b->r;
    SPK_CHECK_HIP(hipSetDevice(r->device));
    int rc;
    if (r->guide) {
        // Guided: passes launched one by one (no graph); a training round's records are turned
        // into Adam steps after its last pass, and the network stays fixed after the rounds.
        spk_guide* g = r->guide;
        if ((rc = launch_clear_film(r->bufs, (uint32_t)b->film_pixels, r->stream))) return rc;
        if ((rc = launch_set_pass(r->bufs.ps, 0, r->stream))) return rc;
        const SceneDev* sd = &b->scenes[0]->dev;
        const size_t np = b->passes.size();
        for (size_t p = 0; p < np; p++) {
            const int32_t rd = b->pass_round[p];
            const bool train = rd >= 0 && g->opts.steps_per_pass > 0;
            if (train && (p == 0 || b->pass_round[p - 1] != rd) &&
                (rc = zero_u32(g->state.as<uint32_t>(), 2, r->stream)))  // rec_count, rec_fin
                return rc;
            GuideNetDev net = guide_net(r, train);
            net.rec_keep = 1.0f / (float)g->opts.train_pass_spp;
            if (r->cfg.regen) rc = run_pass_regen(r, *sd, true, false, &net, nullptr);
            else rc = launch_batch_pass(r->bufs, r->cfg, r->shape, r->stream, sd, &net);
            if (rc) return rc;
            if (train && (p + 1 == np || b->pass_round[p + 1] != rd)) {
                if ((rc = guide_train_round(r, net))) return rc;
                r->train_passes++;
            }
        }
        r->samples_done++;
        if (sync) SPK_CHECK_HIP(hipStreamSynchronize(r->stream));
        return SPK_OK;
    }
    if (r->cfg.regen) {
        // regenerating passes, each driven to completion; graph modes replay captured groups of
        // iterations (a whole-batch graph would need the iteration counts in advance)
        SceneDev none;
        memset(&none, 0, sizeof(none));
        if ((rc = launch_clear_film(r->bufs, (uint32_t)b->film_pixels, r->stream))) return rc;
        if ((rc = launch_set_pass(r->bufs.ps, 0, r->stream))) return rc;
        for (size_t p = 0; p < b->passes.size(); p++)
            if ((rc = run_pass_regen(r, none, true, b->graph != 0, nullptr, nullptr))) return rc;
        r->samples_done++;
        if (sync) SPK_CHECK_HIP(hipStreamSynchronize(r->stream));
        return SPK_OK;
    }
    if (b->graph && !b->exec) {
        hipGraph_t gr;
        SPK_CHECK_HIP(hipStreamBeginCapture(r->stream, hipStreamCaptureModeThreadLocal));
        int crc = b->graph == 2 ? batch_launch_all(b) : launch_batch_pass(r->bufs, r->cfg, r->shape, r->stream);
        hipError_t e = hipStreamEndCapture(r->stream, &gr);
        if (crc) return crc;
        SPK_CHECK_HIP(e);
        e = hipGraphInstantiate(&b->exec, gr, nullptr, nullptr, 0);
        (void)hipGraphDestroy(gr);
        SPK_CHECK_HIP(e);
    }
    if (b->graph == 2) {
        SPK_CHECK_HIP(hipGraphLaunch(b->exec, r->stream));
    } else if (b->graph == 1) {
        if ((rc = launch_clear_film(r->bufs, (uint32_t)b->film_pixels, r->stream))) return rc;
        if ((rc = launch_set_pass(r->bufs.ps, 0, r->stream))) return rc;
        for (size_t p = 0; p < b->passes.size(); p++) SPK_CHECK_HIP(hipGraphLaunch(b->exec, r->stream));
    } else if ((rc = batch_launch_all(b))) {
        return rc;
    }
    r->samples_done++;
    if (sync) SPK_CHECK_HIP(hipStreamSynchronize(r->stream));
    return SPK_OK;
}

int spk_batch_film(spk_batch* b, int32_t image, double* sum_rgb, double* sumsq_rgb, double* count) {
    clear_error();
    if (!b || image < 0 || image >= (int32_t)b->images.size()) return SPK_ERR_INVALID;
    spk_render* r = b->r;
    SPK_CHECK_HIP(hipSetDevice(r->device));
    SPK_CHECK_HIP(hipStreamSynchronize(r->stream));
    const ImageDev& im = b->images[image];
    const size_t P = (size_t)im.cam.W * im.cam.H, off = im.film_off;
    std::vector<float4> s(P), c(P), q(P), qc(P);
    SPK_CHECK_HIP(hipMemcpy(s.data(), r->film_sum.as<float4>() + off, P * 16, hipMemcpyDeviceToHost));
    SPK_CHECK_HIP(hipMemcpy(c.data(), r->film_c.as<float4>() + off, P * 16, hipMemcpyDeviceToHost));
    SPK_CHECK_HIP(hipMemcpy(q.data(), r->film_sq.as<float4>() + off, P * 16, hipMemcpyDeviceToHost));
    SPK_CHECK_HIP(hipMemcpy(qc.data(), r->film_sqc.as<float4>() + off, P * 16, hipMemcpyDeviceToHost));
    for (size_t i = 0; i < P; i++) {
        if (sum_rgb) {
            sum_rgb[3 * i] = (double)s[i].x - (double)c[i].x;
            sum_rgb[3 * i + 1] = (double)s[i].y - (double)c[i].y;
            sum_rgb[3 * i + 2] = (double)s[i].z - (double)c[i].z;
        }
        if (sumsq_rgb) {
            sumsq_rgb[3 * i] = (double)q[i].x - (double)qc[i].x;
            sumsq_rgb[3 * i + 1] = (double)q[i].y - (double)qc[i].y;
            sumsq_rgb[3 * i + 2] = (double)q[i].z - (double)qc[i].z;
        }
        if (count) count[i] = (double)s[i].w;
    }
    return SPK_OK;
}

int spk_batch_info_get(spk_batch* b, spk_batch_info* out) {
    clear_error();
    if (!b || !out) return SPK_ERR_INVALID;
    out->passes = (int64_t)b->passes.size();
    out->chunks = (int64_t)b->chunks.size();
    out->paths = b->paths;
    out->film_pixels = b->film_pixels;
    out->queue = b->r->inflight;
    return SPK_OK;
}

int spk_batch_stats_get(spk_batch* b, spk_render_stats* out) {
    clear_error();
    if (!b) return SPK_ERR_INVALID;
    return spk_render_stats_get(b->r, out);
}

}  // extern "C"
