// spk_device.h -- layouts shared by the host code and the gfx950 kernels.
//
// Everything the kernels read lives in HBM as arrays of 16-byte vectors so each
// per-lane gather is one or more dwordx4 loads:
//   BVH node      64 B  (four float4: both child boxes + child links)
//   triangle      48 B  (v0 | prim id, e1, e2) in leaf order
//   shading rec   16+8 B (geometric normal | material, prim id / light id)
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace spk {

// Material lobes after spica's setScatterFuncs (bsdfs/diffuse.cc:23-32,
// conductor.cc:28-38, dielectric.cc:30-42, roughconductor.cc:38-75, plastic.cc:111-121,
// roughdielectric.cc:42-82, roughplastic.cc:128-158).
enum Lobe : int {
    LOBE_NONE = 0,
    LOBE_LAMBERT = 1,
    LOBE_SPECREFL = 2,
    LOBE_FRESNELSPEC = 3,
    LOBE_MICROREFL = 4,
    LOBE_PLASTIC = 5,      // PlasticBRDF
    LOBE_MICROTRANS = 6,   // MicrofacetTransmission
    LOBE_ROUGHPLASTIC = 7  // RoughPlasticBRDF
};

// a = (R, lobe | distribution << 8), b = (T, etaB), c = (eta, alpha), d = (k, specular
// sampling weight).  R is Kd and T is Ks for the plastics.
struct MatDev {
    float4 a, b, c, d;
};

// One emissive triangle: vertices, normalised vertex normals, area, radiance and
// the geometric normal spica's SurfaceInteraction would carry.
struct LightDev {
    float4 p0;  // xyz, w = area
    float4 e1;  // p1 - p0
    float4 e2;  // p2 - p0
    float4 n0, n1, n2;
    float4 L;   // radiance
    float4 ng;  // geometric normal
};

struct EnvDev {
    int present;
    int w, h;
    float radius;
    const float4* tex;      // level-0 texels (already scaled), [h][w]
    const float* marg_cdf;  // [h+1]
    const float* marg_pdf;  // [h]  func / integral
    const float* cond_cdf;  // [h][w+1]
    const float* cond_pdf;  // [h][w] func / row integral
    const float* pdf2d;     // [h][w] func / marginal integral (Distribution2D::pdf)
    float l2w[9];
    float w2l[9];
};

struct SceneDev {
    const float4* nodes;    // 4 per node
    const float4* tris;     // 3 per leaf-ordered triangle
    const float4* tri_ng;   // geometric normal, w = bits(material | smooth << 31)
    const int2* tri_ids;    // (scene prim id, area light id or -1)
    const float4* tri_n;    // 3 per triangle (vertex normals) when any triangle is smooth
    const MatDev* mats;
    const LightDev* lights;
    int num_areas;
    int num_lights;         // areas + env
    int has_bvh;
    int wide;               // holds Beckmann or plastic / glossy-transmission materials
    EnvDev env;
};

// Child link encoding inside a node: >= 0 interior node index; < 0 leaf with
// ~link = (first_triangle << 3) | (count - 1).
static constexpr int32_t kLeafCountBits = 3;
static constexpr int32_t kMaxLeafTris = 1 << kLeafCountBits;

struct CameraDev {
    float m[12];        // rows of camera-to-world (3x4)
    float sx_scale;     // aspect * tan(fov/2)
    float sy_scale;     // tan(fov/2)
    float lens_r;
    float focal;
    int W, H;
    int pad[2];
};

static constexpr int kMaxDepthCap = 64;

// ---- batched rendering (many images, each its own scene, in shared passes) ----
// Queue entries of a batch carry pixel | image << kBatchPixelBits in their pixel field
// and the absolute sample index in flag bits 7..31 (images up to 4 M pixels, up to 1024 images).
static constexpr uint32_t kBatchPixelBits = 22;
static constexpr uint32_t kBatchMaxImages = 1u << (32 - kBatchPixelBits);

struct ImageDev {
    CameraDev cam;          // cam.W x cam.H film
    uint32_t seed;
    uint32_t film_off;      // first film pixel of the image
    uint32_t pad[2];
};

// k samples of the pixel-order range [pixel_base, pixel_base + n_pix) of one image; its
// paths are [path_off, path_off + k n_pix) and its film pixels [pix_off, pix_off + n_pix)
// of the pass (sample-major, as single-image passes).
struct ChunkDev {
    uint32_t img, pixel_base, n_pix, sample_base, k, path_off, pix_off, pad;
};

struct PassDev {
    uint32_t chunk_begin, chunk_end, n_paths, n_pixels;
};

struct ConfigDev {
    CameraDev cam;
    int W, H;
    uint32_t npix;
    uint32_t queue;         // capacity (paths per pass)
    uint32_t first_sample;  // first sample of this run
    uint32_t end_sample;    // one past the last sample of this run
    uint32_t kmax;          // samples per pass when npix <= queue (else 1)
    uint32_t chunks;        // passes per sample when npix > queue (else 1)
    int max_depth;
    int rr_depth;
    int sampler;
    uint32_t seed;
    int integrator;         // SPK_INTEGRATOR_PATH / SPK_INTEGRATOR_DIRECT
    int wide;               // some scene needs the wide-material shading kernel
    int regen;              // spent paths are replaced by the pass's next camera paths
    uint32_t inflight;      // regen: queue capacity (paths in flight); the pass holds up to `queue`
    float guide_min_alpha;  // guided renders: microfacet reflections below this alpha are not guided
    int guide_depth;        // guided renders: vertices of bounces 0 .. guide_depth - 1 are guided
};

// Device-side per-pass state: queue lengths and work counters per bounce.
struct PassState {
    uint32_t pass_index;
    uint32_t done;
    uint32_t n;             // paths in this pass
    uint32_t sample_base;
    uint32_t pixel_base;    // first pixel-order index
    uint32_t n_pix;         // pixel slots per sample in this pass
    uint32_t k_samples;     // samples in this pass
    uint32_t pad;
    uint32_t chunk_begin;   // batch: chunks of this pass
    uint32_t chunk_end;
    uint32_t n_pixels;      // batch: film pixels touched by this pass
    uint32_t pad2;
    uint32_t countA[kMaxDepthCap + 2];
    uint32_t countS[kMaxDepthCap + 2];
    uint32_t countM[kMaxDepthCap + 2];
    uint32_t workA[kMaxDepthCap + 2];
    uint32_t workS[kMaxDepthCap + 2];
    uint32_t workM[kMaxDepthCap + 2];
    uint32_t countG[kMaxDepthCap + 2];  // vertices handed to the learned proposal
    unsigned long long total_paths;
    unsigned long long total_extend;
    unsigned long long total_shadow;
    unsigned long long total_mis;
    unsigned long long total_samples;
    unsigned long long total_guided;
    unsigned long long count_nodes;  // extend traversal node fetches (counting passes)
    unsigned long long count_tris;   // extend traversal triangle tests (counting passes)
    unsigned long long count_rays;   // extend rays traced in counting passes
    unsigned long long count_nodes_s;  // occlusion traversal node fetches (counting passes)
    unsigned long long count_tris_s;   // occlusion traversal triangle tests (counting passes)
    unsigned long long count_rays_s;   // occlusion rays traced in counting passes
    // regen: per iteration, the continuations carried over (n_cont), the new camera paths
    // appended after them (m_new, path ids from first_new), the next unstarted path of the
    // pass, the extend count, and done = no work left in this iteration
    uint32_t rg_next, rg_ncont, rg_mnew, rg_first, rg_nin, rg_done, rg_pad[2];
};

}  // namespace spk
