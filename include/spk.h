/*
 * spk.h -- C ABI of the MI355X wavefront path tracer (libspk.so).
 *
 * This is the device boundary under spica's plugin layer.  The reference has no
 * device boundary: PathIntegrator::render runs in-process C++
 * (sources/core/integrator.cc:46-110 -> sources/integrators/path/path.cc:42-125).
 * Each entry point below names the reference interface it replaces.  All types
 * are POD (spk_types.h); status codes are returned instead of aborting
 * (spica's Assertion/FatalError call std::abort, sources/core/common.h:69-86).
 *
 * Threading: a scene or render handle may be used by one host thread at a time;
 * different handles are independent.  Each render handle owns one HIP stream.
 */
#ifndef SPK_H_
#define SPK_H_

#include <stddef.h>
#include <stdint.h>

#include "spk_types.h"

#ifdef __cplusplus
extern "C" {
#endif

enum spk_status {
    SPK_OK = 0,
    SPK_ERR_INVALID = 1,   /* bad argument / descriptor */
    SPK_ERR_HIP = 2,       /* HIP runtime error */
    SPK_ERR_NOMEM = 3,     /* device or host allocation failed */
    SPK_ERR_STATE = 4      /* call out of order (e.g. no device) */
};

typedef struct spk_scene spk_scene;
typedef struct spk_render spk_render;

/* Last error message of the calling thread ("" when none). */
const char* spk_last_error(void);

/* Library version string. */
const char* spk_version(void);

/* Selects the HIP device used by handles created afterwards on this thread.
 * Replaces spica's setNumThreads (sources/core/parallel.cc:175-181) as the
 * "where does render run" knob. */
int spk_init(int device);

/* Builds the BVH (binned SAH on the host, flattened to 64-byte nodes) and uploads
 * geometry, materials and lights.  Replaces BVHAccel::BVHAccel/construct
 * (sources/accelerators/bvh.cc:112-148) and Scene::Scene (sources/core/scene.cc:15-20),
 * together with the per-primitive material/light binding of
 * SceneParser::createPrimitive (sources/spica/sceneparser.cc:168-188). */
int spk_scene_create(const spk_scene_desc* desc, spk_scene** out);
int spk_scene_destroy(spk_scene* scene);

typedef struct spk_scene_stats {
    int64_t num_tris;
    int64_t num_nodes;      /* interior 64-byte nodes */
    int64_t num_leaves;
    int32_t max_depth;
    int32_t num_lights;
    double build_ms;        /* host BVH build */
    double upload_ms;
    int64_t device_bytes;
} spk_scene_stats;
int spk_scene_stats_get(const spk_scene* scene, spk_scene_stats* out);

/* Batched closest-hit (any_hit = 0) or occlusion (any_hit = 1) queries on device
 * buffers: org/dir are [n][3] float32, tmax is [n] or NULL (unbounded).  Writes
 * prim_out[n] (scene triangle index, -1 on a miss; for any_hit: 1/0) and t_out[n]
 * (closest t, or NULL).  Directions are normalised first, as spica's Ray does
 * (sources/core/ray.cc:12-20).  Replaces Scene::intersect (sources/core/scene.cc:37-43)
 * -> BVHAccel::intersectBVH (sources/accelerators/bvh.cc:331-387).  `stream` is a
 * hipStream_t (NULL = default stream); the call is asynchronous on it. */
int spk_trace(spk_scene* scene, const float* org, const float* dir, const float* tmax, int64_t n, int32_t any_hit,
              int32_t* prim_out, float* t_out, void* stream);

/* Tuning knobs that spica does not have (its CPU loop needs none). */
typedef struct spk_render_opts {
    int64_t queue_size;     /* paths in flight per pass (0 = auto: 2^21) */
    int32_t block_size;     /* 64, 128 or 256 (0 = auto) */
    int32_t use_graph;      /* capture each pass into a hipGraph and replay it */
    int32_t profile;        /* time every traversal launch with HIP events */
    int32_t count_visits;   /* count node fetches / triangle tests of the closest-hit stage */
} spk_render_opts;

/* Creates a render of `scene` through `camera`: film, path queues and stream.
 * Replaces the set-up half of SamplerIntegrator::render
 * (sources/core/integrator.cc:46-62) and Film::Film (sources/core/film.cc:8-21). */
int spk_render_create(spk_scene* scene, const spk_camera_desc* camera, const spk_render_desc* desc,
                      const spk_render_opts* opts, spk_render** out);
int spk_render_destroy(spk_render* render);

/* Traces `n_spp` more samples per pixel (continuing after the samples already
 * taken, so the film converges progressively as spica's per-pass loop does,
 * sources/core/integrator.cc:64-105).  Asynchronous unless `sync` is nonzero. */
int spk_render_run(spk_render* render, int32_t n_spp, int32_t sync);

/* Waits for the render stream. */
int spk_render_sync(spk_render* render);

/* Film read-back (sources/core/film.cc:23-40): per film pixel (row-major,
 * x mirrored as spica's addPixel does) the compensated sums of L and L^2 and the
 * sample count.  Any pointer may be NULL. */
int spk_render_film(spk_render* render, double* sum_rgb, double* sumsq_rgb, double* count);

/* Film mean Sum/(count + 1e-12) as float RGB (Film::save's division). */
int spk_render_image(spk_render* render, float* rgb);

/* relMSE = mean over pixels/channels of (I - R)^2 / (R^2 + eps), reduced on the GPU
 * in float64 against a reference image `ref_rgb` ([h][w][3] float64, host memory). */
int spk_render_relmse(spk_render* render, const double* ref_rgb, double eps, double* out);

typedef struct spk_render_stats {
    int64_t samples_done;   /* samples per pixel accumulated */
    int64_t paths;          /* camera paths traced */
    int64_t rays_extend;    /* closest-hit rays: camera + continuation */
    int64_t rays_shadow;    /* NEE light-sample occlusion rays */
    int64_t rays_mis;       /* NEE BSDF-sample closest-hit rays */
    int64_t passes;
    double extend_ms;       /* summed closest-hit traversal kernel time (profile = 1) */
    int64_t extend_launches;
    int64_t extend_rays_timed;
    double shadow_ms;
    int64_t shadow_launches;
    double shade_ms;
    int64_t count_nodes;    /* count_visits = 1: node fetches of the closest-hit stage */
    int64_t count_tris;     /* count_visits = 1: triangle tests of the closest-hit stage */
    int64_t count_rays;     /* count_visits = 1: closest-hit rays those counts cover */
} spk_render_stats;
int spk_render_stats_get(spk_render* render, spk_render_stats* out);

/* Raw 32-bit words of the counter-based sample streams for n (pixel, sample, dim)
 * keys (host arrays).  on_device = 0 evaluates on the host, 1 on the GPU with the
 * same code the kernels use.  Replaces Sampler::get1D (sources/core/sampler.h:28-31,
 * sources/samplers/independent.cc:17-19); the unit value is (word >> 8) * 2^-24. */
int spk_sampler_u32(int32_t kind, uint32_t seed, int64_t n, const uint32_t* pixel, const uint32_t* sample,
                    const uint32_t* dim, uint32_t* out, int32_t on_device);

#ifdef __cplusplus
}
#endif

#endif /* SPK_H_ */
