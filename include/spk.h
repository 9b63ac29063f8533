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

/* Builds the BVH (binned SAH on the host, collapsed to 4-wide nodes of 64 bytes) and uploads
 * geometry, materials and lights.  Replaces BVHAccel::BVHAccel/construct
 * (sources/accelerators/bvh.cc:112-148) and Scene::Scene (sources/core/scene.cc:15-20),
 * together with the per-primitive material/light binding of
 * SceneParser::createPrimitive (sources/spica/sceneparser.cc:168-188). */
int spk_scene_create(const spk_scene_desc* desc, spk_scene** out);
/* The same, with a choice of builder and leaf size (opts NULL = defaults: SAH, 4).
 * spica creates its accelerator from the "accelerator" render parameter, fixed to
 * "bvh" (sources/spica/sceneparser.cc:64, 91-95 -> accelerators/bvh.cc:112-148); the
 * LBVH builder runs on the device instead of the host SAH build. */
int spk_scene_build_opts_default(spk_scene_build_opts* opts);
int spk_scene_create_ex(const spk_scene_desc* desc, const spk_scene_build_opts* opts, spk_scene** out);
int spk_scene_destroy(spk_scene* scene);

typedef struct spk_scene_stats {
    int64_t num_tris;
    int64_t num_nodes;      /* interior 4-wide nodes (64 bytes each, child boxes quantised) */
    int64_t num_leaves;
    int32_t max_depth;
    int32_t num_lights;
    double build_ms;        /* BVH build (host SAH, or device LBVH with its transfers) */
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
 * hipStream_t (NULL = default stream); the call is asynchronous on it.  The scene keeps
 * one traversal scratch area (work counter, stack spill): trace one scene on one stream
 * at a time (calls on the same stream are ordered; calls on several streams race). */
int spk_trace(spk_scene* scene, const float* org, const float* dir, const float* tmax, int64_t n, int32_t any_hit,
              int32_t* prim_out, float* t_out, void* stream);

/* spk_trace without allocation.  No counterpart in spica: it is for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc).  Same arguments, kernels and results as spk_trace, bit for bit; the
 * traversal scratch (each persistent thread's stack overflow) is a fixed 32 MiB array per
 * GPU, held from the library's load to the end of the process, instead of per-scene
 * allocations.  A scene whose BVH needs more stack overflow per 128-thread block than the
 * array holds is refused with SPK_ERR_INVALID; deep trees run a smaller grid.  All calls on
 * one GPU share the array: calls that overlap on one GPU race. */
int spk_trace_noalloc(spk_scene* scene, const float* org, const float* dir, const float* tmax, int64_t n,
                      int32_t any_hit, int32_t* prim_out, float* t_out, void* stream);

/* Tuning knobs that spica does not have (its CPU loop needs none). */
typedef struct spk_render_opts {
    int64_t queue_size;     /* paths in flight per pass, 64..2^27 (0 = auto: 2^21) */
    int32_t block_size;     /* 64, 128 or 256 (0 = auto) */
    int32_t use_graph;      /* capture each pass into a hipGraph and replay it */
    int32_t profile;        /* HIP events around every closest-hit launch (1), and also around
                             * every guide, shading and occlusion launch (2) */
    int32_t count_visits;   /* count node fetches / triangle tests of the closest-hit stage */
    int32_t classic_passes; /* 0 (default): a pass covers up to 2^28 paths and keeps queue_size of
                             * them in flight, replacing each finished path by the pass's next camera
                             * path, so every launch sees a full queue; 1: all of a pass's paths
                             * enter at once and the queues shrink bounce by bounce */
} spk_render_opts;

/* Creates a render of `scene` through `camera`: film, path queues and stream.
 * Replaces the set-up half of SamplerIntegrator::render
 * (sources/core/integrator.cc:46-62) and Film::Film (sources/core/film.cc:8-21);
 * desc->integrator picks PathIntegrator::Li (integrators/path/path.cc:42-125) or
 * DirectLightingIntegrator::Li (integrators/directlighting/directlighting.cc:21-57). */
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

/* The same over n film pixels (film indices `pixels`, host int32) against ref_rgb [n][3]:
 * a reference rendered on a pixel subset. */
int spk_render_relmse_pixels(spk_render* render, int64_t n, const int32_t* pixels, const double* ref_rgb, double eps,
                             double* out);

typedef struct spk_render_stats {
    int64_t samples_done;   /* samples per pixel accumulated */
    int64_t paths;          /* camera paths traced */
    int64_t rays_extend;    /* closest-hit rays: camera + continuation */
    int64_t rays_shadow;    /* NEE light-sample occlusion rays */
    int64_t rays_mis;       /* NEE BSDF-sample closest-hit rays */
    int64_t passes;
    double extend_ms;       /* summed closest-hit traversal kernel time (profile >= 1) */
    int64_t extend_launches;
    int64_t extend_rays_timed;
    double shadow_ms;       /* profile = 2: summed occlusion kernel time */
    int64_t shadow_launches;
    double shade_ms;        /* profile = 2: summed shading kernel time */
    int64_t count_nodes;    /* count_visits = 1: node fetches of the closest-hit stage */
    int64_t count_tris;     /* count_visits = 1: triangle tests of the closest-hit stage */
    int64_t count_rays;     /* count_visits = 1: closest-hit rays those counts cover */
    int64_t guided_vertices;/* continuations sampled by the learned proposal (guided renders) */
    int64_t train_passes;   /* sample passes that recorded training data (guided renders) */
    double guide_ms;        /* profile = 2, guided renders: summed guide-kernel (inference + sampling) time */
    int64_t guide_launches;
    int64_t shade_launches; /* profile = 2: shading launches timed in shade_ms */
    int64_t shadow_rays_timed; /* profile >= 1: occlusion rays of the profiled runs */
    int64_t count_nodes_shadow; /* count_visits = 1: node fetches of the occlusion stage */
    int64_t count_tris_shadow;
    int64_t count_rays_shadow;
} spk_render_stats;
int spk_render_stats_get(spk_render* render, spk_render_stats* out);

/* ---- Learned path guiding ------------------------------------------------------------
 * No counterpart in spica: its PathIntegrator samples every continuation from the BSDF
 * (sources/integrators/path/path.cc:83-94).  A guide is a 224-64-64-64-64-40 ReLU network
 * (bf16 weights, fp32 accumulation) over one-blob codes of position, normal and outgoing
 * direction whose 40 outputs define an 8-lobe von Mises-Fisher mixture over incident
 * directions.  Attached to a render it is trained online by Adam on the render's own
 * radiance samples during the first `train_spp` samples per pixel, and continuations at
 * diffuse vertices (and at microfacet reflections of alpha >= glossy_min_alpha) of the first
 * `guided_depth` bounces are drawn from guide_fraction * guide + (1 - guide_fraction) * BSDF
 * with the mixture density in the weight, so the estimator stays unbiased.
 *
 * Network inputs x (7 floats per sample, in [0,1]): position scaled to the scene box (3),
 * then for the normal and the outgoing direction ((v.z + 1)/2, atan2(v.y, v.x)/2pi + 1/2). */
typedef struct spk_guide spk_guide;

typedef struct spk_guide_opts {
    uint32_t seed;           /* weight initialisation */
    int32_t batch_size;      /* training samples per Adam step (default 2^18, multiple of 64) */
    int32_t steps_per_pass;  /* Adam steps after each training round (default 8) */
    int32_t train_spp;       /* samples per pixel of online training (default 64) */
    float learning_rate;     /* default 1e-2 */
    float beta1, beta2, epsilon;  /* default 0.95, 0.95, 1e-8 */
    float guide_fraction;    /* probability of sampling the proposal at a vertex (default 0.5) */
    int32_t train_pass_spp;  /* samples per pixel per training round (default 8: with 8 steps,
                              * 64 Adam steps over the 64 training spp); each round keeps
                              * 1/train_pass_spp of its vertices as training records */
    float glossy_min_alpha;  /* microfacet reflections are guided only where their roughness
                              * alpha is at least this (default 0.5; 0 guides all of them):
                              * narrower lobes are left to BSDF sampling */
    int32_t guided_depth;    /* vertices of bounces 0 .. guided_depth - 1 are guided (default 3,
                              * >= 1): a BSDF-drawn direction the guide holds dark weighs
                              * 1/(1 - guide_fraction) times its plain weight, and over many
                              * guided bounces that factor compounds */
} spk_guide_opts;

int spk_guide_opts_default(spk_guide_opts* out);
int spk_guide_create(int device, const spk_guide_opts* opts, spk_guide** out);
int spk_guide_destroy(spk_guide* guide);
int64_t spk_guide_num_params(void);
/* Flat parameters in order W1[64][224] b1 W2[64][64] b2 W3 b3 W4 b4 W5[40][64] b5 (fp32).
 * Setting them resets the Adam moments and step count. */
int spk_guide_params_get(spk_guide* guide, float* out);
int spk_guide_params_set(spk_guide* guide, const float* in);
/* Host arrays: x [n][7], dirs [n][3] unit vectors.  out_raw [n][40] and log_pdf [n] may be NULL. */
int spk_guide_eval(spk_guide* guide, int64_t n, const float* x, const float* dirs, float* out_raw, float* log_pdf);
/* Directions drawn from the mixture with uniforms u [n][3] (lobe, then the vMF about its
 * mean); dirs [n][3] and their density pdf [n]. */
int spk_guide_sample(spk_guide* guide, int64_t n, const float* x, const float* u, float* dirs, float* pdf);
/* Gradient of the mean over n <= batch_size samples of -weight * log q(dir), and that mean. */
int spk_guide_grad(spk_guide* guide, int64_t n, const float* x, const float* dirs, const float* weight, float* grad,
                   double* loss);
/* `steps` Adam steps on one batch of n <= batch_size samples. */
int spk_guide_train(spk_guide* guide, int64_t n, const float* x, const float* dirs, const float* weight, int32_t steps);

typedef struct spk_guide_stats {
    int64_t steps;           /* Adam steps taken */
    int64_t last_records;    /* training records in the last training pass */
    double last_loss;        /* mean -weight * log q of the last step's batch */
    double last_weight;      /* mean weight of the last step's batch */
} spk_guide_stats;
int spk_guide_stats_get(spk_guide* guide, spk_guide_stats* out);

/* Samples per pixel a render trains the guide for after spk_render_set_guide (the
 * train_spp option); 0 freezes the network, e.g. to reuse it for other views of a scene. */
int spk_guide_set_train_spp(spk_guide* guide, int32_t train_spp);

/* Attach a guide to a render (NULL detaches: plain BSDF sampling again).  The render
 * trains it during its first train_spp samples per pixel after attaching.  One guide
 * serves one render at a time. */
int spk_render_set_guide(spk_render* render, spk_guide* guide);

/* ---- Batched rendering ---------------------------------------------------------------
 * Many images, each with its own scene, camera, sample count and seed, rendered by one
 * set of wavefront queues: their work is cut into chunks (k samples of a pixel range)
 * packed into shared passes, so small images fill the queues together, and a pass is one
 * captured hipGraph replayed (or the whole batch one graph).  Each image's film equals
 * the film spk_render_run gives for it alone, bit for bit (samples are keyed by pixel,
 * sample and dimension, and every pixel accumulates in sample order).  Replaces running
 * spica's PathIntegrator::render once per scene file (sources/integrators/path/path.cc,
 * sources/core/integrator.cc:62-107).  Limits: W * H <= 2^22 per image, up to 1024
 * images; a guide serves a batch whose images render one scene (spk_batch_set_guide). */
typedef struct spk_batch spk_batch;

typedef struct spk_batch_image {
    spk_scene* scene;
    spk_camera_desc camera;
    int32_t spp;
    uint32_t seed;
} spk_batch_image;

typedef struct spk_batch_desc {
    int32_t num_images;
    const spk_batch_image* images;
    int32_t max_depth;
    int32_t rr_depth;
    int32_t sampler;      /* SPK_SAMPLER_INDEPENDENT or SPK_SAMPLER_SOBOL */
    int32_t graph;        /* 0: plain launches, 1: one graph per pass, replayed, 2: the whole batch one graph */
} spk_batch_desc;

int spk_batch_create(const spk_batch_desc* desc, const spk_render_opts* opts, spk_batch** out);
int spk_batch_destroy(spk_batch* batch);
/* Attach a learned-proposal guide to a batch whose images all render one scene (views of it),
 * or detach it (NULL).  The first train_spp samples of every image then run first, as training
 * rounds of train_pass_spp samples taken round by round across the images, each round followed
 * by the guide's Adam steps; the rest of the samples use the trained network.  One network for
 * all the views, trained online as spk_render_set_guide does for one view.  A guided batch
 * launches its passes one by one (no graph replay). */
int spk_batch_set_guide(spk_batch* batch, spk_guide* guide);
/* Renders every image to its sample count (films are cleared first). */
int spk_batch_run(spk_batch* batch, int32_t sync);
int spk_batch_film(spk_batch* batch, int32_t image, double* sum_rgb, double* sumsq_rgb, double* count);
typedef struct spk_batch_info {
    int64_t passes;       /* passes per run */
    int64_t chunks;
    int64_t paths;        /* camera paths per run */
    int64_t film_pixels;
    int64_t queue;        /* paths in flight */
} spk_batch_info;
int spk_batch_info_get(spk_batch* batch, spk_batch_info* out);
int spk_batch_stats_get(spk_batch* batch, spk_render_stats* out);

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
