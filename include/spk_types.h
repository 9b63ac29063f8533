/*
 * spk_types.h -- plain-old-data descriptors shared by the MI355X path tracer's
 * C ABI (spk.h) and its CPU checker.
 *
 * Every struct here is a flat description of what spica's scene description
 * hands to PathIntegrator::render (reference: sources/spica/sceneparser.cc:190-335
 * builds primitives, materials and lights; sources/core/integrator.cc:46-110 reads
 * the camera, film and "sampleCount"; sources/integrators/path/path.cc:68 reads
 * "maxDepth").  No C++ types, no torch types: pointers and sizes only.
 */
#ifndef SPK_TYPES_H_
#define SPK_TYPES_H_

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Material kinds: the BSDF plugins on the path.
 *   DIFFUSE         bsdfs/diffuse.cc:23-32          LambertianReflection(clamp(reflectance))
 *   CONDUCTOR       bsdfs/conductor.cc:28-38        SpecularReflection + FresnelConductor(1, eta, k)
 *   DIELECTRIC      bsdfs/dielectric.cc:30-42       FresnelSpecular(re, tr, 1, gray(intIOR))
 *   ROUGHCONDUCTOR  bsdfs/roughconductor.cc:38-75   MicrofacetReflection + D(alpha) + FresnelConductor
 *   PLASTIC         bsdfs/plastic.cc:20-121         PlasticBRDF(Kd, Ks, 1, gray(intIOR))
 *   ROUGHDIELECTRIC bsdfs/roughdielectric.cc:42-82  MicrofacetTransmission(re, tr, D(alpha), 1, gray(intIOR));
 *                                                   FresnelSpecular when alpha == 0
 *   ROUGHPLASTIC    bsdfs/roughplastic.cc:20-160    RoughPlasticBRDF(Kd, Ks, D(alpha), 1, gray(intIOR))
 * D is the microfacet distribution ("distribution": GGX or Beckmann, spica's default).
 * Spica requires every shape to carry a BSDF (core/material.cc:21-26 dereferences it
 * unconditionally), so there is no "no material" kind. */
enum spk_material_type {
    SPK_MAT_DIFFUSE = 1,
    SPK_MAT_CONDUCTOR = 2,
    SPK_MAT_DIELECTRIC = 3,
    SPK_MAT_ROUGHCONDUCTOR = 4,
    SPK_MAT_PLASTIC = 5,
    SPK_MAT_ROUGHDIELECTRIC = 6,
    SPK_MAT_ROUGHPLASTIC = 7
};

/* "distribution" of the rough materials (core/microfacet.cc:196-356). */
enum spk_microfacet {
    SPK_MICROFACET_GGX = 0,
    SPK_MICROFACET_BECKMANN = 1
};

typedef struct spk_material {
    int32_t type;             /* enum spk_material_type */
    int32_t distribution;     /* enum spk_microfacet (rough materials) */
    double reflectance[3];    /* diffuse "reflectance"; (rough)dielectric "specularReflectance";
                               * (rough)plastic "diffuseReflectance" */
    double transmittance[3];  /* (rough)dielectric "specularTransmittance" */
    double eta[3];            /* (rough)conductor "eta" */
    double k[3];              /* (rough)conductor "k" */
    double int_ior;           /* (rough)dielectric / (rough)plastic "intIOR" */
    double alpha;             /* rough materials' "alpha" (spica default 0.1); 0 -> specular */
    double specular[3];       /* (rough)plastic "specularReflectance" */
} spk_material;

/* Scene: a triangle soup in world space (spica turns every mesh into one Triangle
 * primitive per face, sceneparser.cc:257-270), each triangle with a material and,
 * when emissive, its own area light (sceneparser.cc:173-179).  Light order is the
 * order of area_tri, followed by the environment map when present. */
typedef struct spk_scene_desc {
    int64_t num_tris;
    const float* verts;          /* [num_tris][3][3] world-space vertex positions */
    const float* normals;        /* [num_tris][3][3] per-vertex normals or NULL (face normals);
                                  * a triangle with three zero normals is flat */
    const int32_t* mat_ids;      /* [num_tris] index into materials */
    const int32_t* light_ids;    /* [num_tris] index into area lights or -1 */
    int32_t num_materials;
    int32_t num_area_lights;
    const spk_material* materials;
    const int32_t* area_tri;     /* [num_area_lights] the emitting triangle */
    const double* area_radiance; /* [num_area_lights][3] "radiance" */
    /* Environment map (lights/envmap.cc:17-55); env_width == 0 means none. */
    int32_t env_width;
    int32_t env_height;
    const float* env_rgb;        /* [env_height][env_width][3], row y = theta bin */
    double env_to_world[16];     /* row-major "toWorld"; spica stores its transpose */
    double env_scale;            /* "scale" */
    double env_center[3];        /* "worldCenter" */
    double env_radius;           /* "worldRadius" (spica default 2.0) */
} spk_scene_desc;

/* Acceleration-structure builder.  spica builds one binned-SAH BVH on the CPU
 * (accelerators/bvh.cc:112-148, the "accelerator" parameter "bvh"); SPK_BVH_LBVH builds a linear BVH
 * on the device instead (Morton order, radix tree), cheaper to build and somewhat
 * slower to trace.  Both give the same closest hits. */
enum spk_bvh_builder {
    SPK_BVH_SAH = 0,
    SPK_BVH_LBVH = 1
};

typedef struct spk_scene_build_opts {
    int32_t builder;             /* enum spk_bvh_builder */
    int32_t max_leaf;            /* triangles per leaf, 1..8 (spica: 4) */
    int32_t treelet_passes;      /* LBVH: treelet-restructuring sweeps, 0..16 (default 6; 0 = the
                                  * Morton tree as built) */
} spk_scene_build_opts;

/* Perspective camera (cameras/perspective.cc:23-74): "toWorld", "fov" (degrees,
 * vertical), "apertureRadius", "focusDistance", film "width"/"height". */
typedef struct spk_camera_desc {
    int32_t width;
    int32_t height;
    double to_world[16];         /* row-major camera-to-world (usually a lookAt) */
    double fov_deg;
    double aperture_radius;
    double focus_distance;
} spk_camera_desc;

enum spk_sampler_type {
    SPK_SAMPLER_INDEPENDENT = 0, /* counter-based PCG4D uniform per (pixel, sample, dim) */
    SPK_SAMPLER_SOBOL = 1,       /* shuffled, Owen-scrambled 2D Sobol per dimension pair */
    SPK_SAMPLER_MT19937 = 2      /* CPU checker only: spica's Independent/MT19937 stream */
};

/* The estimator: spica's "path" plugin (integrators/path/path.cc:42-125) or its
 * "directlighting" plugin (integrators/directlighting/directlighting.cc:21-57: emission
 * at the first hit, one-light NEE + MIS at every hit, recursion only through
 * specular reflection / transmission, core/integrator.cc:112-152). */
enum spk_integrator_type {
    SPK_INTEGRATOR_PATH = 0,
    SPK_INTEGRATOR_DIRECT = 1
};

/* Render request: "sampleCount" and "maxDepth" as spica reads them, plus the
 * counter-based sampler key and the integrator plugin. */
typedef struct spk_render_desc {
    int32_t spp;                 /* "sampleCount" */
    int32_t max_depth;           /* "maxDepth" (spica default 16) */
    int32_t sampler;             /* enum spk_sampler_type */
    uint32_t seed;
    int32_t rr_depth;            /* Russian roulette when bounces > rr_depth (spica: 3; path only) */
    int32_t first_sample;        /* sample index of the first pass (resume / split) */
    int32_t integrator;          /* enum spk_integrator_type (0 = path) */
} spk_render_desc;

#ifdef __cplusplus
}
#endif

#endif /* SPK_TYPES_H_ */
