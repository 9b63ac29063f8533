// ref_harness.cpp -- TEST INFRASTRUCTURE ONLY.  A thin extern "C" driver over
// spica's own sources, compiled from /root/reference by oracle/build.py into
// oracle/_ref/libspica_ref.so, used to check the C restatement (spk_oracle.c)
// against the reference itself.  It builds scenes with spica's Triangle /
// GeometricPrimitive / AreaLight / BVHAccel / Scene, and calls
// PathIntegrator::Li through a subclass with an Independent(seed) sampler, so
// the reference's render loop (time(0)-seeded, sources/core/integrator.cc:51,71)
// and its Film/Image (which need stb, absent here) are never touched.
#include <cstdint>
#include <map>
#include <thread>
#include <memory>
#include <vector>

#include "core/bxdf.h"
#include "core/cobject.h"
#include "core/constant.h"
#include "core/fresnel.h"
#include "core/interaction.h"
#include "core/material.h"
#include "core/memory.h"
#include "core/microfacet.h"
#include "core/primitive.h"
#include "core/random.h"
#include "core/renderparams.h"
#include "core/scene.h"
#include "core/triangle.h"

// Each spica plugin header defines the extern "C" plugin entry points; renaming
// them per header keeps one translation unit.
#define createInstance createInstance_bvh
#define getDescription getDescription_bvh
#include "accelerators/bvh.h"
#undef createInstance
#undef getDescription
#define createInstance createInstance_diffuse
#define getDescription getDescription_diffuse
#include "bsdfs/diffuse.h"
#undef createInstance
#undef getDescription
#define createInstance createInstance_conductor
#define getDescription getDescription_conductor
#include "bsdfs/conductor.h"
#undef createInstance
#undef getDescription
#define createInstance createInstance_dielectric
#define getDescription getDescription_dielectric
#include "bsdfs/dielectric.h"
#undef createInstance
#undef getDescription
#define createInstance createInstance_roughconductor
#define getDescription getDescription_roughconductor
#include "bsdfs/roughconductor.h"
#undef createInstance
#undef getDescription
#define createInstance createInstance_path
#define getDescription getDescription_path
#include "integrators/path/path.h"
#undef createInstance
#undef getDescription
#define createInstance createInstance_plastic
#define getDescription getDescription_plastic
#include "bsdfs/plastic.h"
#undef createInstance
#undef getDescription
#define createInstance createInstance_roughdielectric
#define getDescription getDescription_roughdielectric
#include "bsdfs/roughdielectric.h"
#undef createInstance
#undef getDescription
#define createInstance createInstance_roughplastic
#define getDescription getDescription_roughplastic
#include "bsdfs/roughplastic.h"
#undef createInstance
#undef getDescription
#define createInstance createInstance_directlighting
#define getDescription getDescription_directlighting
#include "integrators/directlighting/directlighting.h"
#undef createInstance
#undef getDescription
#define createInstance createInstance_area
#define getDescription getDescription_area
#include "lights/area.h"
#undef createInstance
#undef getDescription
#define createInstance createInstance_independent
#define getDescription getDescription_independent
#include "samplers/independent.h"
#undef createInstance
#undef getDescription

#include "../include/spk_types.h"

using namespace spica;

namespace {

struct RefScene {
    std::vector<std::shared_ptr<Primitive>> prims;
    std::vector<std::shared_ptr<Light>> lights;
    std::map<const Primitive*, int> index;
    std::shared_ptr<Scene> scene;
};

std::shared_ptr<Texture<Spectrum>> tex(const double* v) {
    return std::make_shared<ConstantTexture<Spectrum>>(Spectrum(v[0], v[1], v[2]));
}
std::shared_ptr<Texture<Spectrum>> tex1(double v) {
    return std::make_shared<ConstantTexture<Spectrum>>(Spectrum(v, v, v));
}

const char* dist_name(const spk_material& m) { return m.distribution == SPK_MICROFACET_BECKMANN ? "beckmann" : "ggx"; }

std::shared_ptr<SurfaceMaterial> make_surface(const spk_material& m) {
    switch (m.type) {
    case SPK_MAT_DIFFUSE:
        return std::make_shared<Diffuse>(tex(m.reflectance));
    case SPK_MAT_CONDUCTOR:
        return std::make_shared<Conductor>(tex(m.eta), tex(m.k));
    case SPK_MAT_DIELECTRIC:
        return std::make_shared<Dielectric>(tex(m.reflectance), tex(m.transmittance), tex1(m.int_ior));
    case SPK_MAT_PLASTIC:
        return std::make_shared<Plastic>(tex(m.reflectance), tex(m.specular), tex1(m.int_ior));
    case SPK_MAT_ROUGHDIELECTRIC:
        return std::make_shared<RoughDielectric>(tex(m.reflectance), tex(m.transmittance), tex1(m.alpha), tex1(m.alpha),
                                                 tex1(m.int_ior), dist_name(m));
    case SPK_MAT_ROUGHPLASTIC:
        // the definition (roughplastic.cc:101-108) takes the index before the roughness,
        // as RoughPlastic(RenderParams&) passes them; the header names them the other way
        return std::make_shared<RoughPlastic>(tex(m.reflectance), tex(m.specular), tex1(m.int_ior), tex1(m.alpha),
                                              dist_name(m));
    default:
        return std::make_shared<RoughConductor>(tex(m.eta), tex(m.k), tex1(m.alpha), tex1(m.alpha), dist_name(m));
    }
}

class LiHarness : public PathIntegrator {
public:
    explicit LiHarness(const std::shared_ptr<Sampler>& s) : PathIntegrator(s) {}
    Spectrum li(const Scene& scene, RenderParams& params, const Ray& ray, Sampler& sampler, MemoryArena& arena) const {
        return Li(scene, params, ray, sampler, arena, 0);
    }
};

class DirectHarness : public DirectLightingIntegrator {
public:
    explicit DirectHarness(const std::shared_ptr<Sampler>& s) : DirectLightingIntegrator(s) {}
    Spectrum li(const Scene& scene, RenderParams& params, const Ray& ray, Sampler& sampler, MemoryArena& arena) const {
        return Li(scene, params, ray, sampler, arena, 0);
    }
};

// PathIntegrator::Li (integrator 0) or DirectLightingIntegrator::Li (integrator 1) with
// spica's Independent(seed) sampler, one fresh sampler per path.
template <class H>
void li_loop(RefScene* rs, int64_t n, const double* org, const double* dir, const uint32_t* seeds, int max_depth,
                    double* L_out) {
    RenderParams& params = RenderParams::getInstance();
    params.add("maxDepth", max_depth);
    // on a fresh thread, so the BxDFs' thread_local Random streams start default-seeded
    std::thread worker([&]() {
    auto dummy = std::make_shared<Independent>(0u);
    H h(dummy);
    MemoryArena arena;
    for (int64_t i = 0; i < n; i++) {
        Independent sampler(seeds[i]);
        Ray ray(Point3d(org[3 * i], org[3 * i + 1], org[3 * i + 2]), Vector3d(dir[3 * i], dir[3 * i + 1], dir[3 * i + 2]));
        Spectrum L = h.li(*rs->scene, params, ray, sampler, arena);
        L_out[3 * i] = L.red();
        L_out[3 * i + 1] = L.green();
        L_out[3 * i + 2] = L.blue();
        arena.reset();
    }
    });
    worker.join();
}

}  // namespace

extern "C" {

void* ref_scene_create(const spk_scene_desc* d) {
    RefScene* rs = new RefScene();
    std::vector<std::shared_ptr<Material>> mats;
    for (int m = 0; m < d->num_materials; m++)
        mats.push_back(std::make_shared<Material>(make_surface(d->materials[m]), nullptr));
    for (int64_t i = 0; i < d->num_tris; i++) {
        const float* v = d->verts + 9 * i;
        auto shape = std::make_shared<Triangle>(Point3d(v[0], v[1], v[2]), Point3d(v[3], v[4], v[5]),
                                                Point3d(v[6], v[7], v[8]));
        std::shared_ptr<Light> light = nullptr;
        int li = d->light_ids ? d->light_ids[i] : -1;
        if (li >= 0) {
            const double* L = d->area_radiance + 3 * li;
            light = std::make_shared<AreaLight>(shape, Transform(), Spectrum(L[0], L[1], L[2]));
            rs->lights.push_back(light);
        }
        auto prim = std::make_shared<GeometricPrimitive>(shape, mats[d->mat_ids[i]], light, nullptr);
        rs->index[prim.get()] = (int)i;
        rs->prims.push_back(prim);
    }
    auto accel = std::make_shared<BVHAccel>(rs->prims, false);
    rs->scene = std::make_shared<Scene>(accel, rs->lights);
    return rs;
}

void ref_scene_free(void* p) { delete (RefScene*)p; }

void ref_trace(void* p, int64_t n, const double* org, const double* dir, int any_hit, int32_t* prim_out, double* t_out) {
    RefScene* rs = (RefScene*)p;
    for (int64_t i = 0; i < n; i++) {
        Ray ray(Point3d(org[3 * i], org[3 * i + 1], org[3 * i + 2]), Vector3d(dir[3 * i], dir[3 * i + 1], dir[3 * i + 2]));
        if (any_hit) {
            prim_out[i] = rs->scene->intersect(ray) ? 1 : 0;
            continue;
        }
        SurfaceInteraction isect;
        if (rs->scene->intersect(ray, &isect)) {
            prim_out[i] = rs->index[isect.primitive()];
            t_out[i] = ray.maxDist();
        } else {
            prim_out[i] = -1;
            t_out[i] = -1.0;
        }
    }
}

void ref_li(void* p, int64_t n, const double* org, const double* dir, const uint32_t* seeds, int max_depth, double* L_out) {
    li_loop<LiHarness>((RefScene*)p, n, org, dir, seeds, max_depth, L_out);
}

void ref_li_direct(void* p, int64_t n, const double* org, const double* dir, const uint32_t* seeds, int max_depth,
                   double* L_out) {
    li_loop<DirectHarness>((RefScene*)p, n, org, dir, seeds, max_depth, L_out);
}

// MT19937 stream of spica's Random (core/random.cc:95-97).
void ref_mt(uint32_t seed, int64_t n, double* out) {
    Random r(seed);
    for (int64_t i = 0; i < n; i++) out[i] = r.get1D();
}

// BxDF sample / f / pdf in the local frame for the four lobes on the path.
static BxDF* make_bxdf(const spk_material& m, MemoryArena& arena) {
    switch (m.type) {
    case SPK_MAT_DIFFUSE:
        return arena.allocate<LambertianReflection>(Spectrum(m.reflectance[0], m.reflectance[1], m.reflectance[2]));
    case SPK_MAT_CONDUCTOR: {
        Fresnel* fr = arena.allocate<FresnelConductor>(Spectrum(1.0), Spectrum(m.eta[0], m.eta[1], m.eta[2]),
                                                       Spectrum(m.k[0], m.k[1], m.k[2]));
        return arena.allocate<SpecularReflection>(Spectrum(1.0), fr);
    }
    case SPK_MAT_DIELECTRIC:
        return arena.allocate<FresnelSpecular>(Spectrum(m.reflectance[0], m.reflectance[1], m.reflectance[2]),
                                               Spectrum(m.transmittance[0], m.transmittance[1], m.transmittance[2]), 1.0,
                                               Spectrum(m.int_ior).gray());
    case SPK_MAT_ROUGHDIELECTRIC: {
        MicrofacetDistribution* dist =
            m.distribution == SPK_MICROFACET_BECKMANN
                ? (MicrofacetDistribution*)arena.allocate<BeckmannDistribution>(m.alpha, m.alpha)
                : (MicrofacetDistribution*)arena.allocate<TrowbridgeReitzDistribution>(m.alpha, m.alpha);
        return arena.allocate<MicrofacetTransmission>(Spectrum(m.reflectance[0], m.reflectance[1], m.reflectance[2]),
                                                      Spectrum(m.transmittance[0], m.transmittance[1], m.transmittance[2]),
                                                      dist, 1.0, Spectrum(m.int_ior).gray());
    }
    default: {
        Fresnel* fr = arena.allocate<FresnelConductor>(Spectrum(1.0), Spectrum(m.eta[0], m.eta[1], m.eta[2]),
                                                       Spectrum(m.k[0], m.k[1], m.k[2]));
        MicrofacetDistribution* dist =
            m.distribution == SPK_MICROFACET_BECKMANN
                ? (MicrofacetDistribution*)arena.allocate<BeckmannDistribution>(m.alpha, m.alpha)
                : (MicrofacetDistribution*)arena.allocate<TrowbridgeReitzDistribution>(m.alpha, m.alpha);
        return arena.allocate<MicrofacetReflection>(Spectrum(1.0), dist, fr);
    }
    }
}

int ref_bxdf_sample(const spk_material* m, const double* wo, const double* u, double* wi, double* f, double* pdf) {
    MemoryArena arena;
    BxDF* b = make_bxdf(*m, arena);
    Vector3d w;
    double pd = 0.0;
    BxDFType t = BxDFType::None;
    Spectrum s = b->sample(Vector3d(wo[0], wo[1], wo[2]), &w, Point2d(u[0], u[1]), &pd, &t);
    wi[0] = w.x();
    wi[1] = w.y();
    wi[2] = w.z();
    f[0] = s.red();
    f[1] = s.green();
    f[2] = s.blue();
    *pdf = pd;
    return (int)t;
}

void ref_bxdf_eval(const spk_material* m, const double* wo, const double* wi, double* f, double* pdf) {
    MemoryArena arena;
    BxDF* b = make_bxdf(*m, arena);
    Spectrum s = b->f(Vector3d(wo[0], wo[1], wo[2]), Vector3d(wi[0], wi[1], wi[2]));
    f[0] = s.red();
    f[1] = s.green();
    f[2] = s.blue();
    *pdf = b->pdf(Vector3d(wo[0], wo[1], wo[2]), Vector3d(wi[0], wi[1], wi[2]));
}

}  // extern "C"
