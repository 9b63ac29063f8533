// spica_compat.h -- spica's plugin-facing interface for the GPU path integrator.
//
// Mirrors the names, parameter keys and call sequence of spica's plugin layer so a
// caller written against spica drives the MI355X path tracer the same way:
//   PluginManager::initModule("path")  -> dlopen(".../libspica_path.so")   (core/cobject.cc:12,32-40)
//   createInstance(RenderParams&)      -> new PathIntegrator(params)        (core/cobject.h:56-64, path.h:41)
//   ("directlighting": libspica_directlighting.so -> DirectLightingIntegrator, directlighting.h:28)
//   integrator->render(camera, scene, params)                               (core/integrator.h:34-36)
// Scene objects (Triangle, GeometricPrimitive, Diffuse, Conductor, Dielectric,
// RoughConductor, Plastic, RoughDielectric, RoughPlastic, AreaLight, Envmap,
// PerspectiveCamera, HDRFilm) take spica's
// parameter keys and can describe themselves as the flat descriptors of
// include/spk_types.h, which is what crosses the C ABI into libspk.so.
#pragma once

#include <array>
#include <map>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/spk.h"

namespace spica {

struct Spectrum {
    double r = 0, g = 0, b = 0;
    Spectrum() = default;
    explicit Spectrum(double v) : r(v), g(v), b(v) {}
    Spectrum(double r_, double g_, double b_) : r(r_), g(g_), b(b_) {}
    double gray() const { return 0.2126 * r + 0.7152 * g + 0.0722 * b; }  // core/spectrum.cc:677-679
};

struct Point3d {
    double x = 0, y = 0, z = 0;
};

// Row-major 4x4, applied to column vectors (core/transform.h).
struct Transform {
    std::array<double, 16> m{{1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1}};
    static Transform lookAt(const Point3d& eye, const Point3d& look, const Point3d& up);  // core/transform.cc:167-198
};

class CObject {
public:
    virtual ~CObject() = default;
};

// Typed, string-keyed parameters (core/renderparams.h:27-87).  Missing required
// keys raise spica::ParamError instead of spica's FatalError/abort.
struct ParamError : std::runtime_error {
    using std::runtime_error::runtime_error;
};

class RenderParams {
public:
    static RenderParams& getInstance();
    void clear();
    void add(const std::string& name, int v) { ints_[name] = v; }
    void add(const std::string& name, double v) { doubles_[name] = v; }
    void add(const std::string& name, const std::string& v) { strings_[name] = v; }
    void add(const std::string& name, const char* v) { strings_[name] = v; }
    void add(const std::string& name, const Spectrum& v) { spectra_[name] = v; }
    void add(const std::string& name, const Transform& v) { transforms_[name] = v; }
    void add(const std::string& name, const std::shared_ptr<CObject>& v) { objects_[name] = v; }
    int getInt(const std::string& name, bool remove = false);
    int getInt(const std::string& name, int value, bool remove = false);
    double getDouble(const std::string& name, bool remove = false);
    double getDouble(const std::string& name, double value, bool remove = false);
    std::string getString(const std::string& name, bool remove = false);
    std::string getString(const std::string& name, const std::string& value, bool remove = false);
    Spectrum getSpectrum(const std::string& name, bool remove = false);
    Spectrum getSpectrum(const std::string& name, const Spectrum& value, bool remove = false);
    Transform getTransform(const std::string& name, bool remove = false);
    Transform getTransform(const std::string& name, const Transform& value, bool remove = false);
    std::shared_ptr<CObject> getObject(const std::string& name, bool remove = false);
    std::shared_ptr<CObject> getObject(const std::string& name, const std::shared_ptr<CObject>& value, bool remove = false);

private:
    std::map<std::string, int> ints_;
    std::map<std::string, double> doubles_;
    std::map<std::string, std::string> strings_;
    std::map<std::string, Spectrum> spectra_;
    std::map<std::string, Transform> transforms_;
    std::map<std::string, std::shared_ptr<CObject>> objects_;
};

// ---- film / camera (films/hdrfilm.cc, cameras/perspective.cc) ---------------------
class Film : public CObject {
public:
    Film(int width, int height, const std::string& filename);
    explicit Film(RenderParams& params);  // "width", "height", "outputFile" (hdrfilm.cc:15-19)
    int width() const { return w_; }
    int height() const { return h_; }
    // per film pixel mean RGB (row-major, x mirrored as spica's addPixel does)
    std::vector<double>& image() { return rgb_; }
    const std::vector<double>& image() const { return rgb_; }
    const std::string& filename() const { return filename_; }
    // Radiance RGBE writer (the layout spica's HDRFilm writes, core/image.cc:390-435)
    void saveHDR(const std::string& path) const;
    // Film::save(id) + HDRFilm::saveImage (film.cc:23-40, hdrfilm.cc:21-25): writes
    // sprintf(filename, id) + ".hdr".  The name may hold one integer conversion
    // (%d, %i, %u with flags and width) and %%; anything else raises ParamError.
    std::string save(int id) const;
    static std::string formatName(const std::string& fmt, int id);

private:
    int w_, h_;
    std::string filename_;
    std::vector<double> rgb_;
};

class Camera : public CObject {
public:
    virtual ~Camera() = default;
    virtual spk_camera_desc describe() const = 0;
    std::shared_ptr<Film> film() const { return film_; }

protected:
    std::shared_ptr<Film> film_;
};

class PerspectiveCamera : public Camera {
public:
    PerspectiveCamera(const Transform& toWorld, double fovDeg, const std::shared_ptr<Film>& film,
                      double apertureRadius = 0.0, double focusDistance = 50.0);
    explicit PerspectiveCamera(RenderParams& params);  // "toWorld", "fov", "apertureRadius", "focusDistance", "film"
    spk_camera_desc describe() const override;

private:
    Transform toWorld_;
    double fov_, aperture_, focus_;
};

// ---- materials (bsdfs/*.cc) ----------------------------------------------------------
class SurfaceMaterial : public CObject {
public:
    virtual spk_material describe() const = 0;
};
class Diffuse : public SurfaceMaterial {  // "reflectance"
public:
    explicit Diffuse(const Spectrum& reflectance) : R_(reflectance) {}
    explicit Diffuse(RenderParams& params);
    spk_material describe() const override;

private:
    Spectrum R_;
};
class Conductor : public SurfaceMaterial {  // "eta", "k"
public:
    Conductor(const Spectrum& eta, const Spectrum& k) : eta_(eta), k_(k) {}
    explicit Conductor(RenderParams& params);
    spk_material describe() const override;

private:
    Spectrum eta_, k_;
};
class Dielectric : public SurfaceMaterial {  // "specularReflectance", "specularTransmittance", "intIOR"
public:
    Dielectric(const Spectrum& Kr, const Spectrum& Kt, double intIOR) : Kr_(Kr), Kt_(Kt), ior_(intIOR) {}
    explicit Dielectric(RenderParams& params);
    spk_material describe() const override;

private:
    Spectrum Kr_, Kt_;
    double ior_;
};
// "distribution": "beckmann" (spica's default) or "ggx" (core/microfacet.cc:196-356)
int microfacetFromName(const std::string& name);

class RoughConductor : public SurfaceMaterial {  // "eta", "k", "alpha", "distribution"
public:
    RoughConductor(const Spectrum& eta, const Spectrum& k, double alpha, int distribution = SPK_MICROFACET_GGX)
        : eta_(eta), k_(k), alpha_(alpha), dist_(distribution) {}
    explicit RoughConductor(RenderParams& params);
    spk_material describe() const override;

private:
    Spectrum eta_, k_;
    double alpha_;
    int dist_;
};
class Plastic : public SurfaceMaterial {  // "diffuseReflectance", "specularReflectance", "intIOR" (bsdfs/plastic.cc:96-102)
public:
    Plastic(const Spectrum& Kd, const Spectrum& Ks, double intIOR) : Kd_(Kd), Ks_(Ks), ior_(intIOR) {}
    explicit Plastic(RenderParams& params);
    spk_material describe() const override;

private:
    Spectrum Kd_, Ks_;
    double ior_;
};
// "specularReflectance", "specularTransmittance", "alpha", "intIOR", "distribution" (roughdielectric.cc:31-40)
class RoughDielectric : public SurfaceMaterial {
public:
    RoughDielectric(const Spectrum& Kr, const Spectrum& Kt, double alpha, double intIOR, int distribution)
        : Kr_(Kr), Kt_(Kt), alpha_(alpha), ior_(intIOR), dist_(distribution) {}
    explicit RoughDielectric(RenderParams& params);
    spk_material describe() const override;

private:
    Spectrum Kr_, Kt_;
    double alpha_, ior_;
    int dist_;
};
// "diffuseReflectance", "specularReflectance", "intIOR", "alpha", "distribution" (roughplastic.cc:111-118)
class RoughPlastic : public SurfaceMaterial {
public:
    RoughPlastic(const Spectrum& Kd, const Spectrum& Ks, double alpha, double intIOR, int distribution)
        : Kd_(Kd), Ks_(Ks), alpha_(alpha), ior_(intIOR), dist_(distribution) {}
    explicit RoughPlastic(RenderParams& params);
    spk_material describe() const override;

private:
    Spectrum Kd_, Ks_;
    double alpha_, ior_;
    int dist_;
};

// ---- shapes, primitives, lights --------------------------------------------------------
class Triangle : public CObject {
public:
    Triangle(const Point3d& p0, const Point3d& p1, const Point3d& p2) : p_{{p0, p1, p2}} {}
    const Point3d& operator[](int i) const { return p_[i]; }

private:
    std::array<Point3d, 3> p_;
};

class Light : public CObject {
public:
    virtual ~Light() = default;
};

class AreaLight : public Light {  // "radiance"; one per emissive triangle (sceneparser.cc:173-179)
public:
    AreaLight(const std::shared_ptr<Triangle>& shape, const Spectrum& radiance) : shape_(shape), L_(radiance) {}
    const Spectrum& radiance() const { return L_; }
    const std::shared_ptr<Triangle>& shape() const { return shape_; }

private:
    std::shared_ptr<Triangle> shape_;
    Spectrum L_;
};

class Envmap : public Light {  // lights/envmap.cc:49-55: "toWorld", "scale", "worldCenter", "worldRadius"
public:
    Envmap(int width, int height, std::vector<float> rgb, const Transform& toWorld = Transform(), double scale = 1.0,
           const Point3d& center = Point3d(), double radius = 2.0);
    void fill(spk_scene_desc& d) const;

private:
    int w_, h_;
    std::vector<float> rgb_;
    Transform toWorld_;
    double scale_;
    Point3d center_;
    double radius_;
};

class GeometricPrimitive : public CObject {  // core/primitive.h
public:
    GeometricPrimitive(const std::shared_ptr<Triangle>& shape, const std::shared_ptr<SurfaceMaterial>& material,
                       const std::shared_ptr<AreaLight>& light = nullptr)
        : shape_(shape), material_(material), light_(light) {}
    const std::shared_ptr<Triangle>& shape() const { return shape_; }
    const std::shared_ptr<SurfaceMaterial>& material() const { return material_; }
    const std::shared_ptr<AreaLight>& light() const { return light_; }

private:
    std::shared_ptr<Triangle> shape_;
    std::shared_ptr<SurfaceMaterial> material_;
    std::shared_ptr<AreaLight> light_;
};

// Scene(primitives, lights) (core/scene.h); the GPU BVH is built from the primitives
// when the integrator renders, replacing the "accelerator" plugin.
class Scene {
public:
    Scene(std::vector<std::shared_ptr<GeometricPrimitive>> prims, std::vector<std::shared_ptr<Light>> lights)
        : prims_(std::move(prims)), lights_(std::move(lights)) {}
    const std::vector<std::shared_ptr<GeometricPrimitive>>& primitives() const { return prims_; }
    const std::vector<std::shared_ptr<Light>>& lights() const { return lights_; }

private:
    std::vector<std::shared_ptr<GeometricPrimitive>> prims_;
    std::vector<std::shared_ptr<Light>> lights_;
};

// ---- samplers (samplers/independent.cc + the counter-based Sobol) -----------------------
class Sampler : public CObject {
public:
    explicit Sampler(int kind, unsigned seed = 0) : kind_(kind), seed_(seed) {}
    int kind() const { return kind_; }
    unsigned seed() const { return seed_; }

private:
    int kind_;
    unsigned seed_;
};
class Independent : public Sampler {
public:
    explicit Independent(unsigned seed = 0) : Sampler(SPK_SAMPLER_INDEPENDENT, seed) {}
    explicit Independent(RenderParams& params);  // "seed" (spica's Independent(params) seeds 0)
};
class Sobol : public Sampler {
public:
    explicit Sobol(unsigned seed = 0) : Sampler(SPK_SAMPLER_SOBOL, seed) {}
    explicit Sobol(RenderParams& params);
};

// ---- integrator (core/integrator.h, integrators/path/path.h) ----------------------------
class Integrator : public CObject {
public:
    virtual ~Integrator() = default;
    virtual void render(const std::shared_ptr<const Camera>& camera, const Scene& scene, RenderParams& params) = 0;
};

// SamplerIntegrator::render (core/integrator.cc:46-110) on the GPU, for the estimator a
// subclass names.  Reads "sampleCount" (required), "maxDepth" (default 16), "sampler"
// (object), "accelerator" ("bvh" or "lbvh") and the GPU knobs "queueSize", "blockSize",
// "device", "guiding" (path only), "saveEvery".  Fills camera->film() with the per-pixel
// mean.  spica saves the film after every sample pass as Film::save(samples so far)
// (integrator.cc:100); here it is saved every "saveEvery" samples (default: once, after
// the last, which is the file spica leaves behind) when the film has a non-empty name.
class SamplerIntegrator : public Integrator {
public:
    ~SamplerIntegrator() override;
    void render(const std::shared_ptr<const Camera>& camera, const Scene& scene, RenderParams& params) override;
    // traced-ray counters of the last render (extend, shadow, MIS)
    const spk_render_stats& lastStats() const { return stats_; }

protected:
    SamplerIntegrator(const std::shared_ptr<Sampler>& sampler, int integrator);

private:
    std::shared_ptr<Sampler> sampler_;
    int integrator_;
    spk_render_stats stats_{};
};

// PathIntegrator::Li (integrators/path/path.cc:42-125).
class PathIntegrator : public SamplerIntegrator {
public:
    explicit PathIntegrator(const std::shared_ptr<Sampler>& sampler);
    explicit PathIntegrator(RenderParams& params);
};

// DirectLightingIntegrator::Li (integrators/directlighting/directlighting.cc:21-57).
class DirectLightingIntegrator : public SamplerIntegrator {
public:
    explicit DirectLightingIntegrator(const std::shared_ptr<Sampler>& sampler);
    explicit DirectLightingIntegrator(RenderParams& params);
};

}  // namespace spica

// Plugin entry points with C linkage, as SPICA_EXPORT_PLUGIN emits them (core/cobject.h:56-64).
extern "C" {
spica::CObject* createInstance(spica::RenderParams& params);
const char* getDescription();
}
