// path_plugin.cpp -- the "path" integrator plugin on the MI355X (libspica_path.so); built
// with SPK_PLUGIN_DIRECTLIGHTING it is the "directlighting" plugin
// (libspica_directlighting.so).
//
// SamplerIntegrator::render replaces SamplerIntegrator::render + PathIntegrator::Li
// (sources/core/integrator.cc:46-110, sources/integrators/path/path.cc:42-125): it
// flattens the scene into spk_scene_desc, the camera into spk_camera_desc and the
// parameters into spk_render_desc, then calls the C ABI of libspk.so.
#include <math.h>
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <cstdint>
#include <map>

#include "spica_compat.h"

namespace spica {

// ---- Transform ---------------------------------------------------------------------
Transform Transform::lookAt(const Point3d& eye, const Point3d& look, const Point3d& up) {
    auto nrm = [](double v[3]) {
        double inv = 1.0 / sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
        v[0] *= inv;
        v[1] *= inv;
        v[2] *= inv;
    };
    auto cross = [](const double a[3], const double b[3], double o[3]) {
        o[0] = a[1] * b[2] - a[2] * b[1];
        o[1] = a[2] * b[0] - a[0] * b[2];
        o[2] = a[0] * b[1] - a[1] * b[0];
    };
    double dir[3] = {look.x - eye.x, look.y - eye.y, look.z - eye.z};
    nrm(dir);
    double u[3] = {up.x, up.y, up.z};
    nrm(u);
    double left[3], nup[3];
    cross(u, dir, left);
    if (left[0] == 0 && left[1] == 0 && left[2] == 0) throw ParamError("lookAt: up and viewing direction are parallel");
    nrm(left);
    cross(dir, left, nup);
    Transform t;
    t.m = {{left[0], nup[0], dir[0], eye.x, left[1], nup[1], dir[1], eye.y, left[2], nup[2], dir[2], eye.z, 0, 0, 0, 1}};
    return t;
}

// ---- RenderParams ------------------------------------------------------------------------
RenderParams& RenderParams::getInstance() {
    static RenderParams inst;
    return inst;
}
void RenderParams::clear() {
    ints_.clear();
    doubles_.clear();
    strings_.clear();
    spectra_.clear();
    transforms_.clear();
    objects_.clear();
}

namespace {
template <class M, class V>
V take(M& m, const std::string& name, bool remove) {
    auto it = m.find(name);
    if (it == m.end()) throw ParamError("RenderParams: required parameter \"" + name + "\" is missing");
    V v = it->second;
    if (remove) m.erase(it);
    return v;
}
template <class M, class V>
V take_or(M& m, const std::string& name, const V& def, bool remove) {
    auto it = m.find(name);
    if (it == m.end()) return def;
    V v = it->second;
    if (remove) m.erase(it);
    return v;
}
}  // namespace

int RenderParams::getInt(const std::string& n, bool r) { return take<decltype(ints_), int>(ints_, n, r); }
int RenderParams::getInt(const std::string& n, int v, bool r) { return take_or(ints_, n, v, r); }
double RenderParams::getDouble(const std::string& n, bool r) {
    auto it = doubles_.find(n);
    if (it == doubles_.end() && ints_.count(n)) return (double)getInt(n, r);
    return take<decltype(doubles_), double>(doubles_, n, r);
}
double RenderParams::getDouble(const std::string& n, double v, bool r) {
    if (!doubles_.count(n) && ints_.count(n)) return (double)getInt(n, r);
    return take_or(doubles_, n, v, r);
}
std::string RenderParams::getString(const std::string& n, bool r) { return take<decltype(strings_), std::string>(strings_, n, r); }
std::string RenderParams::getString(const std::string& n, const std::string& v, bool r) { return take_or(strings_, n, v, r); }
Spectrum RenderParams::getSpectrum(const std::string& n, bool r) { return take<decltype(spectra_), Spectrum>(spectra_, n, r); }
Spectrum RenderParams::getSpectrum(const std::string& n, const Spectrum& v, bool r) { return take_or(spectra_, n, v, r); }
Transform RenderParams::getTransform(const std::string& n, bool r) { return take<decltype(transforms_), Transform>(transforms_, n, r); }
Transform RenderParams::getTransform(const std::string& n, const Transform& v, bool r) { return take_or(transforms_, n, v, r); }
std::shared_ptr<CObject> RenderParams::getObject(const std::string& n, bool r) {
    return take<decltype(objects_), std::shared_ptr<CObject>>(objects_, n, r);
}
std::shared_ptr<CObject> RenderParams::getObject(const std::string& n, const std::shared_ptr<CObject>& v, bool r) {
    return take_or(objects_, n, v, r);
}

// ---- Film / camera ------------------------------------------------------------------------
Film::Film(int width, int height, const std::string& filename)
    : w_(width), h_(height), filename_(filename), rgb_((size_t)width * height * 3, 0.0) {
    if (width <= 0 || height <= 0) throw ParamError("Film: non-positive resolution");
}
Film::Film(RenderParams& p)
    : Film(p.getInt("width"), p.getInt("height"), p.getString("outputFile", std::string(""))) {}

void Film::saveHDR(const std::string& path) const {
    // Each scanline: 2, 2, width (big endian), then the R, G, B and exponent planes in
    // literal runs of at most 127 bytes; a pixel is (c * m * 256 / max, e + 128) with
    // max = m * 2^e, and black below 1e-32.
    std::vector<unsigned char> out;
    char head[160];
    snprintf(head, sizeof(head),
             "#?RADIANCE\n# Made with 100%% pure HDR Shop\nFORMAT=32-bit_rle_rgbe\nEXPOSURE=1.0000000000000\n\n-Y %d +X %d\n", h_,
             w_);
    out.insert(out.end(), head, head + strlen(head));
    std::vector<unsigned char> px((size_t)w_ * 4);
    for (int y = 0; y < h_; y++) {
        for (int x = 0; x < w_; x++) {
            const double* c = &rgb_[((size_t)y * w_ + x) * 3];
            const double v = std::max(c[0], std::max(c[1], c[2]));
            unsigned char* e = &px[(size_t)x * 4];
            if (!(v > 1e-32)) {
                e[0] = e[1] = e[2] = e[3] = 0;
            } else {
                int ex;
                const double m = frexp(v, &ex) * 256.0 / v;
                for (int k = 0; k < 3; k++) e[k] = (unsigned char)std::min(255.0, std::max(0.0, c[k] * m));
                e[3] = (unsigned char)(ex + 128);
            }
        }
        out.insert(out.end(), {2, 2, (unsigned char)((w_ >> 8) & 0xFF), (unsigned char)(w_ & 0xFF)});
        for (int k = 0; k < 4; k++)
            for (int x = 0; x < w_;) {
                const int run = std::min(127, w_ - x);
                out.push_back((unsigned char)run);
                for (int j = x; j < x + run; j++) out.push_back(px[(size_t)j * 4 + k]);
                x += run;
            }
    }
    FILE* f = fopen(path.c_str(), "wb");
    if (!f) throw std::runtime_error("Film::saveHDR: cannot open " + path);
    const size_t wrote = fwrite(out.data(), 1, out.size(), f);
    fclose(f);
    if (wrote != out.size()) throw std::runtime_error("Film::saveHDR: short write to " + path);
}

std::string Film::formatName(const std::string& fmt, int id) {
    std::string out;
    bool used = false;
    for (size_t i = 0; i < fmt.size(); i++) {
        if (fmt[i] != '%') {
            out += fmt[i];
            continue;
        }
        if (i + 1 < fmt.size() && fmt[i + 1] == '%') {
            out += '%';
            i++;
            continue;
        }
        size_t j = i + 1;
        while (j < fmt.size() && strchr("-+ 0#", fmt[j])) j++;
        while (j < fmt.size() && fmt[j] >= '0' && fmt[j] <= '9') j++;
        if (used || j >= fmt.size() || !strchr("diu", fmt[j]) || j - i > 8)
            throw ParamError("Film: output name \"" + fmt + "\" needs at most one %d-style conversion");
        char buf[64];
        const std::string spec = fmt.substr(i, j - i + 1);
        snprintf(buf, sizeof(buf), spec.c_str(), id);
        out += buf;
        used = true;
        i = j;
    }
    return out;
}

std::string Film::save(int id) const {
    const std::string path = formatName(filename_, id) + ".hdr";
    saveHDR(path);
    return path;
}

PerspectiveCamera::PerspectiveCamera(const Transform& toWorld, double fovDeg, const std::shared_ptr<Film>& film,
                                     double apertureRadius, double focusDistance)
    : toWorld_(toWorld), fov_(fovDeg), aperture_(apertureRadius), focus_(focusDistance) {
    if (!film) throw ParamError("PerspectiveCamera: no film");
    film_ = film;
}
PerspectiveCamera::PerspectiveCamera(RenderParams& p)
    : PerspectiveCamera(p.getTransform("toWorld", true), p.getDouble("fov", true),
                        std::dynamic_pointer_cast<Film>(p.getObject("film", true)), p.getDouble("apertureRadius", 0.0, true),
                        p.getDouble("focusDistance", 50.0, true)) {}
spk_camera_desc PerspectiveCamera::describe() const {
    spk_camera_desc c;
    memset(&c, 0, sizeof(c));
    c.width = film_->width();
    c.height = film_->height();
    for (int i = 0; i < 16; i++) c.to_world[i] = toWorld_.m[i];
    c.fov_deg = fov_;
    c.aperture_radius = aperture_;
    c.focus_distance = focus_;
    return c;
}

// ---- materials ------------------------------------------------------------------------------
namespace {
spk_material blank(int type) {
    spk_material m;
    memset(&m, 0, sizeof(m));
    m.type = type;
    m.int_ior = 1.333;
    m.alpha = 0.1;
    return m;
}
void put(double* d, const Spectrum& s) {
    d[0] = s.r;
    d[1] = s.g;
    d[2] = s.b;
}
}  // namespace

Diffuse::Diffuse(RenderParams& p) : R_(p.getSpectrum("reflectance")) {}
spk_material Diffuse::describe() const {
    spk_material m = blank(SPK_MAT_DIFFUSE);
    put(m.reflectance, R_);
    return m;
}
Conductor::Conductor(RenderParams& p) : eta_(p.getSpectrum("eta")), k_(p.getSpectrum("k")) {}
spk_material Conductor::describe() const {
    spk_material m = blank(SPK_MAT_CONDUCTOR);
    put(m.eta, eta_);
    put(m.k, k_);
    return m;
}
Dielectric::Dielectric(RenderParams& p)
    : Kr_(p.getSpectrum("specularReflectance", Spectrum(1.0))), Kt_(p.getSpectrum("specularTransmittance", Spectrum(1.0))),
      ior_(p.getDouble("intIOR", 1.333)) {}
spk_material Dielectric::describe() const {
    spk_material m = blank(SPK_MAT_DIELECTRIC);
    put(m.reflectance, Kr_);
    put(m.transmittance, Kt_);
    m.int_ior = ior_;
    return m;
}
int microfacetFromName(const std::string& name) {
    if (name == "beckmann") return SPK_MICROFACET_BECKMANN;
    if (name == "ggx") return SPK_MICROFACET_GGX;
    throw ParamError("unknown microfacet distribution \"" + name + "\" (beckmann, ggx)");
}
RoughConductor::RoughConductor(RenderParams& p)
    : eta_(p.getSpectrum("eta")), k_(p.getSpectrum("k")), alpha_(p.getDouble("alpha", 0.1)),
      dist_(microfacetFromName(p.getString("distribution", std::string("beckmann"), true))) {}
spk_material RoughConductor::describe() const {
    spk_material m = blank(SPK_MAT_ROUGHCONDUCTOR);
    put(m.eta, eta_);
    put(m.k, k_);
    m.alpha = alpha_;
    m.distribution = dist_;
    return m;
}
Plastic::Plastic(RenderParams& p)
    : Kd_(p.getSpectrum("diffuseReflectance")), Ks_(p.getSpectrum("specularReflectance")),
      ior_(p.getDouble("intIOR", 1.5)) {}
spk_material Plastic::describe() const {
    spk_material m = blank(SPK_MAT_PLASTIC);
    put(m.reflectance, Kd_);
    put(m.specular, Ks_);
    m.int_ior = ior_;
    return m;
}
RoughDielectric::RoughDielectric(RenderParams& p)
    : Kr_(p.getSpectrum("specularReflectance")), Kt_(p.getSpectrum("specularTransmittance")),
      alpha_(p.getDouble("alpha", 0.1)), ior_(p.getDouble("intIOR", 1.3333)),
      dist_(microfacetFromName(p.getString("distribution", std::string("beckmann"), true))) {}
spk_material RoughDielectric::describe() const {
    spk_material m = blank(SPK_MAT_ROUGHDIELECTRIC);
    put(m.reflectance, Kr_);
    put(m.transmittance, Kt_);
    m.alpha = alpha_;
    m.int_ior = ior_;
    m.distribution = dist_;
    return m;
}
RoughPlastic::RoughPlastic(RenderParams& p)
    : Kd_(p.getSpectrum("diffuseReflectance")), Ks_(p.getSpectrum("specularReflectance")),
      alpha_(p.getDouble("alpha", 0.1)), ior_(p.getDouble("intIOR", 1.5)),
      dist_(microfacetFromName(p.getString("distribution", std::string("beckmann"), true))) {}
spk_material RoughPlastic::describe() const {
    spk_material m = blank(SPK_MAT_ROUGHPLASTIC);
    put(m.reflectance, Kd_);
    put(m.specular, Ks_);
    m.alpha = alpha_;
    m.int_ior = ior_;
    m.distribution = dist_;
    return m;
}

// ---- environment map ---------------------------------------------------------------------------
Envmap::Envmap(int width, int height, std::vector<float> rgb, const Transform& toWorld, double scale, const Point3d& center,
               double radius)
    : w_(width), h_(height), rgb_(std::move(rgb)), toWorld_(toWorld), scale_(scale), center_(center), radius_(radius) {
    if (width <= 0 || height <= 0 || rgb_.size() != (size_t)width * height * 3) throw ParamError("Envmap: bad image");
}
void Envmap::fill(spk_scene_desc& d) const {
    d.env_width = w_;
    d.env_height = h_;
    d.env_rgb = rgb_.data();
    for (int i = 0; i < 16; i++) d.env_to_world[i] = toWorld_.m[i];
    d.env_scale = scale_;
    d.env_center[0] = center_.x;
    d.env_center[1] = center_.y;
    d.env_center[2] = center_.z;
    d.env_radius = radius_;
}

// ---- samplers ------------------------------------------------------------------------------------
Independent::Independent(RenderParams& p) : Sampler(SPK_SAMPLER_INDEPENDENT, (unsigned)p.getInt("seed", 0)) {}
Sobol::Sobol(RenderParams& p) : Sampler(SPK_SAMPLER_SOBOL, (unsigned)p.getInt("seed", 0)) {}

// ---- integrator ------------------------------------------------------------------------------------
namespace {
void check(int rc, const char* what) {
    if (rc != SPK_OK) throw std::runtime_error(std::string(what) + ": " + spk_last_error());
}
}  // namespace

SamplerIntegrator::SamplerIntegrator(const std::shared_ptr<Sampler>& sampler, int integrator)
    : sampler_(sampler), integrator_(integrator) {
    if (!sampler_) sampler_ = std::make_shared<Independent>(0u);
}
SamplerIntegrator::~SamplerIntegrator() = default;

PathIntegrator::PathIntegrator(const std::shared_ptr<Sampler>& sampler) : SamplerIntegrator(sampler, SPK_INTEGRATOR_PATH) {}
PathIntegrator::PathIntegrator(RenderParams& p)
    : PathIntegrator(std::dynamic_pointer_cast<Sampler>(p.getObject("sampler", nullptr))) {}

DirectLightingIntegrator::DirectLightingIntegrator(const std::shared_ptr<Sampler>& sampler)
    : SamplerIntegrator(sampler, SPK_INTEGRATOR_DIRECT) {}
DirectLightingIntegrator::DirectLightingIntegrator(RenderParams& p)
    : DirectLightingIntegrator(std::dynamic_pointer_cast<Sampler>(p.getObject("sampler", nullptr))) {}

void SamplerIntegrator::render(const std::shared_ptr<const Camera>& camera, const Scene& scene, RenderParams& params) {
    if (!camera) throw ParamError("SamplerIntegrator::render: no camera");
    if (!camera->film()->filename().empty()) Film::formatName(camera->film()->filename(), 0);   // bad names fail first
    const int spp = params.getInt("sampleCount");
    const int maxDepth = params.getInt("maxDepth", 16);
    check(spk_init(params.getInt("device", 0)), "spk_init");

    // Flatten the scene: materials by object identity, one area light per emissive triangle.
    const auto& prims = scene.primitives();
    std::vector<float> verts(prims.size() * 9);
    std::vector<int32_t> mat_ids(prims.size()), light_ids(prims.size(), -1);
    std::vector<spk_material> mats;
    std::map<const SurfaceMaterial*, int> mat_index;
    std::vector<int32_t> area_tri;
    std::vector<double> area_rad;
    for (size_t i = 0; i < prims.size(); i++) {
        const auto& p = prims[i];
        if (!p->shape() || !p->material()) throw ParamError("SamplerIntegrator::render: primitive without shape or BSDF");
        for (int v = 0; v < 3; v++) {
            const Point3d& q = (*p->shape())[v];
            verts[9 * i + 3 * v] = (float)q.x;
            verts[9 * i + 3 * v + 1] = (float)q.y;
            verts[9 * i + 3 * v + 2] = (float)q.z;
        }
        auto it = mat_index.find(p->material().get());
        if (it == mat_index.end()) {
            it = mat_index.emplace(p->material().get(), (int)mats.size()).first;
            mats.push_back(p->material()->describe());
        }
        mat_ids[i] = it->second;
        if (p->light()) {
            light_ids[i] = (int32_t)area_tri.size();
            area_tri.push_back((int32_t)i);
            const Spectrum& L = p->light()->radiance();
            area_rad.insert(area_rad.end(), {L.r, L.g, L.b});
        }
    }
    spk_scene_desc d;
    memset(&d, 0, sizeof(d));
    d.num_tris = (int64_t)prims.size();
    d.verts = verts.data();
    d.normals = nullptr;
    d.mat_ids = mat_ids.data();
    d.light_ids = light_ids.data();
    d.num_materials = (int32_t)mats.size();
    d.materials = mats.data();
    d.num_area_lights = (int32_t)area_tri.size();
    d.area_tri = area_tri.data();
    d.area_radiance = area_rad.data();
    d.env_scale = 1.0;
    d.env_radius = 2.0;
    for (int i = 0; i < 16; i++) d.env_to_world[i] = (i % 5 == 0) ? 1.0 : 0.0;
    for (const auto& l : scene.lights())
        if (auto e = std::dynamic_pointer_cast<Envmap>(l)) e->fill(d);

    // "accelerator" (sceneparser.cc:64, default "bvh"): spica's SAH tree, or "lbvh",
    // a linear BVH built on the device.
    spk_scene_build_opts bo;
    check(spk_scene_build_opts_default(&bo), "spk_scene_build_opts_default");
    const std::string accel = params.getString("accelerator", std::string("bvh"));
    if (accel == "lbvh") bo.builder = SPK_BVH_LBVH;
    else if (accel != "bvh") throw ParamError("SamplerIntegrator::render: unknown accelerator \"" + accel + "\"");
    spk_scene* s = nullptr;
    check(spk_scene_create_ex(&d, &bo, &s), "spk_scene_create_ex");
    spk_camera_desc cd = camera->describe();
    spk_render_desc rd;
    memset(&rd, 0, sizeof(rd));
    rd.spp = spp;
    rd.max_depth = maxDepth;
    rd.sampler = sampler_->kind();
    rd.seed = sampler_->seed();
    rd.rr_depth = 3;
    rd.first_sample = 0;
    rd.integrator = integrator_;
    spk_render_opts o;
    memset(&o, 0, sizeof(o));
    o.queue_size = params.getInt("queueSize", 1 << 24);
    o.block_size = params.getInt("blockSize", 64);
    spk_render* r = nullptr;
    int rc = spk_render_create(s, &cd, &rd, &o, &r);
    if (rc != SPK_OK) {
        spk_scene_destroy(s);
        check(rc, "spk_render_create");
    }
    // "guiding" (GPU key, default 0): the learned proposal, trained online during the first
    // "guideTrainSpp" samples per pixel and MIS-combined with the BSDF.
    spk_guide* g = nullptr;
    if (params.getInt("guiding", 0)) {
        spk_guide_opts go;
        spk_guide_opts_default(&go);
        go.train_spp = params.getInt("guideTrainSpp", go.train_spp);
        go.seed = (uint32_t)params.getInt("guideSeed", (int)go.seed);
        rc = spk_guide_create(params.getInt("device", 0), &go, &g);
        if (rc == SPK_OK) rc = spk_render_set_guide(r, g);
        if (rc != SPK_OK) {
            spk_render_destroy(r);
            if (g) spk_guide_destroy(g);
            spk_scene_destroy(s);
            check(rc, "guiding");
        }
    }
    // spica saves after every pass (integrator.cc:100); "saveEvery" spaces the saves out.
    const bool saving = !camera->film()->filename().empty();
    const int every = std::max(1, std::min(spp, params.getInt("saveEvery", spp)));
    std::vector<double> sum((size_t)cd.width * cd.height * 3), cnt((size_t)cd.width * cd.height);
    std::vector<double>& img = camera->film()->image();
    std::string err;
    for (int done = 0; done < spp && rc == SPK_OK && err.empty();) {
        const int k = saving ? std::min(every, spp - done) : spp;
        rc = spk_render_run(r, k, 1);
        done += k;
        if (rc == SPK_OK && (saving || done == spp)) {
            rc = spk_render_film(r, sum.data(), nullptr, cnt.data());
            // Film::save's division (film.cc:23-40)
            for (size_t i = 0; rc == SPK_OK && i < cnt.size(); i++)
                for (int c = 0; c < 3; c++) img[3 * i + c] = sum[3 * i + c] / (cnt[i] + 1e-12);
            if (rc == SPK_OK && saving) {
                try {
                    camera->film()->save(done);
                } catch (const std::exception& e) {
                    err = e.what();
                }
            }
        }
    }
    if (rc == SPK_OK) rc = spk_render_stats_get(r, &stats_);
    spk_render_destroy(r);
    if (g) spk_guide_destroy(g);
    spk_scene_destroy(s);
    check(rc, "render");
    if (!err.empty()) throw std::runtime_error(err);
}

}  // namespace spica

extern "C" {
#ifdef SPK_PLUGIN_DIRECTLIGHTING
spica::CObject* createInstance(spica::RenderParams& params) { return new spica::DirectLightingIntegrator(params); }
const char* getDescription() { return "Integrate only direct lighting (MI355X wavefront)"; }
#else
spica::CObject* createInstance(spica::RenderParams& params) { return new spica::PathIntegrator(params); }
const char* getDescription() { return "Path tracing integrator (MI355X wavefront)"; }
#endif
}
