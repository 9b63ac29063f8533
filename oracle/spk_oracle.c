/*
 * spk_oracle.c -- CPU checker for the MI355X path tracer.  TEST INFRASTRUCTURE ONLY.
 *
 * A plain-C, double-precision restatement of spica's PathIntegrator path (and of the
 * DirectLightingIntegrator next to it): the
 * estimator, constants, sampling order, tie rules and edge rules of the reference,
 * written from its sources (each function cites the file:line it follows, paths
 * relative to spica's `sources/`).  Only the tests, the smoke check and the CPU
 * baseline leg of the benchmark load this library; the product never does.
 *
 * Pinning (what fixes each part independently of this file's own formulas) is
 * listed in the tests: closed-form KATs of spica's test_geometry.cc, the box.ply
 * known answer, brute force for the BVH, MT19937's standard known answer, BSDF
 * quadrature / reciprocity / chi-square, furnace closed forms for the integrator,
 * and path-by-path agreement with spica's own compiled sources (oracle/_ref).
 */
#include "../include/spk_types.h"

#include <float.h>
#include <math.h>
#include <pthread.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define EXPORT __attribute__((visibility("default")))

/* core/common.h:52-55 */
static const double PI = 3.14159265358979323846;
static const double INV_PI = 1.0 / 3.14159265358979323846;
static const double INFTY = 1.0e32;
static const double EPS = 1.0e-12;

/* ------------------------------------------------------------------------- */
/* Vectors (core/vector3d.h, core/vect_math.h)                               */
/* ------------------------------------------------------------------------- */
typedef struct { double x, y, z; } V3;

static inline V3 v3(double x, double y, double z) { V3 r = {x, y, z}; return r; }
static inline V3 vadd(V3 a, V3 b) { return v3(a.x + b.x, a.y + b.y, a.z + b.z); }
static inline V3 vsub(V3 a, V3 b) { return v3(a.x - b.x, a.y - b.y, a.z - b.z); }
static inline V3 vmul(V3 a, double s) { return v3(a.x * s, a.y * s, a.z * s); }
static inline V3 vneg(V3 a) { return v3(-a.x, -a.y, -a.z); }
static inline V3 vcmul(V3 a, V3 b) { return v3(a.x * b.x, a.y * b.y, a.z * b.z); }
static inline double vdot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
static inline double vabsdot(V3 a, V3 b) { return fabs(vdot(a, b)); }
/* vect_math.h:38-46 */
static inline V3 vcross(V3 a, V3 b) {
    return v3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
static inline double vnorm(V3 a) { return sqrt(a.x * a.x + a.y * a.y + a.z * a.z); }
/* vect_math.h:48-53: divide each component by the norm */
static inline V3 vnormalize(V3 a) {
    double n = sqrt(a.x * a.x + a.y * a.y + a.z * a.z);
    return v3(a.x / n, a.y / n, a.z / n);
}
/* Vector3d/Normal3d::normalized(): multiply by the reciprocal of the norm
 * (vector3d_detail.h:196-200, 134-137; normal3d_detail.h:94-97, 132-134) */
static inline V3 vnormalized(V3 a) {
    double inv = 1.0 / sqrt(a.x * a.x + a.y * a.y + a.z * a.z);
    return v3(a.x * inv, a.y * inv, a.z * inv);
}
static inline double vcomp(V3 a, int i) { return i == 0 ? a.x : (i == 1 ? a.y : a.z); }
static inline int vis_black(V3 a) { return a.x == 0.0 && a.y == 0.0 && a.z == 0.0; }
/* spectrum.cc:677-679 */
static inline double gray(V3 c) { return 0.2126 * c.x + 0.7152 * c.y + 0.0722 * c.z; }
static inline double clampd(double v, double lo, double hi) { return v < lo ? lo : (v > hi ? hi : v); }
/* spectrum.cc:711-713 with defaults lo = 0, hi = INFTY (spectrum.h:78-80) */
static inline V3 spec_clamp(V3 c) {
    return v3(fmax(0.0, fmin(c.x, INFTY)), fmax(0.0, fmin(c.y, INFTY)), fmax(0.0, fmin(c.z, INFTY)));
}
/* vect_math.h:121-128 */
static void coordinate_system(V3 w, V3* u, V3* v) {
    if (fabs(w.x) > fabs(w.y)) {
        *u = vnormalize(v3(-w.z, 0.0, w.x));
    } else {
        *u = vnormalize(v3(0.0, w.z, -w.y));
    }
    *v = vnormalize(vcross(w, *u));
}

/* ------------------------------------------------------------------------- */
/* float.h:23-45 -- spica's one-ulp nudges (note: zero goes down for Up and   */
/* up for Down, because the sign test is "d > 0").                            */
/* ------------------------------------------------------------------------- */
static double next_float_up(double d) {
    if (isinf(d) && d > 0.0) return d;
    if (d == -0.0) d = 0.0;
    uint64_t ui;
    memcpy(&ui, &d, 8);
    if (d > 0.0) ui++; else ui--;
    memcpy(&d, &ui, 8);
    return d;
}
static double next_float_down(double d) {
    if (isinf(d) && d > 0.0) return d;
    if (d == -0.0) d = 0.0;
    uint64_t ui;
    memcpy(&ui, &d, 8);
    if (d > 0.0) ui--; else ui++;
    memcpy(&d, &ui, 8);
    return d;
}

/* interaction.cc:16-30 */
static V3 offset_ray_origin(V3 p, V3 n, V3 w) {
    V3 offset = vmul(n, 1.0e-3);
    if (vdot(w, n) < 0.0) offset = vneg(offset);
    V3 po = vadd(p, offset);
    double r[3];
    for (int i = 0; i < 3; i++) {
        if (vcomp(offset, i) > 0.0) r[i] = next_float_up(vcomp(po, i));
        else r[i] = next_float_down(vcomp(po, i));
    }
    return v3(r[0], r[1], r[2]);
}

/* ------------------------------------------------------------------------- */
/* Ray (ray.cc:12-47)                                                         */
/* ------------------------------------------------------------------------- */
typedef struct { V3 o, d, inv; double maxd; } Ray;

static Ray make_ray(V3 o, V3 d, double maxd) {
    Ray r;
    r.o = o;
    r.d = vnormalized(d);
    r.inv.x = (r.d.x == 0.0) ? INFTY : 1.0 / r.d.x;
    r.inv.y = (r.d.y == 0.0) ? INFTY : 1.0 / r.d.y;
    r.inv.z = (r.d.z == 0.0) ? INFTY : 1.0 / r.d.z;
    r.maxd = maxd;
    return r;
}

/* ------------------------------------------------------------------------- */
/* Triangle (triangle.cc:17-74, 98-197, 212-216)                             */
/* ------------------------------------------------------------------------- */
typedef struct {
    V3 p[3];
    V3 n[3];
    V3 fn;
    double area;
} Tri;

static void tri_init(Tri* t, const float* v, const float* nrm) {
    for (int i = 0; i < 3; i++) t->p[i] = v3(v[3 * i], v[3 * i + 1], v[3 * i + 2]);
    V3 e1 = vsub(t->p[1], t->p[0]);
    V3 e2 = vsub(t->p[2], t->p[0]);
    int zero = nrm != NULL;
    for (int i = 0; nrm && i < 9; i++) zero = zero && nrm[i] == 0.0f;
    if (!nrm || zero) {
        /* triangle.cc:17-31 (three zero normals: a flat triangle among smooth ones) */
        t->fn = vnormalized(vcross(e1, e2));
        t->n[0] = t->n[1] = t->n[2] = t->fn;
    } else {
        /* triangle.cc:33-52 */
        for (int i = 0; i < 3; i++) t->n[i] = vnormalized(v3(nrm[3 * i], nrm[3 * i + 1], nrm[3 * i + 2]));
        V3 fn = vcross(e1, e2);
        if (vnorm(fn) < EPS) fn = vmul(vadd(vadd(t->n[0], t->n[1]), t->n[2]), 1.0 / 3.0);
        t->fn = vnormalize(fn);
    }
    t->area = 0.5 * vnorm(vcross(e1, e2));
}

/* triangle.cc:98-117 / 159-178: Moller-Trumbore, |det| < EPS rejected,
 * accepted when EPS < t <= maxDist (equal t: the later test wins). */
static int tri_intersect(const Tri* t, const Ray* r, double* th, double* uo, double* vo) {
    V3 e1 = vsub(t->p[1], t->p[0]);
    V3 e2 = vsub(t->p[2], t->p[0]);
    V3 pv = vcross(r->d, e2);
    double det = vdot(e1, pv);
    if (det > -EPS && det < EPS) return 0;
    double invdet = 1.0 / det;
    V3 tv = vsub(r->o, t->p[0]);
    double u = vdot(tv, pv) * invdet;
    if (u < 0.0 || u > 1.0) return 0;
    V3 qv = vcross(tv, e1);
    double v = vdot(r->d, qv) * invdet;
    if (v < 0.0 || u + v > 1.0) return 0;
    double tt = vdot(e2, qv) * invdet;
    if (tt <= EPS || tt > r->maxd) return 0;
    *th = tt;
    if (uo) *uo = u;
    if (vo) *vo = v;
    return 1;
}

/* ------------------------------------------------------------------------- */
/* Surface interaction (triangle.cc:119-156, interaction.cc:110-175)          */
/* ------------------------------------------------------------------------- */
typedef struct {
    V3 pos;
    V3 n;        /* geometric normal, normalize(cross(dpdu, dpdv)) */
    V3 ns, ts, bs; /* shading normal, shading.dpdu, shading.dpdv */
    V3 wo;
    int prim;
} Isect;

static void make_isect(const Tri* t, const Ray* r, double th, double u, double v, int prim, Isect* is) {
    is->pos = vadd(r->o, vmul(r->d, th));
    /* no uvs: detUV == 0 -> coordinateSystem(faceNormal) (triangle.cc:127-129) */
    V3 dpdu, dpdv;
    coordinate_system(t->fn, &dpdu, &dpdv);
    is->n = vnormalize(vcross(dpdu, dpdv));
    is->ns = is->n;
    is->ts = dpdu;
    is->bs = dpdv;
    is->wo = vneg(r->d);
    is->prim = prim;
    /* triangle.cc:141-155 */
    V3 ns = vnormalize(vadd(vadd(vmul(t->n[0], 1.0 - u - v), vmul(t->n[1], u)), vmul(t->n[2], v)));
    if (fabs(vdot(ns, t->fn)) < 1.0 - EPS) {
        V3 ss = vnormalize(vcross(ns, t->fn));
        V3 ts = vnormalize(vcross(ns, ss));
        is->ns = vnormalize(vcross(ss, ts)); /* interaction.cc:168-175 */
        is->ts = ss;
        is->bs = ts;
    }
}

/* ------------------------------------------------------------------------- */
/* Materials -> single-lobe BSDF (bsdfs/, core/bsdf.cc, core/bxdf.cc)          */
/* ------------------------------------------------------------------------- */
enum { BX_REFLECTION = 1, BX_TRANSMISSION = 2, BX_DIFFUSE = 4, BX_GLOSSY = 8, BX_SPECULAR = 16, BX_ALL = 31 };
enum {
    LOBE_NONE = 0, LOBE_LAMBERT, LOBE_SPECREFL, LOBE_FRESNELSPEC, LOBE_MICROREFL,
    LOBE_PLASTIC,      /* PlasticBRDF (plastic.cc:20-87) */
    LOBE_MICROTRANS,   /* MicrofacetTransmission (bxdf.cc:301-423) */
    LOBE_ROUGHPLASTIC  /* RoughPlasticBRDF (roughplastic.cc:21-101) */
};

typedef struct {
    int lobe;       /* LOBE_* ; LOBE_NONE = BSDF with zero BxDFs */
    int type;       /* BxDF type bits */
    V3 R;           /* Lambert reflectance / specular ref_ / microfacet ref_ (1) / re_ / Kd */
    V3 T;           /* FresnelSpecular and MicrofacetTransmission tr_ */
    V3 Ks;          /* (rough) plastic Ks */
    double specW;   /* (rough) plastic specularSamplingWeight_ */
    double etaA, etaB;
    V3 cEta, cK;    /* FresnelConductor(etaI = 1, etaT = eta, k) */
    double ax, ay;  /* microfacet alphas */
    int dist;       /* SPK_MICROFACET_GGX / SPK_MICROFACET_BECKMANN */
    V3 n, t, b;     /* BSDF frame (bsdf.cc:14-21) */
} BSDF;

/* The hidden uniform draws of MicrofacetTransmission::sample (bxdf.cc:364),
 * PlasticBRDF::sample (plastic.cc:54) and RoughPlasticBRDF::sample (roughplastic.cc:58):
 * spica takes them from a thread_local Random per source file, default-seeded (0); with
 * the MT19937 sampler they come from three such streams, otherwise from one sampler
 * dimension the caller sets in `u`. */
enum { HID_BXDF = 0, HID_PLASTIC = 1, HID_ROUGHPLASTIC = 2 };
typedef struct MT MT;
typedef struct {
    int use_mt;
    MT* mt[3];
    double u;
} Hidden;
static double hidden_draw(Hidden* h, int stream);

/* bsdfs/diffuse.cc:23-32, conductor.cc:28-38, dielectric.cc:30-42, roughconductor.cc:38-75 */
static void setup_bsdf(const spk_material* m, const Isect* is, BSDF* b) {
    memset(b, 0, sizeof(*b));
    b->n = vnormalized(is->ns);
    b->t = vnormalized(is->ts);
    b->b = vnormalized(is->bs);
    switch (m->type) {
    case SPK_MAT_DIFFUSE: {
        V3 r = spec_clamp(v3(m->reflectance[0], m->reflectance[1], m->reflectance[2]));
        if (!vis_black(r)) {
            b->lobe = LOBE_LAMBERT;
            b->type = BX_REFLECTION | BX_DIFFUSE;
            b->R = r;
        }
        break;
    }
    case SPK_MAT_CONDUCTOR:
        b->lobe = LOBE_SPECREFL;
        b->type = BX_REFLECTION | BX_SPECULAR;
        b->R = v3(1, 1, 1);
        b->cEta = v3(m->eta[0], m->eta[1], m->eta[2]);
        b->cK = v3(m->k[0], m->k[1], m->k[2]);
        break;
    case SPK_MAT_DIELECTRIC: {
        V3 ior = v3(m->int_ior, m->int_ior, m->int_ior);
        double eta = gray(ior);
        V3 re = spec_clamp(v3(m->reflectance[0], m->reflectance[1], m->reflectance[2]));
        V3 tr = spec_clamp(v3(m->transmittance[0], m->transmittance[1], m->transmittance[2]));
        if (vis_black(re) && vis_black(tr)) break;
        b->lobe = LOBE_FRESNELSPEC;
        b->type = BX_REFLECTION | BX_TRANSMISSION | BX_SPECULAR;
        b->R = re;
        b->T = tr;
        b->etaA = 1.0;
        b->etaB = eta;
        break;
    }
    case SPK_MAT_ROUGHCONDUCTOR: {
        V3 al = v3(m->alpha, m->alpha, m->alpha);
        double a = gray(al);
        b->cEta = v3(m->eta[0], m->eta[1], m->eta[2]);
        b->cK = v3(m->k[0], m->k[1], m->k[2]);
        b->R = v3(1, 1, 1);
        b->dist = m->distribution;
        if (a == 0.0) {
            b->lobe = LOBE_SPECREFL;
            b->type = BX_REFLECTION | BX_SPECULAR;
        } else {
            b->lobe = LOBE_MICROREFL;
            b->type = BX_REFLECTION | BX_GLOSSY;
            b->ax = a;
            b->ay = a;
        }
        break;
    }
    case SPK_MAT_PLASTIC: {  /* plastic.cc:111-121: no clamping */
        V3 ior = v3(m->int_ior, m->int_ior, m->int_ior);
        b->lobe = LOBE_PLASTIC;
        b->type = BX_REFLECTION | BX_SPECULAR | BX_DIFFUSE;
        b->R = v3(m->reflectance[0], m->reflectance[1], m->reflectance[2]);
        b->Ks = v3(m->specular[0], m->specular[1], m->specular[2]);
        b->etaA = 1.0;
        b->etaB = gray(ior);
        b->specW = gray(b->Ks) / (gray(b->R) + gray(b->Ks));
        break;
    }
    case SPK_MAT_ROUGHDIELECTRIC: {  /* roughdielectric.cc:42-82 */
        V3 ior = v3(m->int_ior, m->int_ior, m->int_ior);
        V3 al = v3(m->alpha, m->alpha, m->alpha);
        double eta = gray(ior), a = gray(al);
        V3 re = spec_clamp(v3(m->reflectance[0], m->reflectance[1], m->reflectance[2]));
        V3 tr = spec_clamp(v3(m->transmittance[0], m->transmittance[1], m->transmittance[2]));
        if (vis_black(re) && vis_black(tr)) break;
        b->R = re;
        b->T = tr;
        b->etaA = 1.0;
        b->etaB = eta;
        if (a == 0.0) {
            b->lobe = LOBE_FRESNELSPEC;
            b->type = BX_REFLECTION | BX_TRANSMISSION | BX_SPECULAR;
        } else {
            b->lobe = LOBE_MICROTRANS;
            b->type = BX_REFLECTION | BX_TRANSMISSION | BX_GLOSSY;
            b->ax = a;
            b->ay = a;
            b->dist = m->distribution;
        }
        break;
    }
    case SPK_MAT_ROUGHPLASTIC: {  /* roughplastic.cc:128-158 */
        V3 ior = v3(m->int_ior, m->int_ior, m->int_ior);
        V3 al = v3(m->alpha, m->alpha, m->alpha);
        V3 kd = v3(m->reflectance[0], m->reflectance[1], m->reflectance[2]);
        V3 ks = v3(m->specular[0], m->specular[1], m->specular[2]);
        double rough = fmax(0.001, gray(al));
        if (vis_black(ks)) {
            b->lobe = LOBE_LAMBERT;
            b->type = BX_REFLECTION | BX_DIFFUSE;
            b->R = kd;
            break;
        }
        b->lobe = LOBE_ROUGHPLASTIC;
        b->type = BX_REFLECTION | BX_DIFFUSE | BX_GLOSSY;
        b->R = kd;
        b->Ks = ks;
        b->etaA = 1.0;
        b->etaB = gray(ior);
        b->specW = gray(ks) / (gray(kd) + gray(ks));
        b->ax = rough;
        b->ay = rough;
        b->dist = m->distribution;
        break;
    }
    default:
        break;
    }
}

/* fresnel.cc:40-63 */
static V3 fr_conductor(double cosI, V3 etaI, V3 etaT, V3 k) {
    cosI = clampd(cosI, -1.0, 1.0);
    double r[3];
    for (int c = 0; c < 3; c++) {
        double eta = vcomp(etaT, c) / vcomp(etaI, c);
        double etak = vcomp(k, c) / vcomp(etaI, c);
        double cos2 = cosI * cosI;
        double sin2 = 1.0 - cos2;
        double eta2 = eta * eta;
        double etak2 = etak * etak;
        double t0 = eta2 - etak2 - sin2;
        double a2b2 = sqrt(t0 * t0 + 4.0 * eta2 * etak2);
        double t1 = a2b2 + cos2;
        double a = sqrt(0.5 * (a2b2 + t0));
        double t2 = 2.0 * cosI * a;
        double Rs = (t1 - t2) / (t1 + t2);
        double t3 = cos2 * a2b2 + sin2 * sin2;
        double t4 = t2 * sin2;
        double Rp = Rs * (t3 - t4) / (t3 + t4 + EPS);
        r[c] = 0.5 * (Rp + Rs);
    }
    return v3(r[0], r[1], r[2]);
}

/* fresnel.cc:65-85 */
static double fr_dielectric(double cosI, double etaI, double etaT) {
    cosI = clampd(cosI, -1.0, 1.0);
    if (!(cosI > 0.0)) {
        double tmp = etaI;
        etaI = etaT;
        etaT = tmp;
        cosI = fabs(cosI);
    }
    double sinI = sqrt(fmax(0.0, 1.0 - cosI * cosI));
    double sinT = etaI / etaT * sinI;
    if (sinT >= 1.0) return 1.0;
    double cosT = sqrt(fmax(0.0, 1.0 - sinT * sinT));
    double rpar = ((etaT * cosI) - (etaI * cosT)) / ((etaT * cosI) + (etaI * cosT));
    double rper = ((etaI * cosI) - (etaT * cosT)) / ((etaI * cosI) + (etaT * cosT));
    return 0.5 * (rpar * rpar + rper * rper);
}

/* vect_math.h: trig helpers on local directions */
static inline double cos_theta(V3 w) { return w.z; }
static inline double cos2_theta(V3 w) { return w.z * w.z; }
static inline double sin2_theta(V3 w) { return fmax(0.0, 1.0 - cos2_theta(w)); }
static inline double sin_theta(V3 w) { return sqrt(sin2_theta(w)); }
static inline double tan_theta(V3 w) { return sin_theta(w) / (cos_theta(w) + EPS); }
static inline double tan2_theta(V3 w) { return sin2_theta(w) / (cos2_theta(w) + EPS); }
static inline double cos_phi(V3 w) {
    double s = sin_theta(w);
    return (s == 0.0) ? 1.0 : clampd(w.x / s, -1.0, 1.0);
}
static inline double sin_phi(V3 w) {
    double s = sin_theta(w);
    return (s == 0.0) ? 0.0 : clampd(w.y / s, -1.0, 1.0);
}
static inline double cos2_phi(V3 w) { double c = cos_phi(w); return c * c; }
static inline double sin2_phi(V3 w) { double s = sin_phi(w); return s * s; }
static inline int same_hemi(V3 a, V3 b) { return a.z * b.z > 0.0; }
/* vect_math.h:135-140 */
static inline V3 reflect_v(V3 wo, V3 n) { return vadd(vneg(wo), vmul(n, 2.0 * vdot(n, wo))); }
/* vect_math.h:148-158 */
static int refract_v(V3 wi, V3 n, double eta, V3* wt) {
    double cosI = vdot(n, wi);
    double sin2I = fmax(0.0, 1.0 - cosI * cosI);
    double sin2T = eta * eta * sin2I;
    if (sin2T >= 1.0) return 0;
    double cosT = sqrt(1.0 - sin2T);
    *wt = vadd(vmul(vneg(wi), eta), vmul(n, eta * cosI - cosT));
    return 1;
}
static inline V3 faceforward(V3 n, V3 v) { return vdot(n, v) < 0.0 ? vneg(n) : n; }

/* microfacet.cc:205-213 */
static double ggx_D(const BSDF* b, V3 wh) {
    double t2 = tan2_theta(wh);
    if (isinf(t2)) return 0.0;
    double c4 = cos2_theta(wh) * cos2_theta(wh);
    double e = (cos2_phi(wh) / (b->ax * b->ax) + sin2_phi(wh) / (b->ay * b->ay)) * t2;
    return 1.0 / (PI * b->ax * b->ay * c4 * (1.0 + e) * (1.0 + e));
}
/* microfacet.cc:257-266 */
static double ggx_lambda(const BSDF* b, V3 w) {
    double at = fabs(tan_theta(w));
    if (isinf(at)) return 0.0;
    double alpha = sqrt(cos2_phi(w) * b->ax * b->ax + sin2_phi(w) * b->ay * b->ay);
    double a = alpha * at;
    double s = a >= 0.0 ? 1.0 : -1.0;
    return (-1.0 + s * sqrt(1.0 + a * a)) / 2.0;
}
/* microfacet.cc:174-185 */
static double ggx_G1(const BSDF* b, V3 w, V3 wh) {
    if (w.z * vdot(w, wh) <= 0.0) return 0.0;
    return 1.0 / (1.0 + ggx_lambda(b, w));
}
static double ggx_G(const BSDF* b, V3 wo, V3 wi, V3 wh) {
    if (wi.z * vdot(wi, wh) <= 0.0) return 0.0;
    if (wo.z * vdot(wo, wh) <= 0.0) return 0.0;
    return 1.0 / (1.0 + ggx_lambda(b, wo) + ggx_lambda(b, wi));
}
/* microfacet.cc:187-193 (visible-normal pdf) */
static double ggx_pdf(const BSDF* b, V3 wo, V3 wh) {
    return ggx_D(b, wh) * ggx_G1(b, wo, wh) * vabsdot(wo, wh) / fabs(cos_theta(wo));
}
/* microfacet.cc:15-56 (visible-slope sampling with a rational fit for the y slope) */
static void ggx_sample11(double cosT, double u0, double u1, double* sx, double* sy) {
    if (cosT > 0.9999) {
        double r = sqrt(u0 / (1.0 - u0));
        double phi = 2.0 * PI * u1;
        *sx = r * cos(phi);
        *sy = r * sin(phi);
        return;
    }
    double sinT = sqrt(fmax(0.0, 1.0 - cosT * cosT));
    double tanT = sinT / cosT;
    double a = 1.0 / tanT;
    double G1 = 2.0 / (1.0 + sqrt(1.0 + 1.0 / (a * a)));
    double A = 2.0 * u0 / G1 - 1.0;
    double tmp = 1.0 / (A * A - 1.0);
    if (tmp > 1.0e10) tmp = 1.0e10;
    double B = tanT;
    double D = sqrt(fmax(0.0, B * B * tmp * tmp - (A * A - B * B) * tmp));
    double x1 = B * tmp - D;
    double x2 = B * tmp + D;
    *sx = (A < 0.0 || x2 > 1.0 / tanT) ? x1 : x2;
    double S, U;
    if (u1 > 0.5) {
        S = 1.0;
        U = 2.0 * (u1 - 0.5);
    } else {
        S = -1.0;
        U = 2.0 * (0.5 - u1);
    }
    double z = (U * (U * (U * 0.27385 - 0.73369) + 0.46341)) /
               (U * (U * (U * 0.093073 + 0.309420) - 1.0) + 0.597999);
    *sy = S * z * sqrt(1.0 + (*sx) * (*sx));
}
/* microfacet.cc:58-78 */
static V3 ggx_sample_tr(V3 wi, double ax, double ay, double u0, double u1) {
    V3 ws = vnormalized(v3(ax * wi.x, ay * wi.y, wi.z));
    double sx, sy;
    ggx_sample11(cos_theta(ws), u0, u1, &sx, &sy);
    double cp = cos_phi(ws), sp = sin_phi(ws);
    double tmp = cp * sx - sp * sy;
    sy = sp * sx + cp * sy;
    sx = tmp;
    sx *= ax;
    sy *= ay;
    return vnormalized(v3(-sx, -sy, 1.0));
}
/* microfacet.cc:215-249 (sampleVisibleArea = true) */
static V3 ggx_sample_wh(const BSDF* b, V3 wo, double u0, double u1) {
    int flip = wo.z < 0.0;
    V3 wh = ggx_sample_tr(flip ? vneg(wo) : wo, b->ax, b->ay, u0, u1);
    if (flip) wh = vneg(wh);
    return wh;
}

/* math.h:14-42: erfinv, the single-precision coefficients of Giles, "Approximating the
 * erfinv function" (GPU Computing Gems, 2011), evaluated in double as spica does (the
 * float literals are rounded to float first). */
static double math_erfinv(double x) {
    double w, p;
    x = clampd(x, -0.99999, 0.99999);
    w = -log((1 - x) * (1 + x));
    if (w < 5) {
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
        w = sqrt(w) - 3;
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
/* math.h:44-65: erf by Abramowitz & Stegun 7.1.26 (float-rounded coefficients) */
static double math_erf(double x) {
    const double a1 = 0.254829592f, a2 = -0.284496736f, a3 = 1.421413741f;
    const double a4 = -1.453152027f, a5 = 1.061405429f, pp = 0.3275911f;
    int sign = x < 0 ? -1 : 1;
    x = fabs(x);
    double t = 1 / (1 + pp * x);
    double y = 1 - (((((a5 * t + a4) * t) + a3) * t + a2) * t + a1) * t * exp(-x * x);
    return sign * y;
}
/* microfacet.cc:277-288 */
static double bk_D(const BSDF* b, V3 wh) {
    double t2 = tan2_theta(wh);
    if (isinf(t2)) return 0.0;
    double c4 = cos2_theta(wh) * cos2_theta(wh);
    double ab2 = cos2_phi(wh) / (b->ax * b->ax) + sin2_phi(wh) / (b->ay * b->ay);
    return exp(-t2 * ab2) / (PI * b->ax * b->ay * c4);
}
/* microfacet.cc:342-355: Walter et al. 2007's rational fit of the Smith lambda */
static double bk_lambda(const BSDF* b, V3 w) {
    double at = fabs(tan_theta(w));
    if (isinf(at)) return 0.0;
    double alpha = sqrt(cos2_phi(w) * b->ax * b->ax + sin2_phi(w) * b->ay * b->ay);
    double a = 1.0 / (alpha * at);
    if (a >= 1.6) return 0.0;
    return (1.0 - 1.259 * a + 0.396 * a * a) / (3.535 * a + 2.181 * a * a);
}
/* microfacet.cc:80-131: visible slopes by Jakob 2014's bisection-Newton on C22 */
static void bk_sample11(double cosI, double U1, double U2, double* sx, double* sy) {
    if (cosI > 0.9999) {
        double r = sqrt(-log(1.0 - U1));
        double sp = sin(2.0 * PI * U2), cp = cos(2.0 * PI * U2);
        *sx = r * cp;
        *sy = r * sp;
        return;
    }
    double sinI = sqrt(fmax(0.0, 1.0 - cosI * cosI));
    double tanI = sinI / cosI;
    double cotI = 1.0 / tanI;
    double a = -1.0;
    double c = math_erf(cotI);
    double sx0 = fmax(U1, 1.0e-6);
    double thetaI = acos(cosI);
    double fit = 1.0 + thetaI * (-0.876 + thetaI * (0.4265 - 0.0594 * thetaI));
    double bb = c - (1.0 + c) * pow(1.0 - sx0, fit);
    const double SQRT_PI_INV = 1.0 / sqrt(PI);
    double norm = 1.0 / (1.0 + c + SQRT_PI_INV * tanI * exp(-cotI * cotI));
    for (int it = 0; it < 16; it++) {
        if (bb < a || c < bb) bb = 0.5 * (a + c);
        double xm = math_erfinv(bb);
        double value = norm * (1.0 + bb + SQRT_PI_INV * tanI * exp(-xm * xm)) - sx0;
        double deriv = norm * (1.0 - xm * tanI);
        if (fabs(value) < 1.0e-6) break;
        if (value > 0.0) c = bb;
        else a = bb;
        bb -= value / deriv;
    }
    *sx = math_erfinv(bb);
    *sy = math_erfinv(2.0 * fmax(U2, 1.0e-6) - 1.0);
}
/* microfacet.cc:133-157 */
static V3 bk_sample_tr(V3 wi, double ax, double ay, double u0, double u1) {
    V3 ws = vnormalize(v3(ax * wi.x, ay * wi.y, wi.z));
    double sx, sy;
    bk_sample11(cos_theta(ws), u0, u1, &sx, &sy);
    double cp = cos_phi(ws), sp = sin_phi(ws);
    double tmp = cp * sx - sp * sy;
    sy = sp * sx + cp * sy;
    sx = tmp;
    sx *= ax;
    sy *= ay;
    return vnormalize(v3(-sx, -sy, 1.0));
}

/* MicrofacetDistribution with the material's distribution (microfacet.cc:166-193) */
static double mf_D(const BSDF* b, V3 wh) { return b->dist == SPK_MICROFACET_BECKMANN ? bk_D(b, wh) : ggx_D(b, wh); }
static double mf_lambda(const BSDF* b, V3 w) {
    return b->dist == SPK_MICROFACET_BECKMANN ? bk_lambda(b, w) : ggx_lambda(b, w);
}
static double mf_G1(const BSDF* b, V3 w, V3 wh) {
    if (w.z * vdot(w, wh) <= 0.0) return 0.0;
    return 1.0 / (1.0 + mf_lambda(b, w));
}
static double mf_G(const BSDF* b, V3 wo, V3 wi, V3 wh) {
    if (wi.z * vdot(wi, wh) <= 0.0) return 0.0;
    if (wo.z * vdot(wo, wh) <= 0.0) return 0.0;
    return 1.0 / (1.0 + mf_lambda(b, wo) + mf_lambda(b, wi));
}
static double mf_pdf(const BSDF* b, V3 wo, V3 wh) {
    return mf_D(b, wh) * mf_G1(b, wo, wh) * vabsdot(wo, wh) / fabs(cos_theta(wo));
}
static V3 mf_sample_wh(const BSDF* b, V3 wo, double u0, double u1) {
    if (b->dist != SPK_MICROFACET_BECKMANN) return ggx_sample_wh(b, wo, u0, u1);
    int flip = wo.z < 0.0;
    V3 wh = bk_sample_tr(flip ? vneg(wo) : wo, b->ax, b->ay, u0, u1);
    if (flip) wh = vneg(wh);
    return wh;
}

/* fresnel.cc:87-95 */
static double fr_diffuse_reflectance(double eta) {
    if (eta >= 1.0) return -1.4399 / (eta * eta) + 0.7099 / eta + 0.6681 + 0.0636 * eta;
    return -0.4399f + 0.7099 / eta - 0.3319 / (eta * eta) + 0.0636 / (eta * eta * eta);
}

/* bxdf.cc:264-278 */
static V3 micro_f(const BSDF* b, V3 wo, V3 wi) {
    double cO = fabs(cos_theta(wo)), cI = fabs(cos_theta(wi));
    V3 wh = vadd(wi, wo);
    if (cI == 0.0 || cO == 0.0) return v3(0, 0, 0);
    if (wh.x == 0.0 && wh.y == 0.0 && wh.z == 0.0) return v3(0, 0, 0);
    wh = vnormalized(wh);
    V3 F = fr_conductor(fabs(vdot(wi, wh)), v3(1, 1, 1), b->cEta, b->cK);
    double s = mf_D(b, wh) * mf_G(b, wo, wi, wh) / (4.0 * cI * cO);
    return vmul(vcmul(b->R, F), s);
}

/* bxdf.cc:333-354: MicrofacetTransmission::f(wo, wi, wh) */
static V3 mtrans_f_wh(const BSDF* b, V3 wo, V3 wi, V3 wh) {
    double cO = cos_theta(wo), cI = cos_theta(wi);
    if (cO == 0.0 || cI == 0.0) return v3(0, 0, 0);
    if (wo.z * vdot(wo, wh) <= 0.0) return v3(0, 0, 0);
    if (wi.z * vdot(wi, wh) <= 0.0) return v3(0, 0, 0);
    double F = fr_dielectric(vdot(wo, wh), b->etaA, b->etaB);
    if (same_hemi(wo, wi)) return vmul(b->R, F * mf_D(b, wh) * mf_G(b, wo, wi, wh) / (4.0 * cO * cI));
    double eta = vdot(wo, wh) > 0.0 ? (b->etaB / b->etaA) : (b->etaA / b->etaB);
    double factor = 1.0 / eta;
    double sd = vdot(wo, wh) + eta * vdot(wi, wh);
    return vmul(b->T, (1.0 - F) * fabs(mf_D(b, wh) * mf_G(b, wo, wi, wh) * eta * eta * vabsdot(wi, wh) *
                                       vabsdot(wo, wh) * factor * factor / (cI * cO * sd * sd)));
}
/* bxdf.cc:408-423 */
static double mtrans_pdf_wh(const BSDF* b, V3 wo, V3 wi, V3 wh) {
    if (wo.z * vdot(wo, wh) <= 0.0) return 0.0;
    if (wi.z * vdot(wi, wh) <= 0.0) return 0.0;
    double F = fr_dielectric(vdot(wo, wh), b->etaA, b->etaB);
    if (same_hemi(wo, wi)) return F * mf_pdf(b, wo, wh) / (4.0 * vabsdot(wo, wh));
    double eta = vdot(wo, wh) > 0.0 ? (b->etaB / b->etaA) : (b->etaA / b->etaB);
    double sd = vdot(wo, wh) + eta * vdot(wi, wh);
    double dwh = fabs((eta * eta * vdot(wi, wh)) / (sd * sd));
    return (1.0 - F) * mf_pdf(b, wo, wh) * dwh;
}
/* bxdf.cc:315-331, 394-406: the half vector of a (wo, wi) pair, on the +z side */
static V3 mtrans_wh(const BSDF* b, V3 wo, V3 wi) {
    V3 wh;
    if (same_hemi(wo, wi)) {
        wh = vnormalize(vadd(wo, wi));
    } else {
        double eta = cos_theta(wo) > 0.0 ? (b->etaB / b->etaA) : (b->etaA / b->etaB);
        wh = vnormalize(vadd(wo, vmul(wi, eta)));
    }
    return wh.z >= 0.0 ? wh : vneg(wh);
}
/* plastic.cc:28 / roughplastic.cc:28: probability of the specular branch */
static double plastic_prob_spec(const BSDF* b, double Fo) {
    return (Fo * b->specW) / (Fo * b->specW + (1.0 - Fo) * (1.0 - b->specW));
}

/* BxDF::f for the lobe (bxdf.cc:56-58, 86-91, 171-192, 264-278) */
static V3 lobe_f(const BSDF* b, V3 wo, V3 wi) {
    const double deltaEps = 1.0e-4; /* bxdf.cc:19 */
    switch (b->lobe) {
    case LOBE_LAMBERT:
        return vmul(b->R, INV_PI);
    case LOBE_SPECREFL:
        if (vdot(v3(-wo.x, -wo.y, wo.z), wi) > 1.0 - deltaEps) return vmul(b->R, 1.0 / fabs(cos_theta(wi)));
        return v3(0, 0, 0);
    case LOBE_FRESNELSPEC: {
        double F = fr_dielectric(cos_theta(wo), b->etaA, b->etaB);
        if (same_hemi(wo, wi)) {
            if (vdot(v3(-wo.x, -wo.y, wo.z), wi) > 1.0 - deltaEps) return vmul(b->R, F / fabs(cos_theta(wi)));
        } else {
            int entering = cos_theta(wo) > 0.0;
            double etaI = entering ? b->etaA : b->etaB;
            double etaT = entering ? b->etaB : b->etaA;
            V3 wt;
            if (!refract_v(wo, faceforward(v3(0, 0, 1), wo), etaI / etaT, &wt)) return v3(0, 0, 0);
            if (vdot(wi, wt) > 1.0 - deltaEps) return vmul(b->T, (1.0 - F) / fabs(cos_theta(wi)));
        }
        return v3(0, 0, 0);
    }
    case LOBE_MICROREFL:
        return micro_f(b, wo, wi);
    case LOBE_PLASTIC: {  /* plastic.cc:32-46; the "specular" test compares wi with its own mirror image */
        double Fo = fr_dielectric(cos_theta(wo), b->etaA, b->etaB);
        if (vdot(wi, v3(-wi.x, -wi.y, wi.z)) > 1.0 - EPS) return vmul(b->Ks, Fo / fabs(cos_theta(wi)));
        double eta = b->etaB / b->etaA, invEta = 1.0 / eta;
        double Fi = fr_dielectric(cos_theta(wi), b->etaA, b->etaB);
        double Fdr = fr_diffuse_reflectance(eta);
        return vmul(b->R, (1.0 - Fo) * (1.0 - Fi) * invEta * invEta * INV_PI / (1.0 - Fdr));
    }
    case LOBE_MICROTRANS:  /* bxdf.cc:315-331 */
        if (cos_theta(wo) == 0.0 || cos_theta(wi) == 0.0) return v3(0, 0, 0);
        return mtrans_f_wh(b, wo, wi, mtrans_wh(b, wo, wi));
    case LOBE_ROUGHPLASTIC: {  /* roughplastic.cc:35-47 */
        double cO = cos_theta(wo), cI = cos_theta(wi);
        if (cO == 0.0 || cI == 0.0) return v3(0, 0, 0);
        if (!same_hemi(wo, wi)) return v3(0, 0, 0);
        V3 wh = vnormalize(vadd(wi, wo));
        double F = fr_dielectric(vdot(wo, wh), b->etaA, b->etaB);
        V3 spec = vmul(b->Ks, F * mf_D(b, wh) * mf_G(b, wo, wi, wh) / (4.0 * cO * cI));
        V3 diff = vmul(b->R, (1.0 - F) * INV_PI);
        return vadd(spec, diff);
    }
    default:
        return v3(0, 0, 0);
    }
}

/* BxDF::pdf for the lobe (bxdf.cc:42-44, 103-108, 227-248, 293-297) */
static double lobe_pdf(const BSDF* b, V3 wo, V3 wi) {
    const double deltaEps = 1.0e-4;
    switch (b->lobe) {
    case LOBE_LAMBERT:
        return wo.z * wi.z > 0.0 ? fabs(wi.z) * INV_PI : 0.0;
    case LOBE_SPECREFL:
        return vdot(v3(-wo.x, -wo.y, wo.z), wi) > 1.0 - deltaEps ? 1.0 : 0.0;
    case LOBE_FRESNELSPEC: {
        double F = fr_dielectric(cos_theta(wo), b->etaA, b->etaB);
        if (same_hemi(wo, wi)) {
            if (vdot(v3(-wo.x, -wo.y, wo.z), wi) > 1.0 - deltaEps) return F;
        } else {
            int entering = cos_theta(wo) > 0.0;
            double etaI = entering ? b->etaA : b->etaB;
            double etaT = entering ? b->etaB : b->etaA;
            V3 wt;
            if (!refract_v(wo, faceforward(v3(0, 0, 1), wo), etaI / etaT, &wt)) return 0.0;
            if (vdot(wi, wt) > 1.0 - deltaEps) return 1.0 - F;
        }
        return 0.0;
    }
    case LOBE_MICROREFL: {
        if (!same_hemi(wo, wi)) return 0.0;
        V3 wh = vnormalize(vadd(wo, wi));
        return mf_pdf(b, wo, wh) / (4.0 * vdot(wo, wh));
    }
    case LOBE_PLASTIC: {  /* plastic.cc:71-82 */
        double Fo = fr_dielectric(cos_theta(wo), b->etaA, b->etaB);
        double ps = plastic_prob_spec(b, Fo);
        if (vdot(wi, v3(-wi.x, -wi.y, wi.z)) > 1.0 - EPS) return ps;
        return (1.0 - ps) * fabs(cos_theta(wi)) * INV_PI;
    }
    case LOBE_MICROTRANS:  /* bxdf.cc:394-406 */
        return mtrans_pdf_wh(b, wo, wi, mtrans_wh(b, wo, wi));
    case LOBE_ROUGHPLASTIC: {  /* roughplastic.cc:88-97 (F-weighted, not the sampling weights) */
        if (!same_hemi(wo, wi)) return 0.0;
        V3 wh = vnormalize(vadd(wi, wo));
        double F = fr_dielectric(vdot(wo, wh), b->etaA, b->etaB);
        double sp = F * mf_pdf(b, wo, wh) / (4.0 * vabsdot(wo, wh));
        double dp = (1.0 - F) * fabs(cos_theta(wi)) * INV_PI;
        return sp + dp;
    }
    default:
        return 0.0;
    }
}

/* sampling.cc:151-179 */
static void concentric_disk(double u0, double u1, double* dx, double* dy) {
    double ox = 2.0 * u0 - 1.0, oy = 2.0 * u1 - 1.0;
    if (ox == 0.0 && oy == 0.0) {
        *dx = 0.0;
        *dy = 0.0;
        return;
    }
    double theta, r;
    if (fabs(ox) > fabs(oy)) {
        r = ox;
        theta = PI * (oy / ox) / 4.0;
    } else {
        r = oy;
        theta = (PI / 2.0) - PI * (ox / oy) / 4.0;
    }
    *dx = r * cos(theta);
    *dy = r * sin(theta);
}
static V3 cosine_hemisphere(double u0, double u1) {
    double dx, dy;
    concentric_disk(u0, u1, &dx, &dy);
    double z = sqrt(fmax(0.0, 1.0 - dx * dx - dy * dy));
    return v3(dx, dy, z);
}

/* BxDF::sample for the lobe (bxdf.cc:34-40, 93-101, 194-225, 280-291).
 * Returns f; *pdf stays 0 on failure (the caller zeroes it, bsdf.cc:244). */
static V3 lobe_sample(const BSDF* b, V3 wo, V3* wi, double u0, double u1, double* pdf, int* stype, Hidden* hu) {
    switch (b->lobe) {
    case LOBE_LAMBERT: {
        *wi = cosine_hemisphere(u0, u1);
        if (wo.z < 0.0) wi->z *= -1.0;
        *pdf = lobe_pdf(b, wo, *wi);
        return lobe_f(b, wo, *wi);
    }
    case LOBE_SPECREFL: {
        *wi = v3(-wo.x, -wo.y, wo.z);
        *pdf = 1.0;
        double c = cos_theta(*wi);
        V3 F = fr_conductor(fabs(c), v3(1, 1, 1), b->cEta, b->cK);
        return vmul(vcmul(F, b->R), 1.0 / fabs(c));
    }
    case LOBE_FRESNELSPEC: {
        double F = fr_dielectric(cos_theta(wo), b->etaA, b->etaB);
        if (u0 < F) {
            *wi = v3(-wo.x, -wo.y, wo.z);
            *stype = BX_SPECULAR | BX_REFLECTION;
            *pdf = F;
            return vmul(b->R, F / fabs(cos_theta(*wi)));
        }
        int entering = cos_theta(wo) > 0.0;
        double etaI = entering ? b->etaA : b->etaB;
        double etaT = entering ? b->etaB : b->etaA;
        if (!refract_v(wo, faceforward(v3(0, 0, 1), wo), etaI / etaT, wi)) return v3(0, 0, 0);
        V3 ft = vmul(b->T, 1.0 - F);
        ft = vmul(ft, (etaI * etaI) / (etaT * etaT));
        *stype = BX_SPECULAR | BX_TRANSMISSION;
        *pdf = 1.0 - F;
        return vmul(ft, 1.0 / fabs(cos_theta(*wi)));
    }
    case LOBE_MICROREFL: {
        if (wo.z == 0.0) return v3(0, 0, 0);
        V3 wh = mf_sample_wh(b, wo, u0, u1);
        *wi = reflect_v(wo, wh);
        if (!same_hemi(wo, *wi)) return v3(0, 0, 0);
        *pdf = mf_pdf(b, wo, wh) / (4.0 * vdot(wo, wh));
        return micro_f(b, wo, *wi);
    }
    case LOBE_PLASTIC: {  /* plastic.cc:48-69 */
        double Fo = fr_dielectric(cos_theta(wo), b->etaA, b->etaB);
        double ps = plastic_prob_spec(b, Fo);
        if (hidden_draw(hu, HID_PLASTIC) < ps) {
            *wi = v3(-wo.x, -wo.y, wo.z);
            *pdf = ps;
            return vmul(b->Ks, Fo / fabs(cos_theta(*wi)));
        }
        *wi = cosine_hemisphere(u0, u1);
        double eta = b->etaB / b->etaA, invEta = 1.0 / eta;
        double Fi = fr_dielectric(cos_theta(*wi), b->etaA, b->etaB);
        double Fdr = fr_diffuse_reflectance(eta);
        *pdf = (1.0 - ps) * fabs(cos_theta(*wi)) * INV_PI;
        return vmul(b->R, (1.0 - Fo) * (1.0 - Fi) * invEta * invEta * INV_PI / (1.0 - Fdr));
    }
    case LOBE_MICROTRANS: {  /* bxdf.cc:356-392 */
        if (wo.z == 0.0) return v3(0, 0, 0);
        V3 wh = mf_sample_wh(b, wo, u0, u1);
        double cO = vdot(wo, wh.z > 0.0 ? wh : vneg(wh));
        double F = fr_dielectric(cO, b->etaA, b->etaB);
        if (hidden_draw(hu, HID_BXDF) < F) {
            *wi = reflect_v(wo, wh);
            if (!same_hemi(wo, *wi)) return v3(0, 0, 0);
            *stype = BX_REFLECTION | BX_GLOSSY;
        } else {
            double eta = cO > 0.0 ? (b->etaA / b->etaB) : (b->etaB / b->etaA);
            if (!refract_v(wo, wh, eta, wi)) return v3(0, 0, 0);
            if (same_hemi(wo, *wi)) return v3(0, 0, 0);
            *stype = BX_TRANSMISSION | BX_GLOSSY;
        }
        V3 whp = wh.z >= 0.0 ? wh : vneg(wh);
        *pdf = mtrans_pdf_wh(b, wo, *wi, whp);
        return mtrans_f_wh(b, wo, *wi, whp);
    }
    case LOBE_ROUGHPLASTIC: {  /* roughplastic.cc:49-80: the diffuse branch reuses the wh sample's uniforms */
        if (wo.z == 0.0) return v3(0, 0, 0);
        V3 wh = mf_sample_wh(b, wo, u0, u1);
        double Fo = fr_dielectric(vdot(wo, wh), b->etaA, b->etaB);
        double ps = plastic_prob_spec(b, Fo);
        if (hidden_draw(hu, HID_ROUGHPLASTIC) < ps) {
            *wi = reflect_v(wo, wh);
            if (!same_hemi(wo, *wi)) return v3(0, 0, 0);
            *pdf = ps * mf_pdf(b, wo, wh) / (4.0 * vabsdot(wo, wh));
            double cO = cos_theta(wo), cI = cos_theta(*wi);
            return vmul(b->Ks, Fo * mf_D(b, wh) * mf_G(b, wo, *wi, wh) / (4.0 * cO * cI));
        }
        *wi = cosine_hemisphere(u0, u1);
        *pdf = (1.0 - ps) * fabs(cos_theta(*wi)) * INV_PI;
        double eta = b->etaB / b->etaA, invEta = 1.0 / eta;
        double Fi = fr_dielectric(cos_theta(*wi), b->etaA, b->etaB);
        double Fdr = fr_diffuse_reflectance(eta);
        return vmul(b->R, (1.0 - Fo) * (1.0 - Fi) * invEta * invEta * INV_PI / (1.0 - Fdr));
    }
    default:
        return v3(0, 0, 0);
    }
}

static inline V3 to_local(const BSDF* b, V3 v) { return v3(vdot(b->t, v), vdot(b->b, v), vdot(b->n, v)); }
static inline V3 to_world(const BSDF* b, V3 v) {
    return vadd(vadd(vmul(b->t, v.x), vmul(b->b, v.y)), vmul(b->n, v.z));
}
/* bsdf.cc:28-37 with a single BxDF */
static int bsdf_num_components(const BSDF* b, int mask) {
    if (b->lobe == LOBE_NONE) return 0;
    return ((b->type & mask) == b->type) ? 1 : 0;
}
/* bsdf.cc:49-63 */
static V3 bsdf_f(const BSDF* b, V3 woW, V3 wiW, int mask) {
    if (b->lobe == LOBE_NONE) return v3(0, 0, 0);
    V3 wo = to_local(b, woW), wi = to_local(b, wiW);
    int reflect = vdot(wiW, b->n) * vdot(woW, b->n) > 0.0;
    if ((b->type & mask) != 0 && ((reflect && (b->type & BX_REFLECTION)) || (!reflect && (b->type & BX_TRANSMISSION))))
        return lobe_f(b, wo, wi);
    return v3(0, 0, 0);
}
/* bsdf.cc:124-136 */
static double bsdf_pdf(const BSDF* b, V3 woW, V3 wiW, int mask) {
    if (b->lobe == LOBE_NONE) return 0.0;
    V3 wo = to_local(b, woW), wi = to_local(b, wiW);
    if ((b->type & mask) != 0) return lobe_pdf(b, wo, wi);
    return 0.0;
}
/* bsdf.cc:65-122 with a single BxDF.  *pdf is left untouched when wo.z == 0
 * (spica returns before zeroing it); callers initialise it to 0. */
static V3 bsdf_sample(const BSDF* b, V3 woW, V3* wiW, double u0, double u1, double* pdf, int mask, int* stype,
                      Hidden* hu) {
    int match = bsdf_num_components(b, mask);
    if (match == 0) {
        *pdf = 0.0;
        *stype = 0;
        return v3(0, 0, 0);
    }
    double ur0 = fmin(u0 * 1 - 0, 1.0 - EPS);
    V3 wo = vnormalized(to_local(b, woW));
    if (wo.z == 0.0) return v3(0, 0, 0);
    *pdf = 0.0;
    *stype = b->type;
    V3 wi = v3(0, 0, 0);
    V3 f = lobe_sample(b, wo, &wi, ur0, u1, pdf, stype, hu);
    if (*pdf == 0.0) {
        *stype = 0;
        return v3(0, 0, 0);
    }
    *wiW = vnormalized(to_world(b, wi));
    return f;
}

/* ------------------------------------------------------------------------- */
/* Spica-style BVH (accelerators/bvh.cc:139-237, 331-387)                     */
/* ------------------------------------------------------------------------- */
typedef struct { V3 lo, hi; } Box;

static Box box_empty(void) {
    Box b;
    b.lo = v3(DBL_MAX, DBL_MAX, DBL_MAX);
    b.hi = v3(-DBL_MAX, -DBL_MAX, -DBL_MAX);
    return b;
}
static Box box_merge(Box a, Box b) {
    Box r;
    r.lo = v3(fmin(a.lo.x, b.lo.x), fmin(a.lo.y, b.lo.y), fmin(a.lo.z, b.lo.z));
    r.hi = v3(fmax(a.hi.x, b.hi.x), fmax(a.hi.y, b.hi.y), fmax(a.hi.z, b.hi.z));
    return r;
}
static Box box_merge_pt(Box a, V3 p) {
    Box r;
    r.lo = v3(fmin(p.x, a.lo.x), fmin(p.y, a.lo.y), fmin(p.z, a.lo.z));
    r.hi = v3(fmax(p.x, a.hi.x), fmax(p.y, a.hi.y), fmax(p.z, a.hi.z));
    return r;
}
/* bounds3d_detail.h:108-114 */
static double box_area(Box b) {
    V3 d = vsub(b.hi, b.lo);
    return 2.0 * (fabs(d.x * d.y) + fabs(d.y * d.z) + fabs(d.z * d.x));
}
/* bounds3d_detail.h:55-61 */
static int box_max_extent(Box b) {
    V3 d = vsub(b.hi, b.lo);
    if (d.x >= d.y && d.x >= d.z) return 0;
    if (d.y >= d.z) return 1;
    return 2;
}
/* bounds3d_detail.h:63-79 */
static int box_hit(const Box* b, const Ray* r) {
    double t0 = 0.0, t1 = r->maxd;
    for (int i = 0; i < 3; i++) {
        double tt0 = (vcomp(b->lo, i) - vcomp(r->o, i)) * vcomp(r->inv, i);
        double tt1 = (vcomp(b->hi, i) - vcomp(r->o, i)) * vcomp(r->inv, i);
        if (tt0 > tt1) {
            double t = tt0;
            tt0 = tt1;
            tt1 = t;
        }
        t0 = fmax(t0, tt0);
        t1 = fmin(t1, tt1);
        if (t0 > t1) return 0;
    }
    return 1;
}

typedef struct { int prim; V3 c; Box b; } PrimInfo; /* bvh.h:15-26 */
typedef struct { Box b; int left, right, prim; } BNode;

typedef struct {
    BNode* nodes;
    int nnodes, cap;
    int maxdepth;
} BVH;

static int bvh_new_node(BVH* h) {
    if (h->nnodes == h->cap) {
        h->cap = h->cap ? h->cap * 2 : 1024;
        h->nodes = (BNode*)realloc(h->nodes, sizeof(BNode) * (size_t)h->cap);
    }
    return h->nnodes++;
}

static inline int cmp_point(const PrimInfo* a, const PrimInfo* b, int dim) { /* bvh.cc:61-68 */
    return vcomp(a->c, dim) < vcomp(b->c, dim);
}
static inline void swap_pi(PrimInfo* a, PrimInfo* b) {
    PrimInfo t = *a;
    *a = *b;
    *b = t;
}

/* libstdc++ std::nth_element (introselect) as spica calls it for <= 8 primitives
 * (bvh.cc:175-179).  Restated so the median split picks the same elements. */
static void move_median_to_first(PrimInfo* res, PrimInfo* a, PrimInfo* b, PrimInfo* c, int d) {
    if (cmp_point(a, b, d)) {
        if (cmp_point(b, c, d)) swap_pi(res, b);
        else if (cmp_point(a, c, d)) swap_pi(res, c);
        else swap_pi(res, a);
    } else if (cmp_point(a, c, d)) swap_pi(res, a);
    else if (cmp_point(b, c, d)) swap_pi(res, c);
    else swap_pi(res, b);
}
static PrimInfo* unguarded_partition(PrimInfo* first, PrimInfo* last, PrimInfo* pivot, int d) {
    for (;;) {
        while (cmp_point(first, pivot, d)) ++first;
        --last;
        while (cmp_point(pivot, last, d)) --last;
        if (!(first < last)) return first;
        swap_pi(first, last);
        ++first;
    }
}
static void insertion_sort(PrimInfo* first, PrimInfo* last, int d) {
    if (first == last) return;
    for (PrimInfo* i = first + 1; i != last; ++i) {
        if (cmp_point(i, first, d)) {
            PrimInfo val = *i;
            memmove(first + 1, first, (size_t)(i - first) * sizeof(PrimInfo));
            *first = val;
        } else {
            PrimInfo val = *i;
            PrimInfo* l = i;
            PrimInfo* nx = i - 1;
            while (cmp_point(&val, nx, d)) {
                *l = *nx;
                l = nx;
                --nx;
            }
            *l = val;
        }
    }
}
static void nth_element_pi(PrimInfo* first, PrimInfo* nth, PrimInfo* last, int d) {
    if (first == last || nth == last) return;
    int n = (int)(last - first), lg = 0;
    while ((n >> (lg + 1)) > 0) lg++;
    int depth_limit = 2 * lg;
    while (last - first > 3) {
        if (depth_limit == 0) {
            /* heap-select fallback: never reached for the <= 8 elements spica passes */
            insertion_sort(first, last, d);
            return;
        }
        --depth_limit;
        PrimInfo* mid = first + (last - first) / 2;
        move_median_to_first(first, first + 1, mid, last - 1, d);
        PrimInfo* cut = unguarded_partition(first + 1, last, first, d);
        if (cut <= nth) first = cut;
        else last = cut;
    }
    insertion_sort(first, last, d);
}

/* bvh.cc:71-91 */
static inline int to_bucket_le(const PrimInfo* p, int split, int nb, int dim, Box cb) {
    double cmin = vcomp(cb.lo, dim), cmax = vcomp(cb.hi, dim);
    double inv = 1.0 / (fabs(cmax - cmin) + EPS);
    double diff = fabs(vcomp(p->c, dim) - cmin);
    int b = (int)(nb * diff * inv);
    if (b >= nb) b = nb - 1;
    return b <= split;
}
/* libstdc++ std::partition for bidirectional iterators */
static PrimInfo* partition_pi(PrimInfo* first, PrimInfo* last, int split, int nb, int dim, Box cb) {
    for (;;) {
        for (;;) {
            if (first == last) return first;
            if (to_bucket_le(first, split, nb, dim, cb)) ++first;
            else break;
        }
        --last;
        for (;;) {
            if (first == last) return first;
            if (!to_bucket_le(last, split, nb, dim, cb)) --last;
            else break;
        }
        swap_pi(first, last);
        ++first;
    }
}

/* bvh.cc:150-237 */
static int bvh_build_rec(BVH* h, PrimInfo* bd, int start, int end, int depth) {
    if (start == end) return -1;
    if (depth > h->maxdepth) h->maxdepth = depth;
    int node = bvh_new_node(h);
    Box bounds = box_empty();
    for (int i = start; i < end; i++) bounds = box_merge(bounds, bd[i].b);
    int nprims = end - start;
    if (nprims == 1) {
        h->nodes[node].b = bounds;
        h->nodes[node].prim = bd[start].prim;
        h->nodes[node].left = h->nodes[node].right = -1;
        return node;
    }
    Box cb = box_empty();
    for (int i = start; i < end; i++) cb = box_merge_pt(cb, bd[i].c);
    int axis = box_max_extent(cb);
    int mid = (start + end) / 2;
    if (nprims <= 8) {
        nth_element_pi(bd + start, bd + mid, bd + end, axis);
    } else {
        enum { NB = 16 };
        int cnt[NB];
        Box bb[NB];
        for (int i = 0; i < NB; i++) {
            cnt[i] = 0;
            bb[i] = box_empty();
        }
        double cmin = vcomp(cb.lo, axis), cmax = vcomp(cb.hi, axis);
        double idenom = 1.0 / (fabs(cmax - cmin) + EPS);
        for (int i = start; i < end; i++) {
            double numer = vcomp(bd[i].c, axis) - vcomp(cb.lo, axis);
            int b = (int)(NB * fabs(numer) * idenom);
            if (b == NB) b = NB - 1;
            cnt[b]++;
            bb[b] = box_merge(bb[b], bd[i].b);
        }
        double cost[NB - 1];
        for (int i = 0; i < NB - 1; i++) {
            Box b0 = box_empty(), b1 = box_empty();
            int c0 = 0, c1 = 0;
            for (int j = 0; j <= i; j++) {
                b0 = box_merge(b0, bb[j]);
                c0 += cnt[j];
            }
            for (int j = i + 1; j < NB; j++) {
                b1 = box_merge(b1, bb[j]);
                c1 += cnt[j];
            }
            cost[i] = 0.0;
            cost[i] += 0.125 + (c0 * box_area(b0) + c1 * box_area(b1)) / box_area(bounds);
        }
        double minCost = cost[0];
        int minSplit = 0;
        for (int i = 1; i < NB - 1; i++) {
            if (minCost > cost[i]) {
                minCost = cost[i];
                minSplit = i;
            }
        }
        if (minCost < nprims) {
            PrimInfo* it = partition_pi(bd + start, bd + end, minSplit, NB, axis, cb);
            mid = (int)(it - bd);
        }
    }
    int l = bvh_build_rec(h, bd, start, mid, depth + 1);
    int r = bvh_build_rec(h, bd, mid, end, depth + 1);
    h->nodes[node].b = bounds;
    h->nodes[node].left = l;
    h->nodes[node].right = r;
    h->nodes[node].prim = -1;
    return node;
}

/* ------------------------------------------------------------------------- */
/* Environment map: MipMap + Distribution2D (lights/envmap.cc, core/mipmap.cc,*/
/* core/sampling.cc:25-149)                                                   */
/* ------------------------------------------------------------------------- */
typedef struct { int w, h; double* px; } Img; /* rgb, px[(y*w + x)*3 + c] */

typedef struct {
    int n;
    double* func;
    double* cdf;
    double integral;
} Dist1D;

static void dist1d_init(Dist1D* d, const double* f, int n) {
    d->n = n;
    d->func = (double*)malloc(sizeof(double) * (size_t)n);
    d->cdf = (double*)malloc(sizeof(double) * (size_t)(n + 1));
    memcpy(d->func, f, sizeof(double) * (size_t)n);
    d->cdf[0] = 0.0;
    for (int i = 1; i < n + 1; i++) d->cdf[i] = d->cdf[i - 1] + d->func[i - 1] / n;
    d->integral = d->cdf[n];
    if (d->integral == 0.0) {
        for (int i = 1; i < n + 1; i++) d->cdf[i] = (double)i / n;
    } else {
        for (int i = 1; i < n + 1; i++) d->cdf[i] /= d->integral;
    }
}
/* sampling.cc:94-97: upper_bound - 1, capped at size - 1 */
static int find_interval(const double* cdf, int size, double v) {
    int lo = 0, hi = size; /* first index with cdf[i] > v */
    while (lo < hi) {
        int mid = (lo + hi) >> 1;
        if (v < cdf[mid]) hi = mid;
        else lo = mid + 1;
    }
    int ret = lo - 1;
    return ret < size - 1 ? ret : size - 1;
}
/* sampling.cc:70-81 */
static double dist1d_sample(const Dist1D* d, double u, double* pdf, int* off) {
    int o = find_interval(d->cdf, d->n + 1, u);
    if (off) *off = o;
    double du = u - d->cdf[o];
    if (d->cdf[o + 1] - d->cdf[o] > 0.0) du /= (d->cdf[o + 1] - d->cdf[o]);
    *pdf = d->func[o] / d->integral;
    return (o + du) / d->n;
}

typedef struct {
    int levels;
    Img* pyr;
} MipMap;

static inline int round_up_pow2(int v) {
    v--;
    v |= v >> 1;
    v |= v >> 2;
    v |= v >> 4;
    v |= v >> 8;
    v |= v >> 16;
    return v + 1;
}
static inline int is_pow2(int v) { return v && !(v & (v - 1)); }
static inline int log2_int(int v) {
    int r = 0;
    while (v != 0) {
        r++;
        v >>= 1;
    }
    return r;
}
/* mipmap.cc:67-79 (ImageWrap::Repeat) */
static V3 mip_texel(const MipMap* m, int level, int s, int t) {
    const Img* im = &m->pyr[level];
    int w = im->w, h = im->h;
    s = (s % w + w) % w;
    t = (t % h + h) % h;
    const double* p = im->px + ((size_t)t * w + s) * 3;
    return v3(p[0], p[1], p[2]);
}
/* mipmap.cc:37-56 */
static void mip_init(MipMap* m, const Img* base) {
    int rx = base->w, ry = base->h;
    if (!is_pow2(rx) || !is_pow2(ry)) {
        rx = round_up_pow2(rx);
        ry = round_up_pow2(ry);
    }
    m->levels = 1 + log2_int(rx > ry ? rx : ry);
    m->pyr = (Img*)calloc((size_t)m->levels, sizeof(Img));
    m->pyr[0] = *base;
    for (int i = 1; i < m->levels; i++) {
        int sr = m->pyr[i - 1].w / 2;
        int tr = m->pyr[i - 1].h / 2;
        if (sr < 1) sr = 1;
        if (tr < 1) tr = 1;
        m->pyr[i].w = sr;
        m->pyr[i].h = tr;
        m->pyr[i].px = (double*)malloc(sizeof(double) * 3 * (size_t)sr * tr);
        for (int y = 0; y < tr; y++)
            for (int x = 0; x < sr; x++) {
                V3 a = mip_texel(m, i - 1, 2 * x, 2 * y);
                V3 b = mip_texel(m, i - 1, 2 * x + 1, 2 * y);
                V3 c = mip_texel(m, i - 1, 2 * x, 2 * y + 1);
                V3 d = mip_texel(m, i - 1, 2 * x + 1, 2 * y + 1);
                V3 s = vmul(vadd(vadd(vadd(a, b), c), d), 0.25);
                double* p = m->pyr[i].px + ((size_t)y * sr + x) * 3;
                p[0] = s.x;
                p[1] = s.y;
                p[2] = s.z;
            }
    }
}
/* mipmap.cc:71-79: truncation toward zero, so the first half texel extrapolates */
static V3 mip_bilinear(const MipMap* m, int level, double s0, double t0) {
    level = level < 0 ? 0 : (level > m->levels ? m->levels : level);
    const Img* im = &m->pyr[level];
    double s = s0 * im->w - 0.5;
    double t = t0 * im->h - 0.5;
    int si = (int)s, ti = (int)t;
    double ds = s - si, dt = t - ti;
    V3 r = vmul(mip_texel(m, level, si, ti), (1.0 - ds) * (1.0 - dt));
    r = vadd(r, vmul(mip_texel(m, level, si + 1, ti), ds * (1.0 - dt)));
    r = vadd(r, vmul(mip_texel(m, level, si, ti + 1), (1.0 - ds) * dt));
    r = vadd(r, vmul(mip_texel(m, level, si + 1, ti + 1), ds * dt));
    return r;
}
/* mipmap.cc:58-69 */
static V3 mip_lookup(const MipMap* m, double s, double t, double width) {
    double level = m->levels - 1 + log2(fmax(width, 1.0e-8));
    if (level < 0) return mip_bilinear(m, 0, s, t);
    if (level >= m->levels - 1) return mip_texel(m, m->levels - 1, 0, 0);
    int l = (int)level;
    double delta = level - l;
    return vadd(vmul(mip_bilinear(m, l, s, t), 1.0 - delta), vmul(mip_bilinear(m, l + 1, s, t), delta));
}

typedef struct {
    int present;
    int w, h;
    MipMap mip;
    Dist1D* cond; /* h rows */
    Dist1D marg;
    double l2w[3][3], w2l[3][3];
    V3 center;
    double radius;
} Env;

static void mat3_inverse(double a[3][3], double r[3][3]) {
    double det = a[0][0] * (a[1][1] * a[2][2] - a[1][2] * a[2][1]) - a[0][1] * (a[1][0] * a[2][2] - a[1][2] * a[2][0]) +
                 a[0][2] * (a[1][0] * a[2][1] - a[1][1] * a[2][0]);
    double id = 1.0 / det;
    r[0][0] = (a[1][1] * a[2][2] - a[1][2] * a[2][1]) * id;
    r[0][1] = (a[0][2] * a[2][1] - a[0][1] * a[2][2]) * id;
    r[0][2] = (a[0][1] * a[1][2] - a[0][2] * a[1][1]) * id;
    r[1][0] = (a[1][2] * a[2][0] - a[1][0] * a[2][2]) * id;
    r[1][1] = (a[0][0] * a[2][2] - a[0][2] * a[2][0]) * id;
    r[1][2] = (a[0][2] * a[1][0] - a[0][0] * a[1][2]) * id;
    r[2][0] = (a[1][0] * a[2][1] - a[1][1] * a[2][0]) * id;
    r[2][1] = (a[0][1] * a[2][0] - a[0][0] * a[2][1]) * id;
    r[2][2] = (a[0][0] * a[1][1] - a[0][1] * a[1][0]) * id;
}
static inline V3 mat3_apply(const double m[3][3], V3 v) {
    return v3(m[0][0] * v.x + m[0][1] * v.y + m[0][2] * v.z, m[1][0] * v.x + m[1][1] * v.y + m[1][2] * v.z,
              m[2][0] * v.x + m[2][1] * v.y + m[2][2] * v.z);
}

/* envmap.cc:17-47 */
static void env_init(Env* e, const spk_scene_desc* d) {
    memset(e, 0, sizeof(*e));
    if (d->env_width <= 0 || d->env_height <= 0 || !d->env_rgb) return;
    e->present = 1;
    int w = d->env_width, h = d->env_height;
    e->w = w;
    e->h = h;
    Img base;
    base.w = w;
    base.h = h;
    base.px = (double*)malloc(sizeof(double) * 3 * (size_t)w * h);
    for (size_t i = 0; i < (size_t)w * h * 3; i++) base.px[i] = (double)d->env_rgb[i] * d->env_scale;
    mip_init(&e->mip, &base);
    double filter = 1.0 / (w > h ? w : h);
    double* g = (double*)malloc(sizeof(double) * (size_t)w * h);
    for (int v = 0; v < h; v++) {
        double vp = (double)(v + 0.5) / h;
        double st = sin(PI * (v + 0.5) / h);
        for (int u = 0; u < w; u++) {
            double up = (double)(u + 0.5) / w;
            g[(size_t)v * w + u] = gray(mip_lookup(&e->mip, up, vp, filter));
            g[(size_t)v * w + u] *= st;
        }
    }
    e->cond = (Dist1D*)malloc(sizeof(Dist1D) * (size_t)h);
    double* mf = (double*)malloc(sizeof(double) * (size_t)h);
    for (int v = 0; v < h; v++) {
        dist1d_init(&e->cond[v], g + (size_t)v * w, w);
        mf[v] = e->cond[v].integral;
    }
    dist1d_init(&e->marg, mf, h);
    free(mf);
    free(g);
    /* envmap.cc:19: spica keeps the transpose of the given matrix */
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) e->l2w[i][j] = d->env_to_world[j * 4 + i];
    mat3_inverse(e->l2w, e->w2l);
    e->center = v3(d->env_center[0], d->env_center[1], d->env_center[2]);
    e->radius = d->env_radius;
}

/* sampling.cc:136-149 */
static void env_dist_sample(const Env* e, double u0, double u1, double* uu, double* vv, double* pdf) {
    double p1, p0;
    int v;
    double d1 = dist1d_sample(&e->marg, u1, &p1, &v);
    double d0 = dist1d_sample(&e->cond[v], u0, &p0, NULL);
    *pdf = p0 * p1;
    *uu = d0;
    *vv = d1;
}
static double env_dist_pdf(const Env* e, double u, double v) {
    int iu = (int)(u * e->w);
    int iv = (int)(v * e->h);
    iu = iu < 0 ? 0 : (iu > e->w - 1 ? e->w - 1 : iu);
    iv = iv < 0 ? 0 : (iv > e->h - 1 ? e->h - 1 : iv);
    return e->cond[iv].func[iu] / e->marg.integral;
}
/* vect_math.h:55-65 */
static double spherical_phi(V3 d) {
    double p = atan2(d.y, d.x);
    return p < 0.0 ? p + 2.0 * PI : p;
}
static double spherical_theta(V3 d) { return acos(clampd(d.z, -1.0, 1.0)); }
/* envmap.cc:130-134 */
static V3 env_Le(const Env* e, V3 dir) {
    V3 d = vnormalized(mat3_apply(e->w2l, dir));
    return mip_lookup(&e->mip, spherical_phi(d) * (0.5 * INV_PI), spherical_theta(d) * INV_PI, 0.0);
}
/* envmap.cc:81-88 */
static double env_pdfLi(const Env* e, V3 dw) {
    V3 d = mat3_apply(e->w2l, dw);
    double theta = spherical_theta(d);
    double phi = spherical_phi(d);
    double st = sin(theta);
    if (st == 0.0) return 0.0;
    return env_dist_pdf(e, phi * (0.5 * INV_PI), theta * INV_PI) / (2.0 * PI * PI * st);
}
/* envmap.cc:60-79; returns Li, *pdf = 0 when the map pdf vanishes */
static V3 env_sampleLi(const Env* e, V3 p, double u0, double u1, V3* dir, double* pdf, V3* plight) {
    double uu, vv, mpdf;
    env_dist_sample(e, u0, u1, &uu, &vv, &mpdf);
    if (mpdf == 0.0) {
        *pdf = 0.0;
        return v3(0, 0, 0);
    }
    double theta = vv * PI, phi = uu * (2.0 * PI);
    double ct = cos(theta), st = sin(theta), cp = cos(phi), sp = sin(phi);
    *dir = mat3_apply(e->l2w, v3(st * cp, st * sp, ct));
    *pdf = st == 0.0 ? 0.0 : mpdf / (2.0 * PI * PI * st);
    *plight = vadd(p, vmul(*dir, 2.0 * e->radius));
    return mip_lookup(&e->mip, uu, vv, 0.0);
}

/* ------------------------------------------------------------------------- */
/* Scene                                                                      */
/* ------------------------------------------------------------------------- */
typedef struct {
    int ntris;
    Tri* tris;
    int* mat;
    int* light;     /* area light id per triangle or -1 */
    int nmat;
    spk_material* mats;
    int nareas;
    int* area_tri;
    V3* area_L;
    Env env;
    int nlights;    /* nareas + env */
    BVH bvh;
} Scene;

/* ------------------------------------------------------------------------- */
/* Closest / any hit through the spica BVH (bvh.cc:331-387)                  */
/* ------------------------------------------------------------------------- */
typedef struct { int64_t nodes, tris, rays; } Counts;

static int scene_closest(const Scene* s, Ray* r, int* prim, double* th, double* uo, double* vo, Counts* cnt) {
    if (cnt) cnt->rays++;
    if (s->bvh.nnodes == 0) return 0;
    int stackbuf[256];
    int cap = 256;
    int* stack = stackbuf;
    if (s->bvh.maxdepth + 2 > cap) {
        cap = s->bvh.maxdepth + 2;
        stack = (int*)malloc(sizeof(int) * (size_t)cap * 2);
    }
    int sp = 0, hit = 0;
    stack[sp++] = 0;
    while (sp > 0) {
        const BNode* nd = &s->bvh.nodes[stack[--sp]];
        if (nd->prim >= 0) {
            if (cnt) cnt->tris++;
            double t, u, v;
            if (tri_intersect(&s->tris[nd->prim], r, &t, &u, &v)) {
                r->maxd = t; /* primitive.cc:52 */
                *prim = nd->prim;
                *th = t;
                *uo = u;
                *vo = v;
                hit = 1;
            }
        } else {
            if (cnt) cnt->nodes++;
            if (box_hit(&nd->b, r)) {
                if (nd->left >= 0) stack[sp++] = nd->left;
                if (nd->right >= 0) stack[sp++] = nd->right;
            }
        }
    }
    if (stack != stackbuf) free(stack);
    return hit;
}

static int scene_any(const Scene* s, const Ray* r, Counts* cnt) {
    if (cnt) cnt->rays++;
    if (s->bvh.nnodes == 0) return 0;
    int stackbuf[256];
    int cap = 256;
    int* stack = stackbuf;
    if (s->bvh.maxdepth + 2 > cap) {
        cap = s->bvh.maxdepth + 2;
        stack = (int*)malloc(sizeof(int) * (size_t)cap * 2);
    }
    int sp = 0, hit = 0;
    stack[sp++] = 0;
    while (sp > 0) {
        const BNode* nd = &s->bvh.nodes[stack[--sp]];
        if (nd->prim >= 0) {
            if (cnt) cnt->tris++;
            double t;
            if (tri_intersect(&s->tris[nd->prim], r, &t, NULL, NULL)) {
                hit = 1;
                break;
            }
        } else {
            if (cnt) cnt->nodes++;
            if (box_hit(&nd->b, r)) {
                if (nd->left >= 0) stack[sp++] = nd->left;
                if (nd->right >= 0) stack[sp++] = nd->right;
            }
        }
    }
    if (stack != stackbuf) free(stack);
    return hit;
}

/* ------------------------------------------------------------------------- */
/* Samplers                                                                   */
/* ------------------------------------------------------------------------- */
/* MT19937 (random.cc:104-164) */
struct MT { uint32_t mt[624]; int mti; };
static void mt_seed(MT* m, uint32_t s) {
    m->mt[0] = s;
    for (m->mti = 1; m->mti < 624; m->mti++)
        m->mt[m->mti] = (uint32_t)(1812433253u * (m->mt[m->mti - 1] ^ (m->mt[m->mti - 1] >> 30)) + (uint32_t)m->mti);
}
static uint32_t mt_u32(MT* m) {
    static const uint32_t mag01[2] = {0x0u, 0x9908b0dfu};
    uint32_t y;
    if (m->mti >= 624) {
        int kk;
        for (kk = 0; kk < 624 - 397; kk++) {
            y = (m->mt[kk] & 0x80000000u) | (m->mt[kk + 1] & 0x7fffffffu);
            m->mt[kk] = m->mt[kk + 397] ^ (y >> 1) ^ mag01[y & 1u];
        }
        for (; kk < 623; kk++) {
            y = (m->mt[kk] & 0x80000000u) | (m->mt[kk + 1] & 0x7fffffffu);
            m->mt[kk] = m->mt[kk + (397 - 624)] ^ (y >> 1) ^ mag01[y & 1u];
        }
        y = (m->mt[623] & 0x80000000u) | (m->mt[0] & 0x7fffffffu);
        m->mt[623] = m->mt[396] ^ (y >> 1) ^ mag01[y & 1u];
        m->mti = 0;
    }
    y = m->mt[m->mti++];
    y ^= (y >> 11);
    y ^= (y << 7) & 0x9d2c5680u;
    y ^= (y << 15) & 0xefc60000u;
    y ^= (y >> 18);
    return y;
}
static double mt_real2(MT* m) { return mt_u32(m) * (1.0 / 4294967296.0); }

/* PCG4D (Jarzynski & Olano, "Hash Functions for GPU Rendering", JCGT 2020) */
static void pcg4d(uint32_t v[4]) {
    for (int i = 0; i < 4; i++) v[i] = v[i] * 1664525u + 1013904223u;
    v[0] += v[1] * v[3];
    v[1] += v[2] * v[0];
    v[2] += v[0] * v[1];
    v[3] += v[1] * v[2];
    for (int i = 0; i < 4; i++) v[i] ^= v[i] >> 16;
    v[0] += v[1] * v[3];
    v[1] += v[2] * v[0];
    v[2] += v[0] * v[1];
    v[3] += v[1] * v[2];
}
static inline uint32_t reverse_bits32(uint32_t x) {
    x = ((x >> 1) & 0x55555555u) | ((x & 0x55555555u) << 1);
    x = ((x >> 2) & 0x33333333u) | ((x & 0x33333333u) << 2);
    x = ((x >> 4) & 0x0f0f0f0fu) | ((x & 0x0f0f0f0fu) << 4);
    x = ((x >> 8) & 0x00ff00ffu) | ((x & 0x00ff00ffu) << 8);
    return (x >> 16) | (x << 16);
}
/* Burley, "Practical Hash-based Owen Scrambling", JCGT 2020 */
static inline uint32_t lk_permute(uint32_t x, uint32_t seed) {
    x += seed;
    x ^= x * 0x6c50b47cu;
    x ^= x * 0xb82f1e52u;
    x ^= x * 0xc7afe638u;
    x ^= x * 0x8d22f6e6u;
    return x;
}
static inline uint32_t nested_uniform_scramble(uint32_t x, uint32_t seed) {
    return reverse_bits32(lk_permute(reverse_bits32(x), seed));
}
static uint32_t sobol_dir1[32];
static int sobol_ready = 0;
static void sobol_init(void) {
    /* second Sobol dimension: primitive polynomial x + 1, m_1 = 1 */
    uint32_t v = 1u << 31;
    for (int k = 0; k < 32; k++) {
        sobol_dir1[k] = v;
        v ^= v >> 1;
    }
    sobol_ready = 1;
}
static inline uint32_t sobol2(uint32_t idx, int comp) {
    if (comp == 0) return reverse_bits32(idx);
    uint32_t x = 0;
    for (int k = 0; idx; k++, idx >>= 1)
        if (idx & 1u) x ^= sobol_dir1[k];
    return x;
}
/* Dimension layout per path: 0-1 film, 2-3 lens, then 8 per bounce b:
 * 4+8b light pick, +1..2 light uv, +3..4 NEE-BSDF uv, +5..6 continuation uv, +7 RR
 * (the order spica draws them: integrator.cc:84-85, mis.cc:27-30, path.cc:87,119).
 * Sobol pairs: (film) (lens) then per bounce (pick, RR) (light) (nee) (cont). */
static void dim_to_pair(uint32_t dim, uint32_t* pair, int* comp) {
    if (dim < 4) {
        *pair = dim >> 1;
        *comp = (int)(dim & 1u);
        return;
    }
    uint32_t k = (dim - 4) & 7u, b = (dim - 4) >> 3;
    uint32_t base = 2 + 4 * b;
    if (k == 0) { *pair = base; *comp = 0; }
    else if (k == 7) { *pair = base; *comp = 1; }
    else if (k <= 2) { *pair = base + 1; *comp = (int)k - 1; }
    else if (k <= 4) { *pair = base + 2; *comp = (int)k - 3; }
    else { *pair = base + 3; *comp = (int)k - 5; }
}
static uint32_t counter_u32(int kind, uint32_t seed, uint32_t pixel, uint32_t sample, uint32_t dim) {
    if (kind == SPK_SAMPLER_SOBOL) {
        uint32_t pair;
        int comp;
        dim_to_pair(dim, &pair, &comp);
        uint32_t h[4] = {pixel, pair, seed, 0x9e3779b9u};
        pcg4d(h);
        uint32_t idx = nested_uniform_scramble(sample, h[0]);
        uint32_t x = sobol2(idx, comp);
        return nested_uniform_scramble(x, comp == 0 ? h[1] : h[2]);
    }
    uint32_t h[4] = {pixel, sample, dim, seed};
    pcg4d(h);
    return h[0];
}
static inline double u32_to_unit(uint32_t x) { return (double)(x >> 8) * (1.0 / 16777216.0); }

typedef struct {
    int kind;
    uint32_t seed, pixel, sample;
    MT* mt;
    Hidden* hid;   /* the BxDFs' hidden uniforms */
} Sampler;

static inline double smp(Sampler* s, uint32_t dim) {
    if (s->kind == SPK_SAMPLER_MT19937) return mt_real2(s->mt);
    return u32_to_unit(counter_u32(s->kind, s->seed, s->pixel, s->sample, dim));
}

static double hidden_draw(Hidden* h, int stream) { return h->use_mt ? mt_real2(h->mt[stream]) : h->u; }

/* With counter-based samplers the hidden uniform of a BSDF sample is one more
 * dimension: for the path integrator 4 + 8 * 128 + 2 * bounce + (0 NEE, 1 continuation),
 * past the path's own 8 per bounce and the learned proposal's block; for directlighting
 * the last of its 10 per level. */
static inline uint32_t hidden_dim_path(int bounce, int which) { return 4u + 8u * 128u + 2u * (uint32_t)bounce + (uint32_t)which; }
static inline void hidden_set(Sampler* sp, uint32_t dim) {
    if (sp->kind != SPK_SAMPLER_MT19937) sp->hid->u = smp(sp, dim);
}
/* three default-seeded thread-local streams, as spica's Random() (random.cc:70-74) */
static void hidden_init_mt(Hidden* h, MT* storage) {
    h->use_mt = 1;
    for (int k = 0; k < 3; k++) {
        mt_seed(&storage[k], 0u);
        h->mt[k] = &storage[k];
    }
    h->u = 0.0;
}

/* ------------------------------------------------------------------------- */
/* Lights (lights/area.cc:29-45, core/shape.cc:35-48, triangle.cc:180-197)   */
/* ------------------------------------------------------------------------- */
/* Shape::pdf(ref, wi): re-intersect the light triangle from the offset origin */
static double area_pdf(const Scene* s, int li, const Isect* ref, V3 wi) {
    const Tri* t = &s->tris[s->area_tri[li]];
    Ray r = make_ray(offset_ray_origin(ref->pos, ref->n, wi), wi, INFTY);
    double th, u, v;
    if (!tri_intersect(t, &r, &th, &u, &v)) return 0.0;
    Isect li_is;
    make_isect(t, &r, th, u, v, s->area_tri[li], &li_is);
    V3 d = vsub(ref->pos, li_is.pos);
    double ret = vdot(d, d) / (vabsdot(li_is.n, vneg(wi)) * t->area);
    if (isinf(ret)) ret = 0.0;
    return ret;
}

/* interaction.cc:88-93 */
static Ray spawn_ray_to(V3 p, V3 n, V3 q, V3 qn) {
    V3 o = offset_ray_origin(p, n, vsub(q, p));
    V3 tgt = offset_ray_origin(q, qn, vsub(o, q));
    V3 d = vsub(tgt, o);
    return make_ray(o, d, vnorm(d));
}

/* mis.cc:138-141 */
static inline double power_heuristic(double f, double g) { return (f * f) / (f * f + g * g); }

/* mis.cc:35-136 (surface interaction, no media, non-delta lights) */
static V3 estimate_direct(const Scene* s, const Isect* is, const BSDF* b, int li, double ul0, double ul1, double us0,
                          double us1, Counts* cnt, Sampler* sp, uint32_t hdim) {
    const int mask = BX_ALL & ~BX_SPECULAR;
    V3 Ld = v3(0, 0, 0);
    V3 wi = v3(0, 0, 0);
    double lightPdf = 0.0, bsdfPdf = 0.0;
    int is_env = (li >= s->nareas);
    V3 Li = v3(0, 0, 0);
    Ray shadow;
    /* light.sampleLi */
    if (!is_env) {
        const Tri* t = &s->tris[s->area_tri[li]];
        double a0 = ul0, a1 = ul1;
        if (ul0 + ul1 >= 1.0) {
            a0 = 1.0 - ul0;
            a1 = 1.0 - ul1;
        }
        V3 e1 = vsub(t->p[1], t->p[0]), e2 = vsub(t->p[2], t->p[0]);
        V3 lp = vadd(vadd(t->p[0], vmul(e1, a0)), vmul(e2, a1));
        V3 ln = vadd(vadd(vmul(t->n[0], 1.0 - a0 - a1), vmul(t->n[1], a0)), vmul(t->n[2], a1));
        wi = vnormalized(vsub(lp, is->pos));
        lightPdf = area_pdf(s, li, is, wi);
        shadow = spawn_ray_to(is->pos, is->n, lp, ln);
        Li = vdot(ln, vneg(wi)) > 0.0 ? s->area_L[li] : v3(0, 0, 0);
    } else {
        V3 lp;
        Li = env_sampleLi(&s->env, is->pos, ul0, ul1, &wi, &lightPdf, &lp);
        if (lightPdf != 0.0 || !vis_black(Li)) shadow = spawn_ray_to(is->pos, is->n, lp, v3(0, 0, 0));
    }
    if (lightPdf > 0.0 && !vis_black(Li)) {
        V3 f = vmul(bsdf_f(b, is->wo, wi, mask), vabsdot(wi, b->n));
        bsdfPdf = bsdf_pdf(b, is->wo, wi, mask);
        if (!vis_black(f)) {
            if (scene_any(s, &shadow, cnt)) Li = v3(0, 0, 0);
            if (!vis_black(Li)) {
                double w = power_heuristic(lightPdf, bsdfPdf);
                Ld = vadd(Ld, vmul(vcmul(f, Li), w / lightPdf));
            }
        }
    }
    /* BSDF sample */
    {
        int st = 0;
        double pdf = 0.0;
        V3 wi2 = v3(0, 0, 0);
        hidden_set(sp, hdim);
        V3 f = bsdf_sample(b, is->wo, &wi2, us0, us1, &pdf, mask, &st, sp->hid);
        f = vmul(f, vabsdot(wi2, b->n));
        bsdfPdf = pdf;
        int sampledSpecular = (st & BX_SPECULAR) != 0;
        if (!vis_black(f) && bsdfPdf > 0.0) {
            double w = 1.0;
            if (!sampledSpecular) {
                lightPdf = is_env ? env_pdfLi(&s->env, wi2) : area_pdf(s, li, is, wi2);
                if (lightPdf == 0.0) return Ld;
                w = power_heuristic(bsdfPdf, lightPdf);
            }
            Ray r = make_ray(offset_ray_origin(is->pos, is->n, wi2), wi2, INFTY);
            int prim;
            double th, u, v;
            V3 Le = v3(0, 0, 0);
            if (scene_closest(s, &r, &prim, &th, &u, &v, cnt)) {
                if (!is_env && s->light[prim] == li) {
                    Isect hit;
                    make_isect(&s->tris[prim], &r, th, u, v, prim, &hit);
                    Le = vdot(hit.n, vneg(wi2)) > 0.0 ? s->area_L[li] : v3(0, 0, 0);
                }
            } else if (is_env) {
                Le = env_Le(&s->env, r.d);
            }
            if (!vis_black(Le)) Ld = vadd(Ld, vmul(vcmul(f, Le), w / bsdfPdf));
        }
    }
    return Ld;
}

typedef struct { int64_t rays, nodes, tris; } PathStats;

/* path.cc:42-125 with the sampler dimensions of the counter-based streams */
static V3 Li_path(const Scene* s, Ray ray, Sampler* sp, int maxDepth, int rrDepth, PathStats* ps) {
    V3 L = v3(0, 0, 0), beta = v3(1, 1, 1);
    int specular = 0;
    Counts cnt = {0, 0, 0};
    Counts* cp = ps ? &cnt : NULL;
    for (int bounces = 0;; bounces++) {
        uint32_t base = 4u + 8u * (uint32_t)bounces;
        int prim = -1;
        double th = 0, u = 0, v = 0;
        int hit = scene_closest(s, &ray, &prim, &th, &u, &v, cp);
        Isect is;
        if (hit) make_isect(&s->tris[prim], &ray, th, u, v, prim, &is);
        if (bounces == 0 || specular) {
            if (hit) {
                int li = s->light[prim];
                if (li >= 0 && vdot(is.n, vneg(ray.d)) > 0.0) L = vadd(L, vcmul(beta, s->area_L[li]));
            } else if (s->env.present) {
                L = vadd(L, vcmul(beta, env_Le(&s->env, ray.d)));
            }
        }
        if (!hit || bounces >= maxDepth) break;
        BSDF b;
        setup_bsdf(&s->mats[s->mat[prim]], &is, &b);
        if (bsdf_num_components(&b, BX_ALL & ~BX_SPECULAR) > 0 && s->nlights > 0) {
            /* mis.cc:21-33 */
            double upick = smp(sp, base + 0);
            int li = (int)(upick * s->nlights);
            if (li > s->nlights - 1) li = s->nlights - 1;
            double l0 = smp(sp, base + 1), l1 = smp(sp, base + 2);
            double s0 = smp(sp, base + 3), s1 = smp(sp, base + 4);
            V3 Ld = estimate_direct(s, &is, &b, li, l0, l1, s0, s1, cp, sp, hidden_dim_path(bounces, 0));
            L = vadd(L, vcmul(beta, vmul(Ld, (double)s->nlights)));
        }
        V3 wo = vneg(ray.d), wi = v3(0, 0, 0);
        double pdf = 0.0;
        int st = 0;
        double c0 = smp(sp, base + 5), c1 = smp(sp, base + 6);
        hidden_set(sp, hidden_dim_path(bounces, 1));
        V3 f = bsdf_sample(&b, wo, &wi, c0, c1, &pdf, BX_ALL, &st, sp->hid);
        if (vis_black(f) || pdf == 0.0) break;
        beta = vcmul(beta, vmul(f, vabsdot(wi, is.ns) / pdf));
        specular = (st & BX_SPECULAR) != 0;
        ray = make_ray(offset_ray_origin(is.pos, is.n, wi), wi, INFTY);
        if (bounces > rrDepth) {
            double q = fmin(0.95, gray(beta));
            if (smp(sp, base + 7) > q) break;
            beta = vmul(beta, 1.0 / q);
        }
    }
    if (ps) {
        ps->rays += cnt.rays;
        ps->nodes += cnt.nodes;
        ps->tris += cnt.tris;
    }
    return L;
}

/* DirectLightingIntegrator::Li (integrators/directlighting/directlighting.cc:21-57) with
 * SamplerIntegrator::specularReflect / specularTransmit (core/integrator.cc:112-152).
 * `depth` is spica's argument: the recursion passes depth + 1 into specularReflect, which
 * calls Li with its own depth + 1, so each specular bounce adds 2.  Draw order per call
 * (MT19937): light pick, light uv, BSDF uv (uniformSampleOneLight, always drawn when the
 * scene has lights), then while depth + 1 < maxDepth the reflection uv (and its whole
 * subtree), then the transmission uv.  Counter-based dimensions: a block of 10 per
 * recursion level (pick 0, light 1-2, BSDF 3-4, reflection 5-6, transmission 7-8, the
 * NEE BSDF sample's hidden uniform 9); the
 * supported lobes continue through at most one of the two specular calls (a
 * FresnelSpecular lobe, type R|T|S, is no subset of R|S or T|S, so BSDF::sample finds no
 * component for either, bsdf.cc:28-36,65-73), so one block per level suffices. */
static V3 Li_direct(const Scene* s, Ray ray, Sampler* sp, int maxDepth, int depth, int level, Counts* cp) {
    const uint32_t base = 4u + 10u * (uint32_t)level;
    int prim = -1;
    double th = 0, u = 0, v = 0;
    if (!scene_closest(s, &ray, &prim, &th, &u, &v, cp)) {
        /* sum of Le(ray) over the lights: only the environment map emits along a ray */
        return s->env.present ? env_Le(&s->env, ray.d) : v3(0, 0, 0);
    }
    Isect is;
    make_isect(&s->tris[prim], &ray, th, u, v, prim, &is);
    BSDF b;
    setup_bsdf(&s->mats[s->mat[prim]], &is, &b);
    V3 L = v3(0, 0, 0);
    if (depth == 0) {
        int li = s->light[prim];
        if (li >= 0 && vdot(is.n, vneg(ray.d)) > 0.0) L = vadd(L, s->area_L[li]);
    }
    if (s->nlights > 0) {
        double upick = smp(sp, base + 0);
        int li = (int)(upick * s->nlights);
        if (li > s->nlights - 1) li = s->nlights - 1;
        double l0 = smp(sp, base + 1), l1 = smp(sp, base + 2);
        double s0 = smp(sp, base + 3), s1 = smp(sp, base + 4);
        V3 Ld = estimate_direct(s, &is, &b, li, l0, l1, s0, s1, cp, sp, base + 9);
        L = vadd(L, vmul(Ld, (double)s->nlights));
    }
    if (depth + 1 < maxDepth) {
        const int masks[2] = {BX_REFLECTION | BX_SPECULAR, BX_TRANSMISSION | BX_SPECULAR};
        for (int c = 0; c < 2; c++) {
            double u0 = smp(sp, base + 5 + 2 * c), u1 = smp(sp, base + 6 + 2 * c);
            V3 wi = v3(0, 0, 0);
            double pdf = 0.0;
            int st = 0;
            V3 f = bsdf_sample(&b, is.wo, &wi, u0, u1, &pdf, masks[c], &st, sp->hid);
            double cosg = vabsdot(wi, is.n);
            if (pdf > 0.0 && !vis_black(f) && cosg != 0.0) {
                Ray r = make_ray(offset_ray_origin(is.pos, is.n, wi), wi, INFTY);
                V3 Lr = Li_direct(s, r, sp, maxDepth, depth + 2, level + 1, cp);
                L = vadd(L, vmul(vcmul(f, Lr), cosg / pdf));
            }
        }
    }
    return L;
}

/* ------------------------------------------------------------------------- */
/* Camera (cameras/perspective.cc:32-74, core/camera.cc:20-38,               */
/* core/transform.cc:60-76, 205-212)                                          */
/* ------------------------------------------------------------------------- */
typedef struct {
    int W, H;
    double m[4][4];
    double tanHalf, aspect, lensR, focal;
} Cam;

static void cam_init(Cam* c, const spk_camera_desc* d) {
    c->W = d->width;
    c->H = d->height;
    for (int i = 0; i < 4; i++)
        for (int j = 0; j < 4; j++) c->m[i][j] = d->to_world[i * 4 + j];
    c->tanHalf = tan(d->fov_deg * PI / 180.0 / 2.0);
    c->aspect = (double)d->width / d->height;
    c->lensR = d->aperture_radius;
    c->focal = d->focus_distance;
}
/* raster (x+u, y+v) -> screen [-1,1]^2 (y down) -> camera point on the near plane */
static Ray cam_ray(const Cam* c, int x, int y, double f0, double f1, double l0, double l1) {
    double rx = x + f0, ry = y + f1;
    double sx = 2.0 * rx / c->W - 1.0;
    double sy = 1.0 - 2.0 * ry / c->H;
    V3 dir = vnormalize(v3(sx * c->aspect * c->tanHalf, sy * c->tanHalf, 1.0));
    V3 org = v3(0, 0, 0);
    if (c->lensR > 0.0) {
        double dx, dy;
        concentric_disk(l0, l1, &dx, &dy);
        dx *= c->lensR;
        dy *= c->lensR;
        double ft = c->focal / dir.z;
        V3 pf = vmul(dir, ft);
        org = v3(dx, dy, 0.0);
        dir = vnormalize(vsub(pf, org));
    }
    double ow[4];
    for (int i = 0; i < 4; i++) ow[i] = c->m[i][0] * org.x + c->m[i][1] * org.y + c->m[i][2] * org.z + c->m[i][3];
    if (ow[3] != 1.0) {
        ow[0] /= (ow[3] + EPS);
        ow[1] /= (ow[3] + EPS);
        ow[2] /= (ow[3] + EPS);
    }
    V3 dw = v3(c->m[0][0] * dir.x + c->m[0][1] * dir.y + c->m[0][2] * dir.z,
               c->m[1][0] * dir.x + c->m[1][1] * dir.y + c->m[1][2] * dir.z,
               c->m[2][0] * dir.x + c->m[2][1] * dir.y + c->m[2][2] * dir.z);
    return make_ray(v3(ow[0], ow[1], ow[2]), dw, INFTY);
}

/* ------------------------------------------------------------------------- */
/* Exported API (ctypes)                                                      */
/* ------------------------------------------------------------------------- */
EXPORT void* orc_scene_create(const spk_scene_desc* d) {
    if (!sobol_ready) sobol_init();
    Scene* s = (Scene*)calloc(1, sizeof(Scene));
    s->ntris = (int)d->num_tris;
    s->tris = (Tri*)malloc(sizeof(Tri) * (size_t)(s->ntris > 0 ? s->ntris : 1));
    s->mat = (int*)malloc(sizeof(int) * (size_t)(s->ntris > 0 ? s->ntris : 1));
    s->light = (int*)malloc(sizeof(int) * (size_t)(s->ntris > 0 ? s->ntris : 1));
    for (int i = 0; i < s->ntris; i++) {
        tri_init(&s->tris[i], d->verts + (size_t)9 * i, d->normals ? d->normals + (size_t)9 * i : NULL);
        s->mat[i] = d->mat_ids ? d->mat_ids[i] : 0;
        s->light[i] = d->light_ids ? d->light_ids[i] : -1;
    }
    s->nmat = d->num_materials;
    s->mats = (spk_material*)malloc(sizeof(spk_material) * (size_t)(s->nmat > 0 ? s->nmat : 1));
    if (s->nmat > 0) memcpy(s->mats, d->materials, sizeof(spk_material) * (size_t)s->nmat);
    s->nareas = d->num_area_lights;
    s->area_tri = (int*)malloc(sizeof(int) * (size_t)(s->nareas > 0 ? s->nareas : 1));
    s->area_L = (V3*)malloc(sizeof(V3) * (size_t)(s->nareas > 0 ? s->nareas : 1));
    for (int i = 0; i < s->nareas; i++) {
        s->area_tri[i] = d->area_tri[i];
        s->area_L[i] = v3(d->area_radiance[3 * i], d->area_radiance[3 * i + 1], d->area_radiance[3 * i + 2]);
    }
    env_init(&s->env, d);
    s->nlights = s->nareas + (s->env.present ? 1 : 0);
    /* bvh.cc:139-148: primitive order = scene order */
    if (s->ntris > 0) {
        PrimInfo* bd = (PrimInfo*)malloc(sizeof(PrimInfo) * (size_t)s->ntris);
        for (int i = 0; i < s->ntris; i++) {
            const Tri* t = &s->tris[i];
            Box b;
            b.lo = v3(fmin(t->p[0].x, fmin(t->p[1].x, t->p[2].x)), fmin(t->p[0].y, fmin(t->p[1].y, t->p[2].y)),
                      fmin(t->p[0].z, fmin(t->p[1].z, t->p[2].z)));
            b.hi = v3(fmax(t->p[0].x, fmax(t->p[1].x, t->p[2].x)), fmax(t->p[0].y, fmax(t->p[1].y, t->p[2].y)),
                      fmax(t->p[0].z, fmax(t->p[1].z, t->p[2].z)));
            bd[i].prim = i;
            bd[i].b = b;
            bd[i].c = vmul(vadd(b.hi, b.lo), 0.5);
        }
        bvh_build_rec(&s->bvh, bd, 0, s->ntris, 0);
        free(bd);
    }
    return s;
}

EXPORT void orc_scene_free(void* p) {
    Scene* s = (Scene*)p;
    if (!s) return;
    free(s->tris);
    free(s->mat);
    free(s->light);
    free(s->mats);
    free(s->area_tri);
    free(s->area_L);
    free(s->bvh.nodes);
    if (s->env.present) {
        for (int l = 0; l < s->env.mip.levels; l++) free(s->env.mip.pyr[l].px);
        free(s->env.mip.pyr);
        for (int v = 0; v < s->env.h; v++) {
            free(s->env.cond[v].func);
            free(s->env.cond[v].cdf);
        }
        free(s->env.cond);
        free(s->env.marg.func);
        free(s->env.marg.cdf);
    }
    free(s);
}

EXPORT void orc_scene_info(void* p, int64_t* out) {
    Scene* s = (Scene*)p;
    out[0] = s->ntris;
    out[1] = s->bvh.nnodes;
    out[2] = s->bvh.maxdepth;
    out[3] = s->nlights;
}

typedef struct {
    const Scene* s;
    int64_t n;
    const double *o, *d, *tmax;
    int any;
    int32_t* prim;
    double* t;
    int brute;
    int64_t nodes, tris;
    int64_t lo, hi;
} TraceJob;

static void* trace_worker(void* arg) {
    TraceJob* j = (TraceJob*)arg;
    Counts c = {0, 0, 0};
    for (int64_t i = j->lo; i < j->hi; i++) {
        Ray r = make_ray(v3(j->o[3 * i], j->o[3 * i + 1], j->o[3 * i + 2]), v3(j->d[3 * i], j->d[3 * i + 1], j->d[3 * i + 2]),
                         j->tmax ? j->tmax[i] : INFTY);
        int prim = -1;
        double th = -1.0, u, v;
        if (j->brute) {
            /* test_trimesh.cc:10-23 pattern: every triangle, later equal t wins */
            for (int k = 0; k < j->s->ntris; k++) {
                double tt;
                if (tri_intersect(&j->s->tris[k], &r, &tt, &u, &v)) {
                    prim = k;
                    th = tt;
                    if (!j->any) r.maxd = tt;
                    else break;
                }
            }
        } else if (j->any) {
            prim = scene_any(j->s, &r, &c) ? 0 : -1;
        } else {
            if (!scene_closest(j->s, &r, &prim, &th, &u, &v, &c)) {
                prim = -1;
                th = -1.0;
            }
        }
        j->prim[i] = prim;
        if (j->t) j->t[i] = th;
    }
    j->nodes = c.nodes;
    j->tris = c.tris;
    return NULL;
}

/* Closest hit (any_hit = 0) or occlusion (any_hit = 1) for n rays.
 * dir is normalised as spica's Ray does.  prim = -1 on a miss. */
EXPORT void orc_trace(void* p, int64_t n, const double* org, const double* dir, const double* tmax, int any_hit,
                      int brute, int32_t* prim_out, double* t_out, int64_t* counts_out, int nthreads) {
    Scene* s = (Scene*)p;
    if (nthreads < 1) nthreads = 1;
    if (nthreads > 256) nthreads = 256;
    TraceJob jobs[256];
    pthread_t th[256];
    for (int k = 0; k < nthreads; k++) {
        jobs[k].s = s;
        jobs[k].n = n;
        jobs[k].o = org;
        jobs[k].d = dir;
        jobs[k].tmax = tmax;
        jobs[k].any = any_hit;
        jobs[k].prim = prim_out;
        jobs[k].t = t_out;
        jobs[k].brute = brute;
        jobs[k].lo = n * k / nthreads;
        jobs[k].hi = n * (k + 1) / nthreads;
        pthread_create(&th[k], NULL, trace_worker, &jobs[k]);
    }
    int64_t nodes = 0, tris = 0;
    for (int k = 0; k < nthreads; k++) {
        pthread_join(th[k], NULL);
        nodes += jobs[k].nodes;
        tris += jobs[k].tris;
    }
    if (counts_out) {
        counts_out[0] = nodes;
        counts_out[1] = tris;
    }
}

/* Single triangle test for known-answer checks (triangle.cc:98-178). */
EXPORT int orc_tri_intersect(const double* p9, const double* org, const double* dir, double maxd, double* t_out) {
    Tri t;
    float v[9];
    for (int i = 0; i < 9; i++) v[i] = (float)p9[i];
    tri_init(&t, v, NULL);
    for (int i = 0; i < 3; i++) t.p[i] = v3(p9[3 * i], p9[3 * i + 1], p9[3 * i + 2]);
    Ray r = make_ray(v3(org[0], org[1], org[2]), v3(dir[0], dir[1], dir[2]), maxd);
    double th, u, v2;
    int h = tri_intersect(&t, &r, &th, &u, &v2);
    if (h && t_out) *t_out = th;
    return h;
}

/* Slab test (bounds3d_detail.h:63-79) for known-answer checks. */
EXPORT int orc_box_intersect(const double* lo, const double* hi, const double* org, const double* dir, double* tnear,
                             double* tfar) {
    Ray r = make_ray(v3(org[0], org[1], org[2]), v3(dir[0], dir[1], dir[2]), INFTY);
    double t0 = 0.0, t1 = r.maxd;
    for (int i = 0; i < 3; i++) {
        double tt0 = (lo[i] - vcomp(r.o, i)) * vcomp(r.inv, i);
        double tt1 = (hi[i] - vcomp(r.o, i)) * vcomp(r.inv, i);
        if (tt0 > tt1) {
            double t = tt0;
            tt0 = tt1;
            tt1 = t;
        }
        t0 = fmax(t0, tt0);
        t1 = fmin(t1, tt1);
        if (t0 > t1) return 0;
    }
    *tnear = t0;
    *tfar = t1;
    return 1;
}

EXPORT void orc_camera_ray(const spk_camera_desc* d, int x, int y, const double* u4, double* org, double* dir) {
    Cam c;
    cam_init(&c, d);
    Ray r = cam_ray(&c, x, y, u4[0], u4[1], u4[2], u4[3]);
    org[0] = r.o.x;
    org[1] = r.o.y;
    org[2] = r.o.z;
    dir[0] = r.d.x;
    dir[1] = r.d.y;
    dir[2] = r.d.z;
}

/* Raw counter-based sampler words and their unit values. */
EXPORT void orc_sampler_u32(int kind, uint32_t seed, int64_t n, const uint32_t* pixel, const uint32_t* sample,
                            const uint32_t* dim, uint32_t* out) {
    if (!sobol_ready) sobol_init();
    for (int64_t i = 0; i < n; i++) out[i] = counter_u32(kind, seed, pixel[i], sample[i], dim[i]);
}

EXPORT void* orc_mt_new(uint32_t seed) {
    MT* m = (MT*)malloc(sizeof(MT));
    mt_seed(m, seed);
    return m;
}
EXPORT uint32_t orc_mt_u32(void* m) { return mt_u32((MT*)m); }
EXPORT double orc_mt_real(void* m) { return mt_real2((MT*)m); }
EXPORT void orc_mt_free(void* m) { free(m); }

/* BSDF in its local frame (n = +z): f, pdf, and sampling, for the checks. */
static void local_bsdf(const spk_material* m, BSDF* b) {
    Isect is;
    memset(&is, 0, sizeof(is));
    is.ns = v3(0, 0, 1);
    is.ts = v3(1, 0, 0);
    is.bs = v3(0, 1, 0);
    is.n = v3(0, 0, 1);
    setup_bsdf(m, &is, b);
}
EXPORT void orc_bsdf_eval(const spk_material* m, const double* wo, const double* wi, int mask, double* f_out,
                          double* pdf_out) {
    BSDF b;
    local_bsdf(m, &b);
    V3 f = bsdf_f(&b, v3(wo[0], wo[1], wo[2]), v3(wi[0], wi[1], wi[2]), mask);
    f_out[0] = f.x;
    f_out[1] = f.y;
    f_out[2] = f.z;
    *pdf_out = bsdf_pdf(&b, v3(wo[0], wo[1], wo[2]), v3(wi[0], wi[1], wi[2]), mask);
}
/* u3: the two sample uniforms and the BxDF's hidden uniform */
EXPORT int orc_bsdf_sample(const spk_material* m, const double* wo, const double* u3, int mask, double* wi_out,
                           double* f_out, double* pdf_out) {
    BSDF b;
    local_bsdf(m, &b);
    V3 wi = v3(0, 0, 0);
    double pdf = 0.0;
    int st = 0;
    Hidden hu = {0, {NULL, NULL, NULL}, u3[2]};
    V3 f = bsdf_sample(&b, v3(wo[0], wo[1], wo[2]), &wi, u3[0], u3[1], &pdf, mask, &st, &hu);
    wi_out[0] = wi.x;
    wi_out[1] = wi.y;
    wi_out[2] = wi.z;
    f_out[0] = f.x;
    f_out[1] = f.y;
    f_out[2] = f.z;
    *pdf_out = pdf;
    return st;
}
EXPORT double orc_ggx_D(double alpha, const double* wh) {
    BSDF b;
    memset(&b, 0, sizeof(b));
    b.ax = b.ay = alpha;
    return ggx_D(&b, v3(wh[0], wh[1], wh[2]));
}
EXPORT double orc_fr_dielectric(double cosI, double etaI, double etaT) { return fr_dielectric(cosI, etaI, etaT); }
EXPORT void orc_fr_conductor(double cosI, const double* eta, const double* k, double* out) {
    V3 r = fr_conductor(cosI, v3(1, 1, 1), v3(eta[0], eta[1], eta[2]), v3(k[0], k[1], k[2]));
    out[0] = r.x;
    out[1] = r.y;
    out[2] = r.z;
}

/* Environment-map checks: Le(dir), pdfLi(dir), sampleLi(u) */
EXPORT void orc_env_eval(void* p, int64_t n, const double* dirs, double* le_out, double* pdf_out) {
    Scene* s = (Scene*)p;
    for (int64_t i = 0; i < n; i++) {
        V3 d = v3(dirs[3 * i], dirs[3 * i + 1], dirs[3 * i + 2]);
        V3 le = env_Le(&s->env, d);
        le_out[3 * i] = le.x;
        le_out[3 * i + 1] = le.y;
        le_out[3 * i + 2] = le.z;
        pdf_out[i] = env_pdfLi(&s->env, d);
    }
}
EXPORT void orc_env_sample(void* p, int64_t n, const double* u2, double* dir_out, double* pdf_out, double* li_out) {
    Scene* s = (Scene*)p;
    for (int64_t i = 0; i < n; i++) {
        V3 dir = v3(0, 0, 0), lp;
        double pdf = 0.0;
        V3 li = env_sampleLi(&s->env, v3(0, 0, 0), u2[2 * i], u2[2 * i + 1], &dir, &pdf, &lp);
        dir_out[3 * i] = dir.x;
        dir_out[3 * i + 1] = dir.y;
        dir_out[3 * i + 2] = dir.z;
        pdf_out[i] = pdf;
        li_out[3 * i] = li.x;
        li_out[3 * i + 1] = li.y;
        li_out[3 * i + 2] = li.z;
    }
}

/* Paths with spica's MT19937 Independent sampler, one seeded sampler per path, starting
 * from given rays (the draws PathIntegrator::Li, or DirectLightingIntegrator::Li when
 * `direct`, makes: path.cc:42-125, directlighting.cc:21-57).  The BxDFs' hidden streams
 * start default-seeded at the first path and run on across the batch, as spica's
 * thread_local ones do on one fresh thread. */
EXPORT void orc_li_mt_batch(void* p, int64_t n, const double* org, const double* dir, const uint32_t* seeds,
                            int max_depth, int rr_depth, int direct, double* L_out) {
    Scene* s = (Scene*)p;
    MT hmt[3];
    Hidden hid;
    hidden_init_mt(&hid, hmt);
    for (int64_t i = 0; i < n; i++) {
        MT mt;
        mt_seed(&mt, seeds[i]);
        Sampler sp = {SPK_SAMPLER_MT19937, seeds[i], 0, 0, &mt, &hid};
        Ray r = make_ray(v3(org[3 * i], org[3 * i + 1], org[3 * i + 2]), v3(dir[3 * i], dir[3 * i + 1], dir[3 * i + 2]),
                         INFTY);
        V3 L = direct ? Li_direct(s, r, &sp, max_depth, 0, 0, NULL) : Li_path(s, r, &sp, max_depth, rr_depth, NULL);
        L_out[3 * i] = L.x;
        L_out[3 * i + 1] = L.y;
        L_out[3 * i + 2] = L.z;
    }
}

EXPORT void orc_li_mt(void* p, const double* org, const double* dir, uint32_t seed, int max_depth, int rr_depth,
                      double* L_out) {
    orc_li_mt_batch(p, 1, org, dir, &seed, max_depth, rr_depth, 0, L_out);
}

EXPORT void orc_li_mt_direct(void* p, const double* org, const double* dir, uint32_t seed, int max_depth,
                             double* L_out) {
    orc_li_mt_batch(p, 1, org, dir, &seed, max_depth, 3, 1, L_out);
}

/* ------------------------------------------------------------------------- */
/* Render (integrator.cc:46-110, film.cc:23-40, 65-74)                       */
/* ------------------------------------------------------------------------- */
typedef struct {
    const Scene* s;
    Cam cam;
    spk_render_desc rd;
    const int32_t* subset; /* film indices, or NULL = full frame */
    int64_t npix;
    double *sum, *sumsq;
    int64_t* count;
    int64_t next;
    pthread_mutex_t mu;
    PathStats stats;
} RenderJob;

static void render_pixel(RenderJob* j, int64_t k, PathStats* ps, Hidden* hid) {
    const Cam* c = &j->cam;
    int64_t film = j->subset ? j->subset[k] : k;
    int fy = (int)(film / c->W), fx = (int)(film % c->W);
    int x = c->W - 1 - fx, y = fy; /* integrator.cc:88: film pixel (W-1-x, y) */
    uint32_t pix = (uint32_t)(y * c->W + x);
    MT mt;
    for (int si = 0; si < j->rd.spp; si++) {
        uint32_t sample = (uint32_t)(j->rd.first_sample + si);
        Sampler sp = {j->rd.sampler, j->rd.seed, pix, sample, NULL, hid};
        if (j->rd.sampler == SPK_SAMPLER_MT19937) {
            uint32_t h[4] = {pix, sample, j->rd.seed, 0x2545f491u};
            pcg4d(h);
            mt_seed(&mt, h[0]);
            sp.mt = &mt;
        }
        double f0 = smp(&sp, 0), f1 = smp(&sp, 1);
        double l0 = smp(&sp, 2), l1 = smp(&sp, 3);
        Ray r = cam_ray(c, x, y, f0, f1, l0, l1);
        V3 L;
        if (j->rd.integrator == SPK_INTEGRATOR_DIRECT) {
            Counts cnt = {0, 0, 0};
            L = Li_direct(j->s, r, &sp, j->rd.max_depth, 0, 0, ps ? &cnt : NULL);
            if (ps) {
                ps->rays += cnt.rays;
                ps->nodes += cnt.nodes;
                ps->tris += cnt.tris;
            }
        } else {
            L = Li_path(j->s, r, &sp, j->rd.max_depth, j->rd.rr_depth, ps);
        }
        j->sum[3 * k] += L.x;
        j->sum[3 * k + 1] += L.y;
        j->sum[3 * k + 2] += L.z;
        j->sumsq[3 * k] += L.x * L.x;
        j->sumsq[3 * k + 1] += L.y * L.y;
        j->sumsq[3 * k + 2] += L.z * L.z;
        j->count[k] += 1;
    }
}

static void* render_worker(void* arg) {
    RenderJob* j = (RenderJob*)arg;
    PathStats ps = {0, 0, 0};
    MT hmt[3];
    Hidden hid = {0, {NULL, NULL, NULL}, 0.0};
    if (j->rd.sampler == SPK_SAMPLER_MT19937) hidden_init_mt(&hid, hmt);
    for (;;) {
        pthread_mutex_lock(&j->mu);
        int64_t k0 = j->next;
        j->next += 16;
        pthread_mutex_unlock(&j->mu);
        if (k0 >= j->npix) break;
        int64_t k1 = k0 + 16 < j->npix ? k0 + 16 : j->npix;
        for (int64_t k = k0; k < k1; k++) render_pixel(j, k, &ps, &hid);
    }
    pthread_mutex_lock(&j->mu);
    j->stats.rays += ps.rays;
    j->stats.nodes += ps.nodes;
    j->stats.tris += ps.tris;
    pthread_mutex_unlock(&j->mu);
    return NULL;
}

/* Accumulates rd->spp samples per pixel starting at rd->first_sample into
 * sum / sumsq / count (indexed like `subset`, or film order when NULL).
 * stats_out (optional): closest-hit rays traced on the main path, nodes, tris. */
EXPORT void orc_render(void* p, const spk_camera_desc* cd, const spk_render_desc* rd, const int32_t* subset,
                       int64_t nsubset, int nthreads, double* sum, double* sumsq, int64_t* count, int64_t* stats_out) {
    if (!sobol_ready) sobol_init();
    RenderJob j;
    memset(&j, 0, sizeof(j));
    j.s = (Scene*)p;
    cam_init(&j.cam, cd);
    j.rd = *rd;
    j.subset = subset;
    j.npix = subset ? nsubset : (int64_t)cd->width * cd->height;
    j.sum = sum;
    j.sumsq = sumsq;
    j.count = count;
    pthread_mutex_init(&j.mu, NULL);
    if (nthreads < 1) nthreads = 1;
    if (nthreads > 256) nthreads = 256;
    pthread_t th[256];
    for (int k = 0; k < nthreads; k++) pthread_create(&th[k], NULL, render_worker, &j);
    for (int k = 0; k < nthreads; k++) pthread_join(th[k], NULL);
    pthread_mutex_destroy(&j.mu);
    if (stats_out) {
        stats_out[0] = j.stats.rays;
        stats_out[1] = j.stats.nodes;
        stats_out[2] = j.stats.tris;
    }
}
