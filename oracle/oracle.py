"""ctypes front-end of the CPU checker (liboracle.so) and of spica's own compiled
hot path (_ref/libspica_ref.so).  TEST INFRASTRUCTURE ONLY: imported by tests/,
__graft_entry__.smoke() and bench.py's CPU-baseline leg, never by the product.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_LIB = os.path.join(HERE, "liboracle.so")
REF_LIB = os.path.join(HERE, "_ref", "libspica_ref.so")

_d = ctypes.POINTER(ctypes.c_double)
_i32 = ctypes.POINTER(ctypes.c_int32)
_i64 = ctypes.POINTER(ctypes.c_int64)
_u32 = ctypes.POINTER(ctypes.c_uint32)
_vp = ctypes.c_void_p


def _dp(a):
    return a.ctypes.data_as(_d)


class Oracle:
    def __init__(self, path=ORACLE_LIB):
        if not os.path.exists(path):
            raise RuntimeError(f"{path} missing: run oracle/build.py")
        L = ctypes.CDLL(path)
        L.orc_scene_create.restype = _vp
        L.orc_scene_create.argtypes = [_vp]
        L.orc_scene_free.argtypes = [_vp]
        L.orc_scene_info.argtypes = [_vp, _i64]
        L.orc_trace.argtypes = [_vp, ctypes.c_int64, _d, _d, _d, ctypes.c_int, ctypes.c_int, _i32, _d, _i64, ctypes.c_int]
        L.orc_tri_intersect.restype = ctypes.c_int
        L.orc_tri_intersect.argtypes = [_d, _d, _d, ctypes.c_double, _d]
        L.orc_box_intersect.restype = ctypes.c_int
        L.orc_box_intersect.argtypes = [_d, _d, _d, _d, _d, _d]
        L.orc_camera_ray.argtypes = [_vp, ctypes.c_int, ctypes.c_int, _d, _d, _d]
        L.orc_sampler_u32.argtypes = [ctypes.c_int, ctypes.c_uint32, ctypes.c_int64, _u32, _u32, _u32, _u32]
        L.orc_mt_new.restype = _vp
        L.orc_mt_new.argtypes = [ctypes.c_uint32]
        L.orc_mt_u32.restype = ctypes.c_uint32
        L.orc_mt_u32.argtypes = [_vp]
        L.orc_mt_real.restype = ctypes.c_double
        L.orc_mt_real.argtypes = [_vp]
        L.orc_mt_free.argtypes = [_vp]
        L.orc_bsdf_eval.argtypes = [_vp, _d, _d, ctypes.c_int, _d, _d]
        L.orc_bsdf_sample.restype = ctypes.c_int
        L.orc_bsdf_sample.argtypes = [_vp, _d, _d, ctypes.c_int, _d, _d, _d]
        L.orc_ggx_D.restype = ctypes.c_double
        L.orc_ggx_D.argtypes = [ctypes.c_double, _d]
        L.orc_fr_dielectric.restype = ctypes.c_double
        L.orc_fr_dielectric.argtypes = [ctypes.c_double, ctypes.c_double, ctypes.c_double]
        L.orc_fr_conductor.argtypes = [ctypes.c_double, _d, _d, _d]
        L.orc_env_eval.argtypes = [_vp, ctypes.c_int64, _d, _d, _d]
        L.orc_env_sample.argtypes = [_vp, ctypes.c_int64, _d, _d, _d, _d]
        L.orc_li_mt.argtypes = [_vp, _d, _d, ctypes.c_uint32, ctypes.c_int, ctypes.c_int, _d]
        L.orc_li_mt_direct.argtypes = [_vp, _d, _d, ctypes.c_uint32, ctypes.c_int, _d]
        L.orc_li_mt_batch.argtypes = [_vp, ctypes.c_int64, _d, _d, _u32, ctypes.c_int, ctypes.c_int, ctypes.c_int, _d]
        L.orc_render.argtypes = [_vp, _vp, _vp, _i32, ctypes.c_int64, ctypes.c_int, _d, _d, _i64, _i64]
        self.L = L

    # -- scenes --------------------------------------------------------------
    def scene(self, scene):
        return OracleScene(self, scene)

    # -- primitives ------------------------------------------------------------
    def tri_intersect(self, p9, org, dirv, maxd=1e32):
        p9 = np.ascontiguousarray(p9, dtype=np.float64).reshape(9)
        o = np.ascontiguousarray(org, dtype=np.float64)
        d = np.ascontiguousarray(dirv, dtype=np.float64)
        t = ctypes.c_double(-1.0)
        h = self.L.orc_tri_intersect(_dp(p9), _dp(o), _dp(d), maxd, ctypes.byref(t))
        return bool(h), t.value

    def box_intersect(self, lo, hi, org, dirv):
        lo, hi, o, d = (np.ascontiguousarray(x, dtype=np.float64) for x in (lo, hi, org, dirv))
        tn, tf = ctypes.c_double(), ctypes.c_double()
        h = self.L.orc_box_intersect(_dp(lo), _dp(hi), _dp(o), _dp(d), ctypes.byref(tn), ctypes.byref(tf))
        return bool(h), tn.value, tf.value

    def camera_ray(self, camera, x, y, u4):
        cd = camera.desc()
        u = np.ascontiguousarray(u4, dtype=np.float64)
        o = np.zeros(3)
        d = np.zeros(3)
        self.L.orc_camera_ray(ctypes.byref(cd), x, y, _dp(u), _dp(o), _dp(d))
        return o, d

    def sampler_u32(self, kind, seed, pixel, sample, dim):
        pixel, sample, dim = (np.ascontiguousarray(a, dtype=np.uint32) for a in (pixel, sample, dim))
        out = np.zeros(len(pixel), dtype=np.uint32)
        self.L.orc_sampler_u32(kind, seed, len(pixel), pixel.ctypes.data_as(_u32), sample.ctypes.data_as(_u32),
                               dim.ctypes.data_as(_u32), out.ctypes.data_as(_u32))
        return out

    def mt_u32(self, seed, n):
        m = self.L.orc_mt_new(seed)
        out = np.array([self.L.orc_mt_u32(m) for _ in range(n)], dtype=np.uint64)
        self.L.orc_mt_free(m)
        return out

    def mt_real(self, seed, n):
        m = self.L.orc_mt_new(seed)
        out = np.array([self.L.orc_mt_real(m) for _ in range(n)])
        self.L.orc_mt_free(m)
        return out

    def bsdf_eval(self, mat, wo, wi, mask=31):
        from spica_amd.scene import _material_struct

        m = _material_struct(mat)
        wo, wi = (np.ascontiguousarray(x, dtype=np.float64) for x in (wo, wi))
        f = np.zeros(3)
        pdf = ctypes.c_double()
        self.L.orc_bsdf_eval(ctypes.byref(m), _dp(wo), _dp(wi), mask, _dp(f), ctypes.byref(pdf))
        return f, pdf.value

    def bsdf_sample(self, mat, wo, u2, mask=31):
        """u2: the two sample uniforms, optionally a third for the BxDF's hidden draw (0.5)."""
        from spica_amd.scene import _material_struct

        m = _material_struct(mat)
        u2 = list(u2) + [0.5] * (3 - len(u2))
        wo, u2 = (np.ascontiguousarray(x, dtype=np.float64) for x in (wo, u2))
        wi = np.zeros(3)
        f = np.zeros(3)
        pdf = ctypes.c_double()
        st = self.L.orc_bsdf_sample(ctypes.byref(m), _dp(wo), _dp(u2), mask, _dp(wi), _dp(f), ctypes.byref(pdf))
        return wi, f, pdf.value, st

    def ggx_D(self, alpha, wh):
        wh = np.ascontiguousarray(wh, dtype=np.float64)
        return self.L.orc_ggx_D(alpha, _dp(wh))

    def fr_dielectric(self, c, ei, et):
        return self.L.orc_fr_dielectric(c, ei, et)

    def fr_conductor(self, c, eta, k):
        eta, k = (np.ascontiguousarray(x, dtype=np.float64) for x in (eta, k))
        out = np.zeros(3)
        self.L.orc_fr_conductor(c, _dp(eta), _dp(k), _dp(out))
        return out


class OracleScene:
    def __init__(self, orc, scene):
        self.orc = orc
        self.scene = scene
        self._desc = scene.desc()
        self.h = orc.L.orc_scene_create(ctypes.byref(self._desc))
        info = np.zeros(4, dtype=np.int64)
        orc.L.orc_scene_info(self.h, info.ctypes.data_as(_i64))
        self.num_tris, self.num_nodes, self.max_depth, self.num_lights = (int(x) for x in info)

    def trace(self, org, dirv, tmax=None, any_hit=False, brute=False, threads=8):
        org = np.ascontiguousarray(org, dtype=np.float64)
        dirv = np.ascontiguousarray(dirv, dtype=np.float64)
        n = len(org)
        prim = np.zeros(n, dtype=np.int32)
        t = np.zeros(n)
        cnt = np.zeros(2, dtype=np.int64)
        tm = None if tmax is None else np.ascontiguousarray(tmax, dtype=np.float64)
        self.orc.L.orc_trace(self.h, n, _dp(org), _dp(dirv), None if tm is None else _dp(tm), 1 if any_hit else 0,
                             1 if brute else 0, prim.ctypes.data_as(_i32), _dp(t), cnt.ctypes.data_as(_i64), threads)
        return prim, t, cnt

    def render(self, camera, render_desc, threads=8, subset=None, state=None):
        """Accumulates render_desc.spp samples; returns (sum, sumsq, count, stats) in film order."""
        cd = camera.desc()
        npix = camera.width * camera.height if subset is None else len(subset)
        if state is None:
            state = (np.zeros((npix, 3)), np.zeros((npix, 3)), np.zeros(npix, dtype=np.int64))
        s, q, c = state
        stats = np.zeros(3, dtype=np.int64)
        sub = None if subset is None else np.ascontiguousarray(subset, dtype=np.int32)
        self.orc.L.orc_render(self.h, ctypes.byref(cd), ctypes.byref(render_desc),
                              None if sub is None else sub.ctypes.data_as(_i32), npix, threads, _dp(s), _dp(q),
                              c.ctypes.data_as(_i64), stats.ctypes.data_as(_i64))
        return s, q, c, stats

    def env_eval(self, dirs):
        dirs = np.ascontiguousarray(dirs, dtype=np.float64)
        le = np.zeros_like(dirs)
        pdf = np.zeros(len(dirs))
        self.orc.L.orc_env_eval(self.h, len(dirs), _dp(dirs), _dp(le), _dp(pdf))
        return le, pdf

    def env_sample(self, u2):
        u2 = np.ascontiguousarray(u2, dtype=np.float64)
        n = len(u2)
        d = np.zeros((n, 3))
        pdf = np.zeros(n)
        li = np.zeros((n, 3))
        self.orc.L.orc_env_sample(self.h, n, _dp(u2), _dp(d), _dp(pdf), _dp(li))
        return d, pdf, li

    def li_mt(self, org, dirv, seed, max_depth, rr_depth=3):
        o, d = (np.ascontiguousarray(x, dtype=np.float64) for x in (org, dirv))
        L = np.zeros(3)
        self.orc.L.orc_li_mt(self.h, _dp(o), _dp(d), seed, max_depth, rr_depth, _dp(L))
        return L

    def li_mt_batch(self, org, dirv, seeds, max_depth, rr_depth=3, direct=False):
        """Paths from given rays with spica's MT19937 streams, one seed per path; the BxDFs'
        hidden streams run on across the batch (spica's thread_local Random on one thread)."""
        org = np.ascontiguousarray(org, dtype=np.float64)
        dirv = np.ascontiguousarray(dirv, dtype=np.float64)
        seeds = np.ascontiguousarray(seeds, dtype=np.uint32)
        L = np.zeros((len(org), 3))
        self.orc.L.orc_li_mt_batch(self.h, len(org), _dp(org), _dp(dirv), seeds.ctypes.data_as(_u32), max_depth, rr_depth,
                                   1 if direct else 0, _dp(L))
        return L

    def li_mt_direct(self, org, dirv, seed, max_depth):
        """DirectLightingIntegrator::Li with spica's MT19937 stream seeded `seed`."""
        o, d = (np.ascontiguousarray(x, dtype=np.float64) for x in (org, dirv))
        L = np.zeros(3)
        self.orc.L.orc_li_mt_direct(self.h, _dp(o), _dp(d), seed, max_depth, _dp(L))
        return L

    def __del__(self):
        try:
            self.orc.L.orc_scene_free(self.h)
        except Exception:
            pass


class Reference:
    """spica's own PathIntegrator / BVHAccel / BxDF code, compiled by oracle/build.py."""

    def __init__(self, path=REF_LIB):
        if not os.path.exists(path):
            raise RuntimeError(f"{path} missing")
        L = ctypes.CDLL(path, mode=os.RTLD_LAZY)
        L.ref_scene_create.restype = _vp
        L.ref_scene_create.argtypes = [_vp]
        L.ref_scene_free.argtypes = [_vp]
        L.ref_trace.argtypes = [_vp, ctypes.c_int64, _d, _d, ctypes.c_int, _i32, _d]
        L.ref_li.argtypes = [_vp, ctypes.c_int64, _d, _d, _u32, ctypes.c_int, _d]
        L.ref_li_direct.argtypes = [_vp, ctypes.c_int64, _d, _d, _u32, ctypes.c_int, _d]
        L.ref_mt.argtypes = [ctypes.c_uint32, ctypes.c_int64, _d]
        L.ref_bxdf_sample.restype = ctypes.c_int
        L.ref_bxdf_sample.argtypes = [_vp, _d, _d, _d, _d, _d]
        L.ref_bxdf_eval.argtypes = [_vp, _d, _d, _d, _d]
        self.L = L

    def scene(self, scene):
        return RefScene(self, scene)

    def mt_real(self, seed, n):
        out = np.zeros(n)
        self.L.ref_mt(seed, n, _dp(out))
        return out

    def bxdf_sample(self, mat, wo, u2):
        from spica_amd.scene import _material_struct

        m = _material_struct(mat)
        wo, u2 = (np.ascontiguousarray(x, dtype=np.float64) for x in (wo, u2))
        wi, f = np.zeros(3), np.zeros(3)
        pdf = ctypes.c_double()
        t = self.L.ref_bxdf_sample(ctypes.byref(m), _dp(wo), _dp(u2), _dp(wi), _dp(f), ctypes.byref(pdf))
        return wi, f, pdf.value, t

    def bxdf_eval(self, mat, wo, wi):
        from spica_amd.scene import _material_struct

        m = _material_struct(mat)
        wo, wi = (np.ascontiguousarray(x, dtype=np.float64) for x in (wo, wi))
        f = np.zeros(3)
        pdf = ctypes.c_double()
        self.L.ref_bxdf_eval(ctypes.byref(m), _dp(wo), _dp(wi), _dp(f), ctypes.byref(pdf))
        return f, pdf.value


class RefScene:
    def __init__(self, ref, scene):
        self.ref = ref
        self._desc = scene.desc()
        self.h = ref.L.ref_scene_create(ctypes.byref(self._desc))

    def trace(self, org, dirv, any_hit=False):
        org = np.ascontiguousarray(org, dtype=np.float64)
        dirv = np.ascontiguousarray(dirv, dtype=np.float64)
        n = len(org)
        prim = np.zeros(n, dtype=np.int32)
        t = np.zeros(n)
        self.ref.L.ref_trace(self.h, n, _dp(org), _dp(dirv), 1 if any_hit else 0, prim.ctypes.data_as(_i32), _dp(t))
        return prim, t

    def li(self, org, dirv, seeds, max_depth, direct=False):
        """spica's PathIntegrator::Li (or DirectLightingIntegrator::Li) per ray."""
        org = np.ascontiguousarray(org, dtype=np.float64)
        dirv = np.ascontiguousarray(dirv, dtype=np.float64)
        seeds = np.ascontiguousarray(seeds, dtype=np.uint32)
        n = len(org)
        L = np.zeros((n, 3))
        fn = self.ref.L.ref_li_direct if direct else self.ref.L.ref_li
        fn(self.h, n, _dp(org), _dp(dirv), seeds.ctypes.data_as(_u32), max_depth, _dp(L))
        return L

    def __del__(self):
        try:
            self.ref.L.ref_scene_free(self.h)
        except Exception:
            pass


def have_ref():
    return os.path.exists(REF_LIB)
