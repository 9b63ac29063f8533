"""Python handles over the C ABI (include/spk.h).

`GpuScene` owns an spk_scene (BVH + geometry + materials + lights in HBM) and
answers batched ray queries on device tensors; `GpuRender` owns an spk_render
(path queues + film) and accumulates samples per pixel.
"""
import ctypes

import numpy as np

from . import _abi

def init(device=0):
    """Loads the library and makes `device` current for the calling thread (the library, like
    HIP, keeps the current device per thread; the call is cheap, so scenes may be built from
    several threads)."""
    lib = _abi.load()
    _abi.check(lib.spk_init(device), "spk_init")
    return lib


def _ptr(t):
    return ctypes.c_void_p(0 if t is None else t.data_ptr())


class GpuScene:
    """builder: "sah" (host binned SAH, spica's BVH) or "lbvh" (linear BVH built on the device);
    treelet_passes: LBVH treelet-restructuring sweeps (None = the library default, 6)."""

    def __init__(self, scene, device=0, builder="sah", max_leaf=4, treelet_passes=None):
        self.lib = init(device)
        self.scene = scene
        self._desc = scene.desc()
        bo = _abi.SpkSceneBuildOpts()
        _abi.check(self.lib.spk_scene_build_opts_default(ctypes.byref(bo)), "spk_scene_build_opts_default")
        bo.builder = {"sah": _abi.SPK_BVH_SAH, "lbvh": _abi.SPK_BVH_LBVH}[builder]
        bo.max_leaf = max_leaf
        if treelet_passes is not None:
            bo.treelet_passes = int(treelet_passes)
        h = ctypes.c_void_p()
        _abi.check(self.lib.spk_scene_create_ex(ctypes.byref(self._desc), ctypes.byref(bo), ctypes.byref(h)),
                   "spk_scene_create_ex")
        self.handle = h

    def stats(self):
        st = _abi.SpkSceneStats()
        _abi.check(self.lib.spk_scene_stats_get(self.handle, ctypes.byref(st)), "spk_scene_stats_get")
        return {k: getattr(st, k) for k, _ in st._fields_}

    def trace(self, org, dir, tmax=None, any_hit=False, stream=None, noalloc=False):
        """org/dir: float32 [n, 3] device tensors; returns (prim int32 [n], t float32 [n] or None).
        noalloc: trace through spk_trace_noalloc, whose scratch is a fixed per-GPU array (no
        allocation; calls overlapping on one GPU race on it)."""
        import torch

        assert org.is_cuda and dir.is_cuda and org.dtype == torch.float32 and dir.dtype == torch.float32
        org = org.contiguous()
        dir = dir.contiguous()
        n = org.shape[0]
        assert org.shape == (n, 3) and dir.shape == (n, 3)
        if tmax is not None:
            tmax = tmax.contiguous().float()
            assert tmax.shape == (n,)
        prim = torch.empty(n, dtype=torch.int32, device=org.device)
        t = None if any_hit else torch.empty(n, dtype=torch.float32, device=org.device)
        st = torch.cuda.current_stream().cuda_stream if stream is None else stream
        fn, name = (self.lib.spk_trace_noalloc, "spk_trace_noalloc") if noalloc else (self.lib.spk_trace, "spk_trace")
        _abi.check(fn(self.handle, _ptr(org), _ptr(dir), _ptr(tmax), n, 1 if any_hit else 0, _ptr(prim), _ptr(t),
                      ctypes.c_void_p(st)), name)
        return prim, t

    def close(self):
        if self.handle:
            self.lib.spk_scene_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class GpuRender:
    def __init__(self, gscene, camera, render_desc, queue_size=0, profile=False, use_graph=False, count_visits=False,
                 block_size=0, regen=True):
        """queue_size: paths in flight; regen: passes refill their queues with new camera paths as
        paths end (False: classic passes, all paths at once, queues shrinking by bounce); profile:
        HIP events around the closest-hit launches (1 or True) or around every stage (2)."""
        self.lib = gscene.lib
        self.gscene = gscene
        self.camera = camera
        self._cam = camera.desc()
        self._rd = render_desc
        self._opts = _abi.SpkRenderOpts()
        self._opts.queue_size = int(queue_size)
        self._opts.block_size = int(block_size)
        self._opts.profile = int(profile) if profile else 0
        self._opts.use_graph = 1 if use_graph else 0
        self._opts.count_visits = 1 if count_visits else 0
        self._opts.classic_passes = 0 if regen else 1
        h = ctypes.c_void_p()
        _abi.check(self.lib.spk_render_create(gscene.handle, ctypes.byref(self._cam), ctypes.byref(self._rd),
                                              ctypes.byref(self._opts), ctypes.byref(h)), "spk_render_create")
        self.handle = h

    def run(self, n_spp, sync=True):
        _abi.check(self.lib.spk_render_run(self.handle, int(n_spp), 1 if sync else 0), "spk_render_run")

    def sync(self):
        _abi.check(self.lib.spk_render_sync(self.handle), "spk_render_sync")

    def film(self):
        """(sum [h, w, 3], sumsq [h, w, 3], count [h, w]) as float64, film order (x mirrored as spica)."""
        w, h = self.camera.width, self.camera.height
        s = np.zeros((h, w, 3))
        q = np.zeros((h, w, 3))
        c = np.zeros((h, w))
        _abi.check(self.lib.spk_render_film(self.handle, s.ctypes.data_as(_abi.c_d_p), q.ctypes.data_as(_abi.c_d_p),
                                            c.ctypes.data_as(_abi.c_d_p)), "spk_render_film")
        return s, q, c

    def image(self):
        w, h = self.camera.width, self.camera.height
        img = np.zeros((h, w, 3), dtype=np.float32)
        _abi.check(self.lib.spk_render_image(self.handle, img.ctypes.data_as(_abi.c_f_p)), "spk_render_image")
        return img

    def relmse(self, ref, eps=1e-2):
        ref = np.ascontiguousarray(ref, dtype=np.float64)
        out = ctypes.c_double()
        _abi.check(self.lib.spk_render_relmse(self.handle, ref.ctypes.data_as(_abi.c_d_p), eps, ctypes.byref(out)),
                   "spk_render_relmse")
        return out.value

    def relmse_pixels(self, pixels, ref, eps=1e-2):
        """relMSE over the film pixels `pixels` (film indices) against ref [n, 3], reduced on the GPU in float64."""
        pixels = np.ascontiguousarray(pixels, dtype=np.int32)
        ref = np.ascontiguousarray(ref, dtype=np.float64)
        assert ref.shape == (len(pixels), 3)
        out = ctypes.c_double()
        _abi.check(self.lib.spk_render_relmse_pixels(self.handle, len(pixels), pixels.ctypes.data_as(_abi.c_i32_p),
                                                     ref.ctypes.data_as(_abi.c_d_p), eps, ctypes.byref(out)),
                   "spk_render_relmse_pixels")
        return out.value

    def set_guide(self, guide):
        """Attach a Guide (learned proposal, trained online) or None for plain BSDF sampling."""
        self.guide = guide
        _abi.check(self.lib.spk_render_set_guide(self.handle, guide.handle if guide is not None else None),
                   "spk_render_set_guide")

    def stats(self):
        st = _abi.SpkRenderStats()
        _abi.check(self.lib.spk_render_stats_get(self.handle, ctypes.byref(st)), "spk_render_stats_get")
        return {k: getattr(st, k) for k, _ in st._fields_}

    def close(self):
        if self.handle:
            self.lib.spk_render_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Batch:
    """Many images (scene, camera, spp, seed each) rendered through shared wavefront passes."""

    def __init__(self, items, max_depth=16, rr_depth=3, sampler=_abi.SPK_SAMPLER_SOBOL, queue_size=0, block_size=0,
                 graph=1, regen=True):
        """items: list of (GpuScene, Camera, spp, seed)."""
        self.lib = _abi.load()
        self.items = list(items)
        arr = (_abi.SpkBatchImage * len(self.items))()
        for a, (gs, cam, spp, seed) in zip(arr, self.items):
            a.scene = gs.handle
            a.camera = cam.desc()
            a.spp = int(spp)
            a.seed = int(seed)
        self._arr = arr
        d = _abi.SpkBatchDesc()
        d.num_images = len(self.items)
        d.images = arr
        d.max_depth = max_depth
        d.rr_depth = rr_depth
        d.sampler = sampler
        d.graph = graph
        o = _abi.SpkRenderOpts()
        o.queue_size = queue_size
        o.block_size = block_size
        o.classic_passes = 0 if regen else 1
        h = ctypes.c_void_p()
        _abi.check(self.lib.spk_batch_create(ctypes.byref(d), ctypes.byref(o), ctypes.byref(h)), "spk_batch_create")
        self.handle = h

    def set_guide(self, guide):
        """Attach a Guide (all images must render one scene) or None; it trains on the first
        train_spp samples of every image, in rounds across the images."""
        self.guide = guide
        _abi.check(self.lib.spk_batch_set_guide(self.handle, guide.handle if guide is not None else None),
                   "spk_batch_set_guide")

    def run(self, sync=True):
        _abi.check(self.lib.spk_batch_run(self.handle, 1 if sync else 0), "spk_batch_run")

    def film(self, image):
        cam = self.items[image][1]
        w, h = cam.width, cam.height
        s = np.zeros((h, w, 3))
        q = np.zeros((h, w, 3))
        c = np.zeros((h, w))
        _abi.check(self.lib.spk_batch_film(self.handle, image, s.ctypes.data_as(_abi.c_d_p), q.ctypes.data_as(_abi.c_d_p),
                                           c.ctypes.data_as(_abi.c_d_p)), "spk_batch_film")
        return s, q, c

    def info(self):
        st = _abi.SpkBatchInfo()
        _abi.check(self.lib.spk_batch_info_get(self.handle, ctypes.byref(st)), "spk_batch_info_get")
        return {k: getattr(st, k) for k, _ in st._fields_}

    def stats(self):
        st = _abi.SpkRenderStats()
        _abi.check(self.lib.spk_batch_stats_get(self.handle, ctypes.byref(st)), "spk_batch_stats_get")
        return {k: getattr(st, k) for k, _ in st._fields_}

    def close(self):
        if self.handle:
            self.lib.spk_batch_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def guide_opts(**kw):
    o = _abi.SpkGuideOpts()
    _abi.check(_abi.load().spk_guide_opts_default(ctypes.byref(o)), "spk_guide_opts_default")
    names = {f[0] for f in o._fields_}
    for k, v in kw.items():
        if k not in names:
            raise TypeError(f"unknown guide option {k!r}")
        setattr(o, k, v)
    return o


class Guide:
    """The learned path-guiding proposal (one-blob MLP -> 8-lobe vMF mixture) on the GPU."""

    NPARAMS = 29480

    def __init__(self, device=0, **opts):
        self.lib = _abi.load()
        self.opts = guide_opts(**opts)
        h = ctypes.c_void_p()
        _abi.check(self.lib.spk_guide_create(device, ctypes.byref(self.opts), ctypes.byref(h)), "spk_guide_create")
        self.handle = h

    def params(self):
        out = np.zeros(self.NPARAMS, dtype=np.float32)
        _abi.check(self.lib.spk_guide_params_get(self.handle, out.ctypes.data_as(_abi.c_f_p)), "spk_guide_params_get")
        return out

    def set_params(self, th):
        th = np.ascontiguousarray(th, dtype=np.float32)
        assert th.shape == (self.NPARAMS,)
        _abi.check(self.lib.spk_guide_params_set(self.handle, th.ctypes.data_as(_abi.c_f_p)), "spk_guide_params_set")

    @staticmethod
    def _f32(a, shape):
        a = np.ascontiguousarray(a, dtype=np.float32)
        assert a.shape == shape, (a.shape, shape)
        return a

    def eval(self, x, dirs):
        n = len(x)
        x, dirs = self._f32(x, (n, 7)), self._f32(dirs, (n, 3))
        raw = np.zeros((n, 40), dtype=np.float32)
        lp = np.zeros(n, dtype=np.float32)
        _abi.check(self.lib.spk_guide_eval(self.handle, n, x.ctypes.data_as(_abi.c_f_p), dirs.ctypes.data_as(_abi.c_f_p),
                                           raw.ctypes.data_as(_abi.c_f_p), lp.ctypes.data_as(_abi.c_f_p)),
                   "spk_guide_eval")
        return raw, lp

    def sample(self, x, u):
        n = len(x)
        x, u = self._f32(x, (n, 7)), self._f32(u, (n, 3))
        d = np.zeros((n, 3), dtype=np.float32)
        pdf = np.zeros(n, dtype=np.float32)
        _abi.check(self.lib.spk_guide_sample(self.handle, n, x.ctypes.data_as(_abi.c_f_p), u.ctypes.data_as(_abi.c_f_p),
                                             d.ctypes.data_as(_abi.c_f_p), pdf.ctypes.data_as(_abi.c_f_p)),
                   "spk_guide_sample")
        return d, pdf

    def grad(self, x, dirs, weight):
        n = len(x)
        x, dirs, weight = self._f32(x, (n, 7)), self._f32(dirs, (n, 3)), self._f32(weight, (n,))
        g = np.zeros(self.NPARAMS, dtype=np.float32)
        loss = ctypes.c_double()
        _abi.check(self.lib.spk_guide_grad(self.handle, n, x.ctypes.data_as(_abi.c_f_p), dirs.ctypes.data_as(_abi.c_f_p),
                                           weight.ctypes.data_as(_abi.c_f_p), g.ctypes.data_as(_abi.c_f_p),
                                           ctypes.byref(loss)), "spk_guide_grad")
        return loss.value, g

    def train(self, x, dirs, weight, steps=1):
        n = len(x)
        x, dirs, weight = self._f32(x, (n, 7)), self._f32(dirs, (n, 3)), self._f32(weight, (n,))
        _abi.check(self.lib.spk_guide_train(self.handle, n, x.ctypes.data_as(_abi.c_f_p), dirs.ctypes.data_as(_abi.c_f_p),
                                            weight.ctypes.data_as(_abi.c_f_p), int(steps)), "spk_guide_train")

    def set_train_spp(self, n):
        """Training samples per pixel of the next renders it is attached to (0 = frozen)."""
        _abi.check(self.lib.spk_guide_set_train_spp(self.handle, int(n)), "spk_guide_set_train_spp")

    def stats(self):
        st = _abi.SpkGuideStats()
        _abi.check(self.lib.spk_guide_stats_get(self.handle, ctypes.byref(st)), "spk_guide_stats_get")
        return {k: getattr(st, k) for k, _ in st._fields_}

    def close(self):
        if self.handle:
            self.lib.spk_guide_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def sampler_u32(kind, seed, pixel, sample, dim, device=False):
    """Raw 32-bit sample words of the counter-based streams (host or GPU evaluation)."""
    lib = _abi.load()
    pixel = np.ascontiguousarray(pixel, dtype=np.uint32)
    sample = np.ascontiguousarray(sample, dtype=np.uint32)
    dim = np.ascontiguousarray(dim, dtype=np.uint32)
    n = len(pixel)
    out = np.zeros(n, dtype=np.uint32)
    if device:
        init(0)
    _abi.check(lib.spk_sampler_u32(kind, seed, n, pixel.ctypes.data_as(_abi.c_u32_p), sample.ctypes.data_as(_abi.c_u32_p),
                                   dim.ctypes.data_as(_abi.c_u32_p), out.ctypes.data_as(_abi.c_u32_p), 1 if device else 0),
               "spk_sampler_u32")
    return out
