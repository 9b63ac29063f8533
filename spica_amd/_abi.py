"""ctypes mirror of include/spk_types.h and include/spk.h, and the loader of libspk.so.

The library is loaded from the package directory only; a missing or stale build is
an error (there is no CPU fallback for the render path).
"""
import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libspk.so")

SPK_MAT_DIFFUSE = 1
SPK_MAT_CONDUCTOR = 2
SPK_MAT_DIELECTRIC = 3
SPK_MAT_ROUGHCONDUCTOR = 4
SPK_MAT_PLASTIC = 5
SPK_MAT_ROUGHDIELECTRIC = 6
SPK_MAT_ROUGHPLASTIC = 7

SPK_MICROFACET_GGX = 0
SPK_MICROFACET_BECKMANN = 1

SPK_SAMPLER_INDEPENDENT = 0
SPK_SAMPLER_SOBOL = 1
SPK_SAMPLER_MT19937 = 2

SPK_INTEGRATOR_PATH = 0
SPK_INTEGRATOR_DIRECT = 1

SPK_BVH_SAH = 0
SPK_BVH_LBVH = 1

SPK_OK = 0
SPK_ERR_INVALID = 1
SPK_ERR_HIP = 2
SPK_ERR_NOMEM = 3
SPK_ERR_STATE = 4

c_f_p = ctypes.POINTER(ctypes.c_float)
c_d_p = ctypes.POINTER(ctypes.c_double)
c_i32_p = ctypes.POINTER(ctypes.c_int32)
c_u32_p = ctypes.POINTER(ctypes.c_uint32)


class SpkMaterial(ctypes.Structure):
    _fields_ = [
        ("type", ctypes.c_int32),
        ("distribution", ctypes.c_int32),
        ("reflectance", ctypes.c_double * 3),
        ("transmittance", ctypes.c_double * 3),
        ("eta", ctypes.c_double * 3),
        ("k", ctypes.c_double * 3),
        ("int_ior", ctypes.c_double),
        ("alpha", ctypes.c_double),
        ("specular", ctypes.c_double * 3),
    ]


class SpkSceneDesc(ctypes.Structure):
    _fields_ = [
        ("num_tris", ctypes.c_int64),
        ("verts", c_f_p),
        ("normals", c_f_p),
        ("mat_ids", c_i32_p),
        ("light_ids", c_i32_p),
        ("num_materials", ctypes.c_int32),
        ("num_area_lights", ctypes.c_int32),
        ("materials", ctypes.POINTER(SpkMaterial)),
        ("area_tri", c_i32_p),
        ("area_radiance", c_d_p),
        ("env_width", ctypes.c_int32),
        ("env_height", ctypes.c_int32),
        ("env_rgb", c_f_p),
        ("env_to_world", ctypes.c_double * 16),
        ("env_scale", ctypes.c_double),
        ("env_center", ctypes.c_double * 3),
        ("env_radius", ctypes.c_double),
    ]


class SpkCameraDesc(ctypes.Structure):
    _fields_ = [
        ("width", ctypes.c_int32),
        ("height", ctypes.c_int32),
        ("to_world", ctypes.c_double * 16),
        ("fov_deg", ctypes.c_double),
        ("aperture_radius", ctypes.c_double),
        ("focus_distance", ctypes.c_double),
    ]


class SpkRenderDesc(ctypes.Structure):
    _fields_ = [
        ("spp", ctypes.c_int32),
        ("max_depth", ctypes.c_int32),
        ("sampler", ctypes.c_int32),
        ("seed", ctypes.c_uint32),
        ("rr_depth", ctypes.c_int32),
        ("first_sample", ctypes.c_int32),
        ("integrator", ctypes.c_int32),
    ]


class SpkSceneBuildOpts(ctypes.Structure):
    _fields_ = [("builder", ctypes.c_int32), ("max_leaf", ctypes.c_int32), ("treelet_passes", ctypes.c_int32)]


class SpkSceneStats(ctypes.Structure):
    _fields_ = [
        ("num_tris", ctypes.c_int64),
        ("num_nodes", ctypes.c_int64),
        ("num_leaves", ctypes.c_int64),
        ("max_depth", ctypes.c_int32),
        ("num_lights", ctypes.c_int32),
        ("build_ms", ctypes.c_double),
        ("upload_ms", ctypes.c_double),
        ("device_bytes", ctypes.c_int64),
    ]


class SpkRenderOpts(ctypes.Structure):
    _fields_ = [
        ("queue_size", ctypes.c_int64),
        ("block_size", ctypes.c_int32),
        ("use_graph", ctypes.c_int32),
        ("profile", ctypes.c_int32),
        ("count_visits", ctypes.c_int32),
        ("classic_passes", ctypes.c_int32),
    ]


class SpkRenderStats(ctypes.Structure):
    _fields_ = [
        ("samples_done", ctypes.c_int64),
        ("paths", ctypes.c_int64),
        ("rays_extend", ctypes.c_int64),
        ("rays_shadow", ctypes.c_int64),
        ("rays_mis", ctypes.c_int64),
        ("passes", ctypes.c_int64),
        ("extend_ms", ctypes.c_double),
        ("extend_launches", ctypes.c_int64),
        ("extend_rays_timed", ctypes.c_int64),
        ("shadow_ms", ctypes.c_double),
        ("shadow_launches", ctypes.c_int64),
        ("shade_ms", ctypes.c_double),
        ("count_nodes", ctypes.c_int64),
        ("count_tris", ctypes.c_int64),
        ("count_rays", ctypes.c_int64),
        ("guided_vertices", ctypes.c_int64),
        ("train_passes", ctypes.c_int64),
        ("guide_ms", ctypes.c_double),
        ("guide_launches", ctypes.c_int64),
        ("shade_launches", ctypes.c_int64),
        ("shadow_rays_timed", ctypes.c_int64),
        ("count_nodes_shadow", ctypes.c_int64),
        ("count_tris_shadow", ctypes.c_int64),
        ("count_rays_shadow", ctypes.c_int64),
    ]


class SpkBatchImage(ctypes.Structure):
    _fields_ = [
        ("scene", ctypes.c_void_p),
        ("camera", SpkCameraDesc),
        ("spp", ctypes.c_int32),
        ("seed", ctypes.c_uint32),
    ]


class SpkBatchDesc(ctypes.Structure):
    _fields_ = [
        ("num_images", ctypes.c_int32),
        ("images", ctypes.POINTER(SpkBatchImage)),
        ("max_depth", ctypes.c_int32),
        ("rr_depth", ctypes.c_int32),
        ("sampler", ctypes.c_int32),
        ("graph", ctypes.c_int32),
    ]


class SpkBatchInfo(ctypes.Structure):
    _fields_ = [
        ("passes", ctypes.c_int64),
        ("chunks", ctypes.c_int64),
        ("paths", ctypes.c_int64),
        ("film_pixels", ctypes.c_int64),
        ("queue", ctypes.c_int64),
    ]


class SpkGuideOpts(ctypes.Structure):
    _fields_ = [
        ("seed", ctypes.c_uint32),
        ("batch_size", ctypes.c_int32),
        ("steps_per_pass", ctypes.c_int32),
        ("train_spp", ctypes.c_int32),
        ("learning_rate", ctypes.c_float),
        ("beta1", ctypes.c_float),
        ("beta2", ctypes.c_float),
        ("epsilon", ctypes.c_float),
        ("guide_fraction", ctypes.c_float),
        ("train_pass_spp", ctypes.c_int32),
        ("glossy_min_alpha", ctypes.c_float),
        ("guided_depth", ctypes.c_int32),
    ]


class SpkGuideStats(ctypes.Structure):
    _fields_ = [
        ("steps", ctypes.c_int64),
        ("last_records", ctypes.c_int64),
        ("last_loss", ctypes.c_double),
        ("last_weight", ctypes.c_double),
    ]


# name -> (restype, argtypes); every entry point declared in include/spk.h
SIGNATURES = {
    "spk_last_error": (ctypes.c_char_p, []),
    "spk_version": (ctypes.c_char_p, []),
    "spk_init": (ctypes.c_int, [ctypes.c_int]),
    "spk_scene_create": (ctypes.c_int, [ctypes.POINTER(SpkSceneDesc), ctypes.POINTER(ctypes.c_void_p)]),
    "spk_scene_build_opts_default": (ctypes.c_int, [ctypes.POINTER(SpkSceneBuildOpts)]),
    "spk_scene_create_ex": (ctypes.c_int, [ctypes.POINTER(SpkSceneDesc), ctypes.POINTER(SpkSceneBuildOpts),
                                           ctypes.POINTER(ctypes.c_void_p)]),
    "spk_scene_destroy": (ctypes.c_int, [ctypes.c_void_p]),
    "spk_scene_stats_get": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(SpkSceneStats)]),
    "spk_trace": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int64,
                                 ctypes.c_int32, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]),
    "spk_trace_noalloc": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                                         ctypes.c_int64, ctypes.c_int32, ctypes.c_void_p, ctypes.c_void_p,
                                         ctypes.c_void_p]),
    "spk_render_create": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(SpkCameraDesc), ctypes.POINTER(SpkRenderDesc),
                                         ctypes.POINTER(SpkRenderOpts), ctypes.POINTER(ctypes.c_void_p)]),
    "spk_render_destroy": (ctypes.c_int, [ctypes.c_void_p]),
    "spk_render_run": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int32, ctypes.c_int32]),
    "spk_render_sync": (ctypes.c_int, [ctypes.c_void_p]),
    "spk_render_film": (ctypes.c_int, [ctypes.c_void_p, c_d_p, c_d_p, c_d_p]),
    "spk_render_image": (ctypes.c_int, [ctypes.c_void_p, c_f_p]),
    "spk_render_relmse": (ctypes.c_int, [ctypes.c_void_p, c_d_p, ctypes.c_double, c_d_p]),
    "spk_render_relmse_pixels": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int64, c_i32_p, c_d_p, ctypes.c_double, c_d_p]),
    "spk_render_stats_get": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(SpkRenderStats)]),
    "spk_sampler_u32": (ctypes.c_int, [ctypes.c_int32, ctypes.c_uint32, ctypes.c_int64, c_u32_p, c_u32_p, c_u32_p,
                                       c_u32_p, ctypes.c_int32]),
    "spk_guide_opts_default": (ctypes.c_int, [ctypes.POINTER(SpkGuideOpts)]),
    "spk_guide_create": (ctypes.c_int, [ctypes.c_int, ctypes.POINTER(SpkGuideOpts), ctypes.POINTER(ctypes.c_void_p)]),
    "spk_guide_destroy": (ctypes.c_int, [ctypes.c_void_p]),
    "spk_guide_num_params": (ctypes.c_int64, []),
    "spk_guide_params_get": (ctypes.c_int, [ctypes.c_void_p, c_f_p]),
    "spk_guide_params_set": (ctypes.c_int, [ctypes.c_void_p, c_f_p]),
    "spk_guide_eval": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int64, c_f_p, c_f_p, c_f_p, c_f_p]),
    "spk_guide_sample": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int64, c_f_p, c_f_p, c_f_p, c_f_p]),
    "spk_guide_grad": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int64, c_f_p, c_f_p, c_f_p, c_f_p, c_d_p]),
    "spk_guide_train": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int64, c_f_p, c_f_p, c_f_p, ctypes.c_int32]),
    "spk_guide_stats_get": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(SpkGuideStats)]),
    "spk_guide_set_train_spp": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int32]),
    "spk_render_set_guide": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p]),
    "spk_batch_create": (ctypes.c_int, [ctypes.POINTER(SpkBatchDesc), ctypes.POINTER(SpkRenderOpts),
                                        ctypes.POINTER(ctypes.c_void_p)]),
    "spk_batch_destroy": (ctypes.c_int, [ctypes.c_void_p]),
    "spk_batch_set_guide": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p]),
    "spk_batch_run": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int32]),
    "spk_batch_film": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int32, c_d_p, c_d_p, c_d_p]),
    "spk_batch_info_get": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(SpkBatchInfo)]),
    "spk_batch_stats_get": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(SpkRenderStats)]),
}

_lib = None


class SpkError(RuntimeError):
    pass


def load():
    """Loads libspk.so (built in-tree by __graft_entry__.build()); raises if absent."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise SpkError(f"{LIB_PATH} is missing: build the HIP extension first (python -c 'import __graft_entry__ as g; g.build()')")
    # torch ships its own HIP runtime.  Imported first, its libamdhip64 satisfies the
    # library's dependency and the process holds one runtime; loading the library first
    # brings in the system runtime as a second one, and whichever of the two initialises
    # later finds no device.
    import torch  # noqa: F401
    lib = ctypes.CDLL(LIB_PATH)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def check(rc, what):
    if rc != SPK_OK:
        msg = load().spk_last_error().decode(errors="replace")
        raise SpkError(f"{what} failed (status {rc}): {msg}")
