"""Scene description: the flat form of what spica's scene parser builds.

spica turns an XML file into a vector of Triangle primitives, each with a material
and (for emitters) its own AreaLight (sources/spica/sceneparser.cc:168-188,
257-270), plus optional environment-map emitters (sources/lights/envmap.cc:49-55).
`Scene` holds the same thing as numpy arrays: world-space float32 vertices, one
material per triangle, one area light per emissive triangle, and lays it out as
the POD descriptor of include/spk_types.h.
"""
import ctypes
import math

import numpy as np

from . import _abi


# ---------------------------------------------------------------------------
# materials (parameter names follow sources/bsdfs/*.cc)
# ---------------------------------------------------------------------------
def diffuse(reflectance):
    """bsdfs/diffuse.cc: "reflectance"."""
    return dict(type=_abi.SPK_MAT_DIFFUSE, reflectance=tuple(map(float, reflectance)))


def conductor(eta, k):
    """bsdfs/conductor.cc: "eta", "k" (no defaults in spica)."""
    return dict(type=_abi.SPK_MAT_CONDUCTOR, eta=tuple(map(float, eta)), k=tuple(map(float, k)))


def dielectric(int_ior=1.333, reflectance=(1.0, 1.0, 1.0), transmittance=(1.0, 1.0, 1.0)):
    """bsdfs/dielectric.cc: "intIOR" (default 1.333), "specularReflectance", "specularTransmittance"."""
    return dict(type=_abi.SPK_MAT_DIELECTRIC, int_ior=float(int_ior), reflectance=tuple(map(float, reflectance)),
                transmittance=tuple(map(float, transmittance)))


_DISTRIBUTIONS = {"ggx": _abi.SPK_MICROFACET_GGX, "beckmann": _abi.SPK_MICROFACET_BECKMANN}


def _distribution(name):
    if name not in _DISTRIBUTIONS:
        raise ValueError(f"microfacet distribution {name!r} is not one of {sorted(_DISTRIBUTIONS)}")
    return _DISTRIBUTIONS[name]


def roughconductor(alpha=0.1, eta=(1.0, 1.0, 1.0), k=(0.0, 0.0, 0.0), distribution="ggx"):
    """bsdfs/roughconductor.cc: "alpha" (default 0.1), "eta", "k", "distribution" ("ggx" here by
    default; spica's own default, "beckmann", is what its scene files get when they name none)."""
    return dict(type=_abi.SPK_MAT_ROUGHCONDUCTOR, alpha=float(alpha), eta=tuple(map(float, eta)), k=tuple(map(float, k)),
                distribution=_distribution(distribution))


def plastic(diffuse_reflectance=(0.5, 0.5, 0.5), specular_reflectance=(1.0, 1.0, 1.0), int_ior=1.5):
    """bsdfs/plastic.cc: "diffuseReflectance", "specularReflectance", "intIOR" (default 1.5)."""
    return dict(type=_abi.SPK_MAT_PLASTIC, reflectance=tuple(map(float, diffuse_reflectance)),
                specular=tuple(map(float, specular_reflectance)), int_ior=float(int_ior))


def roughdielectric(alpha=0.1, int_ior=1.3333, reflectance=(1.0, 1.0, 1.0), transmittance=(1.0, 1.0, 1.0),
                    distribution="beckmann"):
    """bsdfs/roughdielectric.cc: "specularReflectance", "specularTransmittance", "alpha" (0.1),
    "intIOR" (1.3333), "distribution" ("beckmann")."""
    return dict(type=_abi.SPK_MAT_ROUGHDIELECTRIC, alpha=float(alpha), int_ior=float(int_ior),
                reflectance=tuple(map(float, reflectance)), transmittance=tuple(map(float, transmittance)),
                distribution=_distribution(distribution))


def roughplastic(alpha=0.1, diffuse_reflectance=(0.5, 0.5, 0.5), specular_reflectance=(1.0, 1.0, 1.0), int_ior=1.5,
                 distribution="beckmann"):
    """bsdfs/roughplastic.cc: "diffuseReflectance", "specularReflectance", "intIOR" (1.5),
    "alpha" (0.1, at least 0.001), "distribution" ("beckmann")."""
    return dict(type=_abi.SPK_MAT_ROUGHPLASTIC, alpha=float(alpha), int_ior=float(int_ior),
                reflectance=tuple(map(float, diffuse_reflectance)), specular=tuple(map(float, specular_reflectance)),
                distribution=_distribution(distribution))


# Conductor optical constants (RGB) used by the procedural scenes: silver for
# mirrors, gold for the GGX spheres.
SILVER = dict(eta=(0.155, 0.117, 0.138), k=(4.828, 3.122, 2.146))
GOLD = dict(eta=(0.143, 0.374, 1.442), k=(3.983, 2.385, 1.603))


def _material_struct(m):
    s = _abi.SpkMaterial()
    s.type = m["type"]
    s.distribution = m.get("distribution", _abi.SPK_MICROFACET_GGX)
    for key in ("reflectance", "transmittance", "eta", "k", "specular"):
        v = m.get(key, (0.0, 0.0, 0.0))
        getattr(s, key)[:] = v
    s.int_ior = m.get("int_ior", 1.333)
    s.alpha = m.get("alpha", 0.1)
    return s


# ---------------------------------------------------------------------------
# transforms (sources/core/transform.cc:167-198)
# ---------------------------------------------------------------------------
def look_at(eye, target, up):
    """spica's Transform::lookAt: columns (left = up x dir, newUp = dir x left, dir), origin eye."""
    eye = np.asarray(eye, dtype=np.float64)
    target = np.asarray(target, dtype=np.float64)
    up = np.asarray(up, dtype=np.float64)
    d = target - eye
    d = d * (1.0 / np.linalg.norm(d))
    upn = up * (1.0 / np.linalg.norm(up))
    left = np.cross(upn, d)
    if np.linalg.norm(left) == 0.0:
        raise ValueError("look_at: up and viewing direction are parallel")
    left = left * (1.0 / np.linalg.norm(left))
    new_up = np.cross(d, left)
    m = np.zeros((4, 4))
    m[:3, 0] = left
    m[:3, 1] = new_up
    m[:3, 2] = d
    m[:3, 3] = eye
    m[3, 3] = 1.0
    return m


class Camera:
    """Perspective sensor (sources/cameras/perspective.cc:23-30): toWorld, fov (deg), film size."""

    def __init__(self, width, height, to_world, fov_deg, aperture_radius=0.0, focus_distance=50.0):
        self.width = int(width)
        self.height = int(height)
        self.to_world = np.asarray(to_world, dtype=np.float64).reshape(4, 4)
        self.fov_deg = float(fov_deg)
        self.aperture_radius = float(aperture_radius)
        self.focus_distance = float(focus_distance)

    def desc(self):
        c = _abi.SpkCameraDesc()
        c.width = self.width
        c.height = self.height
        c.to_world[:] = [float(x) for x in self.to_world.reshape(-1)]
        c.fov_deg = self.fov_deg
        c.aperture_radius = self.aperture_radius
        c.focus_distance = self.focus_distance
        return c


# ---------------------------------------------------------------------------
# scene
# ---------------------------------------------------------------------------
class Scene:
    """Triangle soup + materials + area lights (+ optional environment map)."""

    def __init__(self):
        self._verts = []
        self._mats = []
        self._mat_ids = []
        self._normals = []
        self._emit = []  # per chunk: radiance or None
        self.materials = []
        self.env = None
        self._frozen = None

    # -- building --------------------------------------------------------
    def add_material(self, m):
        self.materials.append(m)
        return len(self.materials) - 1

    def add_triangles(self, verts, material, radiance=None, normals=None):
        """verts: [n, 3, 3]; material: index or dict; radiance: emitter radiance or None;
        normals: per-corner shading normals [n, 3, 3] (smooth mesh) or None (flat)."""
        v = np.asarray(verts, dtype=np.float32).reshape(-1, 3, 3)
        mid = material if isinstance(material, int) else self.add_material(material)
        if normals is not None:
            normals = np.asarray(normals, dtype=np.float32).reshape(v.shape)
        self._normals.append(normals)
        self._verts.append(v)
        self._mat_ids.append(np.full(len(v), mid, dtype=np.int32))
        self._emit.append(None if radiance is None else np.asarray(radiance, dtype=np.float64).reshape(3))
        self._frozen = None
        return self

    def set_envmap(self, rgb, to_world=None, scale=1.0, center=(0.0, 0.0, 0.0), radius=2.0):
        """lights/envmap.cc:49-55: image [h][w][3], "toWorld", "scale", "worldCenter", "worldRadius"."""
        rgb = np.ascontiguousarray(rgb, dtype=np.float32)
        assert rgb.ndim == 3 and rgb.shape[2] == 3
        self.env = dict(rgb=rgb, to_world=np.eye(4) if to_world is None else np.asarray(to_world, dtype=np.float64),
                        scale=float(scale), center=tuple(map(float, center)), radius=float(radius))
        self._frozen = None
        return self

    # -- frozen arrays ---------------------------------------------------
    def freeze(self):
        if self._frozen is not None:
            return self._frozen
        verts = np.concatenate(self._verts) if self._verts else np.zeros((0, 3, 3), np.float32)
        mat_ids = np.concatenate(self._mat_ids) if self._mat_ids else np.zeros(0, np.int32)
        light_ids = np.full(len(verts), -1, dtype=np.int32)
        area_tri, area_rad = [], []
        off = 0
        for v, e in zip(self._verts, self._emit):
            if e is not None:
                for t in range(len(v)):
                    light_ids[off + t] = len(area_tri)
                    area_tri.append(off + t)
                    area_rad.append(e)
            off += len(v)
        fz = dict(
            verts=np.ascontiguousarray(verts, dtype=np.float32),
            mat_ids=np.ascontiguousarray(mat_ids, dtype=np.int32),
            light_ids=light_ids,
            area_tri=np.asarray(area_tri, dtype=np.int32),
            area_radiance=np.ascontiguousarray(np.asarray(area_rad, dtype=np.float64).reshape(-1, 3)),
        )
        if any(nm is not None for nm in self._normals):
            fz["normals"] = np.ascontiguousarray(np.concatenate(
                [np.zeros(v.shape, np.float32) if nm is None else nm for v, nm in zip(self._verts, self._normals)]))
        self._frozen = fz
        return fz

    @property
    def num_tris(self):
        return len(self.freeze()["verts"])

    @property
    def num_lights(self):
        return len(self.freeze()["area_tri"]) + (1 if self.env is not None else 0)

    def bounds(self):
        v = self.freeze()["verts"].reshape(-1, 3)
        return v.min(axis=0), v.max(axis=0)

    def save_bin(self, path, camera, spp, max_depth, sampler, seed):
        """Flat scene + render request for the C++ plugin driver (spica_amd/host/plugin_demo)."""
        fz = self.freeze()
        with open(path, "wb") as f:
            f.write(b"SPKS")
            f.write(np.int32(1).tobytes())
            f.write(np.int64(len(fz["verts"])).tobytes())
            f.write(fz["verts"].tobytes())
            f.write(fz["mat_ids"].tobytes())
            f.write(fz["light_ids"].tobytes())
            f.write(np.int32(len(self.materials)).tobytes())
            for m in self.materials:
                f.write(bytes(_material_struct(m)))
            f.write(np.int32(len(fz["area_tri"])).tobytes())
            f.write(fz["area_tri"].tobytes())
            f.write(fz["area_radiance"].tobytes())
            f.write(bytes(camera.desc()))
            f.write(np.array([spp, max_depth, sampler], dtype=np.int32).tobytes())
            f.write(np.uint32(seed).tobytes())

    def desc(self):
        """SpkSceneDesc pointing into this scene's arrays (keep the Scene alive while used)."""
        fz = self.freeze()
        d = _abi.SpkSceneDesc()
        n = len(fz["verts"])
        d.num_tris = n
        d.verts = fz["verts"].ctypes.data_as(_abi.c_f_p)
        d.normals = fz["normals"].ctypes.data_as(_abi.c_f_p) if "normals" in fz else ctypes.cast(None, _abi.c_f_p)
        d.mat_ids = fz["mat_ids"].ctypes.data_as(_abi.c_i32_p)
        d.light_ids = fz["light_ids"].ctypes.data_as(_abi.c_i32_p)
        mats = (_abi.SpkMaterial * max(1, len(self.materials)))()
        for i, m in enumerate(self.materials):
            mats[i] = _material_struct(m)
        self._mat_array = mats
        d.num_materials = len(self.materials)
        d.materials = ctypes.cast(mats, ctypes.POINTER(_abi.SpkMaterial))
        d.num_area_lights = len(fz["area_tri"])
        d.area_tri = fz["area_tri"].ctypes.data_as(_abi.c_i32_p)
        d.area_radiance = fz["area_radiance"].ctypes.data_as(_abi.c_d_p)
        if self.env is not None:
            e = self.env
            d.env_height, d.env_width = e["rgb"].shape[:2]
            d.env_rgb = e["rgb"].ctypes.data_as(_abi.c_f_p)
            d.env_to_world[:] = [float(x) for x in e["to_world"].reshape(-1)]
            d.env_scale = e["scale"]
            d.env_center[:] = e["center"]
            d.env_radius = e["radius"]
        else:
            d.env_width = 0
            d.env_height = 0
            d.env_rgb = ctypes.cast(None, _abi.c_f_p)
            d.env_to_world[:] = [float(x) for x in np.eye(4).reshape(-1)]
            d.env_scale = 1.0
            d.env_radius = 2.0
        self._desc = d
        return d


# ---------------------------------------------------------------------------
# geometry helpers
# ---------------------------------------------------------------------------
def quad(p0, p1, p2, p3):
    """Two triangles (p0,p1,p2), (p0,p2,p3); face normal = normalize((p1-p0) x (p2-p0))."""
    p0, p1, p2, p3 = (np.asarray(p, dtype=np.float64) for p in (p0, p1, p2, p3))
    return np.stack([np.stack([p0, p1, p2]), np.stack([p0, p2, p3])])


def rot_y(deg):
    a = math.radians(deg)
    c, s = math.cos(a), math.sin(a)
    return np.array([[c, 0.0, s], [0.0, 1.0, 0.0], [-s, 0.0, c]])


def box(center, size, rot_y_deg=0.0):
    """Axis-aligned box rotated about +y, 12 outward-facing triangles."""
    hx, hy, hz = (0.5 * s for s in size)
    c = np.asarray(center, dtype=np.float64)
    R = rot_y(rot_y_deg)
    corners = np.array([[sx * hx, sy * hy, sz * hz] for sx in (-1, 1) for sy in (-1, 1) for sz in (-1, 1)])
    P = corners @ R.T + c

    def v(ix, iy, iz):
        return P[(ix << 2) | (iy << 1) | iz]

    faces = [
        quad(v(0, 0, 0), v(0, 0, 1), v(0, 1, 1), v(0, 1, 0)),  # -x
        quad(v(1, 0, 0), v(1, 1, 0), v(1, 1, 1), v(1, 0, 1)),  # +x
        quad(v(0, 0, 0), v(1, 0, 0), v(1, 0, 1), v(0, 0, 1)),  # -y
        quad(v(0, 1, 0), v(0, 1, 1), v(1, 1, 1), v(1, 1, 0)),  # +y
        quad(v(0, 0, 0), v(0, 1, 0), v(1, 1, 0), v(1, 0, 0)),  # -z
        quad(v(0, 0, 1), v(1, 0, 1), v(1, 1, 1), v(0, 1, 1)),  # +z
    ]
    return np.concatenate(faces)


def uv_sphere(center, radius, slices=20, stacks=17):
    """Closed UV sphere with pole fans: 2 * slices * (stacks - 1) triangles (640 for 20 x 17)."""
    c = np.asarray(center, dtype=np.float64)
    th = np.pi * np.arange(stacks + 1) / stacks
    ph = 2.0 * np.pi * np.arange(slices) / slices

    def p(j, i):
        i %= slices
        return c + radius * np.array([np.sin(th[j]) * np.cos(ph[i]), np.cos(th[j]), np.sin(th[j]) * np.sin(ph[i])])

    tris = []
    for i in range(slices):
        tris.append([p(0, 0), p(1, i + 1), p(1, i)])
    for j in range(1, stacks - 1):
        for i in range(slices):
            a, b, cc, d = p(j, i), p(j, i + 1), p(j + 1, i + 1), p(j + 1, i)
            tris.append([a, b, cc])
            tris.append([a, cc, d])
    for i in range(slices):
        tris.append([p(stacks, 0), p(stacks - 1, i), p(stacks - 1, i + 1)])
    return np.asarray(tris)


_SPHERE_CACHE = {}


def unit_sphere(slices=20, stacks=17):
    key = (slices, stacks)
    if key not in _SPHERE_CACHE:
        _SPHERE_CACHE[key] = uv_sphere((0.0, 0.0, 0.0), 1.0, slices, stacks)
    return _SPHERE_CACHE[key]
