"""Procedural workloads.

C0  Cornell box: 5 wall quads, two boxes (0.3^3 rotated 15 deg, 0.3x0.6x0.3 rotated
    -18 deg), a 0.25 x 0.25 ceiling light = 36 triangles; 256x256, fov 40, 64 spp,
    maxDepth 8, Russian roulette after bounce 3, independent sampler seed 0.
C1  Sphere field: seed-1 generator, 512 spheres of 640 triangles (UV sphere 20 x 17
    with pole fans; "subdiv 4 / 640 tris" cannot both hold for an icosphere) with
    centres U[-5,5]^3 and radius log-U[0.05,0.4] on a 20x20 ground quad; 40% Lambert
    U[0.1,0.9], 20% mirror, 20% dielectric IOR 1.5, 20% GGX alpha U[0.05,0.5]; two
    quad area lights; 1920x1080, 256 spp, maxDepth 16, Sobol.
C2  Batch: 64 images; image i draws its scene from the C1 generator with seed i and
    U{16..2048} spheres (10k-1.3M triangles), a W x H film with W, H ~ U{128..1024} and
    U{16..512} samples per pixel; maxDepth 16, Sobol seed i.
C4  Heavy combined: seed 11; 6000 spheres of the C1 generator + two noise-displaced
    height-field grids of 1,002,528 triangles each (ground and back wall), 4 quad lights
    and a 1024x512 synthetic HDR sky (gradient 0.5-2.0, sun disc of radiance 1e4),
    8 views at 1920x1080, 2048 spp, maxDepth 32.
C3  Occluded glossy: the same generator with seed 7, 60% GGX alpha U[0.1,0.3], one
    quad light inside a box open on its top that lights the scene via the ceiling;
    1280x720, 1024 spp, maxDepth 16.
"""
import math

import numpy as np

from . import _abi
from .scene import (GOLD, SILVER, Camera, Scene, box, conductor, dielectric, diffuse, look_at, plastic, quad,
                    roughconductor, roughdielectric, roughplastic, unit_sphere)


class Workload:
    def __init__(self, name, scene, camera, spp, max_depth, sampler, seed, rr_depth=3, integrator=_abi.SPK_INTEGRATOR_PATH):
        self.name = name
        self.integrator = integrator
        self.scene = scene
        self.camera = camera
        self.spp = spp
        self.max_depth = max_depth
        self.sampler = sampler
        self.seed = seed
        self.rr_depth = rr_depth

    def render_desc(self, spp=None, first_sample=0, sampler=None, seed=None, max_depth=None, integrator=None):
        d = _abi.SpkRenderDesc()
        d.integrator = self.integrator if integrator is None else integrator
        d.spp = self.spp if spp is None else spp
        d.max_depth = self.max_depth if max_depth is None else max_depth
        d.sampler = self.sampler if sampler is None else sampler
        d.seed = self.seed if seed is None else seed
        d.rr_depth = self.rr_depth
        d.first_sample = first_sample
        return d


WHITE = (0.75, 0.75, 0.75)
RED = (0.63, 0.065, 0.05)
GREEN = (0.14, 0.45, 0.09)


def cornell_box(width=256, height=256, spp=64, max_depth=8, light=(17.0, 12.0, 4.0)):
    s = Scene()
    white = s.add_material(diffuse(WHITE))
    red = s.add_material(diffuse(RED))
    green = s.add_material(diffuse(GREEN))
    black = s.add_material(diffuse((0.0, 0.0, 0.0)))
    # room [0,1]^3, open at z = 0
    s.add_triangles(quad((0, 0, 0), (0, 0, 1), (1, 0, 1), (1, 0, 0)), white)   # floor
    s.add_triangles(quad((0, 1, 0), (1, 1, 0), (1, 1, 1), (0, 1, 1)), white)   # ceiling
    s.add_triangles(quad((0, 0, 1), (0, 1, 1), (1, 1, 1), (1, 0, 1)), white)   # back
    s.add_triangles(quad((1, 0, 0), (1, 0, 1), (1, 1, 1), (1, 1, 0)), red)     # +x wall
    s.add_triangles(quad((0, 0, 0), (0, 1, 0), (0, 1, 1), (0, 0, 1)), green)   # -x wall
    s.add_triangles(box((0.62, 0.15, 0.35), (0.3, 0.3, 0.3), 15.0), white)
    s.add_triangles(box((0.35, 0.3, 0.65), (0.3, 0.6, 0.3), -18.0), white)
    # light quad just under the ceiling, emitting downwards (face normal -y)
    lo, hi, y = 0.375, 0.625, 0.999
    s.add_triangles(quad((lo, y, lo), (hi, y, lo), (hi, y, hi), (lo, y, hi)), black, radiance=light)
    cam = Camera(width, height, look_at((0.5, 0.5, -1.37), (0.5, 0.5, 0.0), (0.0, 1.0, 0.0)), 40.0)
    return Workload("C0 cornell box", s, cam, spp, max_depth, _abi.SPK_SAMPLER_INDEPENDENT, 0)


def _sphere_materials(rng, mix):
    """mix: list of (probability, kind) with kind in lambert/mirror/dielectric/ggx(lo,hi), and
    beckmann(lo,hi) (rough gold), plastic, roughdielectric(lo,hi,dist), roughplastic(lo,hi,dist)."""
    u = rng.random()
    acc = 0.0
    for p, kind in mix:
        acc += p
        if u < acc:
            break
    if kind[0] == "lambert":
        return diffuse(rng.uniform(0.1, 0.9, size=3))
    if kind[0] == "mirror":
        return conductor(**SILVER)
    if kind[0] == "dielectric":
        return dielectric(1.5)
    if kind[0] == "beckmann":
        return roughconductor(alpha=rng.uniform(kind[1], kind[2]), distribution="beckmann", **GOLD)
    if kind[0] == "plastic":
        return plastic(rng.uniform(0.1, 0.9, size=3), (1.0, 1.0, 1.0), 1.5)
    if kind[0] == "roughdielectric":
        return roughdielectric(alpha=rng.uniform(kind[1], kind[2]), int_ior=1.5, distribution=kind[3])
    if kind[0] == "roughplastic":
        return roughplastic(alpha=rng.uniform(kind[1], kind[2]), diffuse_reflectance=rng.uniform(0.1, 0.9, size=3),
                            distribution=kind[3])
    return roughconductor(alpha=rng.uniform(kind[1], kind[2]), **GOLD)


C1_MIX = [(0.4, ("lambert",)), (0.2, ("mirror",)), (0.2, ("dielectric",)), (0.2, ("ggx", 0.05, 0.5))]
# every BSDF kind, both microfacet distributions
WIDE_MIX = [(0.15, ("lambert",)), (0.1, ("mirror",)), (0.1, ("dielectric",)), (0.1, ("ggx", 0.05, 0.5)),
            (0.1, ("beckmann", 0.05, 0.5)), (0.1, ("plastic",)), (0.1, ("roughdielectric", 0.05, 0.4, "beckmann")),
            (0.05, ("roughdielectric", 0.05, 0.4, "ggx")), (0.1, ("roughplastic", 0.05, 0.4, "beckmann")),
            (0.1, ("roughplastic", 0.05, 0.4, "ggx"))]
C3_MIX = [(0.6, ("ggx", 0.1, 0.3)), (0.25, ("lambert",)), (0.1, ("mirror",)), (0.05, ("dielectric",))]


def sphere_field(seed=1, n_spheres=512, mix=None, lights="quads", ground=True):
    """The sphere-field generator (C1; C2, C3 and C4 reuse it)."""
    rng = np.random.default_rng(seed)
    mix = C1_MIX if mix is None else mix
    s = Scene()
    unit = unit_sphere()
    centers = rng.uniform(-5.0, 5.0, size=(n_spheres, 3))
    radii = np.exp(rng.uniform(math.log(0.05), math.log(0.4), size=n_spheres))
    for i in range(n_spheres):
        m = _sphere_materials(rng, mix)
        s.add_triangles(unit * radii[i] + centers[i], m)
    g = -5.5
    if ground:
        s.add_triangles(quad((-10, g, -10), (-10, g, 10), (10, g, 10), (10, g, -10)), diffuse((0.5, 0.5, 0.5)))
    black = s.add_material(diffuse((0.0, 0.0, 0.0)))
    if lights == "quads":
        for cx in (-4.0, 4.0):
            y, h = 9.0, 1.5
            s.add_triangles(quad((cx - h, y, -h), (cx + h, y, -h), (cx + h, y, h), (cx - h, y, h)), black,
                            radiance=(12.0, 12.0, 12.0))
    return s, black


def c1_spheres(width=1920, height=1080, spp=256, max_depth=16, seed=1, n_spheres=512):
    s, _ = sphere_field(seed, n_spheres)
    cam = Camera(width, height, look_at((0.0, 3.0, -18.0), (0.0, -1.0, 0.0), (0.0, 1.0, 0.0)), 40.0)
    return Workload("C1 sphere field", s, cam, spp, max_depth, _abi.SPK_SAMPLER_SOBOL, seed)


class BatchWorkload:
    def __init__(self, name, items, max_depth, sampler):
        self.name = name
        self.items = items          # list of (Scene, Camera, spp, seed)
        self.max_depth = max_depth
        self.sampler = sampler

    def paths(self):
        return sum(c.width * c.height * spp for _, c, spp, _ in self.items)


def c2_batch(n_images=64, draw_seed=2, max_depth=16, spheres=(16, 2048), res=(128, 1024), spp=(16, 512)):
    """The batch workload: per image an independent sphere field, film size and sample count."""
    rng = np.random.default_rng(draw_seed)
    items = []
    for i in range(n_images):
        n = int(rng.integers(spheres[0], spheres[1] + 1))
        w = int(rng.integers(res[0], res[1] + 1))
        h = int(rng.integers(res[0], res[1] + 1))
        k = int(rng.integers(spp[0], spp[1] + 1))
        sc, _ = sphere_field(seed=i, n_spheres=n)
        cam = Camera(w, h, look_at((0.0, 3.0, -18.0), (0.0, -1.0, 0.0), (0.0, 1.0, 0.0)), 40.0)
        items.append((sc, cam, k, i))
    return BatchWorkload(f"C2 batch of {n_images}", items, max_depth, _abi.SPK_SAMPLER_SOBOL)


def c3_occluded_glossy(width=1280, height=720, spp=1024, max_depth=16, seed=7, n_spheres=512):
    s, black = sphere_field(seed, n_spheres, mix=C3_MIX, lights=None)
    white = s.add_material(diffuse((0.8, 0.8, 0.8)))
    # ceiling that receives the light
    y = 10.0
    s.add_triangles(quad((-10, y, -10), (10, y, -10), (10, y, 10), (-10, y, 10)), white)
    # a 2 x 1 x 2 box open on its top (+y) around the light
    cx, cy, cz, h, hy = 0.0, 7.0, 0.0, 1.0, 0.5
    lo_y, hi_y = cy - hy, cy + hy
    s.add_triangles(quad((cx - h, lo_y, cz - h), (cx + h, lo_y, cz - h), (cx + h, lo_y, cz + h), (cx - h, lo_y, cz + h)), white)
    s.add_triangles(quad((cx - h, lo_y, cz - h), (cx - h, hi_y, cz - h), (cx + h, hi_y, cz - h), (cx + h, lo_y, cz - h)), white)
    s.add_triangles(quad((cx - h, lo_y, cz + h), (cx + h, lo_y, cz + h), (cx + h, hi_y, cz + h), (cx - h, hi_y, cz + h)), white)
    s.add_triangles(quad((cx - h, lo_y, cz - h), (cx - h, lo_y, cz + h), (cx - h, hi_y, cz + h), (cx - h, hi_y, cz - h)), white)
    s.add_triangles(quad((cx + h, lo_y, cz - h), (cx + h, hi_y, cz - h), (cx + h, hi_y, cz + h), (cx + h, lo_y, cz + h)), white)
    # light on the box floor, facing up (+y)
    l, ly = 0.8, lo_y + 0.01
    s.add_triangles(quad((cx - l, ly, cz - l), (cx - l, ly, cz + l), (cx + l, ly, cz + l), (cx + l, ly, cz - l)), black,
                    radiance=(60.0, 60.0, 60.0))
    cam = Camera(width, height, look_at((0.0, 2.0, -18.0), (0.0, -1.0, 0.0), (0.0, 1.0, 0.0)), 40.0)
    return Workload("C3 occluded glossy", s, cam, spp, max_depth, _abi.SPK_SAMPLER_SOBOL, seed)


def _noise(rng, u, v, octaves=5):
    """Smooth deterministic noise on [0,1]^2: a sum of randomly oriented sinusoids."""
    h = np.zeros_like(u)
    amp, freq = 1.0, 2.0
    for _ in range(octaves):
        for _ in range(3):
            th = rng.uniform(0, 2 * math.pi)
            ph = rng.uniform(0, 2 * math.pi)
            h += amp * np.sin(freq * 2 * math.pi * (u * math.cos(th) + v * math.sin(th)) + ph)
        amp *= 0.5
        freq *= 2.1
    return h / 3.0


def noise_grid(rng, n, origin, eu, ev, normal_dir, amplitude):
    """n x n quads (2 n^2 triangles) over origin + s eu + t ev, displaced along normal_dir."""
    t = np.linspace(0.0, 1.0, n + 1)
    U, V = np.meshgrid(t, t, indexing="ij")
    H = amplitude * _noise(rng, U, V)
    P = (np.asarray(origin)[None, None, :] + U[..., None] * np.asarray(eu)[None, None, :]
         + V[..., None] * np.asarray(ev)[None, None, :] + H[..., None] * np.asarray(normal_dir)[None, None, :])
    a, b, c, d = P[:-1, :-1], P[1:, :-1], P[1:, 1:], P[:-1, 1:]
    t1 = np.stack([a, b, c], axis=2).reshape(-1, 3, 3)
    t2 = np.stack([a, c, d], axis=2).reshape(-1, 3, 3)
    return np.concatenate([t1, t2]).astype(np.float32)


def sky_envmap(w=1024, h=512, sun_theta_deg=40.0, sun_phi_deg=60.0, sun_radius_deg=0.75, sun=1.0e4):
    """Lat-long z-up sky (u = phi / 2pi, v = theta / pi): luminance 2.0 at the horizon to 0.5 at
    the zenith (blue-tinted), 0.5 below the horizon, and a sun disc of radiance `sun`."""
    v = (np.arange(h) + 0.5) / h
    u = (np.arange(w) + 0.5) / w
    theta = v * math.pi
    up = np.clip(np.cos(theta), 0.0, 1.0)
    lum = np.where(theta <= math.pi / 2, 2.0 - 1.5 * up, 0.5)
    img = np.stack([0.8 * lum, 0.9 * lum, 1.1 * lum], axis=-1)[:, None, :].repeat(w, axis=1)
    st, sp = math.radians(sun_theta_deg), math.radians(sun_phi_deg)
    sdir = np.array([math.sin(st) * math.cos(sp), math.sin(st) * math.sin(sp), math.cos(st)])
    TH, PH = np.meshgrid(theta, u * 2 * math.pi, indexing="ij")
    d = np.stack([np.sin(TH) * np.cos(PH), np.sin(TH) * np.sin(PH), np.cos(TH)], axis=-1)
    img[(d @ sdir) >= math.cos(math.radians(sun_radius_deg))] = sun
    return img.astype(np.float32)


C4_VIEWS = 8


def c4_scene(seed=11, n_spheres=6000, grid_n=708):
    s, black = sphere_field(seed, n_spheres, lights=None, ground=False)
    rng = np.random.default_rng(seed + 1000)
    s.add_triangles(noise_grid(rng, grid_n, (-12.0, -5.5, -12.0), (24.0, 0.0, 0.0), (0.0, 0.0, 24.0), (0.0, 1.0, 0.0), 0.4),
                    diffuse((0.5, 0.5, 0.5)))
    s.add_triangles(noise_grid(rng, grid_n, (-12.0, -5.5, 9.0), (0.0, 16.0, 0.0), (24.0, 0.0, 0.0), (0.0, 0.0, 1.0), 0.6),
                    diffuse((0.6, 0.55, 0.5)))
    for cx, cz in ((-4.0, -4.0), (4.0, -4.0), (-4.0, 4.0), (4.0, 4.0)):
        y, hh = 9.0, 0.75
        s.add_triangles(quad((cx - hh, y, cz - hh), (cx + hh, y, cz - hh), (cx + hh, y, cz + hh), (cx - hh, y, cz + hh)),
                        black, radiance=(12.0, 12.0, 12.0))
    # z-up sky onto the y-up world
    s.set_envmap(sky_envmap(), to_world=np.array([[1, 0, 0, 0], [0, 0, 1, 0], [0, 1, 0, 0], [0, 0, 0, 1]], float),
                 radius=60.0)
    return s


def c4_cameras(width=1920, height=1080, views=C4_VIEWS):
    cams = []
    for i in range(views):
        a = math.radians(-90.0 + (i - (views - 1) / 2.0) * 12.0)
        eye = (18.0 * math.cos(a), 3.0 + 0.5 * (i % 3), 18.0 * math.sin(a))
        cams.append(Camera(width, height, look_at(eye, (0.0, -1.0, 0.0), (0.0, 1.0, 0.0)), 40.0))
    return cams


def c4_heavy(width=1920, height=1080, spp=2048, max_depth=32, seed=11, n_spheres=6000, grid_n=708, views=C4_VIEWS):
    """The heavy combined workload: one scene, `views` cameras (Sobol seed = 11 + view)."""
    s = c4_scene(seed, n_spheres, grid_n)
    return [Workload(f"C4 heavy view {i}", s, cam, spp, max_depth, _abi.SPK_SAMPLER_SOBOL, seed + i)
            for i, cam in enumerate(c4_cameras(width, height, views))]
