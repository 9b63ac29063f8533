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
C3  Occluded glossy: the same generator with seed 7, 60% GGX alpha U[0.1,0.3], one
    quad light inside a box open on its top that lights the scene via the ceiling;
    1280x720, 1024 spp, maxDepth 16.
"""
import math

import numpy as np

from . import _abi
from .scene import GOLD, SILVER, Camera, Scene, box, conductor, dielectric, diffuse, look_at, quad, roughconductor, unit_sphere


class Workload:
    def __init__(self, name, scene, camera, spp, max_depth, sampler, seed, rr_depth=3):
        self.name = name
        self.scene = scene
        self.camera = camera
        self.spp = spp
        self.max_depth = max_depth
        self.sampler = sampler
        self.seed = seed
        self.rr_depth = rr_depth

    def render_desc(self, spp=None, first_sample=0, sampler=None, seed=None, max_depth=None):
        d = _abi.SpkRenderDesc()
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
    """mix: list of (probability, kind) with kind in lambert/mirror/dielectric/ggx(lo,hi)."""
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
    return roughconductor(alpha=rng.uniform(kind[1], kind[2]), **GOLD)


C1_MIX = [(0.4, ("lambert",)), (0.2, ("mirror",)), (0.2, ("dielectric",)), (0.2, ("ggx", 0.05, 0.5))]
C3_MIX = [(0.6, ("ggx", 0.1, 0.3)), (0.25, ("lambert",)), (0.1, ("mirror",)), (0.05, ("dielectric",))]


def sphere_field(seed=1, n_spheres=512, mix=None, lights="quads"):
    """The sphere-field generator (C1; C2 and C3 reuse it)."""
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
