"""Shared test helpers: seeded ray sets, statistics, small scenes."""
import numpy as np

from spica_amd import configs
from spica_amd.scene import Camera, Scene, diffuse, look_at, quad


def unit_vectors(rng, n):
    z = rng.uniform(-1.0, 1.0, n)
    phi = rng.uniform(0.0, 2.0 * np.pi, n)
    r = np.sqrt(np.maximum(0.0, 1.0 - z * z))
    return np.stack([r * np.cos(phi), r * np.sin(phi), z], axis=1)


def ray_set(scene, camera, oracle_scene, n, seed=1):
    """The hit-parity ray mix: 1/3 uniform origins in 1.5x the scene box with
    uniform-sphere directions, 1/3 camera rays, 1/3 first-bounce rays
    (cosine-hemisphere about the face normal of the camera-ray hit)."""
    rng = np.random.default_rng(seed)
    n1 = n // 3
    n2 = n // 3
    n3 = n - n1 - n2
    lo, hi = scene.bounds()
    c = 0.5 * (lo + hi)
    h = 0.75 * (hi - lo)
    o1 = rng.uniform(c - h, c + h, size=(n1, 3))
    d1 = unit_vectors(rng, n1)
    # camera rays (pinhole; spica's raster -> camera mapping)
    W, H = camera.width, camera.height
    px = rng.uniform(0, W, n2 + n3)
    py = rng.uniform(0, H, n2 + n3)
    t = np.tan(np.radians(camera.fov_deg) / 2.0)
    sx = (2.0 * px / W - 1.0) * (W / H) * t
    sy = (1.0 - 2.0 * py / H) * t
    dc = np.stack([sx, sy, np.ones_like(sx)], axis=1)
    dc /= np.linalg.norm(dc, axis=1, keepdims=True)
    M = camera.to_world
    dw = dc @ M[:3, :3].T
    dw /= np.linalg.norm(dw, axis=1, keepdims=True)
    ow = np.repeat(M[:3, 3][None, :], n2 + n3, axis=0)
    o2, d2 = ow[:n2], dw[:n2]
    # first bounce: hit points of more camera rays
    prim, th, _ = oracle_scene.trace(ow[n2:], dw[n2:])
    fz = scene.freeze()
    v = fz["verts"].astype(np.float64)
    ok = prim >= 0
    p = ow[n2:] + dw[n2:] * th[:, None]
    e1 = v[prim, 1] - v[prim, 0]
    e2 = v[prim, 2] - v[prim, 0]
    nrm = np.cross(e1, e2)
    nrm /= np.maximum(np.linalg.norm(nrm, axis=1, keepdims=True), 1e-30)
    nrm = np.where((np.sum(nrm * dw[n2:], axis=1) > 0)[:, None], -nrm, nrm)
    u1, u2 = rng.random(n3), rng.random(n3)
    r = np.sqrt(u1)
    loc = np.stack([r * np.cos(2 * np.pi * u2), r * np.sin(2 * np.pi * u2), np.sqrt(np.maximum(0, 1 - u1))], axis=1)
    a = np.where(np.abs(nrm[:, :1]) > 0.9, np.array([[0.0, 1.0, 0.0]]), np.array([[1.0, 0.0, 0.0]]))
    tx = np.cross(a, nrm)
    tx /= np.linalg.norm(tx, axis=1, keepdims=True)
    ty = np.cross(nrm, tx)
    d3 = loc[:, :1] * tx + loc[:, 1:2] * ty + loc[:, 2:] * nrm
    o3 = p + 1e-3 * nrm
    # camera rays that escaped stay as they are
    o3 = np.where(ok[:, None], o3, ow[n2:])
    d3 = np.where(ok[:, None], d3, dw[n2:])
    org = np.concatenate([o1, o2, o3]).astype(np.float32)
    dirv = np.concatenate([d1, d2, d3]).astype(np.float32)
    return org, dirv


def tri_intersect64(v, org, dirv):
    """Moller-Trumbore in float64 over arrays (v: [n,3,3]); returns t, u, v, det (no range checks)."""
    e1 = v[:, 1] - v[:, 0]
    e2 = v[:, 2] - v[:, 0]
    d = dirv / np.linalg.norm(dirv, axis=1, keepdims=True)
    p = np.cross(d, e2)
    det = np.sum(e1 * p, axis=1)
    inv = 1.0 / np.where(det == 0, 1e-300, det)
    tv = org - v[:, 0]
    u = np.sum(tv * p, axis=1) * inv
    q = np.cross(tv, e1)
    w = np.sum(d * q, axis=1) * inv
    t = np.sum(e2 * q, axis=1) * inv
    return t, u, w, det


def classify_mismatches(verts, org, dirv, p_gpu, t_gpu, p_cpu, t_cpu, rel_t=1e-6, edge=1e-6):
    """Mismatched closest-hit ids, classified:
    near_tie  both triangles hit (float64) at t within rel_t of each other;
    edge      the disagreement is explained by a hit within the float32 reach of a
              triangle edge, either side: barycentric distance below
              max(edge, 32 * 2^-24 * (|o| + t) / (sqrt(2 * area) * |cos|)), i.e. a few
              float32 rounding steps of the hit position measured in triangle sizes,
              divided by the cosine between ray and normal (the determinant of the
              ray/triangle system shrinks with it, so grazing rays lose precision);
    other     anything else (a real error)."""
    idx = np.nonzero(p_gpu != p_cpu)[0]
    out = {"near_tie": 0, "edge": 0, "other": 0, "other_idx": []}
    if len(idx) == 0:
        return out
    o = org[idx].astype(np.float64)
    d = dirv[idx].astype(np.float64)
    v = verts.astype(np.float64)
    def reach(k, t):
        tri = v[k]
        nrm = np.cross(tri[1] - tri[0], tri[2] - tri[0])
        area2 = np.linalg.norm(nrm)
        cos = abs(np.dot(nrm, d[j])) / max(area2 * np.linalg.norm(d[j]), 1e-300)
        return max(edge, 32.0 * 2.0 ** -24 * (np.linalg.norm(o[j]) + abs(t)) /
                   (np.sqrt(max(area2, 1e-30)) * max(cos, 1e-3)))

    for j, i in enumerate(idx):
        pg, pc = int(p_gpu[i]), int(p_cpu[i])
        tg = tc = None
        bary_g = bary_c = None
        eg = ec = edge
        if pg >= 0:
            t, u, w, _ = tri_intersect64(v[pg][None], o[j][None], d[j][None])
            tg, bary_g = t[0], min(u[0], w[0], 1.0 - u[0] - w[0])
            eg = reach(pg, tg)
        if pc >= 0:
            t, u, w, _ = tri_intersect64(v[pc][None], o[j][None], d[j][None])
            tc, bary_c = t[0], min(u[0], w[0], 1.0 - u[0] - w[0])
            ec = reach(pc, tc)
        if pg >= 0 and pc >= 0 and bary_g > -eg and bary_c > -ec and abs(tg - tc) <= rel_t * max(abs(tc), 1e-30):
            out["near_tie"] += 1
        elif (bary_g is not None and abs(bary_g) < eg) or (bary_c is not None and abs(bary_c) < ec):
            out["edge"] += 1
        else:
            out["other"] += 1
            out["other_idx"].append(int(i))
            out.setdefault("other_detail", []).append(
                dict(ray=int(i), gpu=(pg, float(t_gpu[i]), None if bary_g is None else float(bary_g)),
                     cpu=(pc, float(t_cpu[i]), None if bary_c is None else float(bary_c)),
                     t_gpu_tri64=None if tg is None else float(tg), t_cpu_tri64=None if tc is None else float(tc)))
    return out


# Precision of a float32 film mean against a float64 one.  Relative: per-sample radiance in
# float32 carries a few ulp (1e-7) per bounce, the Kahan film sums keep the mean to ~1e-7.
# Absolute: products of the estimator's 1e-12 regularisers (Fresnel, tan theta) leave
# radiance of 1e-35 .. 1e-45 (float32 denormals) where float64 gives 1e-99; such a pixel is
# zero in both.  A pixel whose samples are all equal (an emitter seen directly, black
# background, underflow residue) has zero sample variance on both sides and its means differ
# only by this precision; the floor keeps z finite there and is far below any statistical
# standard error of the scenes' pixels (their means are 1e-3 and up).
FILM_REL_PRECISION = 1e-5
FILM_ABS_PRECISION = 1e-12


def zstats(sum_a, sq_a, n_a, sum_b, sq_b, n_b):
    """Per pixel-channel z of the difference of two sample means (Welch), with the float
    precision floor added to the standard error."""
    ma = sum_a / n_a
    mb = sum_b / n_b
    va = np.maximum(sq_a / n_a - ma * ma, 0.0) * n_a / (n_a - 1)
    vb = np.maximum(sq_b / n_b - mb * mb, 0.0) * n_b / (n_b - 1)
    floor = FILM_REL_PRECISION * np.maximum(np.abs(ma), np.abs(mb)) + FILM_ABS_PRECISION
    se = np.sqrt(va / n_a + vb / n_b + floor * floor)
    z = np.where(se > 0, (ma - mb) / np.where(se > 0, se, 1.0), np.where(np.abs(ma - mb) > 0, np.inf, 0.0))
    return z, ma, mb


def worst_pixels(z, sum_a, sq_a, n_a, sum_b, sq_b, n_b, k=5):
    """The k largest |z| pixel-channels with both sides' mean, sample deviation and count."""
    flat = np.argsort(-np.abs(z).reshape(-1))[:k]
    out = []
    for f in flat:
        p, c = divmod(int(f), z.shape[1])
        na, nb = float(np.broadcast_to(n_a, z.shape)[p, c]), float(np.broadcast_to(n_b, z.shape)[p, c])
        ma, mb = sum_a[p, c] / na, sum_b[p, c] / nb
        sa = np.sqrt(max(sq_a[p, c] / na - ma * ma, 0.0))
        sb = np.sqrt(max(sq_b[p, c] / nb - mb * mb, 0.0))
        out.append(f"px {p} ch {c}: z {z[p, c]:.2f} a {ma:.6g} (sd {sa:.4g}, n {na:.0f}) b {mb:.6g} (sd {sb:.4g}, n {nb:.0f})")
    return out


def furnace_room(albedo=0.5, Le=1.0, size=1.0):
    """Closed cube whose 12 triangles all emit Le inwards and reflect Lambert `albedo`:
    the expected radiance with maxDepth D is Le * sum_{k=0..D} albedo^k everywhere."""
    s = Scene()
    m = s.add_material(diffuse((albedo, albedo, albedo)))
    a, b = 0.0, size
    inward = [
        quad((a, a, a), (b, a, a), (b, a, b), (a, a, b)),  # floor, normal +y
        quad((a, b, a), (a, b, b), (b, b, b), (b, b, a)),  # ceiling, -y
        quad((a, a, a), (a, a, b), (a, b, b), (a, b, a)),  # x = 0, +x
        quad((b, a, a), (b, b, a), (b, b, b), (b, a, b)),  # x = 1, -x
        quad((a, a, a), (a, b, a), (b, b, a), (b, a, a)),  # z = 0, +z
        quad((a, a, b), (b, a, b), (b, b, b), (a, b, b)),  # z = 1, -z
    ]
    centre = np.full(3, 0.5 * size)
    for q in inward:
        for tri in q:
            nrm = np.cross(tri[1] - tri[0], tri[2] - tri[0])
            if np.dot(nrm, centre - tri.mean(axis=0)) < 0:  # emission is one-sided: face the inside
                tri = tri[[0, 2, 1]]
            s.add_triangles(tri[None], m, radiance=(Le, Le, Le))
    cam = Camera(32, 32, look_at((0.5 * size, 0.5 * size, 0.1 * size), (0.5 * size, 0.45 * size, size), (0, 1, 0)), 60.0)
    return s, cam


def wide_materials_scene():
    """A 4 x 3 wall of large spheres, one BSDF kind each (plastic, rough dielectric, rough
    plastic, both microfacet distributions, and the path's original four), over a diffuse
    floor under two quad lights; the camera sees mostly spheres."""
    from spica_amd.scene import (GOLD, SILVER, conductor, dielectric, diffuse, plastic, quad, roughconductor,
                                 roughdielectric, roughplastic, unit_sphere)
    s = Scene()
    mats = [diffuse((0.7, 0.3, 0.2)), conductor(**SILVER), dielectric(1.5), roughconductor(0.2, **GOLD),
            roughconductor(0.3, distribution="beckmann", **GOLD), plastic((0.2, 0.5, 0.8), (1.0, 1.0, 1.0), 1.5),
            roughdielectric(0.2, 1.5, distribution="beckmann"), roughdielectric(0.1, 1.5, distribution="ggx"),
            roughplastic(0.15, (0.8, 0.6, 0.2), distribution="beckmann"), roughplastic(0.3, (0.3, 0.8, 0.3), distribution="ggx"),
            plastic((0.9, 0.9, 0.9), (0.5, 0.5, 0.5), 1.3), roughdielectric(0.4, 1.33, distribution="beckmann")]
    unit = unit_sphere()
    for i, m in enumerate(mats):
        cx, cy = -3.0 + 2.0 * (i % 4), -1.0 + 2.0 * (i // 4)
        s.add_triangles(unit * 0.95 + np.array([cx, cy, 0.0]), m)
    s.add_triangles(quad((-8, -2.2, -8), (-8, -2.2, 8), (8, -2.2, 8), (8, -2.2, -8)), diffuse((0.5, 0.5, 0.5)))
    black = s.add_material(diffuse((0.0, 0.0, 0.0)))
    for cx in (-3.0, 3.0):
        y, h = 5.0, 1.0
        s.add_triangles(quad((cx - h, y, -h - 2), (cx + h, y, -h - 2), (cx + h, y, h - 2), (cx - h, y, h - 2)), black,
                        radiance=(15.0, 15.0, 15.0))
    cam = Camera(48, 32, look_at((0.0, 0.5, -9.0), (0.0, 0.0, 0.0), (0.0, 1.0, 0.0)), 40.0)
    return s, cam


def small_mixed_scene(seed=3, n_spheres=24):
    """Sphere field with every material kind, for whole-path parity at small size."""
    s, _ = configs.sphere_field(seed=seed, n_spheres=n_spheres)
    cam = Camera(48, 32, look_at((0.0, 3.0, -18.0), (0.0, -1.0, 0.0), (0.0, 1.0, 0.0)), 40.0)
    return s, cam
