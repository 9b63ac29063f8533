"""Renders the float64 CPU references used for relMSE (the checker's own path, spica's
estimator, independent sampler with a seed no GPU run uses).

    python tests/golden/make_reference.py c0            # 256x256 full frame, 65536 spp
    python tests/golden/make_reference.py c3 [nx]        # 1280x720, 160x90 stratified pixels, 65536 spp
    python tests/golden/make_reference.py c1             # 1920x1080, 160x90 stratified pixels, 65536 spp
    python tests/golden/make_reference.py c4             # view 0 of the heavy scene, 80x45 stratified pixels, 65536 spp
    python tests/golden/make_reference.py c2 [i ...]     # batch images i (default 0 1 2), 40 x ~40h/w pixels, 65536 spp

Writes tests/golden/<name>_ref.npz with the film indices, per-pixel mean and the
standard error of the mean.  CPU-only; run in the build container.
"""
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from oracle.oracle import Oracle  # noqa: E402
from spica_amd import configs  # noqa: E402

SEED = 0x5EED5EED


def subset_pixels(w, h, nx, ny):
    """Stratified film pixels: one per cell of an nx x ny grid, centred."""
    xs = ((np.arange(nx) + 0.5) * w / nx).astype(int)
    ys = ((np.arange(ny) + 0.5) * h / ny).astype(int)
    return np.array([y * w + x for y in ys for x in xs], dtype=np.int32)


def render(name, wl, subset, spp, chunk, threads, write=True):
    sc = Oracle().scene(wl.scene)
    npix = len(subset) if subset is not None else wl.camera.width * wl.camera.height
    state = (np.zeros((npix, 3)), np.zeros((npix, 3)), np.zeros(npix, dtype=np.int64))
    t0 = time.time()
    for first in range(0, spp, chunk):
        rd = wl.render_desc(spp=min(chunk, spp - first), first_sample=first, sampler=0, seed=SEED)
        sc.render(wl.camera, rd, threads=threads, subset=subset, state=state)
        print(f"{name}: {first + chunk}/{spp} spp, {time.time() - t0:.0f} s", flush=True)
    s, q, c = state
    mean = s / c[:, None]
    var = np.maximum(q / c[:, None] - mean ** 2, 0.0) * c[:, None] / (c[:, None] - 1)
    res = dict(pixels=np.arange(npix) if subset is None else subset, mean=mean,
               sem=np.sqrt(var / c[:, None]).astype(np.float32), spp=spp, seed=SEED)
    if write:
        out = os.path.join(os.path.dirname(os.path.abspath(__file__)), f"{name}_ref.npz")
        np.savez_compressed(out, **res)
        print("wrote", out)
    return res


def render_batch(images, threads):
    """References for some images of the batch workload, one npz holding all of them."""
    wb = configs.c2_batch()
    out = {"images": np.array(images, dtype=np.int32), "spp": 65536, "seed": SEED}
    for i in images:
        sc, cam, _, _ = wb.items[i]
        wl = configs.Workload(f"C2 image {i}", sc, cam, 65536, wb.max_depth, 0, SEED)
        nx = 40
        ny = max(8, int(round(nx * cam.height / cam.width)))
        r = render(f"c2[{i}]", wl, subset_pixels(cam.width, cam.height, nx, ny), 65536, 4096, threads, write=False)
        for k in ("pixels", "mean", "sem"):
            out[f"{k}_{i}"] = r[k]
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "c2_ref.npz")
    np.savez_compressed(path, **out)
    print("wrote", path)


if __name__ == "__main__":
    which = sys.argv[1]
    threads = int(os.environ.get("REF_THREADS", os.cpu_count() or 8))
    if which == "c0":
        render("c0", configs.cornell_box(), None, 65536, 1024, threads)
    elif which == "c3":
        n = int(sys.argv[2]) if len(sys.argv) > 2 else 0
        wl = configs.c3_occluded_glossy()
        nx, ny = (160, 90) if n == 0 else (n, max(1, n * 9 // 16))
        render("c3", wl, subset_pixels(wl.camera.width, wl.camera.height, nx, ny), 65536, 1024, threads)
    elif which == "c1":
        wl = configs.c1_spheres()
        render("c1", wl, subset_pixels(wl.camera.width, wl.camera.height, 160, 90), 65536, 1024, threads)
    elif which == "c2":
        render_batch([int(a) for a in sys.argv[2:]] or [0, 1, 2], threads)
    elif which == "c4":
        wl = configs.c4_heavy()[0]
        render("c4", wl, subset_pixels(wl.camera.width, wl.camera.height, 80, 45), 65536, 1024, threads)
