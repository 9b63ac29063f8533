"""The 65536-spp float64 CPU references (tests/golden/*_ref.npz, written by
tests/golden/make_reference.py) that the benchmark's relMSE figures are measured against:
they load without pickle, cover valid film pixels of their workloads, are precise enough for
a relMSE of 1e-3, and a fresh checker render of some of their pixels (another seed) agrees
with them within the combined standard errors."""
import os

import numpy as np
import pytest

from spica_amd import configs

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _load(name):
    return np.load(os.path.join(GOLDEN, f"{name}_ref.npz"), allow_pickle=False)


def _check(pixels, mean, sem, w, h, own_max=1e-4):
    assert pixels.ndim == 1 and len(np.unique(pixels)) == len(pixels)
    assert pixels.min() >= 0 and pixels.max() < w * h
    assert mean.shape == (len(pixels), 3) and sem.shape == mean.shape
    assert np.all(np.isfinite(mean)) and np.all(mean >= 0.0)
    assert np.all(np.isfinite(sem)) and np.all(sem >= 0.0)
    # precision: the reference's own relative error, in the relMSE's metric
    own = float(np.mean(sem.astype(np.float64) ** 2 / (mean ** 2 + 1e-2)))
    print(f"reference's own relMSE {own:.3g}")
    assert own < own_max, own


# far below the 1e-3 the time-to-relMSE figures aim at; C4's sun caustics through glass leave
# its reference at ~2.7e-3, which is why C4 is reported as relMSE at equal samples instead
@pytest.mark.parametrize("name,wl,own_max", [("c0", lambda: configs.cornell_box(), 1e-4),
                                             ("c1", lambda: configs.c1_spheres(), 1e-4),
                                             ("c3", lambda: configs.c3_occluded_glossy(), 1e-4),
                                             ("c4", lambda: configs.c4_heavy(views=1)[0], 5e-3)])
def test_reference_files(name, wl, own_max):
    d = _load(name)
    assert int(d["spp"]) == 65536
    cam = wl().camera
    _check(d["pixels"], d["mean"], d["sem"], cam.width, cam.height, own_max)


def test_batch_reference_file():
    d = _load("c2")
    assert int(d["spp"]) == 65536 and list(d["images"]) == [0, 1]
    wb = configs.c2_batch()
    for i in d["images"]:
        cam = wb.items[int(i)][1]
        _check(d[f"pixels_{i}"], d[f"mean_{i}"], d[f"sem_{i}"], cam.width, cam.height)


def test_cornell_reference_matches_a_fresh_checker_render():
    from oracle.oracle import Oracle

    d = _load("c0")
    wl = configs.cornell_box()
    rng = np.random.default_rng(5)
    pick = np.sort(rng.choice(len(d["pixels"]), size=24, replace=False))
    subset = d["pixels"][pick].astype(np.int32)
    sc = Oracle().scene(wl.scene)
    n = len(subset)
    state = (np.zeros((n, 3)), np.zeros((n, 3)), np.zeros(n, dtype=np.int64))
    spp = 4096
    sc.render(wl.camera, wl.render_desc(spp=spp, sampler=0, seed=12345), threads=4, subset=subset, state=state)
    s, q, c = state
    m = s / c[:, None]
    var = np.maximum(q / c[:, None] - m ** 2, 0.0) * c[:, None] / (c[:, None] - 1)
    se = np.sqrt(var / c[:, None] + d["sem"][pick].astype(np.float64) ** 2)
    z = np.where(se > 0, (m - d["mean"][pick]) / np.maximum(se, 1e-30), 0.0)
    print(f"fresh render vs reference: max |z| {np.abs(z).max():.2f}")
    assert np.mean(np.abs(z) > 3) <= 0.05 and np.abs(z).max() < 5
