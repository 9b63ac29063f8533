"""CPU-baseline calibration -- test infrastructure, runs where oracle/_ref is built.

bench.py times the float64 restatement (liboracle.so) on the GPU host's cores.  This
script times spica's own compiled path (oracle/_ref: PathIntegrator::Li with an
Independent MT19937 sampler, spica's BVH) against the restatement on the same C0
camera paths, one thread each, so the bench's CPU figure maps to "spica's CPU path":
spica rate = restatement rate x (spica / restatement) from here.

    python -m oracle.calibrate [paths]
"""
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path[0] = ROOT  # not this directory: its oracle.py would shadow the package

from oracle.oracle import Oracle, Reference, have_ref  # noqa: E402
from spica_amd import configs  # noqa: E402


def main(n=20000):
    if not have_ref():
        raise SystemExit("oracle/_ref is not built (needs the spica sources)")
    w = configs.cornell_box()
    orc, ref = Oracle(), Reference()
    osc, rsc = orc.scene(w.scene), ref.scene(w.scene)
    rng = np.random.default_rng(3)
    cam = w.camera
    px = rng.choice(cam.width * cam.height, size=n, replace=False).astype(np.int32)
    orgs, dirs = [], []
    for p in px:
        y, x = divmod(int(p), cam.width)
        o, d = orc.camera_ray(cam, x, y, rng.random(4))
        orgs.append(o)
        dirs.append(d)
    seeds = rng.integers(1, 2 ** 31, n).astype(np.uint32)
    t0 = time.perf_counter()
    rsc.li(np.array(orgs), np.array(dirs), seeds, w.max_depth)
    t_ref = time.perf_counter() - t0
    # the restatement on the same pixels (film order) at 1 spp, one thread, its own render loop
    film = (px // cam.width) * cam.width + (cam.width - 1 - px % cam.width)
    t0 = time.perf_counter()
    osc.render(cam, w.render_desc(spp=1, sampler=0, seed=9), threads=1, subset=film.astype(np.int32))
    t_orc = time.perf_counter() - t0
    r_ref, r_orc = n / t_ref / 1e6, n / t_orc / 1e6
    print(f"C0, {n} paths, 1 thread: spica (compiled) {r_ref:.4f} Mpaths/s, restatement {r_orc:.4f} Mpaths/s, "
          f"spica / restatement = {r_ref / r_orc:.3f}")


if __name__ == "__main__":
    main(int(sys.argv[1]) if len(sys.argv) > 1 else 20000)
