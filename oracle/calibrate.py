"""CPU-baseline calibration -- test infrastructure, runs where oracle/_ref is built.

bench.py times the float64 restatement (liboracle.so) on the GPU host's cores.  This
script times spica's own compiled path (oracle/_ref: PathIntegrator::Li with an
Independent MT19937 sampler, spica's BVH) against the restatement on the same camera
paths, one thread each, so the bench's CPU figure maps to "spica's CPU path":
spica rate = restatement rate x (spica / restatement) from here.  Both workloads are
measured: C0 (36 triangles, maxDepth 8) and C1 (327,686 triangles, maxDepth 16), since the
ratio depends on how much of a path's time goes to traversal.

    python -m oracle.calibrate [--write] [c0_paths c1_paths]
"""
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path[0] = ROOT  # not this directory: its oracle.py would shadow the package

from oracle.oracle import Oracle, Reference, have_ref  # noqa: E402
from spica_amd import configs  # noqa: E402

OUT = os.path.join(ROOT, "oracle", "calibration.json")


def measure(w, n, seed=3):
    """Mpaths/s of spica's compiled Li and of the restatement's render loop on the same n
    random pixels of workload w, one thread each."""
    orc, ref = Oracle(), Reference()
    t0 = time.perf_counter()
    osc = orc.scene(w.scene)
    t_build_orc = time.perf_counter() - t0
    t0 = time.perf_counter()
    rsc = ref.scene(w.scene)
    t_build_ref = time.perf_counter() - t0
    rng = np.random.default_rng(seed)
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
    return dict(paths=n, spica_mpaths_per_s=round(r_ref, 5), restatement_mpaths_per_s=round(r_orc, 5),
                ratio=round(r_ref / r_orc, 3), bvh_build_s=dict(spica=round(t_build_ref, 2),
                                                                restatement=round(t_build_orc, 2)))


def main(argv):
    if not have_ref():
        raise SystemExit("oracle/_ref is not built (needs the spica sources)")
    write = "--write" in argv
    nums = [int(a) for a in argv if not a.startswith("--")]
    n0, n1 = (nums + [20000, 4000])[:2] if nums else (20000, 4000)
    res = {}
    for name, w, n in (("c0", configs.cornell_box(), n0), ("c1", configs.c1_spheres(), n1)):
        r = measure(w, n)
        res[name] = r
        print(f"{name}: {n} paths, 1 thread: spica (compiled) {r['spica_mpaths_per_s']:.4f} Mpaths/s, restatement "
              f"{r['restatement_mpaths_per_s']:.4f} Mpaths/s, spica / restatement = {r['ratio']:.3f}", flush=True)
    if write:
        out = dict(
            ratios=res,
            workloads=dict(c0="C0 Cornell box (36 triangles), random camera paths, maxDepth 8",
                           c1="C1 sphere field (327,686 triangles), random camera paths, maxDepth 16"),
            spica="oracle/_ref (spica's PathIntegrator::Li, BVHAccel, Independent MT19937; g++ -O2)",
            restatement="oracle/liboracle.so (gcc -O2), its own render loop, independent sampler",
            host=cpu_model(), command="python -m oracle.calibrate --write")
        with open(OUT, "w") as f:
            json.dump(out, f, indent=1)
        print(f"wrote {OUT}")


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


if __name__ == "__main__":
    main(sys.argv[1:])
