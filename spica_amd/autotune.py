"""Autotuner for the wavefront render: paths in flight (queue) and traversal block size.

Passes regenerate paths: a pass covers up to 2^28 paths (whole samples of the film) and keeps
`queue` of them in flight, starting the pass's next camera path in every slot a finished path
frees, so every launch works on a full queue of mixed-depth paths.  More paths in flight hide
more latency per launch and need fewer launches, at ~272 B of queue memory each; the
candidates span 2^20..2^27 (the largest queue a render takes) for single renders and
2^20..2^23 for batches.  The tuner times a pass per candidate, keeps the fastest by traced rays
per second and caches the choice per (scene size, film, depth, device).
"""
import json
import os
import time

DEFAULT_QUEUES = tuple(1 << k for k in range(20, 28))
BATCH_QUEUES = tuple(1 << k for k in range(20, 24))
PASS_PATHS = 1 << 28  # paths per regenerating pass
DEFAULT_BLOCKS = (64, 128, 256)
CACHE = os.path.join(os.path.expanduser("~"), ".cache", "spica_amd", "autotune.json")


def _key(gscene, camera, desc):
    import torch

    st = gscene.stats()
    return (f"{torch.cuda.get_device_name(0)}|tris={st['num_tris']}|{camera.width}x{camera.height}|depth={desc.max_depth}"
            f"|regen28")


def _load():
    try:
        with open(CACHE) as f:
            return json.load(f)
    except (OSError, ValueError):
        return {}


def _store(db):
    try:
        os.makedirs(os.path.dirname(CACHE), exist_ok=True)
        with open(CACHE, "w") as f:
            json.dump(db, f, indent=1)
    except OSError:
        pass


def tune(gscene, camera, desc, queues=DEFAULT_QUEUES, blocks=DEFAULT_BLOCKS, spp=8, use_cache=True, log=None,
         max_spp=64):
    """Returns (queue_size, block_size, table) for this scene/film/depth."""
    import torch

    from . import render

    key = _key(gscene, camera, desc)
    db = _load() if use_cache else {}
    if key in db:
        e = db[key]
        return e["queue"], e["block"], e.get("table", [])
    npix = camera.width * camera.height
    table = []
    best = None
    per_pass = min(max(1, PASS_PATHS // npix), max(1, desc.spp))
    for q in queues:
        n = min(max(spp, 2 * per_pass), max_spp)  # two full passes where they fit in max_spp
        for b in blocks:
            gr = render.GpuRender(gscene, camera, desc, queue_size=q, block_size=b)
            gr.run(per_pass)
            st0 = gr.stats()
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            gr.run(n)
            torch.cuda.synchronize()
            dt = time.perf_counter() - t0
            st1 = gr.stats()
            gr.close()
            rays = sum(st1[k] - st0[k] for k in ("rays_extend", "rays_shadow", "rays_mis"))
            rate = rays / dt / 1e6
            table.append(dict(queue=q, block=b, mrays_per_s=round(rate, 1), ms_per_spp=round(dt * 1e3 / n, 3)))
            if log:
                log(f"autotune: queue 2^{q.bit_length() - 1} block {b}: {rate:.1f} Mrays/s")
            if best is None or rate > best[0]:
                best = (rate, q, b)
    db[key] = dict(queue=best[1], block=best[2], table=table)
    if use_cache:
        _store(db)
    return best[1], best[2], table


def tune_batch(items, max_depth, sampler, block, queues=BATCH_QUEUES, paths=1 << 27, log=None):
    """Paths in flight for a batch: times every candidate on the whole batch's images (each
    image's own scene and film) with their sample counts scaled so the probe holds about
    `paths` paths, keeps the fastest by traced rays per second.  Returns (queue, table)."""
    import torch

    from . import render

    total = sum(cam.width * cam.height * spp for _, cam, spp, _ in items)
    f = min(1.0, paths / max(total, 1))
    probe = [(g, cam, max(1, int(round(spp * f))), seed) for g, cam, spp, seed in items]
    table = []
    best = None
    for q in queues:
        b = render.Batch(probe, max_depth=max_depth, sampler=sampler, queue_size=q, block_size=block, graph=1)
        b.run(sync=True)  # warm: first launches, graph capture
        st0 = b.stats()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        b.run(sync=True)
        torch.cuda.synchronize()
        dt = time.perf_counter() - t0
        st1 = b.stats()
        b.close()
        rays = sum(st1[k] - st0[k] for k in ("rays_extend", "rays_shadow", "rays_mis"))
        rate = rays / dt / 1e6
        table.append(dict(queue=q, mrays_per_s=round(rate, 1), seconds=round(dt, 4)))
        if log:
            log(f"autotune batch: queue 2^{q.bit_length() - 1}: {rate:.1f} Mrays/s")
        if best is None or rate > best[0]:
            best = (rate, q)
    return best[1], table
