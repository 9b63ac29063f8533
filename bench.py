"""Benchmark: plain wavefront path tracing of the 512-sphere field on one MI355X.

Workload: seed-1 sphere field, 512 spheres x 640 triangles
on a ground quad (327,686 triangles), two quad lights, 1920x1080, maxDepth 16,
Russian roulette after bounce 3, Sobol sampler.  One step = one sample per pixel
over the whole frame (2,073,600 paths), i.e. one pass of the hot path; the full
256-spp render is 256 steps.

Printed (one JSON line): total traced rays per second (camera + continuation +
shadow + MIS rays, all traversals), the closest-hit traversal kernel against the
HBM roofline, the relMSE of the rendered C1 image against its 65536-spp float64 CPU
reference, the wall time to relMSE 1e-3 of plain and guided tracing on the
occluded-glossy scene (C3, online-trained proposal included in the guided time), the
64-image batch (C2) throughput, the heavy scene (C4: 8-view throughput, plain and guided
relMSE of view 0), and
spica's CPU path (the FP64 C restatement, the reference's estimator and BVH) timed
on this host's cores on a bounded sample.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "time to relMSE 1e-3 (guided vs plain PT, 1 GPU); Mrays/s + roofline frac"
HBM_PEAK_GBS = 8000.0
B_RAY_IO = 32 + 16          # ray in (o, d, tmax, id) + hit out (t, u, v, prim)
B_NODE = 64                 # one 4-wide BVH node: four 8-bit-quantised child boxes against the node's box + links
B_TRI = 48                  # one triangle: v0 | id, e1, e2


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def read_traffic(profile_dir, rays_per_launch):
    """HBM bytes per closest-hit launch: the committed rocprofv3 PMC figure per ray times
    this run's rays per launch (the profile states its own queue/block), or None."""
    path = os.path.join(profile_dir, "extend_c1_traffic.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        d = json.load(f)
    if "bytes_per_ray" not in d:
        return None
    return round(d["bytes_per_ray"] * rays_per_launch)


def host_cpus():
    """What this host gives the process: CPU model, logical CPUs, the affinity set and the
    cgroup CPU quota (cores' worth of time, None when unlimited)."""
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    quota = None
    try:
        with open("/sys/fs/cgroup/cpu.max") as f:
            q, per = f.read().split()[:2]
            if q != "max":
                quota = round(int(q) / int(per), 2)
    except (OSError, ValueError):
        pass
    return dict(model=model, logical_cpus=os.cpu_count(), affinity=len(os.sched_getaffinity(0)), cgroup_quota_cpus=quota)


def cpu_rate(sc, w, threads, budget_s, seed):
    """Rays and paths per second of the CPU checker over random pixels of w, ~budget_s seconds."""
    import numpy as np

    rng = np.random.default_rng(seed)
    npix = w.camera.width * w.camera.height
    n = 2000 * max(1, threads // 4)
    rays = paths = 0
    elapsed = 0.0
    while elapsed < budget_s:
        subset = rng.choice(npix, size=n, replace=False).astype(np.int32)
        rd = w.render_desc(spp=1, first_sample=int(rng.integers(0, 1 << 20)))
        t0 = time.perf_counter()
        _, _, _, st = sc.render(w.camera, rd, threads=threads, subset=subset)
        dt = time.perf_counter() - t0
        elapsed += dt
        rays += int(st[0])
        paths += n
        n = int(min(1000000, max(2000, n * max(1.0, 0.5 * (budget_s - elapsed) / max(dt, 1e-3)))))
    return rays, paths, elapsed


def cpu_baseline(w, threads, budget_s):
    """The CPU checker (spica's estimator, FP64, spica's BVH) on a bounded pixel sample, on
    every CPU the process may use, plus one thread for the per-core rate; spica's own compiled
    path runs at the calibrated ratio of the checker's speed (oracle/calibration.json)."""
    from oracle.oracle import Oracle

    orc = Oracle()
    t0 = time.perf_counter()
    sc = orc.scene(w.scene)
    build_s = time.perf_counter() - t0
    rays, paths, elapsed = cpu_rate(sc, w, threads, budget_s, 0)
    r1, p1, e1 = cpu_rate(sc, w, 1, min(5.0, budget_s), 1)
    value = rays / elapsed / 1e6
    one = r1 / e1 / 1e6
    cal = None
    cal_path = os.path.join(ROOT, "oracle", "calibration.json")
    if os.path.exists(cal_path):
        with open(cal_path) as f:
            c = json.load(f)
        r = c["ratios"]["c1"]["ratio"]
        cal = dict(spica_over_port=r, spica_over_port_c0=c["ratios"]["c0"]["ratio"],
                   spica_equivalent_mrays_per_s=round(r * value, 3),
                   measured=f"spica's compiled path vs the checker, one thread each, C1 ({c['ratios']['c1']['paths']} "
                            f"camera paths) and C0 ({c['ratios']['c0']['paths']}); {c['host']}")
    return dict(value=round(value, 3), unit="Mrays/s", cores=threads, kind="port", calibration=cal,
                host=host_cpus(), one_thread_mrays_per_s=round(one, 4),
                scaling_over_one_thread=round(value / one, 2),
                sample=f"{paths} random pixels of the 1920x1080 frame at 1 spp ({rays} rays, {elapsed:.1f} s) on "
                       f"{threads} threads; one thread: {p1} pixels, {e1:.1f} s; spica-style BVH build "
                       f"{build_s:.1f} s excluded",
                mpaths_per_s=round(paths / elapsed / 1e6, 4))


def load_reference(name):
    """65536-spp float64 CPU reference on a stratified pixel subset (tests/golden/make_reference.py)."""
    import numpy as np

    path = os.path.join(ROOT, "tests", "golden", f"{name}_ref.npz")
    if not os.path.exists(path):
        return None
    d = np.load(path)
    return d["pixels"], d["mean"], int(d["spp"])


def build_scenes(scene_list, builder="sah", workers=4):
    """GpuScenes of many scene descriptions, built by a few host threads at once (each host
    SAH build is itself threaded, the device LBVH builds queue on the GPU)."""
    from concurrent.futures import ThreadPoolExecutor

    from spica_amd import render

    with ThreadPoolExecutor(max_workers=workers) as ex:
        return list(ex.map(lambda sc: render.GpuScene(sc, builder=builder), scene_list))


def load_batch_reference():
    """65536-spp float64 CPU references of some batch images (tests/golden/make_reference.py c2)."""
    import numpy as np

    path = os.path.join(ROOT, "tests", "golden", "c2_ref.npz")
    if not os.path.exists(path):
        return None
    d = np.load(path)
    return dict(spp=int(d["spp"]), images={int(i): (d[f"pixels_{i}"], d[f"mean_{i}"]) for i in d["images"]})


def time_to_relmse(gs, w, ref, guide_cfg, target, max_spp, queue, block, chunk=32):
    """Wall time (render calls only, synchronised) until relMSE against the reference
    falls to `target`, checked every `chunk` samples per pixel on the GPU in float64."""
    from spica_amd import render

    pix, refm, _ = ref
    gr = render.GpuRender(gs, w.camera, w.render_desc(spp=max_spp), queue_size=queue, block_size=block)
    g = None
    if guide_cfg is not None:
        g = render.Guide(**guide_cfg)
        gr.set_guide(g)
    t = 0.0
    spp = 0
    e = float("inf")
    hit = None
    trace = []
    while spp < max_spp:
        k = min(chunk, max_spp - spp)
        t0 = time.perf_counter()
        gr.run(k, sync=True)
        t += time.perf_counter() - t0
        spp += k
        e = gr.relmse_pixels(pix, refm)
        if spp in (64, 256, 1024) or e <= target:
            trace.append([spp, round(t, 4), float(f"{e:.4g}")])
        if e <= target:
            hit = (spp, t)
            break
    st = gr.stats()
    gr.close()
    if g is not None:
        g.close()
    return dict(seconds=None if hit is None else round(hit[1], 4), spp=None if hit is None else hit[0],
                last_relmse=float(f"{e:.4g}"), trace=trace, guided_vertices=st["guided_vertices"])


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=64)
    ap.add_argument("--warmup", type=int, default=8)
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--cpu-threads", type=int, default=len(os.sched_getaffinity(0)),
                    help="CPU-baseline threads (default: every CPU this process may run on)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--queue", type=int, default=0, help="paths in flight (0 = autotune)")
    ap.add_argument("--block", type=int, default=0, help="traversal block size (0 = autotune)")
    ap.add_argument("--no-guide", action="store_true", help="skip the guided-vs-plain time-to-relMSE leg")
    ap.add_argument("--no-batch", action="store_true", help="skip the 64-image batch leg")
    ap.add_argument("--no-heavy", action="store_true", help="skip the heavy-scene (C4) leg")
    ap.add_argument("--heavy-spp", type=int, default=256, help="samples per pixel of the C4 relMSE renders")
    ap.add_argument("--copy-probe", type=int, default=0,
                    help="only copy this many bytes device-to-device (calibrates profiler byte counters)")
    args = ap.parse_args()
    if args.copy_probe:
        import torch

        n = args.copy_probe // 4
        a = torch.ones(n, dtype=torch.float32, device="cuda")
        b = torch.empty_like(a)
        for _ in range(3):
            b.copy_(a)
        torch.cuda.synchronize()
        log(f"bench: copied {n * 4} bytes x3")
        return
    if args.gpus != 1:
        print(f"bench.py: this build measures one GPU only (asked for --gpus {args.gpus}); "
              "paths shard by pixel/sample range with one film reduce, not run here", file=sys.stderr)
        sys.exit(2)

    import torch

    from spica_amd import configs, render

    log("bench: building C1 scene")
    w = configs.c1_spheres()
    gs = render.GpuScene(w.scene)
    sst = gs.stats()
    log(f"bench: {sst['num_tris']} tris, {sst['num_nodes']} nodes, BVH build {sst['build_ms']:.0f} ms")

    # Autotune (outside the timed region): paths in flight and traversal block size.
    table = []
    if args.queue == 0 or args.block == 0:
        from spica_amd import autotune

        q, b, table = autotune.tune(gs, w.camera, w.render_desc(), log=log,
                                    queues=(args.queue,) if args.queue else autotune.DEFAULT_QUEUES,
                                    blocks=(args.block,) if args.block else autotune.DEFAULT_BLOCKS)
        args.queue, args.block = q, b
    log(f"bench: queue {args.queue} block {args.block}")

    # Counting pass (outside the timed region): node fetches / triangle tests per closest-hit ray.
    gc = render.GpuRender(gs, w.camera, w.render_desc(spp=4), queue_size=args.queue, block_size=args.block,
                          count_visits=True)
    gc.run(4)
    cst = gc.stats()
    gc.close()
    v_node = cst["count_nodes"] / max(1, cst["count_rays"])
    v_tri = cst["count_tris"] / max(1, cst["count_rays"])
    b_ray = B_RAY_IO + B_NODE * v_node + B_TRI * v_tri
    log(f"bench: V_node {v_node:.2f}  V_tri {v_tri:.2f}  B_ray {b_ray:.0f} B")

    gr = render.GpuRender(gs, w.camera, w.render_desc(spp=args.warmup + args.steps), queue_size=args.queue,
                          block_size=args.block, profile=True)
    gr.run(args.warmup)
    st0 = gr.stats()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    gr.run(args.steps, sync=True)
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    st1 = gr.stats()
    rays = (st1["rays_extend"] - st0["rays_extend"]) + (st1["rays_shadow"] - st0["rays_shadow"]) + \
           (st1["rays_mis"] - st0["rays_mis"])
    paths = st1["paths"] - st0["paths"]
    ext_ms = st1["extend_ms"] - st0["extend_ms"]
    ext_launches = st1["extend_launches"] - st0["extend_launches"]
    ext_rays = st1["extend_rays_timed"] - st0["extend_rays_timed"]
    value = rays / dt / 1e6
    avg_launch_s = ext_ms / 1e3 / max(1, ext_launches)
    bytes_per_launch = b_ray * ext_rays / max(1, ext_launches)
    achieved = bytes_per_launch / avg_launch_s / 1e9
    log(f"bench: {args.steps} steps in {dt*1e3:.1f} ms: {value:.1f} Mrays/s, {paths/dt/1e6:.1f} Mpaths/s; "
        f"closest-hit {ext_rays} rays in {ext_ms:.1f} ms over {ext_launches} launches")

    c1_ref = load_reference("c1")
    c1_relmse = None
    if c1_ref is not None:
        c1_relmse = dict(value=float(f"{gr.relmse_pixels(c1_ref[0], c1_ref[1]):.5g}"),
                         spp=args.warmup + args.steps,
                         reference=f"{c1_ref[2]}-spp float64 CPU render, {len(c1_ref[0])} stratified pixels")
        log(f"bench: C1 relMSE at {args.warmup + args.steps} spp: {c1_relmse['value']}")
    gr.close()

    ttr = None
    c3_ref = load_reference("c3")
    if not args.no_guide and c3_ref is not None:
        w3 = configs.c3_occluded_glossy()
        gs3 = render.GpuScene(w3.scene)
        guide_cfg = dict(batch_size=1 << 18, train_spp=64, train_pass_spp=4, steps_per_pass=2, learning_rate=1e-2,
                         guide_fraction=0.5)
        warm = render.GpuRender(gs3, w3.camera, w3.render_desc(spp=16), queue_size=args.queue, block_size=args.block)
        warm.run(16)
        warm.close()
        log("bench: C3 time to relMSE 1e-3, plain")
        plain = time_to_relmse(gs3, w3, c3_ref, None, 1e-3, 16384, args.queue, args.block)
        log(f"bench: plain {plain}")
        log("bench: C3 time to relMSE 1e-3, guided")
        guided = time_to_relmse(gs3, w3, c3_ref, guide_cfg, 1e-3, 16384, args.queue, args.block)
        log(f"bench: guided {guided}")
        ratio = None
        if plain["seconds"] and guided["seconds"]:
            ratio = round(guided["seconds"] / plain["seconds"], 4)
        ttr = dict(workload="C3 occluded glossy: seed-7 sphere field (60% GGX alpha 0.1-0.3), light inside a box open "
                            "towards the ceiling, 1280x720, maxDepth 16, Sobol seed 7",
                   target=1e-3, eps=1e-2,
                   reference=f"{c3_ref[2]}-spp float64 CPU render (spica's estimator), {len(c3_ref[0])} stratified pixels",
                   plain=plain, guided=guided, guided_over_plain=ratio,
                   guide=dict(guide_cfg, network="one-blob(7x32) -> 64x4 ReLU -> 8 vMF lobes, bf16 MFMA, fp32 accum"))
        gs3.close()

    batch = None
    if not args.no_batch:
        log("bench: C2 batch: generating 64 scenes")
        t0 = time.perf_counter()
        wb = configs.c2_batch()
        gen_s = time.perf_counter() - t0
        t0 = time.perf_counter()
        scenes = build_scenes([sc for sc, _, _, _ in wb.items])
        scene_s = time.perf_counter() - t0
        tris = sum(g.stats()["num_tris"] for g in scenes)
        bvh_s = sum(g.stats()["build_ms"] for g in scenes) / 1e3  # summed over builds running side by side on 4 threads
        setup_s = gen_s + scene_s
        items = [(g, cam, spp, seed) for g, (_, cam, spp, seed) in zip(scenes, wb.items)]
        b = render.Batch(items, max_depth=wb.max_depth, sampler=wb.sampler, queue_size=args.queue,
                         block_size=args.block, graph=1)
        st0 = b.stats()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        b.run(sync=True)
        torch.cuda.synchronize()
        bdt = time.perf_counter() - t0
        st1 = b.stats()
        brays = sum(st1[k] - st0[k] for k in ("rays_extend", "rays_shadow", "rays_mis"))
        info = b.info()
        batch = dict(workload="C2: 64 images, image i = C1 generator seed i with U{16..2048} spheres, W,H ~ U{128..1024}, "
                              "spp ~ U{16..512}, maxDepth 16, Sobol seed i; one per-pass hipGraph replayed over shared passes",
                     images=len(items), triangles=int(tris), paths=int(info["paths"]), passes=int(info["passes"]),
                     seconds=round(bdt, 3), mrays_per_s=round(brays / bdt / 1e6, 3),
                     mpaths_per_s=round(info["paths"] / bdt / 1e6, 3), setup_s=round(setup_s, 1),
                     generate_s=round(gen_s, 1), scenes_s=round(scene_s, 1), bvh_build_thread_s=round(bvh_s, 1),
                     queue=int(info["queue"]), block=args.block)
        c2_ref = load_batch_reference()
        if c2_ref is not None:
            import numpy as np

            rel = {}
            for i, (pix, refm) in c2_ref["images"].items():
                s, _, c = b.film(i)
                m = (s / np.maximum(c, 1)[..., None]).reshape(-1, 3)[pix]
                rel[str(i)] = dict(spp=int(wb.items[i][2]),
                                   relmse=float(f"{float(((m - refm) ** 2 / (refm ** 2 + 1e-2)).mean()):.4g}"))
            batch["relmse"] = dict(images=rel, eps=1e-2,
                                   reference=f"{c2_ref['spp']}-spp float64 CPU render, stratified pixels per image")
        b.close()
        for g in scenes:
            g.close()
        # the same batch over trees built on the device: the time to the first pixel
        t0 = time.perf_counter()
        scenes = build_scenes([sc for sc, _, _, _ in wb.items], builder="lbvh")
        lscene_s = time.perf_counter() - t0
        lbvh_s = sum(g.stats()["build_ms"] for g in scenes) / 1e3
        items = [(g, cam, spp, seed) for g, (_, cam, spp, seed) in zip(scenes, wb.items)]
        b = render.Batch(items, max_depth=wb.max_depth, sampler=wb.sampler, queue_size=args.queue,
                         block_size=args.block, graph=1)
        st0 = b.stats()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        b.run(sync=True)
        torch.cuda.synchronize()
        ldt = time.perf_counter() - t0
        st1 = b.stats()
        lrays = sum(st1[k] - st0[k] for k in ("rays_extend", "rays_shadow", "rays_mis"))
        batch["lbvh"] = dict(scenes_s=round(lscene_s, 1), bvh_build_thread_s=round(lbvh_s, 1), seconds=round(ldt, 3),
                             mrays_per_s=round(lrays / ldt / 1e6, 3))
        log(f"bench: batch {batch}")
        b.close()
        for g in scenes:
            g.close()

    heavy = None
    if not args.no_heavy:
        log("bench: C4 heavy scene")
        t0 = time.perf_counter()
        views = configs.c4_heavy()
        gs4 = render.GpuScene(views[0].scene)
        st4 = gs4.stats()
        setup_s = time.perf_counter() - t0
        vspp = 16
        items = [(gs4, v.camera, vspp, v.seed) for v in views]
        b = render.Batch(items, max_depth=views[0].max_depth, sampler=views[0].sampler, queue_size=args.queue,
                         block_size=args.block, graph=1)
        b.run()
        st0 = b.stats()
        t0 = time.perf_counter()
        b.run(sync=True)
        bdt = time.perf_counter() - t0
        st1 = b.stats()
        brays = sum(st1[k] - st0[k] for k in ("rays_extend", "rays_shadow", "rays_mis"))
        bpaths = st1["paths"] - st0["paths"]
        b.close()
        # the same batch over a linear BVH built on the device
        gl = render.GpuScene(views[0].scene, builder="lbvh")
        stl = gl.stats()
        bl = render.Batch([(gl,) + it[1:] for it in items], max_depth=views[0].max_depth, sampler=views[0].sampler,
                          queue_size=args.queue, block_size=args.block, graph=1)
        bl.run()
        sl0 = bl.stats()
        t0 = time.perf_counter()
        bl.run(sync=True)
        ldt = time.perf_counter() - t0
        sl1 = bl.stats()
        lrays = sum(sl1[k] - sl0[k] for k in ("rays_extend", "rays_shadow", "rays_mis"))
        bl.close()
        gl.close()
        heavy = dict(workload="C4: seed 11, 6000 spheres + 2 noise-displaced 1M-triangle grids, 1024x512 HDR sky "
                              "(sun 1e4, gradient 0.5-2.0), 4 quad lights; 8 views 1920x1080, maxDepth 32, Sobol",
                     triangles=int(st4["num_tris"]), bvh_build_ms=round(st4["build_ms"], 1), setup_s=round(setup_s, 1),
                     batch=dict(views=len(items), spp=vspp, seconds=round(bdt, 3), mrays_per_s=round(brays / bdt / 1e6, 3),
                                mpaths_per_s=round(bpaths / bdt / 1e6, 3), rays_per_path=round(brays / bpaths, 3)),
                     lbvh=dict(bvh_build_ms=round(stl["build_ms"], 1), nodes=int(stl["num_nodes"]),
                               max_depth=int(stl["max_depth"]), batch_seconds=round(ldt, 3),
                               mrays_per_s=round(lrays / ldt / 1e6, 3)))
        c4_ref = load_reference("c4")
        if c4_ref is not None:
            pix, refm, rspp = c4_ref
            res = {}
            for name, guided in (("plain", False), ("guided", True)):
                v = views[0]
                gr = render.GpuRender(gs4, v.camera, v.render_desc(spp=args.heavy_spp), queue_size=args.queue,
                                      block_size=args.block)
                g = None
                if guided:
                    g = render.Guide(train_spp=64, train_pass_spp=4, steps_per_pass=2)
                    gr.set_guide(g)
                t0 = time.perf_counter()
                gr.run(args.heavy_spp, sync=True)
                vdt = time.perf_counter() - t0
                res[name] = dict(spp=args.heavy_spp, seconds=round(vdt, 3),
                                 relmse=float(f"{gr.relmse_pixels(pix, refm):.5g}"))
                gr.close()
                if g is not None:
                    g.close()
            heavy["view0"] = dict(res, reference=f"{rspp}-spp float64 CPU render, {len(pix)} stratified pixels of view 0")
        log(f"bench: heavy {heavy}")
        gs4.close()

    cpu = None
    if not args.no_cpu:
        log(f"bench: CPU baseline on {args.cpu_threads} threads for ~{args.cpu_seconds:.0f} s")
        cpu = cpu_baseline(w, args.cpu_threads, args.cpu_seconds)

    out = {
        "metric": METRIC,
        "value": round(value, 3),
        "unit": "Mrays/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(dt * 1e3 / args.steps, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "C1 sphere field: 512 spheres x 640 tris + ground + 2 quad lights (327686 tris), "
                               "1920x1080, 1 spp per step, maxDepth 16, RR after 3, Sobol seed 1",
                   "queue": args.queue, "block": args.block, "autotune": table},
        "mpaths_per_s": round(paths / dt / 1e6, 3),
        "rays_per_path": round(rays / max(1, paths), 3),
        "roofline": {
            "bound": "hbm",
            "kernel": "k_extend (closest-hit traversal)",
            "achieved": round(achieved, 2),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": read_traffic(os.path.join(ROOT, "profiles"), ext_rays / max(1, ext_launches)),
            "b_ray": round(b_ray, 1),
            "v_node": round(v_node, 3),
            "v_tri": round(v_tri, 3),
            "avg_launch_us": round(avg_launch_s * 1e6, 2),
            "rays_per_launch": round(ext_rays / max(1, ext_launches), 1),
        },
        "relmse": c1_relmse,
        "time_to_relmse": ttr,
        "batch": batch,
        "heavy": heavy,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)
    gs.close()


if __name__ == "__main__":
    main()
