"""Benchmark: plain wavefront path tracing of the 512-sphere field on one MI355X.

Workload (the plain tracer at scale): seed-1 sphere field, 512 spheres x 640 triangles
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
import copy
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
BF16_PEAK_TFLOPS = 2500.0   # MI355X dense bf16 MFMA
# guide network 224-64-64-64-64-40: FLOPs per vertex of one forward pass (2 per multiply-add)
GUIDE_FWD_FLOP = 2 * (224 * 64 + 3 * 64 * 64 + 64 * 40)
# one training sample: forward, back-propagated deltas of layers 5..2, and all weight gradients
GUIDE_TRAIN_FLOP = GUIDE_FWD_FLOP + 2 * (64 * 40 + 3 * 64 * 64) + GUIDE_FWD_FLOP


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


def guide_roofline(gs3, w3, queue, block, guide_cfg):
    """The guide kernel (one-blob MLP inference on MFMA + mixture/BSDF sampling per vertex) and the
    training step, against the bf16 MFMA peak: a profiled guided render (HIP events around every
    guide launch) over the 64 samples after the 64 training samples, the regime the rest of a
    time-to-relMSE render runs in (training passes hold 4 samples each and drain often), and
    timed training steps on one batch."""
    import numpy as np

    from spica_amd import render

    gr = render.GpuRender(gs3, w3.camera, w3.render_desc(spp=128), queue_size=queue, block_size=block, profile=2)
    g = render.Guide(**guide_cfg)
    gr.set_guide(g)
    gr.run(64, sync=True)
    st0 = gr.stats()
    gr.run(64, sync=True)
    st1 = gr.stats()
    gr.close()
    g.close()
    launches = max(1, st1["guide_launches"] - st0["guide_launches"])
    verts = (st1["guided_vertices"] - st0["guided_vertices"]) / launches
    avg_s = (st1["guide_ms"] - st0["guide_ms"]) / 1e3 / launches
    achieved = GUIDE_FWD_FLOP * verts / avg_s / 1e12
    # training: steps on one host batch of batch_size records; the first call includes the upload
    b = guide_cfg["batch_size"]
    rng = np.random.default_rng(5)
    x = rng.random((b, 7)).astype(np.float32)
    d = rng.normal(size=(b, 3)).astype(np.float32)
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    wgt = rng.random(b).astype(np.float32)
    g = render.Guide(**guide_cfg)
    g.train(x, d, wgt, steps=1)
    t0 = time.perf_counter()
    g.train(x, d, wgt, steps=1)
    t1 = time.perf_counter()
    g.train(x, d, wgt, steps=21)
    t2 = time.perf_counter()
    g.close()
    step_s = ((t2 - t1) - (t1 - t0)) / 20
    train_tflops = GUIDE_TRAIN_FLOP * b / step_s / 1e12
    return dict(bound="mfma", kernel="k_guide (inference + per-vertex mixture/BSDF sampling)",
                flop_per_vertex=GUIDE_FWD_FLOP, vertices_per_launch=round(verts, 1),
                avg_launch_us=round(avg_s * 1e6, 2), achieved=round(achieved, 2), peak=BF16_PEAK_TFLOPS,
                unit="TFLOP/s", frac=round(achieved / BF16_PEAK_TFLOPS, 4),
                training=dict(flop_per_sample=GUIDE_TRAIN_FLOP, batch=b, step_us=round(step_s * 1e6, 1),
                              achieved=round(train_tflops, 2), frac=round(train_tflops / BF16_PEAK_TFLOPS, 4),
                              kernels="k_guide_fb + k_guide_dw + k_guide_psum + k_guide_adam"))


def c3_leg(args, c3_ref):
    """C3 time to relMSE 1e-3, plain and guided, over render seeds 7..10 (the Sobol seed; the
    scene is seed 7's), with the ratio's decomposition: cost per sample (after training) and the
    relMSE at 1024 spp (the error at equal samples)."""
    from spica_amd import configs, render

    w3 = configs.c3_occluded_glossy()
    gs3 = render.GpuScene(w3.scene)
    guide_cfg = dict(batch_size=1 << 18, train_spp=64, train_pass_spp=8, steps_per_pass=8, learning_rate=1e-2,
                     guide_fraction=0.5)
    warm = render.GpuRender(gs3, w3.camera, w3.render_desc(spp=16), queue_size=args.queue, block_size=args.block)
    warm.run(16)
    warm.close()
    runs = []
    for seed in args.c3_seeds:
        ws = copy.copy(w3)  # the same scene and camera, another Sobol seed
        ws.seed = seed
        log(f"bench: C3 time to relMSE 1e-3, render seed {seed}, plain")
        plain = time_to_relmse(gs3, ws, c3_ref, None, 1e-3, 16384, args.queue, args.block)
        log(f"bench: plain {plain}")
        log(f"bench: C3 time to relMSE 1e-3, render seed {seed}, guided")
        guided = time_to_relmse(gs3, ws, c3_ref, guide_cfg, 1e-3, 16384, args.queue, args.block)
        log(f"bench: guided {guided}")
        ratio = None
        if plain["seconds"] and guided["seconds"]:
            ratio = round(guided["seconds"] / plain["seconds"], 4)
        runs.append(dict(seed=seed, plain=plain, guided=guided, guided_over_plain=ratio))
    roof = guide_roofline(gs3, w3, args.queue, args.block, guide_cfg)
    log(f"bench: guide roofline {roof}")
    gs3.close()

    def at(trace, spp):
        for k, t, e in trace:
            if k == spp:
                return t, e
        return None, None

    cost, var = [], []
    for r in runs:
        (tp, ep), (tg, eg) = at(r["plain"]["trace"], 1024), at(r["guided"]["trace"], 1024)
        (tp64, _), (tg64, _) = at(r["plain"]["trace"], 64), at(r["guided"]["trace"], 64)
        if None not in (tp, ep, tg, eg, tp64, tg64):
            # per-sample cost after the 64 training samples, and the error at equal samples
            cost.append(((tg - tg64) / 960) / ((tp - tp64) / 960))
            var.append(eg / ep)
    ratios = [r["guided_over_plain"] for r in runs if r["guided_over_plain"]]
    mean = lambda v: round(sum(v) / len(v), 4) if v else None  # noqa: E731
    first = runs[0]
    return dict(workload="C3 occluded glossy: seed-7 sphere field (60% GGX alpha 0.1-0.3), light inside a box open "
                         "towards the ceiling, 1280x720, maxDepth 16, Sobol",
                target=1e-3, eps=1e-2,
                reference=f"{c3_ref[2]}-spp float64 CPU render (spica's estimator), {len(c3_ref[0])} stratified pixels",
                plain=first["plain"], guided=first["guided"], guided_over_plain=first["guided_over_plain"],
                seeds=[dict(seed=r["seed"], plain_s=r["plain"]["seconds"], guided_s=r["guided"]["seconds"],
                            guided_over_plain=r["guided_over_plain"]) for r in runs],
                guided_over_plain_mean=mean(ratios),
                decomposition=dict(cost_per_sample_guided_over_plain=mean(cost),
                                   relmse_at_1024spp_guided_over_plain=mean(var),
                                   note="guided/plain time ~ cost ratio x error ratio at equal samples"),
                guide=dict(guide_cfg, network="one-blob(7x32) -> 64x4 ReLU -> 8 vMF lobes, bf16 MFMA, fp32 accum"),
                guide_roofline=roof)


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
    ap.add_argument("--batch-queue", type=int, default=0, help="paths in flight of the C2 batch (0 = autotune)")
    ap.add_argument("--no-guide", action="store_true", help="skip the guided-vs-plain time-to-relMSE leg")
    ap.add_argument("--no-batch", action="store_true", help="skip the 64-image batch leg")
    ap.add_argument("--no-heavy", action="store_true", help="skip the heavy-scene (C4) leg")
    ap.add_argument("--no-batch-graph", action="store_true",
                    help="skip the C2 batch as one whole-batch graph launch (a graph of ~50k kernels that "
                         "rocprofv3's kernel trace does not accept)")
    ap.add_argument("--heavy-spp", type=int, default=2048,
                    help="samples per pixel of the C4 relMSE renders (the configured 2048; at <= 512 the guided "
                         "batch also runs with guide seeds 2 and 3)")
    ap.add_argument("--c3-seeds", type=int, nargs="+", default=[7, 8, 9, 10],
                    help="render seeds of the C3 time-to-relMSE leg")
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
    # occlusion rays: ray in (origin | path, direction | tmax) + contribution in, and the any-hit visits
    v_node_s = cst["count_nodes_shadow"] / max(1, cst["count_rays_shadow"])
    v_tri_s = cst["count_tris_shadow"] / max(1, cst["count_rays_shadow"])
    b_shadow = 48 + 16 + B_NODE * v_node_s + B_TRI * v_tri_s
    log(f"bench: V_node {v_node:.2f}  V_tri {v_tri:.2f}  B_ray {b_ray:.0f} B; occlusion V_node {v_node_s:.2f} "
        f"V_tri {v_tri_s:.2f} B {b_shadow:.0f} B")

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
    # The shading and occlusion kernels, from a render with events around every stage (after the
    # timed region, which carries the closest-hit events only).
    gp = render.GpuRender(gs, w.camera, w.render_desc(spp=args.warmup + args.steps), queue_size=args.queue,
                          block_size=args.block, profile=2)
    gp.run(args.warmup)
    sp0 = gp.stats()
    gp.run(args.steps, sync=True)
    sp1 = gp.stats()
    gp.close()
    # shading: per hit the queue entry (48 B) and hit (16 B) in, the triangle's shading record
    # (normal | material 16 B, ids 8 B), the path radiance read and written (32 B); per appended
    # shadow, MIS or continuation ray 48 B out.  Materials and lights stay in cache.
    sh_ms = sp1["shade_ms"] - sp0["shade_ms"]
    sh_hits = sp1["extend_rays_timed"] - sp0["extend_rays_timed"]
    sh_bytes = 120 * sh_hits + 48 * ((sp1["rays_shadow"] - sp0["rays_shadow"]) + (sp1["rays_mis"] - sp0["rays_mis"]) +
                                     (sp1["rays_extend"] - sp0["rays_extend"]) - (sp1["paths"] - sp0["paths"]))
    so_ms = sp1["shadow_ms"] - sp0["shadow_ms"]
    so_rays = sp1["shadow_rays_timed"] - sp0["shadow_rays_timed"]
    other_roofs = {
        "k_shade": dict(bound="hbm", unit="GB/s", peak=HBM_PEAK_GBS, bytes_per_hit=round(sh_bytes / max(1, sh_hits), 1),
                        avg_launch_us=round(sh_ms * 1e3 / max(1, sp1["shade_launches"] - sp0["shade_launches"]), 2),
                        achieved=round(sh_bytes / (sh_ms / 1e3) / 1e9, 2),
                        frac=round(sh_bytes / (sh_ms / 1e3) / 1e9 / HBM_PEAK_GBS, 4)),
        "k_shadow": dict(bound="hbm", unit="GB/s", peak=HBM_PEAK_GBS, b_ray=round(b_shadow, 1),
                         v_node=round(v_node_s, 3), v_tri=round(v_tri_s, 3),
                         avg_launch_us=round(so_ms * 1e3 / max(1, sp1["shadow_launches"] - sp0["shadow_launches"]), 2),
                         achieved=round(b_shadow * so_rays / (so_ms / 1e3) / 1e9, 2),
                         frac=round(b_shadow * so_rays / (so_ms / 1e3) / 1e9 / HBM_PEAK_GBS, 4)),
    }
    log(f"bench: other kernels {other_roofs}")
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
        ttr = c3_leg(args, c3_ref)

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
        bq, btable = args.batch_queue, []
        if bq == 0:
            from spica_amd import autotune

            bq, btable = autotune.tune_batch(items, wb.max_depth, wb.sampler, args.block, log=log)
        b = render.Batch(items, max_depth=wb.max_depth, sampler=wb.sampler, queue_size=bq, block_size=args.block,
                         graph=1)
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
                              "spp ~ U{16..512}, maxDepth 16, Sobol seed i; regenerating passes, one captured group of iterations replayed",
                     images=len(items), triangles=int(tris), paths=int(info["paths"]), passes=int(info["passes"]),
                     seconds=round(bdt, 3), mrays_per_s=round(brays / bdt / 1e6, 3),
                     mpaths_per_s=round(info["paths"] / bdt / 1e6, 3), setup_s=round(setup_s, 1),
                     generate_s=round(gen_s, 1), scenes_s=round(scene_s, 1), bvh_build_thread_s=round(bvh_s, 1),
                     queue=int(info["queue"]), block=args.block, autotune=btable)
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
        if args.no_batch_graph:
            b.close()
        else:
            # the same batch as one whole-batch hipGraph launch: classic passes (a fixed launch sequence
            # per pass, so the graph holds every pass), films checked against the regenerating run's
            import numpy as np

            films = [b.film(i) for i in (0, 1)]
            b.close()
            bg = render.Batch(items, max_depth=wb.max_depth, sampler=wb.sampler, queue_size=bq, block_size=args.block,
                              graph=2, regen=False)
            bg.run(sync=True)  # capture and first launch
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            bg.run(sync=True)
            torch.cuda.synchronize()
            gdt = time.perf_counter() - t0
            same = all(np.array_equal(x, y) for i, f in zip((0, 1), films) for x, y in zip(bg.film(i), f))
            ginfo = bg.info()
            bg.close()
            batch["whole_batch_graph"] = dict(seconds=round(gdt, 3), mrays_per_s=round(brays / gdt / 1e6, 3),
                                              passes=int(ginfo["passes"]), queue=int(ginfo["queue"]),
                                              films_equal_images_0_1=bool(same),
                                              note="classic passes, the whole batch captured as one graph and launched once")
            log(f"bench: C2 whole-batch graph {batch['whole_batch_graph']}")
        for g in scenes:
            g.close()
        # the same batch over trees built on the device: the time to the first pixel
        t0 = time.perf_counter()
        scenes = build_scenes([sc for sc, _, _, _ in wb.items], builder="lbvh")
        lscene_s = time.perf_counter() - t0
        lbvh_s = sum(g.stats()["build_ms"] for g in scenes) / 1e3
        items = [(g, cam, spp, seed) for g, (_, cam, spp, seed) in zip(scenes, wb.items)]
        b = render.Batch(items, max_depth=wb.max_depth, sampler=wb.sampler, queue_size=bq, block_size=args.block,
                         graph=1)
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
        # the 8 views as one batch, plain and with one guide for all views (trained on the first
        # 64 samples of every view, rounds of 8 across the views), to heavy_spp samples each; the
        # guided batch again with guide seeds 2 and 3 (its 256-spp relMSE is decided by single
        # sun-caustic samples, so one trained network is one draw)
        res = {}
        runs = [("plain", False), ("guided", 1)]
        if args.heavy_spp <= 512:
            runs += [("guided_seed2", 2), ("guided_seed3", 3)]
        for name, guided in runs:
            items = [(gs4, v.camera, args.heavy_spp, v.seed) for v in views]
            bh = render.Batch(items, max_depth=views[0].max_depth, sampler=views[0].sampler, queue_size=args.queue,
                              block_size=args.block, graph=1)
            g = None
            if guided:
                g = render.Guide(batch_size=1 << 18, train_spp=64, train_pass_spp=8, steps_per_pass=8, seed=guided)
                bh.set_guide(g)
            st0 = bh.stats()
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            bh.run(sync=True)
            vdt = time.perf_counter() - t0
            st1 = bh.stats()
            hrays = sum(st1[k] - st0[k] for k in ("rays_extend", "rays_shadow", "rays_mis"))
            res[name] = dict(views=len(items), spp=args.heavy_spp, seconds=round(vdt, 3),
                             mrays_per_s=round(hrays / vdt / 1e6, 3),
                             mpaths_per_s=round((st1["paths"] - st0["paths"]) / vdt / 1e6, 3))
            if c4_ref is not None:
                import numpy as np

                pix, refm, _ = c4_ref
                sm, _, cn = bh.film(0)
                m = (sm / np.maximum(cn, 1)[..., None]).reshape(-1, 3)[pix]
                e = ((m - refm) ** 2 / (refm ** 2 + 1e-2)).mean(axis=1)
                res[name]["relmse_view0"] = float(f"{float(e.mean()):.5g}")
                # at 256 spp single sun-caustic samples decide the mean: also the share of its worst
                # pixel and the mean without the 4 worst of the reference pixels (0.1%)
                es = np.sort(e)
                res[name]["relmse_view0_worst_pixel_share"] = float(f"{float(es[-1] / max(e.sum(), 1e-30)):.3g}")
                res[name]["relmse_view0_without_worst_0.1pct"] = float(f"{float(es[:-4].mean()):.5g}")
            if g is not None:
                res[name]["guide_steps"] = int(g.stats()["steps"])
                g.close()
            bh.close()
            log(f"bench: C4 8-view batch {name}: {res[name]}")
        heavy["views_batch"] = dict(res, reference=None if c4_ref is None else
                                    f"{c4_ref[2]}-spp float64 CPU render, {len(c4_ref[0])} stratified pixels of view 0")
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
        "roofline_other": other_roofs,
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
