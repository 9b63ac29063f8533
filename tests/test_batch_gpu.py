"""Batched rendering: every image of a batch (own scene, film size, spp and seed) must
come out bit for bit as its single-image render, with plain launches, a replayed
per-pass graph and a whole-batch graph."""
import numpy as np
import pytest

from spica_amd import configs

pytestmark = pytest.mark.gpu


def _small_batch():
    wl = configs.c2_batch(n_images=5, draw_seed=7, spheres=(4, 40), res=(24, 96), spp=(3, 40))
    # one image lit by an environment map as well (each image's scene carries its own)
    wl.items[1][0].set_envmap(configs.sky_envmap(64, 32, sun=200.0),
                              to_world=np.array([[1, 0, 0, 0], [0, 0, 1, 0], [0, 1, 0, 0], [0, 0, 0, 1]], float),
                              radius=40.0)
    return wl


@pytest.fixture(scope="module")
def batch_setup(gpu):
    from spica_amd import render
    wl = _small_batch()
    scenes = [render.GpuScene(sc) for sc, _, _, _ in wl.items]
    singles = []
    for gs, (_, cam, spp, seed) in zip(scenes, wl.items):
        desc = configs.Workload("x", None, cam, spp, wl.max_depth, wl.sampler, seed).render_desc()
        gr = render.GpuRender(gs, cam, desc, queue_size=1 << 16, block_size=64)
        gr.run(spp)
        singles.append(gr.film())
        gr.close()
    yield wl, scenes, singles
    for gs in scenes:
        gs.close()


@pytest.mark.parametrize("regen", [True, False])
@pytest.mark.parametrize("graph", [0, 1, 2])
@pytest.mark.parametrize("queue", [1 << 12, 1 << 16])
def test_batch_equals_single_renders(batch_setup, graph, queue, regen):
    """Also with regenerating passes (the queue keeps `queue` paths in flight and refills) and
    classic ones (every path of a pass enters at once)."""
    from spica_amd import render
    wl, scenes, singles = batch_setup
    items = [(gs, cam, spp, seed) for gs, (_, cam, spp, seed) in zip(scenes, wl.items)]
    b = render.Batch(items, max_depth=wl.max_depth, sampler=wl.sampler, queue_size=queue, block_size=64, graph=graph,
                     regen=regen)
    info = b.info()
    assert info["paths"] == wl.paths()
    for rep in range(2):  # a run starts from cleared films
        b.run()
        for i, (s_ref, q_ref, c_ref) in enumerate(singles):
            s, q, c = b.film(i)
            assert np.array_equal(c, c_ref), (graph, queue, i)
            assert np.array_equal(s, s_ref), (graph, queue, i, np.abs(s - s_ref).max())
            assert np.array_equal(q, q_ref), (graph, queue, i)
    st = b.stats()
    assert st["paths"] == 2 * wl.paths()
    print(f"graph {graph} queue {queue} regen {regen}: {info['passes']} passes, {info['chunks']} chunks, "
          f"{info['paths']} paths")
    b.close()


def test_single_render_graph_replay_equals_launches(gpu):
    """use_graph: one captured pass replayed per pass gives the same film as plain launches,
    across runs with different sample ranges."""
    from spica_amd import render
    wl = configs.cornell_box(width=48, height=40, max_depth=8)
    gs = render.GpuScene(wl.scene)
    films = []
    for use_graph, regen in ((False, False), (True, False), (False, True), (True, True)):
        gr = render.GpuRender(gs, wl.camera, wl.render_desc(spp=64), queue_size=1 << 13, block_size=64,
                              use_graph=use_graph, regen=regen)
        gr.run(20)
        gr.run(33)
        films.append(gr.film())
        gr.close()
    for f in films[1:]:
        for a, b in zip(films[0], f):
            assert np.array_equal(a, b)
    assert films[0][2].min() == 53
    gs.close()


def _views(n=3, w=96, h=64):
    """n views of the occluded-glossy scene (light in a box under the ceiling) around its centre."""
    import math

    from spica_amd.scene import Camera, look_at

    wl = configs.c3_occluded_glossy(width=w, height=h)
    cams = []
    for i in range(n):
        a = math.radians(-90.0 + 20.0 * (i - (n - 1) / 2.0))
        cams.append(Camera(w, h, look_at((18.0 * math.cos(a), 2.0, 18.0 * math.sin(a)), (0.0, -1.0, 0.0), (0, 1, 0)),
                           40.0))
    return wl, cams


def test_guided_batch_is_unbiased_and_deterministic(gpu):
    """Three views of one scene in a batch with one guide (trained on the first 32 samples of
    every view, in rounds across the views): every view agrees with the plain batch per pixel
    (at most 0.5% of pixel-channels beyond 3 sigma, none beyond 6), the guide took its Adam steps,
    and a second guided run gives the same films bit for bit."""
    from spica_amd import render
    from tests.helpers import zstats

    wl, cams = _views()
    gs = render.GpuScene(wl.scene)
    spp = 512
    items = [(gs, cam, spp, 7 + i) for i, cam in enumerate(cams)]
    plain = render.Batch(items, max_depth=16, queue_size=1 << 20, block_size=64, graph=1)
    plain.run()
    films_plain = [plain.film(i) for i in range(len(cams))]
    plain.close()
    runs = []
    for _ in range(2):
        b = render.Batch(items, max_depth=16, queue_size=1 << 20, block_size=64, graph=1)
        g = render.Guide(batch_size=1 << 14, train_spp=32, train_pass_spp=4, steps_per_pass=4)
        b.set_guide(g)
        b.run()
        st = b.stats()
        assert st["guided_vertices"] > 0 and st["train_passes"] == 8
        assert g.stats()["steps"] == 32
        runs.append([b.film(i) for i in range(len(cams))])
        b.close()
        g.close()
    for i in range(len(cams)):
        (sg, qg, cg), (sp, qp, cp) = runs[0][i], films_plain[i]
        z, mg, mp = zstats(sg.reshape(-1, 3), qg.reshape(-1, 3), cg.reshape(-1, 1), sp.reshape(-1, 3),
                           qp.reshape(-1, 3), cp.reshape(-1, 1))
        a = np.abs(z)
        print(f"view {i}: guided vs plain |z|>3 {np.mean(a > 3):.4%}, max |z| {a.max():.2f}, means "
              f"{mg.mean():.5f} {mp.mean():.5f}", flush=True)
        assert np.mean(a > 3) <= 0.005 and a.max() < 6.0
        for x, y in zip(runs[0][i], runs[1][i]):
            assert np.array_equal(x, y)
    gs.close()


def test_guided_batch_needs_one_scene(gpu):
    from spica_amd import render

    wl = _small_batch()
    scenes = [render.GpuScene(sc) for sc, _, _, _ in wl.items[:2]]
    b = render.Batch([(g, cam, spp, seed) for g, (_, cam, spp, seed) in zip(scenes, wl.items[:2])], max_depth=8,
                     queue_size=1 << 16, block_size=64)
    g = render.Guide()
    with pytest.raises(Exception, match="one scene"):
        b.set_guide(g)
    g.close()
    b.close()
    for s in scenes:
        s.close()
