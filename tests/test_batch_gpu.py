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


@pytest.mark.parametrize("graph", [0, 1, 2])
@pytest.mark.parametrize("queue", [1 << 12, 1 << 16])
def test_batch_equals_single_renders(batch_setup, graph, queue):
    from spica_amd import render
    wl, scenes, singles = batch_setup
    items = [(gs, cam, spp, seed) for gs, (_, cam, spp, seed) in zip(scenes, wl.items)]
    b = render.Batch(items, max_depth=wl.max_depth, sampler=wl.sampler, queue_size=queue, block_size=64, graph=graph)
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
    print(f"graph {graph} queue {queue}: {info['passes']} passes, {info['chunks']} chunks, {info['paths']} paths")
    b.close()


def test_single_render_graph_replay_equals_launches(gpu):
    """use_graph: one captured pass replayed per pass gives the same film as plain launches,
    across runs with different sample ranges."""
    from spica_amd import render
    wl = configs.cornell_box(width=48, height=40, max_depth=8)
    gs = render.GpuScene(wl.scene)
    films = []
    for use_graph in (False, True):
        gr = render.GpuRender(gs, wl.camera, wl.render_desc(spp=64), queue_size=1 << 13, block_size=64,
                              use_graph=use_graph)
        gr.run(20)
        gr.run(33)
        films.append(gr.film())
        gr.close()
    for a, b in zip(*films):
        assert np.array_equal(a, b)
    assert films[0][2].min() == 53
    gs.close()
