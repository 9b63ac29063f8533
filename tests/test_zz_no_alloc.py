"""spk_trace_noalloc (no allocation; its traversal scratch is a fixed device array) against
spk_trace on the same scene and rays: identical hit ids and t, bit for bit, on two calls in a
row, the second with fewer rays, for closest hits and occlusion, on a scene whose traversal
stacks spill past the LDS slots into the scratch."""
import numpy as np
import pytest

from spica_amd import _abi, configs
from spica_amd.scene import Scene, diffuse, unit_sphere

pytestmark = pytest.mark.gpu


def _scenes():
    deep = Scene()
    unit = unit_sphere()
    for r in range(1, 41):  # nested spheres seen from the centre: stacks spill past 16 slots
        deep.add_triangles(unit * float(r), diffuse((0.5, 0.5, 0.5)))
    return [("c1-128", configs.c1_spheres(n_spheres=128).scene, 6.0), ("nested", deep, 0.01)]


@pytest.mark.parametrize("which", [0, 1])
def test_noalloc_trace_equals_trace(gpu, which):
    import torch

    name, scene, spread = _scenes()[which]
    gs = gpu.GpuScene(scene)
    rng = np.random.default_rng(21 + which)
    for call, n in enumerate((1_500_000, 200_000)):
        print(f"{name}: call {call}, {n} rays", flush=True)
        org = torch.from_numpy((spread * rng.normal(size=(n, 3))).astype(np.float32)).cuda()
        dirv = torch.from_numpy(rng.normal(size=(n, 3)).astype(np.float32)).cuda()
        tmax = torch.from_numpy(rng.uniform(0.5, 30.0, n).astype(np.float32)).cuda()
        pa, ta = gs.trace(org, dirv)
        pb, tb = gs.trace(org, dirv, noalloc=True)
        torch.cuda.synchronize()
        assert torch.equal(pa, pb) and torch.equal(ta.view(torch.int32), tb.view(torch.int32))
        oa, _ = gs.trace(org, dirv, tmax=tmax, any_hit=True)
        ob, _ = gs.trace(org, dirv, tmax=tmax, any_hit=True, noalloc=True)
        torch.cuda.synchronize()
        assert torch.equal(oa, ob)
        print(f"{name}: call {call}: hits {(pa >= 0).sum().item()}, occluded {(oa == 1).sum().item()}", flush=True)
    gs.close()


def test_noalloc_rejects_bad_arguments(gpu):
    lib = _abi.load()
    assert lib.spk_trace_noalloc(None, None, None, None, 1, 0, None, None, None) == _abi.SPK_ERR_INVALID
    assert b"bad arguments" in lib.spk_last_error()
