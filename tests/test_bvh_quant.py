"""Host check of the 64-byte BVH node quantiser: every child box decoded exactly as the
traversal decodes it (fmaf(q, scale, origin)) contains the child's float box, on real trees
(Cornell box, a sphere field) and on hand-made flat, nearly flat, far-off and huge boxes."""
import os
import subprocess

import numpy as np
import pytest

from spica_amd import configs

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = os.path.join(ROOT, "tests", "cpp", "quant_check.cpp")
BVH = os.path.join(ROOT, "spica_amd", "csrc", "bvh_build.cpp")


@pytest.fixture(scope="module")
def quant_check(tmp_path_factory):
    exe = str(tmp_path_factory.mktemp("qc") / "quant_check")
    cmd = ["/opt/rocm/bin/hipcc", "-O2", "-std=c++17", "-D__HIP_PLATFORM_AMD__", "-I/opt/rocm/include", SRC, BVH,
           "-o", exe, "-lpthread"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-2000:]
    return exe


def _run(exe, arg):
    r = subprocess.run([exe, arg], capture_output=True, text=True, timeout=120)
    print(r.stdout.strip())
    return r


@pytest.mark.parametrize("scene", ["cornell", "spheres"])
def test_quantised_boxes_contain_children(quant_check, tmp_path, scene):
    wl = configs.cornell_box() if scene == "cornell" else configs.c1_spheres(n_spheres=96)
    d = wl.scene.desc()
    v = np.ctypeslib.as_array(d.verts, shape=(d.num_tris * 9,)).astype(np.float32)
    path = tmp_path / "tris.bin"
    path.write_bytes(np.int64(d.num_tris).tobytes() + v.tobytes())
    r = _run(quant_check, str(path))
    assert r.returncode == 0 and "violations 0" in r.stdout, r.stdout + r.stderr
    infl = float(r.stdout.split()[-1])
    assert 1.0 <= infl < 2.0


def test_quantised_edge_boxes(quant_check):
    r = _run(quant_check, "--edge")
    assert r.returncode == 0 and "violations 0" in r.stdout, r.stdout + r.stderr
