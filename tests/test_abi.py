"""The C ABI library: loads without a GPU, exports every entry point include/spk.h
declares, and rejects malformed descriptors with an error status (no abort)."""
import ctypes
import os
import re

import numpy as np

from spica_amd import _abi, configs
from tests.conftest import ROOT


def _declared():
    names = set()
    for h in ("spk.h",):
        txt = open(os.path.join(ROOT, "include", h)).read()
        txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
        for m in re.finditer(r"^\s*(?:const\s+)?\w+\*?\s+\*?(spk_\w+)\s*\(", txt, flags=re.M):
            names.add(m.group(1))
    return names


def test_exports_every_declared_symbol():
    lib = _abi.load()
    declared = _declared()
    assert "spk_trace" in declared and "spk_render_run" in declared
    for name in declared:
        assert hasattr(lib, name), name
    assert set(_abi.SIGNATURES) == declared


def test_version_and_error_string():
    lib = _abi.load()
    assert b"gfx950" in lib.spk_version()
    assert lib.spk_last_error() is not None


def test_bad_descriptors_are_rejected():
    lib = _abi.load()
    w = configs.cornell_box()
    d = w.scene.desc()
    bad = np.full(w.scene.num_tris, 99, dtype=np.int32)
    d.mat_ids = bad.ctypes.data_as(_abi.c_i32_p)
    h = ctypes.c_void_p()
    rc = lib.spk_scene_create(ctypes.byref(d), ctypes.byref(h))
    assert rc == _abi.SPK_ERR_INVALID and b"material" in lib.spk_last_error()
    d = w.scene.desc()
    lid = w.scene.freeze()["light_ids"].copy()
    lid[0] = 0  # triangle 0 claims light 0, which names another triangle
    d.light_ids = lid.ctypes.data_as(_abi.c_i32_p)
    rc = lib.spk_scene_create(ctypes.byref(d), ctypes.byref(h))
    assert rc == _abi.SPK_ERR_INVALID and b"light" in lib.spk_last_error()
    s = configs.cornell_box().scene
    proj = np.eye(4)
    proj[3, 0] = 0.5  # a projective environment transform
    s.set_envmap(np.ones((4, 8, 3), np.float32), to_world=proj)
    d = s.desc()
    rc = lib.spk_scene_create(ctypes.byref(d), ctypes.byref(h))
    assert rc == _abi.SPK_ERR_INVALID and b"affine" in lib.spk_last_error()
    assert lib.spk_scene_create(None, ctypes.byref(h)) == _abi.SPK_ERR_INVALID
    assert lib.spk_render_run(None, 1, 0) == _abi.SPK_ERR_INVALID
    assert lib.spk_sampler_u32(7, 0, 0, None, None, None, None, 0) == _abi.SPK_ERR_INVALID


def test_scene_build_options_are_checked():
    lib = _abi.load()
    bo = _abi.SpkSceneBuildOpts()
    assert lib.spk_scene_build_opts_default(ctypes.byref(bo)) == _abi.SPK_OK
    assert (bo.builder, bo.max_leaf, bo.treelet_passes) == (_abi.SPK_BVH_SAH, 4, 6)
    d = configs.cornell_box().scene.desc()
    h = ctypes.c_void_p()
    for builder, leaf, tp in ((2, 4, 6), (-1, 4, 6), (_abi.SPK_BVH_LBVH, 0, 6), (_abi.SPK_BVH_LBVH, 9, 6),
                              (_abi.SPK_BVH_SAH, 100, 6), (_abi.SPK_BVH_LBVH, 4, -1), (_abi.SPK_BVH_LBVH, 4, 17)):
        bo.builder, bo.max_leaf, bo.treelet_passes = builder, leaf, tp
        rc = lib.spk_scene_create_ex(ctypes.byref(d), ctypes.byref(bo), ctypes.byref(h))
        assert rc == _abi.SPK_ERR_INVALID and b"build options" in lib.spk_last_error()
    assert lib.spk_scene_build_opts_default(None) == _abi.SPK_ERR_INVALID


def _c_name(cls):
    return re.sub(r"(?<!^)([A-Z])", r"_\1", cls.__name__).lower()


def test_struct_layouts_match_the_header(tmp_path):
    """Every ctypes mirror in the bindings has the size and field offsets of its C struct in
    include/ (compiled with the host compiler), so no option or statistic is read from the
    wrong bytes."""
    import shutil
    import subprocess

    import pytest

    cc = shutil.which("gcc") or shutil.which("cc")
    if cc is None:
        pytest.skip("no host C compiler")
    structs = [v for v in vars(_abi).values()
               if isinstance(v, type) and issubclass(v, ctypes.Structure) and v.__name__.startswith("Spk")]
    assert len(structs) >= 13
    lines = ['#include <stdio.h>', '#include <stddef.h>', '#include "spk.h"', "int main(void) {"]
    for cls in structs:
        c = _c_name(cls)
        lines.append(f'    printf("{c} %zu\\n", sizeof({c}));')
        for f in cls._fields_:
            lines.append(f'    printf("{c}.{f[0]} %zu\\n", offsetof({c}, {f[0]}));')
    lines += ["    return 0;", "}"]
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines) + "\n")
    exe = tmp_path / "layout"
    subprocess.run([cc, "-std=c99", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)], check=True)
    got = dict(line.split() for line in subprocess.run([str(exe)], check=True, capture_output=True,
                                                       text=True).stdout.splitlines())
    for cls in structs:
        c = _c_name(cls)
        assert int(got[c]) == ctypes.sizeof(cls), c
        for f in cls._fields_:
            assert int(got[f"{c}.{f[0]}"]) == getattr(cls, f[0]).offset, f"{c}.{f[0]}"


def test_guide_option_defaults():
    from spica_amd import render

    o = render.guide_opts()
    assert (o.batch_size, o.train_spp, o.train_pass_spp, o.steps_per_pass) == (1 << 18, 64, 8, 8)
    assert abs(o.learning_rate - 1e-2) < 1e-9 and o.guide_fraction == 0.5 and o.glossy_min_alpha == 0.5
    assert abs(o.beta1 - 0.95) < 1e-7 and abs(o.beta2 - 0.95) < 1e-7 and o.guided_depth == 3
    try:
        render.guide_opts(glossy_min_alpa=0.0)
    except TypeError as e:
        assert "glossy_min_alpa" in str(e)
    else:
        raise AssertionError("a misspelt guide option was accepted")


def test_guide_options_are_checked():
    """Malformed guide options are rejected before any device is touched."""
    from spica_amd import render

    lib = _abi.load()
    h = ctypes.c_void_p()
    bad = [dict(batch_size=100), dict(batch_size=0), dict(train_pass_spp=0), dict(learning_rate=0.0),
           dict(guide_fraction=1.0), dict(beta2=1.0), dict(beta1=-0.1), dict(glossy_min_alpha=-0.5),
           dict(glossy_min_alpha=float("nan")), dict(guided_depth=0), dict(steps_per_pass=-1)]
    for kw in bad:
        o = render.guide_opts(**kw)
        rc = lib.spk_guide_create(0, ctypes.byref(o), ctypes.byref(h))
        msg = lib.spk_last_error()
        assert rc == _abi.SPK_ERR_INVALID and (b"bad options" in msg or b"batch_size" in msg), (kw, rc, msg)
        assert not h.value
