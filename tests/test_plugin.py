"""spica's plugin flow on the GPU path: dlopen the "path" plugin, createInstance(params),
integrator->render(camera, scene, params) (sources/core/cobject.cc:12-66,
sources/spica/sceneparser.cc:87-105)."""
import ctypes
import os
import subprocess

import numpy as np
import pytest

from spica_amd import build, configs

PLUGIN = build.PLUGIN
DEMO = build.DEMO


def test_plugin_exports_spica_entry_points():
    from spica_amd import _abi

    _abi.load()   # binds libspk.so to torch's HIP runtime before the plugin pulls it in
    for path, text in ((PLUGIN, b"Path tracing"), (build.PLUGIN_DIRECT, b"direct lighting")):
        lib = ctypes.CDLL(path)
        lib.getDescription.restype = ctypes.c_char_p
        assert text in lib.getDescription()
        assert hasattr(lib, "createInstance")


def test_plugin_reports_errors_without_aborting(tmp_path):
    """No GPU here: the render fails with an error message and a status, not an abort."""
    import torch

    if torch.cuda.is_available():
        pytest.skip("only meaningful without a GPU")
    w = configs.cornell_box(width=16, height=16)
    sf = tmp_path / "s.bin"
    w.scene.save_bin(sf, w.camera, 4, 8, 0, 0)
    r = subprocess.run([DEMO, str(sf), str(tmp_path / "o.f64"), PLUGIN], capture_output=True, text=True, timeout=120)
    assert r.returncode == 4 and "render failed" in r.stderr


@pytest.mark.gpu
@pytest.mark.parametrize("accel", ["bvh", "lbvh"])
def test_plugin_render_equals_c_abi_render(gpu, tmp_path, accel):
    """The C++ plugin path and the Python handles over the same C ABI give the same film,
    bit for bit, with either accelerator."""
    w = configs.cornell_box(width=64, height=48)
    sf = tmp_path / "s.bin"
    w.scene.save_bin(sf, w.camera, 32, 8, 0, 0)
    out = tmp_path / "img.f64"
    r = subprocess.run([DEMO, str(sf), str(out), PLUGIN] + (["lbvh"] if accel == "lbvh" else []), capture_output=True,
                       text=True, timeout=300)
    print(r.stdout, r.stderr)
    assert r.returncode == 0
    img = np.fromfile(out).reshape(48, 64, 3)
    gs = gpu.GpuScene(w.scene, builder="sah" if accel == "bvh" else "lbvh")
    gr = gpu.GpuRender(gs, w.camera, w.render_desc(spp=32))
    gr.run(32)
    s, q, c = gr.film()
    want = s / (c[..., None] + 1e-12)
    assert np.array_equal(img, want)
    assert "rays extend" in r.stdout


@pytest.mark.gpu
def test_plugin_guiding_parameter(gpu, tmp_path):
    """"guiding" = 1 attaches the learned proposal inside the plugin's render: guided
    vertices are reported and the image agrees with the plain render in the mean."""
    w = configs.cornell_box(width=64, height=48)
    sf = tmp_path / "s.bin"
    w.scene.save_bin(sf, w.camera, 256, 8, 0, 0)
    out = tmp_path / "img.f64"
    r = subprocess.run([DEMO, str(sf), str(out), PLUGIN, "guiding"], capture_output=True, text=True, timeout=300)
    print(r.stdout, r.stderr)
    assert r.returncode == 0
    guided = int(r.stdout.split("guided")[1].split()[0])
    assert guided > 0
    img = np.fromfile(out).reshape(48, 64, 3)
    gs = gpu.GpuScene(w.scene)
    gr = gpu.GpuRender(gs, w.camera, w.render_desc(spp=256, seed=99))
    gr.run(256)
    s, q, c = gr.film()
    plain = s / (c[..., None] + 1e-12)
    assert abs(img.mean() / plain.mean() - 1) < 0.01


@pytest.mark.gpu
def test_plugin_saves_like_film_save(gpu, tmp_path):
    """"saveEvery" = 4 with outputFile "img%02d": the plugin writes img04.hdr and img08.hdr
    (Film::save(id) + ".hdr", film.cc:23-40, hdrfilm.cc:21-25), each byte-equal to the RGBE
    encoding of the render stopped at that sample count; a name with a non-integer
    conversion is refused before rendering."""
    from spica_amd import sceneio

    w = configs.cornell_box(width=40, height=30)
    sf = tmp_path / "s.bin"
    w.scene.save_bin(sf, w.camera, 8, 8, 0, 0)
    out = tmp_path / "img.f64"
    name = str(tmp_path / "img%02d")
    r = subprocess.run([DEMO, str(sf), str(out), PLUGIN, "out=" + name, "every=4"], capture_output=True, text=True,
                       timeout=300)
    print(r.stdout, r.stderr)
    assert r.returncode == 0
    for k in (4, 8):
        gs = gpu.GpuScene(w.scene)
        gr = gpu.GpuRender(gs, w.camera, w.render_desc(spp=k))
        gr.run(k)
        s, q, c = gr.film()
        gr.close()
        gs.close()
        sceneio.write_rgbe(tmp_path / f"want{k}.hdr", s / (c[..., None] + 1e-12))
        assert (tmp_path / f"img{k:02d}.hdr").read_bytes() == (tmp_path / f"want{k}.hdr").read_bytes()
    r = subprocess.run([DEMO, str(sf), str(out), PLUGIN, "out=" + str(tmp_path / "bad%s")], capture_output=True,
                       text=True, timeout=300)
    assert r.returncode == 4 and "conversion" in r.stderr


@pytest.mark.gpu
def test_directlighting_plugin_equals_c_abi_render(gpu, tmp_path):
    """spica's "directlighting" plugin (libspica_directlighting.so) renders the film the C ABI
    gives with integrator = directlighting, bit for bit."""
    w = configs.cornell_box(width=64, height=48)
    sf = tmp_path / "s.bin"
    w.scene.save_bin(sf, w.camera, 16, 8, 0, 0)
    out = tmp_path / "img.f64"
    r = subprocess.run([DEMO, str(sf), str(out), build.PLUGIN_DIRECT], capture_output=True, text=True, timeout=300)
    print(r.stdout, r.stderr)
    assert r.returncode == 0 and "direct lighting" in r.stdout
    img = np.fromfile(out).reshape(48, 64, 3)
    gs = gpu.GpuScene(w.scene)
    gr = gpu.GpuRender(gs, w.camera, w.render_desc(spp=16, integrator=1))
    gr.run(16)
    s, q, c = gr.film()
    gr.close()
    gs.close()
    assert np.array_equal(img, s / (c[..., None] + 1e-12))


@pytest.mark.gpu
def test_plugin_wide_materials_equal_c_abi_render(gpu, tmp_path):
    """Plastic, rough dielectric, rough plastic and Beckmann materials built through the
    plugin's spica-style classes give the C ABI's film bit for bit."""
    from tests.helpers import wide_materials_scene

    s, cam = wide_materials_scene()
    sf = tmp_path / "s.bin"
    s.save_bin(sf, cam, 16, 16, 0, 0)
    out = tmp_path / "img.f64"
    r = subprocess.run([DEMO, str(sf), str(out), PLUGIN], capture_output=True, text=True, timeout=300)
    print(r.stdout, r.stderr)
    assert r.returncode == 0
    img = np.fromfile(out).reshape(cam.height, cam.width, 3)
    from spica_amd import configs as cf

    wl = cf.Workload("wide", s, cam, 16, 16, 0, 0)
    gs = gpu.GpuScene(s)
    gr = gpu.GpuRender(gs, cam, wl.render_desc())
    gr.run(16)
    sm, q, c = gr.film()
    gr.close()
    gs.close()
    assert np.array_equal(img, sm / (c[..., None] + 1e-12))
