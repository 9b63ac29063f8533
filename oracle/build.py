"""Build recipes for the CPU checker (test infrastructure).

* liboracle.so        -- the C restatement (spk_oracle.c), gcc.
* _ref/libspica_ref.so -- spica's own hot-path sources compiled straight from
  /root/reference/sources with g++ (no reference build system), plus the
  extern "C" driver ref_harness.cpp.  Only files that compile from the tree as
  shipped are used; core/image.cc, core/meshio.cc and spica/sceneparser.cc need
  stb / tinyobjloader / tinyxml2, which are not vendored, and are left out.  The
  Image symbols that Film and MipMap reference stay unresolved and are never
  called, so the library is loaded with RTLD_LAZY.  Built only where
  /root/reference exists; the product, its GPU tests and the benchmark never load it.
"""
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_LIB = os.path.join(HERE, "liboracle.so")
REF_DIR = os.path.join(HERE, "_ref")
REF_LIB = os.path.join(REF_DIR, "libspica_ref.so")
REF_SRC = "/root/reference/sources"

REF_UNITS = [
    "core/triangle", "core/primitive", "core/shape", "core/ray", "core/interaction", "core/bsdf", "core/bxdf",
    "core/microfacet", "core/fresnel", "core/sampling", "core/mis", "core/random", "core/integrator", "core/parallel",
    "core/memory", "core/spectrum", "core/transform", "core/matrix4x4", "core/light", "core/camera", "core/film",
    "core/filter", "core/cobject", "core/renderparams", "core/scene", "core/visibility_tester", "core/texture",
    "core/material", "core/medium", "core/phase", "core/bssrdf", "core/mipmap", "core/interpolation", "core/float",
    "core/exception", "core/uv", "core/constant", "core/tmo", "core/birateral", "core/quaternion",
    "accelerators/bvh", "integrators/path/path", "integrators/directlighting/directlighting", "samplers/independent", "bsdfs/diffuse", "bsdfs/conductor",
    "bsdfs/dielectric", "bsdfs/roughconductor", "bsdfs/plastic", "bsdfs/roughdielectric", "bsdfs/roughplastic", "lights/area", "cameras/perspective", "filters/box", "films/hdrfilm",
]


def build_oracle():
    src = os.path.join(HERE, "spk_oracle.c")
    tmp = ORACLE_LIB + ".tmp"
    cmd = ["gcc", "-O2", "-std=c11", "-Wall", "-fPIC", "-shared", src, "-o", tmp, "-lm", "-lpthread"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("gcc failed building liboracle.so")
    os.replace(tmp, ORACLE_LIB)
    return ORACLE_LIB


def build_ref(jobs=8):
    """Returns the library path, or None when the reference tree is absent."""
    if not os.path.isdir(REF_SRC):
        return None
    obj_dir = os.path.join(REF_DIR, "obj")
    os.makedirs(obj_dir, exist_ok=True)
    flags = ["g++", "-std=c++17", "-O2", "-fPIC", "-w", "-I", REF_SRC]

    def compile_one(unit):
        src = os.path.join(REF_SRC, unit + ".cc")
        obj = os.path.join(obj_dir, unit.replace("/", "_") + ".o")
        if os.path.exists(obj) and os.path.getmtime(obj) >= os.path.getmtime(src):
            return obj
        r = subprocess.run(flags + ["-c", src, "-o", obj], capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError(f"g++ failed on {src}:\n{r.stderr[-2000:]}")
        return obj

    with ThreadPoolExecutor(max_workers=jobs) as ex:
        objs = list(ex.map(compile_one, REF_UNITS))
    harness = os.path.join(HERE, "ref_harness.cpp")
    hobj = os.path.join(obj_dir, "ref_harness.o")
    r = subprocess.run(flags + ["-c", harness, "-o", hobj], capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"g++ failed on ref_harness.cpp:\n{r.stderr[-2000:]}")
    tmp = REF_LIB + ".tmp"
    # Every spica plugin defines createInstance/getDescription; only one copy is kept.
    cmd = ["g++", "-shared", "-pthread", "-Wl,--allow-multiple-definition", "-o", tmp] + objs + [hobj, "-lpthread"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link of libspica_ref.so failed:\n{r.stderr[-2000:]}")
    os.replace(tmp, REF_LIB)
    return REF_LIB


if __name__ == "__main__":
    print(build_oracle())
    print(build_ref())
