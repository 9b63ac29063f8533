"""In-tree build of libspk.so for gfx950 (hipcc, no JIT cache)."""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libspk.so")

SOURCES = ["spk_kernels.hip", "spk_guide.hip", "spk_api.cpp", "spk_guide_api.cpp", "bvh_build.cpp", "lbvh.hip",
           "spk_noalloc.hip"]
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"

# Packed FP32 arithmetic is disabled for all device code: hipcc would otherwise
# emit it from plain float code.  Multiply-adds are fused only within one source
# expression (-ffp-contract=on), so a function's rounding does not depend on the kernel
# it is inlined into (batched and single renders then agree bit for bit).
FLAGS = ["-O3", f"--offload-arch={ARCH}", "-fPIC", "-shared", "-std=c++17", "-Wall", "-ffp-contract=on",
         "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]


OBJ_DIR = os.path.join(HERE, "_build")

# The stage kernels divide with a reciprocal and a multiply (within ~2 ulp of the correctly
# rounded quotient, against ~10 instructions for it) and flush single-precision denormals,
# which lets that reciprocal go without range scaling: shading and the learned proposal are
# bound by their arithmetic (the static VALU count of k_shade and k_guide drops by a third).
# The traversal's own divisions (inverse ray direction, triangle determinant) stay correctly
# rounded by going through double precision.
EXTRA_FLAGS = {"spk_kernels.hip": ["-fno-hip-fp32-correctly-rounded-divide-sqrt", "-fgpu-flush-denormals-to-zero"]}


def build_lib(verbose=False, jobs=8):
    """Compiles every source to an object (in parallel; always, so a header change is never
    missed), then links libspk.so."""
    from concurrent.futures import ThreadPoolExecutor

    os.makedirs(OBJ_DIR, exist_ok=True)
    compile_flags = [f for f in FLAGS if f != "-shared"]

    def compile_one(src):
        obj = os.path.join(OBJ_DIR, src + ".o")
        cmd = [HIPCC] + compile_flags + EXTRA_FLAGS.get(src, []) + ["-c", os.path.join(CSRC, src), "-o", obj]
        if verbose:
            print(" ".join(cmd), flush=True)
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            sys.stderr.write(r.stdout + r.stderr)
            raise RuntimeError(f"hipcc failed on {src}")
        return obj

    with ThreadPoolExecutor(max_workers=jobs) as ex:
        objs = list(ex.map(compile_one, SOURCES))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC"] + objs + ["-o", tmp]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("hipcc failed linking libspk.so")
    os.replace(tmp, LIB)
    return LIB


HOST = os.path.join(HERE, "host")
PLUGIN = os.path.join(HERE, "libspica_path.so")
PLUGIN_DIRECT = os.path.join(HERE, "libspica_directlighting.so")
DEMO = os.path.join(HOST, "plugin_demo")


def build_plugin(verbose=False):
    """The spica-compatible "path" and "directlighting" plugins (C++ over the C ABI) and
    the demo driver."""
    cmds = [
        ["g++", "-std=c++17", "-O2", "-Wall", "-fPIC", "-shared", os.path.join(HOST, "path_plugin.cpp"),
         "-L" + HERE, "-lspk", "-Wl,-rpath,$ORIGIN", "-Wl,-rpath-link,/opt/rocm/lib", "-o", PLUGIN + ".tmp"],
        ["g++", "-std=c++17", "-O2", "-Wall", "-fPIC", "-shared", "-DSPK_PLUGIN_DIRECTLIGHTING",
         os.path.join(HOST, "path_plugin.cpp"), "-L" + HERE, "-lspk", "-Wl,-rpath,$ORIGIN",
         "-Wl,-rpath-link,/opt/rocm/lib", "-o", PLUGIN_DIRECT + ".tmp"],
        ["g++", "-std=c++17", "-O2", "-Wall", os.path.join(HOST, "plugin_demo.cpp"), "-L" + HERE, "-lspica_path",
         "-lspk", "-ldl", "-Wl,-rpath,$ORIGIN/..", "-Wl,-rpath-link,/opt/rocm/lib", "-o", DEMO + ".tmp"],
    ]
    for i, cmd in enumerate(cmds):
        if verbose:
            print(" ".join(cmd), flush=True)
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            sys.stderr.write(r.stdout + r.stderr)
            raise RuntimeError("g++ failed building the path plugin")
        if i == 0:
            os.replace(PLUGIN + ".tmp", PLUGIN)
        if i == 1:
            os.replace(PLUGIN_DIRECT + ".tmp", PLUGIN_DIRECT)
    os.replace(DEMO + ".tmp", DEMO)
    return PLUGIN


if __name__ == "__main__":
    print(build_lib(verbose=True))
    print(build_plugin(verbose=True))
