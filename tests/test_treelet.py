"""Host check of the LBVH treelet restructuring's topology search (spica_amd/csrc/spk_treelet.h):
its dynamic programme over leaf subsets finds the cheapest of every binary tree over 2..7 leaves
(exhaustive enumeration on random treelets), and the tree rebuilt from its recorded splits
costs that much."""
import os
import subprocess

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = os.path.join(ROOT, "tests", "cpp", "treelet_check.cpp")


def test_treelet_topology_is_optimal(tmp_path):
    exe = str(tmp_path / "treelet_check")
    cmd = ["/opt/rocm/bin/hipcc", "-O2", "-std=c++17", "-D__HIP_PLATFORM_AMD__", "-I/opt/rocm/include", SRC, "-o", exe]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-2000:]
    r = subprocess.run([exe, "600"], capture_output=True, text=True, timeout=120)
    print(r.stdout.strip())
    assert r.returncode == 0 and "mismatches 0" in r.stdout, r.stdout + r.stderr
