"""The device tree build's algorithm on the host: tests/native/build_levels.cpp runs gbuild.hip's level plan
serially with its per-item and per-node steps (gbuild_core.h) and compares every tree with the host build's bytes
(wide.cpp), so the plan is checked without a GPU (test_gpu_build.py checks the kernels)."""
import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "..", "threert_amd", "csrc")


def test_the_level_plan_builds_the_host_trees(tmp_path):
    exe = str(tmp_path / "build_levels")
    cmd = ["g++", "-O2", "-std=c++17", "-ffp-contract=off", "-I", CSRC, "-w",
           os.path.join(HERE, "native", "build_levels.cpp"), os.path.join(CSRC, "wide.cpp"), "-o", exe]
    subprocess.run(cmd, check=True, timeout=300)
    r = subprocess.run([exe], capture_output=True, text=True, timeout=300)
    print(r.stdout)
    assert r.returncode == 0 and r.stdout.strip().endswith("ALL EQUAL"), r.stdout + r.stderr
    assert r.stdout.count("equal") >= 13
