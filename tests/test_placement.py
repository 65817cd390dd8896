"""Host placement helpers (csrc/pm_placement.cpp, role of the reference's
host_affinity.h:14-46 and host_cpu_topology.h:15-36), built with g++ and run
on the CPU: cpulist parsing, the worker order over synthetic topologies (one
CPU per physical core before any second hyperthread, the caller's core last)
and pinning the calling thread."""
import os
import subprocess

from conftest import PKG_DIR, ROOT


def test_placement_helpers(tmp_path):
    exe = tmp_path / "placement_check"
    src = os.path.join(ROOT, "tests", "native", "placement_check.cpp")
    lib = os.path.join(PKG_DIR, "csrc", "pm_placement.cpp")
    subprocess.run(["g++", "-std=c++17", "-O1", "-g", "-Wall", "-I", os.path.join(PKG_DIR, "csrc"), src, lib,
                    "-o", str(exe)], check=True, timeout=120)
    out = subprocess.run([str(exe)], capture_output=True, text=True, timeout=60)
    assert out.returncode == 0, out.stdout + out.stderr
    assert out.stdout.strip().endswith("ok")
