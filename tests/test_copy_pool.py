"""The host copy pool of the host-buffer path (csrc/pm_copy_pool.cpp), built
with ThreadSanitizer and run on the CPU: chunks land whole, flags carry each
call's value, and no data race between the caller and the workers."""
import os
import subprocess

import pytest

from conftest import PKG_DIR, ROOT


@pytest.mark.parametrize("workers", [0, 1, 3])
def test_copy_pool_under_tsan(tmp_path, workers):
    exe = tmp_path / "copy_pool_check"
    src = os.path.join(ROOT, "tests", "native", "copy_pool_check.cpp")
    pool = os.path.join(PKG_DIR, "csrc", "pm_copy_pool.cpp")
    placement = os.path.join(PKG_DIR, "csrc", "pm_placement.cpp")
    subprocess.run(["g++", "-std=c++17", "-O1", "-g", "-fsanitize=thread", "-I", os.path.join(PKG_DIR, "csrc"),
                    src, pool, placement, "-lpthread", "-o", str(exe)], check=True, timeout=120)
    out = subprocess.run([str(exe), str(workers), "200"], capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stdout + out.stderr
    assert "ok: 200 calls" in out.stdout
    assert "WARNING: ThreadSanitizer" not in out.stderr, out.stderr
