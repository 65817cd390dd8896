"""pm_place_near_device on the GPU: the calling thread ends up restricted to
CPUs of the device's NUMA node that the process may use (reference
host_affinity.h:14-46), or, when the topology is unknown, is left as it was.
Runs in a child process: the call changes the caller's affinity."""
import json
import os
import subprocess
import sys

import pytest

from conftest import ROOT

pytestmark = pytest.mark.gpu

CHILD = r"""
import json, os, sys
sys.path.insert(0, {tests!r})
from conftest import load_package
pm = load_package()
before = sorted(os.sched_getaffinity(0))
status, node, cpus = pm.place_near_device(0)
print(json.dumps({{"status": status, "node": node, "cpus": cpus, "before": before,
                   "after": sorted(os.sched_getaffinity(0))}}))
"""


def node_cpus(node):
    text = open(f"/sys/devices/system/node/node{node}/cpulist").read().strip()
    out = set()
    for part in text.split(","):
        if "-" in part:
            a, b = part.split("-")
            out.update(range(int(a), int(b) + 1))
        elif part:
            out.add(int(part))
    return out


def test_place_near_device():
    out = subprocess.run([sys.executable, "-c", CHILD.format(tests=os.path.join(ROOT, "tests"))], capture_output=True, text=True,
                         timeout=120)
    assert out.returncode == 0, out.stderr
    r = json.loads(out.stdout.strip().splitlines()[-1])
    print("placement:", r, flush=True)
    if r["status"] == 0:
        after = set(r["after"])
        assert after and after <= set(r["before"]), r
        assert after <= node_cpus(r["node"]), r
        assert len(after) == r["cpus"], r
    else:
        assert r["status"] == 2, r  # PM_EUNSUPPORTED: topology unknown, placement unchanged
        assert r["after"] == r["before"], r
