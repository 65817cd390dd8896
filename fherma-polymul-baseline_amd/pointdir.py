"""Point directories of the `solution` harness (the reference's main.cpp wire
format, main.cpp:3-11,42-47,106-133): `manifest.json` with the point (N, W, L)
and the case count, `point/q.bin` (28 little-endian u32), one
`cases/NNNNNN/{a,b}.bin` per case (N x 28 u32), and after a run
`out/NNNNNN/c.bin` plus `out/results.json` ({"init_s", "cases": [{"i",
"seconds", "status"}]}).  Reading and writing only; no arithmetic.
"""
import json
import os
import subprocess

import numpy as np

LIMBS = 28
WIDTH = 868


def _u32_bytes(x):
    return np.ascontiguousarray(np.asarray(x, dtype="<u4")).tobytes()


def write_point(path, q, n, cases, labels=None):
    """Writes a point directory: modulus q (28 u32), N = n and the input pairs
    `cases` ([(a, b)], each [n, 28] uint32).  Optional `labels` (one name per
    case) go to labels.json beside the manifest."""
    os.makedirs(os.path.join(path, "point"), exist_ok=True)
    with open(os.path.join(path, "point", "q.bin"), "wb") as f:
        f.write(_u32_bytes(q))
    for i, (a, b) in enumerate(cases):
        if np.shape(a) != (n, LIMBS) or np.shape(b) != (n, LIMBS):
            raise ValueError(f"case {i}: a and b must be [{n}, {LIMBS}]")
        d = os.path.join(path, "cases", f"{i:06d}")
        os.makedirs(d, exist_ok=True)
        with open(os.path.join(d, "a.bin"), "wb") as f:
            f.write(_u32_bytes(a))
        with open(os.path.join(d, "b.bin"), "wb") as f:
            f.write(_u32_bytes(b))
    with open(os.path.join(path, "manifest.json"), "w") as f:
        json.dump({"point": {"N": int(n), "W": WIDTH, "L": LIMBS}, "cases": len(cases)}, f)
    if labels is not None:
        with open(os.path.join(path, "labels.json"), "w") as f:
            json.dump(list(labels), f)


def read_manifest(path):
    """(N, case count) of a point directory."""
    with open(os.path.join(path, "manifest.json")) as f:
        m = json.load(f)
    return int(m["point"]["N"]), int(m["cases"])


def read_modulus(path):
    return np.fromfile(os.path.join(path, "point", "q.bin"), dtype="<u4").astype(np.uint32)


def read_case(path, i, n):
    """Input pair of case i, each [n, 28] uint32."""
    d = os.path.join(path, "cases", f"{i:06d}")
    a = np.fromfile(os.path.join(d, "a.bin"), dtype="<u4").astype(np.uint32)
    b = np.fromfile(os.path.join(d, "b.bin"), dtype="<u4").astype(np.uint32)
    return a.reshape(n, LIMBS), b.reshape(n, LIMBS)


def read_results(path):
    with open(os.path.join(path, "out", "results.json")) as f:
        return json.load(f)


def read_output(path, i, n, where="out"):
    """c of case i from out/ (or another folder laid out the same way, such as
    expected/); None when the file is absent or has the wrong size."""
    f = os.path.join(path, where, f"{i:06d}", "c.bin")
    if not os.path.exists(f) or os.path.getsize(f) != n * LIMBS * 4:
        return None
    return np.fromfile(f, dtype="<u4").astype(np.uint32).reshape(n, LIMBS)


def write_output(path, i, c, where="expected"):
    d = os.path.join(path, where, f"{i:06d}")
    os.makedirs(d, exist_ok=True)
    with open(os.path.join(d, "c.bin"), "wb") as f:
        f.write(_u32_bytes(c))


def solution_binary():
    """The drop-in harness built beside this package (make -C this folder)."""
    return os.path.join(os.path.dirname(os.path.abspath(__file__)), "solution")


def run_solution(path, binary=None, timeout=600, quiet=True, backend=None):
    """Runs `solution <path>` (fherma_init once, each fherma_run timed alone)
    and returns the parsed out/results.json; backend="quartic" sets
    PM_BACKEND for the process (the special-modulus backend)."""
    env = None
    if backend is not None:
        env = dict(os.environ, PM_BACKEND=backend)
    subprocess.run([binary or solution_binary(), str(path)], check=True, timeout=timeout,
                   stdout=subprocess.DEVNULL if quiet else None, env=env)
    return read_results(path)
