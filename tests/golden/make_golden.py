"""Writes tests/golden/golden.json from GMP Kronecker products.

Runs in the build container only (GMP is under /opt/conda there). For each
case it builds gen_golden.c against oracle/oracle.c (for the seeded input
streams only), computes c by Kronecker substitution with GMP, and records
SHA-256 digests of q, a, b and c; for N <= 64 the full c words as hex too.
The digests are the fixtures: tests rebuild the same inputs and check that
the oracle and the GPU path reproduce c bit for bit.
"""
import hashlib
import json
import os
import subprocess
import sys
import tempfile
from pathlib import Path

HERE = Path(__file__).resolve().parent
ROOT = HERE.parents[1]
GMP = "/opt/conda"

FAMILIES = ["dense-random-1", "dense-random-2", "zero", "identity", "wrap-sign",
            "all-q-minus-one", "alternating", "limb-boundaries", "random-after-edges"]
# (n, modulus spec, families, input seed)
PLAN = [
    (32, "seed:1", FAMILIES, 7),
    (32, "delta:11993087", FAMILIES, 7),
    (64, "seed:2", ["dense-random-1"], 9),
    (1024, "seed:3", FAMILIES, 11),
    (4096, "seed:4", ["dense-random-1", "all-q-minus-one", "limb-boundaries"], 13),
    (8192, "seed:5", ["dense-random-1"], 15),
    (16384, "seed:6", ["dense-random-1"], 17),
    (32768, "seed:1", FAMILIES, 7),
    (32768, "delta:11993087", ["dense-random-1", "all-q-minus-one", "limb-boundaries"], 7),
    (65536, "seed:8", ["dense-random-1", "all-q-minus-one"], 19),
    (131072, "seed:9", ["dense-random-1", "all-q-minus-one"], 21),
]


def sha(path):
    return hashlib.sha256(Path(path).read_bytes()).hexdigest()


def main():
    with tempfile.TemporaryDirectory() as tmp:
        exe = Path(tmp) / "gen_golden"
        subprocess.run(["gcc", "-O2", "-std=c99", "-D_POSIX_C_SOURCE=200809L", "-pthread", f"-I{ROOT/'oracle'}", f"-I{GMP}/include",
                        str(HERE / "gen_golden.c"), str(ROOT / "oracle" / "oracle.c"),
                        f"-L{GMP}/lib", "-lgmp", f"-Wl,-rpath,{GMP}/lib", "-o", str(exe)],
                       check=True)
        cases = []
        for n, modulus, families, seed in PLAN:
            for family in families:
                out = Path(tmp) / "case"
                out.mkdir(exist_ok=True)
                subprocess.run([str(exe), str(out), str(n), modulus, family, str(seed)], check=True)
                record = {"n": n, "modulus": modulus, "family": family, "seed": seed,
                          "sha256": {k: sha(out / f"{k}.bin") for k in "qabc"}}
                if n <= 64:
                    record["c_hex"] = (out / "c.bin").read_bytes().hex()
                cases.append(record)
                print(n, modulus, family, record["sha256"]["c"][:16], flush=True)
    doc = {"generator": "tests/golden/gen_golden.c (GMP Kronecker substitution)",
           "layout": "little-endian uint32, 28 limbs per coefficient, row-major",
           "cases": cases}
    (HERE / "golden.json").write_text(json.dumps(doc, indent=1) + "\n")


if __name__ == "__main__":
    sys.exit(main())
