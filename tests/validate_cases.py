"""Local validation of the `solution` drop-in over point directories (role of
the reference's tests/validate.py:17-140): make the validator's nine case
families for one N, write their exact products beside them, run a harness
binary over the directory and compare its outputs byte for byte.

    python tests/validate_cases.py --n 1024 [--delta D] [--binary PATH] [--dir DIR]
                                   [--existing] [--verify-only]

Moduli follow the reference validator: the largest prime q = 2^868 - 2N*j + 1
(so 2N | q - 1), or q = 2^868 - delta with --delta.  Inputs follow its streams:
Python's random.Random(20260916 + N), a and b drawn first, then the second
dense pair (tests/validate.py:46-59), so a directory made here holds the same
cases as the reference's for the same N.  Expected products come from the
oracle's RNS port (pinned against the GMP Kronecker fixtures), and for N <= 32
also from schoolbook, as the reference checks (tests/validate.py:72-73).  The
report goes to validation.json in the directory.  Test tooling: it uses the
oracle, so it lives under tests/.
"""
import argparse
import hashlib
import json
import os
import random
import statistics
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
sys.path.insert(0, HERE)

from conftest import load_package  # noqa: E402

WIDTH, LIMBS = 868, 28
FAMILIES = ["dense-random-1", "dense-random-2", "zero", "identity", "wrap-sign", "all-q-minus-one",
            "alternating", "limb-boundaries", "random-after-edges"]
_SMALL_PRIMES = [p for p in range(3, 1000) if all(p % d for d in range(2, int(p ** 0.5) + 1))]


def probable_prime(n, rounds=40):
    """Miller-Rabin with the first `rounds` odd primes as bases (gmpy2's
    is_prime(q, 40) in the reference validator)."""
    if n < 2:
        return False
    for p in _SMALL_PRIMES:
        if n % p == 0:
            return n == p
    d, s = n - 1, 0
    while d % 2 == 0:
        d //= 2
        s += 1
    for a in _SMALL_PRIMES[:rounds]:
        x = pow(a, d, n)
        if x in (1, n - 1):
            continue
        for _ in range(s - 1):
            x = x * x % n
            if x == n - 1:
                break
        else:
            return False
    return True


def validator_modulus(n, delta=None):
    """tests/validate.py:19-23: the largest probable prime 2^868 - 2N*j + 1, or
    2^868 - delta."""
    if delta is not None:
        return (1 << WIDTH) - delta
    q = (1 << WIDTH) - 2 * n + 1
    while not probable_prime(q):
        q -= 2 * n
    return q


def family_inputs(n, q):
    """The nine (name, a, b) families of tests/validate.py:46-59 as Python
    integers, drawn from the same stream."""
    rng = random.Random(20260916 + n)
    a = [rng.randrange(q) for _ in range(n)]
    b = [rng.randrange(q) for _ in range(n)]
    a2 = [rng.randrange(q) for _ in range(n)]
    b2 = [rng.randrange(q) for _ in range(n)]
    edge = [0, 1, q - 1, q - 2, 1 << 864, (1 << 864) - 1, 1 << 32, (1 << 32) - 1]
    one = [1] + [0] * (n - 1)
    wrap_b = ([0, q - 1] + [0] * (n - 2)) if n > 1 else [0]
    return [
        ("dense-random-1", a, b),
        ("dense-random-2", a2, b2),
        ("zero", a, [0] * n),
        ("identity", a, one),
        ("wrap-sign", [0] * (n - 1) + [q - 1], wrap_b),
        ("all-q-minus-one", [q - 1] * n, [q - 1] * n),
        ("alternating", [q - 1 if i % 2 else 0 for i in range(n)], a),
        ("limb-boundaries", [edge[i % len(edge)] for i in range(n)], b),
        ("random-after-edges", b, a),
    ]


def to_limbs(values):
    raw = b"".join(int(v).to_bytes(4 * LIMBS, "little") for v in values)
    return np.frombuffer(raw, dtype="<u4").astype(np.uint32).reshape(len(values), LIMBS)


def default_dir(n, delta):
    return os.path.join(ROOT, "local", f"n{n}-w{WIDTH}" + (f"-c{delta}" if delta is not None else ""))


def make(path, n, delta=None, log=print):
    """Writes the point directory, the expected products (expected/NNNNNN/c.bin)
    and labels.json; returns q as an integer."""
    import oracle_lib
    pm = load_package()
    orc = oracle_lib.Oracle()
    q = validator_modulus(n, delta)
    q_limbs = to_limbs([q])[0]
    primes = orc.primes_general(n, q_limbs)
    cases, labels = [], []
    for name, a, b in family_inputs(n, q):
        t0 = time.monotonic()
        A, B = to_limbs(a), to_limbs(b)
        c = orc.rns_multiply(q_limbs, A, B, primes)
        if n <= 32 and not np.array_equal(c, orc.schoolbook(q_limbs, A, B)):
            raise AssertionError(f"{name}: oracle and schoolbook disagree")
        pm.pointdir.write_output(path, len(cases), c, where="expected")
        cases.append((A, B))
        labels.append(name)
        log(f"make N={n} {name}: {time.monotonic() - t0:.3f}s")
    pm.pointdir.write_point(path, q_limbs, n, cases, labels=labels)
    return q


def verify(path, binary=None):
    """Compares out/ with expected/ case by case; writes validation.json and
    returns the report."""
    pm = load_package()
    n, count = pm.pointdir.read_manifest(path)
    labels = json.load(open(os.path.join(path, "labels.json")))
    results = pm.pointdir.read_results(path)
    q = int.from_bytes(pm.pointdir.read_modulus(path).astype("<u4").tobytes(), "little")
    verdicts = []
    for i in range(count):
        got = pm.pointdir.read_output(path, i, n)
        want = pm.pointdir.read_output(path, i, n, where="expected")
        result = results["cases"][i] if i < len(results["cases"]) else {"status": "missing"}
        passed = result.get("status") == "ok" and got is not None and want is not None and np.array_equal(got, want)
        verdicts.append({"case": labels[i] if i < len(labels) else str(i), "passed": bool(passed),
                         "seconds": result.get("seconds")})
    times = [v["seconds"] for v in verdicts if v["seconds"] is not None]
    report = {
        "kind": "local-oracle",
        "point": {"N": n, "W": WIDTH, "L": LIMBS},
        "q_delta": (1 << WIDTH) - q,
        "passed": all(v["passed"] for v in verdicts) and len(verdicts) == count,
        "verdicts": verdicts,
        "binary": binary,
        "binary_sha256": hashlib.sha256(open(binary, "rb").read()).hexdigest() if binary else None,
        "median_seconds": statistics.median(times) if times else None,
        "init_seconds": results.get("init_s"),
    }
    with open(os.path.join(path, "validation.json"), "w") as f:
        json.dump(report, f, indent=2)
    return report


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("--n", type=int, default=32)
    ap.add_argument("--delta", type=int, help="q = 2^868 - delta (0 < delta < 2^28)")
    ap.add_argument("--dir", help="point directory (default: local/n<N>-w868[-c<delta>])")
    ap.add_argument("--binary", help="harness to run over the directory (default: the package's solution)")
    ap.add_argument("--existing", action="store_true", help="reuse the cases already in the directory")
    ap.add_argument("--backend", choices=("general", "quartic"),
                    help="product backend the harness's fherma_init selects (PM_BACKEND)")
    ap.add_argument("--verify-only", action="store_true")
    args = ap.parse_args(argv)
    if args.delta is not None and not 0 < args.delta < (1 << 28):
        ap.error("delta must be in (0, 2^28)")
    if args.n < 2 or args.n & (args.n - 1):
        ap.error("N must be a power of two >= 2")
    path = args.dir or default_dir(args.n, args.delta)
    pm = load_package()
    binary = args.binary or pm.pointdir.solution_binary()
    if not args.verify_only:
        if not args.existing:
            make(path, args.n, args.delta)
        pm.pointdir.run_solution(path, binary=binary, backend=args.backend)
    report = verify(path, binary)
    for v in report["verdicts"]:
        print(v["case"], "PASS" if v["passed"] else "FAIL", v["seconds"])
    print(json.dumps({k: v for k, v in report.items() if k != "verdicts"}, indent=2))
    return 0 if report["passed"] else 1


if __name__ == "__main__":
    sys.exit(main())
