"""Pins the CPU oracle (oracle/oracle.c) before anything is compared with it.

Each check is chosen so that a plausible mistake in the oracle (a dropped
term, a wrong sign, a transposed operand, a wrong root order) fails it:
  * schoolbook vs Python big integers (independent arithmetic);
  * RNS restatement vs schoolbook on every validator family;
  * both vs the GMP Kronecker fixtures of tests/golden (independent algorithm);
  * the reference's prime rule vs its published fixture
    (results/rns-arithmetic.json of the reference);
  * NTTs vs direct evaluation at odd powers of psi, inverse(forward) = N*x and
    the convolution theorem;
  * CRT vs the exact signed convolution reduced mod q.
"""
import hashlib
import json
import os

import numpy as np
import pytest

from conftest import ROOT

GOLDEN = json.load(open(os.path.join(ROOT, "tests", "golden", "golden.json")))


def to_int(w):
    return int.from_bytes(np.ascontiguousarray(w, dtype="<u4").tobytes(), "little")


def py_negacyclic(a, b, q):
    n = len(a)
    A = [to_int(x) for x in a]
    B = [to_int(x) for x in b]
    c = [0] * n
    for i in range(n):
        for j in range(n):
            if i + j < n:
                c[i + j] += A[i] * B[j]
            else:
                c[i + j - n] -= A[i] * B[j]
    return c  # signed, not reduced


def sha(x):
    return hashlib.sha256(np.ascontiguousarray(x, dtype="<u4").tobytes()).hexdigest()


def modulus_of(spec, workload):
    kind, value = spec.split(":")
    return workload.modulus(int(value)) if kind == "seed" else workload.special_modulus(int(value))


def test_generators_agree(orc, workload):
    for seed in (1, 2, 99):
        q = orc.modulus(seed)
        assert (q == workload.modulus(seed)).all()
        assert to_int(q).bit_length() == 868 and to_int(q) % 2 == 1
        assert (orc.poly(seed, 3, 500, q) == workload.poly(seed, 3, 500, q)).all()
    assert to_int(orc.special_modulus(11993087)) == (1 << 868) - 11993087
    assert (orc.special_modulus(11993087) == workload.special_modulus(11993087)).all()


@pytest.mark.parametrize("family", ["dense-random-1", "all-q-minus-one", "limb-boundaries", "wrap-sign"])
def test_schoolbook_vs_python(orc, workload, family):
    q = workload.modulus(3)
    a, b = workload.family(family, 16, q, 5)
    Q = to_int(q)
    want = [x % Q for x in py_negacyclic(a, b, q)]
    got = orc.schoolbook(q, a, b)
    assert [to_int(x) for x in got] == want


@pytest.mark.parametrize("modulus", ["seed:1", "delta:11993087"])
def test_rns_restatement_vs_schoolbook(orc, workload, modulus):
    q = modulus_of(modulus, workload)
    ref = orc.primes_reference(57)
    gen = orc.primes_general(1 << 17, q)
    for family in workload.FAMILIES:
        a, b = workload.family(family, 64, q, 11)
        want = orc.schoolbook(q, a, b)
        assert (orc.rns_multiply(q, a, b, ref) == want).all(), family
        assert (orc.rns_multiply(q, a, b, gen) == want).all(), family


def test_reference_prime_fixture(orc):
    # results/rns-arithmetic.json of the reference: prime_count 57,
    # prime_min 2099838977, prime_max 2147352577, product_bits 1767
    primes = orc.primes_reference(57)
    assert len(primes) == 57
    assert primes.min() == 2099838977 and primes.max() == 2147352577
    assert orc.product_bits(primes) == 1767
    assert all(p % 65536 == 1 for p in primes)


def test_golden_inputs_and_small_outputs(orc, workload):
    for case in GOLDEN["cases"]:
        if case["n"] > 64:
            continue
        q = modulus_of(case["modulus"], workload)
        a, b = workload.family(case["family"], case["n"], q, case["seed"])
        assert sha(q) == case["sha256"]["q"]
        assert sha(a) == case["sha256"]["a"], case["family"]
        assert sha(b) == case["sha256"]["b"], case["family"]
        c = orc.schoolbook(q, a, b)
        assert c.astype("<u4").tobytes().hex() == case["c_hex"], (case["n"], case["family"])


@pytest.mark.parametrize("n", [1024, 4096, 8192, 16384, 32768, 65536, 131072])
def test_golden_rns_oracle(orc, workload, n):
    """The RNS restatement reproduces GMP's Kronecker products bit for bit."""
    cases = [c for c in GOLDEN["cases"] if c["n"] == n]
    assert cases
    if n == 32768:
        cases = cases[:4] + [c for c in cases if c["modulus"].startswith("delta")][:1]
    for case in cases:
        q = modulus_of(case["modulus"], workload)
        a, b = workload.family(case["family"], n, q, case["seed"])
        assert sha(a) == case["sha256"]["a"] and sha(b) == case["sha256"]["b"]
        c = orc.rns_multiply(q, a, b, orc.reference_primes_for(n, q))
        assert sha(c) == case["sha256"]["c"], (n, case["family"])


@pytest.mark.parametrize("n,threads", [(2, 5), (64, 3), (4096, 8), (32768, 16)])
def test_threaded_rns_oracle(orc, workload, n, threads):
    """The oracle product split over threads (per-prime units, then CRT
    blocks; more threads than coefficients at N = 2) equals the one-thread
    product word for word, equals schoolbook at N = 64 and the GMP fixture at
    N = 2^15."""
    q = workload.modulus(3)
    a, b = workload.family("dense-random-1", n, q, 7)
    primes = orc.reference_primes_for(n, q)
    got = orc.rns_multiply_threads(q, a, b, primes, threads)
    assert (got == orc.rns_multiply(q, a, b, primes)).all()
    if n <= 64:
        assert (got == orc.schoolbook(q, a, b)).all()
    if n == 32768:
        case = [c for c in GOLDEN["cases"] if c["n"] == n and c["family"] == "dense-random-1"][0]
        q = modulus_of(case["modulus"], workload)
        a, b = workload.family(case["family"], n, q, case["seed"])
        assert sha(orc.rns_multiply_threads(q, a, b, orc.reference_primes_for(n, q), threads)) == case["sha256"]["c"]


def test_general_primes(orc, workload):
    q = workload.modulus(1)
    primes = orc.primes_general(1 << 17, q)
    assert len(primes) == 59
    assert all(p < (1 << 30) and p % (1 << 18) == 1 for p in primes)
    assert list(primes) == sorted(primes, reverse=True)
    assert orc.product_bits(primes) > 2 + 17 + 2 * 868


def direct_eval(x, p, psi, n):
    """x(psi^(2*brv(r)+1)) for r < n, as Python integers."""
    logn = n.bit_length() - 1
    out = []
    for r in range(n):
        br = int(format(r, f"0{logn}b")[::-1], 2) if logn else 0
        root = pow(psi, 2 * br + 1, p)
        out.append(sum(int(v) * pow(root, j, p) for j, v in enumerate(x)) % p)
    return out


@pytest.mark.parametrize("n", [2, 4, 8, 32])
def test_ntt_direct_evaluation(orc, n):
    p = 1073479681  # 4095*2^18 + 1
    psi = orc.find_psi(p, n)
    assert pow(psi, n, p) == p - 1
    rng = np.random.default_rng(n)
    x = rng.integers(0, p, n, dtype=np.uint64).astype(np.uint32)
    assert list(orc.ntt_forward(p, psi, x)) == direct_eval(x, p, psi, n)
    back = orc.ntt_inverse(p, psi, orc.ntt_forward(p, psi, x))
    assert list(back) == [int(v) * n % p for v in x]


def test_ntt_convolution_theorem(orc):
    p, n = 1073479681, 256
    psi = orc.find_psi(p, n)
    rng = np.random.default_rng(5)
    a = rng.integers(0, p, n, dtype=np.uint64).astype(np.uint32)
    b = rng.integers(0, p, n, dtype=np.uint64).astype(np.uint32)
    fa, fb = orc.ntt_forward(p, psi, a), orc.ntt_forward(p, psi, b)
    prod = ((fa.astype(np.uint64) * fb.astype(np.uint64)) % p).astype(np.uint32)
    c = orc.ntt_inverse(p, psi, prod)
    inv_n = pow(n, p - 2, p)
    got = [int(v) * inv_n % p for v in c]
    want = [0] * n
    for i in range(n):
        for j in range(n):
            s = 1 if i + j < n else -1
            want[(i + j) % n] = (want[(i + j) % n] + s * int(a[i]) * int(b[j])) % p
    assert got == want


def test_crt_of_exact_convolution(orc, workload):
    q = workload.modulus(7)
    a, b = workload.family("dense-random-1", 16, q, 3)
    y = py_negacyclic(a, b, q)
    primes = orc.primes_general(16, q)
    res = np.array([[v % int(p) for v in y] for p in primes], dtype=np.uint32)
    got = orc.crt(q, primes, res)
    Q = to_int(q)
    assert [to_int(x) for x in got] == [v % Q for v in y]


def test_planted_mistake_is_caught(orc, workload):
    """A dropped wrap-around sign in a candidate result is caught by the schoolbook check."""
    q = workload.modulus(2)
    a, b = workload.family("dense-random-1", 8, q, 1)
    Q = to_int(q)
    y = py_negacyclic(a, b, q)
    wrong = []
    for k in range(8):  # cyclic instead of negacyclic
        s = sum(to_int(a[i]) * to_int(b[(k - i) % 8]) for i in range(8))
        wrong.append(s % Q)
    assert wrong != [v % Q for v in y]
    assert [to_int(x) for x in orc.schoolbook(q, a, b)] == [v % Q for v in y]


def test_residues_against_python_modulo(orc, workload):
    """orc_residues (the restatement of rns/solve.cu:138-157: even/odd Shoup
    sums of the 28 limbs against 2^(32j) mod p) against Python's exact
    int % p, on values at every limb boundary: 2^(32j) - 1, 2^(32j),
    2^(32j) + 1, all-ones limbs, q - 1, and random coefficients."""
    import random
    q = workload.modulus(5)
    qi = workload.to_int(q)
    vals = [0, 1, qi - 1, (1 << 867) + 12345]
    for j in range(1, 28):
        vals += [(1 << (32 * j)) - 1, 1 << (32 * j), (1 << (32 * j)) + 1]
    vals.append(sum(0xFFFFFFFF << (32 * j) for j in range(27)))  # 27 all-ones limbs
    rng = random.Random(7)
    vals += [rng.randrange(qi) for _ in range(64)]
    x = np.stack([workload.from_int(v) for v in vals])
    for primes in (orc.primes_reference(), orc.primes_general(1 << 17, q)):
        got = orc.residues(x, primes)
        want = np.array([[v % int(p) for v in vals] for p in primes], dtype=np.uint64)
        assert (got.astype(np.uint64) == want).all()
