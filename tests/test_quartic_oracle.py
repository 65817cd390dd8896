"""The special-modulus quartic path, on the CPU: the oracle's restatement of
the reference's quartic algorithm (quartic/host_setup.h:41-110,
quartic/solve.cu:289-316 and 906-1040) pinned by the reference's own
figures for its prime rule and by exact products (GMP Kronecker fixtures,
schoolbook), and the library's quartic plan (pm_plan_quartic)."""
import hashlib
import json
import os

import numpy as np
import pytest

from conftest import ROOT

GOLDEN = json.load(open(os.path.join(ROOT, "tests", "golden", "golden.json")))
DELTA = 11993087  # the reference's published configuration (EXPERIMENTS_RU.md:94)


def sha(x):
    return hashlib.sha256(np.ascontiguousarray(x, dtype="<u4").tobytes()).hexdigest()


def test_reference_prime_rule_figures(orc):
    """research/LAZY_NTT_RU.md:6-7: at N = 2^15, delta = 11993087 gives 30-bit
    primes, a 480-bit P and a 475-bit bound; delta = 2^28 - 1 needs the
    31-bit set, a 496-bit P against a 479-bit bound."""
    primes, pbits, bbits = orc.quartic_primes_reference(32768, DELTA)
    assert len(primes) == 16 and (primes < (1 << 30)).all() and (pbits, bbits) == (480, 475)
    primes, pbits, bbits = orc.quartic_primes_reference(32768, (1 << 28) - 1)
    assert len(primes) == 16 and (primes >= (1 << 30)).any() and (pbits, bbits) == (496, 479)
    for p in primes:
        p = int(p)
        assert pow((1 << 28) - 1, (p - 1) // 4, p) == 1 and p % 65536 == 1


def test_quartic_delta(orc, workload):
    assert orc.quartic_delta(workload.special_modulus(DELTA)) == DELTA
    assert orc.quartic_delta(workload.special_modulus((1 << 28) - 1)) == (1 << 28) - 1
    assert orc.quartic_delta(workload.special_modulus(1 << 28)) is None  # delta must be below 2^28
    assert orc.quartic_delta(workload.modulus(1)) is None


@pytest.mark.parametrize("rule", ["reference", "library"])
def test_quartic_oracle_vs_golden_n32(pm, orc, workload, rule):
    """Every case family of the reference validator at N = 32 under q =
    2^868 - 11993087: the quartic product equals the GMP Kronecker fixture,
    with the reference's primes and with the library's plan."""
    q = workload.special_modulus(DELTA)
    primes = orc.quartic_primes_reference(32, DELTA)[0] if rule == "reference" else pm.plan_quartic(q, 1 << 15)[1]
    cases = [c for c in GOLDEN["cases"] if c["n"] == 32 and c["modulus"] == f"delta:{DELTA}"]
    assert len(cases) == 9
    for case in cases:
        a, b = workload.family(case["family"], 32, q, case["seed"])
        assert sha(orc.quartic_multiply(q, a, b, primes)) == case["sha256"]["c"], case["family"]


def test_quartic_oracle_vs_schoolbook(pm, orc, workload):
    """Random and extreme operands at small N and several deltas (the
    largest, 2^28 - 1, included): quartic product == schoolbook."""
    for delta in (1, 3, DELTA, (1 << 28) - 1):
        q = workload.special_modulus(delta)
        for n in (4, 16, 64):
            primes = pm.plan_quartic(q, max(n, 1 << 10))[1]
            for family in ("dense-random-1", "all-q-minus-one", "wrap-sign", "limb-boundaries"):
                a, b = workload.family(family, n, q, 5)
                assert (orc.quartic_multiply(q, a, b, primes) == orc.schoolbook(q, a, b)).all(), (delta, n, family)


def test_quartic_oracle_vs_golden_n32768(pm, orc, workload):
    q = workload.special_modulus(DELTA)
    primes = pm.plan_quartic(q, 1 << 15)[1]
    cases = [c for c in GOLDEN["cases"] if c["n"] == 32768 and c["modulus"] == f"delta:{DELTA}"][:1]
    for case in cases:
        a, b = workload.family(case["family"], 32768, q, case["seed"])
        assert sha(orc.quartic_multiply(q, a, b, primes)) == case["sha256"]["c"], case["family"]


def test_library_quartic_plan(pm, workload):
    """pm_plan_quartic: primes c 2^18 + 1 < 2^30 with delta a fourth power,
    r^4 = delta, P > 4 max_n (2^217 - 1)^2 (1 + 3 delta); a general q refused."""
    for delta, max_n in ((DELTA, 1 << 15), ((1 << 28) - 1, 1 << 17), (1, 1 << 10)):
        q = workload.special_modulus(delta)
        k, primes, roots, bits = pm.plan_quartic(q, max_n)
        P = 1
        for p, r in zip(primes, roots):
            p, r = int(p), int(r)
            assert p % (1 << 18) == 1 and p < (1 << 30) and pow(r, 4, p) == delta % p
            P *= p
        bound = max_n * ((1 << 217) - 1) ** 2 * (1 + 3 * delta)
        assert P > 4 * bound and P.bit_length() == bits and k <= 17
    with pytest.raises(pm.PolymulError):
        pm.plan_quartic(workload.modulus(1), 1 << 15)
