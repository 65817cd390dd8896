"""CPU model of the CRT's final reduction modulo q (pm_kernels.hip:
crt_finish; the constant floor(2^906/q) of pm_params.cpp).

X = sum_i t_i (M_i mod q) + ((-K P) mod q) is below 2^905 for residues
t_i < 2 p_i.  The kernel estimates Qh = floor(floor(X/2^863) floor(2^906/q)
/ 2^43), forms X - Qh q and subtracts q once if that is at least q, so it
needs 0 <= X - Qh q < 2q.  This checks:
  * the fixture's residue vectors (tests/golden/crt_barrett_cases.json, an
    865-bit q just above 2^864) put X - Qh q at 2q or above with an estimate
    from floor(X/2^864) and floor(2^908/q), and inside [0, 2q) with the
    kernel's;
  * the kernel's estimate stays inside [0, 2q) on X drawn at the largest
    X/2^863 and with X mod 2^863 at its top, for q near 2^864, near 2^868
    and the reference's 2^868 - 11993087;
  * X < 2^905 for the largest residues and q.
"""
import json
import os
import random
import sys

import pytest

from conftest import ROOT

sys.path.insert(0, os.path.join(ROOT, "tests", "golden"))
import make_crt_barrett_cases as gen  # noqa: E402

FIXTURE = json.load(open(os.path.join(ROOT, "tests", "golden", "crt_barrett_cases.json")))


def fixture_q():
    return sum(w << (32 * i) for i, w in enumerate(FIXTURE["q_words"]))


def rest(X, q, shift, mu_exp):
    mu = (1 << mu_exp) // q
    assert mu < 1 << 64  # two words
    return X - (((X >> shift) * mu) >> (mu_exp - shift)) * q


def test_fixture_modulus_and_primes_are_the_planned_ones():
    q = fixture_q()
    assert q.bit_length() == 865 and q > 1 << 864 and q & 1
    primes, _ = gen.planned_primes(q, FIXTURE["max_n"])
    assert primes == FIXTURE["primes"]


def test_fixture_cases_break_the_2_864_estimate_not_the_kernels():
    q = fixture_q()
    primes, product = gen.planned_primes(q, FIXTURE["max_n"])
    consts = gen.crt_constants(primes, q, product)
    for t in FIXTURE["residues"]:
        assert all(0 <= ti < p for ti, p in zip(t, primes))
        assert gen.barrett_rest(t, primes, q, product, 864, 908, consts) >= 2 * q
        r = gen.barrett_rest(t, primes, q, product, 863, 906, consts)
        assert 0 <= r < 2 * q


@pytest.mark.parametrize("which", ["just-above-2^864", "866-bit", "near-2^868", "reference"])
def test_kernel_estimate_in_range_at_the_extremes(which):
    rng = random.Random(which)
    qs = {
        "just-above-2^864": [(1 << 864) + rng.getrandbits(850) | 1 for _ in range(4)],
        "866-bit": [(1 << 865) + rng.getrandbits(864) | 1 for _ in range(4)],
        "near-2^868": [(1 << 868) - rng.getrandbits(800) | 1 for _ in range(4)],
        "reference": [(1 << 868) - 11993087],
    }[which]
    for q in qs:
        for i in range(4000):
            xh = (1 << 42) - 1 - rng.getrandbits(20 if i % 2 else 41)
            xl = (1 << 863) - 1 - rng.getrandbits(rng.choice((8, 200, 862)))
            X = min((xh << 863) | xl, (1 << 905) - 1)
            r = rest(X, q, 863, 906)
            assert 0 <= r < 2 * q, (hex(q), X)


def test_crt_sum_below_2_905():
    q = (1 << 868) - 1
    primes, product = gen.planned_primes((1 << 867) + 1, 1 << 17)
    # residues t_i < 2 p_i (lazy), every M_i mod q and (-K P) mod q below q
    assert sum(2 * p - 1 for p in primes) * (q - 1) + q < 1 << 905
    assert len(primes) <= 64
