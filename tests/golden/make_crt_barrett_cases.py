"""Writes tests/golden/crt_barrett_cases.json: canonical residue vectors whose
CRT lands where a quotient estimate from floor(X / 2^864) and
floor(2^908 / q) would leave X - Qh q in [2q, 3q).

The CRT of canonical residues t_i (pm_kernels.hip: crt_finish) forms
    K = round(sum_i t_i floor(2^42/p_i) / 2^42),
    X = sum_i t_i (M_i mod q) + ((-K P) mod q)        (X < 2^905),
and reduces X modulo q with one Barrett step and one conditional
subtraction.  With the estimate from X's bits above 2^864 the error
X/q - Qh reaches 1 + 2^864/q + X/2^908, past 2 for an 865-bit q just above
2^864; with the bits above 2^863 and floor(2^906/q) it stays below 2.  This
script draws such a q (865 bits, q - 2^864 < 2^850), the primes the library
plans for it at N = 2^15 (largest c 2^18 + 1 below 2^30 first, until
P > 4 N (q - 1)^2), and random residue vectors near p_i - 1 until it has
found four whose estimate from 2^864 fails.  The vectors are the fixture;
the expected product words follow from them by exact integer arithmetic in
the tests.  Pure Python, a few minutes.
"""
import json
import random
from pathlib import Path

HERE = Path(__file__).resolve().parent
N_MAX = 1 << 15


def is_prime(n):
    if n < 2:
        return False
    d, s = n - 1, 0
    while d % 2 == 0:
        d //= 2
        s += 1
    for a in (2, 3, 5, 7, 11, 13, 17, 19, 23, 29, 31, 37):
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


def planned_primes(q, max_n):
    bound = 4 * max_n * (q - 1) ** 2
    out, product = [], 1
    for c in range((1 << 12) - 1, 0, -1):
        p = (c << 18) + 1
        if not is_prime(p):
            continue
        out.append(p)
        product *= p
        if product > bound:
            break
    return out, product


def crt_constants(primes, q, product):
    return [(product // p) % q for p in primes], [(1 << 42) // p for p in primes]


def barrett_rest(t, primes, q, product, shift, mu_exp, consts):
    """X - Qh q of the kernel's CRT finish for residues t."""
    k = len(primes)
    m_mod_q, rounding = consts
    est = sum(ti * r for ti, r in zip(t, rounding))
    K = min((est + (1 << 41)) >> 42, k)
    X = sum(ti * m for ti, m in zip(t, m_mod_q)) + (-K * product) % q
    mu = (1 << mu_exp) // q
    qh = ((X >> shift) * mu) >> (mu_exp - shift)
    return X - qh * q


def main():
    rng = random.Random(11)
    q = (1 << 864) + rng.getrandbits(850) | 1
    primes, product = planned_primes(q, N_MAX)
    consts = crt_constants(primes, q, product)
    cases = []
    draws = random.Random(7)
    while len(cases) < 4:
        t = [p - 1 - draws.getrandbits(draws.choice((4, 12, 20, 29))) % p for p in primes]
        if barrett_rest(t, primes, q, product, 864, 908, consts) >= 2 * q:
            assert barrett_rest(t, primes, q, product, 863, 906, consts) < 2 * q
            cases.append(t)
    q_words = [(q >> (32 * w)) & 0xFFFFFFFF for w in range(28)]
    out = {"q_words": q_words, "max_n": N_MAX, "primes": primes, "residues": cases}
    (HERE / "crt_barrett_cases.json").write_text(json.dumps(out) + "\n")
    print(f"wrote {len(cases)} cases, k = {len(primes)}")


if __name__ == "__main__":
    main()
