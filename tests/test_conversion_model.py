"""CPU model of the int8-MFMA residue conversion's recombination
(pm_kernels.hip: conv_residue), run on the primes the library plans.

The kernel forms, per (coefficient, prime), the four digit sums of the signed
bytes (a_j - 128) against balanced base-256 digits of 256^j mod p, and then
    v = (acc0 + (acc1 << 8)) + fold.x + Shoup((acc2 + (acc3 << 8)) + fold.y, 2^16)
in wrapping 32-bit arithmetic, where fold = {128 (C0 + 256 C1) + p,
128 (C2 + 256 C3) + 2p} (C_d the digit sums of the table).  It relies on:
  * v, computed exactly (no wrap), lies in [0, 2^32) and is congruent to the
    coefficient mod p;
  * after one conditional subtraction of 2p it lies below 4p (the NTT's
    lazy input range).
This checks both on random bytes and on the byte patterns that drive the
recombination's halves to their extremes, for every prime of the form
c 2^18 + 1 the planner can choose (the 64 largest below 2^30).  A wrong fold
constant, a dropped correction or a bound that does not hold fails here.
"""
import numpy as np
import pytest

P32 = 1 << 32


def candidate_primes(count=64):
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

    out = []
    for c in range(((1 << 30) - 1) >> 18, 0, -1):
        p = (c << 18) + 1
        if p < (1 << 30) and is_prime(p):
            out.append(p)
            if len(out) == count:
                break
    return out


def digits(p):
    """Balanced base-256 digits (d0..d3 in [-128,127]) of 256^j mod p, j < 112."""
    rows = []
    r = 1
    for _ in range(112):
        v, ds = r, []
        for _ in range(4):
            d = ((v + 128) & 255) - 128
            ds.append(d)
            v = (v - d) // 256
        assert v == 0
        rows.append(ds)
        r = r * 256 % p
    return np.array(rows, dtype=np.int64)  # [112][4]


def shoup(a, w, p):
    """The kernel's Shoup product (wrapping 32-bit), w < p."""
    ws = (w << 32) // p
    return (a * w - ((a * ws) >> 32) * p) % P32


def model(p, D, byte_rows):
    """v per input (exact integer) and the kernel's wrapped value."""
    C = 128 * D.sum(axis=0)  # corrections
    fold_x = (int(C[0]) + 256 * int(C[1]) + p) % P32
    fold_y = (int(C[2]) + 256 * int(C[3]) + 2 * p) % P32
    out = []
    for a in byte_rows:
        acc = [int(x) for x in ((a.astype(np.int64) - 128) @ D)]  # the MFMA's exact int32 sums
        u_raw = (acc[0] + (acc[1] << 8)) % P32
        w = (acc[2] + (acc[3] << 8) + fold_y) % P32
        v_kernel = (u_raw + fold_x + shoup(w, 1 << 16, p)) % P32
        # exact integers, no wrap
        T = [acc[d] + int(C[d]) for d in range(4)]
        u, wt = T[0] + 256 * T[1], T[2] + 256 * T[3]
        assert 0 <= wt + 2 * p < P32  # the Shoup input is the true w + 2p
        x = shoup(wt + 2 * p, 1 << 16, p)
        assert 0 <= x < 2 * p
        v = u + p + x
        out.append((v, v_kernel, int(sum(int(b) << (8 * j) for j, b in enumerate(a)))))
    return out


@pytest.mark.parametrize("p", candidate_primes()[::7] + candidate_primes()[-3:])
def test_recombination_range_and_congruence(p):
    D = digits(p)
    rng = np.random.default_rng(p)
    rows = [rng.integers(0, 256, 112, dtype=np.int64) for _ in range(24)]
    e, f = D[:, 0] + 256 * D[:, 1], D[:, 2] + 256 * D[:, 3]
    for sel in (e, f):  # each half at its largest and smallest
        rows.append(np.where(sel > 0, 255, 0))
        rows.append(np.where(sel < 0, 255, 0))
    rows += [np.full(112, 255), np.zeros(112, np.int64)]
    for v, v_kernel, value in model(p, D, rows):
        assert 0 <= v < P32, "recombined value outside 32 bits"
        assert v_kernel == v, "wrapped arithmetic differs from the exact value"
        assert v % p == value % p, "not congruent to the coefficient"
        r = v - 2 * p if v >= 2 * p else v
        assert r < 4 * p, "one subtraction of 2p does not reach the NTT's input range"


def test_bounds_hold_for_every_plannable_prime():
    """The exact extremes over any 112 bytes (what pm_params.cpp checks when it
    builds the tables) hold with margin for the 64 largest primes."""
    for p in candidate_primes():
        D = digits(p)
        e, f = D[:, 0] + 256 * D[:, 1], D[:, 2] + 256 * D[:, 3]
        u_lo, u_hi = 255 * int(e[e < 0].sum()), 255 * int(e[e > 0].sum())
        w_lo, w_hi = 255 * int(f[f < 0].sum()), 255 * int(f[f > 0].sum())
        assert u_lo >= -p and u_hi + 3 * p <= P32 and u_hi < 3 * p
        assert w_lo >= -2 * p and w_hi + 2 * p < P32


def test_library_plans_only_checked_primes(pm, workload):
    """Every prime the library plans is among the candidates checked above."""
    cand = set(candidate_primes())
    for seed in (1, 2):
        for max_n in (1 << 12, 1 << 17):
            _, primes, _, _ = pm.plan_rns(workload.modulus(seed), max_n)
            assert len(primes) > 0 and set(int(p) for p in primes) <= cand
