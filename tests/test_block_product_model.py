"""CPU model of the fused tile's incomplete transform and block product
(pm_kernels.hip: ct_bfly, block_product, redc64, gs_bfly; the constants of
pm_params.cpp: block_product_constants), in the kernel's own 32/64-bit
wrapping arithmetic, on the 30-bit primes c 2^18 + 1 the planner chooses.

The tile's forward Cooley-Tukey stages stop at span 4: block G (positions
4G..4G+3) then holds its operand modulo X^4 - zeta_G, zeta_G = +w (G even) or
-w (G odd) for the forward twiddle w = psi^brv of index 2^(logn-3) + G/2
(zeta = -1 for N = 4).  Each block is multiplied as
    c_k = REDC(sum_{i<=k} A_i B_{k-i} + sum_{i>k} A_i Z_{k-i+4}),
A_i = a_i reduced to [0,p), B_j = Shoup(b_j, S), Z_j = Shoup(b_j, zeta S),
S = 4 N^-1 2^32 (times any constant), and the Gentleman-Sande stages from
span 4 upward undo the forward ones.  This checks:
  * the result equals the negacyclic product mod p (schoolbook) for
    N = 4 .. 256, on random inputs in the tile's input range [0,4p) and on
    all-maximal inputs, so a wrong zeta (sign, index, the N = 4 case) or a
    wrong scale fails;
  * the bounds the kernel relies on hold for every candidate prime: the
    64-bit REDC sum T + m p stays below 2^64, its result below 3p, and every
    stage's outputs inside the ranges the next stage assumes;
  * the inverse's complemented operands (gs_bfly_cn / _cc / _nc: a value that
    is the next stage's Y operand held as 2p - Y) give the same product for
    any choice of the stages that complement, extremes included.
"""
import numpy as np
import pytest

M32 = (1 << 32) - 1


def candidate_primes(count=64):
    out, c = [], (1 << 12) - 1
    while len(out) < count and c > 0:
        p = c * (1 << 18) + 1
        if p < (1 << 30) and all(pow(a, p - 1, p) == 1 for a in (2, 3, 5, 7, 11, 13, 17, 19, 23)):
            out.append(p)
        c -= 1
    return out


PRIMES = candidate_primes()


def brv(x, bits):
    return int(format(x, f"0{bits}b")[::-1], 2) if bits else 0


def psi_for(p, n):
    """A primitive 2n-th root of unity mod p."""
    for g in range(2, 200):
        psi = pow(g, (p - 1) // (2 * n), p)
        if pow(psi, n, p) == p - 1:
            return psi
    raise AssertionError("no root")


def shoup_pair(w, p):
    return w, (w << 32) // p


def csub(x, m):
    t = (x - m) & M32
    return t if t < x else x


def mulhi(a, b):
    return (a * b) >> 32


def ct_bfly(X, Y, w, p):
    """The kernel's forward butterfly with the twiddle stored negated."""
    assert X < 4 * p
    x = csub(X, 2 * p)
    value, shoup = (-w) & M32, shoup_pair(w, p)[1]
    nt = (mulhi(Y, shoup) * p + Y * value) & M32
    y = (x + 2 * p + nt) & M32
    X2 = (x - nt) & M32
    assert X2 < 4 * p and y < 4 * p
    return X2, y


def shoup_np(a, w, p):
    value, shoup = shoup_pair(w, p)
    r = (a * value + mulhi(a, shoup) * ((-p) & M32)) & M32
    assert r < 2 * p and r % p == a * w % p
    return r


def gs_bfly(X, Y, w, p):
    assert X < 2 * p and Y < 2 * p
    s = X + Y
    d = (X - Y + 2 * p) & M32
    return csub(s, 2 * p), shoup_np(d, w, p)


def gs_bfly_cn(X, Yc, w, p):
    """Complemented Y in (Yc = 2p - Y in [0,2p]), normal outputs."""
    assert X < 2 * p and Yc <= 2 * p
    e = (X - Yc) & M32
    d = X + Yc
    Xo = min(e, (e + 2 * p) & M32)
    assert Xo < 2 * p
    return Xo, shoup_np(d, w, p)


def gs_bfly_cc(X, Yc, w, p):
    """Complemented Y in, complemented outputs (twiddle p - w)."""
    assert X < 2 * p and Yc <= 2 * p
    f = (Yc - X) & M32
    d = X + Yc
    Xo = min(f, (f + 2 * p) & M32)
    assert Xo <= 2 * p
    return Xo, shoup_np(d, (p - w) % p, p)


def gs_bfly_nc(X, Y, w, p):
    """Normal inputs, complemented outputs (twiddle p - w)."""
    assert X < 2 * p and Y < 2 * p
    u = 4 * p - (X + Y)
    d = (X - Y + 2 * p) & M32
    Xo = csub(u, 2 * p)
    assert Xo <= 2 * p
    return Xo, shoup_np(d, (p - w) % p, p)


def redc64(T, p):
    neg_inv = (-pow(p, -1, 1 << 32)) & M32
    m = (T * neg_inv) & M32
    total = m * p + T
    assert total < (1 << 64), "REDC sum overflows 64 bits"
    r = total >> 32
    assert r < 3 * p
    return r


def zeta_of(g, logn, fwd, p):
    if logn == 2:
        return p - 1
    w = fwd[(1 << (logn - 3)) + g // 2]
    return (p - w) % p if g & 1 else w


def tile_product(a, b, p, complemented=()):
    """complemented: the inverse stages beta that leave the positions with bit
    beta + 1 set complemented (2p - v) for stage beta + 1 (kBlockLog - 1 = 1
    stands for the block product complementing the positions with bit 2 set)."""
    n = len(a)
    logn = n.bit_length() - 1
    psi = psi_for(p, n)
    fwd = [pow(psi, brv(j, logn), p) for j in range(n)]
    inv = [pow(psi, -brv(j, logn), p) for j in range(n)]
    a, b = list(a), list(b)
    for beta in range(logn - 1, 1, -1):  # forward stages down to span 4
        m = 1 << (logn - 1 - beta)
        for x in range(n):
            if x >> beta & 1:
                continue
            w = fwd[m + (x >> (beta + 1))]
            y = x | 1 << beta
            a[x], a[y] = ct_bfly(a[x], a[y], w, p)
            b[x], b[y] = ct_bfly(b[x], b[y], w, p)
    S = 4 * pow(n, -1, p) * pow(2, 32, p) % p
    for g in range(n // 4):
        zs = zeta_of(g, logn, fwd, p) * S % p
        A = [csub(csub(a[4 * g + i], 2 * p), p) for i in range(4)]
        B = [shoup_np(b[4 * g + j], S, p) for j in range(4)]
        Z = [0] + [shoup_np(b[4 * g + j], zs, p) for j in range(1, 4)]
        for k in range(4):
            T = sum(A[i] * (B[k - i] if i <= k else Z[k - i + 4]) for i in range(4))
            assert T < 8 * p * p
            t = redc64(T, p)
            comp = 1 in complemented and logn > 2 and (4 * g + k) >> 2 & 1
            a[4 * g + k] = csub(3 * p - t, p) if comp else csub(t, p)
    for beta in range(2, logn):  # inverse stages from span 4 upward
        m = 1 << (logn - 1 - beta)
        cin = (beta - 1) in complemented
        for x in range(n):
            if x >> beta & 1:
                continue
            w = inv[m + (x >> (beta + 1))]
            y = x | 1 << beta
            cout = beta in complemented and beta + 1 < logn and x >> (beta + 1) & 1
            bf = {(False, False): gs_bfly, (True, False): gs_bfly_cn,
                  (True, True): gs_bfly_cc, (False, True): gs_bfly_nc}[(cin, bool(cout))]
            a[x], a[y] = bf(a[x], a[y], w, p)
    assert all(v < 2 * p for v in a)  # the last stage leaves normal values
    return [csub(v, p) for v in a]


def negacyclic(a, b, p):
    n = len(a)
    out = [0] * n
    for i in range(n):
        for j in range(n):
            if i + j < n:
                out[i + j] += a[i] * b[j]
            else:
                out[i + j - n] -= a[i] * b[j]
    return [v % p for v in out]


@pytest.mark.parametrize("n", [4, 8, 16, 64, 256])
@pytest.mark.parametrize("which", [0, 1, -1])
def test_block_product_equals_negacyclic_product(n, which):
    p = PRIMES[which]
    rng = np.random.default_rng(n * 7 + which)
    for case in ("random", "maximal"):
        if case == "random":
            a = [int(v) for v in rng.integers(0, 4 * p, size=n)]
            b = [int(v) for v in rng.integers(0, 4 * p, size=n)]
        else:
            a = [4 * p - 1] * n
            b = [4 * p - 1] * n
        got = tile_product(a, b, p)
        want = negacyclic([v % p for v in a], [v % p for v in b], p)
        assert got == want, (n, p, case)


@pytest.mark.parametrize("n", [8, 64, 256])
@pytest.mark.parametrize("which", [0, -1])
def test_complemented_inverse_stages_give_the_same_product(n, which):
    """The kernel complements, per stage, the next stage's Y operands when that
    bit is held in registers; any such choice of stages (none, all, every
    other, random) gives the same exact product."""
    p = PRIMES[which]
    logn = n.bit_length() - 1
    rng = np.random.default_rng(n + which)
    a = [int(v) for v in rng.integers(0, 4 * p, size=n)]
    b = [int(v) for v in rng.integers(0, 4 * p, size=n)]
    want = negacyclic([v % p for v in a], [v % p for v in b], p)
    stages = list(range(1, logn - 1))
    choices = [(), tuple(stages), tuple(stages[::2]), tuple(stages[1::2]),
               tuple(s for s in stages if rng.integers(0, 2))]
    for chosen in choices:
        assert tile_product(a, b, p, chosen) == want, chosen
    # extremes: every residue at the top of its range
    top = [4 * p - 1] * n
    want_top = negacyclic([v % p for v in top], [v % p for v in top], p)
    assert tile_product(top, top, p, tuple(stages)) == want_top


def test_block_product_bounds_hold_for_every_candidate_prime():
    for p in PRIMES:
        assert 4 * p < (1 << 32)
        t_max = 4 * (p - 1) * (2 * p - 1)  # A < p, B and Z < 2p, four terms
        total = t_max + M32 * p  # m < 2^32
        assert total < (1 << 64)
        assert (total >> 32) < 3 * p
        # after one conditional subtraction the inverse stages get [0, 2p)
        assert (total >> 32) - p < 2 * p


def test_zeta_of_n4_is_minus_one_and_blocks_pair_up():
    p = PRIMES[0]
    assert zeta_of(0, 2, None, p) == p - 1
    n = 64
    psi = psi_for(p, n)
    fwd = [pow(psi, brv(j, 6), p) for j in range(n)]
    for g in range(0, n // 4, 2):
        z0, z1 = zeta_of(g, 6, fwd, p), zeta_of(g + 1, 6, fwd, p)
        assert (z0 + z1) % p == 0
        # the blocks' moduli X^4 - zeta are the 16 factors of X^64 + 1
        assert pow(z0, n // 4, p) == p - 1
