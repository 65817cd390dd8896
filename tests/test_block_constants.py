"""The tile's block-product and complement tables as pm_params.cpp builds
them (block_product_constants, complement_twiddle), built with g++ and run on
the CPU (tests/native/block_constants_dump.cpp), checked against direct
polynomial arithmetic rather than their defining formulas:
  * for every N = 2^2 .. 2^6, a random polynomial taken through the forward
    Cooley-Tukey stages down to span 4 with the library's own twiddle table
    leaves in block G exactly A mod (X^4 - zeta_G), where zeta_G is the table
    entry divided by the tile's scale (long division of A by X^4 - zeta_G);
  * every zeta_G is a root of Y^(N/4) = -1 and they are distinct;
  * the tile's scale is 4 times the post-inverse scale for N >= 4;
  * a Shoup product with the complement twiddle is congruent to minus the
    product with the inverse twiddle, for random operands.
"""
import json
import os
import random
import subprocess

from conftest import PKG_DIR, ROOT

M32 = (1 << 32) - 1


def dump(tmp_path):
    exe = tmp_path / "block_constants_dump"
    src = os.path.join(ROOT, "tests", "native", "block_constants_dump.cpp")
    lib = os.path.join(PKG_DIR, "csrc", "pm_params.cpp")
    subprocess.run(["g++", "-std=c++17", "-O1", "-Wall", "-I", os.path.join(PKG_DIR, "csrc"), src, lib,
                    "-o", str(exe)], check=True, timeout=120)
    out = subprocess.run([str(exe)], capture_output=True, text=True, timeout=60)
    assert out.returncode == 0, out.stdout + out.stderr
    return json.loads(out.stdout)


def forward_to_blocks(a, fwd, p):
    """Cooley-Tukey stages of span N/2 .. 4 with the tile's twiddle indexing
    (stage with m groups, group g: table entry m + g), exact arithmetic."""
    n = len(a)
    logn = n.bit_length() - 1
    a = list(a)
    for beta in range(logn - 1, 1, -1):
        m = 1 << (logn - 1 - beta)
        for x in range(n):
            if x >> beta & 1:
                continue
            w = fwd[m + (x >> (beta + 1))]
            y = x | 1 << beta
            t = a[y] * w % p
            a[x], a[y] = (a[x] + t) % p, (a[x] - t) % p
    return a


def reduce_mod_x4_minus(a, zeta, p):
    """A mod (X^4 - zeta): X^(4k + r) -> zeta^k X^r."""
    out = [0] * 4
    for e, c in enumerate(a):
        out[e % 4] = (out[e % 4] + c * pow(zeta, e // 4, p)) % p
    return out


def test_zeta_table_matches_the_transform(tmp_path):
    rng = random.Random(2)
    for prime in dump(tmp_path)["primes"]:
        p, fwd = prime["p"], prime["forward"]
        for logn in range(2, 7):
            n = 1 << logn
            S = prime["scale4"][logn][0]
            assert S == 4 * prime["scale"][logn] % p
            assert prime["scale4"][logn][1] == (S << 32) // p
            S_inv = pow(S, p - 2, p)
            zetas = []
            for g in range(n // 4):
                value, shoup = prime["zeta"][(1 << (logn - 2)) + g]
                assert shoup == (value << 32) // p
                zetas.append(value * S_inv % p)
            assert len(set(zetas)) == n // 4
            assert all(pow(z, n // 4, p) == p - 1 for z in zetas)
            a = [rng.randrange(p) for _ in range(n)]
            blocks = forward_to_blocks(a, fwd, p)
            for g, z in enumerate(zetas):
                assert blocks[4 * g:4 * g + 4] == reduce_mod_x4_minus(a, z, p), (p, logn, g)


def test_complement_twiddles_negate_the_product(tmp_path):
    rng = random.Random(3)
    for prime in dump(tmp_path)["primes"]:
        p = prime["p"]
        for w, (cv, cs) in zip(prime["inverse"], prime["complement"]):
            assert cs == (cv << 32) // p and cv < p
            for _ in range(8):
                d = rng.getrandbits(32)
                # the kernel's Shoup product d*cv + floor(d*cs/2^32)*(2^32 - p), mod 2^32
                r = (d * cv + ((d * cs) >> 32) * ((-p) & M32)) & M32
                assert r < 2 * p and (r + d * w) % p == 0
