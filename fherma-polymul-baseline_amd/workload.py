"""Seeded synthetic workloads for the 868-bit negacyclic product.

SplitMix64 streams (public-domain finaliser constants):
  value64(seed, tag, i) = mix64(mix64(seed ^ tag) + i)
  modulus(seed): 14 u64 words value64(seed, 0x51, w); bits >= 868 cleared,
                 bit 867 and bit 0 set.
  poly(seed, poly, n, q): coefficient i, attempt t < 64 uses u64 words
                 value64(seed, 0xA0000000 + poly, (i*64 + t)*16 + w), w < 14,
                 masked to 868 bits; the first attempt below q is taken
                 (after 64 rejections: value - q).
Coefficients are returned as [n, 28] little-endian uint32 limbs, the wire
layout of fherma.h (tensor<N x L x u32>).
"""
import numpy as np

LIMBS = 28
WIDTH = 868
_GOLDEN = np.uint64(0x9E3779B97F4A7C15)
_C1 = np.uint64(0xBF58476D1CE4E5B9)
_C2 = np.uint64(0x94D049BB133111EB)
TAG_MODULUS = 0x51
TAG_POLY = 0xA0000000
_MASK36 = np.uint64((1 << 36) - 1)


def mix64(x):
    x = np.asarray(x, dtype=np.uint64)
    with np.errstate(over="ignore"):
        z = x + _GOLDEN
        z = (z ^ (z >> np.uint64(30))) * _C1
        z = (z ^ (z >> np.uint64(27))) * _C2
        return z ^ (z >> np.uint64(31))


def value64(seed, tag, index):
    key = mix64(np.uint64(seed) ^ np.uint64(tag))
    with np.errstate(over="ignore"):
        return mix64(key + np.asarray(index, dtype=np.uint64))


def _words(v64):
    """[rows, 14] uint64 -> [rows, 28] uint32 (little-endian limbs)."""
    return np.ascontiguousarray(v64).view("<u4").reshape(v64.shape[0], LIMBS)


def modulus(seed):
    v = value64(seed, TAG_MODULUS, np.arange(14, dtype=np.uint64)).reshape(1, 14).copy()
    v[0, 13] &= _MASK36
    v[0, 13] |= np.uint64(1 << 35)
    v[0, 0] |= np.uint64(1)
    return _words(v)[0].copy()


def special_modulus(delta):
    """q = 2^868 - delta, 0 < delta < 2^32."""
    return from_int((1 << WIDTH) - int(delta))


def _less(words, q):
    """rows of [rows, 28] uint32 strictly below q (28 words)."""
    less = np.zeros(words.shape[0], dtype=bool)
    decided = np.zeros(words.shape[0], dtype=bool)
    for w in range(LIMBS - 1, -1, -1):
        col = words[:, w]
        lt = (col < q[w]) & ~decided
        gt = (col > q[w]) & ~decided
        less |= lt
        decided |= lt | gt
    return less


def poly(seed, which, n, q):
    q = np.asarray(q, dtype=np.uint32)
    out = np.zeros((n, LIMBS), dtype=np.uint32)
    pending = np.arange(n, dtype=np.uint64)
    tag = TAG_POLY + int(which)
    w = np.arange(14, dtype=np.uint64)
    last = None
    for t in range(64):
        if pending.size == 0:
            break
        idx = ((pending * np.uint64(64) + np.uint64(t)) * np.uint64(16))[:, None] + w[None, :]
        v = value64(seed, tag, idx)
        v[:, 13] &= _MASK36
        words = _words(v)
        ok = _less(words, q)
        out[pending[ok].astype(np.int64)] = words[ok]
        last = (pending[~ok], words[~ok])
        pending = pending[~ok]
    if pending.size:
        rows, words = last
        for r, wv in zip(rows.astype(np.int64), words):
            out[r] = from_int(to_int(wv) - to_int(q))
    return out


def to_int(words):
    return int.from_bytes(np.ascontiguousarray(words, dtype="<u4").tobytes(), "little")


def from_int(x, limbs=LIMBS):
    return np.frombuffer(int(x).to_bytes(4 * limbs, "little"), dtype="<u4").copy()


def family(name, n, q, seed):
    """Input pair of one of the reference validator's case families
    (tests/validate.py:46-59 of the reference), with our seeded streams."""
    q = np.asarray(q, dtype=np.uint32)
    qi = to_int(q)
    r0, r1 = poly(seed, 0, n, q), poly(seed, 1, n, q)
    zero = np.zeros((n, LIMBS), dtype=np.uint32)
    qm1 = from_int(qi - 1)
    if name == "dense-random-1":
        return r0, r1
    if name == "dense-random-2":
        return poly(seed, 2, n, q), poly(seed, 3, n, q)
    if name == "zero":
        return r0, zero
    if name == "identity":
        b = zero.copy()
        b[0, 0] = 1
        return r0, b
    if name == "wrap-sign":
        a, b = zero.copy(), zero.copy()
        a[n - 1] = qm1
        if n > 1:
            b[1] = qm1
        return a, b
    if name == "all-q-minus-one":
        return np.tile(qm1, (n, 1)), np.tile(qm1, (n, 1))
    if name == "alternating":
        a = zero.copy()
        a[1::2] = qm1
        return a, r0
    if name == "limb-boundaries":
        edge = [0, 1, qi - 1, qi - 2, 1 << 864, (1 << 864) - 1, 1 << 32, (1 << 32) - 1]
        a = np.stack([from_int(edge[i % 8]) for i in range(n)])
        return a, r1
    if name == "random-after-edges":
        return r1, r0
    raise ValueError(f"unknown family {name}")


FAMILIES = ["dense-random-1", "dense-random-2", "zero", "identity", "wrap-sign",
            "all-q-minus-one", "alternating", "limb-boundaries", "random-after-edges"]
