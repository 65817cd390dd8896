"""ctypes binding of oracle/liboracle.so (the CPU checker)."""
import ctypes
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIMBS = 28
U32 = np.ctypeslib.ndpointer(np.uint32, flags="C")


class Oracle:
    def __init__(self, path=None):
        path = path or os.path.join(ROOT, "oracle", "liboracle.so")
        lib = ctypes.CDLL(path)
        u32, u64 = ctypes.c_uint32, ctypes.c_uint64
        lib.orc_gen_modulus.argtypes = [u64, U32]
        lib.orc_quartic_delta.argtypes = [U32, ctypes.POINTER(u32)]
        lib.orc_quartic_primes_reference.argtypes = [u32, u32, U32, ctypes.POINTER(u32), ctypes.POINTER(u32)]
        lib.orc_quartic_primes_reference.restype = u32
        lib.orc_quartic_multiply.argtypes = [u32, U32, u32, U32, U32, U32, U32]
        lib.orc_special_modulus.argtypes = [u32, U32]
        lib.orc_gen_poly.argtypes = [u64, u32, u32, U32, U32]
        lib.orc_schoolbook.argtypes = [u32, U32, U32, U32, U32]
        lib.orc_primes_reference.argtypes = [u32, U32]
        lib.orc_primes_general.argtypes = [u32, U32, U32]
        lib.orc_primes_general.restype = u32
        lib.orc_find_psi.argtypes = [u32, u32]
        lib.orc_find_psi.restype = u32
        lib.orc_crt_bound_ok.argtypes = [u32, U32, u32, U32]
        lib.orc_product_bits.argtypes = [u32, U32]
        lib.orc_product_bits.restype = u32
        lib.orc_pow_mod.argtypes = [u32, u64, u32]
        lib.orc_pow_mod.restype = u32
        lib.orc_residues.argtypes = [u32, U32, u32, U32, U32]
        lib.orc_ntt_forward.argtypes = [u32, u32, u32, U32]
        lib.orc_ntt_inverse.argtypes = [u32, u32, u32, U32]
        lib.orc_crt.argtypes = [u32, U32, u32, U32, U32, U32]
        lib.orc_crt_constants.argtypes = [U32, u32, U32, U32, U32, U32]
        lib.orc_rns_multiply.argtypes = [u32, U32, u32, U32, U32, U32, U32]
        lib.orc_rns_multiply_threads.argtypes = [u32, U32, u32, U32, U32, U32, U32, u32]
        self.lib = lib

    @staticmethod
    def _u32(x):
        return np.ascontiguousarray(np.asarray(x, dtype=np.uint32))

    def modulus(self, seed):
        q = np.zeros(LIMBS, np.uint32)
        self.lib.orc_gen_modulus(seed, q)
        return q

    def special_modulus(self, delta):
        q = np.zeros(LIMBS, np.uint32)
        self.lib.orc_special_modulus(delta, q)
        return q

    def poly(self, seed, which, n, q):
        out = np.zeros(n * LIMBS, np.uint32)
        self.lib.orc_gen_poly(seed, which, n, self._u32(q), out)
        return out.reshape(n, LIMBS)

    def schoolbook(self, q, a, b):
        n = a.shape[0]
        c = np.zeros(n * LIMBS, np.uint32)
        assert self.lib.orc_schoolbook(n, self._u32(q), self._u32(a).ravel(), self._u32(b).ravel(), c) == 0
        return c.reshape(n, LIMBS)

    def primes_reference(self, k=57):
        out = np.zeros(64, np.uint32)
        assert self.lib.orc_primes_reference(k, out) == 0
        return out[:k].copy()

    def primes_general(self, max_n, q):
        out = np.zeros(64, np.uint32)
        k = self.lib.orc_primes_general(max_n, self._u32(q), out)
        return out[:k].copy()

    def product_bits(self, primes):
        return self.lib.orc_product_bits(len(primes), self._u32(primes))

    def find_psi(self, p, n):
        return self.lib.orc_find_psi(p, n)

    def pow_mod(self, a, e, p):
        return self.lib.orc_pow_mod(a, e, p)

    def residues(self, x, primes):
        n = x.shape[0]
        out = np.zeros(len(primes) * n, np.uint32)
        self.lib.orc_residues(n, self._u32(x).ravel(), len(primes), self._u32(primes), out)
        return out.reshape(len(primes), n)

    def ntt_forward(self, p, psi, x):
        x = self._u32(x).copy()
        self.lib.orc_ntt_forward(p, psi, int(x.size).bit_length() - 1, x)
        return x

    def ntt_inverse(self, p, psi, x):
        x = self._u32(x).copy()
        self.lib.orc_ntt_inverse(p, psi, int(x.size).bit_length() - 1, x)
        return x

    def crt(self, q, primes, residues):
        k, n = residues.shape
        c = np.zeros(n * LIMBS, np.uint32)
        assert self.lib.orc_crt(n, self._u32(q), k, self._u32(primes), self._u32(residues).ravel(), c) == 0
        return c.reshape(n, LIMBS)

    def crt_constants(self, q, primes):
        k = len(primes)
        bases = np.zeros(k * LIMBS, np.uint32)
        inv = np.zeros(k, np.uint32)
        pmq = np.zeros(LIMBS, np.uint32)
        assert self.lib.orc_crt_constants(self._u32(q), k, self._u32(primes), bases, inv, pmq) == 0
        return bases.reshape(k, LIMBS), inv, pmq

    def quartic_delta(self, q):
        d = ctypes.c_uint32()
        return d.value if self.lib.orc_quartic_delta(self._u32(q), ctypes.byref(d)) == 0 else None

    def quartic_primes_reference(self, n, delta):
        """(primes, product bits, bound bits) by the reference's quartic rule."""
        out = np.zeros(32, np.uint32)
        pb, bb = ctypes.c_uint32(), ctypes.c_uint32()
        k = self.lib.orc_quartic_primes_reference(n, delta, out, ctypes.byref(pb), ctypes.byref(bb))
        return out[:k].copy(), pb.value, bb.value

    def quartic_multiply(self, q, a, b, primes):
        n = a.shape[0]
        c = np.zeros(n * LIMBS, np.uint32)
        rc = self.lib.orc_quartic_multiply(n, self._u32(q), len(primes), self._u32(primes),
                                           self._u32(a).ravel(), self._u32(b).ravel(), c)
        assert rc == 0, rc
        return c.reshape(n, LIMBS)

    def rns_multiply(self, q, a, b, primes):
        n = a.shape[0]
        c = np.zeros(n * LIMBS, np.uint32)
        rc = self.lib.orc_rns_multiply(n, self._u32(q), len(primes), self._u32(primes),
                                       self._u32(a).ravel(), self._u32(b).ravel(), c)
        assert rc == 0, rc
        return c.reshape(n, LIMBS)

    def rns_multiply_threads(self, q, a, b, primes, threads):
        """rns_multiply with its per-prime and per-coefficient-block units on
        `threads` threads (word-for-word the same result)."""
        n = a.shape[0]
        c = np.zeros(n * LIMBS, np.uint32)
        rc = self.lib.orc_rns_multiply_threads(n, self._u32(q), len(primes), self._u32(primes),
                                               self._u32(a).ravel(), self._u32(b).ravel(), c, threads)
        assert rc == 0, rc
        return c.reshape(n, LIMBS)

    def reference_primes_for(self, n, q):
        """The reference's own prime set when it covers N, else the general rule."""
        ref = self.primes_reference(57)
        if n <= 32768 and self.lib.orc_crt_bound_ok(n, self._u32(q), len(ref), ref):
            return ref
        return self.primes_general(n, q)
