"""C-ABI checks that need no GPU: the library loads, exports every symbol
include/polymul.h declares, plans the RNS on the host, and rejects
out-of-coverage requests with a reason."""
import os
import re
import subprocess

import numpy as np
import pytest

from conftest import PKG_DIR, ROOT


def declared(header):
    text = open(os.path.join(ROOT, "include", header)).read()
    return sorted(set(re.findall(r"^\w[\w\s\*]*?\b(pm_\w+)\s*\(", text, re.M)))


def test_header_symbols_exported(pm):
    lib = pm.library()
    names = declared("polymul.h")
    assert "pm_multiply" in names and "pm_create" in names
    for name in names:
        assert hasattr(lib, name), name
    assert sorted(pm.EXPORTED) == sorted(n for n in names if n in pm.EXPORTED)
    assert set(names) <= set(pm.EXPORTED) | set(getattr(pm, "EXPORTED_EXTRA", []))


def test_fherma_entry_points_exported():
    so = os.path.join(PKG_DIR, "libfherma_polymul.so")
    out = subprocess.run(["nm", "-D", "--defined-only", "-C", so], capture_output=True, text=True, check=True).stdout
    for sig in ["fherma_init(fherma::Point const&)", "fherma_run(void*, fherma::Inputs const&)",
                "fherma_free(void*)"]:
        assert sig in out, sig


def test_plan_matches_oracle_rule(pm, orc, workload):
    for seed in (1, 4, 9):
        q = workload.modulus(seed)
        k, primes, psi, bits = pm.plan_rns(q, 1 << 17)
        assert list(primes) == list(orc.primes_general(1 << 17, q))
        assert bits == orc.product_bits(primes)
        for p, g in zip(primes, psi):  # primitive 2^18-th roots
            p, g = int(p), int(g)
            assert pow(g, 1 << 17, p) == p - 1


@pytest.mark.parametrize("max_n,q_fn,why", [
    (3, lambda w: w.modulus(1), "power-of-two"),
    (1 << 18, lambda w: w.modulus(1), "power-of-two"),
    (1024, lambda w: np.zeros(28, np.uint32), "odd"),
    (1024, lambda w: w.from_int((1 << 800) + 2), "odd"),
    # the CRT's one-step Barrett finish needs q > 2^864 (an 864-bit q gives
    # wrong products in a CPU emulation of that step)
    (1024, lambda w: w.from_int((1 << 863) + 12345), "865 to 868 bits"),
    (1024, lambda w: w.from_int((1 << 99) + 1), "865 to 868 bits"),
])
def test_plan_rejects(pm, workload, max_n, q_fn, why):
    with pytest.raises(pm.PolymulError) as err:
        pm.plan_rns(q_fn(workload), max_n)
    assert why in str(err.value)


def test_plan_rejects_wide_modulus(pm, workload):
    q = np.full(28, 0xFFFFFFFF, dtype=np.uint32)  # 896 bits
    with pytest.raises(pm.PolymulError) as err:
        pm.plan_rns(q, 1024)
    assert "2^868" in str(err.value)


def test_plan_accepts_smallest_modulus(pm, workload):
    k, primes, psi, bits = pm.plan_rns(workload.from_int((1 << 864) + 1), 1024)  # 865 bits
    assert k > 0 and bits > 2 * 865
