"""The oracle's parameter math against the reference's own rns::setup
(rns/host_setup.h:61-126), compiled from the reference tree by oracle/Makefile
into oracle/_ref/ref_setup.  Skipped where that driver was not built (the
reference tree is not present on every machine)."""
import json
import os
import subprocess

import numpy as np
import pytest

from conftest import ROOT

REF = os.path.join(ROOT, "oracle", "_ref", "ref_setup")
pytestmark = pytest.mark.skipif(not os.path.exists(REF), reason="oracle/_ref not built here")


def qhex(q):
    return "".join(f"{int(w):08x}" for w in reversed(q))


@pytest.mark.parametrize("modulus", ["delta:11993087", "seed:1"])
def test_setup_constants_match_reference(orc, workload, modulus):
    kind, v = modulus.split(":")
    q = workload.special_modulus(int(v)) if kind == "delta" else workload.modulus(int(v))
    n = 4096
    out = subprocess.run([REF, str(n), qhex(q)], check=True, capture_output=True, text=True, timeout=300)
    ref = json.loads(out.stdout)
    primes = orc.primes_reference(57)
    assert ref["primes"] == [int(p) for p in primes]
    assert ref["reciprocal"] == [(1 << 64) // int(p) for p in primes]
    assert ref["psi"] == [orc.find_psi(int(p), n) for p in primes]
    bases, inv, pmq = orc.crt_constants(q, primes)
    assert ref["bases_mod_q"] == [[int(w) for w in row] for row in bases]
    assert ref["product_mod_q"] == [int(w) for w in pmq]
    # scale[prime][0] = N^-1 * (M_i mod p_i)^-1 (rns/host_setup.h:111)
    assert ref["scale0"] == [pow(n, int(p) - 2, int(p)) * int(i) % int(p) for p, i in zip(primes, inv)]
