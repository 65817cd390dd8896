"""Case/wire tooling (tests/validate_cases.py, role of the reference's
tests/validate.py): the validator's modulus rule and case families, point
directories in main.cpp's format with exact expected products, and (GPU) the
`solution` harness run over them and verified byte for byte."""
import json
import os
import random

import numpy as np
import pytest

import validate_cases as vc


def test_probable_prime_against_sympy():
    sympy = pytest.importorskip("sympy")
    rng = random.Random(5)
    for _ in range(200):
        x = rng.randrange(1 << 100) | 1
        assert vc.probable_prime(x) == sympy.isprime(x), x
    # Carmichael numbers and prime squares are composite
    for x in (561, 1105, 1729, 2465, 41041, 825265, 7919 * 7919):
        assert not vc.probable_prime(x)
    assert vc.probable_prime((1 << 127) - 1) and not vc.probable_prime((1 << 128) - 1)


def test_validator_modulus_rule():
    sympy = pytest.importorskip("sympy")
    for n in (2, 32, 1024):
        q = vc.validator_modulus(n)
        assert q.bit_length() == 868 and (q - 1) % (2 * n) == 0
        assert sympy.isprime(q)
        # largest such prime: no prime 2^868 - 2N*j + 1 above it
        x = (1 << 868) - 2 * n + 1
        while x > q:
            assert not sympy.isprime(x)
            x -= 2 * n
    assert vc.validator_modulus(32, 11993087) == (1 << 868) - 11993087


def test_family_streams_and_shapes():
    q = vc.validator_modulus(8, 11993087)
    fam = vc.family_inputs(8, q)
    assert [name for name, _, _ in fam] == vc.FAMILIES
    rng = random.Random(20260916 + 8)
    a = [rng.randrange(q) for _ in range(8)]
    b = [rng.randrange(q) for _ in range(8)]
    d = dict((name, (x, y)) for name, x, y in fam)
    assert d["dense-random-1"] == (a, b)
    assert d["random-after-edges"] == (b, a)
    assert d["identity"][1] == [1] + [0] * 7 and d["zero"][1] == [0] * 8
    assert d["wrap-sign"] == ([0] * 7 + [q - 1], [0, q - 1] + [0] * 6)
    limbs = vc.to_limbs([q - 1, 1 << 864])
    assert limbs.shape == (2, 28) and int.from_bytes(limbs[1].astype("<u4").tobytes(), "little") == 1 << 864


def test_make_writes_point_and_exact_products(tmp_path, pm):
    n = 32
    q = vc.make(str(tmp_path), n, log=lambda *_: None)  # checks the oracle against schoolbook for N <= 32
    m = json.loads((tmp_path / "manifest.json").read_text())
    assert m == {"point": {"N": n, "W": 868, "L": 28}, "cases": 9}
    assert json.loads((tmp_path / "labels.json").read_text()) == vc.FAMILIES
    assert int.from_bytes((tmp_path / "point" / "q.bin").read_bytes(), "little") == q
    for i in range(9):
        assert (tmp_path / "cases" / f"{i:06d}" / "a.bin").stat().st_size == n * 112
    # closed forms of three families: zero, identity and wrap-sign (-X^(N-1) * -X = X^N = -1)
    zero = pm.pointdir.read_output(str(tmp_path), 2, n, where="expected")
    ident = pm.pointdir.read_output(str(tmp_path), 3, n, where="expected")
    wrap = pm.pointdir.read_output(str(tmp_path), 4, n, where="expected")
    a, _ = pm.pointdir.read_case(str(tmp_path), 3, n)
    assert not zero.any()
    assert np.array_equal(ident, a)
    want = np.zeros((n, 28), np.uint32)
    want[0] = vc.to_limbs([q - 1])[0]
    assert np.array_equal(wrap, want)


def test_verify_flags_a_wrong_output(tmp_path, pm):
    """verify() compares out/ with expected/: copies of the expected files pass,
    one flipped bit fails that case only, and a missing output fails."""
    n = 16
    vc.make(str(tmp_path), n, 11993087, log=lambda *_: None)
    for i in range(9):
        c = pm.pointdir.read_output(str(tmp_path), i, n, where="expected")
        pm.pointdir.write_output(str(tmp_path), i, c, where="out")
    cases = [{"i": i, "seconds": 0.001, "status": "ok"} for i in range(9)]
    (tmp_path / "out" / "results.json").write_text(json.dumps({"init_s": 0.5, "cases": cases}))
    assert vc.verify(str(tmp_path))["passed"]
    c = pm.pointdir.read_output(str(tmp_path), 5, n)
    c[3, 7] ^= 1
    pm.pointdir.write_output(str(tmp_path), 5, c, where="out")
    rep = vc.verify(str(tmp_path))
    assert not rep["passed"] and [v["passed"] for v in rep["verdicts"]] == [True] * 5 + [False] + [True] * 3
    assert rep["q_delta"] == 11993087
    os.remove(tmp_path / "out" / "000001" / "c.bin")
    assert not vc.verify(str(tmp_path))["verdicts"][1]["passed"]


@pytest.mark.gpu
@pytest.mark.parametrize("n,delta,backend", [(32, None, None), (1024, None, None), (4096, 11993087, None),
                                             (32768, None, None), (32, 11993087, "quartic"),
                                             (4096, 11993087, "quartic"), (32768, 11993087, "quartic")])
def test_solution_passes_validation(tmp_path, n, delta, backend):
    """The drop-in `solution` over the validator's cases: every case exact,
    also with the quartic backend chosen through PM_BACKEND; at N = 2^15 the
    validator's modulus rule gives the reference's published q = 2^868 -
    11993087."""
    print("validate", n, delta, backend, flush=True)
    assert vc.main(["--n", str(n), "--dir", str(tmp_path)] + (["--delta", str(delta)] if delta else [])
                   + (["--backend", backend] if backend else [])) == 0
    rep = json.loads((tmp_path / "validation.json").read_text())
    assert rep["passed"] and len(rep["verdicts"]) == 9
