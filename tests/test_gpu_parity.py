"""GPU parity: each HIP kernel against the oracle, and the whole product
against the GMP Kronecker fixtures, bit for bit (integer work: exact)."""
import hashlib
import json
import os
import subprocess
import sys

import numpy as np
import pytest
import torch

from conftest import PKG_DIR, ROOT

pytestmark = pytest.mark.gpu
GOLDEN = json.load(open(os.path.join(ROOT, "tests", "golden", "golden.json")))


def sha(x):
    return hashlib.sha256(np.ascontiguousarray(x, dtype="<u4").tobytes()).hexdigest()


def modulus_of(spec, workload):
    kind, value = spec.split(":")
    return workload.modulus(int(value)) if kind == "seed" else workload.special_modulus(int(value))


@pytest.fixture(scope="module")
def torch_cuda():
    assert torch.cuda.is_available(), "GPU tests need a GPU"
    return torch


def to_dev(torch, x):
    return torch.from_numpy(np.ascontiguousarray(x, dtype=np.uint32).view(np.int32)).cuda()


def to_host(t):
    return t.cpu().numpy().view(np.uint32)


_CTX = {}


def context(pm, q, max_n):
    key = (bytes(np.asarray(q, dtype=np.uint32)), max_n)
    if key not in _CTX:
        _CTX[key] = pm.Context(q, max_n)
    return _CTX[key]


def say(*a):
    print(*a, flush=True)


def expected_t(orc, ctx, q, a, b):
    """Scaled inverse-NTT residues t_i = N^-1 (M_i mod p_i)^-1 (a*b mod X^N+1) mod p_i."""
    n = a.shape[0]
    max_n = ctx.info()["max_n"]
    primes = [p for p, _ in ctx.primes()]
    psis = [g for _, g in ctx.primes()]
    _, basis_inv, _ = orc.crt_constants(q, np.array(primes, dtype=np.uint32))
    ra = orc.residues(a, np.array(primes, dtype=np.uint32))
    rb = orc.residues(b, np.array(primes, dtype=np.uint32))
    out = np.zeros((len(primes), n), dtype=np.uint32)
    for i, (p, g) in enumerate(zip(primes, psis)):
        psi_n = pow(int(g), max_n // n, p)
        fa, fb = orc.ntt_forward(p, psi_n, ra[i]), orc.ntt_forward(p, psi_n, rb[i])
        prod = ((fa.astype(np.uint64) * fb.astype(np.uint64)) % p).astype(np.uint32)
        c = orc.ntt_inverse(p, psi_n, prod).astype(np.uint64)
        scale = pow(n, p - 2, p) * int(basis_inv[i]) % p
        out[i] = (c * np.uint64(scale) % np.uint64(p)).astype(np.uint32)
    return out, ra, rb


@pytest.mark.parametrize("n", [32, 1024, 4096])
def test_convert_stage(pm, orc, workload, torch_cuda, n):
    say("convert", n)
    q = workload.modulus(11)
    a, b = workload.family("dense-random-1", n, q, 2)
    a[0] = workload.from_int(workload.to_int(q) - 1)  # largest valid coefficient
    ctx = context(pm, q, 1 << 15)
    primes = np.array([p for p, _ in ctx.primes()], dtype=np.uint32)
    k = len(primes)
    da, db = to_dev(torch_cuda, a), to_dev(torch_cuda, b)
    ra = torch_cuda.zeros(k * n, dtype=torch_cuda.int32, device="cuda")
    rb = torch_cuda.zeros(k * n, dtype=torch_cuda.int32, device="cuda")
    ctx.stage_convert(n, da, db, ra, rb)
    ctx.synchronize()
    ga, gb = to_host(ra).reshape(k, n), to_host(rb).reshape(k, n)
    # the NTT's lazy input range: below 4p (one conditional subtraction)
    assert (ga < 4 * primes[:, None].astype(np.uint64)).all() and (gb < 4 * primes[:, None].astype(np.uint64)).all()
    assert ((ga % primes[:, None]) == orc.residues(a, primes)).all()
    assert ((gb % primes[:, None]) == orc.residues(b, primes)).all()


@pytest.mark.parametrize("n", [32, 4096])
def test_convert_stage_scalar_matches_mfma(pm, orc, workload, torch_cuda, n):
    """Scalar limb reduction and int8 MFMA conversion agree mod p (and with the oracle)."""
    say("convert scalar", n)
    q = workload.modulus(11)
    a, b = workload.family("limb-boundaries", n, q, 2)
    ctx = context(pm, q, 1 << 15)
    primes = np.array([p for p, _ in ctx.primes()], dtype=np.uint32)
    k = len(primes)
    da, db = to_dev(torch_cuda, a), to_dev(torch_cuda, b)
    bufs = [torch_cuda.zeros(k * n, dtype=torch_cuda.int32, device="cuda") for _ in range(4)]
    ctx.stage_convert(n, da, db, bufs[0], bufs[1])
    ctx.stage_convert_scalar(n, da, db, bufs[2], bufs[3])
    ctx.synchronize()
    g = [to_host(x).reshape(k, n) for x in bufs]
    assert all((x < 4 * primes[:, None].astype(np.uint64)).all() for x in g[:2])  # MFMA: below 4p
    assert all((x < 2 * primes[:, None]).all() for x in g[2:])  # scalar limb reduction: below 2p
    assert ((g[2] % primes[:, None]) == orc.residues(a, primes)).all()
    assert ((g[3] % primes[:, None]) == orc.residues(b, primes)).all()
    assert ((g[0] % primes[:, None]) == (g[2] % primes[:, None])).all()
    assert ((g[1] % primes[:, None]) == (g[3] % primes[:, None])).all()


@pytest.mark.parametrize("n,outer", [(32, 0), (1024, 0), (8192, 0), (32768, 0), (32768, 1), (32768, 2)])
def test_ntt_stage(pm, orc, workload, torch_cuda, n, outer):
    say("ntt", n, outer)
    q = workload.modulus(12)
    a, b = workload.family("dense-random-1", n, q, 3)
    ctx = context(pm, q, 1 << 17)
    want, ra, rb = expected_t(orc, ctx, q, a, b)
    dra, drb = to_dev(torch_cuda, ra), to_dev(torch_cuda, rb)
    ctx.stage_ntt(n, dra, drb, outer=outer)
    ctx.synchronize()
    got = to_host(dra).reshape(want.shape)
    assert (got == want).all()


@pytest.mark.parametrize("n", [32, 4096])
def test_crt_stage(pm, orc, workload, torch_cuda, n):
    say("crt", n)
    for family in ("dense-random-1", "all-q-minus-one", "zero", "limb-boundaries"):
        q = workload.modulus(13)
        a, b = workload.family(family, n, q, 4)
        ctx = context(pm, q, 1 << 15)
        t, _, _ = expected_t(orc, ctx, q, a, b)
        c = torch_cuda.zeros(n * 28, dtype=torch_cuda.int32, device="cuda")
        ctx.stage_crt(n, to_dev(torch_cuda, t), c)
        ctx.synchronize()
        want = orc.rns_multiply(q, a, b, orc.reference_primes_for(n, q))
        assert (to_host(c).reshape(n, 28) == want).all(), family


@pytest.mark.parametrize("n", [2, 64, 32768])
def test_crt_mfma_matches_scalar(pm, workload, torch_cuda, n):
    """The int8-MFMA CRT and the vector-ALU CRT give identical words for
    canonical residues: random, all p_i - 1 (largest sums and quotient),
    all zero, and per-coefficient mixtures of both extremes."""
    say("crt mfma vs scalar", n)
    q = workload.modulus(17)
    ctx = context(pm, q, max(n, 1 << 15))
    primes = np.asarray([p for p, _ in ctx.primes()], dtype=np.uint64)
    k = len(primes)
    rng = np.random.default_rng(n)
    cases = {
        "random": (rng.integers(0, 1 << 62, size=(k, n), dtype=np.uint64) % primes[:, None]),
        "max": np.repeat((primes - 1)[:, None], n, axis=1),
        "zero": np.zeros((k, n), dtype=np.uint64),
        "mixed": np.where(rng.integers(0, 2, size=(k, n)) == 1, (primes - 1)[:, None], 0),
    }
    for name, t in cases.items():
        t_dev = to_dev(torch_cuda, np.ascontiguousarray(t.astype(np.uint32)))
        c_mfma = torch_cuda.zeros(n * 28, dtype=torch_cuda.int32, device="cuda")
        c_scalar = torch_cuda.ones(n * 28, dtype=torch_cuda.int32, device="cuda")
        ctx.stage_crt(n, t_dev, c_mfma)
        ctx.stage_crt_scalar(n, t_dev, c_scalar)
        ctx.synchronize()
        assert (to_host(c_mfma) == to_host(c_scalar)).all(), name


def test_crt_reduction_edge_vs_integers(pm, torch_cuda):
    """Residue vectors (tests/golden/crt_barrett_cases.json) whose CRT sum
    for an 865-bit q just above 2^864 sits where a quotient estimate from the
    bits above 2^864 leaves X - Qh q in [2q, 3q): both CRTs must give the
    words of (sum_i t_i M_i - K P) mod q, K the kernel's nearest-integer
    quotient, computed here with Python integers; zeros and random residues
    fill the other coefficients."""
    fx = json.load(open(os.path.join(ROOT, "tests", "golden", "crt_barrett_cases.json")))
    say("crt reduction edge")
    q_words = np.asarray(fx["q_words"], dtype=np.uint32)
    q = sum(int(w) << (32 * i) for i, w in enumerate(fx["q_words"]))
    ctx = context(pm, q_words, fx["max_n"])
    primes = [p for p, _ in ctx.primes()]
    assert primes == fx["primes"]
    k, n = len(primes), 64
    product = 1
    for p in primes:
        product *= p
    rng = np.random.default_rng(5)
    t = np.zeros((k, n), dtype=np.uint64)
    t[:, 32:] = rng.integers(0, 1 << 62, size=(k, n - 32), dtype=np.uint64) % np.asarray(primes, dtype=np.uint64)[:, None]
    for j, case in zip((0, 9, 17, 31), fx["residues"]):
        t[:, j] = case
    want = np.zeros((n, 28), dtype=np.uint32)
    for j in range(n):
        tj = [int(v) for v in t[:, j]]
        est = sum(ti * ((1 << 42) // p) for ti, p in zip(tj, primes))
        K = (est + (1 << 41)) >> 42
        c = (sum(ti * (product // p) for ti, p in zip(tj, primes)) - K * product) % q
        want[j] = [(c >> (32 * w)) & 0xFFFFFFFF for w in range(28)]
    t_dev = to_dev(torch_cuda, np.ascontiguousarray(t.astype(np.uint32)))
    for stage in (ctx.stage_crt, ctx.stage_crt_scalar):
        c_dev = torch_cuda.zeros(n * 28, dtype=torch_cuda.int32, device="cuda")
        stage(n, t_dev, c_dev)
        ctx.synchronize()
        assert (to_host(c_dev).reshape(n, 28) == want).all(), stage.__name__


@pytest.mark.parametrize("n", [32, 64, 1024, 4096, 8192, 16384, 32768, 65536, 131072])
def test_product_vs_golden(pm, workload, n):
    for case in GOLDEN["cases"]:
        if case["n"] != n:
            continue
        say("golden", n, case["modulus"], case["family"])
        q = modulus_of(case["modulus"], workload)
        a, b = workload.family(case["family"], n, q, case["seed"])
        ctx = context(pm, q, max(n, 1 << 15))
        c = ctx.multiply(a, b)
        assert sha(c) == case["sha256"]["c"], (n, case["modulus"], case["family"])


@pytest.mark.parametrize("n", [2, 4, 8, 16, 128, 256, 512, 2048])
def test_product_small_sizes_vs_schoolbook(pm, orc, workload, n):
    say("small", n)
    q = workload.modulus(21)
    ctx = context(pm, q, 1 << 15)
    for family in ("dense-random-1", "all-q-minus-one", "wrap-sign"):
        a, b = workload.family(family, n, q, 9)
        want = orc.schoolbook(q, a, b) if n <= 256 else orc.rns_multiply(q, a, b, orc.primes_reference())
        assert (ctx.multiply(a, b) == want).all(), (n, family)


@pytest.mark.parametrize("n", [64, 4096])
def test_product_865_bit_modulus_vs_oracle(pm, orc, workload, n):
    """An 865-bit q just above 2^864 (the smallest the library accepts)
    through the whole product, against the oracle's RNS port (schoolbook for
    N = 64) on the families with the largest sums."""
    say("865-bit q", n)
    import random
    rng = random.Random(865)
    q = workload.from_int((1 << 864) + rng.getrandbits(850) | 1)
    ctx = context(pm, q, 1 << 15)
    for family in ("dense-random-1", "all-q-minus-one", "limb-boundaries"):
        a, b = workload.family(family, n, q, 5)
        want = orc.schoolbook(q, a, b) if n <= 256 else orc.rns_multiply(q, a, b, orc.primes_reference())
        assert (ctx.multiply(a, b) == want).all(), (n, family)


def test_random_moduli_sweep(pm, orc, workload):
    """Twelve random odd moduli of 865 to 868 bits, each with its own
    context, at sizes 2 to 4096 and a different validator family each, plus
    the all-(q-1) inputs whose sums are largest: against schoolbook up to
    N = 256 and the oracle's RNS port with the reference's 31-bit primes above
    (a prime set unrelated to the library's 30-bit one)."""
    import random
    rng = random.Random(20261204)
    sizes = [2, 8, 32, 64, 256, 1024, 4096]
    for i in range(12):
        bits = 865 + i % 4
        q = workload.from_int((1 << (bits - 1)) | rng.getrandbits(bits - 1) | 1)
        n = sizes[i % len(sizes)]
        first = workload.FAMILIES[i % len(workload.FAMILIES)]
        families = (first, "all-q-minus-one" if first != "all-q-minus-one" else "limb-boundaries")
        say("random q", i, bits, n, families[0])
        ctx = pm.Context(q, 4096)
        try:
            for family in families:
                a, b = workload.family(family, n, q, 100 + i)
                want = orc.schoolbook(q, a, b) if n <= 256 else orc.rns_multiply(q, a, b, orc.primes_reference())
                assert (ctx.multiply(a, b) == want).all(), (i, bits, n, family)
        finally:
            ctx.close()


def test_outer_schedules_agree(pm, workload, torch_cuda):
    say("outer schedules")
    case = [c for c in GOLDEN["cases"] if c["n"] == 32768 and c["family"] == "dense-random-1"][0]
    q = modulus_of(case["modulus"], workload)
    a, b = workload.family("dense-random-1", 32768, q, case["seed"])
    ctx = context(pm, q, 1 << 17)
    k = ctx.info()["k"]
    for outer in (0, 1, 2):
        ra = torch_cuda.zeros(k * 32768, dtype=torch_cuda.int32, device="cuda")
        rb = torch_cuda.zeros_like(ra)
        c = torch_cuda.zeros(32768 * 28, dtype=torch_cuda.int32, device="cuda")
        ctx.stage_convert(32768, to_dev(torch_cuda, a), to_dev(torch_cuda, b), ra, rb)
        ctx.stage_ntt(32768, ra, rb, outer=outer)
        ctx.stage_crt(32768, ra, c)
        ctx.synchronize()
        assert sha(to_host(c)) == case["sha256"]["c"], outer


def test_device_multiply_and_reuse(pm, orc, workload, torch_cuda):
    say("device multiply")
    q = workload.modulus(5)
    ctx = context(pm, q, 1 << 15)
    for seed, n in ((1, 4096), (2, 1024), (3, 4096)):  # state reuse across sizes
        a, b = workload.family("dense-random-1", n, q, seed)
        c = torch_cuda.zeros(n * 28, dtype=torch_cuda.int32, device="cuda")
        ctx.multiply_device(n, to_dev(torch_cuda, a), to_dev(torch_cuda, b), c)
        ctx.synchronize()
        assert (to_host(c).reshape(n, 28) == orc.rns_multiply(q, a, b, orc.primes_reference())).all()


def test_solver_mirror(pm, workload):
    say("solver")
    case = [c for c in GOLDEN["cases"] if c["n"] == 1024 and c["family"] == "limb-boundaries"][0]
    q = modulus_of(case["modulus"], workload)
    a, b = workload.family(case["family"], 1024, q, case["seed"])
    solver = pm.Solver(1024, 868, 28, q)
    assert sha(solver.run(a, b)) == case["sha256"]["c"]
    with pytest.raises(RuntimeError):
        solver.run(a, b[:512])
    solver.free()


def test_solution_harness(workload, tmp_path):
    say("solution harness")
    n = 1024
    cases = [c for c in GOLDEN["cases"] if c["n"] == n][:3]
    q = modulus_of(cases[0]["modulus"], workload)
    (tmp_path / "point").mkdir()
    (tmp_path / "point" / "q.bin").write_bytes(q.astype("<u4").tobytes())
    for i, case in enumerate(cases):
        d = tmp_path / "cases" / f"{i:06d}"
        d.mkdir(parents=True)
        a, b = workload.family(case["family"], n, q, case["seed"])
        (d / "a.bin").write_bytes(a.astype("<u4").tobytes())
        (d / "b.bin").write_bytes(b.astype("<u4").tobytes())
    # a truncated case is reported as crashed, the others still run
    bad = tmp_path / "cases" / f"{len(cases):06d}"
    bad.mkdir(parents=True)
    (bad / "a.bin").write_bytes(b"\0" * 100)
    (bad / "b.bin").write_bytes(b"\0" * 100)
    (tmp_path / "manifest.json").write_text(json.dumps({"point": {"N": n, "W": 868, "L": 28},
                                                        "cases": len(cases) + 1}))
    subprocess.run([os.path.join(PKG_DIR, "solution"), str(tmp_path)], check=True, timeout=120)
    report = json.loads((tmp_path / "out" / "results.json").read_text())
    assert [c["status"] for c in report["cases"]] == ["ok"] * len(cases) + ["crashed"]
    assert report["init_s"] >= 0
    for i, case in enumerate(cases):
        got = (tmp_path / "out" / f"{i:06d}" / "c.bin").read_bytes()
        assert hashlib.sha256(got).hexdigest() == case["sha256"]["c"]


def test_plain_products_on_two_streams_share_workspace_safely(pm, workload, torch_cuda):
    """Plain products enqueued back to back on two streams (and then a
    host-buffer product) share the context's residue workspace: each is
    ordered after the previous one, so every result equals the same product
    computed alone."""
    say("two streams")
    n = 32768
    q = workload.modulus(21)
    ctx = context(pm, q, 1 << 15)
    ops = [workload.family("dense-random-1", n, q, 40 + i) for i in range(3)]
    alone = [ctx.multiply(a, b) for a, b in ops]
    s1, s2 = torch_cuda.cuda.Stream(), torch_cuda.cuda.Stream()
    dev = [(to_dev(torch_cuda, a), to_dev(torch_cuda, b)) for a, b in ops]
    outs = [torch_cuda.zeros(n * 28, dtype=torch_cuda.int32, device="cuda") for _ in ops]
    torch_cuda.cuda.synchronize()
    for rep in range(3):
        for i, ((da, db), dc) in enumerate(zip(dev, outs)):
            ctx.multiply_device(n, da, db, dc, (s1 if i % 2 == 0 else s2).cuda_stream)
        host = ctx.multiply(*ops[1])  # the context's own stream, same workspace
        torch_cuda.cuda.synchronize()
        for i, dc in enumerate(outs):
            assert (to_host(dc).reshape(n, 28) == alone[i]).all(), (rep, i)
        assert (host == alone[1]).all(), rep


def test_fherma_run_input_size():
    """fherma_run refuses a and b that are not exactly N x L words for the
    N of fherma_init ("RNS input size", rns/solve.cu:723-724) and answers the
    next well-formed call (tests/native/fherma_size_check.cpp)."""
    say("fherma_run input size")
    out = subprocess.run([os.path.join(PKG_DIR, "fherma_size_check")], capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stdout + out.stderr
    assert out.stdout.strip().endswith("ok")


def test_plain_schedule_and_autotune(pm, workload, tmp_path):
    """Every plain-path schedule (outer passes 0..2, and on the device path
    conversion slices 1, 2, 4) gives the same product; the autotuner keeps
    the fastest per size and its JSON loads back."""
    q = workload.modulus(23)
    ctx = pm.Context(q, 1 << 15)
    say("schedules")
    for n in (64, 4096, 32768):
        a, b = workload.family("dense-random-1", n, q, 9)
        want = None
        for outer in (0, 1, 2):
            ctx.set_outer(n, outer)
            assert ctx.get_outer(n) == outer
            c = ctx.multiply(a, b)
            want = c if want is None else want
            assert (c == want).all(), (n, outer)
            da, db = to_dev(torch, a), to_dev(torch, b)
            dc = torch.empty_like(da)
            for sl in (1, 2, 4, 0):
                ctx.set_conversion_slices(n, sl)
                assert ctx.get_conversion_slices(n) == sl
                ctx.multiply_device(n, da, db, dc)
                ctx.synchronize()
                assert (dc.cpu().numpy().view(np.uint32).reshape(n, 28) == want).all(), (n, outer, sl)
    with pytest.raises(pm.PolymulError):
        ctx.set_conversion_slices(4096, 3)
    with pytest.raises(pm.PolymulError):
        ctx.set_outer(32768, 3)
    with pytest.raises(pm.PolymulError):
        ctx.set_outer(2, 1)  # would leave a tile of one element
    path = tmp_path / "schedule.json"
    res = pm.autotune_plain(ctx, [1024, 32768], reps=3, path=str(path))
    assert set(res) == {1024, 32768}
    for n, r in res.items():
        assert r["outer"] == ctx.get_outer(n) and r["outer"] in r["ms"]
        assert r["slices"] == ctx.get_conversion_slices(n) and set(r["slices_ms"]) == {1, 2, 4}
    ctx.set_outer(32768, 0 if res[32768]["outer"] else 1)
    ctx.set_conversion_slices(32768, 0)
    applied = ctx.load_schedule(str(path))
    assert applied["outer"][32768] == res[32768]["outer"] == ctx.get_outer(32768)
    assert applied["conversion_slices"][32768] == res[32768]["slices"] == ctx.get_conversion_slices(32768)
    ctx.close()


def test_host_path_chunks_and_split(pm, workload, torch_cuda):
    """The host-buffer path (inputs read from page-locked staging over PCIe,
    CRT chunks written back into it) equals the device-resident path for
    every chunk count, on consecutive products through the same staging; the
    split submit/collect API gives the same product and frees the inputs at
    submit."""
    q = workload.modulus(29)
    ctx = pm.Context(q, 1 << 15)
    for n in (64, 4096, 32768):
        want = {}
        for seed in (1, 2):
            a, b = workload.family("dense-random-1", n, q, seed)
            c = torch_cuda.zeros(n * 28, dtype=torch_cuda.int32, device="cuda")
            ctx.multiply_device(n, to_dev(torch_cuda, a), to_dev(torch_cuda, b), c)
            ctx.synchronize()
            want[seed] = (a, b, to_host(c).reshape(n, 28))
        for chunks in (0, 1, 2, 4, 8, 16):
            if chunks and n // chunks < 64:
                with pytest.raises(pm.PolymulError):
                    ctx.set_io_chunks(n, chunks)
                continue
            say("host path", n, "chunks", chunks)
            ctx.set_io_chunks(n, chunks)
            for seed in (1, 2, 1):  # alternate inputs: no stale staging
                a, b, c = want[seed]
                assert (ctx.multiply(a, b) == c).all(), (n, chunks, seed)
        ctx.set_io_chunks(n, 0)
        a, b, c = want[2]
        a2, b2 = a.copy(), b.copy()
        ctx.submit(a2, b2)
        a2[:] = 0  # inputs may be reused once submit returns
        b2[:] = 0
        assert (ctx.collect() == c).all(), n
    with pytest.raises(pm.PolymulError):
        ctx.set_io_chunks(4096, 3)
    lib = pm.library()
    c = np.empty((64, 28), dtype=np.uint32)
    assert lib.pm_multiply_collect(ctx.handle, c.ctypes.data) == 1  # nothing submitted
    a, b, _ = want[1]
    ctx.submit(a, b)
    assert lib.pm_multiply_submit(ctx.handle, a.shape[0], a.ctypes.data, b.ctypes.data) == 1  # one pending
    assert (ctx.collect() == want[1][2]).all()
    ctx.close()


def test_host_path_stall_times_out_and_recovers(pm, workload, torch_cuda):
    """A host that never raises one input chunk's ready flag (injected from the
    host, so every kernel still ends normally): the device wait for it runs out
    after the wait limit, collect reports PM_ETIMEOUT, and the next products on
    the same context are exact again."""
    q = workload.modulus(29)
    ctx = pm.Context(q, 1 << 15)
    n = 1 << 15
    a, b = workload.family("dense-random-1", n, q, 3)
    ctx.set_io_chunks(n, 8)  # a|b in 8 input chunks
    want = ctx.multiply(a, b)
    with pytest.raises(pm.PolymulError):
        ctx.inject_io_stall(pm.MAX_IO_CHUNKS)
    for stall in (0, 5, 7):
        say("stall input chunk", stall)
        ctx.set_io_wait_limit(2000)  # 2 ms
        ctx.inject_io_stall(stall)
        ctx.submit(a, b)
        with pytest.raises(pm.PolymulError) as err:
            ctx.collect()
        assert err.value.status == pm.PM_ETIMEOUT, err.value
        ctx.set_io_wait_limit(0)  # the default bound again
        assert (ctx.multiply(a, b) == want).all(), stall
        a2, b2 = workload.family("dense-random-1", n, q, 4)
        c2 = torch_cuda.zeros(n * 28, dtype=torch_cuda.int32, device="cuda")
        ctx.multiply_device(n, to_dev(torch_cuda, a2), to_dev(torch_cuda, b2), c2)
        ctx.synchronize()
        assert (ctx.multiply(a2, b2) == to_host(c2).reshape(n, 28)).all(), stall
    ctx.close()
