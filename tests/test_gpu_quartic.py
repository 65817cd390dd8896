"""GPU parity of the quartic backend (q = 2^868 - delta): the conversion
stage against a Python model of the channels, the whole product against the
GMP Kronecker fixtures of the reference's special modulus (every case
family at N = 32, the N = 2^15 cases) on the host-buffer, device and batched
paths, against the general backend on uneven batches up to N = 2^17, and the
serving loop."""
import hashlib
import json
import os

import numpy as np
import pytest
import torch

from conftest import ROOT

pytestmark = pytest.mark.gpu
GOLDEN = json.load(open(os.path.join(ROOT, "tests", "golden", "golden.json")))
DELTA = 11993087


def sha(x):
    return hashlib.sha256(np.ascontiguousarray(x, dtype="<u4").tobytes()).hexdigest()


def say(*a):
    print(*a, flush=True)


def to_dev(x):
    return torch.from_numpy(np.ascontiguousarray(x, dtype=np.uint32).view(np.int32)).cuda()


def test_quartic_refuses_general_modulus(pm, workload):
    ctx = pm.Context(workload.modulus(1), 1024)
    with pytest.raises(pm.PolymulError) as err:
        ctx.set_backend("quartic")
    assert "quartic" in str(err.value)
    assert ctx.backend()[0] == "general"
    ctx.close()


def test_quartic_conversion_stage(pm, workload):
    """Channel j of prime i at row 4i + j holds sum_m a_m (r i^j)^m mod p for
    the 217-bit components a_m (closed form, Python integers)."""
    n = 64
    q = workload.special_modulus(DELTA)
    ctx = pm.Context(q, 1 << 15)
    ctx.set_backend("quartic")
    k, primes, roots, _, iotas = pm.plan_quartic(q, 1 << 15, iotas=True)
    assert ctx.backend() == ("quartic", 4 * k)
    for family in ("dense-random-1", "all-q-minus-one", "limb-boundaries"):
        say("quartic convert", family)
        a, b = workload.family(family, n, q, 3)
        ra = torch.zeros(4 * k * n, dtype=torch.int32, device="cuda")
        rb = torch.zeros_like(ra)
        ctx.stage_convert(n, to_dev(a), to_dev(b), ra, rb)
        ctx.synchronize()
        got = ra.cpu().numpy().view(np.uint32).reshape(4 * k, n)
        for m in range(n):
            x = workload.to_int(a[m])
            parts = [(x >> (217 * j)) & ((1 << 217) - 1) for j in range(4)]
            for i in range(k):
                p, r, io = int(primes[i]), int(roots[i]), int(iotas[i])
                for ch in range(4):
                    y = r * pow(io, ch, p) % p
                    want = sum(parts[j] * pow(y, j, p) for j in range(4)) % p
                    assert int(got[4 * i + ch, m]) == want, (family, m, i, ch)
    ctx.close()


@pytest.mark.parametrize("n", [32, 32768])
def test_quartic_products_vs_golden(pm, workload, n):
    """Host-buffer (pm_multiply), device and batched products of the quartic
    backend against the GMP fixtures of q = 2^868 - 11993087."""
    q = workload.special_modulus(DELTA)
    ctx = pm.Context(q, 1 << 15)
    ctx.set_backend("quartic")
    cases = [c for c in GOLDEN["cases"] if c["n"] == n and c["modulus"] == f"delta:{DELTA}"]
    assert cases
    ops = [workload.family(c["family"], n, q, c["seed"]) for c in cases]
    batch = pm.Batch(ctx, [n] * len(cases))
    for i, (a, b) in enumerate(ops):
        batch.upload(i, a, b)
    batch.run()
    batch.wait()
    for i, (case, (a, b)) in enumerate(zip(cases, ops)):
        say("quartic golden", n, case["family"])
        assert sha(ctx.multiply(a, b)) == case["sha256"]["c"], ("host", case["family"])
        da, db = to_dev(a), to_dev(b)
        dc = torch.zeros_like(da)
        ctx.multiply_device(n, da, db, dc)
        ctx.synchronize()
        assert sha(dc.cpu().numpy().view(np.uint32)) == case["sha256"]["c"], ("device", case["family"])
        assert sha(batch.download(i)) == case["sha256"]["c"], ("batch", case["family"])
    batch.close()
    ctx.close()


def test_quartic_batch_matches_general(pm, workload):
    """Uneven sizes 2^5..2^17 (outer stages as passes of their own above the
    2^15 tile) and the largest delta: the quartic batch equals the general
    backend's batch word for word, on two replays."""
    for delta in (DELTA, (1 << 28) - 1):
        q = workload.special_modulus(delta)
        ns = [131072, 32, 4096, 32768, 65536, 1024, 16384, 256]
        say("quartic vs general batch", delta, ns)
        ctx = pm.Context(q, 1 << 17)
        ops = [workload.family("dense-random-1" if i % 3 else "all-q-minus-one", n, q, 900 + i)
               for i, n in enumerate(ns)]
        outs = {}
        for backend in ("general", "quartic"):
            ctx.set_backend(backend)
            batch = pm.Batch(ctx, ns)
            for i, (a, b) in enumerate(ops):
                batch.upload(i, a, b)
            batch.run()
            batch.wait()
            first = [batch.download(i) for i in range(len(ns))]
            batch.run()
            batch.wait()
            assert all((batch.download(i) == first[i]).all() for i in range(len(ns))), backend
            outs[backend] = first
            batch.close()
        for i, n in enumerate(ns):
            assert (outs["quartic"][i] == outs["general"][i]).all(), (delta, n)
        # the quartic schedule kept apart from the general one: 2^13 tiles
        # and two outer passes at 2^15 on the quartic backend only
        ctx.set_backend("quartic")
        ctx.set_batch_outer(32768, 2)
        assert ctx.get_batch_outer(32768) == 2
        ctx.set_backend("general")
        assert ctx.get_batch_outer(32768) == 2  # the general default, a 2^13 tile too
        ctx.set_batch_outer(32768, 1)
        ctx.set_backend("quartic")
        assert ctx.get_batch_outer(32768) == 2
        batch = pm.Batch(ctx, ns)
        for i, (a, b) in enumerate(ops):
            batch.upload(i, a, b)
        batch.run()
        batch.wait()
        assert all((batch.download(i) == outs["general"][i]).all() for i in range(len(ns))), "quartic outer 2"
        batch.close()
        # tiles at 16 residues per thread run the quartic channel rows too
        for tile_n in (256, 1024, 4096, 8192, 16384):
            ctx.set_batch_residues_per_thread(tile_n, 16)
        batch = pm.Batch(ctx, ns)
        for i, (a, b) in enumerate(ops):
            batch.upload(i, a, b)
        batch.run()
        batch.wait()
        assert all((batch.download(i) == outs["general"][i]).all() for i in range(len(ns))), "quartic 16 per thread"
        batch.close()
        ctx.close()


def test_quartic_random_deltas(pm, orc, workload):
    """Deltas other than the reference's (1, 3, five random ones below
    2^28): each a plan of its own primes; quartic products against schoolbook
    at N = 4 and 64 and against the oracle's quartic restatement and the
    general backend at N = 4096."""
    import random
    rng = random.Random(2868)
    deltas = [1, 3] + [rng.randrange(1, 1 << 28) | 1 for _ in range(5)]
    for j, delta in enumerate(deltas):
        q = workload.special_modulus(delta)
        ctx = pm.Context(q, 4096)
        try:
            ctx.set_backend("quartic")
            primes = pm.plan_quartic(q, 4096)[1]
            for i, n in enumerate((4, 64, 4096)):
                family = ("dense-random-1", "all-q-minus-one", "limb-boundaries", "wrap-sign")[(j + i) % 4]
                say("quartic delta", delta, n, family)
                a, b = workload.family(family, n, q, 60 + j)
                got = ctx.multiply(a, b)
                if n <= 64:
                    assert (got == orc.schoolbook(q, a, b)).all(), (delta, n, family)
                else:
                    assert (got == orc.quartic_multiply(q, a, b, primes)).all(), (delta, n, family)
                    ctx.set_backend("general")
                    assert (ctx.multiply(a, b) == got).all(), (delta, n, family, "general")
                    ctx.set_backend("quartic")
        finally:
            ctx.close()


def test_quartic_serving(pm, workload):
    q = workload.special_modulus(DELTA)
    ctx = pm.Context(q, 1 << 14)
    ops = [workload.family("dense-random-1", n, q, 40 + i) for i, n in enumerate([1024, 16384, 64, 4096, 8192])]
    want = [ctx.multiply(a, b) for a, b in ops]  # general backend
    ctx.set_backend("quartic")
    server = pm.Server(ctx, 1 << 30, max_batch=4, max_inflight=2)
    ids = [server.submit(a, b) for a, b in ops]
    for rid, w in zip(ids, want):
        c, _ = server.wait(rid)
        assert (c == w).all()
    server.close()
    ctx.close()


def test_autotune_backend(pm, workload, tmp_path):
    """The backend autotuner times both backends at the special modulus,
    keeps one, writes it to the schedule file, and a fresh context that loads
    the file runs that backend with unchanged products; at a general modulus
    it keeps the general backend."""
    q = workload.special_modulus(DELTA)
    n = 1 << 12
    ctx = pm.Context(q, n)
    a, b = workload.family("dense-random-1", n, q, 3)
    want = ctx.multiply(a, b)
    path = str(tmp_path / "schedule.json")
    got = pm.autotune_backend(ctx, n, count=4, reps=2, path=path)
    say("autotune", got)
    assert set(got["ms"]) == {"general", "quartic"} and got["backend"] in got["ms"]
    assert ctx.backend()[0] == got["backend"]
    ctx.close()
    ctx = pm.Context(q, n)
    assert ctx.load_schedule(path)["backend"] == got["backend"]
    assert ctx.backend()[0] == got["backend"]
    assert (ctx.multiply(a, b) == want).all()
    ctx.close()
    ctx = pm.Context(workload.modulus(1), n)
    assert pm.autotune_backend(ctx, n, count=2, reps=1)["backend"] == "general"
    ctx.close()
