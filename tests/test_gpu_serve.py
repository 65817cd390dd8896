"""GPU tests of the continuous-batching server: every result is the same as
the product computed alone (and as the oracle), whatever it was mixed with;
admission respects the HBM budget."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


def request_bytes(n, k):
    al = lambda x: (x + 255) // 256 * 256
    return al(3 * n * 112) + al(2 * k * n * 4)


def test_served_results_equal_alone(pm, orc, workload):
    q = workload.modulus(3)
    ctx = pm.Context(q, 1 << 16)
    rng = np.random.default_rng(5)
    ns = [1 << int(x) for x in rng.integers(10, 17, size=40)]
    say("serve mix", ns)
    reqs = []
    for i, n in enumerate(ns):
        a, b = workload.family("dense-random-1", n, q, 1000 + i)
        reqs.append((n, a, b))
    server = pm.Server(ctx, 4 << 30, max_batch=8, max_inflight=2)
    ids = [server.submit(a, b) for _, a, b in reqs]
    results = [server.wait(r) for r in ids]
    stats = server.stats()
    say("stats", stats)
    server.close()
    assert stats["completed"] == len(ns)
    for i, ((n, a, b), (c, lat)) in enumerate(zip(reqs, results)):
        alone = ctx.multiply(a, b)
        assert (c == alone).all(), (i, n)
        assert lat > 0
    # every served result against the oracle as well (its C code runs outside
    # the GIL, so the 40 products are checked on the host cores in parallel)
    from concurrent.futures import ThreadPoolExecutor

    def oracle_equal(i):
        n, a, b = reqs[i]
        return bool((results[i][0] == orc.rns_multiply(q, a, b, orc.reference_primes_for(n, q))).all())

    with ThreadPoolExecutor(16) as pool:
        agree = list(pool.map(oracle_equal, range(len(reqs))))
    assert all(agree), [i for i, ok in enumerate(agree) if not ok]
    # the same mix served with tiles at 16 residues per thread
    for tile_n in (256, 1024, 2048, 4096, 8192, 16384):
        ctx.set_batch_residues_per_thread(tile_n, 16)
    server = pm.Server(ctx, 4 << 30, max_batch=8, max_inflight=2)
    ids = [server.submit(a, b) for _, a, b in reqs]
    again = [server.wait(r)[0] for r in ids]
    server.close()
    assert all((x == y[0]).all() for x, y in zip(again, results))
    ctx.close()


def test_budget_admission_queues(pm, workload):
    q = workload.modulus(4)
    ctx = pm.Context(q, 1 << 16)
    k = ctx.info()["k"]
    budget = request_bytes(1 << 16, k) + request_bytes(1 << 12, k)
    server = pm.Server(ctx, budget, max_batch=4, max_inflight=2)
    reqs = []
    for i, n in enumerate([1 << 16, 1 << 16, 1 << 12, 1 << 16, 1 << 12]):
        a, b = workload.family("dense-random-1", n, q, 50 + i)
        reqs.append((a, b, server.submit(a, b)))
    for a, b, rid in reqs:
        c, _ = server.wait(rid)
        assert (c == ctx.multiply(a, b)).all()
    stats = server.stats()
    assert stats["peak_bytes"] <= budget
    assert stats["launches"] >= 3  # two N=2^16 products never fit together
    server.close()
    with pytest.raises(pm.PolymulError):
        pm.Server(ctx, request_bytes(1 << 12, k), max_batch=4, max_inflight=1)
    ctx.close()


def test_served_page_locked_outputs(pm, workload):
    """Outputs in page-locked memory come back on the launch stream; mixed
    with pageable outputs in the same launches, every result is the alone one."""
    import torch
    q = workload.modulus(6)
    ctx = pm.Context(q, 1 << 15)
    ns = [1 << 15, 1 << 11, 1 << 13, 1 << 15, 1 << 12, 1 << 14]
    server = pm.Server(ctx, 2 << 30, max_batch=4, max_inflight=2)
    reqs = []
    for i, n in enumerate(ns):
        a, b = workload.family("dense-random-1", n, q, 70 + i)
        out = None
        if i % 2 == 0:
            out = torch.empty(n * 28, dtype=torch.int32).pin_memory().numpy().view(np.uint32).reshape(n, 28)
            out.fill(0xFFFFFFFF)
        reqs.append((a, b, out, server.submit(a, b, out=out)))
    for i, (a, b, out, rid) in enumerate(reqs):
        c, _ = server.wait(rid)
        if out is not None:
            assert c is out
        assert (c == ctx.multiply(a, b)).all(), i
    server.close()
    with pytest.raises(ValueError):
        server2 = pm.Server(ctx, 1 << 30, max_batch=1, max_inflight=1)
        try:
            a, b = workload.family("dense-random-1", 64, q, 1)
            server2.submit(a, b, out=np.empty((64, 27), dtype=np.uint32))
        finally:
            server2.close()
    ctx.close()


def test_served_page_locked_inputs_reused(pm, workload):
    """Page-locked inputs are read by the launch's gather kernel and page-locked
    outputs written by its scatter kernel.  The same page-locked buffers are
    refilled between rounds (a stale read would repeat the last round's
    product), one pair sits 4 bytes past a 16-byte boundary (the word-wise
    copy), and pageable requests run in the same launches."""
    import torch
    q = workload.modulus(8)
    ctx = pm.Context(q, 1 << 15)
    ns = [1 << 15, 1 << 10, 1 << 13, 1 << 12, 1 << 14, 1 << 11]

    def pinned_words(n, offset):
        t = torch.empty(n * 28 + offset, dtype=torch.int32).pin_memory()
        return t.numpy().view(np.uint32)[offset:].reshape(n, 28)

    bufs = [(pinned_words(n, 1 if i == 1 else 0), pinned_words(n, 1 if i == 1 else 0),
             pinned_words(n, 1 if i == 1 else 0)) for i, n in enumerate(ns)]
    server = pm.Server(ctx, 2 << 30, max_batch=4, max_inflight=2)
    for rnd in range(3):
        reqs = []
        for i, n in enumerate(ns):
            a, b = workload.family("dense-random-1", n, q, 300 + 10 * rnd + i)
            if i % 3 == 2:  # pageable inputs and output
                reqs.append((a, b, None, server.submit(a, b)))
                continue
            pa, pb, pc = bufs[i]
            pa[...] = a
            pb[...] = b
            pc.fill(0xFFFFFFFF)
            reqs.append((a, b, pc, server.submit(pa, pb, out=pc)))
        say("round", rnd, "submitted", len(reqs))
        for i, (a, b, out, rid) in enumerate(reqs):
            c, _ = server.wait(rid)
            if out is not None:
                assert c is out
            assert (c == ctx.multiply(a, b)).all(), (rnd, i, ns[i])
    server.close()
    ctx.close()
