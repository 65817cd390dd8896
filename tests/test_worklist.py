"""Host checks of the batch work list (no GPU): every conversion block, NTT
unit and CRT block of every product appears exactly once, phases are
contiguous and ordered so each item depends only on earlier ones, and every
index a kernel derives from an item stays inside that product's buffers."""
import numpy as np
import pytest

CONV, OUTER_FWD, TILE, OUTER_INV, CRT = range(5)
BLOCK = 1024       # coefficients of a CRT item
CONV_BLOCK = 2048  # coefficients of a conversion item


def default_outer(logn):
    """Default batch schedule: outer stages down to a 2^13 tile, at most two
    (run inside the conversion and the CRT, so no outer-pass items)."""
    return 0 if logn <= 13 else min(2, logn - 13)


@pytest.mark.parametrize("ns", [
    [32768] * 5,
    [4096, 32768, 1024, 8192, 32, 16384, 2048, 1024, 65536, 131072, 512],
    [32, 64, 128, 256, 512],
    [131072, 131072, 65536],
])
def test_worklist_coverage(pm, ns):
    k = 59
    items, units = pm.plan_worklist(k, ns)
    types = items[:, 0]
    # phases in order, each contiguous
    assert list(types) == sorted(types)
    tiles = items[types == TILE]
    tlogs = tiles[:, 2] & 0xFF
    assert list(tlogs) == sorted(tlogs, reverse=True)
    seen_conv, seen_unit, seen_crt = set(), set(), set()
    for t, j, a0, a1 in items:
        n = ns[j] if t != TILE else None
        if t == CONV:
            assert a0 in (0, 1) and a1 * CONV_BLOCK < n
            seen_conv.add((j, a0, a1))
        elif t in (OUTER_FWD, OUTER_INV):
            raise AssertionError("default outer stages are fused: no separate outer-pass items")
        elif t == CRT:
            assert a0 * BLOCK < n
            seen_crt.add((j, a0))
        elif t == TILE:
            tlog, count = a0 & 0xF, a0 >> 8
            assert 5 <= tlog <= 15 and 1 <= count <= (1 << (15 - tlog))
            for u in units[a1:a1 + count]:
                job, ps = int(u[0]), int(u[1])
                logn = ns[job].bit_length() - 1
                outer = default_outer(logn)
                assert logn - outer == tlog
                prime, s = ps >> 2, ps & 3
                assert prime < k and s < (1 << outer)
                key = (job, prime, s)
                assert key not in seen_unit
                seen_unit.add(key)
    for j, n in enumerate(ns):
        assert all((j, p, b) in seen_conv for p in (0, 1) for b in range(-(-n // CONV_BLOCK)))
        assert all((j, b) in seen_crt for b in range(-(-n // BLOCK)))
        outer = default_outer(n.bit_length() - 1)
        assert all((j, p, s) in seen_unit for p in range(k) for s in range(1 << outer))
    assert len(seen_unit) == sum(k << default_outer(n.bit_length() - 1) for n in ns)


def test_worklist_rejects(pm):
    with pytest.raises(pm.PolymulError):
        pm.plan_worklist(59, [16])
    with pytest.raises(pm.PolymulError):
        pm.plan_worklist(59, [1 << 18])
