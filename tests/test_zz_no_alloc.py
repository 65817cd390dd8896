"""The allocation-free entry point pm_multiply_device_noalloc, called through
the C-ABI: the same words as pm_multiply_device (which uses the context's
workspace) on two calls in a row, the second smaller, with the static scratch
holding the first call's residues; sizes beyond the scratch are refused."""
import ctypes

import numpy as np
import pytest
import torch

from conftest import load_package

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


def test_noalloc_matches_device_product(workload):
    pm = load_package()
    lib = pm.library()
    q = workload.modulus(31)
    ctx = pm.Context(q, 1 << 16)
    stream = torch.cuda.Stream()
    for n, seed in ((32768, 1), (1024, 2)):  # second call smaller: stale residues must not matter
        say("noalloc", n)
        a, b = workload.family("dense-random-1", n, q, seed)
        da = torch.from_numpy(a.view(np.int32)).cuda()
        db = torch.from_numpy(b.view(np.int32)).cuda()
        c_ref = torch.zeros_like(da)
        c_new = torch.full_like(da, -1)
        torch.cuda.synchronize()
        st = lib.pm_multiply_device(ctx.handle, n, da.data_ptr(), db.data_ptr(), c_ref.data_ptr(),
                                    stream.cuda_stream)
        assert st == pm.PM_OK
        st = lib.pm_multiply_device_noalloc(ctx.handle, n, da.data_ptr(), db.data_ptr(), c_new.data_ptr(),
                                            stream.cuda_stream)
        assert st == pm.PM_OK, lib.pm_last_error(ctx.handle)
        stream.synchronize()
        got = c_new.cpu().numpy().view(np.uint32)
        want = c_ref.cpu().numpy().view(np.uint32)
        assert np.array_equal(got, want), n
    # 2 x 59 primes x 2^16 residues exceed the 16 MiB scratch: refused, nothing launched
    n = 1 << 16
    buf = torch.zeros(n * 28, dtype=torch.int32, device="cuda")
    st = lib.pm_multiply_device_noalloc(ctx.handle, n, buf.data_ptr(), buf.data_ptr(), buf.data_ptr(), None)
    assert st == pm.PM_EUNSUPPORTED
    assert b"static scratch" in lib.pm_last_error(ctx.handle)
    ctx.close()


def test_noalloc_quartic_backend(workload):
    """The same check with the special-modulus backend selected (64 channel
    rows of residues per operand: 2 x 64 x 2^15 words fill the scratch
    exactly), two calls in a row, the second smaller."""
    pm = load_package()
    lib = pm.library()
    q = workload.special_modulus(11993087)
    ctx = pm.Context(q, 1 << 15)
    ctx.set_backend("quartic")
    stream = torch.cuda.Stream()
    for n, seed in ((32768, 3), (2048, 4)):
        say("noalloc quartic", n)
        a, b = workload.family("all-q-minus-one" if n == 2048 else "dense-random-2", n, q, seed)
        da = torch.from_numpy(a.view(np.int32)).cuda()
        db = torch.from_numpy(b.view(np.int32)).cuda()
        c_ref = torch.zeros_like(da)
        c_new = torch.full_like(da, -1)
        torch.cuda.synchronize()
        st = lib.pm_multiply_device(ctx.handle, n, da.data_ptr(), db.data_ptr(), c_ref.data_ptr(),
                                    stream.cuda_stream)
        assert st == pm.PM_OK
        st = lib.pm_multiply_device_noalloc(ctx.handle, n, da.data_ptr(), db.data_ptr(), c_new.data_ptr(),
                                            stream.cuda_stream)
        assert st == pm.PM_OK, lib.pm_last_error(ctx.handle)
        stream.synchronize()
        assert np.array_equal(c_new.cpu().numpy().view(np.uint32), c_ref.cpu().numpy().view(np.uint32)), n
    ctx.close()
