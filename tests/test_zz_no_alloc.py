"""lg_grand_product_noalloc against lg_grand_product through the C ABI:
two calls in a row (the second smaller, so it runs on what the first left in
the shared scratch array), bit-exact; a size above the scratch limit is
refused with a message."""
import ctypes
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

P = 2**31 - 2**24 + 1


def log(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _e6_cols(rng, n, nonzero=False):
    lo = 1 if nonzero else 0
    return rng.integers(lo, P, size=6 * n, dtype=np.uint64).astype(np.uint32)


def test_grand_product_noalloc_equals_grand_product():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    from loom_amd import _lib
    from loom_amd.prover import Context
    L = _lib.lib()
    ctx = Context(0)
    dev = torch.device("cuda", 0)
    rng = np.random.Generator(np.random.PCG64(77))
    try:
        for n in (1 << 17, 1 << 10):
            log("grand product noalloc: n = %d" % n)
            num = torch.from_numpy(_e6_cols(rng, n).view(np.int32)).to(dev)
            den = torch.from_numpy(_e6_cols(rng, n, nonzero=True).view(np.int32)).to(dev)
            ref = torch.zeros(6 * n, dtype=torch.int32, device=dev)
            got = torch.full((6 * n,), -1, dtype=torch.int32, device=dev)
            torch.cuda.synchronize()
            p = lambda t: ctypes.cast(t.data_ptr(), _lib.u32p)
            _lib.check(L.lg_grand_product(ctx.h, p(num), p(den), p(ref), n), "lg_grand_product")
            _lib.check(L.lg_grand_product_noalloc(ctx.h, p(num), p(den), p(got), n), "lg_grand_product_noalloc")
            a, b = ref.cpu().numpy(), got.cpu().numpy()
            assert np.array_equal(a, b), "n = %d: noalloc result differs" % n
            assert a[0] != 0  # gp[0] = 1 in Montgomery form
        too_big = (1 << 18) * 2
        rc = L.lg_grand_product_noalloc(ctx.h, p(num), p(den), p(got), too_big)
        assert rc == -1
        assert "exceeds the scratch limit" in L.lg_last_error().decode()
    finally:
        ctx.close()
