"""The allocation-free speed pass (d4_speed_pass_dev_noalloc, include/d4.h) through the C ABI against
d4_speed_pass_dev on the same device buffers: the same bits, on two calls in a row (the second one
smaller, so the shared scratch holds the first call's larger contents), and the size limit."""
import numpy as np
import pytest

from helpers import random_config

pytestmark = pytest.mark.gpu


def _call(ctx, fn, xy_t, tau, tol_t=None, cap2=0.0):
    import torch
    n = xy_t.shape[0]
    v = torch.full((n,), float("nan"), dtype=torch.float64, device=xy_t.device)
    ctx.use_torch_stream()
    rc = getattr(ctx.L, fn)(ctx.h, n, xy_t.data_ptr(), float(tau), v.data_ptr(),
                            tol_t.data_ptr() if tol_t is not None else None, float(cap2))
    torch.cuda.synchronize()
    return rc, v.cpu().numpy()


def test_noalloc_speed_pass_equals_speed_pass(ctx):
    import torch
    dev = torch.device("cuda:0")
    cap2 = 1e-4
    for n, seed, tau in ((8192, 5, 0.1), (1000, 6, -0.05), (3, 7, 0.0)):
        print(f"[noalloc] n={n}", flush=True)
        xy = torch.from_numpy(random_config(n, seed)).to(dev)
        t0 = torch.full((n,), cap2, dtype=torch.float64, device=dev)
        t1 = t0.clone()
        rc0, v0 = _call(ctx, "d4_speed_pass_dev", xy, tau, t0, cap2)
        rc1, v1 = _call(ctx, "d4_speed_pass_dev_noalloc", xy, tau, t1, cap2)
        assert rc0 == 0 and rc1 == 0
        np.testing.assert_array_equal(v1, v0)
        assert torch.equal(t0, t1)
        rc2, v2 = _call(ctx, "d4_speed_pass_dev_noalloc", xy, tau)
        assert rc2 == 0
        np.testing.assert_array_equal(v2, v0)


def test_noalloc_speed_pass_rejects_sizes_beyond_its_scratch(ctx):
    import torch
    n = 8193
    xy = torch.zeros((n, 2), dtype=torch.float64, device="cuda:0")
    v = torch.zeros(n, dtype=torch.float64, device="cuda:0")
    rc = ctx.L.d4_speed_pass_dev_noalloc(ctx.h, n, xy.data_ptr(), 0.0, v.data_ptr(), None, 0.0)
    assert rc == -1
    assert "exceeds 8192" in ctx.L.d4_last_error(ctx.h).decode()
    assert torch.count_nonzero(v).item() == 0
