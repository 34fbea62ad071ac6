"""gjw_normali_noalloc (no allocation, static device scratch) against
gjw_normali through the C ABI: two calls in a row, the second smaller, each
bit for bit equal to the allocating entry point on the same input (the same
kernels and partition); sizes beyond the scratch and host pointers are
refused."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _images(rng, count, m, n):
    a = rng.standard_normal((count, n, m))  # (count, n, m) C-order = column-major images
    a[0] *= 1e-3
    if count > 1:
        a[1, 3, 4] = np.nan        # the guard zeroes this image
    if count > 2:
        a[2] = 0.0                 # a zero maximum is taken as 1
    return a


def _run(fn, ctx, a, shape, guard):
    from gerjoii_amd import _native as N
    import torch
    t = torch.from_numpy(a.copy()).cuda()
    s = torch.from_numpy(shape).cuda() if shape is not None else None
    torch.cuda.synchronize()
    count, n, m = a.shape
    N.check(fn(ctx.h, N.ptr(t), m, n, count, guard, N.ptr(s)), fn.__name__)
    torch.cuda.synchronize()
    return t.cpu().numpy()


def test_noalloc_matches_normali_twice(gpu_ctx):
    from gerjoii_amd import _native as N
    L = N.load()
    rng = np.random.default_rng(7)
    for count, m, n, guard in ((3, 600, 700, 1), (2, 40, 50, 0)):   # second call smaller
        a = _images(rng, count, m, n)
        shape = (rng.uniform(size=(n, m)) > 0.3).astype(np.float64)
        want = _run(L.gjw_normali, gpu_ctx, a, shape, guard)
        got = _run(L.gjw_normali_noalloc, gpu_ctx, a, shape, guard)
        print("noalloc", count, m, n, guard, flush=True)
        np.testing.assert_array_equal(got, want)
        got2 = _run(L.gjw_normali_noalloc, gpu_ctx, a, None, guard)
        np.testing.assert_array_equal(got2, _run(L.gjw_normali, gpu_ctx, a, None, guard))


def test_noalloc_refuses_oversize_and_host_memory(gpu_ctx):
    import torch
    from gerjoii_amd import _native as N
    L = N.load()
    t = torch.zeros(16, dtype=torch.float64, device="cuda")
    # 300 images of a million cells need 300 x 256 partials > 65536
    rc = L.gjw_normali_noalloc(gpu_ctx.h, N.ptr(t), 1000, 1000, 300, 0, None)
    assert rc != 0 and b"scratch" in L.gjw_last_error()
    h = np.zeros((2, 4, 4))
    rc = L.gjw_normali_noalloc(gpu_ctx.h, N.ptr(h), 4, 4, 2, 0, None)
    assert rc != 0 and b"device memory" in L.gjw_last_error()
