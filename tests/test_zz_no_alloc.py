"""tka_gradient_dev_noalloc (the gradient with every buffer taken from a fixed
device region, for callers that cannot allocate) against tka_gradient_dev,
called through the C-ABI on device tensors: dY and Z must agree bit for bit
(the same kernels run in the same order).  Two calls in a row, the second on
fewer points, for each theta: the region keeps the first call's contents,
which the second must not depend on."""
import ctypes as C

import numpy as np
import pytest

from tests.conftest import rng

pytestmark = pytest.mark.gpu


def ring_p(n):
    """A valid CSR P: each point's successor on a ring, weight 1/n."""
    row = np.arange(n + 1, dtype=np.int32)
    col = ((np.arange(n) + 1) % n).astype(np.int32)
    val = np.full(n, 1.0 / n, np.float32)
    return row, col, val


def call(engine, name, Y, row, col, val, theta):
    import torch
    L = engine._lib.lib()
    dY = torch.full_like(Y, float("nan"))
    z = C.c_double(float("nan"))
    stream = torch.cuda.current_stream().cuda_stream
    engine._lib.check(getattr(L, name)(Y.data_ptr(), Y.shape[0], Y.shape[1], row.data_ptr(), col.data_ptr(),
                                       val.data_ptr(), float(theta), dY.data_ptr(), C.byref(z), stream))
    torch.cuda.synchronize()
    return dY.cpu().numpy(), z.value


@pytest.mark.parametrize("theta,sizes", [(0.5, (131072, 6000)), (0.0, (20000, 3000))])
def test_noalloc_gradient_equals_gradient_dev(engine, theta, sizes):
    import torch
    dev = torch.device("cuda:0")
    for k, n in enumerate(sizes):
        print(f"[noalloc] theta={theta} call {k + 1}: N={n}", flush=True)
        row, col, val = (torch.from_numpy(a).to(dev) for a in ring_p(n))
        Y = torch.tensor(rng(10 + k).normal(size=(n, 2)) * 5.0, dtype=torch.float32, device=dev)
        ref, zr = call(engine, "tka_gradient_dev", Y, row, col, val, theta)
        got, zg = call(engine, "tka_gradient_dev_noalloc", Y, row, col, val, theta)
        assert np.all(np.isfinite(ref)) and np.isfinite(zr) and zr > 0
        np.testing.assert_array_equal(got, ref)
        assert zg == zr


def test_noalloc_rejects_more_points_than_its_region(engine):
    import torch
    dev = torch.device("cuda:0")
    n = engine._lib.NOALLOC_MAX_POINTS + 1
    row, col, val = (torch.from_numpy(a).to(dev) for a in ring_p(n))
    Y = torch.zeros((n, 2), dtype=torch.float32, device=dev)
    with pytest.raises(engine._lib.WrongParameterError, match="TKA_NOALLOC_MAX_POINTS"):
        call(engine, "tka_gradient_dev_noalloc", Y, row, col, val, 0.5)
