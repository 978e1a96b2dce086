"""KL(P||Q) as a torch.autograd.Function: backward = 4 x the engine's force
kernel, checked against central finite differences of the FP64 KL and
against the oracle's gradient at a fixed Y."""
import numpy as np
import pytest

from tests.conftest import rng

pytestmark = pytest.mark.gpu


def mixture(n, d, k=5, seed=0):
    g = rng(seed)
    c = g.normal(0, 5, size=(k, d))
    return c[np.arange(n) % k] + g.normal(size=(n, d))


@pytest.fixture(scope="module")
def problem(oracle):
    X = mixture(500, 16)
    row, col, val, _ = oracle.build_p(X, 12.0, knn_mode=1)
    Y = rng(2).normal(size=(500, 2)) * 2.0
    return row, col, val, Y


def test_fp64_gradient_matches_central_differences(engine, problem):
    row, col, val, Y = problem
    dY, _ = engine.gradient(Y, row, col, val, 0.0, precision="fp64")
    h = 1e-6  # FP64 central differences
    for (i, d) in [(0, 0), (17, 1), (250, 0), (499, 1)]:
        Yp, Ym = Y.copy(), Y.copy()
        Yp[i, d] += h
        Ym[i, d] -= h
        fd = (engine.kl(Yp, row, col, val, mode="exact") - engine.kl(Ym, row, col, val, mode="exact")) / (2 * h)
        assert abs(fd - 4 * dY[i, d]) < 1e-6 * max(1.0, abs(fd)), (i, d, fd, 4 * dY[i, d])


@pytest.mark.parametrize("theta", [0.0, 0.5])
def test_autograd_backward_is_4x_reference_gradient(engine, oracle, problem, theta):
    import torch

    from tapkee_amd.autograd import kl_divergence
    row, col, val, Y = problem
    Y = Y - Y.mean(0)
    dev = torch.device("cuda:0")
    Yt = torch.tensor(Y, dtype=torch.float32, device=dev, requires_grad=True)
    r = torch.tensor(row, device=dev)
    c = torch.tensor(col, device=dev)
    v = torch.tensor(val, dtype=torch.float32, device=dev)
    loss = kl_divergence(Yt, r, c, v, theta)
    loss.backward()
    g = Yt.grad.double().cpu().numpy()
    if theta == 0.0:
        ref, _ = oracle.gradient_exact_sparse(row, col, val, Y)
        kl_ref = oracle.kl_exact_sparse(row, col, val, Y)
        assert abs(loss.item() - kl_ref) < 1e-4 * abs(kl_ref)
    else:
        ref, _ = oracle.gradient_bh(row, col, val, Y, theta)
    err = np.linalg.norm(g - 4 * ref) / np.linalg.norm(4 * ref)
    assert err < 1e-4, err


def test_autograd_drives_an_optimiser(engine):
    import torch

    from tapkee_amd.autograd import TSNELoss
    X = mixture(800, 20)
    loss_fn = TSNELoss(X, perplexity=15.0, theta=0.5, dist_dtype="fp32")
    torch.manual_seed(0)
    Y = torch.nn.Parameter(torch.randn(800, 2, device="cuda") * 1e-2)
    opt = torch.optim.Adam([Y], lr=0.1)
    first = None
    for _ in range(150):
        opt.zero_grad()
        loss = loss_fn(Y)
        loss.backward()
        opt.step()
        first = loss.item() if first is None else first
    assert loss.item() < 0.5 * first


def test_gradient_at_workload5_size(engine, oracle):
    """N=10,000, D=784 mixture (workload 5): the engine's FP64 exact
    gradient against the oracle's, against central differences of the exact
    KL (h=1e-6), and the Barnes-Hut gradient against the oracle's pointer
    quadtree at theta=0.5."""
    from tapkee_amd import datasets
    X, _ = datasets.autograd_mixture()
    X = X.astype(np.float64)
    N = X.shape[0]
    row, col, val = engine.build_p(X, 30.0)
    Y = rng(3).normal(size=(N, 2)) * 3.0
    Y -= Y.mean(0)
    dY, z = engine.gradient(Y, row, col, val, 0.0, precision="fp64")
    ref, zr = oracle.gradient_exact_sparse(row, col, val, Y)
    assert np.linalg.norm(dY - ref) / np.linalg.norm(ref) < 1e-10
    assert abs(z - zr) / zr < 1e-12
    h = 1e-6
    for (i, d) in [(0, 0), (4321, 1), (9999, 0)]:
        Yp, Ym = Y.copy(), Y.copy()
        Yp[i, d] += h
        Ym[i, d] -= h
        fd = (engine.kl(Yp, row, col, val, mode="exact") - engine.kl(Ym, row, col, val, mode="exact")) / (2 * h)
        assert abs(fd - 4 * dY[i, d]) < 1e-6 * max(1.0, abs(fd)), (i, d, fd, 4 * dY[i, d])
    gb, _ = engine.gradient(Y, row, col, val, 0.5, precision="fp64")
    rb, _ = oracle.gradient_bh(row, col, val, Y, 0.5)
    assert np.linalg.norm(gb - rb) / np.linalg.norm(rb) < 1e-9


def test_autograd_fp64_gradcheck(engine):
    """torch.autograd.gradcheck of the FP64 Function at theta=0 (central
    differences of the forward against the backward, every coordinate)."""
    import torch

    from tapkee_amd.autograd import kl_divergence
    X = mixture(48, 6)
    row, col, val = engine.build_p(X, 5.0)
    dev = torch.device("cuda:0")
    Y = torch.tensor(rng(7).normal(size=(48, 2)) * 2.0, dtype=torch.float64, device=dev, requires_grad=True)
    r, c = torch.tensor(row, device=dev), torch.tensor(col, device=dev)
    v = torch.tensor(val, dtype=torch.float64, device=dev)
    assert torch.autograd.gradcheck(lambda y: kl_divergence(y, r, c, v, 0.0), (Y,), eps=1e-6, atol=1e-7, rtol=1e-5)


def test_autograd_fp64_at_workload5_size(engine, oracle):
    """N=10,000, D=784 (workload 5) through the FP64 Function: the backward
    against 4x the oracle's exact gradient, and against central differences
    (h=1e-6) of the Function's own forward; the forward against the
    oracle's KL."""
    import torch

    from tapkee_amd import datasets
    from tapkee_amd.autograd import kl_divergence
    X, _ = datasets.autograd_mixture()
    X = X.astype(np.float64)
    N = X.shape[0]
    row, col, val = engine.build_p(X, 30.0)
    Yh = rng(3).normal(size=(N, 2)) * 3.0
    dev = torch.device("cuda:0")
    r, c = torch.tensor(row, device=dev), torch.tensor(col, device=dev)
    v = torch.tensor(val, dtype=torch.float64, device=dev)
    Y = torch.tensor(Yh, dtype=torch.float64, device=dev, requires_grad=True)
    loss = kl_divergence(Y, r, c, v, 0.0)
    loss.backward()
    g = Y.grad.cpu().numpy()
    ref, _ = oracle.gradient_exact_sparse(row, col, val, Yh)
    assert np.linalg.norm(g - 4 * ref) / np.linalg.norm(4 * ref) < 1e-10
    assert abs(loss.item() - oracle.kl_exact_sparse(row, col, val, Yh)) < 1e-10 * abs(loss.item())
    h = 1e-6
    with torch.no_grad():
        for (i, d) in [(0, 0), (4321, 1), (9999, 0)]:
            Yp, Ym = Y.detach().clone(), Y.detach().clone()
            Yp[i, d] += h
            Ym[i, d] -= h
            fd = (kl_divergence(Yp, r, c, v, 0.0).item() - kl_divergence(Ym, r, c, v, 0.0).item()) / (2 * h)
            assert abs(fd - g[i, d]) < 1e-6 * max(1.0, abs(fd)), (i, d, fd, g[i, d])
