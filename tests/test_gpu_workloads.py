"""Workloads 1 and 2 at their specified sizes.

The oracle cannot run 1000 iterations at these sizes in seconds, so each
test pins what it can against the oracle (P, the gradient at a fixed Y) and
checks the full run through size-independent properties (finite, zero-mean
output, decreasing KL, certified kNN, symmetric P summing to one).
"""
import numpy as np
import pytest

from tests.conftest import rng

pytestmark = pytest.mark.gpu


def test_workload1_exact_dense(engine, oracle):
    """N=2000, D=50 mixture of 10 Gaussians, perplexity 30, theta=0 with the
    reference's dense P: dense P and the exact gradient at a fixed Y against
    the oracle, then the full 1000-iteration run."""
    from tapkee_amd import datasets
    X, _ = datasets.gaussian_mixture()
    Xn = oracle.normalize(X)
    P = engine.perplexity_dense(Xn, 30.0)
    DD = oracle.sqdist(Xn)
    rP, _, _ = oracle.perplexity_dense(DD, 30.0)
    np.testing.assert_allclose(P, rP, rtol=1e-10, atol=1e-300)
    Ps = (rP + rP.T) / 2
    Ps /= Ps.sum()
    Y = rng(4).normal(size=(2000, 2))
    dY, z = engine.gradient_dense(Ps, Y, precision="fp64")
    ref, zr = oracle.gradient_exact_dense(Ps, Y)
    assert np.linalg.norm(dY - ref) / np.linalg.norm(ref) < 1e-10
    assert abs(z - zr) / zr < 1e-12
    Yr, st = engine.tsne(X, perplexity=30.0, theta=0.0, seed=1)
    assert np.all(np.isfinite(Yr)) and abs(Yr.mean(0)).max() < 1e-6
    assert st["kl"][-1] < 0.5 * st["kl"][0]
    assert st["kl_final_exact"] > 0


def test_workload2_prototypes(engine, oracle):
    """N=20000, D=784 clipped noisy copies of 10 sparse prototypes, perplexity
    30, theta=0.5: certified kNN identical for BF16 and FP32 tiles, a
    symmetric P that sums to one, the Barnes-Hut gradient at a fixed Y
    against the oracle's pointer quadtree, then the full run."""
    from tapkee_amd import datasets
    X, _ = datasets.image_prototypes()
    X = X.astype(np.float64)
    i32, d32, f32 = engine.knn(engine.normalize(X), 90, dist_dtype="fp32")
    i16, d16, f16 = engine.knn(engine.normalize(X), 90, dist_dtype="bf16")
    np.testing.assert_array_equal(i16, i32)
    np.testing.assert_array_equal(d16, d32)
    assert f32 == 0
    row, col, val = engine.build_p(X, 30.0)
    assert abs(val.sum() - 1.0) < 1e-12
    import scipy.sparse as sp
    M = sp.csr_matrix((val, col, row), shape=(20000, 20000))
    assert abs(M - M.T).max() < 1e-18
    Y = rng(6).normal(size=(20000, 2)) * 5.0
    Y -= Y.mean(0)
    g, _ = engine.gradient(Y, row, col, val, 0.5, precision="fp64")
    r, _ = oracle.gradient_bh(row, col, val, Y, 0.5)
    assert np.linalg.norm(g - r) / np.linalg.norm(r) < 1e-9
    Yr, st = engine.tsne(X, perplexity=30.0, theta=0.5, seed=1)
    assert np.all(np.isfinite(Yr)) and abs(Yr.mean(0)).max() < 1e-4
    assert st["kl"][-1] < 0.5 * st["kl"][0]
    assert st["certificate_failures"] == 0
