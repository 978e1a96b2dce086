"""GPU parity of the 3-D Barnes-Hut path (octree) against the oracle's
pointer octree -- the reference's quadtree (quadtree.hpp) with a third axis,
itself pinned to the quadtree on coordinate planes and to the exact gradient
(tests/test_oracle.py).

Tolerances: FP64 gradient relative L2 <= 1e-9 and Z <= 1e-10 (same cells;
centres of mass from prefix sums instead of online means); FP32 <= 1e-4.
"""
import numpy as np
import pytest

from tests.conftest import rng

pytestmark = pytest.mark.gpu


def mixture(n=600, d=20, k=6, seed=0):
    g = rng(seed)
    c = g.normal(0, 5, size=(k, d))
    return c[np.arange(n) % k] + g.normal(size=(n, d))


def rel_l2(a, b):
    return float(np.linalg.norm(np.asarray(a) - np.asarray(b)) / max(np.linalg.norm(np.asarray(b)), 1e-300))


@pytest.fixture(scope="module")
def problem3(oracle):
    X = mixture(700, 20)
    row, col, val, _ = oracle.build_p(X, 20.0, knn_mode=1)
    Y = rng(7).normal(size=(700, 3)) * 3.0
    return row, col, val, Y - Y.mean(0)


@pytest.mark.parametrize("theta", [0.2, 0.5, 0.8])
def test_octree_gradient_fp64_matches_oracle(engine, oracle, problem3, theta):
    row, col, val, Y = problem3
    dY, z = engine.gradient(Y, row, col, val, theta, precision="fp64")
    rdY, rz = oracle.gradient_bh(row, col, val, Y, theta)
    assert rel_l2(dY, rdY) < 1e-9
    assert abs(z - rz) / rz < 1e-10


def test_octree_gradient_fp32(engine, oracle, problem3):
    row, col, val, Y = problem3
    dY, z = engine.gradient(Y, row, col, val, 0.5, precision="fp32")
    rdY, rz = oracle.gradient_bh(row, col, val, Y, 0.5)
    assert rel_l2(dY, rdY) < 1e-4
    assert abs(z - rz) / rz < 1e-5


def test_octree_theta0_is_exact(engine, oracle, problem3):
    row, col, val, Y = problem3
    # theta -> 0 through the tree kernel: only leaves are accepted
    dY, z = engine.gradient(Y, row, col, val, 1e-300, precision="fp64")
    rdY, rz = oracle.gradient_exact_sparse(row, col, val, Y)
    assert rel_l2(dY, rdY) < 1e-12
    assert abs(z - rz) / rz < 1e-12


@pytest.mark.parametrize("plane", [(0, 1), (0, 2), (1, 2)], ids=["xy", "xz", "yz"])
def test_octree_on_a_plane_equals_quadtree(engine, problem3, plane):
    # points on a coordinate plane: the octree's cells are the quadtree's
    row, col, val, Y3 = problem3
    Y2 = Y3[:, :2] - Y3[:, :2].mean(0)
    Yp = np.zeros_like(Y3)
    Yp[:, plane[0]] = Y2[:, 0]
    Yp[:, plane[1]] = Y2[:, 1]
    g2, z2 = engine.gradient(Y2, row, col, val, 0.5, precision="fp64")
    g3, z3 = engine.gradient(Yp, row, col, val, 0.5, precision="fp64")
    assert abs(z3 - z2) / z2 < 1e-13
    np.testing.assert_allclose(g3[:, plane[0]], g2[:, 0], rtol=1e-11, atol=1e-15)
    np.testing.assert_allclose(g3[:, plane[1]], g2[:, 1], rtol=1e-11, atol=1e-15)
    assert np.abs(g3[:, 3 - plane[0] - plane[1]]).max() < 1e-15


def test_octree_many_tiles(engine, oracle):
    # 30k points: many scan tiles and pairs workgroups, cells whose key range
    # leaves the pairs kernel's LDS window (searched in global memory)
    X = mixture(30_000, 10, k=8, seed=2)
    row, col, val = engine.build_p(X, 10.0, dist_dtype="fp32")
    Y = rng(11).normal(size=(30_000, 3)) * np.array([10.0, 3.0, 1.0])
    Y = Y - Y.mean(0)
    for prec, tol in (("fp64", 1e-9), ("fp32", 1e-4)):
        dY, z = engine.gradient(Y, row, col, val, 0.5, precision=prec)
        rdY, rz = oracle.gradient_bh(row, col, val, Y, 0.5)
        assert rel_l2(dY, rdY) < tol, prec
        assert abs(z - rz) / rz < tol, prec


def test_octree_tight_clusters(engine, oracle):
    # clusters of points within 1e-8 of each other (far below 2^-21 of the
    # box): equal Morton keys, one level-21 node per cluster whose children
    # are its points; the oracle separates them in deeper cells, with the
    # same forces (a point's own cluster is never accepted as a summary)
    g = rng(13)
    centres = g.normal(size=(40, 3)) * 5.0
    Y = np.repeat(centres, 25, axis=0) + g.normal(size=(1000, 3)) * 1e-8
    Y = Y - Y.mean(0)
    X = mixture(1000, 8, k=5, seed=3)
    row, col, val = engine.build_p(X, 10.0, dist_dtype="fp32")
    dY, z = engine.gradient(Y, row, col, val, 0.5, precision="fp64")
    rdY, rz = oracle.gradient_bh(row, col, val, Y, 0.5)
    assert rel_l2(dY, rdY) < 1e-8
    assert abs(z - rz) / rz < 1e-10


@pytest.mark.parametrize("m", [300, 3000])
def test_octree_long_runs(engine, oracle, m):
    # m points (and 40 more elsewhere) within ~1e-5 of each other in a box of
    # width ~9: their keys share the sorted top 24 bits (level-8 cells), so
    # their full-key order comes from the run sort: a wave's LDS sort for the
    # 300-point run, an in-place sort in global memory for the 3000-point
    # one, ranks for the 40
    g = rng(19)
    Y = g.normal(size=(12000, 3)) * np.array([3.0, 2.0, 1.5])
    Y[:m] = g.normal(size=(m, 3)) * 1e-5 + [0.3, -0.2, 0.1]
    Y[m:m + 40] = g.normal(size=(40, 3)) * 1e-5 + [-1.1, 0.7, 0.4]
    Y = Y - Y.mean(0)
    X = mixture(12000, 8, k=6, seed=5)
    row, col, val = engine.build_p(X, 10.0, dist_dtype="fp32")
    dY, z = engine.gradient(Y, row, col, val, 0.5, precision="fp64")
    rdY, rz = oracle.gradient_bh(row, col, val, Y, 0.5)
    assert rel_l2(dY, rdY) < 1e-8
    assert abs(z - rz) / rz < 1e-10


def test_octree_kl_entry_points(engine, oracle, problem3):
    row, col, val, Y = problem3
    kl = engine.kl(Y, row, col, val, mode="reference", theta=0.5)
    assert abs(kl - oracle.kl_bh(row, col, val, Y, 0.5)) < 1e-9 * abs(kl)


def _exact_kl(oracle, X, Y, perp):
    row, col, val, _ = oracle.build_p(X, perp, knn_mode=1)
    return oracle.kl_exact_sparse(row, col, val, Y)


def test_octree_full_run_within_oracle_spread(engine, oracle):
    """1000 iterations of 3-D Barnes-Hut t-SNE from the same P and Y0 (the
    trajectory is chaotic: six oracle runs from Y0 perturbed by 1e-15 to
    1e-12 give the spread); the engine's final exact-Z KL must lie within it,
    widened by 2%, for the FP64 and the FP32 loop."""
    X = mixture(500, 20)
    y0 = rng(9).normal(size=(500, 3)) * 1e-4
    ref = []
    for eps in (0.0, 1e-15, 3e-15, 1e-14, 1e-13, 1e-12):
        Yr, _ = oracle.tsne_run(X, no_dims=3, perplexity=15.0, theta=0.5, max_iter=1000, knn_mode=1,
                                y0=y0 * (1 + eps))
        ref.append(_exact_kl(oracle, X, Yr, 15.0))
    lo, hi = min(ref) * 0.98, max(ref) * 1.02
    for prec, tol in (("fp64", 1e-9), ("fp32", 1e-7)):
        Y, st = engine.tsne(X, no_dims=3, perplexity=15.0, theta=0.5, y0=y0, loop_precision=prec,
                            dist_dtype="fp32")
        assert Y.shape == (500, 3) and np.all(np.isfinite(Y)) and abs(Y.mean(0)).max() < 1e-6
        kl = st["kl_final_exact"]
        assert abs(kl - _exact_kl(oracle, X, Y, 15.0)) < tol * kl
        assert lo <= kl <= hi, (prec, kl, ref)


def test_octree_session_sort_fault_is_reported(engine):
    X = mixture(3000, 16, k=4)
    s = engine.Session(X, no_dims=3, perplexity=20.0, theta=0.5, seed=3)
    s.inject_sort_fault()
    s.step(60)
    with pytest.raises(engine._lib.EngineError, match="radix sort"):
        s.embedding()
    s.close()
    s = engine.Session(X, no_dims=3, perplexity=20.0, theta=0.5, seed=3)
    s.step(60)
    Y = s.embedding()
    assert Y.shape == (3000, 3) and np.all(np.isfinite(Y))
    s.close()
