"""The benchmark configurations at their full sizes against the oracle.

Config 3: Barnes-Hut t-SNE on the 50-D swiss roll, N=100,000 (the headline).
Config 4: kNN + P of a 50-cluster anisotropic mixture, N=250,000, D=128,
          with FP32, BF16 and FP8 e4m3 distance tiles.
Config 1 (kNN-P variant): N=2,000, D=50, theta=0 with K=90 neighbours.
Config 5 (parametric half): a 784-256-256-2 MLP trained on the KL.

The oracle cannot run the whole problem at these sizes in seconds, so it is
run on what it can finish: exact neighbours of a strided sample of query
rows (bit-exact indices and FP64 distances expected), the perplexity rows of
that sample (same bisection step counts, values to 1e-12), the Barnes-Hut
gradient at a spread embedding (one pointer-quadtree gradient), and the
normalisation of the whole input (bit-exact).  The symmetrisation at full
size is checked against scipy's (P + P^T) / 2 of the engine's own
conditional P, and the full runs through the final KL of the FP32 loop
against the FP64 loop (within 1%).
"""
import numpy as np
import pytest
import scipy.sparse as sp

from tests.conftest import rng

pytestmark = pytest.mark.gpu


def rel_l2(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def check_knn_p_sample(engine, oracle, Xn, K, perp, rows, dtypes, max_resolved=0):
    """Sampled rows of the kNN (each dtype) and of the conditional P against
    the oracle; returns the engine's FP32-tile idx, dist and the FP8 recall.
    At most max_resolved rows may fail their certificate (they are re-solved
    with every distance in FP64; every row is still compared with FP32's)."""
    print(f"[config] oracle brute force on {len(rows)} query rows of {Xn.shape}", flush=True)
    ridx, rdist = oracle.knn_brute_rows(Xn, K, rows)
    out = {}
    for dtype in dtypes:
        print(f"[config] engine kNN {dtype} on {Xn.shape}", flush=True)
        idx, dist, fails = engine.knn(Xn, K, dist_dtype=dtype)
        out[dtype] = (idx, dist, fails)
        if dtype == "fp8":
            continue   # approximate mode: rows failing the bound are counted, not re-solved
        limit = max_resolved if dtype == "bf16" else 0
        print(f"[config] {dtype}: {fails} rows not certified (re-solved exactly)", flush=True)
        assert fails <= limit, f"{dtype}: {fails} rows not certified"
        np.testing.assert_array_equal(idx[rows], ridx)
        np.testing.assert_array_equal(dist[rows], rdist)
    idx, dist, _ = out["fp32"]
    for dtype in dtypes:
        if dtype not in ("fp32", "fp8"):
            np.testing.assert_array_equal(out[dtype][0], idx)   # every row, not only the sample
    val, beta, iters, unconv = engine.perplexity(dist, perp)
    rval, rbeta, riters, runconv = oracle.perplexity_sparse(rdist, perp)
    np.testing.assert_array_equal(iters[rows], riters)
    np.testing.assert_allclose(beta[rows], rbeta, rtol=1e-12)
    np.testing.assert_allclose(val[rows], rval, rtol=1e-12, atol=1e-300)
    recall = None
    if "fp8" in out:
        i8 = out["fp8"][0]
        recall = float(np.mean([len(set(a) & set(b)) / K for a, b in zip(i8[rows], ridx)]))
        print(f"[config] FP8 recall@{K} on the sample: {recall:.4f}, "
              f"rows failing the bound: {out['fp8'][2]}", flush=True)
    return idx, dist, val, recall


def check_symmetric_p(engine, X, idx, val, perp, K):
    """build_p (normalise, kNN, P, symmetrise, /sum) equals scipy's (P+P^T)/2
    of the conditional P, divided by its sum."""
    N = X.shape[0]
    row, col, sval = engine.build_p(X, perp, dist_dtype="bf16")
    Pc = sp.csr_matrix((val.ravel(), idx.ravel(), np.arange(N + 1) * K), shape=(N, N))
    S = ((Pc + Pc.T) / 2).tocsr()
    S.sort_indices()
    S.data /= S.data.sum()
    np.testing.assert_array_equal(row, S.indptr)
    np.testing.assert_array_equal(col, S.indices)
    np.testing.assert_allclose(sval, S.data, rtol=1e-12, atol=0)
    return row, col, sval


# ---------------------------------------------------------------- config 3
@pytest.fixture(scope="module")
def swiss():
    from tapkee_amd import datasets
    X, _ = datasets.swiss_roll_50d()
    return X


def test_config3_knn_and_p(engine, oracle, swiss):
    X = swiss
    Xn = engine.normalize(X)
    np.testing.assert_array_equal(Xn, oracle.normalize(X))
    rows = np.arange(0, X.shape[0], 125)  # 800 query rows
    idx, dist, val, _ = check_knn_p_sample(engine, oracle, Xn, 90, 30.0, rows, ("bf16", "fp32"))
    check_symmetric_p(engine, X, idx, val, 30.0, 90)


@pytest.fixture(scope="module")
def swiss_runs(engine, swiss):
    """The full 1000-iteration runs with the FP32 and FP64 loops (seed 1)."""
    out = {}
    for prec in ("fp32", "fp64"):
        Y, st = engine.tsne(swiss, perplexity=30.0, theta=0.5, seed=1, loop_precision=prec)
        out[prec] = (Y, st)
    return out


def test_config3_full_run_fp32_vs_fp64_loop(swiss_runs):
    kl = {}
    for prec, (Y, st) in swiss_runs.items():
        assert np.all(np.isfinite(Y)) and st["certificate_failures"] == 0
        assert st["kl"][-1] < 0.5 * st["kl"][0]
        kl[prec] = st["kl_final_exact"]
    rel = kl["fp32"] / kl["fp64"] - 1
    print(f"[config] final exact KL: FP32 loop {kl['fp32']:.6f}, FP64 loop {kl['fp64']:.6f} ({rel:+.2e})",
          flush=True)
    assert abs(rel) < 0.01


def test_config3_gradient_at_final_embedding(engine, oracle, swiss, swiss_runs):
    """The gradient at the final embedding of the 1000-iteration FP32 run:
    FP64 and FP32 Barnes-Hut against the oracle's pointer quadtree, and
    Barnes-Hut against the exact all-pairs gradient."""
    X = swiss
    row, col, val = engine.build_p(X, 30.0, dist_dtype="bf16")
    Y = swiss_runs["fp32"][0]
    print(f"[config] oracle Barnes-Hut gradient at N={X.shape[0]}, |Y|max={np.abs(Y).max():.1f}", flush=True)
    ref, zr = oracle.gradient_bh(row, col, val, Y, 0.5)
    g64, z64 = engine.gradient(Y, row, col, val, 0.5, precision="fp64")
    g32, z32 = engine.gradient(Y, row, col, val, 0.5, precision="fp32")
    e64, e32 = rel_l2(g64, ref), rel_l2(g32, ref)
    gx, zx = engine.gradient(Y, row, col, val, 0.0, precision="fp64")
    ebh = rel_l2(ref, gx)
    print(f"[config] gradient rel L2: FP64 BH vs oracle {e64:.2e}, FP32 BH vs oracle {e32:.2e}, "
          f"BH (theta=0.5) vs exact {ebh:.2e}; Z BH/exact - 1 = {zr / zx - 1:.2e}", flush=True)
    assert e64 < 1e-9 and abs(z64 - zr) < 1e-10 * zr, (e64, z64, zr)
    assert e32 < 1e-4 and abs(z32 - zr) < 1e-5 * zr, (e32, z32, zr)
    # near convergence the gradient is a small difference of attraction and
    # repulsion, so the Barnes-Hut error (the reference's cell test compares
    # the half-width with theta) is large relative to it; it must shrink
    # with theta
    g2, _ = engine.gradient(Y, row, col, val, 0.2, precision="fp64")
    e2 = rel_l2(g2, gx)
    print(f"[config] BH vs exact at theta=0.2: {e2:.2e}", flush=True)
    assert e2 < 0.5 * ebh and ebh > 1e-6


# ---------------------------------------------------------------- config 4
def test_config4_precision_sweep(engine, oracle):
    from tapkee_amd import datasets
    X, _ = datasets.anisotropic_mixture()
    X = X.astype(np.float64)
    Xn = engine.normalize(X)
    np.testing.assert_array_equal(Xn, oracle.normalize(X))
    rows = np.arange(7, X.shape[0], 1563)  # 160 query rows
    # the block-centred BF16 tiles leave a few high-dimensional rows (50
    # overlapping clusters in 128-D) to the FP64 re-solve: at most 0.05%
    idx, dist, val, recall = check_knn_p_sample(engine, oracle, Xn, 90, 30.0, rows, ("fp32", "bf16", "fp8"),
                                                max_resolved=Xn.shape[0] // 2000)
    assert recall > 0.9
    check_symmetric_p(engine, X, idx, val, 30.0, 90)


# ---------------------------------------------------------------- config 1, kNN-P variant
def test_config1_knn_p_exact_loop(engine, oracle):
    """N=2000, D=50, theta=0 with the sparse kNN P (K=90): P as sets against
    the oracle, the exact gradient at a fixed Y, and the full run's final KL
    for the FP32 and FP64 loops."""
    from tapkee_amd import datasets
    X, labels = datasets.gaussian_mixture()
    row, col, val = engine.build_p(X, 30.0, knn_k=90, dist_dtype="bf16")
    rrow, rcol, rval, _ = oracle.build_p(X, 30.0, K=90, knn_mode=1)
    np.testing.assert_array_equal(row, rrow)
    for i in range(X.shape[0]):   # the oracle keeps the reference's in-row insertion order
        a = slice(row[i], row[i + 1])
        o = np.argsort(rcol[a], kind="stable")
        np.testing.assert_array_equal(col[a], rcol[a][o])
        np.testing.assert_allclose(val[a], rval[a][o], rtol=1e-12, atol=0)
    Y = rng(5).normal(size=(2000, 2)) * 4.0
    g, z = engine.gradient(Y, row, col, val, 0.0, precision="fp64")
    ref, zr = oracle.gradient_exact_sparse(rrow, rcol, rval, Y)
    assert rel_l2(g, ref) < 1e-10 and abs(z - zr) < 1e-12 * zr
    g32, _ = engine.gradient(Y, row, col, val, 0.0, precision="fp32")
    assert rel_l2(g32, ref) < 1e-5
    kl = {}
    for prec in ("fp32", "fp64"):
        Yr, st = engine.tsne(X, perplexity=30.0, theta=0.0, exact_p="knn", knn_k=90, seed=1, loop_precision=prec)
        assert np.all(np.isfinite(Yr)) and st["kl"][-1] < 0.5 * st["kl"][0]
        kl[prec] = st["kl_final_exact"]
        cent = np.stack([Yr[labels == c].mean(0) for c in range(10)])
        d = ((Yr[:, None, :] - cent[None]) ** 2).sum(-1)
        assert (d.argmin(1) == labels).mean() > 0.95   # the ten clusters come out separated
    assert abs(kl["fp32"] / kl["fp64"] - 1) < 0.01


# ---------------------------------------------------------------- config 5, parametric half
def test_config5_mlp_kl_falls(engine):
    """The 784-256-256-2 MLP trained by Adam on the engine's KL (exact Z,
    theta=0) for 60 full-batch steps: the KL falls and stays finite."""
    import torch  # noqa: F401

    from tapkee_amd import datasets, parametric
    X, _ = datasets.autograd_mixture()
    net, hist = parametric.train(X, perplexity=30.0, theta=0.0, steps=60, lr=1e-3, log_every=10)
    kl = hist["kl"]
    print(f"[config] MLP KL trace: {[round(v, 4) for v in kl]}", flush=True)
    assert np.all(np.isfinite(kl)) and kl[-1] < 0.8 * kl[0]
