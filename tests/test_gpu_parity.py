"""GPU parity: every engine step against the FP64 oracle on the same inputs.

Tolerances (stated per test): integer/index outputs bit-exact; FP64 kernels
to a few ulps of accumulated rounding; FP32 kernels relative L2 <= 1e-5
(exact sums) or 1e-4 (Barnes-Hut, same tree).
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


def csr_from_oracle(oracle, X, perp):
    row, col, val, Xn = oracle.build_p(X, perp, knn_mode=1)
    return row, col, val, Xn


# ---------------------------------------------------------------- normalise
@pytest.mark.parametrize("shape", [(500, 20), (1001, 50), (64, 784)])
def test_normalize_bit_exact(engine, oracle, shape):
    X = rng(1).normal(size=shape) * 3 + 1
    np.testing.assert_array_equal(engine.normalize(X), oracle.normalize(X))


# ---------------------------------------------------------------- kNN
@pytest.mark.parametrize("dtype", ["fp32", "bf16"])
# BF16 up to d = 158 runs on the block-centred tiles (d + 2 values per image
# row in vectors of 16: d = 3 -> 1, 20 -> 2, 50 -> 4, 93 -> 6, 107 -> 7,
# 128 -> 9, 148 and 150 -> 10, the widest); beyond, on hi/lo splits streamed
# in k-stages (333, 784).  FP32: d = 128 gives rows of 17 vectors (D+1
# k-values); 333 (42 -> 44) and 784 stream through LDS in k-stages
@pytest.mark.parametrize("n,d,K", [(700, 20, 30), (1500, 50, 90), (300, 784, 20), (129, 3, 128), (600, 128, 40),
                                   (500, 93, 30), (400, 107, 25), (350, 148, 20), (450, 150, 20), (260, 333, 30)])
def test_knn_bit_exact(engine, oracle, dtype, n, d, K):
    X = oracle.normalize(mixture(n, d))
    idx, dist, fails = engine.knn(X, K, dist_dtype=dtype)
    ridx, rdist = oracle.knn_brute(X, K)
    np.testing.assert_array_equal(idx, ridx)
    np.testing.assert_array_equal(dist, rdist)


@pytest.mark.parametrize("dtype", ["fp32", "bf16"])
def test_knn_wide_rows_many_blocks(engine, oracle, dtype):
    # N = 33,000 rows of d = 160 (BF16 31 vectors per half row, FP32 21):
    # enough query blocks for the four-wave launch of the k-staged tiles;
    # every row's K = 30 neighbours are exact (checked on a sample of rows
    # against brute force), and both precisions give the same graph
    n, d, K = 33_000, 160, 30
    X = oracle.normalize(mixture(n, d, k=40, seed=7))
    idx, dist, fails = engine.knn(X, K, dist_dtype=dtype)
    rows = np.unique(np.concatenate([np.arange(0, n, 997), [n - 1]]))
    ridx, rdist = oracle.knn_brute_rows(X, K, rows)
    np.testing.assert_array_equal(idx[rows], ridx)
    np.testing.assert_array_equal(dist[rows], rdist)
    assert np.all(np.diff(dist, axis=1) >= 0)


@pytest.mark.parametrize("dtype", ["fp32", "bf16"])
@pytest.mark.parametrize("n,d,K", [(900, 20, 450), (1300, 7, 1299), (600, 130, 385)])
def test_knn_large_k_bit_exact(engine, oracle, dtype, n, d, K):
    # K beyond the candidate buffers (384): every distance in FP64 and an
    # exact selection per row, whatever the tile dtype; K = N - 1 takes
    # every other row
    X = oracle.normalize(mixture(n, d))
    idx, dist, fails = engine.knn(X, K, dist_dtype=dtype)
    ridx, rdist = oracle.knn_brute(X, K)
    np.testing.assert_array_equal(idx, ridx)
    np.testing.assert_array_equal(dist, rdist)
    assert fails == 0


@pytest.mark.parametrize("dtype", ["fp32", "bf16"])
def test_knn_uncertified_rows_are_resolved_exactly(engine, oracle, dtype):
    # k' = K leaves little margin for the error bound: rows fail their
    # certificate and are re-solved with every distance in FP64; the result
    # is still the exact neighbour list (the tiles' rows are in locality
    # order, the re-solve works in the caller's: ties by the caller's index)
    X = oracle.normalize(mixture(1500, 50))
    idx, dist, fails = engine.knn(X, 60, dist_dtype=dtype, candidates=60)
    assert fails > 0  # (the compaction keeps somewhat more than k', so most rows still certify)
    ridx, rdist = oracle.knn_brute(X, 60)
    np.testing.assert_array_equal(idx, ridx)
    np.testing.assert_array_equal(dist, rdist)


def test_knn_large_k_ties(engine, oracle):
    # a lattice with duplicated points: many equal distances around the K-th
    g = np.stack(np.meshgrid(np.arange(20.0), np.arange(20.0)), -1).reshape(-1, 2)
    X = np.concatenate([g, g[:100]])
    for K in (400, 499):
        idx, dist, _ = engine.knn(X, K, dist_dtype="fp32")
        ridx, rdist = oracle.knn_brute(X, K)
        np.testing.assert_array_equal(idx, ridx)
        np.testing.assert_array_equal(dist, rdist)


def test_knn_fp8_is_approximate_but_close(engine, oracle):
    # FP8 e4m3 tiles keep 4 significant bits; the FP64 re-rank of 4K
    # candidates recovers most true neighbours, and reported distances are
    # exact for the indices returned
    X = oracle.normalize(mixture(1200, 32))
    idx, dist, fails = engine.knn(X, 30, dist_dtype="fp8", candidates=120)
    ridx, rdist = oracle.knn_brute(X, 30)
    recall = np.mean([len(set(a) & set(b)) / 30 for a, b in zip(idx, ridx)])
    assert recall > 0.95, recall
    D = ((X[:, None, :] - X[None, :, :]) ** 2).sum(-1)
    np.testing.assert_allclose(dist, np.take_along_axis(D, idx, 1), rtol=1e-12, atol=1e-15)
    assert np.all(np.diff(dist, axis=1) >= 0)


def test_knn_ties_broken_by_index(engine, oracle):
    # duplicated points and a lattice: many exactly equal distances
    g = np.stack(np.meshgrid(np.arange(12.0), np.arange(12.0)), -1).reshape(-1, 2)
    X = np.concatenate([g, g[:40]])
    for dtype in ("fp32", "bf16"):
        idx, dist, _ = engine.knn(X, 10, dist_dtype=dtype)
        ridx, rdist = oracle.knn_brute(X, 10)
        np.testing.assert_array_equal(idx, ridx)
        np.testing.assert_array_equal(dist, rdist)


# ---------------------------------------------------------------- perplexity
# K = 90: four rows per wave; 300: one wave per row; 900: one workgroup per row
@pytest.mark.parametrize("perp", [30.0, 100.0, 300.0])
def test_perplexity_rows(engine, oracle, perp):
    X = oracle.normalize(mixture(1000, 20))
    _, dist = oracle.knn_brute(X, int(3 * perp))
    val, beta, iters, unconv = engine.perplexity(dist, perp)
    rval, rbeta, riters, runconv = oracle.perplexity_sparse(dist, perp)
    # same bisection path: identical iteration counts and betas (dyadic updates)
    np.testing.assert_array_equal(iters, riters)
    np.testing.assert_array_equal(beta, rbeta)
    assert unconv == runconv
    np.testing.assert_allclose(val, rval, rtol=1e-12, atol=1e-300)


@pytest.mark.parametrize("compat", [False, True])
def test_perplexity_dense(engine, oracle, compat):
    X = oracle.normalize(mixture(300, 10))
    P = engine.perplexity_dense(X, 20.0, dd_mode="tapkee_compat" if compat else "textbook")
    RP, _, _ = oracle.perplexity_dense(oracle.sqdist(X, compat), 20.0)
    np.testing.assert_allclose(P, RP, rtol=1e-11, atol=1e-300)


# ---------------------------------------------------------------- symmetrise
def test_symmetrize_matches_reference_as_sets(engine, oracle):
    X = oracle.normalize(mixture(900, 16))
    idx, dist = oracle.knn_brute(X, 45)
    val, *_ = oracle.perplexity_sparse(dist, 15.0)
    row, col, sval = engine.symmetrize(idx, val)
    rrow, rcol, rsval = oracle.symmetrize(idx, val)
    np.testing.assert_array_equal(row, rrow)  # same counts per row
    for i in range(len(row) - 1):
        a = slice(row[i], row[i + 1])
        o = np.argsort(rcol[a], kind="stable")
        np.testing.assert_array_equal(col[a], rcol[a][o])
        np.testing.assert_array_equal(sval[a], rsval[a][o])


def test_symmetrize_many_tiles(engine, oracle):
    # N*K = 1.8M edges of a random graph (few mutual pairs): 1.8M one-way
    # edges radix-sorted and merged into 20k rows
    g = rng(5)
    N, K = 20_000, 90
    idx = np.stack([g.choice(N - 1, K, replace=False) for _ in range(N)]).astype(np.int64)
    idx += idx >= np.arange(N)[:, None]  # no self edges
    val = g.random((N, K))
    val /= val.sum(1, keepdims=True)
    row, col, sval = engine.symmetrize(idx, val)
    rrow, rcol, rsval = oracle.symmetrize(idx, val)
    np.testing.assert_array_equal(row, rrow)
    rid = np.repeat(np.arange(N), np.diff(rrow))
    o = np.lexsort((rcol, rid))
    np.testing.assert_array_equal(col, rcol[o])
    np.testing.assert_array_equal(sval, rsval[o])


@pytest.mark.parametrize("K", [200, 300])
def test_symmetrize_wide_rows(engine, oracle, K):
    """K = 200 (rows of 256 sorted in a wave) and K = 300 (the sort path)."""
    g = rng(9)
    N = 3000
    idx = np.stack([g.choice(N - 1, K, replace=False) for _ in range(N)]).astype(np.int64)
    idx += idx >= np.arange(N)[:, None]
    val = g.random((N, K))
    row, col, sval = engine.symmetrize(idx, val)
    rrow, rcol, rsval = oracle.symmetrize(idx, val)
    np.testing.assert_array_equal(row, rrow)
    rid = np.repeat(np.arange(N), np.diff(rrow))
    o = np.lexsort((rcol, rid))
    np.testing.assert_array_equal(col, rcol[o])
    np.testing.assert_array_equal(sval, rsval[o])


def test_symmetrize_hub_rows(engine, oracle):
    """Points listed by thousands of rows they do not list themselves: a row
    whose reverse-only entries exceed what a wave stages in LDS (1024)."""
    g = rng(10)
    N, K = 5000, 20
    idx = np.stack([g.choice(N - 1, K, replace=False) for _ in range(N)]).astype(np.int64)
    idx += idx >= np.arange(N)[:, None]
    # rows 10.. list point 0 (in-degree ~2500) and rows 3000.. list point 1
    for i in range(10, N, 2):
        if 0 not in idx[i]:
            idx[i, 0] = 0
    for i in range(3000, N):
        if 1 not in idx[i] and i != 1:
            idx[i, 1] = 1
    assert all(len(set(r)) == K for r in idx)
    val = g.random((N, K))
    row, col, sval = engine.symmetrize(idx, val)
    rrow, rcol, rsval = oracle.symmetrize(idx, val)
    np.testing.assert_array_equal(row, rrow)
    assert row[1] - row[0] > 1024
    rid = np.repeat(np.arange(N), np.diff(rrow))
    o = np.lexsort((rcol, rid))
    np.testing.assert_array_equal(col, rcol[o])
    np.testing.assert_array_equal(sval, rsval[o])


@pytest.mark.parametrize("dtype", ["fp32", "bf16"])
def test_build_p_matches_oracle(engine, oracle, dtype):
    """The whole input-similarity stage: identical pattern, values to 1e-13."""
    X = mixture(500, 20)
    row, col, val = engine.build_p(X, 15.0, dist_dtype=dtype)
    rrow, rcol, rval, _ = oracle.build_p(X, 15.0, knn_mode=1)
    np.testing.assert_array_equal(row, rrow)
    for i in range(len(row) - 1):
        a = slice(row[i], row[i + 1])
        o = np.argsort(rcol[a], kind="stable")
        np.testing.assert_array_equal(col[a], rcol[a][o])
        np.testing.assert_allclose(val[a], rval[a][o], rtol=1e-13, atol=0)


def test_build_p_large_perplexity_matches_oracle(engine, oracle):
    """Perplexity 150 (K = 450, beyond the tiles' candidate buffers): the
    exact kNN path, the workgroup-per-row bisection and the symmetrisation
    of 2NK = 630k edges; identical pattern, values to 1e-13."""
    X = mixture(700, 12)
    row, col, val = engine.build_p(X, 150.0, dist_dtype="bf16")
    rrow, rcol, rval, _ = oracle.build_p(X, 150.0, knn_mode=1)
    np.testing.assert_array_equal(row, rrow)
    for i in range(len(row) - 1):
        a = slice(row[i], row[i + 1])
        o = np.argsort(rcol[a], kind="stable")
        np.testing.assert_array_equal(col[a], rcol[a][o])
        np.testing.assert_allclose(val[a], rval[a][o], rtol=1e-13, atol=0)


def test_session_large_perplexity_runs(engine):
    X = mixture(1500, 10, k=4)
    Y, st = engine.tsne(X, perplexity=200.0, theta=0.5, seed=2, max_iter=300)
    assert np.all(np.isfinite(Y)) and st["knn_k"] == 600
    assert st["kl_final_exact"] > 0.0


# ---------------------------------------------------------------- gradient
@pytest.fixture(scope="module")
def small_problem(oracle):
    X = mixture(700, 20)
    row, col, val, _ = csr_from_oracle(oracle, X, 20.0)
    Y = rng(7).normal(size=(700, 2)) * 3.0
    return row, col, val, Y


def test_gradient_exact_fp64(engine, oracle, small_problem):
    row, col, val, Y = small_problem
    dY, z = engine.gradient(Y, row, col, val, 0.0, precision="fp64")
    rdY, rz = oracle.gradient_exact_sparse(row, col, val, Y)
    assert rel_l2(dY, rdY) < 1e-12
    assert abs(z - rz) / rz < 1e-13


def test_gradient_exact_fp32(engine, oracle, small_problem):
    row, col, val, Y = small_problem
    dY, z = engine.gradient(Y, row, col, val, 0.0, precision="fp32")
    rdY, rz = oracle.gradient_exact_sparse(row, col, val, Y)
    assert rel_l2(dY, rdY) < 1e-5
    assert abs(z - rz) / rz < 1e-5


@pytest.mark.parametrize("theta", [0.2, 0.5, 0.8])
def test_gradient_bh_fp64_matches_reference_tree(engine, oracle, small_problem, theta):
    row, col, val, Y = small_problem
    Y = Y - Y.mean(0)
    dY, z = engine.gradient(Y, row, col, val, theta, precision="fp64")
    rdY, rz = oracle.gradient_bh(row, col, val, Y, theta)
    assert rel_l2(dY, rdY) < 1e-9
    assert abs(z - rz) / rz < 1e-10


def test_gradient_bh_fp32(engine, oracle, small_problem):
    row, col, val, Y = small_problem
    Y = Y - Y.mean(0)
    dY, z = engine.gradient(Y, row, col, val, 0.5, precision="fp32")
    rdY, rz = oracle.gradient_bh(row, col, val, Y, 0.5)
    assert rel_l2(dY, rdY) < 1e-4
    assert abs(z - rz) / rz < 1e-5


def test_gradient_bh_theta0_is_exact(engine, oracle, small_problem):
    row, col, val, Y = small_problem
    # theta = 0 through the tree kernel: only leaves are accepted
    dY_tree, z_tree = engine.gradient(Y, row, col, val, 1e-300, precision="fp64")
    rdY, rz = oracle.gradient_exact_sparse(row, col, val, Y)
    assert rel_l2(dY_tree, rdY) < 1e-12
    assert abs(z_tree - rz) / rz < 1e-12


@pytest.mark.parametrize("dims", [2, 3])
def test_duplicate_embedding_points(engine, oracle, small_problem, dims):
    """Exact duplicates in Y (documented deviation): every point is its own
    leaf and skips only itself, so each duplicate sees the others at D = 0
    (q = 1), as the exact gradient has it -- the limit of distinct points
    that coincide.  The reference instead stores one point per leaf and only
    counts the others (quadtree.hpp:218-234), and loses that count when the
    leaf later splits.  Checked: through the tree kernel with theta -> 0 the
    gradient equals the exact one; at theta = 0.5 the forces and Z equal the
    oracle's for the same points with each duplicate moved by eps (distinct
    points, which the reference's tree handles), to O(eps)."""
    row, col, val, Y = small_problem
    Y = rng(17).normal(size=(len(Y), dims)) * 3.0
    Y[1::7] = Y[0::7][: len(Y[1::7])]  # every 7th point duplicated into the next
    Y[5] = Y[0]                         # a triple
    Y = Y - Y.mean(0)
    g0, z0 = engine.gradient(Y, row, col, val, 1e-300, precision="fp64")
    r0, rz0 = oracle.gradient_exact_sparse(row, col, val, Y)
    assert rel_l2(g0, r0) < 1e-12
    assert abs(z0 - rz0) / rz0 < 1e-12
    # displaced copies: separate Morton keys (above 2^-32 of the box in 2-D,
    # 2^-21 in 3-D), so the oracle's tree separates them too
    eps = 1e-6 if dims == 2 else 1e-4
    Yd = Y.copy()
    seen = {}
    for i in range(len(Y)):
        key = Y[i].tobytes()
        k = seen.get(key, 0)
        Yd[i] += eps * k * np.arange(1, dims + 1)
        seen[key] = k + 1
    assert sum(v > 1 for v in seen.values()) > 50
    g, z = engine.gradient(Y, row, col, val, 0.5, precision="fp64")
    r, rz = oracle.gradient_bh(row, col, val, Yd, 0.5)
    tol = 1e-5 if dims == 2 else 1e-3
    assert rel_l2(g, r) < tol
    assert abs(z - rz) / rz < tol


@pytest.mark.parametrize("compat", [False, True])
def test_gradient_dense(engine, oracle, compat):
    X = oracle.normalize(mixture(250, 10))
    P, _, _ = oracle.perplexity_dense(oracle.sqdist(X), 15.0)
    P = P + P.T
    P /= P.sum()
    Y = rng(3).normal(size=(250, 2)) * (0.05 if compat else 2.0)
    mode = "tapkee_compat" if compat else "textbook"
    dY, z = engine.gradient_dense(P, Y, dd_mode=mode, precision="fp64")
    rdY, rz = oracle.gradient_exact_dense(P, Y, compat=compat)
    assert rel_l2(dY, rdY) < 1e-11
    assert abs(z - rz) / abs(rz) < 1e-12


# ---------------------------------------------------------------- tree
@pytest.mark.parametrize("n", [400, 60_000, 400_000])
def test_bh_tree_matches_reference_tree(engine, oracle, n):
    # 400 points sort as a single tile; the larger sizes chain 100-400 tiles
    # through the radix sort's look-back
    Y = rng(11).normal(size=(n, 2))
    Y -= Y.mean(0)
    t = engine.bh_tree(Y)
    r = oracle.quadtree(Y)
    # compress the reference tree: drop empty nodes and one-child chain nodes
    keep = (r["cum"] > 0) & ((r["leaf"] == 1) | (r["nchild"] >= 2))
    rl = r["level"][keep]
    rc = r["cum"][keep]
    rleaf = r["leaf"][keep] == 1
    # pre-order sequences agree node for node
    assert len(rl) == len(t["count"])
    np.testing.assert_array_equal(rc, t["count"])
    np.testing.assert_array_equal(rleaf, t["leaf_point"] >= 0)
    inner = ~rleaf
    np.testing.assert_array_equal(rl[inner], t["level"][inner])
    np.testing.assert_array_equal(r["index"][keep][rleaf], t["leaf_point"][rleaf])
    # centres of mass: engine from FP64 prefix-sum differences, reference
    # from running means; both within a few ulp of the prefix magnitude
    atol = 1e-12 if n < 1000 else 1e-9
    np.testing.assert_allclose(t["comx"], r["comx"][keep], rtol=0, atol=atol)
    np.testing.assert_allclose(t["comy"], r["comy"][keep], rtol=0, atol=atol)
    np.testing.assert_allclose(t["root_box"][2:], r["root_box"][2:], rtol=1e-15)


@pytest.mark.parametrize("m", [300, 3000])
def test_bh_tree_tight_clusters(engine, oracle, m):
    # m points (and 40 more elsewhere) within ~1e-5 of each other in a box of
    # width ~9: their keys share the sorted top bits (level-8 to level-16
    # cells for 2 to 4 radix passes), so the full-key order comes from the
    # run sort: ranks for runs of at most 64 points (the 40), a wave's LDS
    # sort for runs of 65..1024 points, an in-place sort in global memory for
    # the 3000-point run.  The engine builds the tree with each radix width
    # and checks that they agree.  The 64-bit keys are all distinct (checked
    # when the case was made).
    r = rng(17)
    Y = r.normal(size=(20000, 2))
    Y[:m] = r.normal(size=(m, 2)) * 1e-5 + [0.3, -0.2]
    Y[m:m + 40] = r.normal(size=(40, 2)) * 1e-5 + [-1.1, 0.7]
    Y -= Y.mean(0)
    t = engine.bh_tree(Y)
    ref = oracle.quadtree(Y)
    keep = (ref["cum"] > 0) & ((ref["leaf"] == 1) | (ref["nchild"] >= 2))
    rleaf = ref["leaf"][keep] == 1
    assert len(ref["level"][keep]) == len(t["count"])
    np.testing.assert_array_equal(ref["cum"][keep], t["count"])
    np.testing.assert_array_equal(rleaf, t["leaf_point"] >= 0)
    np.testing.assert_array_equal(ref["level"][keep][~rleaf], t["level"][~rleaf])
    np.testing.assert_array_equal(ref["index"][keep][rleaf], t["leaf_point"][rleaf])
    np.testing.assert_allclose(t["comx"], ref["comx"][keep], rtol=0, atol=1e-9)
    np.testing.assert_allclose(t["comy"], ref["comy"][keep], rtol=0, atol=1e-9)


# ---------------------------------------------------------------- KL
def test_kl_variants(engine, oracle, small_problem):
    row, col, val, Y = small_problem
    Y = Y - Y.mean(0)
    assert abs(engine.kl(Y, row, col, val, mode="exact") -
               oracle.kl_exact_sparse(row, col, val, Y)) < 1e-11
    a = engine.kl(Y, row, col, val, theta=0.5, mode="reference")
    b = oracle.kl_bh(row, col, val, Y, 0.5)
    assert abs(a - b) < 1e-10 * max(1.0, abs(b))


def test_kl_dense(engine, oracle):
    X = oracle.normalize(mixture(200, 8))
    P, _, _ = oracle.perplexity_dense(oracle.sqdist(X), 12.0)
    P = (P + P.T) / (P + P.T).sum()
    Y = rng(5).normal(size=(200, 2))
    assert abs(engine.kl_dense(P, Y) - oracle.kl_exact_dense(P, Y)) < 1e-11


# ---------------------------------------------------------------- update
def test_update_recurrence(engine, oracle):
    g = rng(4)
    n = 300
    dY = g.normal(size=(n, 2))
    dY[::7] = 0.0
    uY = g.normal(size=(n, 2))
    uY[::5] = 0.0
    gains = g.uniform(0.005, 2, size=(n, 2))
    Y = g.normal(size=(n, 2))
    a = engine.update(dY, uY, gains, Y, 0.5, precision="fp64")
    b = oracle.update(dY, uY, gains, Y, 0.5)
    for x, y in zip(a, b):
        np.testing.assert_allclose(x, y, rtol=1e-14, atol=1e-14)


# ---------------------------------------------------------------- profiling
@pytest.mark.parametrize("dims", [2, 3])
def test_profiled_interaction_count(engine, dims):
    # With theta far below any cell's size ratio no cell is accepted: every
    # point interacts with each of the other N-1 leaves, so each profiled
    # traversal launch counts N(N-1) interactions (summed from the
    # traversal's workgroups).
    X = mixture(3000, 16, k=5)
    n = X.shape[0]
    s = engine.Session(X, no_dims=dims, perplexity=20.0, theta=1e-9, seed=2)
    s.step(3)
    s.profile(True)
    s.step(2)
    kt = s.kernel_timing()
    s.profile(False)
    s.close()
    assert kt["launches_repulsion"] >= 2 and kt["launches_repulsion"] % 2 == 0
    assert kt["interactions"] == kt["launches_repulsion"] * n * (n - 1)


# ---------------------------------------------------------------- full run
@pytest.mark.parametrize("dims", [2, 3])
def test_renumbered_run_returns_caller_order(engine, dims):
    # points are renumbered every 50 iterations (quadtree and octree loops);
    # embedding() must undo it.  After 120 iterations rows must still belong
    # to their points: nearby inputs stay nearby.
    X = mixture(2000, 16, k=4)
    labels = np.arange(2000) % 4
    s = engine.Session(X, no_dims=dims, perplexity=20.0, theta=0.5, seed=3)
    s.step(120)
    Y = s.embedding()
    s.close()
    cent = np.stack([Y[labels == c].mean(0) for c in range(4)])
    d = ((Y[:, None, :] - cent[None]) ** 2).sum(-1)
    assert (d.argmin(1) == labels).mean() > 0.95


def test_sort_fault_is_reported(engine):
    # The Morton sort's look-back is bounded; when it gives up it sets a
    # fault word.  The word is set here from the host (the sort itself stays
    # valid): steps must run to the end without touching the GPU's memory out
    # of bounds (no walk, identity renumbering at iteration 50, a KL
    # evaluation at 50) and the embedding and the KL must then raise.
    X = mixture(3000, 16, k=4)
    s = engine.Session(X, perplexity=20.0, theta=0.5, seed=3)
    s.inject_sort_fault()
    s.step(60)
    with pytest.raises(engine._lib.EngineError, match="radix sort"):
        s.embedding()
    with pytest.raises(engine._lib.EngineError, match="radix sort"):
        s.kl()
    s.close()
    # a session without the fault is unaffected
    s = engine.Session(X, perplexity=20.0, theta=0.5, seed=3)
    s.step(60)
    assert np.all(np.isfinite(s.embedding()))
    s.close()


@pytest.mark.parametrize("n", [3000, 70_000, 600_000])
def test_attraction_inside_build_equals_attraction_alone(engine, n):
    # Untimed steps run the attraction rows in extra workgroups of the tree
    # build's launches; profiling mode runs the build without them and the
    # attraction as a launch of its own.  Each row is computed by the same
    # code either way, so the two loops must agree bit for bit (70k points:
    # several sort tiles and row groups per launch; 600k points: more than
    # 512 sort tiles, so tiles are numbered in arrival order and the extra
    # workgroups take no number).
    X = mixture(n, 16, k=6)
    out = []
    for prof in (False, True):
        s = engine.Session(X, perplexity=20.0, theta=0.5, seed=4, log_every=0)
        s.profile(prof)
        s.step(60)
        out.append(s.embedding())
        s.close()
    assert np.array_equal(out[0], out[1])


def test_y0_replays_std_rand(engine, oracle):
    X = mixture(300, 10)
    # max_iter = 0: the output is the (centred) initial solution
    oracle.srand(12345)
    Yr, _ = oracle.tsne_run(X, perplexity=10.0, theta=0.5, max_iter=0)
    Y, _ = engine.tsne(X, perplexity=10.0, theta=0.5, max_iter=0, seed=12345, loop_precision="fp64")
    np.testing.assert_allclose(Y, Yr - Yr.mean(0), rtol=0, atol=1e-18)


def _exact_kl_oracle(oracle, X, Y, perp, theta):
    if theta > 0:
        row, col, val, _ = oracle.build_p(X, perp, knn_mode=1)
        return oracle.kl_exact_sparse(row, col, val, Y)
    Xn = oracle.normalize(X)
    P, _, _ = oracle.perplexity_dense(oracle.sqdist(Xn), perp)
    P = P + P.T
    P /= P.sum()
    return oracle.kl_exact_dense(P, Y)


@pytest.mark.parametrize("theta", [0.5, 0.0])
def test_full_run_final_kl_within_oracle_spread(engine, oracle, theta):
    """1000 iterations from the same P and Y0.  The trajectory is chaotic: a
    1e-15 relative change of Y0 moves the oracle's own KL at iteration 300 by
    several percent, at 1000 iterations by 0.7% (theta=0) to 2% (theta=0.5);
    the engine's summation orders (points renumbered for locality, blocked
    sums) are such changes.  The engine's final exact-Z KL must lie within
    the spread of six perturbed oracle runs widened by 2%."""
    X = mixture(500, 20)
    y0 = rng(9).normal(size=(500, 2)) * 1e-4
    ref = []
    for eps in (0.0, 1e-15, 3e-15, 1e-14, 1e-13, 1e-12):
        Yr, _ = oracle.tsne_run(X, perplexity=15.0, theta=theta, max_iter=1000, knn_mode=1, y0=y0 * (1 + eps))
        ref.append(_exact_kl_oracle(oracle, X, Yr, 15.0, theta))
    lo, hi = min(ref) * 0.98, max(ref) * 1.02
    # The engine's own final KL is taken on the P its loop holds: FP64, or
    # rounded to FP32 (relative 6e-8 per entry) for the FP32 loop.  The
    # embedding it returns is in the caller's row order (points are
    # renumbered inside), so the oracle's KL of it must agree.
    for prec, tol in (("fp64", 1e-9), ("fp32", 1e-7)):
        Y, st = engine.tsne(X, perplexity=15.0, theta=theta, y0=y0, loop_precision=prec, dist_dtype="fp32")
        assert np.all(np.isfinite(Y)) and abs(Y.mean(0)).max() < 1e-6
        kl = st["kl_final_exact"]
        kl_o = _exact_kl_oracle(oracle, X, Y, 15.0, theta)
        assert abs(kl - kl_o) < tol * kl, f"{prec}: engine {kl!r} oracle {kl_o!r}"
        assert lo <= kl <= hi, (prec, kl, ref)


def test_session_from_precomputed_knn_equals_features_path(engine):
    # the neighbour graph the features path computes, handed over: the same
    # P and the same Y0 give the same trajectory (the features path starts
    # with the points in the kNN's locality order, the handed-over graph in
    # the caller's order, so sums run in different orders: equal to rounding)
    X = mixture(3000, 16, k=5)
    y0 = rng(9).normal(size=(3000, 2)) * 1e-4
    idx, dist, _ = engine.knn(engine.normalize(X), 45, dist_dtype="fp32")
    ref = engine.Session(X, perplexity=15.0, theta=0.5, y0=y0, max_iter=300)
    ref.step(300)
    a = engine.Session.from_knn(idx, dist, perplexity=15.0, theta=0.5, y0=y0, max_iter=300)
    a.step(300)
    np.testing.assert_allclose(a.stats()["kl"], ref.stats()["kl"], rtol=1e-4)
    a.close()
    ref.close()


@pytest.mark.parametrize("theta", [0.5, 0.0])
def test_session_from_precomputed_p(engine, oracle, theta):
    # a joint P (here the engine's own, scaled by 3: the session divides by
    # the sum) gives the features path's trajectory up to the rounding of
    # that division; at theta=0 the sparse P runs with the exact repulsion
    X = mixture(2000, 12, k=4)
    y0 = rng(4).normal(size=(2000, 2)) * 1e-4
    row, col, val = engine.build_p(X, 12.0, dist_dtype="fp32")
    a = engine.Session.from_p(row, col, 3.0 * val, theta=theta, y0=y0, max_iter=200)
    a.step(200)
    Ya = a.embedding()
    kla = a.stats()["kl"]
    a.close()
    kw = dict(exact_p="knn") if theta == 0.0 else {}
    b = engine.Session(X, perplexity=12.0, theta=theta, y0=y0, max_iter=200, **kw)
    b.step(200)
    klb = b.stats()["kl"]
    b.close()
    np.testing.assert_allclose(kla, klb, rtol=1e-4)
    # the gradient at the result against the oracle's with the same P
    g, _ = engine.gradient(Ya, row, col, val, theta, precision="fp64")
    r, _ = (oracle.gradient_bh(row, col, val, Ya, theta) if theta > 0
            else oracle.gradient_exact_sparse(row, col, val, Ya))
    assert rel_l2(g, r) < 1e-9
