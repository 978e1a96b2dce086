"""Pins the FP64 oracle (oracle/tsne_oracle.cpp) against what the mathematics
fixes: closed forms, invariants, independent library routines, brute force
and finite differences.  The reference ships no golden vectors for t-SNE
(its only t-SNE test checks the output shape, test/unit/methods.cpp:119-122),
so these checks are the pins.  CPU only.
"""
import numpy as np
import pytest
import scipy.optimize
import scipy.sparse
import scipy.spatial.distance as ssd

from tests.conftest import rng


def mixture(n=300, d=10, k=4, seed=0):
    g = rng(seed)
    c = g.normal(0, 5, size=(k, d))
    return c[np.arange(n) % k] + g.normal(size=(n, d))


# ---------------------------------------------------------------- RNG
def test_glibc_rand_known_answer(oracle):
    # glibc rand() after srand(1): the classic first values
    oracle.srand(1)
    assert [oracle.rand() for _ in range(5)] == [1804289383, 846930886, 1681692777, 1714636915, 1957747793]


def test_gaussian_random_is_polar_method(oracle):
    # restate the Marsaglia polar step on the same rand() stream
    oracle.srand(7)
    r = [oracle.rand() for _ in range(400)]
    oracle.srand(7)
    g = oracle.gaussian_random(50)
    out, k = [], 0
    RM = 2147483647.0
    while len(out) < 50:
        x = 2 * (r[k] / (RM + 1)) - 1
        y = 2 * (r[k + 1] / (RM + 1)) - 1
        k += 2
        s = x * x + y * y
        if s >= 1.0 or s == 0.0:
            continue
        out.append(x * np.sqrt(-2 * np.log(s) / s))
    np.testing.assert_array_equal(g, np.array(out))


def test_vptree_draw_count(oracle):
    # one draw per VP-tree node holding more than one item
    def count(n):
        if n <= 1:
            return 0
        return 1 + count(n // 2 - 1) + count(n - n // 2)
    for n in (0, 1, 2, 3, 10, 97, 1000):
        assert oracle.vptree_rand_calls(n) == count(n)
    # and the VP-tree build really consumes that many
    X = mixture(257, 4)
    oracle.srand(3)
    oracle.knn_vptree(X, 5)
    after = oracle.rand()
    oracle.srand(3)
    for _ in range(oracle.vptree_rand_calls(257)):
        oracle.rand()
    assert oracle.rand() == after


# ---------------------------------------------------------------- normalise
def test_normalize_closed_form(oracle):
    X = rng(1).normal(size=(200, 7)) * 4 + 2
    Z = X - X.mean(0)
    np.testing.assert_allclose(oracle.normalize(X), Z / Z.max(), rtol=1e-13, atol=1e-15)


# ---------------------------------------------------------------- distances
def test_sqdist_modes(oracle):
    X = rng(2).normal(size=(60, 5))
    np.testing.assert_allclose(oracle.sqdist(X), ssd.cdist(X, X, "sqeuclidean"), rtol=1e-12, atol=1e-12)
    # the reference's overwrite at tsne.hpp:948-950 leaves -2 X X^T
    np.testing.assert_allclose(oracle.sqdist(X, compat=True), -2 * X @ X.T, rtol=1e-12, atol=1e-12)


def test_knn_brute_matches_scipy(oracle):
    X = mixture(400, 6)
    K = 25
    idx, dist = oracle.knn_brute(X, K)
    D = ssd.cdist(X, X, "sqeuclidean")
    np.fill_diagonal(D, np.inf)
    order = np.lexsort((np.broadcast_to(np.arange(400), D.shape), D), axis=1)[:, :K]
    np.testing.assert_array_equal(idx, order)
    np.testing.assert_allclose(dist, np.take_along_axis(D, order, 1), rtol=1e-12)
    # the sampled-rows form answers the same for the rows asked
    rows = np.array([399, 0, 17, 17, 250])
    ri, rd = oracle.knn_brute_rows(X, K, rows)
    np.testing.assert_array_equal(ri, order[rows])
    np.testing.assert_array_equal(rd, dist[rows])


def test_knn_brute_tiny_by_hand(oracle):
    X = np.array([[0.0], [1.0], [3.0], [4.0], [10.0]])
    idx, dist = oracle.knn_brute(X, 2)
    np.testing.assert_array_equal(idx, [[1, 2], [0, 2], [3, 1], [2, 1], [3, 2]])
    np.testing.assert_array_equal(dist, [[1, 9], [1, 4], [1, 4], [1, 9], [36, 49]])


def test_vptree_is_approximate_subset(oracle):
    X = oracle.normalize(mixture(500, 8))
    K = 20
    oracle.srand(1)
    vi, vd = oracle.knn_vptree(X, K)
    bi, bd = oracle.knn_brute(X, K)
    D = ssd.cdist(X, X, "sqeuclidean")
    # every reported distance is the true squared distance, ascending
    np.testing.assert_allclose(vd, np.take_along_axis(D, vi, 1), rtol=1e-12)
    assert np.all(np.diff(vd, axis=1) >= 0)
    recall = np.mean([len(set(a) & set(b)) / K for a, b in zip(vi, bi)])
    # pruning with triangle-inequality tests on squared distances
    # (vptree.hpp:273-292) misses true neighbours: recall is well below 1
    assert 0.6 < recall < 1.0


# ---------------------------------------------------------------- perplexity
def entropy(p):
    p = p[p > 0]
    return -np.sum(p * np.log(p))


def test_perplexity_invariants_and_root(oracle):
    X = oracle.normalize(mixture(300, 10))
    _, dist = oracle.knn_brute(X, 45)
    val, beta, iters, unconv = oracle.perplexity_sparse(dist, 15.0)
    assert unconv == 0
    np.testing.assert_allclose(val.sum(1), 1.0, rtol=1e-14)
    H = np.array([entropy(r) for r in val])
    assert np.all(np.abs(H - np.log(15.0)) < 1e-5)
    # brentq on H(beta) = log(perplexity) brackets the bisection's answer
    for n in range(0, 300, 37):
        d = dist[n]

        def h(b):
            p = np.exp(-b * d)
            p /= p.sum()
            return entropy(p) - np.log(15.0)
        b0 = scipy.optimize.brentq(h, 1e-8, 1e8, xtol=1e-14)
        assert abs(beta[n] - b0) / b0 < 1e-3


def test_perplexity_equidistant_row_is_uniform(oracle):
    dist = np.full((3, 30), 2.5)
    val, beta, iters, unconv = oracle.perplexity_sparse(dist, 10.0)
    # entropy is log(30) until exp approaches underflow, where the DBL_MIN
    # seed of sum_P (tsne.hpp:666) bends it; the row stays uniform throughout
    assert np.all(iters > 20)
    assert np.all(val == val[:, :1])
    assert np.all(val.sum(1) <= 1.0) and np.all(val.sum(1) > 0.99)
    # a row whose distances already give perplexity 10 at beta=1 stops at once
    d = np.concatenate([np.zeros(10), np.full(20, 800.0)])[None, :]
    val, beta, iters, unconv = oracle.perplexity_sparse(d, 10.0)
    assert iters[0] == 1 and beta[0] == 1.0 and unconv == 0
    np.testing.assert_allclose(val[0, :10], 0.1, rtol=1e-12)


def test_perplexity_dense_diagonal(oracle):
    X = oracle.normalize(mixture(120, 6))
    P, beta, unconv = oracle.perplexity_dense(oracle.sqdist(X), 10.0)
    assert unconv == 0
    np.testing.assert_allclose(P.sum(1), 1.0, rtol=1e-13)
    assert np.all(np.diag(P) < 1e-300)


# ---------------------------------------------------------------- symmetrise
def test_symmetrize_matches_scipy(oracle):
    X = oracle.normalize(mixture(250, 6))
    idx, dist = oracle.knn_brute(X, 30)
    val, *_ = oracle.perplexity_sparse(dist, 10.0)
    row, col, sval = oracle.symmetrize(idx, val)
    N = 250
    P = scipy.sparse.csr_matrix((val.ravel(), idx.ravel(), np.arange(N + 1) * 30), shape=(N, N))
    S = ((P + P.T) / 2).toarray()
    mine = scipy.sparse.csr_matrix((sval, col, row), shape=(N, N)).toarray()
    np.testing.assert_allclose(mine, S, rtol=1e-15, atol=0)
    # pattern is exactly the union E u E^T
    assert row[-1] == np.count_nonzero(S)


# ---------------------------------------------------------------- gradients and KL
def _kl_torch(P_dense, Y):
    import torch
    Yt = torch.tensor(Y, dtype=torch.float64, requires_grad=True)
    Pt = torch.tensor(P_dense, dtype=torch.float64)
    D2 = ((Yt[:, None, :] - Yt[None, :, :]) ** 2).sum(-1)
    q = 1.0 / (1.0 + D2)
    q = q - torch.diag(torch.diag(q))
    Q = q / q.sum()
    mask = Pt > 0
    kl = (Pt[mask] * torch.log(Pt[mask] / Q[mask])).sum()
    kl.backward()
    return kl.item(), Yt.grad.numpy()


def _csr_problem(oracle, n=150, perp=10.0):
    row, col, val, _ = oracle.build_p(mixture(n, 6), perp, knn_mode=1)
    Y = rng(3).normal(size=(n, 2)) * 2
    P = scipy.sparse.csr_matrix((val, col, row), shape=(n, n)).toarray()
    return row, col, val, Y, P


def test_exact_gradient_is_quarter_of_autograd(oracle):
    # the reference drops the factor 4 of dKL/dY (tsne.hpp:346, :393-396)
    row, col, val, Y, P = _csr_problem(oracle)
    dC, z = oracle.gradient_exact_sparse(row, col, val, Y)
    kl, g = _kl_torch(P, Y)
    np.testing.assert_allclose(4 * dC, g, rtol=1e-10, atol=1e-13)
    assert abs(oracle.kl_exact_sparse(row, col, val, Y) - kl) < 1e-12
    # dense formula with the same P gives the same gradient
    dC2, z2 = oracle.gradient_exact_dense(P, Y)
    np.testing.assert_allclose(dC2, dC, rtol=1e-10, atol=1e-14)
    assert abs(z - z2) < 1e-9 * z


def test_kl_central_differences(oracle):
    row, col, val, Y, P = _csr_problem(oracle, n=60)
    dC, _ = oracle.gradient_exact_sparse(row, col, val, Y)
    h = 1e-6
    for (i, d) in [(0, 0), (7, 1), (33, 0), (59, 1)]:
        Yp, Ym = Y.copy(), Y.copy()
        Yp[i, d] += h
        Ym[i, d] -= h
        fd = (oracle.kl_exact_sparse(row, col, val, Yp) - oracle.kl_exact_sparse(row, col, val, Ym)) / (2 * h)
        assert abs(fd - 4 * dC[i, d]) < 1e-6 * max(1.0, abs(fd))


def test_bh_theta0_equals_exact(oracle):
    row, col, val, Y, _ = _csr_problem(oracle)
    a, za = oracle.gradient_bh(row, col, val, Y, 0.0)
    b, zb = oracle.gradient_exact_sparse(row, col, val, Y)
    np.testing.assert_allclose(a, b, rtol=1e-10, atol=1e-14)
    assert abs(za - zb) < 1e-10 * zb


def test_bh_error_shrinks_with_theta(oracle):
    row, col, val, Y, _ = _csr_problem(oracle, n=400)
    exact, _ = oracle.gradient_exact_sparse(row, col, val, Y)
    errs = [np.linalg.norm(oracle.gradient_bh(row, col, val, Y, t)[0] - exact) for t in (0.1, 0.4, 0.8)]
    assert errs[0] < errs[1] < errs[2]
    assert errs[1] / np.linalg.norm(exact) < 0.05


# ---------------------------------------------------------------- octree (3-D)
# The octree restatement has no counterpart in the reference (its tree is
# 2-D).  It is pinned by reducing to the reference's quadtree: points on a
# coordinate plane through the mean sit on every split plane of the third
# axis, go to its lower half at every level, and see the quadtree's cells of
# the other two axes (same child order, same centres of mass, max of the
# half-widths unchanged by the third axis' 1e-5): the 3-D gradient and Z
# equal the 2-D ones bit for bit, for each of the three planes.  Its theta=0
# limit is the exact gradient.
@pytest.mark.parametrize("plane", [(0, 1), (0, 2), (1, 2)], ids=["xy", "xz", "yz"])
def test_octree_on_a_plane_is_the_quadtree(oracle, plane):
    row, col, val, Y2, _ = _csr_problem(oracle, n=300)
    Y2 = Y2 - Y2.mean(0)  # the mean of the planar coordinates is exactly 0 on the third axis anyway
    Y3 = np.zeros((len(Y2), 3))
    Y3[:, plane[0]] = Y2[:, 0]
    Y3[:, plane[1]] = Y2[:, 1]
    for theta in (0.3, 0.5, 1.0):
        g2, z2 = oracle.gradient_bh(row, col, val, Y2, theta)
        g3, z3 = oracle.gradient_bh(row, col, val, Y3, theta)
        assert z3 == z2
        np.testing.assert_array_equal(g3[:, plane[0]], g2[:, 0])
        np.testing.assert_array_equal(g3[:, plane[1]], g2[:, 1])
        assert np.all(g3[:, 3 - plane[0] - plane[1]] == 0.0)
        assert oracle.kl_bh(row, col, val, Y3, theta) == oracle.kl_bh(row, col, val, Y2, theta)


def test_octree_theta0_equals_exact(oracle):
    row, col, val, _, _ = _csr_problem(oracle)
    Y = rng(4).normal(size=(len(row) - 1, 3)) * 2
    a, za = oracle.gradient_bh(row, col, val, Y, 0.0)
    b, zb = oracle.gradient_exact_sparse(row, col, val, Y)
    np.testing.assert_allclose(a, b, rtol=1e-10, atol=1e-14)
    assert abs(za - zb) < 1e-10 * zb


def test_octree_error_shrinks_with_theta(oracle):
    row, col, val, _, _ = _csr_problem(oracle, n=400)
    Y = rng(5).normal(size=(400, 3)) * 3
    exact, _ = oracle.gradient_exact_sparse(row, col, val, Y)
    errs = [np.linalg.norm(oracle.gradient_bh(row, col, val, Y, t)[0] - exact) for t in (0.1, 0.4, 0.8)]
    assert errs[0] < errs[1] < errs[2]
    assert errs[1] / np.linalg.norm(exact) < 0.05


def test_octree_structure_by_hand(oracle):
    # two points in opposite corners: the root splits once into two leaves;
    # eight points, one per octant: one root and eight leaves
    n, per = oracle.octree_stats(np.array([[-1.0, -1.0, -1.0], [1.0, 1.0, 1.0]]))
    assert n == 3 and per[0] == 1 and per[1] == 2
    corners = np.array([[x, y, z] for z in (-1.0, 1.0) for y in (-1.0, 1.0) for x in (-1.0, 1.0)])
    n, per = oracle.octree_stats(corners)
    assert n == 9 and per[1] == 8 and per[2:].sum() == 0
    # x of the mean is 0.85/3 and hw = 1 + 0.85/3 + 1e-5: the level-1 cell of
    # the pair spans x in [0.2833, 1.5667] and splits at 0.925, between them
    P = np.array([[0.9, 0.9, 0.9], [0.95, 0.9, 0.9], [-1.0, -1.0, -1.0]])
    n, per = oracle.octree_stats(P)
    assert n == 5 and list(per[:3]) == [1, 2, 2] and per[3:].sum() == 0
    # with (-1, 1, 1) added the x-range of the pair's cells is [-0.0475,
    # 0.91001] at level 1, halved towards the pair: splits at 0.43125,
    # 0.6706, 0.7903, 0.8502, 0.8801, 0.8951 keep 0.9 and 0.91 together, the
    # level-7 split at 0.90256 parts them into two level-8 leaves
    P = np.array([[0.9, 0.9, 0.9], [0.91, 0.9, 0.9], [-1.0, -1.0, -1.0], [-1.0, 1.0, 1.0]])
    n, per = oracle.octree_stats(P)
    assert list(per[:10]) == [1, 3, 1, 1, 1, 1, 1, 1, 2, 0] and n == 12


# ---------------------------------------------------------------- Barnes-Hut acceptance
# Eight points on the integer grid.  Their mean (-1.375, 4.75) is exact, so
# the root box is exact up to the +1e-5 of quadtree.hpp:128-129: it is taller
# than wide (hw = 17.625 + 1e-5, hh = 34.75 + 1e-5).  Points 1 and 3,
# (10, 36) and (13, 36), share a level-2 cell C with half-sizes (hw/4, hh/4)
# whose two children are single-point leaves; C's centre of mass (11.5, 36)
# is exact (two points: the online update multiplies by 0 and 1/2).  For the
# query points p in BH_QUERIES every other cell decision of every point lies
# more than 5% away from C's ratio, so moving theta across that ratio flips
# exactly one decision: p sees C as one summary or as its two leaves.
BH_Y = np.array([[-1, 0], [10, 36], [-19, -30], [13, 36], [-10, -16], [15, -6], [-9, 26], [-10, -8]], float)
BH_QUERIES = (2, 4, 5, 7)


def _bh_neg_z(oracle, Y, theta):
    # with an empty P, dC = -neg_f / sum_Q (tsne.hpp:346)
    row = np.zeros(len(Y) + 1, np.int32)
    dC, z = oracle.gradient_bh(row, np.zeros(0, np.int32), np.zeros(0), Y, theta)
    return -dC * z, z


@pytest.mark.parametrize("transpose", [False, True], ids=["tall_root", "wide_root"])
@pytest.mark.parametrize("p", BH_QUERIES)
def test_bh_acceptance_boundary_by_hand(oracle, p, transpose):
    """quadtree.hpp:335: a cell is a summary iff max(hh, hw) / sqrt(D) < theta,
    D the squared distance to its centre of mass.  Worked out by hand for
    cell C and query p: the change of Z and of p's repulsion when theta
    crosses C's ratio is one summary term minus two leaf terms.  A cell test
    with the full width, with only hw (or only hh), with D for sqrt(D), or with
    <= for < fails here; the transposed point set makes the root wider than
    tall."""
    Y = BH_Y[:, ::-1].copy() if transpose else BH_Y.copy()
    m = Y.mean(0)
    half = (np.abs(Y - m).max(0) + 1e-5) / 4.0        # C is two levels down
    com = 0.5 * (Y[1] + Y[3])
    d = Y[p] - com
    D = d[0] * d[0] + d[1] * d[1]
    r = max(half[0], half[1]) / np.sqrt(D)            # C's acceptance ratio

    def q(a):
        return 1.0 / (1.0 + np.sum((Y[p] - a) ** 2))

    dz = 2 * q(com) - q(Y[1]) - q(Y[3])
    dn = 2 * q(com) ** 2 * (Y[p] - com) - q(Y[1]) ** 2 * (Y[p] - Y[1]) - q(Y[3]) ** 2 * (Y[p] - Y[3])
    assert abs(dz) > 1e-4 * q(com)                    # the flip is visible

    n_lo, z_lo = _bh_neg_z(oracle, Y, r * (1 - 1e-9))  # C opened: two leaves
    n_hi, z_hi = _bh_neg_z(oracle, Y, r * (1 + 1e-9))  # C accepted: one summary
    n_at, z_at = _bh_neg_z(oracle, Y, r)               # strict <: still opened
    np.testing.assert_allclose(z_hi - z_lo, dz, rtol=1e-8)
    np.testing.assert_allclose((n_hi - n_lo)[p], dn, rtol=1e-8)
    others = np.arange(len(Y)) != p
    np.testing.assert_allclose((n_hi - n_lo)[others], 0.0, atol=1e-15)
    assert z_at == z_lo and np.array_equal(n_at, n_lo)

    # the reference KL (tsne.hpp:454-486) uses the same Z: one edge p -> 0
    v = 0.25
    row = np.zeros(len(Y) + 1, np.int32)
    row[p + 1:] = 1
    fm = np.finfo(np.float32).tiny
    for theta, z in ((r * (1 - 1e-9), z_lo), (r * (1 + 1e-9), z_hi), (r, z_at)):
        C = oracle.kl_bh(row, np.array([0], np.int32), np.array([v]), Y, theta)
        qp = 1.0 / (1.0 + np.sum((Y[p] - Y[0]) ** 2))
        assert abs(C - v * np.log((v + fm) / (qp / z + fm))) < 1e-13


def test_bh_root_summary_closed_form(oracle):
    """With theta above every root ratio each point sees the whole tree as
    one summary of N points at the mean, itself included (quadtree.hpp:335-343):
    Z = sum_p N q(p, mean), neg_f[p] = N q(p, mean)^2 (p - mean)."""
    Y = BH_Y.copy()
    m = Y.mean(0)
    qm = 1.0 / (1.0 + np.sum((Y - m) ** 2, axis=1))
    neg, z = _bh_neg_z(oracle, Y, 1e6)
    np.testing.assert_allclose(z, len(Y) * qm.sum(), rtol=1e-13)
    np.testing.assert_allclose(neg, len(Y) * (qm ** 2)[:, None] * (Y - m), rtol=1e-12, atol=1e-16)
    # and just below the smallest root ratio the root is opened for that point
    s = np.abs(Y - m).max() + 1e-5
    rmin = (s / np.sqrt(np.sum((Y - m) ** 2, axis=1))).min()
    _, z2 = _bh_neg_z(oracle, Y, rmin * (1 - 1e-9))
    assert abs(z2 - z) > 1e-6 * z


def test_quadtree_cells_by_membership(oracle):
    Y = rng(5).normal(size=(300, 2))
    t = oracle.quadtree(Y)
    cx, cy, hw, hh = t["root_box"]
    np.testing.assert_allclose([cx, cy], Y.mean(0), atol=1e-14)
    np.testing.assert_allclose(hw, max(Y[:, 0].max() - cx, cx - Y[:, 0].min()) + 1e-5, rtol=1e-15)
    np.testing.assert_allclose(hh, max(Y[:, 1].max() - cy, cy - Y[:, 1].min()) + 1e-5, rtol=1e-15)
    tx = (Y[:, 0] - (cx - hw)) / (2 * hw)
    ty = (Y[:, 1] - (cy - hh)) / (2 * hh)
    pos = [0]
    n_leaf = [0]

    # the export is pre-order with all four children (NW, NE, SW, SE) of
    # every split cell; points on a split line belong to the west/north cell
    def walk(level, ix, iy):
        k = pos[0]
        pos[0] += 1
        assert t["level"][k] == level
        qx = np.ceil(tx * 2.0 ** level) - 1
        qy = np.ceil(ty * 2.0 ** level) - 1
        members = np.flatnonzero((qx == ix) & (qy == iy))
        assert t["cum"][k] == len(members)
        if len(members):
            np.testing.assert_allclose([t["comx"][k], t["comy"][k]], Y[members].mean(0), rtol=0, atol=1e-13)
        if t["leaf"][k]:
            assert len(members) <= 1
            if len(members):
                assert t["index"][k] == members[0]
                n_leaf[0] += 1
        else:
            assert len(members) >= 2
            for c in range(4):
                walk(level + 1, 2 * ix + (c & 1), 2 * iy + (c >> 1))

    walk(0, 0, 0)
    assert pos[0] == len(t["level"]) and n_leaf[0] == 300


def test_kl_dense_closed_form(oracle):
    row, col, val, Y, P = _csr_problem(oracle, n=80)
    C = oracle.kl_exact_dense(P, Y)
    D2 = ssd.cdist(Y, Y, "sqeuclidean")
    q = 1 / (1 + D2)
    np.fill_diagonal(q, 0)
    sq = np.finfo(float).tiny + q.sum()
    Q = q / sq
    np.fill_diagonal(Q, np.finfo(float).tiny / sq)
    ref = np.sum(P * np.log((P + 1e-9) / (Q + 1e-9)))
    assert abs(C - ref) < 1e-10


def test_kl_bh_formula(oracle):
    row, col, val, Y, P = _csr_problem(oracle, n=120)
    # at theta=0 the BH estimate of Z is exact; the sum runs over the pattern
    C = oracle.kl_bh(row, col, val, Y, 0.0)
    D2 = ssd.cdist(Y, Y, "sqeuclidean")
    q = 1 / (1 + D2)
    np.fill_diagonal(q, 0)
    Z = q.sum()
    fm = np.finfo(np.float32).tiny
    ref = 0.0
    for i in range(120):
        for e in range(row[i], row[i + 1]):
            Qv = q[i, col[e]] / Z
            ref += val[e] * np.log((val[e] + fm) / (Qv + fm))
    assert abs(C - ref) < 1e-11


def test_update_recurrence(oracle):
    dY = np.array([[1.0, -2.0], [0.0, 3.0], [-1.0, 0.0]])
    uY = np.array([[0.5, -0.5], [1.0, 0.0], [0.0, 0.0]])
    gains = np.array([[1.0, 1.0], [0.01, 0.5], [0.012, 1.0]])
    Y = np.array([[1.0, 2.0], [3.0, 4.0], [5.0, 6.0]])
    u, g, y = oracle.update(dY, uY, gains, Y, 0.5)
    # sign(dY) != sign(uY) -> +0.2 else *0.8, floor 0.01; sign(0) = 0
    eg = np.array([[0.8, 0.8], [1.0 * 0.01 + 0.2 - 0.0, 0.7], [0.212, 0.8]])
    eg[1, 0] = 0.01 + 0.2   # sign(0)=0 != sign(1)=1
    eg[1, 1] = 0.5 + 0.2    # sign(3)=1 != sign(0)=0
    eg[2, 0] = 0.012 + 0.2  # sign(-1) != sign(0)
    eg[2, 1] = 0.8          # sign(0) == sign(0)
    np.testing.assert_allclose(g, eg, rtol=1e-15)
    eu = 0.5 * uY - 200.0 * eg * dY
    np.testing.assert_allclose(u, eu, rtol=1e-15)
    ey = Y + eu
    np.testing.assert_allclose(y, ey - ey.mean(0), rtol=1e-14, atol=1e-12)


def test_theta0_reference_path_breaks(oracle):
    # documents the reference defect: with its DD overwrite the exact path
    # diverges (NaN KL), with textbook distances it converges
    X = mixture(150, 8)
    _, info_c = oracle.tsne_run(X, perplexity=10.0, theta=0.0, max_iter=120, dd_compat=True, log_every=50,
                                y0=rng(0).normal(size=(150, 2)) * 1e-4)
    _, info_t = oracle.tsne_run(X, perplexity=10.0, theta=0.0, max_iter=120, log_every=50,
                                y0=rng(0).normal(size=(150, 2)) * 1e-4)
    assert np.all(np.isfinite(info_t["kl"]))
    assert not np.all(np.isfinite(info_c["kl"])) or info_c["kl"][-1] > 2 * info_t["kl"][-1]


def test_full_run_decreases_kl(oracle):
    X = mixture(200, 8)
    Y, info = oracle.tsne_run(X, perplexity=10.0, theta=0.5, max_iter=400, knn_mode=1,
                              y0=rng(1).normal(size=(200, 2)) * 1e-4)
    kl = info["kl"]
    assert kl[-1] < kl[4] and kl[-1] < 2.0
    assert abs(Y.mean(0)).max() < 1e-12
