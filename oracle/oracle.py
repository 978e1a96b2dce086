"""ctypes front-end of the FP64 CPU restatement (oracle/tsne_oracle.cpp).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, always as the checker or the timed CPU baseline,
never by the product package.  Every function mirrors one of tapkee's
(include/tapkee/external/barnes_hut_sne/tsne.hpp, quadtree.hpp, vptree.hpp);
the citation is on the C++ side.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_SO = os.path.join(_HERE, "_build", "liboracle_tsne.so")
_lib = None

_dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_ip = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
_f64p = C.POINTER(C.c_double)


def build(force: bool = False) -> str:
    """Compile the restatement with its Makefile (needs g++)."""
    cmd = ["make", "-C", _HERE]
    if force:
        cmd.append("-B")
    subprocess.run(cmd, check=True, stdout=subprocess.DEVNULL)
    return _SO


def use_native() -> str:
    """Build the restatement for this host's CPU (-march=native, same
    -ffp-contract=off arithmetic) and load it instead of the portable build;
    for timing on the machine that runs it.  Returns the library path."""
    global _SO, _lib
    subprocess.run(["make", "-B", "-C", _HERE, "native"], check=True, stdout=subprocess.DEVNULL)
    _SO = os.path.join(_HERE, "_build", "liboracle_tsne_native.so")
    _lib = None
    lib()
    return _SO


class RunOptions(C.Structure):
    _fields_ = [
        ("max_iter", C.c_int), ("stop_lying_iter", C.c_int), ("mom_switch_iter", C.c_int),
        ("knn_mode", C.c_int), ("exact_p_mode", C.c_int), ("dd_compat", C.c_int),
        ("y0_mode", C.c_int), ("y0", C.POINTER(C.c_double)), ("log_every", C.c_int),
        ("kl_trace", C.POINTER(C.c_double)), ("kl_iter", C.POINTER(C.c_int)),
        ("kl_cap", C.c_int), ("kl_count", C.c_int),
        ("ms_p", C.c_double), ("ms_loop", C.c_double), ("nnz", C.c_longlong),
    ]


def lib():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(_SO):
        build()
    L = C.CDLL(_SO)
    L.or_gaussian_random.restype = C.c_double
    L.or_uniform_random.restype = C.c_double
    L.or_srand.argtypes = [C.c_uint]
    L.or_rand.restype = C.c_int
    L.or_vptree_rand_calls.argtypes = [C.c_longlong]
    L.or_vptree_rand_calls.restype = C.c_longlong
    L.or_zero_mean.argtypes = [_dp, C.c_int, C.c_int]
    L.or_normalize.argtypes = [_dp, C.c_int, C.c_int]
    L.or_sqdist.argtypes = [_dp, C.c_int, C.c_int, _dp, C.c_int]
    L.or_knn_brute.argtypes = [_dp, C.c_int, C.c_int, C.c_int, _ip, _dp]
    L.or_knn_brute_rows.argtypes = [_dp, C.c_int, C.c_int, C.c_int, _ip, C.c_int, _ip, _dp]
    L.or_vptree_knn.argtypes = [_dp, C.c_int, C.c_int, C.c_int, _ip, _dp]
    L.or_vptree_knn_sample.argtypes = [_dp, C.c_int, C.c_int, C.c_int, C.c_int, _ip, _dp, _f64p, _f64p]
    L.or_vptree_knn_sample.restype = C.c_int
    L.or_perplexity_sparse.argtypes = [_dp, C.c_int, C.c_int, C.c_double, _dp, _dp, _ip]
    L.or_perplexity_sparse.restype = C.c_int
    L.or_perplexity_dense.argtypes = [_dp, C.c_int, C.c_double, _dp, _dp]
    L.or_perplexity_dense.restype = C.c_int
    L.or_symmetrize.argtypes = [_ip, _ip, _dp, C.c_int, _ip, _ip, _dp]
    L.or_symmetrize.restype = C.c_longlong
    L.or_gradient_bh.argtypes = [_ip, _ip, _dp, _dp, C.c_int, C.c_double, _dp]
    L.or_gradient_bh.restype = C.c_double
    L.or_gradient_bh3.argtypes = [_ip, _ip, _dp, _dp, C.c_int, C.c_double, _dp]
    L.or_gradient_bh3.restype = C.c_double
    L.or_kl_bh3.argtypes = [_ip, _ip, _dp, _dp, C.c_int, C.c_double]
    L.or_kl_bh3.restype = C.c_double
    L.or_octree_stats.argtypes = [_dp, C.c_int, C.c_int, _ip]
    L.or_octree_stats.restype = C.c_int
    L.or_quadtree_export.argtypes = [_dp, C.c_int, C.c_int, _ip, _ip, _ip, _ip, _dp, _dp, _ip, _dp]
    L.or_quadtree_export.restype = C.c_int
    L.or_gradient_exact_dense.argtypes = [_dp, _dp, C.c_int, C.c_int, _dp, C.c_int]
    L.or_gradient_exact_dense.restype = C.c_double
    L.or_gradient_exact_sparse.argtypes = [_ip, _ip, _dp, _dp, C.c_int, C.c_int, _dp]
    L.or_gradient_exact_sparse.restype = C.c_double
    L.or_kl_exact_dense.argtypes = [_dp, _dp, C.c_int, C.c_int, C.c_int]
    L.or_kl_exact_dense.restype = C.c_double
    L.or_kl_bh.argtypes = [_ip, _ip, _dp, _dp, C.c_int, C.c_double]
    L.or_kl_bh.restype = C.c_double
    L.or_kl_exact_sparse.argtypes = [_ip, _ip, _dp, _dp, C.c_int, C.c_int]
    L.or_kl_exact_sparse.restype = C.c_double
    L.or_update.argtypes = [_dp, _dp, _dp, _dp, C.c_int, C.c_int, C.c_double, C.c_double]
    L.or_tsne_run.argtypes = [_dp, C.c_int, C.c_int, _dp, C.c_int, C.c_double, C.c_double, C.POINTER(RunOptions)]
    L.or_tsne_run.restype = C.c_int
    L.or_build_p.argtypes = [_dp, C.c_int, C.c_int, C.c_double, C.c_int, C.c_int, _ip, _ip, _dp]
    L.or_build_p.restype = C.c_longlong
    L.or_run_options_size.restype = C.c_int
    assert L.or_run_options_size() == C.sizeof(RunOptions), "RunOptions layout mismatch"
    _lib = L
    return L


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def _i32(a):
    return np.ascontiguousarray(a, dtype=np.int32)


# ---------------------------------------------------------------- RNG
def srand(seed: int) -> None:
    lib().or_srand(seed)


def rand() -> int:
    return lib().or_rand()


def gaussian_random(n: int) -> np.ndarray:
    L = lib()
    return np.array([L.or_gaussian_random() for _ in range(n)])


def vptree_rand_calls(n: int) -> int:
    return lib().or_vptree_rand_calls(n)


# ---------------------------------------------------------------- steps
def normalize(X):
    X = _f64(X).copy()
    lib().or_normalize(X, X.shape[0], X.shape[1])
    return X


def sqdist(X, compat=False):
    X = _f64(X)
    DD = np.empty((X.shape[0], X.shape[0]))
    lib().or_sqdist(X, X.shape[0], X.shape[1], DD, int(compat))
    return DD


def knn_brute(X, K):
    X = _f64(X)
    N = X.shape[0]
    idx = np.empty((N, K), np.int32)
    dist = np.empty((N, K))
    lib().or_knn_brute(X, N, X.shape[1], K, idx, dist)
    return idx, dist


def knn_brute_rows(X, K, rows):
    """Exact kNN of the query rows `rows` only (FP64, ties by index)."""
    X = _f64(X)
    rows = _i32(rows)
    idx = np.empty((len(rows), K), np.int32)
    dist = np.empty((len(rows), K))
    lib().or_knn_brute_rows(X, X.shape[0], X.shape[1], K, rows, len(rows), idx, dist)
    return idx, dist


def knn_vptree(X, K):
    X = _f64(X)
    N = X.shape[0]
    idx = np.empty((N, K), np.int32)
    dist = np.empty((N, K))
    lib().or_vptree_knn(X, N, X.shape[1], K, idx, dist)
    return idx, dist


def knn_vptree_sample(X, K, stride):
    """VP-tree over all rows, searches of rows 0, stride, 2*stride, ... only.
    Returns (idx, dist, seconds_build, seconds_search) for the timed sample."""
    X = _f64(X)
    N = X.shape[0]
    nq = (N + stride - 1) // stride
    idx = np.empty((nq, K), np.int32)
    dist = np.empty((nq, K))
    tb, ts = C.c_double(0), C.c_double(0)
    lib().or_vptree_knn_sample(X, N, X.shape[1], K, stride, idx, dist, C.byref(tb), C.byref(ts))
    return idx, dist, tb.value, ts.value


def perplexity_sparse(dist, perplexity):
    dist = _f64(dist)
    N, K = dist.shape
    val = np.empty((N, K))
    beta = np.empty(N)
    iters = np.empty(N, np.int32)
    unconv = lib().or_perplexity_sparse(dist, N, K, perplexity, val, beta, iters)
    return val, beta, iters, unconv


def perplexity_dense(DD, perplexity):
    DD = _f64(DD)
    N = DD.shape[0]
    P = np.empty((N, N))
    beta = np.empty(N)
    unconv = lib().or_perplexity_dense(DD, N, perplexity, P, beta)
    return P, beta, unconv


def symmetrize(idx, val):
    """CSR symmetrisation of an N x K neighbour list (reference row layout)."""
    idx = _i32(idx)
    val = _f64(val)
    N, K = idx.shape
    row = _i32(np.arange(N + 1) * K)
    srow = np.empty(N + 1, np.int32)
    scol = np.empty(2 * N * K, np.int32)
    sval = np.empty(2 * N * K)
    nnz = lib().or_symmetrize(row, idx.reshape(-1), val.reshape(-1), N, srow, scol, sval)
    return srow, scol[:nnz].copy(), sval[:nnz].copy()


def build_p(X, perplexity, K=None, knn_mode=1):
    """normalise -> kNN -> perplexity -> symmetrise -> /sum (no x12)."""
    X = _f64(X).copy()
    N, D = X.shape
    K = int(3 * perplexity) if K is None else K
    row = np.empty(N + 1, np.int32)
    col = np.empty(2 * N * K, np.int32)
    val = np.empty(2 * N * K)
    nnz = lib().or_build_p(X, N, D, perplexity, K, knn_mode, row, col, val)
    return row, col[:nnz].copy(), val[:nnz].copy(), X


def gradient_bh(row, col, val, Y, theta):
    """Barnes-Hut gradient (no factor 4) and Z: the reference's quadtree for
    N x 2 embeddings, its octree generalisation for N x 3."""
    Y = _f64(Y)
    dC = np.empty_like(Y)
    fn = lib().or_gradient_bh3 if Y.shape[1] == 3 else lib().or_gradient_bh
    assert Y.shape[1] in (2, 3)
    z = fn(_i32(row), _i32(col), _f64(val), Y, Y.shape[0], theta, dC)
    return dC, z


def gradient_exact_dense(P, Y, compat=False):
    Y = _f64(Y)
    dC = np.empty_like(Y)
    z = lib().or_gradient_exact_dense(_f64(P), Y, Y.shape[0], Y.shape[1], dC, int(compat))
    return dC, z


def gradient_exact_sparse(row, col, val, Y):
    Y = _f64(Y)
    dC = np.empty_like(Y)
    z = lib().or_gradient_exact_sparse(_i32(row), _i32(col), _f64(val), Y, Y.shape[0], Y.shape[1], dC)
    return dC, z


def kl_exact_dense(P, Y, compat=False):
    Y = _f64(Y)
    return lib().or_kl_exact_dense(_f64(P), Y, Y.shape[0], 2, int(compat))


def kl_bh(row, col, val, Y, theta):
    Y = _f64(Y)
    assert Y.shape[1] in (2, 3)
    fn = lib().or_kl_bh3 if Y.shape[1] == 3 else lib().or_kl_bh
    return fn(_i32(row), _i32(col), _f64(val), Y, Y.shape[0], theta)


def octree_stats(Y, cap=64):
    """Octree of N x 3 points: (non-empty node count, non-empty cells per level)."""
    Y = _f64(Y)
    per = np.empty(cap, np.int32)
    n = lib().or_octree_stats(Y, Y.shape[0], cap, per)
    return n, per


def kl_exact_sparse(row, col, val, Y):
    Y = _f64(Y)
    return lib().or_kl_exact_sparse(_i32(row), _i32(col), _f64(val), Y, Y.shape[0], Y.shape[1])


def update(dY, uY, gains, Y, momentum, eta=200.0):
    uY = _f64(uY).copy()
    gains = _f64(gains).copy()
    Y = _f64(Y).copy()
    lib().or_update(_f64(dY), uY, gains, Y, Y.shape[0], Y.shape[1], momentum, eta)
    return uY, gains, Y


def quadtree(Y):
    """Pointer-quadtree export: dict of per-node arrays in pre-order."""
    Y = _f64(Y)
    N = Y.shape[0]
    cap = 64 * N + 64
    out = {k: np.empty(cap, np.int32) for k in ("level", "cum", "leaf", "index", "nchild")}
    comx = np.empty(cap)
    comy = np.empty(cap)
    box = np.empty(4)
    n = lib().or_quadtree_export(Y, N, cap, out["level"], out["cum"], out["leaf"], out["index"],
                                 comx, comy, out["nchild"], box)
    assert n <= cap
    res = {k: v[:n].copy() for k, v in out.items()}
    res["comx"] = comx[:n].copy()
    res["comy"] = comy[:n].copy()
    res["root_box"] = box
    return res


def tsne_run(X, no_dims=2, perplexity=30.0, theta=0.5, max_iter=1000, knn_mode=0, exact_p_mode=0,
             dd_compat=False, y0=None, log_every=50, stop_lying_iter=250, mom_switch_iter=250):
    """Full reference schedule (tsne.hpp:61-205).  Returns (Y, info)."""
    X = _f64(X).copy()
    N, D = X.shape
    Y = np.empty((N, no_dims))
    o = RunOptions()
    o.max_iter, o.stop_lying_iter, o.mom_switch_iter = max_iter, stop_lying_iter, mom_switch_iter
    o.knn_mode, o.exact_p_mode, o.dd_compat = knn_mode, exact_p_mode, int(dd_compat)
    keep = []
    if y0 is not None:
        y0 = _f64(y0).reshape(N, no_dims)
        keep.append(y0)
        o.y0_mode = 1
        o.y0 = y0.ctypes.data_as(C.POINTER(C.c_double))
    o.log_every = log_every
    cap = 64
    kl = np.zeros(cap)
    kli = np.zeros(cap, np.int32)
    o.kl_trace = kl.ctypes.data_as(C.POINTER(C.c_double))
    o.kl_iter = kli.ctypes.data_as(C.POINTER(C.c_int))
    o.kl_cap = cap
    lib().or_tsne_run(X, N, D, Y, no_dims, perplexity, theta, C.byref(o))
    n = min(o.kl_count, cap)
    return Y, {"kl": kl[:n].copy(), "kl_iter": kli[:n].copy(), "ms_p": o.ms_p, "ms_loop": o.ms_loop,
               "nnz": o.nnz}
