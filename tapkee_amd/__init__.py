"""MI355X t-SNE engine with tapkee's Python interface.

`embed(data, method='t-sne', sne_perplexity=..., sne_theta=..., target_dimension=...)`
mirrors tapkee's Python binding (src/python/nanobind_extension.cpp:28-96):
`data` is features x samples, the result is samples x target_dimension.
The low-level chain `withParameters(ParametersSet).embedUsing(matrix)` mirrors
nanobind_extension.cpp:153-176 / include/tapkee/chain_interface.hpp:471-500.
All work runs on the GPU through libtapkee_amd.so; there is no CPU path.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib
from ._lib import (CancelledError, DeviceError, EngineError, MissedParameterError, MultipleParameterError,
                   NoDataError, NotEnoughMemoryError, UnsupportedMethodError, WrongParameterError)

__all__ = [
    "embed", "tsne", "Session", "withParameters", "ParametersSet", "Parameter", "TapkeeOutput", "DimensionReductionMethod",
    "parse_reduction_method", "knn", "normalize", "perplexity", "perplexity_dense", "symmetrize", "build_p",
    "gradient",
    "gradient_dense", "kl", "kl_dense", "update", "bh_tree", "find_neighbors", "neighbors_connected", "set_logger",
    "EngineError", "WrongParameterError", "NoDataError",
    "NotEnoughMemoryError", "CancelledError", "UnsupportedMethodError", "DeviceError", "MissedParameterError",
    "MultipleParameterError",
]


# ------------------------------------------------------------------ methods
class DimensionReductionMethod:
    """A method object (include/tapkee/defines/methods.hpp); only t-SNE runs here."""

    def __init__(self, name: str, supported: bool):
        self.name = name
        self.supported = supported

    def __eq__(self, other):
        return isinstance(other, DimensionReductionMethod) and other.name == self.name

    def __hash__(self):
        return hash(self.name)

    def __repr__(self):
        return f"DimensionReductionMethod({self.name!r})"


tDistributedStochasticNeighborEmbedding = DimensionReductionMethod(
    "t-distributed Stochastic Neighbor Embedding (t-SNE)", True)

# Names accepted by tapkee's parser (include/tapkee/defines/method_names.hpp);
# every method but t-SNE is outside this engine and raises on use.
_OTHER = {
    "local_tangent_space_alignment": "Local Tangent Space Alignment", "ltsa": "Local Tangent Space Alignment",
    "locally_linear_embedding": "Locally Linear Embedding", "lle": "Locally Linear Embedding",
    "hessian_locally_linear_embedding": "Hessian Locally Linear Embedding", "hlle": "Hessian Locally Linear Embedding",
    "multidimensional_scaling": "Multidimensional Scaling", "mds": "Multidimensional Scaling",
    "landmark_multidimensional_scaling": "Landmark Multidimensional Scaling",
    "l-mds": "Landmark Multidimensional Scaling", "isomap": "Isomap", "landmark_isomap": "Landmark Isomap",
    "l-isomap": "Landmark Isomap", "diffusion_map": "Diffusion map", "dm": "Diffusion map",
    "kernel_pca": "Kernel Principal Component Analysis", "kpca": "Kernel Principal Component Analysis",
    "pca": "Principal Component Analysis", "random_projection": "Random Projection", "ra": "Random Projection",
    "laplacian_eigenmaps": "Laplacian Eigenmaps", "la": "Laplacian Eigenmaps",
    "locality_preserving_projections": "Locality Preserving Projections",
    "lpp": "Locality Preserving Projections",
    "neighborhood_preserving_embedding": "Neighborhood Preserving Embedding",
    "npe": "Neighborhood Preserving Embedding",
    "linear_local_tangent_space_alignment": "Linear Local Tangent Space Alignment",
    "lltsa": "Linear Local Tangent Space Alignment",
    "stochastic_proximity_embedding": "Stochastic Proximity Embedding", "spe": "Stochastic Proximity Embedding",
    "passthru": "passthru", "factor_analysis": "Factor Analysis", "fa": "Factor Analysis",
    "manifold_sculpting": "manifold sculpting", "ms": "manifold sculpting",
}
_METHODS = {"t-stochastic_proximity_embedding": tDistributedStochasticNeighborEmbedding,
            "t-sne": tDistributedStochasticNeighborEmbedding}
for _k, _v in _OTHER.items():
    _METHODS[_k] = DimensionReductionMethod(_v, False)


def parse_reduction_method(name: str) -> DimensionReductionMethod:
    """Method name -> method object; unknown names raise (src/cli/util.hpp:192-195)."""
    try:
        return _METHODS[name]
    except KeyError:
        raise ValueError(f"Unknown dimension reduction method '{name}'") from None


# ------------------------------------------------------------------ parameters
# Keyword names and defaults of include/tapkee/defines/keywords.hpp.
METHOD = "dimension reduction method"
TARGET_DIMENSION = "target dimension"
SNE_PERPLEXITY = "SNE perplexity"
SNE_THETA = "SNE theta"
CANCEL_FUNCTION = "cancel function"
_DEFAULTS = {TARGET_DIMENSION: 2, SNE_PERPLEXITY: 30.0, SNE_THETA: 0.5, CANCEL_FUNCTION: None}


class Parameter:
    """A named value (stichwort::Parameter)."""

    def __init__(self, name: str, value):
        self.name = name
        self.value = value

    @staticmethod
    def create(name: str, value) -> "Parameter":
        return Parameter(name, value)


class ParametersSet:
    """Keyword set with duplicate detection (stichwort/parameter.hpp:232-311)."""

    def __init__(self):
        self._map = {}
        self._dups = []

    def add(self, p: Parameter) -> "ParametersSet":
        if p.name in self._map:
            self._dups.append(p.name)
        self._map[p.name] = p
        return self

    def check(self) -> None:
        if self._dups:
            raise MultipleParameterError("The following parameters are set more than once: " +
                                         " ".join(self._dups) + " ")

    def __contains__(self, name):
        return name in self._map

    def __getitem__(self, name):
        if name not in self._map:
            raise MissedParameterError(name + " is missed")
        return self._map[name].value

    def merged(self, defaults: dict) -> "ParametersSet":
        out = ParametersSet()
        out._map = dict(self._map)
        for k, v in defaults.items():
            if k not in out._map:
                out._map[k] = Parameter(k, v)
        return out


class TapkeeOutput:
    """Embedding (samples x target_dimension) plus engine statistics."""

    def __init__(self, embedding: np.ndarray, stats: dict | None = None):
        self.embedding = embedding
        self.stats = stats or {}


class ParametersInitializedState:
    def __init__(self, parameters: ParametersSet, **engine_options):
        self.parameters = parameters
        self.engine_options = engine_options

    def embedUsing(self, matrix) -> TapkeeOutput:  # noqa: N802 (tapkee's name)
        """matrix: features x samples (DenseMatrix layout)."""
        p = self.parameters
        p.check()                                   # embed.hpp:103
        if METHOD not in p:                          # parameter.hpp:282-292
            raise MissedParameterError(METHOD + " is missed")
        q = p.merged(_DEFAULTS)                      # embed.hpp:104
        m = q[METHOD]
        if isinstance(m, str):
            m = parse_reduction_method(m)
        if not (isinstance(m, DimensionReductionMethod) and m.supported):
            raise UnsupportedMethodError(f"{getattr(m, 'name', m)} is not provided by this engine "
                                         "(only t-SNE runs on it)")
        cancel = q[CANCEL_FUNCTION]
        if cancel is not None and cancel():         # methods.hpp:30-31
            raise CancelledError("cancelled")
        X = np.asarray(matrix, dtype=np.float64)
        if X.ndim != 2 or X.shape[1] == 0:
            raise NoDataError("No data to embed")
        Y, stats = tsne(X.T, no_dims=int(q[TARGET_DIMENSION]), perplexity=float(q[SNE_PERPLEXITY]),
                        theta=float(q[SNE_THETA]), **self.engine_options)
        return TapkeeOutput(Y, stats)


def withParameters(parameters: ParametersSet, **engine_options) -> ParametersInitializedState:  # noqa: N802
    """tapkee::with (chain_interface.hpp:497-500)."""
    return ParametersInitializedState(parameters, **engine_options)


def embed(data, method: str = "lle", neighbors_method=None, eigen_method=None, num_neighbors=None,
          target_dimension=None, gaussian_kernel_width=None, landmark_ratio=None, max_iteration=None,
          diffusion_map_timesteps=None, sne_perplexity=None, sne_theta=None, squishing_rate=None,
          spe_global_strategy=None, spe_num_updates=None, spe_tolerance=None, nullspace_shift=None,
          klle_shift=None, fa_epsilon=None, check_connectivity=None, **engine_options) -> np.ndarray:
    """tapkee.embed (nanobind_extension.cpp:28-96): features x samples -> samples x target_dimension.

    Keywords that only other methods use are accepted and ignored, as in the
    reference.  Extra keywords are engine options (see `tsne`).
    """
    params = ParametersSet()
    params.add(Parameter.create(METHOD, parse_reduction_method(method)))
    if target_dimension is not None:
        params.add(Parameter.create(TARGET_DIMENSION, int(target_dimension)))
    if sne_perplexity is not None:
        params.add(Parameter.create(SNE_PERPLEXITY, float(sne_perplexity)))
    if sne_theta is not None:
        params.add(Parameter.create(SNE_THETA, float(sne_theta)))
    return withParameters(params, **engine_options).embedUsing(data).embedding


# ------------------------------------------------------------------ engine
def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def _i32(a):
    return np.ascontiguousarray(a, dtype=np.int32)


def _options(N, no_dims, dist_dtype="bf16", loop_precision="fp32", candidates=0, knn_k=0, max_iter=1000,
             stop_lying_iter=250, mom_switch_iter=250, exact_p="dense", dd_mode="textbook", y0=None, seed=None,
             log_every=50, kl_mode="reference", verbose=False, device=0, device_mem_cap=0, cancel=None):
    o = _lib.default_options()
    o.max_iter, o.stop_lying_iter, o.mom_switch_iter = max_iter, stop_lying_iter, mom_switch_iter
    o.dist_dtype = _lib.DIST_NAMES[dist_dtype]
    o.loop_precision = {"fp32": _lib.FP32, "fp64": _lib.FP64}[loop_precision]
    o.candidates = int(candidates)
    o.knn_k = int(knn_k)
    o.exact_p_mode = {"dense": _lib.EXACT_P_DENSE, "knn": _lib.EXACT_P_KNN}[exact_p]
    o.dd_mode = {"textbook": _lib.DD_TEXTBOOK, "tapkee_compat": _lib.DD_TAPKEE_COMPAT}[dd_mode]
    o.log_every = int(log_every)
    o.kl_mode = {"reference": _lib.KL_AS_REFERENCE, "exact": _lib.KL_EXACT_Z}[kl_mode]
    o.verbose = int(bool(verbose))
    o.device = int(device)
    o.device_mem_cap = int(device_mem_cap)
    keep = []
    if y0 is not None:
        y0 = _f64(y0).reshape(N, no_dims)
        keep.append(y0)
        o.y0_mode = _lib.Y0_EXPLICIT
        o.y0 = y0.ctypes.data_as(C.POINTER(C.c_double))
    elif seed is not None:
        o.y0_mode = _lib.Y0_SEEDED
        o.seed = int(seed)
    if cancel is not None:
        cb = _lib.CANCEL_FN(lambda: 1 if cancel() else 0)
        keep.append(cb)
        o.cancel = cb
    return o, keep


def _stats_dict(st):
    n = min(st.kl_count, _lib.MAX_KL_TRACE)
    d = {k: getattr(st, k) for k in ("ms_normalize", "ms_knn", "ms_perplexity", "ms_symmetrize", "ms_loop",
                                     "ms_total", "nnz", "knn_k", "bisection_unconverged", "certificate_failures",
                                     "kl_final_exact", "ms_upload")}
    d["kl_iter"] = list(st.kl_iter[:n])
    d["kl"] = list(st.kl_value[:n])
    return d


def tsne(X, no_dims=2, perplexity=30.0, theta=0.5, **options):
    """Full t-SNE on the GPU (tka_tsne_run, replacing tsne::TSNE::run).

    X: N x D (one point per row).  Returns (Y N x no_dims, stats dict).
    Options: dist_dtype ('bf16'|'fp32'), loop_precision ('fp32'|'fp64'),
    candidates, knn_k, max_iter, stop_lying_iter, mom_switch_iter,
    exact_p ('dense'|'knn'), dd_mode ('textbook'|'tapkee_compat'),
    y0 / seed (None replays std::rand() exactly like tapkee; seed=s calls
    srand(s) first; an array y0 is used as given), log_every,
    kl_mode ('reference'|'exact'), verbose, device, device_mem_cap, cancel.
    """
    X = _f64(X)
    if X.ndim != 2:
        raise WrongParameterError("X must be 2-D (N x D)")
    N, D = X.shape
    o, keep = _options(N, no_dims, **options)
    Y = np.empty((N, no_dims))
    st = _lib.Stats()
    _lib.check(_lib.lib().tka_tsne_run(X, N, D, Y, no_dims, perplexity, theta, C.byref(o), C.byref(st)))
    del keep
    return Y, _stats_dict(st)


class Session:
    """A t-SNE problem on the device, advanced step by step (tka_session_*)."""

    def __init__(self, X, no_dims=2, perplexity=30.0, theta=0.5, **options):
        X = _f64(X)
        N, D = X.shape
        self.N, self.dims = N, no_dims
        o, self._keep = _options(N, no_dims, **options)
        h = C.c_void_p()
        _lib.check(_lib.lib().tka_session_create(X, N, D, no_dims, perplexity, theta, C.byref(o), C.byref(h)))
        self._h = h

    @classmethod
    def from_knn(cls, idx, dist, no_dims=2, perplexity=30.0, theta=0.5, **options):
        """A session on a precomputed neighbour graph (tka_session_create_knn):
        idx (N x K, the object itself excluded) and dist (N x K squared
        distances) replace the kNN search."""
        idx = _i32(idx)
        dist = _f64(dist)
        if idx.ndim != 2 or idx.shape != dist.shape:
            raise WrongParameterError("idx and dist must both be N x K")
        N, K = idx.shape
        self = cls.__new__(cls)
        self.N, self.dims = N, no_dims
        o, self._keep = _options(N, no_dims, **options)
        h = C.c_void_p()
        _lib.check(_lib.lib().tka_session_create_knn(idx, dist, N, K, no_dims, perplexity, theta, C.byref(o),
                                                     C.byref(h)))
        self._h = h
        return self

    @classmethod
    def from_p(cls, row, col, val, no_dims=2, theta=0.5, **options):
        """A session on a precomputed joint P in CSR (tka_session_create_p),
        used after division by its sum."""
        row, col, val = _i32(row), _i32(col), _f64(val)
        N = len(row) - 1
        if len(col) != len(val) or len(col) < row[-1]:
            raise WrongParameterError("col and val must hold row_ptr[N] entries")
        self = cls.__new__(cls)
        self.N, self.dims = N, no_dims
        o, self._keep = _options(N, no_dims, **options)
        h = C.c_void_p()
        _lib.check(_lib.lib().tka_session_create_p(row, col, val, N, no_dims, theta, C.byref(o), C.byref(h)))
        self._h = h
        return self

    def step(self, iterations=1):
        _lib.check(_lib.lib().tka_session_step(self._h, int(iterations)))

    def synchronize(self):
        _lib.check(_lib.lib().tka_session_synchronize(self._h))

    def embedding(self):
        Y = np.empty((self.N, self.dims))
        _lib.check(_lib.lib().tka_session_embedding(self._h, Y))
        return Y

    def inject_sort_fault(self):
        """Test hook: mark the Barnes-Hut tree's sort as failed (see
        tka_session_inject_sort_fault); embedding() and kl() then raise."""
        _lib.check(_lib.lib().tka_session_inject_sort_fault(self._h))

    def kl(self):
        v = C.c_double(0)
        _lib.check(_lib.lib().tka_session_kl(self._h, C.byref(v)))
        return v.value

    def stats(self):
        st = _lib.Stats()
        _lib.check(_lib.lib().tka_session_stats(self._h, C.byref(st)))
        return _stats_dict(st)

    def profile(self, enable=True):
        _lib.check(_lib.lib().tka_session_profile(self._h, int(bool(enable))))

    def kernel_timing(self):
        t = _lib.KernelTiming()
        _lib.check(_lib.lib().tka_session_kernel_timing(self._h, C.byref(t)))
        return {k: getattr(t, k) for k, _ in _lib.KernelTiming._fields_}

    def close(self):
        if getattr(self, "_h", None):
            _lib.lib().tka_session_destroy(self._h)
            self._h = None

    def __del__(self):
        # at interpreter exit the module's globals may be cleared before the
        # session goes; the process is ending and frees the device anyway
        if _lib is not None:
            self.close()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()


_logger_keep = None


def set_logger(fn=None, levels=("warning", "error")):
    """Route the engine's messages (tka_set_logger, mirroring tapkee's
    Logging singleton): fn(level, message) receives each message of an
    enabled level, level one of "info", "warning", "debug", "error",
    "benchmark"; fn=None restores the default logger, which writes
    "[level] message" to stderr.  The defaults enable warning and error, as
    tapkee's (include/tapkee/utils/logging.hpp:96-98)."""
    global _logger_keep
    mask = 0
    for name in levels:
        mask |= 1 << _lib.LOG_LEVEL_NAMES.index(name)
    if fn is None:
        cb = C.cast(None, _lib.LOG_FN)
    else:
        cb = _lib.LOG_FN(lambda lvl, msg, user: fn(_lib.LOG_LEVEL_NAMES[lvl], msg.decode(errors="replace")))
    _lib.lib().tka_set_logger(cb, None, mask)
    _logger_keep = cb


def normalize(X):
    X = _f64(X).copy()
    _lib.check(_lib.lib().tka_normalize(X, X.shape[0], X.shape[1]))
    return X


def knn(X, K, dist_dtype="fp32", candidates=0):
    """Exact K nearest neighbours (self excluded): (idx N x K int32, dist N x K, certificate failures)."""
    X = _f64(X)
    N, D = X.shape
    idx = np.empty((N, K), np.int32)
    dist = np.empty((N, K))
    fails = C.c_int32(0)
    _lib.check(_lib.lib().tka_knn(X, N, D, K, _lib.DIST_NAMES[dist_dtype], candidates, idx, dist,
                                  C.byref(fails)))
    return idx, dist, fails.value


def neighbors_connected(idx):
    """tapkee's is_connected (include/tapkee/neighbors/connected.hpp:19-61):
    DFS from object 0 along the directed neighbour lists idx (N x k)."""
    idx = _i32(idx)
    N, k = idx.shape
    out = C.c_int32(0)
    _lib.check(_lib.lib().tka_neighbors_connected(idx, N, k, C.byref(out)))
    return bool(out.value)


def _log(level, msg):
    _lib.lib().tka_log(level, msg.encode())


def find_neighbors(X, k, check_connectivity=True, method="Brute-force", dist_dtype="bf16"):
    """tapkee's find_neighbors (include/tapkee/neighbors/neighbors.hpp:172-216)
    for the rows of X (N x D) under the Euclidean distance, on the GPU: the
    exact k nearest other rows of each row, nearest first (ties by index),
    whatever the method named.  k above N-1 is lowered with the reference's
    warning; with check_connectivity a graph that is not connected is
    recomputed with 2k neighbours.  Returns (idx N x k_used, k_used)."""
    X = _f64(X)
    N = X.shape[0]
    if N == 0:
        raise NoDataError("no data to search")
    if k > N - 1:
        _log(_lib.LOG_WARNING, "Number of neighbors is greater than number of objects to embed. "
                               "Using greatest possible number of neighbors.")
        k = N - 1
    _log(_lib.LOG_INFO, f"Using the {method} neighbors computation method.")
    idx, _, _ = knn(X, k, dist_dtype=dist_dtype)
    if check_connectivity and not neighbors_connected(idx):
        _log(_lib.LOG_WARNING, f"The neighborhood graph with {k} neighbors is not connected. "
                               f"Recomputing with a larger number of neighbors {2 * k}.")
        return find_neighbors(X, 2 * k, check_connectivity, method, dist_dtype)
    _log(_lib.LOG_INFO, f"The neighborhood graph with {k} neighbors is connected.")
    return idx, k


def perplexity(dist, perp):
    dist = _f64(dist)
    N, K = dist.shape
    val = np.empty((N, K))
    beta = np.empty(N)
    iters = np.empty(N, np.int32)
    unconv = C.c_int32(0)
    _lib.check(_lib.lib().tka_perplexity(dist, N, K, perp, val, beta, iters, C.byref(unconv)))
    return val, beta, iters, unconv.value


def perplexity_dense(X, perp, dd_mode="textbook"):
    X = _f64(X)
    N, D = X.shape
    P = np.empty((N, N))
    _lib.check(_lib.lib().tka_perplexity_dense(X, N, D, perp, 1 if dd_mode == "tapkee_compat" else 0, P))
    return P


def symmetrize(idx, val):
    idx = _i32(idx)
    val = _f64(val)
    N, K = idx.shape
    row = np.empty(N + 1, np.int32)
    col = np.empty(2 * N * K, np.int32)
    sval = np.empty(2 * N * K)
    nnz = C.c_int64(0)
    _lib.check(_lib.lib().tka_symmetrize(idx, val, N, K, row, col, sval, C.byref(nnz)))
    return row, col[:nnz.value].copy(), sval[:nnz.value].copy()


def build_p(X, perp, knn_k=0, dist_dtype="bf16"):
    """Normalised symmetric P (CSR: row_ptr, col, val) of data X (N x D), no exaggeration."""
    X = _f64(X)
    N, D = X.shape
    K = knn_k if knn_k > 0 else int(3 * perp)
    row = np.empty(N + 1, np.int32)
    col = np.empty(2 * N * max(K, 1), np.int32)
    val = np.empty(2 * N * max(K, 1))
    nnz = C.c_int64(0)
    _lib.check(_lib.lib().tka_build_p(X, N, D, perp, knn_k, _lib.DIST_NAMES[dist_dtype], row, col, val,
                                      C.byref(nnz)))
    return row, col[:nnz.value].copy(), val[:nnz.value].copy()


def gradient(Y, row, col, val, theta, precision="fp32"):
    """dY without the factor 4 (tsne.hpp:325-351 / :353-407) and Z."""
    Y = _f64(Y)
    N, d = Y.shape
    dY = np.empty_like(Y)
    z = C.c_double(0)
    _lib.check(_lib.lib().tka_gradient(Y, N, d, _i32(row), _i32(col), _f64(val), theta,
                                       _lib.FP64 if precision == "fp64" else _lib.FP32, dY, C.byref(z)))
    return dY, z.value


def gradient_dense(P, Y, dd_mode="textbook", precision="fp32"):
    Y = _f64(Y)
    N, d = Y.shape
    dY = np.empty_like(Y)
    z = C.c_double(0)
    _lib.check(_lib.lib().tka_gradient_dense(_f64(P), Y, N, d, 1 if dd_mode == "tapkee_compat" else 0,
                                             _lib.FP64 if precision == "fp64" else _lib.FP32, dY, C.byref(z)))
    return dY, z.value


def kl(Y, row, col, val, theta=0.0, mode="exact"):
    Y = _f64(Y)
    out = C.c_double(0)
    _lib.check(_lib.lib().tka_kl(Y, Y.shape[0], Y.shape[1], _i32(row), _i32(col), _f64(val), theta,
                                 _lib.KL_EXACT_Z if mode == "exact" else _lib.KL_AS_REFERENCE, C.byref(out)))
    return out.value


def kl_dense(P, Y, dd_mode="textbook"):
    Y = _f64(Y)
    out = C.c_double(0)
    _lib.check(_lib.lib().tka_kl_dense(_f64(P), Y, Y.shape[0], Y.shape[1],
                                       1 if dd_mode == "tapkee_compat" else 0, C.byref(out)))
    return out.value


def update(dY, uY, gains, Y, momentum, eta=200.0, precision="fp64"):
    uY = _f64(uY).copy()
    gains = _f64(gains).copy()
    Y = _f64(Y).copy()
    _lib.check(_lib.lib().tka_update(_f64(dY), uY, gains, Y, Y.shape[0], Y.shape[1], momentum, eta,
                                     _lib.FP64 if precision == "fp64" else _lib.FP32))
    return uY, gains, Y


def bh_tree(Y, precision="fp64"):
    """Compressed pre-order Barnes-Hut tree of Y (N x 2) as a dict of arrays."""
    Y = _f64(Y)
    N = Y.shape[0]
    cap = 2 * N
    n = C.c_int32(0)
    out = {k: np.empty(cap, np.int32) for k in ("level", "count", "leaf_point", "skip")}
    cx = np.empty(cap)
    cy = np.empty(cap)
    box = np.empty(4)
    _lib.check(_lib.lib().tka_bh_tree(Y, N, _lib.FP64 if precision == "fp64" else _lib.FP32, C.byref(n),
                                      out["level"], out["count"], out["leaf_point"], out["skip"], cx, cy, box))
    m = n.value
    res = {k: v[:m].copy() for k, v in out.items()}
    res["comx"] = cx[:m].copy()
    res["comy"] = cy[:m].copy()
    res["root_box"] = box
    return res
