"""Boundary checks that need no GPU: the C-ABI library loads and exports every
symbol the header declares, the Python front end mirrors tapkee's keyword
surface and error behaviour, and the C++ drop-in header compiles and raises
tapkee's exception types (checked by tests/cpp/test_dropin.cpp)."""
import os
import re
import subprocess

import numpy as np
import pytest

import tapkee_amd
from tapkee_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    text = open(os.path.join(ROOT, "include", "tapkee_amd.h")).read()
    return set(re.findall(r"TKA_API\s+[\w\s\*]+?\b(tka_\w+)\s*\(", text))


def test_header_declares_the_bound_symbols():
    assert declared_symbols() == set(_lib.SYMBOLS)


def test_library_exports_every_declared_symbol():
    L = _lib.lib()
    for name in declared_symbols():
        assert hasattr(L, name), name
    out = subprocess.run(["nm", "-D", "--defined-only", _lib.LIB_PATH], capture_output=True, text=True).stdout
    exported = set(re.findall(r" T (tka_\w+)", out))
    assert declared_symbols() <= exported


def test_default_options_are_the_reference_schedule():
    o = _lib.default_options()
    assert o.struct_size == __import__("ctypes").sizeof(_lib.Options)
    # tsne.hpp:71-73, :129, :184
    assert (o.max_iter, o.stop_lying_iter, o.mom_switch_iter) == (1000, 250, 250)
    assert (o.momentum, o.final_momentum, o.eta, o.exaggeration) == (0.5, 0.8, 200.0, 12.0)
    assert o.log_every == 50


def X100():
    g = np.random.default_rng(0)
    return g.normal(size=(3, 100))  # features x samples


# Each case fails in validation, before any device work (tka_tsne_run checks
# its arguments first), so these run on a machine without a GPU.
@pytest.mark.parametrize("kwargs,exc", [
    (dict(sne_perplexity=40.0), tapkee_amd.WrongParameterError),         # methods/tsne.hpp:20
    (dict(sne_perplexity=10.0, sne_theta=-1.0), tapkee_amd.WrongParameterError),  # :21
    (dict(sne_perplexity=10.0, target_dimension=0), tapkee_amd.WrongParameterError),  # base.hpp:43
    (dict(sne_perplexity=10.0, target_dimension=100), tapkee_amd.WrongParameterError),
    (dict(sne_perplexity=10.0, target_dimension=4), tapkee_amd.WrongParameterError),  # BH trees: 2-D or 3-D
    (dict(sne_perplexity=0.1, sne_theta=0.5), tapkee_amd.WrongParameterError),  # K = 0
])
def test_embed_validation(kwargs, exc):
    with pytest.raises(exc):
        tapkee_amd.embed(X100(), method="t-sne", **kwargs)
    # tapkee's errors surface as RuntimeError in its Python binding
    assert issubclass(exc, RuntimeError)


def test_unknown_and_unsupported_methods():
    with pytest.raises(ValueError):
        tapkee_amd.parse_reduction_method("unknown")  # test/test_nanobind_extension.py:4-9
    with pytest.raises(tapkee_amd.UnsupportedMethodError):
        tapkee_amd.embed(X100(), method="isomap", num_neighbors=10)
    assert tapkee_amd.parse_reduction_method("t-sne") is tapkee_amd.parse_reduction_method(
        "t-stochastic_proximity_embedding")


def test_parameters_set_errors():
    ps = tapkee_amd.ParametersSet()
    ps.add(tapkee_amd.Parameter.create("SNE perplexity", 10.0))
    with pytest.raises(tapkee_amd.MissedParameterError):
        tapkee_amd.withParameters(ps).embedUsing(X100())
    ps.add(tapkee_amd.Parameter.create("dimension reduction method", tapkee_amd.parse_reduction_method("t-sne")))
    ps.add(tapkee_amd.Parameter.create("SNE perplexity", 5.0))
    with pytest.raises(tapkee_amd.MultipleParameterError):
        tapkee_amd.withParameters(ps).embedUsing(X100())


def test_empty_input_and_cancel():
    with pytest.raises(tapkee_amd.NoDataError):
        tapkee_amd.embed(np.zeros((3, 0)), method="t-sne")
    ps = tapkee_amd.ParametersSet()
    ps.add(tapkee_amd.Parameter.create("dimension reduction method", tapkee_amd.parse_reduction_method("t-sne")))
    ps.add(tapkee_amd.Parameter.create("SNE perplexity", 10.0))
    ps.add(tapkee_amd.Parameter.create("cancel function", lambda: True))
    with pytest.raises(tapkee_amd.CancelledError):
        tapkee_amd.withParameters(ps).embedUsing(X100())


def test_non_finite_input_rejected():
    X = X100()
    X[1, 5] = np.inf
    with pytest.raises(tapkee_amd.WrongParameterError):
        tapkee_amd.embed(X, method="t-sne", sne_perplexity=10.0)


def test_cpp_dropin_header_cpu_cases(tmp_path):
    exe = tmp_path / "test_dropin"
    src = os.path.join(ROOT, "tests", "cpp", "test_dropin.cpp")
    libdir = os.path.dirname(_lib.LIB_PATH)
    subprocess.run(["g++", "-std=c++17", "-O1", "-Wall", "-Werror", "-I", os.path.join(ROOT, "include"), src,
                    "-L", libdir, "-ltapkee_amd", f"-Wl,-rpath,{libdir}", "-o", str(exe)], check=True)
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "all drop-in checks passed" in r.stdout


def test_noalloc_limit_matches_header():
    text = open(os.path.join(ROOT, "include", "tapkee_amd.h")).read()
    m = re.search(r"#define TKA_NOALLOC_MAX_POINTS (\d+)", text)
    assert m and int(m.group(1)) == _lib.NOALLOC_MAX_POINTS


def test_logger_receives_the_reference_warning():
    # tsne.hpp:604-605 warns when perplexity > K; the message is emitted in
    # validation, before any device work (on a machine without a GPU the
    # call then fails)
    got = []
    tapkee_amd.set_logger(lambda lvl, msg: got.append((lvl, msg)), levels=("warning", "error"))
    try:
        with pytest.raises(_lib.EngineError):
            tapkee_amd.tsne(np.random.default_rng(0).normal(size=(200, 4)), perplexity=30.0, knn_k=20,
                            device=1 << 20)
    finally:
        tapkee_amd.set_logger(None)
    assert ("warning", "Perplexity should be lower than K!") in got
    # a disabled level is not delivered
    got.clear()
    tapkee_amd.set_logger(lambda lvl, msg: got.append((lvl, msg)), levels=("error",))
    try:
        with pytest.raises(_lib.EngineError):
            tapkee_amd.tsne(np.random.default_rng(0).normal(size=(200, 4)), perplexity=30.0, knn_k=20,
                            device=1 << 20)
    finally:
        tapkee_amd.set_logger(None)
    assert got == []


def test_cpp_neighbors_header_cpu_cases(tmp_path):
    exe = tmp_path / "test_neighbors"
    src = os.path.join(ROOT, "tests", "cpp", "test_neighbors.cpp")
    libdir = os.path.dirname(_lib.LIB_PATH)
    subprocess.run(["g++", "-std=c++17", "-O1", "-Wall", "-Werror", "-I", os.path.join(ROOT, "include"), src,
                    "-L", libdir, "-ltapkee_amd", f"-Wl,-rpath,{libdir}", "-o", str(exe)], check=True)
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "all neighbour checks passed" in r.stdout


def test_neighbors_connected_fixed_graphs():
    # connected.hpp:19-61: DFS from object 0 along the directed lists
    assert tapkee_amd.neighbors_connected(np.array([[1], [2], [0]]))
    assert not tapkee_amd.neighbors_connected(np.array([[1], [0], [0]]))
    assert not tapkee_amd.neighbors_connected(np.array([[1, 1], [0, 0], [3, 3], [2, 2]]))
    with pytest.raises(tapkee_amd.WrongParameterError):
        tapkee_amd.neighbors_connected(np.array([[1], [5], [0]]))


def test_precomputed_inputs_are_checked_before_device_work():
    row = np.array([0, 1, 2, 3], np.int32)
    with pytest.raises(tapkee_amd.WrongParameterError, match="column index"):
        tapkee_amd.Session.from_p(row, np.array([1, 2, 7]), np.array([0.2, 0.3, 0.5]))
    with pytest.raises(tapkee_amd.WrongParameterError, match="non-negative"):
        tapkee_amd.Session.from_p(row, np.array([1, 2, 0]), np.array([0.2, -0.3, 0.5]))
    idx = np.array([[1], [2], [2], [0]])
    with pytest.raises(tapkee_amd.WrongParameterError, match="object itself"):
        tapkee_amd.Session.from_knn(idx, np.ones((4, 1)), perplexity=0.3, theta=0.0)
    with pytest.raises(tapkee_amd.WrongParameterError, match="finite"):
        tapkee_amd.Session.from_knn(np.array([[1], [2], [3], [0]]), np.array([[1.0], [np.nan], [1.0], [1.0]]),
                                    perplexity=0.3, theta=0.0)
    with pytest.raises(tapkee_amd.WrongParameterError, match="listed twice"):
        tapkee_amd.Session.from_knn(np.array([[1, 2], [2, 0], [3, 3], [0, 1]]), np.ones((4, 2)),
                                    perplexity=0.6, theta=0.0)
