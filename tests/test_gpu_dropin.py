"""The C++ drop-in header (include/tapkee_amd/tapkee.hpp) and the Python
front end on a real device: tapkee's t-SNE smoke test shape
(test/unit/methods.cpp:16-32,119-122) plus the out-of-memory mapping."""
import os
import subprocess

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_cpp_dropin_on_gpu(engine, tmp_path):
    from tapkee_amd import _lib
    exe = tmp_path / "test_dropin"
    libdir = os.path.dirname(_lib.LIB_PATH)
    subprocess.run(["g++", "-std=c++17", "-O1", "-I", os.path.join(ROOT, "include"),
                    os.path.join(ROOT, "tests", "cpp", "test_dropin.cpp"), "-L", libdir, "-ltapkee_amd",
                    f"-Wl,-rpath,{libdir}", "-o", str(exe)], check=True)
    r = subprocess.run([str(exe), "--gpu"], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "all drop-in checks passed (with GPU)" in r.stdout


def test_python_embed_smoke_shape(engine):
    g = np.random.default_rng(0)
    t = (3 * np.pi / 4) * (g.uniform(-1, 1, 100) + 0.5)
    h = g.uniform(-1, 1, 100) - 0.5
    X = np.stack([t * np.cos(t), 10 * h, t * np.sin(t)])  # features x samples
    Y = engine.embed(X, method="t-sne", target_dimension=2, num_neighbors=20, sne_perplexity=10.0)
    assert Y.shape == (100, 2) and np.all(np.isfinite(Y))
    with pytest.raises(engine.NotEnoughMemoryError):
        engine.embed(X, method="t-sne", sne_perplexity=10.0, device_mem_cap=4096)


def test_session_steps_match_one_shot_run(engine):
    g = np.random.default_rng(1)
    X = g.normal(size=(600, 12))
    y0 = g.normal(size=(600, 2)) * 1e-4
    Y1, st1 = engine.tsne(X, perplexity=12.0, theta=0.5, max_iter=400, y0=y0, loop_precision="fp64")
    with engine.Session(X, perplexity=12.0, theta=0.5, max_iter=400, y0=y0, loop_precision="fp64") as s:
        s.step(123)
        s.step(277)
        Y2 = s.embedding()
        assert s.stats()["kl"] == st1["kl"]
    np.testing.assert_array_equal(Y1, Y2)   # deterministic: same kernels, same order


def test_phase_timers_and_iteration_lines_reach_the_logger(engine):
    # tsne.hpp:87 and :142 (timed_context, utils/time.hpp:29-42) at the
    # benchmark level, tsne.hpp:68 and :191 at the info level
    import re
    got = []
    engine.set_logger(lambda lvl, msg: got.append((lvl, msg)), levels=("info", "warning", "error", "benchmark"))
    try:
        X = np.random.default_rng(0).normal(size=(600, 8))
        engine.tsne(X, perplexity=10.0, theta=0.5, seed=1, max_iter=120)
    finally:
        engine.set_logger(None)
    msgs = [m for _, m in got]
    assert ("info", "Using Barnes-Hut-SNE algorithm") in got
    assert any(lvl == "benchmark" and re.fullmatch(r"Input similarities computation took [0-9.e+-]+ seconds\.", m)
               for lvl, m in got), msgs
    assert any(lvl == "benchmark" and re.fullmatch(r"Main t-SNE loop took [0-9.e+-]+ seconds\.", m)
               for lvl, m in got), msgs
    its = [int(re.match(r"Iteration (\d+): error is ", m).group(1)) for lvl, m in got
           if lvl == "info" and m.startswith("Iteration ")]
    assert its == [50, 100, 119]   # tsne.hpp:186: every 50 iterations and the last


def test_cpp_neighbors_backend_on_gpu(engine, tmp_path):
    from tapkee_amd import _lib
    exe = tmp_path / "test_neighbors"
    libdir = os.path.dirname(_lib.LIB_PATH)
    subprocess.run(["g++", "-std=c++17", "-O1", "-I", os.path.join(ROOT, "include"),
                    os.path.join(ROOT, "tests", "cpp", "test_neighbors.cpp"), "-L", libdir, "-ltapkee_amd",
                    f"-Wl,-rpath,{libdir}", "-o", str(exe)], check=True)
    r = subprocess.run([str(exe), "--gpu"], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "all neighbour checks passed (with GPU)" in r.stdout


def test_python_find_neighbors_matches_brute_force(engine, oracle):
    # the reference's neighbour test data (test/unit/neighbors.cpp:17-50) and
    # a 2-D mixture: exact neighbours (ties by index) against the oracle
    line = np.arange(100, dtype=float)[:, None]
    idx, k = engine.find_neighbors(line, 10)
    assert k == 10
    for i in range(100):
        s = set(idx[i])
        assert len(s) == 10 and i not in s
        assert all(i - j - 1 in s for j in range(min(i, 5))) and all(i + j + 1 in s for j in range(min(99 - i, 5)))
    X = np.random.default_rng(2).normal(size=(3000, 7))
    idx, k = engine.find_neighbors(X, 12, check_connectivity=False)
    ridx, _ = oracle.knn_brute(X, 12)
    np.testing.assert_array_equal(idx, ridx)
    # the reference's test follows the directed lists from object 0, so an
    # object that is nobody's neighbour leaves the graph "not connected"
    # (random 7-D points have such objects at k = 12): k doubles until none
    # is left
    idx, k = engine.find_neighbors(X, 12)
    assert k > 12 and engine.neighbors_connected(idx)
    assert not engine.neighbors_connected(idx[:, :k // 2])
    # two far groups: recomputed with 2k until connected (neighbors.hpp:196-205)
    got = []
    engine.set_logger(lambda lvl, msg: got.append((lvl, msg)), levels=("info", "warning"))
    try:
        two = np.concatenate([np.arange(50.0), 1000.0 + np.arange(50.0)])[:, None]
        idx, k = engine.find_neighbors(two, 10)
    finally:
        engine.set_logger(None)
    assert k == 80 and engine.neighbors_connected(idx)
    assert sum("is not connected" in m for _, m in got) == 3
