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
