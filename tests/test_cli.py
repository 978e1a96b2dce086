"""The command-line front end (tapkee_amd/bin/tapkee, built from
tapkee_amd/cli/tapkee_cli.cpp) against tapkee's CLI (src/cli/main.cpp,
src/cli/util.hpp): option checks and messages on the CPU, a seeded
embedding from a CSV on the GPU."""
import os
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CLI = os.path.join(ROOT, "tapkee_amd", "bin", "tapkee")


def run(*args, stdin=None):
    return subprocess.run([CLI, *args], capture_output=True, text=True, input=stdin, timeout=120)


def test_cli_help_and_option_errors(tmp_path):
    r = run("--help")
    assert r.returncode == 1 and "--sne-perplexity" in r.stdout and "--seed" in r.stdout
    r = run("-m", "lle")
    assert r.returncode == 1 and "[error] Method lle is not provided by this engine" in r.stderr
    r = run("-m", "t-snee", "--verbose")   # util.hpp:171-189 suggests the closest name
    assert r.returncode == 1 and "Did you mean `t-sne`?" in r.stderr and "[error] Unknown method t-snee" in r.stderr
    r = run("-td", "0")
    assert r.returncode == 1 and "positive target dimensionality" in r.stderr
    r = run("-k", "2")
    assert r.returncode == 1 and "number of neighbors is too small" in r.stderr
    r = run("--nm", "kdtree")
    assert r.returncode == 1 and "Unknown neighbors method kdtree" in r.stderr
    bad = tmp_path / "bad.csv"
    bad.write_text("1,2,3\n4,5\n")
    r = run("-i", str(bad))
    assert r.returncode == 1 and "Some error occured: Wrong data at line 1" in r.stderr
    r = run("--no-such-option")
    assert r.returncode == 1 and "unknown option" in r.stderr


@pytest.mark.gpu
def test_cli_seeded_embedding_from_csv(engine, tmp_path):
    g = np.random.default_rng(3)
    X = np.concatenate([g.normal(c, 1.0, size=(150, 6)) for c in (0.0, 8.0, 16.0)])
    src = tmp_path / "x.csv"
    np.savetxt(src, X, delimiter=",", fmt="%.17g")
    outs = []
    for k in range(2):
        o = tmp_path / f"y{k}.csv"
        r = run("-i", str(src), "-o", str(o), "-m", "t-sne", "--sne-perplexity", "15", "--seed", "7", "--benchmark")
        assert r.returncode == 0, r.stderr
        assert "[benchmark] Input similarities computation took" in r.stderr
        assert "[benchmark] Main t-SNE loop took" in r.stderr
        outs.append(np.loadtxt(o, delimiter=","))
    np.testing.assert_array_equal(outs[0], outs[1])   # --seed makes runs repeatable
    Y, _ = engine.tsne(X, perplexity=15.0, seed=7)
    # the CLI writes with the stream's default 6 significant digits
    np.testing.assert_allclose(outs[0], Y, rtol=1e-5, atol=1e-5 * np.abs(Y).max())
    # --transpose-input / --transpose-output with another delimiter
    srct = tmp_path / "xt.tsv"
    np.savetxt(srct, X.T, delimiter=";", fmt="%.17g")
    o = tmp_path / "yt.tsv"
    r = run("-i", str(srct), "-o", str(o), "-d", ";", "--transpose-input", "--transpose-output",
            "--sne-perplexity", "15", "--seed", "7")
    assert r.returncode == 0, r.stderr
    np.testing.assert_array_equal(np.loadtxt(o, delimiter=";").T, outs[0])
