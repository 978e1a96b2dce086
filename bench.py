#!/usr/bin/env python3
"""t-SNE hot-path benchmark on one MI355X.

Workload: Barnes-Hut t-SNE at N=100,000 points, D=50 (a 3-D Swiss roll
mapped to 50-D by orthonormal rows plus N(0, 0.05^2) noise), perplexity 30
(K=90 neighbours), theta=0.5, 2-D output; seeded synthetic data.

A "step" is one iteration of the main loop (tree build, Barnes-Hut
repulsion, CSR attraction, gains/momentum update), on P and Y resident in
HBM; `value` is iterations per second.  The same JSON line also reports the
kNN + P build time with BF16 and with FP32 distance tiles, the full
1000-iteration run and its final KL against an FP64 run from the same P and
Y0, the roofline of the dominant kernel (HIP-event timings on the engine's
stream) and the CPU restatement of the reference timed on this host.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--n N]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)

METRIC = "t-SNE iters/sec + kNN/P build ms at N=100k (BF16 vs FP32), ΔKL vs FP64"
PEAK_FP32_TFLOPS = 157.3
PEAK_HBM_GBS = 8000.0
FLOPS_PER_INTERACTION = 14   # dx, dy, D, q = 1/(1+D), z += c q, m = c q^2, f += m d (2)


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def cpu_baseline(row, col, val, Y0, theta, budget_s=20.0, max_iters=40):
    """The FP64 single-thread restatement of the reference loop (oracle),
    timed on this host: tree build + forces + update per iteration."""
    import numpy as np

    from oracle import oracle as O
    O.lib()
    Y = Y0.copy()
    uY = np.zeros_like(Y)
    gains = np.ones_like(Y)
    v = val * 1.0   # after the exaggeration phase
    n = 0
    t0 = time.perf_counter()
    while n < max_iters and (time.perf_counter() - t0) < budget_s:
        dY, _ = O.gradient_bh(row, col, v, Y, theta)
        uY, gains, Y = O.update(dY, uY, gains, Y, 0.8)
        n += 1
    dt = time.perf_counter() - t0
    return n / dt, n, dt


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=300)
    ap.add_argument("--warmup", type=int, default=30)
    ap.add_argument("--n", type=int, default=100000)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--no-pipeline", action="store_true", help="skip the full 1000-iteration runs")
    a = ap.parse_args()
    if a.gpus != 1:
        print(f"bench.py runs on one GPU only (got --gpus {a.gpus}); this engine has no multi-GPU path",
              file=sys.stderr)
        sys.exit(2)

    import numpy as np

    import tapkee_amd
    from tapkee_amd import datasets

    N = a.n
    X, _ = datasets.swiss_roll_50d(n=N)
    perp, theta = 30.0, 0.5
    log(f"[bench] workload: swiss roll N={N} D={X.shape[1]} perplexity={perp} theta={theta}")

    # ---- kNN + P build, BF16 vs FP32 distance tiles (setup of a session)
    build = {}
    for dt in ("fp32", "bf16"):
        s = tapkee_amd.Session(X, perplexity=perp, theta=theta, dist_dtype=dt, seed=1)
        st = s.stats()
        build[dt] = {"knn_ms": st["ms_knn"], "p_ms": st["ms_perplexity"] + st["ms_symmetrize"],
                     "normalize_ms": st["ms_normalize"],
                     "total_ms": st["ms_normalize"] + st["ms_knn"] + st["ms_perplexity"] + st["ms_symmetrize"],
                     "certificate_failures": st["certificate_failures"], "nnz": st["nnz"]}
        log(f"[bench] build {dt}: {build[dt]}")
        if dt == "fp32":
            s.close()
    sess = s  # bf16 session

    # ---- timed steps (FP32 loop)
    W, K = a.warmup, a.steps
    sess.step(W)
    sess.synchronize()
    t0 = time.perf_counter()
    sess.step(K)
    sess.synchronize()
    t1 = time.perf_counter()
    ms_per_step = (t1 - t0) * 1e3 / K
    value = K / (t1 - t0)
    log(f"[bench] {K} steps: {ms_per_step:.4f} ms/step, {value:.1f} iters/s")

    # ---- per-kernel timing (HIP events on the engine stream)
    sess.profile(True)
    sess.step(min(K, 100))
    kt = sess.kernel_timing()
    sess.profile(False)
    log(f"[bench] kernel timing: {kt}")
    rep_ms = kt["ms_repulsion"] / max(kt["launches_repulsion"], 1)
    build_ms = kt["ms_build"] / max(kt["launches_build"], 1)
    upd_ms = kt["ms_step"] / max(kt["launches_step"], 1)
    att_ms = kt["ms_attraction"] / max(kt["launches_attraction"], 1)
    inter = kt["interactions"] / max(kt["launches_repulsion"], 1)
    achieved_tflops = inter * FLOPS_PER_INTERACTION / (rep_ms * 1e-3) / 1e12
    nnz = kt["nnz"]
    # attraction: per edge col (4) + P (4) + y_col (8); per row y_i (8), row_ptr (4), pos (8)
    att_bytes = nnz * (4 + 4 + 8) + N * (8 + 4 + 8)
    att_gbs = att_bytes / (att_ms * 1e-3) / 1e9
    roofline = {"kernel": "bh_traverse", "bound": "mfma", "achieved": round(achieved_tflops, 3),
                "peak": PEAK_FP32_TFLOPS, "unit": "TFLOP/s", "frac": round(achieved_tflops / PEAK_FP32_TFLOPS, 5),
                "traffic": None, "interactions_per_launch": int(inter), "ms_per_launch": round(rep_ms, 5),
                "note": "FP32 VALU work (14 flop per cell interaction) against the published FP32 peak"}
    kernels = {"tree_build_ms": round(build_ms, 5), "repulsion_ms": round(rep_ms, 5),
               "update_ms": round(upd_ms, 5),
               "attraction_ms": round(att_ms, 5), "attraction_concurrent_with": "tree build + traversal",
               "attraction_GBps": round(att_gbs, 1),
               "attraction_hbm_frac": round(att_gbs / PEAK_HBM_GBS, 4)}

    # ---- full runs: 1000 iterations FP32 loop vs FP64 loop from the same P and Y0
    pipeline = None
    if not a.no_pipeline:
        y0 = np.random.Generator(np.random.PCG64(5)).normal(size=(N, 2)) * 1e-4
        runs = {}
        for prec in ("fp32", "fp64"):
            t = time.perf_counter()
            Y, st = tapkee_amd.tsne(X, perplexity=perp, theta=theta, dist_dtype="bf16", loop_precision=prec, y0=y0)
            wall = time.perf_counter() - t
            runs[prec] = {"wall_s": round(wall, 3), "loop_ms": round(st["ms_loop"], 1),
                          "kl_final_exact": st["kl_final_exact"], "kl_trace_last": st["kl"][-1]}
            log(f"[bench] full run {prec}: {runs[prec]}")
        d_kl = (runs["fp32"]["kl_final_exact"] - runs["fp64"]["kl_final_exact"]) / runs["fp64"]["kl_final_exact"]
        pipeline = {"fp32_loop": runs["fp32"], "fp64_loop": runs["fp64"], "delta_kl_rel_vs_fp64": d_kl,
                    "iters": 1000}

    # ---- CPU baseline: the reference loop restated in FP64, one core
    cpu = None
    if not a.no_cpu:
        row, col, val = tapkee_amd.build_p(X, perp, dist_dtype="bf16")
        Ycpu = sess.embedding()
        rate, n_it, dt = cpu_baseline(row, col, val, Ycpu, theta)
        cpu = {"value": round(rate, 4), "unit": "iters/s", "cores": 1, "kind": "port",
               "sample": f"{n_it} Barnes-Hut iterations (quadtree build + forces + update, FP64, one thread) "
                         f"at N={N} on the engine's P, from the embedding after {W + K + min(K, 100)} "
                         f"iterations; {dt:.1f} s"}
        log(f"[bench] cpu baseline: {cpu}")
    sess.close()

    out = {
        "metric": METRIC, "value": round(value, 3), "unit": "iters/s", "n_gpus": 1, "steps": K, "warmup": W,
        "ms_per_step": round(ms_per_step, 5), "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "fp32", "data": "synthetic",
        "config": {"workload": "bh-tsne swiss-roll N=100000 D=50 perplexity=30 K=90 theta=0.5 2-D",
                   "n": N, "d": int(X.shape[1]), "perplexity": perp, "theta": theta, "knn_dtype": "bf16",
                   "loop_dtype": "fp32"},
        "knn_p_build_ms": {k: round(v["total_ms"], 3) for k, v in build.items()},
        "knn_p_build": build,
        "kernels": kernels,
        "roofline": roofline,
        "cpu_baseline": cpu,
        "pipeline": pipeline,
    }
    if cpu:
        out["speedup_vs_cpu_iters"] = round(value / cpu["value"], 1)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
