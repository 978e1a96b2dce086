"""Parametric t-SNE: a small MLP X -> Y trained on KL(P||Q) through the
engine's autograd function (workload 5).

The P of the data is built once on the GPU (kNN, perplexity, symmetrise);
each step evaluates KL and its gradient with the engine's force kernels and
back-propagates through the MLP.  The result is compared with the KL of the
non-parametric embedding of the same data.
"""
from __future__ import annotations

import time

import numpy as np
import torch

from .autograd import TSNELoss


def mlp(d_in: int, hidden=(256, 256), d_out: int = 2) -> torch.nn.Sequential:
    layers, d = [], d_in
    for h in hidden:
        layers += [torch.nn.Linear(d, h), torch.nn.ReLU()]
        d = h
    layers.append(torch.nn.Linear(d, d_out))
    return torch.nn.Sequential(*layers)


def train(X, perplexity=30.0, theta=0.0, hidden=(256, 256), steps=500, lr=1e-3, seed=0, device="cuda",
          dist_dtype="bf16", log_every=50):
    """Full-batch Adam on KL(P||Q) of mlp(X).  Returns (model, history dict)."""
    X = np.asarray(X, dtype=np.float64)
    loss_fn = TSNELoss(X, perplexity=perplexity, theta=theta, dist_dtype=dist_dtype, device=device)
    torch.manual_seed(seed)
    net = mlp(X.shape[1], hidden).to(device)
    Xt = torch.tensor((X - X.mean(0)) / np.abs(X - X.mean(0)).max(), dtype=torch.float32, device=device)
    opt = torch.optim.Adam(net.parameters(), lr=lr)
    hist = {"step": [], "kl": []}
    t0 = time.perf_counter()
    for it in range(steps):
        opt.zero_grad()
        loss = loss_fn(net(Xt))
        loss.backward()
        opt.step()
        if it % log_every == 0 or it == steps - 1:
            hist["step"].append(it)
            hist["kl"].append(loss.item())
    torch.cuda.synchronize()
    hist["seconds"] = time.perf_counter() - t0
    with torch.no_grad():
        hist["final_kl"] = loss_fn(net(Xt)).item()
    return net, hist
