"""Seeded synthetic inputs for the five t-SNE workloads.

Each generator returns an (N, D) float array, one point per row (tapkee's
DenseMatrix is D x N column-major, i.e. the same bytes).  Seeds default to
20261203 + workload number so every run sees identical inputs; `digest()`
gives the SHA-256 that reports quote.
"""
from __future__ import annotations

import hashlib

import numpy as np

SEED0 = 20261203


def _rng(seed):
    return np.random.Generator(np.random.PCG64(seed))


def gaussian_mixture(n=2000, d=50, clusters=10, centre_sd=5.0, sd=1.0, seed=SEED0 + 1):
    """Workload 1: equal-size isotropic Gaussian clusters, float64."""
    g = _rng(seed)
    centres = g.normal(0.0, centre_sd, size=(clusters, d))
    labels = np.arange(n) % clusters
    X = centres[labels] + g.normal(0.0, sd, size=(n, d))
    return X.astype(np.float64), labels


def image_prototypes(n=20000, d=784, protos=10, p_zero=0.8, noise=0.1, seed=SEED0 + 2):
    """Workload 2: clipped noisy copies of sparse [0,1] prototypes, float32."""
    g = _rng(seed)
    P = g.uniform(0.0, 1.0, size=(protos, d))
    P[g.uniform(size=(protos, d)) < p_zero] = 0.0
    labels = np.arange(n) % protos
    X = np.clip(P[labels] + g.normal(0.0, noise, size=(n, d)), 0.0, 1.0)
    return X.astype(np.float32), labels


def swiss_roll_50d(n=100000, d=50, noise=0.05, seed=SEED0 + 3):
    """Workload 3: 3-D Swiss roll mapped to d dims by orthonormal rows, + noise."""
    g = _rng(seed)
    t = g.uniform(1.5 * np.pi, 4.5 * np.pi, size=n)
    h = g.uniform(0.0, 10.0, size=n)
    P3 = np.stack([t * np.cos(t), h, t * np.sin(t)], axis=1)
    Q, _ = np.linalg.qr(g.normal(size=(d, 3)))  # d x 3, orthonormal columns
    X = P3 @ Q.T + g.normal(0.0, noise, size=(n, d))
    return X.astype(np.float64), t


def anisotropic_mixture(n=250000, d=128, clusters=50, centre_sd=4.0, seed=SEED0 + 4):
    """Workload 4: Gaussian clusters with per-cluster diagonal scales U(0.5,2)."""
    g = _rng(seed)
    centres = g.normal(0.0, centre_sd, size=(clusters, d))
    scales = g.uniform(0.5, 2.0, size=(clusters, d))
    labels = np.arange(n) % clusters
    X = centres[labels] + g.normal(size=(n, d)) * scales[labels]
    return X.astype(np.float32), labels


def autograd_mixture(n=10000, d=784, seed=SEED0 + 5):
    """Workload 5: workload-1-style mixture at N=10k, D=784."""
    return gaussian_mixture(n=n, d=d, seed=seed)


def digest(a: np.ndarray) -> str:
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


WORKLOADS = {
    1: gaussian_mixture,
    2: image_prototypes,
    3: swiss_roll_50d,
    4: anisotropic_mixture,
    5: autograd_mixture,
}
