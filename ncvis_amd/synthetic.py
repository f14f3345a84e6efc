"""Synthetic datasets for the NCVis fit-path workloads (seeded, float32, C-contiguous).

Shapes and distributions follow the workload definitions used by the tests and the benchmark:

* ``digits_like``   -- N=5,620, D=64, integer values 0..16 from 10 isotropic Gaussian clusters
  (centres uniform in [2,14], sigma 2, rounded and clipped). The shape of the UCI optical-digits
  set the reference's README times (README.md:35-47).
* ``normalized_mixture`` -- 50-cluster mixture with N(0,1) centres, within-cluster sigma 0.3,
  L2-normalised rows (cosine workload, N=100,000, D=300).
* ``anisotropic_mixture`` -- 100 clusters, N(0,1) centres, per-cluster diagonal scales
  log-uniform in [0.1, 1] (the N=1,000,000, D=128 precision workload).
* ``gaussian_mixture`` -- generic isotropic mixture (N=200,000, D=64, 20 clusters; and the
  multi-dataset batches).
"""
from __future__ import annotations

import numpy as np


def _labels(rng, N, n_clusters, equal=True):
    if equal:
        lab = np.repeat(np.arange(n_clusters), -(-N // n_clusters))[:N]
        return lab
    return rng.integers(0, n_clusters, N)


def digits_like(N: int = 5620, D: int = 64, n_clusters: int = 10, seed: int = 0):
    rng = np.random.default_rng(seed)
    centers = rng.uniform(2.0, 14.0, (n_clusters, D))
    lab = _labels(rng, N, n_clusters)
    X = np.rint(centers[lab] + rng.normal(0.0, 2.0, (N, D)))
    X = np.clip(X, 0, 16).astype(np.float32)
    return np.ascontiguousarray(X), lab


def normalized_mixture(N: int = 100_000, D: int = 300, n_clusters: int = 50, sigma: float = 0.3,
                       seed: int = 0):
    rng = np.random.default_rng(seed)
    centers = rng.normal(0.0, 1.0, (n_clusters, D)).astype(np.float32)
    lab = rng.integers(0, n_clusters, N)
    X = centers[lab] + rng.normal(0.0, sigma, (N, D)).astype(np.float32)
    X /= np.linalg.norm(X, axis=1, keepdims=True)
    return np.ascontiguousarray(X, np.float32), lab


def anisotropic_mixture(N: int = 1_000_000, D: int = 128, n_clusters: int = 100, seed: int = 0,
                        chunk: int = 200_000):
    rng = np.random.default_rng(seed)
    centers = rng.normal(0.0, 1.0, (n_clusters, D)).astype(np.float32)
    scales = np.exp(rng.uniform(np.log(0.1), np.log(1.0), (n_clusters, D))).astype(np.float32)
    lab = rng.integers(0, n_clusters, N)
    X = np.empty((N, D), np.float32)
    for s in range(0, N, chunk):
        e = min(N, s + chunk)
        z = rng.standard_normal((e - s, D), dtype=np.float32)
        X[s:e] = centers[lab[s:e]] + z * scales[lab[s:e]]
    return X, lab


def gaussian_mixture(N: int, D: int, n_clusters: int, sigma: float = 1.0, center_scale: float = 4.0,
                     seed: int = 0):
    rng = np.random.default_rng(seed)
    centers = rng.normal(0.0, center_scale, (n_clusters, D)).astype(np.float32)
    lab = rng.integers(0, n_clusters, N)
    X = centers[lab] + rng.normal(0.0, sigma, (N, D)).astype(np.float32)
    return np.ascontiguousarray(X, np.float32), lab

