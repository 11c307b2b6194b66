"""Seeded synthetic workloads (numpy PCG64 streams; same bytes on every host).

The reference's own generators (internal/testutil/data.go) draw from Go's
math/rand, which cannot be reproduced here, so every workload is generated
by these functions and fed identically to the GPU library and the oracle.
"""
import hashlib

import numpy as np


def sha256(*arrays):
    h = hashlib.sha256()
    for a in arrays:
        h.update(np.ascontiguousarray(a).tobytes())
    return h.hexdigest()


def gaussian_mixture(n, d, components, seed, center_low=-10.0, center_high=10.0, std=1.0, centers=None):
    """n points: uniform component choice, center + N(0, std^2 I).
    Shape of internal/testutil/data.go:19-39 (ClusteredData)."""
    rng = np.random.default_rng(seed)
    if centers is None:
        centers = rng.uniform(center_low, center_high, (components, d)).astype(np.float32)
    comp = rng.integers(0, components, n)
    x = centers[comp] + rng.standard_normal((n, d), dtype=np.float32) * np.float32(std)
    return x.astype(np.float32), centers


def clustered_queries(data, q, seed, noise=0.1):
    """queries near random data points (internal/testutil/data.go:44-61)"""
    rng = np.random.default_rng(seed)
    base = data[rng.integers(0, data.shape[0], q)]
    return (base + rng.standard_normal(base.shape, dtype=np.float32) * np.float32(noise)).astype(np.float32)


def config1(n=20000, d=64, q=200):
    """Plain PQIVF CPU-reference config: 32-component mixture, centers
    U[-10,10], std 1.0; queries from fresh mixture draws."""
    x, centers = gaussian_mixture(n, d, 32, seed=1)
    qs, _ = gaussian_mixture(q, d, 32, seed=2, centers=centers)
    return x, qs


def config2(n=1_000_000, d=128, q=10_000, seed_rates=3, seed_base=4, seed_queries=5):
    """Integer-valued data 0..255: per-dim Poisson rate ~ U[10,30], clipped."""
    rng = np.random.default_rng(seed_rates)
    lam = rng.uniform(10, 30, d)
    rb = np.random.default_rng(seed_base)
    base = np.minimum(rb.poisson(lam, (n, d)), 255).astype(np.float32)
    rq = np.random.default_rng(seed_queries)
    qs = np.minimum(rq.poisson(lam, (q, d)), 255).astype(np.float32)
    return base, qs


def normalized_mixture(n, d, components, std, seed, centers=None, chunk=1_000_000):
    """L2-normalised Gaussian-mixture vectors (centers N(0, I_d)), generated in
    chunks so that 10M x 96 fits in host memory once."""
    rng = np.random.default_rng(seed)
    if centers is None:
        centers = rng.standard_normal((components, d), dtype=np.float32)
        centers /= np.linalg.norm(centers, axis=1, keepdims=True)
    out = np.empty((n, d), np.float32)
    for s in range(0, n, chunk):
        e = min(n, s + chunk)
        comp = rng.integers(0, components, e - s)
        x = centers[comp] + rng.standard_normal((e - s, d), dtype=np.float32) * np.float32(std)
        x /= np.linalg.norm(x, axis=1, keepdims=True)
        out[s:e] = x
    return out, centers
