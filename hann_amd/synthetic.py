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


def config3(n=1_000_000, d=128, components=1024, seed=5, chunk=250_000):
    """Training at scale: a 1,024-component mixture, centers N(0, 25 I),
    per-dimension std 1.0."""
    rng = np.random.default_rng(seed)
    centers = (rng.standard_normal((components, d)) * 5).astype(np.float32)
    x = np.empty((n, d), np.float32)
    for s in range(0, n, chunk):
        e = min(n, s + chunk)
        x[s:e] = centers[rng.integers(0, components, e - s)] + rng.standard_normal((e - s, d), dtype=np.float32)
    return x, centers


# the batched ragged search workload: 10M x 96 L2-normalised mixture vectors
# (4,096 components, per-dimension std 0.06), 16,384 lists, PQ 24 x 256
CONFIG4 = {"n": 10_000_000, "dim": 96, "components": 4096, "sigma": 0.06, "nlist": 16384, "m": 24, "pq_k": 256,
           "kmeans_iters": 10, "train_sample": 1_000_000, "pq_train_sample": 65536, "seed": 1,
           "candidate_clusters": 8, "data_seed": 7}


def config4(n=None):
    """The vectors of the batched search workload (and the mixture centers)."""
    c = CONFIG4
    return normalized_mixture(c["n"] if n is None else n, c["dim"], c["components"], c["sigma"], seed=c["data_seed"])


def config4_batch(centers, nq, seed, sigma=None):
    """One query batch: queries from the same mixture, each with its own k in
    {10, 50, 100} and nprobe in {8, 16, 32, 64, 128}."""
    sigma = CONFIG4["sigma"] if sigma is None else sigma
    q, _ = normalized_mixture(nq, centers.shape[1], centers.shape[0], sigma, seed, centers=centers)
    rng = np.random.default_rng(seed + 1)
    k = rng.choice(np.array([10, 50, 100], np.int32), nq).astype(np.int32)
    nprobe = rng.choice(np.array([8, 16, 32, 64, 128], np.int32), nq).astype(np.int32)
    return q, k, nprobe
