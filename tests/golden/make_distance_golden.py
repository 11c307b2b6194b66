"""Write tests/golden/distance_golden.npz from the reference's own C kernels.

The reference's core/simd_distance.c and core/simd_ops.c are compiled from
source by oracle/Makefile into oracle/_ref/libhannref.so. This script calls
the per-ISA symbols of that library (the AVX2+FMA tier, the AVX tier and the
scalar fallback, core/simd_distance.c:151-196) on seeded inputs and stores
inputs and outputs as a small fixture, so the oracle and the HIP kernels are
pinned to the reference's bits without the reference present.

Run from the repository root after `make -C oracle ref`:
    python tests/golden/make_distance_golden.py
"""
import ctypes
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
LIB = os.path.join(ROOT, "oracle", "_ref", "libhannref.so")
OUT = os.path.join(ROOT, "tests", "golden", "distance_golden.npz")

DIMS = [1, 2, 3, 4, 5, 6, 7, 8, 9, 15, 16, 17, 24, 25, 31, 32, 33, 40, 64, 96, 128, 200, 784]
PAIRS_PER_DIM = 6
TIERS = ["avx2", "avx", "fallback"]
METRICS = ["squared_euclidean", "euclidean", "manhattan", "cosine_distance"]
BATCH_SHAPES = [(25, 1), (25, 7), (128, 64), (784, 7), (8, 256), (4, 256), (96, 33)]


def main():
    if not os.path.exists(LIB):
        sys.exit("oracle/_ref/libhannref.so is missing: run `make -C oracle ref` first")
    lib = ctypes.CDLL(LIB)
    fp = ctypes.POINTER(ctypes.c_float)
    dp = ctypes.POINTER(ctypes.c_double)
    for m in METRICS:
        for t in TIERS:
            f = getattr(lib, f"{m}_{t}")
            f.restype = ctypes.c_float
            f.argtypes = [fp, fp, ctypes.c_size_t]
            g = getattr(lib, f"{m}_batch_{t}")
            g.restype = None
            g.argtypes = [fp, fp, ctypes.c_size_t, ctypes.c_size_t, dp]

    rng = np.random.default_rng(20240611)
    out = {}
    # per-pair vectors: uniform [-10, 10) like core/distance_test.go:213-214,
    # plus integer-valued, zero and near-identical cases
    a_all, b_all, dims_all = [], [], []
    for d in DIMS:
        for t in range(PAIRS_PER_DIM):
            if t < 3:
                a = (rng.random(d, dtype=np.float32) * 20 - 10).astype(np.float32)
                b = (rng.random(d, dtype=np.float32) * 20 - 10).astype(np.float32)
            elif t == 3:
                a = rng.integers(0, 256, d).astype(np.float32)
                b = rng.integers(0, 256, d).astype(np.float32)
            elif t == 4:
                a = np.zeros(d, np.float32)
                b = (rng.standard_normal(d) * 3).astype(np.float32)
            else:
                a = (rng.standard_normal(d) * 100).astype(np.float32)
                b = (a + rng.standard_normal(d).astype(np.float32) * 1e-3).astype(np.float32)
            a_all.append(a)
            b_all.append(b)
            dims_all.append(d)
    out["pair_dims"] = np.array(dims_all, np.int64)
    out["pair_a"] = np.concatenate(a_all)
    out["pair_b"] = np.concatenate(b_all)
    for m in METRICS:
        for t in TIERS:
            f = getattr(lib, f"{m}_{t}")
            vals = np.empty(len(dims_all), np.float32)
            for i, (a, b) in enumerate(zip(a_all, b_all)):
                vals[i] = f(a.ctypes.data_as(fp), b.ctypes.data_as(fp), a.size)
            out[f"pair_{m}_{t}"] = vals

    for si, (dim, n) in enumerate(BATCH_SHAPES):
        q = (rng.random(dim, dtype=np.float32) * 20 - 10).astype(np.float32)
        flat = (rng.random(dim * n, dtype=np.float32) * 20 - 10).astype(np.float32)
        out[f"batch{si}_q"] = q
        out[f"batch{si}_flat"] = flat
        for m in METRICS:
            res = np.empty(n, np.float64)
            getattr(lib, f"{m}_batch_avx2")(q.ctypes.data_as(fp), flat.ctypes.data_as(fp), dim, n,
                                            res.ctypes.data_as(dp))
            out[f"batch{si}_{m}_avx2"] = res
    out["batch_shapes"] = np.array(BATCH_SHAPES, np.int64)
    np.savez_compressed(OUT, **out)
    print(f"wrote {OUT}: {len(dims_all)} pairs, {len(BATCH_SHAPES)} batches")


if __name__ == "__main__":
    main()
