"""Write tests/golden/normalize_golden.npz from the reference's own C kernel.

core/simd_ops.c is compiled from source by oracle/Makefile into
oracle/_ref/libhannref.so together with core/simd_distance.c. This script
calls the per-ISA normalize symbols of that library (normalize_avx2, the tier
hann_cpu_init(2) installs, core/simd_ops.c:58-76, plus normalize_avx and the
scalar normalize_fallback) on seeded inputs and stores inputs and outputs as a
small fixture, so the oracle and the HIP kernel are pinned to the reference's
bits without the reference present.

Run from the repository root after `make -C oracle ref`:
    python tests/golden/make_normalize_golden.py
"""
import ctypes
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
LIB = os.path.join(ROOT, "oracle", "_ref", "libhannref.so")
OUT = os.path.join(ROOT, "tests", "golden", "normalize_golden.npz")

LENS = [1, 2, 3, 4, 7, 8, 9, 15, 16, 17, 24, 31, 32, 33, 40, 63, 64, 65, 96, 100, 128, 200, 784, 1000]
TIERS = ["avx2", "avx", "fallback"]


def main():
    if not os.path.exists(LIB):
        sys.exit("oracle/_ref/libhannref.so is missing: run `make -C oracle ref` first")
    lib = ctypes.CDLL(LIB)
    fp = ctypes.POINTER(ctypes.c_float)
    for t in TIERS:
        f = getattr(lib, f"normalize_{t}")
        f.restype = None
        f.argtypes = [fp, ctypes.c_size_t]
    rng = np.random.default_rng(20241204)
    vecs = []
    for n in LENS:
        vecs.append((rng.random(n, dtype=np.float32) * 20 - 10).astype(np.float32))      # uniform [-10, 10)
        vecs.append(rng.integers(0, 256, n).astype(np.float32))                          # integer-valued
        vecs.append((rng.standard_normal(n) * 1e-3).astype(np.float32))                  # small
        vecs.append((rng.standard_normal(n) * 3e4).astype(np.float32))                   # large
        vecs.append(np.zeros(n, np.float32))                                              # zero: unchanged
    out = {"lens": np.array([v.size for v in vecs], np.int64), "inputs": np.concatenate(vecs)}
    for t in TIERS:
        f = getattr(lib, f"normalize_{t}")
        res = []
        for v in vecs:
            w = v.copy()
            f(w.ctypes.data_as(fp), w.size)
            res.append(w)
        out[f"out_{t}"] = np.concatenate(res)
    np.savez_compressed(OUT, **out)
    diff = int(np.sum(out["out_avx2"].view(np.uint32) != out["out_fallback"].view(np.uint32)))
    print(f"wrote {OUT}: {len(vecs)} vectors, {out['inputs'].size} floats; avx2 vs fallback differ in {diff} floats")


if __name__ == "__main__":
    main()
