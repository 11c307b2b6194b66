"""CPU oracle of the hann hot path -- TEST INFRASTRUCTURE ONLY.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
import this package, and only as the checker (or as the timed CPU baseline).
The product (hann_amd) never imports, links or calls it.

liboracle.so is the C restatement in oracle.c (built by oracle/Makefile);
_ref/libhannref.so, when present, is the reference's own C distance code
compiled from source (used to write and re-check the golden vectors).
"""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")
REF_PATH = os.path.join(HERE, "_ref", "libhannref.so")

SQEUCLIDEAN, EUCLIDEAN, MANHATTAN, COSINE = 0, 1, 2, 3

_fp = ctypes.POINTER(ctypes.c_float)
_dp = ctypes.POINTER(ctypes.c_double)
_i32p = ctypes.POINTER(ctypes.c_int32)
_i64p = ctypes.POINTER(ctypes.c_int64)
_u8p = ctypes.POINTER(ctypes.c_uint8)
_szp = ctypes.POINTER(ctypes.c_size_t)
_u64p = ctypes.POINTER(ctypes.c_uint64)
_sz = ctypes.c_size_t
_int = ctypes.c_int
_u64 = ctypes.c_uint64

_SIGS = {
    "or_hash64": (_u64, [_u64]),
    "or_rng_next": (_u64, [_u64p]),
    "or_rng_bounded": (_u64, [_u64p, _u64]),
    "or_rng_perm": (None, [_u64p, ctypes.c_int64, _i64p]),
    "or_stream_seed": (_u64, [_u64, _u64, _u64]),
    "or_sqdiff": (ctypes.c_float, [_fp, _fp, _sz]),
    "or_normalize": (None, [_fp, _sz]),
    "or_euclidean": (ctypes.c_float, [_fp, _fp, _sz]),
    "or_manhattan": (ctypes.c_float, [_fp, _fp, _sz]),
    "or_cosine": (ctypes.c_float, [_fp, _fp, _sz]),
    "or_pair": (ctypes.c_float, [_int, _fp, _fp, _sz]),
    "or_batch": (None, [_int, _fp, _fp, _sz, _sz, _dp]),
    "or_batch_topk": (None, [_int, _fp, _sz, _fp, _sz, _sz, _sz, _int, _i64p, _fp]),
    "or_assign": (None, [_fp, _sz, _sz, _fp, _sz, _int, _i32p]),
    "or_kmeans_update": (None, [_fp, _sz, _sz, _i32p, _sz, _u64p, _fp]),
    "or_kmeans": (_sz, [_fp, _sz, _sz, _sz, _int, _u64p, _int, _fp]),
    "or_kmeanspp_seed": (None, [_fp, _sz, _sz, _sz, _u64p, _fp]),
    "or_kmeans_init": (_sz, [_fp, _sz, _sz, _sz, _int, _u64p, _int, _int, _fp]),
    "or_encode": (_int, [_fp, _fp, _sz, _sz, _fp, _szp, _sz, _i32p]),
    "or_train": (_int, [_fp, _sz, _sz, _sz, _sz, _sz, _int, _u64, _u64, _int, _fp, _szp, _i32p,
                        _fp, _szp, _i32p]),
    "or_search": (_int, [_fp, _sz, _fp, _sz, _i64p, _i64p, _fp, _i32p, _u8p, _fp, _szp, _sz, _sz,
                         _int, _int, _int, _int, _i64p, _dp, ctypes.POINTER(_int)]),
    "or_search_rerank": (_int, [_fp, _sz, _fp, _sz, _i64p, _i64p, _fp, _i32p, _u8p, _fp, _szp, _sz, _sz,
                                _int, _int, _int, _int, _int, _i64p, _dp, ctypes.POINTER(_int), _fp]),
    "or_search_rerank_many": (None, [_fp, _sz, _sz, _fp, _sz, _i64p, _i64p, _fp, _i32p, _u8p, _fp, _szp, _sz,
                                     _sz, _int, _i32p, _int, _i32p, _int, _int, _int, _i64p, _dp, _i32p, _fp]),
    "or_search_many": (None, [_fp, _sz, _sz, _fp, _sz, _i64p, _i64p, _fp, _i32p, _u8p, _fp, _szp, _sz,
                              _sz, _int, _int, _int, _int, _int, _i64p, _dp, _i32p]),
}

_lib = None


def build():
    """Compile liboracle.so (and _ref when the reference sources exist)."""
    subprocess.run(["make", "-s", "-C", HERE, "all"], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        l = ctypes.CDLL(LIB_PATH)
        for n, (r, a) in _SIGS.items():
            f = getattr(l, n)
            f.restype = r
            f.argtypes = a
        _lib = l
    return _lib


def f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


def pair(metric, a, b):
    a, b = f32(a), f32(b)
    return np.float32(lib().or_pair(metric, a.ctypes.data_as(_fp), b.ctypes.data_as(_fp), a.size))


def normalize(vec):
    """A normalized copy of vec (core/simd_kernels.inc.h:267-315, AVX2 tier)."""
    v = f32(vec).copy()
    lib().or_normalize(v.ctypes.data_as(_fp), v.size)
    return v


def batch(metric, q, flat, dim):
    q, flat = f32(q), f32(flat)
    n = flat.size // dim
    out = np.empty(n, np.float64)
    lib().or_batch(metric, q.ctypes.data_as(_fp), flat.ctypes.data_as(_fp), dim, n, out.ctypes.data_as(_dp))
    return out


def batch_topk(metric, queries, base, k, nthreads=1):
    queries, base = f32(queries), f32(base)
    nq, d = queries.shape
    ids = np.empty((nq, k), np.int64)
    rank = np.empty((nq, k), np.float32)
    lib().or_batch_topk(metric, queries.ctypes.data_as(_fp), nq, base.ctypes.data_as(_fp), base.shape[0], d, k,
                        nthreads, ids.ctypes.data_as(_i64p), rank.ctypes.data_as(_fp))
    return ids, rank


def assign(data, cents, nthreads=1):
    data, cents = f32(data), f32(cents)
    n, d = data.shape
    out = np.empty(n, np.int32)
    lib().or_assign(data.ctypes.data_as(_fp), n, d, cents.ctypes.data_as(_fp), cents.shape[0], nthreads,
                    out.ctypes.data_as(_i32p))
    return out


def kmeans_update(data, assign_, k, cents, rng_state):
    data = f32(data)
    cents = f32(cents).copy()
    a = np.ascontiguousarray(assign_, np.int32)
    st = (ctypes.c_uint64 * 1)(rng_state)
    lib().or_kmeans_update(data.ctypes.data_as(_fp), data.shape[0], data.shape[1], a.ctypes.data_as(_i32p), k,
                           st, cents.ctypes.data_as(_fp))
    return cents, st[0]


def kmeans(data, k, iters, rng_state, nthreads=1):
    data = f32(data)
    n, d = data.shape
    out = np.zeros((max(min(k, n), 1), d), np.float32)
    st = (ctypes.c_uint64 * 1)(rng_state)
    kk = lib().or_kmeans(data.ctypes.data_as(_fp), n, d, k, iters, st, nthreads, out.ctypes.data_as(_fp))
    return out[:kk], st[0]


def kmeans_init(data, k, iters, rng_state, init, nthreads=1):
    """or_kmeans_init: init 0 (permutation, the reference) or 1 (k-means++)."""
    data = f32(data)
    n, d = data.shape
    out = np.zeros((max(min(k, n), 1), d), np.float32)
    st = (ctypes.c_uint64 * 1)(rng_state)
    kk = lib().or_kmeans_init(data.ctypes.data_as(_fp), n, d, k, iters, st, init, nthreads, out.ctypes.data_as(_fp))
    return out[:kk], st[0]


def stream_seed(seed, epoch, role):
    return lib().or_stream_seed(seed, epoch, role)


def perm(state, n):
    st = (ctypes.c_uint64 * 1)(state)
    out = np.zeros(n, np.int64)
    lib().or_rng_perm(st, n, out.ctypes.data_as(_i64p))
    return out, st[0]


def encode(vec, centroid, m, codebooks, cb_len):
    vec, centroid, codebooks = f32(vec), f32(centroid), f32(codebooks)
    cbl = np.ascontiguousarray(cb_len, np.uint64)
    codes = np.zeros(m, np.int32)
    rc = lib().or_encode(vec.ctypes.data_as(_fp), centroid.ctypes.data_as(_fp), vec.size, m,
                         codebooks.ctypes.data_as(_fp), cbl.ctypes.data_as(_szp), codebooks.shape[1],
                         codes.ctypes.data_as(_i32p))
    return codes if rc == 0 else None


class TrainedIndex:
    """Output of or_train plus the CSR list layout that or_search reads."""

    def __init__(self, ids, vecs, coarse, assign_, codebooks, cb_len, codes, m):
        self.m = m
        self.coarse = coarse
        self.codebooks = codebooks
        self.cb_len = cb_len
        k = coarse.shape[0]
        order = np.argsort(assign_, kind="stable")  # list order = ascending id within a list
        self.list_off = np.zeros(k + 1, np.int64)
        np.add.at(self.list_off, assign_.astype(np.int64) + 1, 1)
        self.list_off = np.cumsum(self.list_off).astype(np.int64)
        self.ids = np.ascontiguousarray(ids[order], np.int64)
        self.vecs = f32(vecs[order])
        self.codes = np.ascontiguousarray(codes[order], np.int32)
        self.has_codes = np.ones(len(ids), np.uint8)
        self.assign = assign_


def train(ids, vecs, coarse_k, m, pq_k, iters, seed, epoch=0, nthreads=8):
    """or_train over vectors sorted by id; returns a TrainedIndex."""
    order = np.argsort(ids, kind="stable")
    ids = np.ascontiguousarray(np.asarray(ids)[order], np.int64)
    vecs = f32(np.asarray(vecs)[order])
    n, d = vecs.shape
    coarse = np.zeros((coarse_k, d), np.float32)
    kc = (ctypes.c_size_t * 1)()
    assign_ = np.zeros(n, np.int32)
    s = d // m
    codebooks = np.zeros((m, pq_k, s), np.float32)
    cb_len = np.zeros(m, np.uint64)
    codes = np.zeros((n, m), np.int32)
    rc = lib().or_train(vecs.ctypes.data_as(_fp), n, d, coarse_k, m, pq_k, iters, seed, epoch, nthreads,
                        coarse.ctypes.data_as(_fp), kc, assign_.ctypes.data_as(_i32p),
                        codebooks.ctypes.data_as(_fp), cb_len.ctypes.data_as(_szp), codes.ctypes.data_as(_i32p))
    if rc != 0:
        raise RuntimeError(f"or_train failed: {rc}")
    coarse = coarse[:kc[0]]
    return TrainedIndex(ids, vecs, coarse, assign_, codebooks, cb_len, codes, m)


def search_many(ix, queries, k, nprobe, allow_fallback=True, nthreads=1, have_codebooks=True):
    queries = f32(queries)
    nq, d = queries.shape
    ids = np.full((nq, k), -1, np.int64)
    dist = np.full((nq, k), np.inf, np.float64)
    cnt = np.zeros(nq, np.int32)
    lib().or_search_many(queries.ctypes.data_as(_fp), nq, d, ix.coarse.ctypes.data_as(_fp), ix.coarse.shape[0],
                         ix.list_off.ctypes.data_as(_i64p), ix.ids.ctypes.data_as(_i64p),
                         ix.vecs.ctypes.data_as(_fp), ix.codes.ctypes.data_as(_i32p),
                         ix.has_codes.ctypes.data_as(_u8p), ix.codebooks.ctypes.data_as(_fp),
                         ix.cb_len.ctypes.data_as(_szp), ix.codebooks.shape[1], ix.m, 1 if have_codebooks else 0,
                         nprobe, 1 if allow_fallback else 0, k, nthreads, ids.ctypes.data_as(_i64p),
                         dist.ctypes.data_as(_dp), cnt.ctypes.data_as(_i32p))
    return ids, dist, cnt


def search_rerank_many(ix, queries, k, nprobe, rerank, allow_fallback=True, nthreads=1, have_codebooks=True):
    """or_search_rerank for every query (k and nprobe scalars or per query).
    Returns ids [nq][kmax], distances, counts and each query's candidate
    bound (the PQ rank of the last of its rerank*k candidates, or +inf)."""
    queries = f32(queries)
    nq, d = queries.shape
    ks = np.ascontiguousarray(np.broadcast_to(np.asarray(k, np.int32), (nq,)), np.int32)
    nps = np.ascontiguousarray(np.broadcast_to(np.asarray(nprobe, np.int32), (nq,)), np.int32)
    kmax = int(ks.max()) if nq else 1
    ids = np.full((nq, kmax), -1, np.int64)
    dist = np.full((nq, kmax), np.inf, np.float64)
    cnt = np.zeros(nq, np.int32)
    bound = np.zeros(nq, np.float32)
    lib().or_search_rerank_many(queries.ctypes.data_as(_fp), nq, d, ix.coarse.ctypes.data_as(_fp),
                                ix.coarse.shape[0], ix.list_off.ctypes.data_as(_i64p), ix.ids.ctypes.data_as(_i64p),
                                ix.vecs.ctypes.data_as(_fp), ix.codes.ctypes.data_as(_i32p),
                                ix.has_codes.ctypes.data_as(_u8p), ix.codebooks.ctypes.data_as(_fp),
                                ix.cb_len.ctypes.data_as(_szp), ix.codebooks.shape[1], ix.m,
                                1 if have_codebooks else 0, nps.ctypes.data_as(_i32p), 1 if allow_fallback else 0,
                                ks.ctypes.data_as(_i32p), kmax, int(rerank), nthreads, ids.ctypes.data_as(_i64p),
                                dist.ctypes.data_as(_dp), cnt.ctypes.data_as(_i32p), bound.ctypes.data_as(_fp))
    return ids, dist, cnt, bound


class CSRIndex:
    """The CSR view or_search reads: list_off [K+1]; per entry ids, codes
    [N][m] (rows of -1: no codes), vectors [N][d]; coarse centroids,
    codebooks [m][P][d/m] and their valid row counts."""

    def __init__(self, list_off, ids, codes, vecs, coarse, codebooks, cb_len):
        self.list_off = np.ascontiguousarray(list_off, np.int64)
        self.ids = np.ascontiguousarray(ids, np.int64)
        self.codes = np.ascontiguousarray(codes, np.int32)
        self.m = self.codes.shape[1]
        self.vecs = f32(vecs)
        self.coarse = f32(coarse)
        self.codebooks = f32(codebooks)
        self.cb_len = np.ascontiguousarray(cb_len, np.uint64)
        self.has_codes = (self.codes[:, 0] >= 0).astype(np.uint8) if self.codes.size else np.zeros(0, np.uint8)
