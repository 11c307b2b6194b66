"""Host-side mirror of the reference's core package (core/metric.go,
core/distance.go, core/vector_ops.go) over the hann_amd C-ABI.

Names, argument meaning and error messages follow the Go API, so the tests
read like the reference's own: Metric.Distance / Rank / FromRank /
DistanceBatch / RankBatch, and the built-ins Euclidean, SquaredEuclidean,
Manhattan and Cosine (core/metric.go:150-165). Every computation runs on the
GPU through libhann_amd.so.
"""
import math

import numpy as np

from . import _lib
from ._lib import HannError


def _f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


class Metric:
    """core.Metric (core/metric.go:14-22) for the four built-in metrics."""

    def __init__(self, name, metric_id, rank_id, from_rank, normalize):
        self._name = name
        self._id = metric_id
        self._rank_id = rank_id
        self._from_rank = from_rank
        self._normalize = normalize

    def Name(self):
        return self._name

    def Normalizes(self):
        return self._normalize

    @property
    def id(self):
        return self._id

    def _pair(self, mid, a, b):
        # core/distance.go:19-33: length check, then 0 for empty vectors
        a, b = _f32(a), _f32(b)
        if a.size != b.size:
            raise HannError(_lib.HANN_ERR_INVALID_ARG, "vectors must have the same length")
        if a.size == 0:
            return 0.0
        lib = _lib.load()
        fn = {_lib.METRIC_SQEUCLIDEAN: lib.simd_squared_euclidean,
              _lib.METRIC_EUCLIDEAN: lib.simd_euclidean,
              _lib.METRIC_MANHATTAN: lib.simd_manhattan,
              _lib.METRIC_COSINE: lib.simd_cosine_distance}[mid]
        v = fn(_lib.fptr(a), _lib.fptr(b), a.size)
        if math.isnan(v) and lib.hann_last_error():
            _lib.check(_lib.HANN_ERR_DEVICE)
        return float(v)

    def Distance(self, a, b):
        return self._pair(self._id, a, b)

    def Rank(self, a, b):
        # core/metric.go:41-46
        return self._pair(self._rank_id, a, b)

    def FromRank(self, rank):
        # core/metric.go:49-54
        return self._from_rank(rank) if self._from_rank else rank

    @staticmethod
    def _batch_size(query, flat, out):
        # core/metric.go:104-117
        if query.size == 0:
            raise HannError(_lib.HANN_ERR_INVALID_ARG, "query must not be empty")
        if flat.size % query.size != 0:
            raise HannError(_lib.HANN_ERR_INVALID_ARG,
                            f"flat length {flat.size} is not a multiple of the query length {query.size}")
        n = flat.size // query.size
        if out.size != n:
            raise HannError(_lib.HANN_ERR_INVALID_ARG,
                            f"out length {out.size} does not match the candidate count {n}")
        return n

    def _batch(self, rank, query, flat, out):
        query, flat = _f32(query), _f32(flat)
        if not isinstance(out, np.ndarray) or out.dtype != np.float64 or not out.flags.c_contiguous:
            raise HannError(_lib.HANN_ERR_INVALID_ARG, "out must be a contiguous float64 array")
        n = self._batch_size(query, flat, out)
        if n == 0:
            return
        _lib.check(_lib.load().hann_distance_batch(self._id, 1 if rank else 0, _lib.fptr(query), 1,
                                                   _lib.fptr(flat), n, query.size, _lib.dptr(out)))

    def DistanceBatch(self, query, flat, out):
        """core/metric.go:64-74: out[i] = Distance(query, row i of flat)."""
        self._batch(False, query, flat, out)

    def RankBatch(self, query, flat, out):
        """core/metric.go:82-100: out[i] = Rank(query, row i of flat)."""
        self._batch(True, query, flat, out)

    def distance_matrix(self, queries, base, rank=False):
        """Many queries at once: [nq][nb] float64 (build extension)."""
        queries, base = _f32(queries), _f32(base)
        nq, d = queries.shape
        nb = base.shape[0]
        out = np.empty((nq, nb), np.float64)
        if nq and nb:
            _lib.check(_lib.load().hann_distance_batch(self._id, 1 if rank else 0, _lib.fptr(queries), nq,
                                                       _lib.fptr(base), nb, d, _lib.dptr(out)))
        return out


Euclidean = Metric("euclidean", _lib.METRIC_EUCLIDEAN, _lib.METRIC_SQEUCLIDEAN, math.sqrt, False)
SquaredEuclidean = Metric("squared_euclidean", _lib.METRIC_SQEUCLIDEAN, _lib.METRIC_SQEUCLIDEAN, None, False)
Manhattan = Metric("manhattan", _lib.METRIC_MANHATTAN, _lib.METRIC_MANHATTAN, None, False)
Cosine = Metric("cosine", _lib.METRIC_COSINE, _lib.METRIC_COSINE, None, True)

_REGISTRY = {m.Name(): m for m in (Euclidean, SquaredEuclidean, Manhattan, Cosine)}


def MetricByName(name):
    """core/metric.go:197-202"""
    m = _REGISTRY.get(name)
    return m, m is not None


def batch_topk(metric, queries, base, k, precision=_lib.PREC_FP32, overfetch=0):
    """Exact top-k rows of `base` for every query: (ids [nq][k], dists [nq][k]).

    Ordered by (rank, row); ids are -1 and dists inf where base has fewer
    than k rows."""
    queries, base = _f32(queries), _f32(base)
    if queries.ndim != 2 or base.ndim != 2 or queries.shape[1] != base.shape[1]:
        raise HannError(_lib.HANN_ERR_INVALID_ARG, "queries and base must be 2-D with equal widths")
    nq, d = queries.shape
    ids = np.empty((nq, k), np.int64)
    dist = np.empty((nq, k), np.float64)
    _lib.check(_lib.load().hann_batch_topk(metric.id, _lib.fptr(queries), nq, _lib.fptr(base), base.shape[0],
                                           d, k, precision, overfetch, _lib.i64ptr(ids), _lib.dptr(dist)))
    return ids, dist


def ExhaustiveFallback(metric, query, node_ids, node_vecs, k, candidate_ids=()):
    """The exhaustive fallback of HNSW's Search (hnsw/index.go:1198-1319) on
    the GPU batch top-k: the k nearest nodes by metric.Rank that are not
    already graph candidates, from the nodes taken in ascending id order.

    Returns (ids, ranks): at most k ids with their rank values (FromRank not
    applied: the reference merges them with the graph candidates and sorts
    by (rank, id) before FromRank), ordered by (rank, id). The reference
    keeps them in per-worker heaps whose choice among equal ranks at the
    k-th place is unspecified; here equal ranks go by id. Rows of the base
    are the nodes in ascending id order, so the top-k's (rank, row) order is
    (rank, id): the k + |candidates| best rows hold the k best
    non-candidates."""
    q = _f32(query).reshape(1, -1)
    ids = np.ascontiguousarray(node_ids, np.int64)
    vecs = _f32(node_vecs)
    if vecs.ndim != 2 or vecs.shape[0] != ids.size or vecs.shape[1] != q.shape[1]:
        raise HannError(_lib.HANN_ERR_INVALID_ARG, "node vectors must be [len(node_ids)][dim]")
    order = np.argsort(ids, kind="stable")
    ids_sorted, rows = ids[order], np.ascontiguousarray(vecs[order])
    cand = np.asarray(list(candidate_ids), np.int64)
    n = ids.size - int(np.isin(cand, ids).sum()) if cand.size else ids.size
    if k <= 0 or n <= 0:
        return np.zeros(0, np.int64), np.zeros(0, np.float64)
    kk = min(ids.size, k + cand.size)
    # Rank: the metric's rank metric (squared Euclidean for Euclidean,
    # core/metric.go:155-158), so the values are ranks
    rank_metric = {_lib.METRIC_EUCLIDEAN: SquaredEuclidean}.get(metric.id, metric)
    rid, rank = batch_topk(rank_metric, q, rows, kk)
    rid, rank = rid[0], rank[0]
    got = ids_sorted[rid[rid >= 0]]
    rank = rank[rid >= 0]
    keep = ~np.isin(got, cand) if cand.size else np.ones(got.size, bool)
    return got[keep][:k], rank[keep][:k]


def NormalizeVector(vec):
    """core/vector_ops.go:16-22: scale a float32 array to unit norm in place
    (a zero vector is left as it is; an empty one is a no-op)."""
    if not (isinstance(vec, np.ndarray) and vec.dtype == np.float32 and vec.flags.c_contiguous):
        raise HannError(_lib.HANN_ERR_INVALID_ARG, "NormalizeVector needs a contiguous float32 array")
    if vec.size == 0:
        return
    _lib.load().simd_normalize(_lib.fptr(vec), vec.size)


def NormalizeBatch(vecs):
    """core/vector_ops.go:24-54: normalize every vector in place. The
    reference runs one simd_normalize per vector on a worker pool; here all
    vectors of one length go to the GPU in one launch."""
    by_len = {}
    for i, v in enumerate(vecs):
        if not (isinstance(v, np.ndarray) and v.dtype == np.float32 and v.flags.c_contiguous):
            raise HannError(_lib.HANN_ERR_INVALID_ARG, "NormalizeBatch needs contiguous float32 arrays")
        if v.size:
            by_len.setdefault(v.size, []).append(i)
    for dim, idx in by_len.items():
        flat = np.ascontiguousarray(np.stack([vecs[i] for i in idx]))
        _lib.check(_lib.load().hann_normalize_batch(_lib.fptr(flat), len(idx), dim))
        for j, i in enumerate(idx):
            vecs[i][:] = flat[j]
