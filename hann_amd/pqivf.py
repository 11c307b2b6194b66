"""Host-side mirror of the reference's pqivf package over the hann_amd C-ABI.

pqivf.New(dim, *options) and the Index methods keep the Go names, argument
meaning and error messages of pqivf/index.go, so the tests read like the
reference's own (pqivf/index_test.go). The bookkeeping runs in the C++
runtime inside libhann_amd.so and every computation on the GPU.
"""
import collections
import ctypes
import os
import time

import numpy as np

from . import _lib
from ._lib import HannError

Neighbor = collections.namedtuple("Neighbor", ["ID", "Distance"])          # core/index.go:125-128
IndexStats = collections.namedtuple("IndexStats", ["Count", "Dimension", "Distance", "FallbackSearches"])

_BIT = {"coarse_k": 0, "num_subquantizers": 1, "pq_k": 2, "kmeans_iters": 3, "candidate_clusters": 4}


def GetSeed():
    """core/utils.go:12-20: HANN_SEED when set, otherwise the current time."""
    s = os.environ.get("HANN_SEED")
    if s is not None:
        try:
            return int(s) & 0xFFFFFFFFFFFFFFFF
        except ValueError:
            pass
    return time.time_ns() & 0xFFFFFFFFFFFFFFFF


def _opt(field, value):
    def apply(o):
        setattr(o, field, value)
        if field in _BIT:
            o.explicit_mask |= 1 << _BIT[field]
    return apply


def WithCoarseK(n):
    return _opt("coarse_k", n)


def WithNumSubquantizers(n):
    return _opt("num_subquantizers", n)


def WithPQK(n):
    return _opt("pq_k", n)


def WithKMeansIters(n):
    return _opt("kmeans_iters", n)


def WithCandidateClusters(n):
    return _opt("candidate_clusters", n)


def WithBruteForceFallback(allow):
    def apply(o):
        o.bruteforce_fallback = 1 if allow else 0
        o.explicit_mask |= 1 << 5
    return apply


# build-side extensions
def WithSeed(seed):
    return _opt("seed", seed)


def WithTrainSample(n):
    return _opt("train_sample", n)


def WithPQTrainSample(n):
    return _opt("pq_train_sample", n)


def WithKMeansInit(init):
    """Build extension: 0 = permutation seeding (the reference), 1 = k-means++."""
    def f(o):
        o.kmeans_init = int(init)
    return f


def WithTrainPrecision(p):
    """_lib.PREC_BF16: certified BF16 MFMA assignment in Train (same results)."""
    return _opt("train_precision", p)


def New(dimension, *opts):
    """pqivf.New (pqivf/index.go:106-146)."""
    lib = _lib.load()
    o = _lib.PQIVFOptions()
    lib.hann_pqivf_default_options(ctypes.byref(o))
    o.seed = GetSeed()
    for f in opts:
        f(o)
    h = ctypes.c_void_p()
    _lib.check(lib.hann_pqivf_new(int(dimension), ctypes.byref(o), ctypes.byref(h)))
    return Index(h, int(dimension))


def _vec(v):
    return np.ascontiguousarray(v, dtype=np.float32).reshape(-1)


def _pack(vectors):
    """map[int][]float32 -> (ids, flat, lens) in the caller's order."""
    ids = np.fromiter((int(k) for k in vectors.keys()), dtype=np.int64, count=len(vectors))
    vs = [_vec(vectors[k]) for k in vectors.keys()]
    lens = np.array([v.size for v in vs], dtype=np.uint64)
    flat = np.concatenate(vs) if vs else np.zeros(1, np.float32)
    return ids, np.ascontiguousarray(flat, np.float32), lens


class Index:
    """pqivf.Index: core.Index + core.BulkIndex + core.Trainer."""

    def __init__(self, handle, dim):
        self._h = handle
        self._dim = dim
        self._lib = _lib.load()

    def __del__(self):
        h, self._h = getattr(self, "_h", None), None
        if h:
            try:
                self._lib.hann_pqivf_free(h)
            except Exception:
                pass

    @property
    def handle(self):
        return self._h

    # -- mutations
    def Add(self, id, vector):
        v = _vec(vector)
        _lib.check(self._lib.hann_pqivf_add(self._h, int(id), _lib.fptr(v), v.size))

    def BulkAdd(self, vectors):
        ids, flat, lens = _pack(vectors)
        _lib.check(self._lib.hann_pqivf_bulk_add(self._h, _lib.i64ptr(ids), _lib.fptr(flat), _lib.sizeptr(lens),
                                                 ids.size))

    def bulk_add_arrays(self, ids, vecs):
        """BulkAdd from an id array and a [n][dim] float32 matrix."""
        ids = np.ascontiguousarray(ids, np.int64)
        vecs = np.ascontiguousarray(vecs, np.float32)
        _lib.check(self._lib.hann_pqivf_bulk_add(self._h, _lib.i64ptr(ids), _lib.fptr(vecs), None, ids.size))

    def Delete(self, id):
        _lib.check(self._lib.hann_pqivf_delete(self._h, int(id)))

    def BulkDelete(self, ids):
        a = np.ascontiguousarray(list(ids) if not isinstance(ids, np.ndarray) else ids, np.int64)
        _lib.check(self._lib.hann_pqivf_bulk_delete(self._h, _lib.i64ptr(a) if a.size else None, a.size))

    def Update(self, id, vector):
        v = _vec(vector)
        _lib.check(self._lib.hann_pqivf_update(self._h, int(id), _lib.fptr(v), v.size))

    def BulkUpdate(self, updates):
        ids, flat, lens = _pack(updates)
        _lib.check(self._lib.hann_pqivf_bulk_update(self._h, _lib.i64ptr(ids), _lib.fptr(flat), _lib.sizeptr(lens),
                                                    ids.size))

    def Train(self):
        _lib.check(self._lib.hann_pqivf_train(self._h))

    # -- queries
    def Search(self, query, k):
        """Search (pqivf/index.go:850-949): Neighbors in non-decreasing distance."""
        q = _vec(query)
        kk = max(int(k), 1)
        ids = np.empty(kk, np.int64)
        dist = np.empty(kk, np.float64)
        cnt = ctypes.c_int64()
        _lib.check(self._lib.hann_pqivf_search(self._h, _lib.fptr(q), q.size, int(k), _lib.i64ptr(ids),
                                               _lib.dptr(dist), ctypes.byref(cnt)))
        return [Neighbor(int(ids[i]), float(dist[i])) for i in range(cnt.value)]

    def search_batch(self, queries, k, nprobe=None, scan_mode=_lib.SCAN_EXACT, rerank=0,
                     coarse_precision=_lib.PREC_FP32, fused=0, list_config=0):
        """Batched search (build extension): k and nprobe scalars or per-query;
        list_config pins a list-major scan configuration (0: tuned)."""
        q = np.ascontiguousarray(queries, np.float32)
        nq = q.shape[0]
        ks = np.ascontiguousarray(np.broadcast_to(np.asarray(k, np.int32), (nq,)), np.int32)
        kmax = int(ks.max()) if nq else 1
        nps = None if nprobe is None else np.ascontiguousarray(np.broadcast_to(np.asarray(nprobe, np.int32), (nq,)),
                                                                np.int32)
        p = _lib.SearchParams(scan_mode, coarse_precision, rerank, fused, list_config)
        ids = np.full((nq, kmax), -1, np.int64)
        dist = np.full((nq, kmax), np.inf, np.float64)
        cnt = np.zeros(nq, np.int32)
        if nq:
            _lib.check(self._lib.hann_pqivf_search_batch(
                self._h, _lib.fptr(q), nq, _lib.i32ptr(ks), _lib.i32ptr(nps) if nps is not None else None,
                ctypes.byref(p), _lib.i64ptr(ids), _lib.dptr(dist), _lib.i32ptr(cnt)))
        return ids, dist, cnt

    def Stats(self):
        s = _lib.IndexStats()
        _lib.check(self._lib.hann_pqivf_stats(self._h, ctypes.byref(s)))
        return IndexStats(int(s.count), int(s.dimension), s.distance.decode(), int(s.fallback_searches))

    # -- persistence
    def save_bytes(self):
        buf = ctypes.c_void_p()
        n = ctypes.c_size_t()
        _lib.check(self._lib.hann_pqivf_save(self._h, ctypes.byref(buf), ctypes.byref(n)))
        try:
            return ctypes.string_at(buf, n.value)
        finally:
            self._lib.hann_free_buffer(buf)

    def load_bytes(self, data):
        b = ctypes.create_string_buffer(bytes(data), len(data))
        _lib.check(self._lib.hann_pqivf_load(self._h, ctypes.cast(b, ctypes.c_void_p), len(data)))

    def _take(self, fn, *args):
        buf = ctypes.c_void_p()
        n = ctypes.c_size_t()
        _lib.check(fn(self._h, *args, ctypes.byref(buf), ctypes.byref(n)))
        try:
            return ctypes.string_at(buf, n.value)
        finally:
            self._lib.hann_free_buffer(buf)

    def _give(self, fn, data):
        data = bytes(data)
        b = ctypes.create_string_buffer(data, len(data))
        _lib.check(fn(self._h, ctypes.cast(b, ctypes.c_void_p), len(data)))

    def GobEncode(self, format_version=1):
        """GobEncode (pqivf/index.go:991-1017): the gob stream of serializedPQIVF."""
        return self._take(self._lib.hann_pqivf_gob_encode, format_version)

    def GobDecode(self, data):
        """GobDecode (pqivf/index.go:1020-1084)."""
        self._give(self._lib.hann_pqivf_gob_decode, data)

    def Save(self, w, format_version=1):
        """Save (pqivf/index.go:1086-1092): the reference's gob file."""
        w.write(self._take(self._lib.hann_pqivf_save_gob, format_version))

    def Load(self, r):
        """Load (pqivf/index.go:1094-1100): a gob file written by Save here or by Go hann."""
        self._give(self._lib.hann_pqivf_load_gob, r.read())

    # -- measurement (build extension)
    def set_profile(self, on):
        _lib.check(self._lib.hann_pqivf_profile(self._h, 1 if on else 0))

    def stage_times(self):
        """(coarse_ms, scan_ms, rerank_ms, searches) accumulated since the last call."""
        a, b, c = ctypes.c_double(), ctypes.c_double(), ctypes.c_double()
        n = ctypes.c_int64()
        _lib.check(self._lib.hann_pqivf_stage_times(self._h, ctypes.byref(a), ctypes.byref(b), ctypes.byref(c),
                                                    ctypes.byref(n)))
        return a.value, b.value, c.value, n.value

    def last_scan(self):
        """(mode, fused) of the scan the last search ran: a SCAN_* mode (the
        list modes only when the list-major scan ran) and whether the
        persistent table scan ran."""
        m, f = ctypes.c_int(), ctypes.c_int()
        _lib.check(self._lib.hann_pqivf_last_scan(self._h, ctypes.byref(m), ctypes.byref(f)))
        return m.value, f.value

    def scan_kernel_time(self):
        """Milliseconds in the main scan kernel over the profiled searches
        (call after stage_times)."""
        v = ctypes.c_double()
        _lib.check(self._lib.hann_pqivf_scan_kernel_time(self._h, ctypes.byref(v)))
        return v.value

    def last_tune(self):
        """The autotuner's last choice on this index, as text."""
        buf = ctypes.create_string_buffer(512)
        _lib.check(self._lib.hann_pqivf_last_tune(self._h, buf, 512))
        return buf.value.decode()

    def scan_kernel_times(self):
        """(filter_ms, sample_ms): milliseconds in the main scan kernel (the
        list-major scan's filter pass) and in the list-major sample pass over
        the profiled searches (call after stage_times)."""
        f, s = ctypes.c_double(), ctypes.c_double()
        _lib.check(self._lib.hann_pqivf_scan_kernel_times(self._h, ctypes.byref(f), ctypes.byref(s)))
        return f.value, s.value

    def export_lists(self, with_vectors=True):
        """(list_off [k+1], ids [N], codes [N][m], vecs [N][dim] or None)"""
        k, m, _, d = self.shape()
        off = np.zeros(k + 1, np.int64)
        _lib.check(self._lib.hann_pqivf_export_lists(self._h, _lib.i64ptr(off), None, None, None))
        n = int(off[-1])
        ids = np.zeros(max(n, 1), np.int64)
        codes = np.zeros((max(n, 1), m), np.int32)
        vecs = np.zeros((max(n, 1), d), np.float32) if with_vectors else None
        _lib.check(self._lib.hann_pqivf_export_lists(self._h, None, _lib.i64ptr(ids), _lib.i32ptr(codes),
                                                     _lib.fptr(vecs) if with_vectors else None))
        return off, ids[:n], codes[:n], (vecs[:n] if with_vectors else None)

    def search_batch_dev(self, q_ptr, nq, k_ptr, np_ptr, kmax, max_nprobe, out_ids_ptr, out_dist_ptr,
                         out_cnt_ptr, stream, scan_mode=_lib.SCAN_ADC_F32, rerank=0,
                         coarse_precision=_lib.PREC_FP32, fused=0):
        """Device-resident batch search (pointers are device addresses)."""
        p = _lib.SearchParams(scan_mode, coarse_precision, rerank, fused)
        _lib.check(self._lib.hann_pqivf_search_batch_dev(self._h, q_ptr, nq, k_ptr, np_ptr, kmax, max_nprobe,
                                                         ctypes.byref(p), out_ids_ptr, out_dist_ptr, out_cnt_ptr,
                                                         stream))

    def search_graph(self, q_ptr, nq, k_ptr, np_ptr, kmax, max_nprobe, out_ids_ptr, out_dist_ptr, out_cnt_ptr,
                     stream, scan_mode=_lib.SCAN_ADC_F32, rerank=0, coarse_precision=_lib.PREC_FP32, fused=0):
        """A device search bound to these buffers, replayed as a HIP graph."""
        p = _lib.SearchParams(scan_mode, coarse_precision, rerank, fused)
        h = ctypes.c_void_p()
        _lib.check(self._lib.hann_pqivf_search_graph_create(self._h, q_ptr, nq, k_ptr, np_ptr, kmax, max_nprobe,
                                                            ctypes.byref(p), out_ids_ptr, out_dist_ptr,
                                                            out_cnt_ptr, stream, ctypes.byref(h)))
        return SearchGraph(self, h)

    # -- test hooks (pqivf/export_test.go)
    def shape(self):
        k, m, p, d = ctypes.c_int64(), ctypes.c_int64(), ctypes.c_int64(), ctypes.c_int64()
        _lib.check(self._lib.hann_pqivf_shape(self._h, ctypes.byref(k), ctypes.byref(m), ctypes.byref(p),
                                              ctypes.byref(d)))
        return k.value, m.value, p.value, d.value

    def trained(self):
        return bool(self._lib.hann_pqivf_is_trained(self._h))


class SearchGraph:
    """hann_search_graph: launch() searches the bound buffers again."""

    def __init__(self, index, handle):
        self._index = index   # the index must outlive the graph
        self._lib = index._lib
        self._h = handle

    def launch(self):
        _lib.check(self._lib.hann_search_graph_launch(self._h))

    def counts(self):
        c, r, e = ctypes.c_int64(), ctypes.c_int64(), ctypes.c_int64()
        _lib.check(self._lib.hann_search_graph_counts(self._h, ctypes.byref(c), ctypes.byref(r), ctypes.byref(e)))
        return {"captures": c.value, "replays": r.value, "eager": e.value}

    def close(self):
        if self._h:
            self._lib.hann_search_graph_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def AssignPoints(data, centroids, precision=_lib.PREC_FP32):
    """assignPoints (pqivf/index.go:738-789) on the GPU."""
    data = np.ascontiguousarray(data, np.float32)
    cents = np.ascontiguousarray(centroids, np.float32)
    n = data.shape[0]
    out = np.empty(n, np.int32)
    if n == 0:
        return out
    _lib.check(_lib.load().hann_assign_points(_lib.fptr(data), n, data.shape[1], _lib.fptr(cents), cents.shape[0],
                                              precision, _lib.i32ptr(out)))
    return out


def KMeans(data, k, iters, seed, precision=_lib.PREC_FP32, init=0):
    """runKMeans; init 0 = the reference's permutation seeding, 1 = k-means++."""
    data = np.ascontiguousarray(data, np.float32)
    n, d = data.shape
    out = np.zeros((max(min(k, n), 1), d), np.float32)
    kk = ctypes.c_size_t()
    _lib.check(_lib.load().hann_kmeans(_lib.fptr(data), n, d, k, iters, seed, init, precision, _lib.fptr(out),
                                       ctypes.byref(kk)))
    return out[:kk.value]


def EncodeVector(idx, vector, cluster):
    v = _vec(vector)
    _, m, _, _ = idx.shape()
    codes = np.zeros(m, np.int32)
    _lib.check(_lib.load().hann_pqivf_encode_vector(idx.handle, _lib.fptr(v), v.size, int(cluster),
                                                    _lib.i32ptr(codes)))
    return codes.tolist()


def CoarseCentroids(idx):
    k, m, p, d = idx.shape()
    out = np.zeros((k, d), np.float32)
    if k:
        _lib.check(_lib.load().hann_pqivf_coarse_centroids(idx.handle, _lib.fptr(out)))
    return out


def Codebooks(idx):
    """[m][p][dim/m] codebooks and the valid row count of each."""
    k, m, p, d = idx.shape()
    out = np.zeros((m, max(p, 0), d // m if m else 0), np.float32)
    cbl = np.zeros(max(m, 1), np.int64)
    _lib.check(_lib.load().hann_pqivf_codebooks(idx.handle, _lib.fptr(out) if out.size else None,
                                                _lib.i64ptr(cbl)))
    return out, cbl[:m]


def InvertedIDs(idx):
    """{cluster: [ids in storage order]} for every present list key."""
    k, _, _, _ = idx.shape()
    lib = _lib.load()
    res = {}
    n = ctypes.c_int64()
    for c in range(k):
        _lib.check(lib.hann_pqivf_list_ids(idx.handle, c, None, ctypes.byref(n)))
        if n.value < 0:
            continue
        ids = np.zeros(max(n.value, 1), np.int64)
        _lib.check(lib.hann_pqivf_list_ids(idx.handle, c, _lib.i64ptr(ids), ctypes.byref(n)))
        res[c] = ids[:n.value].tolist()
    return res


def InvertedCodes(idx):
    k, m, _, _ = idx.shape()
    lib = _lib.load()
    res = {}
    n = ctypes.c_int64()
    for c in range(k):
        _lib.check(lib.hann_pqivf_list_codes(idx.handle, c, None, ctypes.byref(n)))
        if n.value < 0:
            continue
        codes = np.zeros((max(n.value, 1), m), np.int32)
        _lib.check(lib.hann_pqivf_list_codes(idx.handle, c, _lib.i32ptr(codes), ctypes.byref(n)))
        res[c] = codes[:n.value]
    return res
