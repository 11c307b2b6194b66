"""Request loop over a device-resident PQIVF index (hann_server_* C-ABI).

Server(index) starts the library's continuous-batching loop; submit() queues
a request of one or more queries with their own k and nprobe and returns a
ticket; result(ticket) waits for it. replay() drives an open-loop arrival
schedule from a native thread and returns per-request latencies.
"""
import ctypes

import numpy as np

from . import _lib


class Server:
    def __init__(self, index, max_batch=4096, slots=2, hbm_budget=1 << 30, scan_mode=_lib.SCAN_ADC_F32, rerank=0,
                 fused=1, coarse_precision=_lib.PREC_FP32):
        self._lib = _lib.load()
        o = _lib.ServerOptions()
        self._lib.hann_server_default_options(ctypes.byref(o))
        o.max_batch, o.slots, o.hbm_budget = max_batch, slots, hbm_budget
        o.params = _lib.SearchParams(scan_mode, coarse_precision, rerank, fused)
        self._index = index   # keeps the index alive
        self._h = ctypes.c_void_p()
        _lib.check(self._lib.hann_server_create(index.handle, ctypes.byref(o), ctypes.byref(self._h)))
        self._k = {}

    def close(self):
        h, self._h = self._h, None
        if h:
            self._lib.hann_server_free(h)

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def submit(self, queries, k, nprobe=None):
        q = np.ascontiguousarray(np.atleast_2d(queries), np.float32)
        nq = q.shape[0]
        ks = np.ascontiguousarray(np.broadcast_to(np.asarray(k, np.int32), (nq,)), np.int32)
        nps = None if nprobe is None else np.ascontiguousarray(np.broadcast_to(np.asarray(nprobe, np.int32), (nq,)),
                                                                np.int32)
        t = ctypes.c_uint64()
        _lib.check(self._lib.hann_server_submit(self._h, _lib.fptr(q), nq, _lib.i32ptr(ks),
                                                _lib.i32ptr(nps) if nps is not None else None, ctypes.byref(t)))
        self._k[t.value] = (nq, int(ks.max()))
        return t.value

    def submit_bulk_add(self, ids, vecs):
        ids = np.ascontiguousarray(ids, np.int64)
        vecs = np.ascontiguousarray(vecs, np.float32)
        t = ctypes.c_uint64()
        _lib.check(self._lib.hann_server_submit_bulk_add(self._h, _lib.i64ptr(ids), _lib.fptr(vecs), ids.size,
                                                         ctypes.byref(t)))
        self._k[t.value] = (0, 1)
        return t.value

    def submit_bulk_delete(self, ids):
        ids = np.ascontiguousarray(ids, np.int64)
        t = ctypes.c_uint64()
        _lib.check(self._lib.hann_server_submit_bulk_delete(self._h, _lib.i64ptr(ids), ids.size, ctypes.byref(t)))
        self._k[t.value] = (0, 1)
        return t.value

    def result(self, ticket):
        """(ids [nq][kmax], dists [nq][kmax], counts [nq], latency_ms)"""
        nq, kmax = self._k.pop(ticket)
        ids = np.full((max(nq, 1), kmax), -1, np.int64)
        dist = np.full((max(nq, 1), kmax), np.inf, np.float64)
        cnt = np.zeros(max(nq, 1), np.int32)
        lat = ctypes.c_double()
        _lib.check(self._lib.hann_server_poll(self._h, ticket, 1, _lib.i64ptr(ids), _lib.dptr(dist), _lib.i32ptr(cnt),
                                              ctypes.byref(lat)))
        return ids[:nq], dist[:nq], cnt[:nq], lat.value

    def stats(self):
        c = _lib.ServerCounters()
        _lib.check(self._lib.hann_server_stats(self._h, ctypes.byref(c)))
        return {f: getattr(c, f) for f, _ in _lib.ServerCounters._fields_}

    def replay(self, at_us, kind, q_off, q_cnt, queries, k, nprobe, add_ids=None, add_vecs=None, del_ids=None,
               kmax_out=10):
        at_us = np.ascontiguousarray(at_us, np.float64)
        kind = np.ascontiguousarray(kind, np.int32)
        q_off = np.ascontiguousarray(q_off, np.int64)
        q_cnt = np.ascontiguousarray(q_cnt, np.int32)
        queries = np.ascontiguousarray(queries, np.float32)
        k = np.ascontiguousarray(k, np.int32)
        nprobe = np.ascontiguousarray(nprobe, np.int32)
        add_ids = np.ascontiguousarray(add_ids if add_ids is not None else np.zeros(1), np.int64)
        add_vecs = np.ascontiguousarray(add_vecs if add_vecs is not None else np.zeros((1, queries.shape[1])),
                                        np.float32)
        del_ids = np.ascontiguousarray(del_ids if del_ids is not None else np.zeros(1), np.int64)
        nreq = at_us.size
        out_ids = np.full((queries.shape[0], kmax_out), -1, np.int64)
        out_dist = np.full((queries.shape[0], kmax_out), np.inf, np.float64)
        lat = np.zeros(nreq, np.float64)
        wall = ctypes.c_double()
        _lib.check(self._lib.hann_server_replay(
            self._h, nreq, _lib.dptr(at_us), _lib.i32ptr(kind), _lib.i64ptr(q_off), _lib.i32ptr(q_cnt),
            _lib.fptr(queries), _lib.i32ptr(k), _lib.i32ptr(nprobe), _lib.i64ptr(add_ids), _lib.fptr(add_vecs),
            _lib.i64ptr(del_ids), kmax_out, _lib.i64ptr(out_ids), _lib.dptr(out_dist), _lib.dptr(lat),
            ctypes.byref(wall)))
        return out_ids, out_dist, lat, wall.value
