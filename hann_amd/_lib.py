"""ctypes binding of libhann_amd.so (the C-ABI declared in include/hann_amd.h).

The library is built in-tree by `python -m hann_amd.build`. Loading fails
loudly when it is missing: there is no CPU fallback.
"""
import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libhann_amd.so")

HANN_OK = 0
HANN_ERR_INVALID_ARG = 1
HANN_ERR_NOT_TRAINED = 2
HANN_ERR_EMPTY = 3
HANN_ERR_DUPLICATE_ID = 4
HANN_ERR_NOT_FOUND = 5
HANN_ERR_OOM = 6
HANN_ERR_DEVICE = 7
HANN_ERR_FORMAT = 8
HANN_ERR_UNSUPPORTED = 9

METRIC_SQEUCLIDEAN, METRIC_EUCLIDEAN, METRIC_MANHATTAN, METRIC_COSINE = 0, 1, 2, 3
PREC_FP32, PREC_BF16 = 0, 1
SCAN_EXACT, SCAN_ADC_F32, SCAN_ADC_F16, SCAN_LIST_BF16, SCAN_LIST_BF16X3, SCAN_LIST_FP8 = 0, 1, 2, 3, 4, 5


class HannError(Exception):
    """A failed library call: `code` is the hann_status, str() the message."""

    def __init__(self, code, msg):
        super().__init__(msg)
        self.code = code


c_float_p = ctypes.POINTER(ctypes.c_float)
c_double_p = ctypes.POINTER(ctypes.c_double)
c_int32_p = ctypes.POINTER(ctypes.c_int32)
c_int64_p = ctypes.POINTER(ctypes.c_int64)
c_size_p = ctypes.POINTER(ctypes.c_size_t)
vp = ctypes.c_void_p
sz = ctypes.c_size_t
i32 = ctypes.c_int32
i64 = ctypes.c_int64
u64 = ctypes.c_uint64
cint = ctypes.c_int


class PQIVFOptions(ctypes.Structure):
    _fields_ = [
        ("coarse_k", i64), ("num_subquantizers", i64), ("pq_k", i64), ("kmeans_iters", i64),
        ("candidate_clusters", i64), ("bruteforce_fallback", i32), ("explicit_mask", ctypes.c_uint32),
        ("seed", u64), ("train_precision", i32), ("kmeans_init", i32), ("train_sample", i64),
        ("pq_train_sample", i64),
    ]


class SearchParams(ctypes.Structure):
    _fields_ = [("scan_mode", i32), ("coarse_precision", i32), ("rerank", i32), ("fused", i32), ("list_config", i32)]


class ServerOptions(ctypes.Structure):
    _fields_ = [("max_batch", i32), ("slots", i32), ("hbm_budget", i64), ("params", SearchParams)]


class ServerCounters(ctypes.Structure):
    _fields_ = [("completed_queries", i64), ("completed_requests", i64), ("batches", i64), ("mutations", i64),
                ("queue_depth", i64), ("in_flight", i64), ("budget_in_use", i64), ("budget_peak", i64),
                ("held_peak", i64), ("scratch_over_bound", i64)]


class IndexStats(ctypes.Structure):
    _fields_ = [("count", i64), ("dimension", i64), ("fallback_searches", i64),
                ("distance", ctypes.c_char * 32)]


# name -> (restype, argtypes)
SIGNATURES = {
    "hann_last_error": (ctypes.c_char_p, []),
    "hann_version": (ctypes.c_char_p, []),
    "hann_init": (cint, [cint]),
    "hann_cpu_init": (None, [cint]),
    "simd_euclidean": (ctypes.c_float, [c_float_p, c_float_p, sz]),
    "simd_squared_euclidean": (ctypes.c_float, [c_float_p, c_float_p, sz]),
    "simd_manhattan": (ctypes.c_float, [c_float_p, c_float_p, sz]),
    "simd_cosine_distance": (ctypes.c_float, [c_float_p, c_float_p, sz]),
    "simd_euclidean_batch": (None, [c_float_p, c_float_p, sz, sz, c_double_p]),
    "simd_squared_euclidean_batch": (None, [c_float_p, c_float_p, sz, sz, c_double_p]),
    "simd_manhattan_batch": (None, [c_float_p, c_float_p, sz, sz, c_double_p]),
    "simd_cosine_distance_batch": (None, [c_float_p, c_float_p, sz, sz, c_double_p]),
    "simd_normalize": (None, [c_float_p, sz]),
    "hann_normalize_batch": (cint, [c_float_p, sz, sz]),
    "hann_normalize_batch_dev": (cint, [vp, sz, sz, vp]),
    "hann_distance_batch": (cint, [cint, cint, c_float_p, sz, c_float_p, sz, sz, c_double_p]),
    "hann_distance_batch_dev": (cint, [cint, cint, vp, sz, vp, sz, sz, vp, vp]),
    "hann_batch_topk": (cint, [cint, c_float_p, sz, c_float_p, sz, sz, sz, cint, sz, c_int64_p,
                               c_double_p]),
    "hann_batch_topk_dev": (cint, [cint, vp, sz, vp, sz, sz, sz, cint, sz, vp, vp, vp]),
    "hann_batch_topk_noalloc_dev": (cint, [cint, vp, sz, vp, sz, sz, sz, vp, vp, vp]),
    "hann_assign_points": (cint, [c_float_p, sz, sz, c_float_p, sz, cint, c_int32_p]),
    "hann_assign_points_dev": (cint, [vp, sz, sz, vp, sz, cint, vp, vp]),
    "hann_kmeans": (cint, [c_float_p, sz, sz, sz, cint, u64, cint, cint, c_float_p, c_size_p]),
    "hann_pqivf_default_options": (None, [ctypes.POINTER(PQIVFOptions)]),
    "hann_pqivf_new": (cint, [i64, ctypes.POINTER(PQIVFOptions), ctypes.POINTER(vp)]),
    "hann_pqivf_free": (None, [vp]),
    "hann_pqivf_add": (cint, [vp, i64, c_float_p, sz]),
    "hann_pqivf_bulk_add": (cint, [vp, c_int64_p, c_float_p, c_size_p, sz]),
    "hann_pqivf_delete": (cint, [vp, i64]),
    "hann_pqivf_bulk_delete": (cint, [vp, c_int64_p, sz]),
    "hann_pqivf_update": (cint, [vp, i64, c_float_p, sz]),
    "hann_pqivf_bulk_update": (cint, [vp, c_int64_p, c_float_p, c_size_p, sz]),
    "hann_pqivf_train": (cint, [vp]),
    "hann_search_default_params": (None, [ctypes.POINTER(SearchParams)]),
    "hann_pqivf_search": (cint, [vp, c_float_p, sz, i64, c_int64_p, c_double_p, c_int64_p]),
    "hann_pqivf_search_batch": (cint, [vp, c_float_p, sz, c_int32_p, c_int32_p,
                                       ctypes.POINTER(SearchParams), c_int64_p, c_double_p,
                                       c_int32_p]),
    "hann_pqivf_search_batch_dev": (cint, [vp, vp, sz, vp, vp, i32, i32, ctypes.POINTER(SearchParams),
                                           vp, vp, vp, vp]),
    "hann_pqivf_search_graph_create": (cint, [vp, vp, sz, vp, vp, i32, i32, ctypes.POINTER(SearchParams),
                                              vp, vp, vp, vp, ctypes.POINTER(vp)]),
    "hann_search_graph_launch": (cint, [vp]),
    "hann_search_graph_counts": (cint, [vp, c_int64_p, c_int64_p, c_int64_p]),
    "hann_search_graph_destroy": (None, [vp]),
    "hann_pqivf_stats": (cint, [vp, ctypes.POINTER(IndexStats)]),
    "hann_pqivf_save": (cint, [vp, ctypes.POINTER(vp), c_size_p]),
    "hann_pqivf_load": (cint, [vp, vp, sz]),
    "hann_pqivf_gob_encode": (cint, [vp, cint, ctypes.POINTER(vp), c_size_p]),
    "hann_pqivf_gob_decode": (cint, [vp, vp, sz]),
    "hann_pqivf_save_gob": (cint, [vp, cint, ctypes.POINTER(vp), c_size_p]),
    "hann_pqivf_load_gob": (cint, [vp, vp, sz]),
    "hann_free_buffer": (None, [vp]),
    "hann_pqivf_profile": (cint, [vp, cint]),
    "hann_pqivf_stage_times": (cint, [vp, c_double_p, c_double_p, c_double_p, c_int64_p]),
    "hann_pqivf_scan_kernel_time": (cint, [vp, c_double_p]),
    "hann_pqivf_scan_kernel_times": (cint, [vp, c_double_p, c_double_p]),
    "hann_pqivf_last_tune": (cint, [vp, ctypes.c_char_p, ctypes.c_size_t]),
    "hann_pqivf_last_scan": (cint, [vp, c_int32_p, c_int32_p]),
    "hann_pqivf_export_lists": (cint, [vp, c_int64_p, c_int64_p, c_int32_p, c_float_p]),
    "hann_server_default_options": (None, [ctypes.POINTER(ServerOptions)]),
    "hann_server_create": (cint, [vp, ctypes.POINTER(ServerOptions), ctypes.POINTER(vp)]),
    "hann_server_free": (None, [vp]),
    "hann_server_submit": (cint, [vp, c_float_p, sz, c_int32_p, c_int32_p, ctypes.POINTER(u64)]),
    "hann_server_submit_bulk_add": (cint, [vp, c_int64_p, c_float_p, sz, ctypes.POINTER(u64)]),
    "hann_server_submit_bulk_delete": (cint, [vp, c_int64_p, sz, ctypes.POINTER(u64)]),
    "hann_server_poll": (cint, [vp, u64, cint, c_int64_p, c_double_p, c_int32_p, c_double_p]),
    "hann_server_stats": (cint, [vp, ctypes.POINTER(ServerCounters)]),
    "hann_server_replay": (cint, [vp, sz, c_double_p, c_int32_p, c_int64_p, c_int32_p, c_float_p, c_int32_p,
                                  c_int32_p, c_int64_p, c_float_p, c_int64_p, i32, c_int64_p, c_double_p,
                                  c_double_p, c_double_p]),
    "hann_pqivf_is_trained": (cint, [vp]),
    "hann_pqivf_shape": (cint, [vp, c_int64_p, c_int64_p, c_int64_p, c_int64_p]),
    "hann_pqivf_coarse_centroids": (cint, [vp, c_float_p]),
    "hann_pqivf_codebooks": (cint, [vp, c_float_p, c_int64_p]),
    "hann_pqivf_list_ids": (cint, [vp, i64, c_int64_p, c_int64_p]),
    "hann_pqivf_list_codes": (cint, [vp, i64, c_int32_p, c_int64_p]),
    "hann_pqivf_encode_vector": (cint, [vp, c_float_p, sz, i64, c_int32_p]),
}

_lib = None


def load():
    """Load libhann_amd.so once and attach the C signatures."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"{LIB_PATH} is missing: build it with `python -m hann_amd.build` "
                          "(hann_amd has no CPU fallback)")
    lib = ctypes.CDLL(LIB_PATH)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name, None)
        if fn is None:
            continue
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def exported_symbols():
    return list(SIGNATURES)


def check(status):
    if status != HANN_OK:
        raise HannError(status, load().hann_last_error().decode())


def fptr(a):
    return a.ctypes.data_as(c_float_p)


def dptr(a):
    return a.ctypes.data_as(c_double_p)


def i64ptr(a):
    return a.ctypes.data_as(c_int64_p)


def i32ptr(a):
    return a.ctypes.data_as(c_int32_p)


def sizeptr(a):
    return a.ctypes.data_as(c_size_p)
