"""Host logic of the PQIVF runtime through the C-ABI, without GPU compute.

An untrained index keeps its vectors in the pending map and never touches
the device, so the reference's behavioural rules for New, Add, BulkAdd,
Delete, BulkDelete, Update, BulkUpdate, Stats, Save/Load and the Search
validation order are checked here on the CPU (pqivf/index_test.go)."""
import ctypes
import re
import struct
import subprocess

import numpy as np
import pytest

from hann_amd import HannError, _lib, pqivf


def test_library_exports_every_declared_symbol():
    hdr = open(_lib.HERE + "/../include/hann_amd.h").read()
    declared = set(re.findall(r"\b((?:hann|simd)_\w+)\s*\(", hdr))
    out = subprocess.run(["nm", "-D", "--defined-only", _lib.LIB_PATH], capture_output=True, text=True).stdout
    exported = {line.split()[-1] for line in out.splitlines()}
    assert declared and declared <= exported, sorted(declared - exported)
    lib = _lib.load()
    for name in declared:
        assert getattr(lib, name) is not None
    assert set(_lib.SIGNATURES) == declared


def new(dim, *opts):
    return pqivf.New(dim, *opts)


def test_new_validation_messages():
    """pqivf/index_test.go:502-547"""
    with pytest.raises(HannError, match=r"dimension \(0\) must be positive"):
        new(0)
    with pytest.raises(HannError, match=r"coarseK \(0\) must be positive"):
        new(4, pqivf.WithCoarseK(0))
    with pytest.raises(HannError, match=r"coarseK \(-1\) must be positive"):
        new(4, pqivf.WithCoarseK(-1))
    with pytest.raises(HannError, match=r"pqK \(0\) must be positive"):
        new(4, pqivf.WithPQK(0))
    with pytest.raises(HannError, match=r"kMeansIters \(0\) must be positive"):
        new(4, pqivf.WithKMeansIters(0))
    with pytest.raises(HannError, match=r"numSubquantizers \(0\) must be positive"):
        new(4, pqivf.WithNumSubquantizers(0))
    with pytest.raises(HannError, match=r"dimension \(10\) must be divisible by numSubquantizers \(4\)"):
        new(10, pqivf.WithNumSubquantizers(4))


@pytest.mark.parametrize("dim,want", [(16, 8), (12, 4), (6, 2), (7, 1), (64, 8), (96, 8)])
def test_default_num_subquantizers(dim, want):
    """pqivf/index.go:150-157, pqivf/index_test.go:541-548"""
    idx = new(dim)
    assert idx.shape()[1] == want


def test_basic_pending_operations():
    """pqivf/index_test.go:33-66 on an untrained index"""
    idx = new(6)
    idx.Add(1, [1, 2, 3, 4, 5, 6])
    s = idx.Stats()
    assert (s.Count, s.Dimension, s.Distance, s.FallbackSearches) == (1, 6, "euclidean", 0)
    idx.Update(1, [6, 5, 4, 3, 2, 1])
    idx.Delete(1)
    assert idx.Stats().Count == 0


def test_add_errors():
    idx = new(4)
    idx.Add(7, [1, 2, 3, 4])
    with pytest.raises(HannError, match="id 7 already exists in pending vectors"):
        idx.Add(7, [1, 2, 3, 4])
    with pytest.raises(HannError, match="vector dimension 3 does not match index dimension 4"):
        idx.Add(8, [1, 2, 3])
    with pytest.raises(HannError, match="id 9 not found"):
        idx.Delete(9)


def test_bulk_add_validation_is_partial_in_id_order():
    """pqivf/index_test.go:643-677: ids below the failing one stay in."""
    idx = new(4)
    idx.Add(5, [0, 0, 0, 0])
    with pytest.raises(HannError, match="id 5 already exists in pending vectors"):
        idx.BulkAdd({9: [1, 1, 1, 1], 5: [2, 2, 2, 2], 1: [3, 3, 3, 3]})
    assert idx.Stats().Count == 2   # id 1 went in before id 5 failed; 9 did not
    with pytest.raises(HannError, match="vector dimension 3 does not match index dimension 4 for id 3"):
        idx.BulkAdd({3: [1, 2, 3], 2: [1, 1, 1, 1]})
    assert idx.Stats().Count == 3


def test_bulk_delete_skips_unknown_and_keeps_caller_order():
    """pqivf/index_test.go:679-695, 982-999"""
    idx = new(4)
    idx.BulkAdd({1: [1, 1, 1, 1], 2: [2, 2, 2, 2]})
    ids = [2, 99, 1]
    idx.BulkDelete(ids)
    assert ids == [2, 99, 1]
    assert idx.Stats().Count == 0
    idx.BulkDelete([])


def test_update_failure_keeps_entry():
    """pqivf/index_test.go:455-476"""
    idx = new(4)
    idx.Add(1, [1, 2, 3, 4])
    with pytest.raises(HannError, match="does not match index dimension"):
        idx.Update(1, [1, 2])
    assert idx.Stats().Count == 1
    with pytest.raises(HannError, match="id 2 not found"):
        idx.Update(2, [1, 2, 3, 4])
    idx.Update(1, [4, 3, 2, 1])
    assert idx.Stats().Count == 1


def test_bulk_update_validates_before_touching():
    """pqivf/index_test.go:478-500, 697-710"""
    idx = new(4)
    idx.BulkAdd({1: [1, 1, 1, 1], 2: [2, 2, 2, 2]})
    with pytest.raises(HannError, match="invalid vector for id 2: vector dimension 2 does not match"):
        idx.BulkUpdate({1: [0, 0, 0, 0], 2: [0, 0]})
    assert idx.Stats().Count == 2
    with pytest.raises(HannError, match="id 3 not found"):
        idx.BulkUpdate({3: [0, 0, 0, 0]})
    assert idx.Stats().Count == 2


def test_search_validation_order_untrained():
    """pqivf/index.go:854-864, pqivf/index_test.go:361-369"""
    idx = new(4)
    idx.Add(1, [1, 2, 3, 4])
    with pytest.raises(HannError, match="k must be positive"):
        idx.Search([1, 2, 3, 4], 0)
    with pytest.raises(HannError, match="not trained"):
        idx.Search([1, 2, 3, 4], 1)
    with pytest.raises(HannError, match="not trained"):
        idx.Search([1, 2], 1)


def test_train_needs_coarse_k_vectors():
    idx = new(4, pqivf.WithCoarseK(4))
    idx.BulkAdd({i: [i, i, i, i] for i in range(3)})
    with pytest.raises(HannError, match=r"not enough vectors \(3\) to train coarse quantizer with 4 clusters"):
        idx.Train()


def test_save_load_pending_round_trip():
    idx = new(4, pqivf.WithCandidateClusters(5))
    vecs = {i: [i, i + 1, i + 2, i + 3] for i in range(10)}
    idx.BulkAdd(vecs)
    data = idx.save_bytes()
    other = new(4)
    other.load_bytes(data)
    assert other.Stats().Count == 10
    assert not other.trained()
    with pytest.raises(HannError, match="id 3 already exists in pending vectors"):
        other.Add(3, [0, 0, 0, 0])
    other.Add(11, [0, 0, 0, 0])
    assert other.Stats().Count == 11


def _header(data):
    return data[:12], data[12:]


def test_load_rejects_newer_format_and_bad_parameters():
    """pqivf/index_test.go:741-764, 944-978"""
    idx = new(4)
    data = bytearray(idx.save_bytes())
    future = bytearray(data)
    struct.pack_into("<I", future, 8, 2)
    other = new(4)
    with pytest.raises(HannError, match="format version 2, but this build supports up to 1"):
        other.load_bytes(bytes(future))
    # field layout: magic(8) version(4) dim coarseK m pqK iters nprobe (int64 each)
    for off, val, msg in [(12, 0, "invalid dimension 0"), (28, 3, "invalid subquantizer count 3 for dimension 4"),
                          (20, 0, "invalid training parameters: coarseK 0"),
                          (36, -1, "pqK -1"), (44, 0, "kMeansIters 0")]:
        bad = bytearray(data)
        struct.pack_into("<q", bad, off, val)
        with pytest.raises(HannError, match=msg):
            other.load_bytes(bytes(bad))
    with pytest.raises(HannError, match="not a hann PQIVF index file"):
        other.load_bytes(b"garbage-garbage-garbage")
    with pytest.raises(HannError, match="truncated"):
        other.load_bytes(bytes(data[:30]))


def test_load_keeps_configured_nprobe_when_file_has_none():
    """pqivf/index.go:1067-1074: a zero probe count keeps the configured one."""
    idx = new(4)
    data = bytearray(idx.save_bytes())
    struct.pack_into("<q", data, 52, 0)   # nprobe field
    other = new(4, pqivf.WithCandidateClusters(7))
    other.load_bytes(bytes(data))
    again = bytearray(other.save_bytes())
    assert struct.unpack_from("<q", again, 52)[0] == 7


def test_stats_distance_name_and_default_fallback_count():
    idx = new(8)
    s = idx.Stats()
    assert s.Distance == "euclidean" and s.FallbackSearches == 0 and s.Count == 0


def test_seed_from_environment(monkeypatch):
    monkeypatch.setenv("HANN_SEED", "1234")
    assert pqivf.GetSeed() == 1234
    monkeypatch.delenv("HANN_SEED")
    assert pqivf.GetSeed() > 0


def test_null_handle_is_an_error():
    lib = _lib.load()
    rc = lib.hann_pqivf_delete(None, 1)
    assert rc == _lib.HANN_ERR_INVALID_ARG
    assert b"null" in lib.hann_last_error()
    s = _lib.IndexStats()
    assert lib.hann_pqivf_stats(None, ctypes.byref(s)) == _lib.HANN_ERR_INVALID_ARG
