"""The reference's on-disk format (Go encoding/gob, pqivf/index.go:968-1101)
through the C-ABI, on the CPU (decoding and encoding are host work).

Pins:
  * tests/golden/index_v1.gob is a file Go hann itself wrote
    (pqivf/testdata/index_v1.gob). Loading it gives the state that the
    independent reader tests/golden/gob2json.py decodes from it
    (index_v1.json), and saving that state again in the file's layout gives
    the file back byte for byte: Go's own encoding of every value type,
    type definition and message frame the format uses.
  * The current layout (with FormatVersion and NumCandidateClusters) is
    read back by the independent reader, value for value.
  * The reference's decode tests (pqivf/index_test.go:741-793, 944-978):
    garbage, a newer format version, a legacy file without PendingVectors,
    invalid parameters; gob streams for them are written here by a small
    encoder of structs of int fields, written from the gob wire format.
"""
import io
import os
import sys

import numpy as np
import pytest

from conftest import GOLDEN, bits_to_f32
from hann_amd import HannError, _lib, pqivf

sys.path.insert(0, GOLDEN)
import gob2json  # noqa: E402


def gob_file():
    with open(os.path.join(GOLDEN, "index_v1.gob"), "rb") as fh:
        return fh.read()


def decode_stream(data):
    """The independent reader: every value of one gob stream."""
    return gob2json.Decoder().stream(data)


# -- a minimal gob encoder for structs of int (or string) fields -----------
def g_uint(x):
    if x < 128:
        return bytes([x])
    b = x.to_bytes((x.bit_length() + 7) // 8, "big")
    return bytes([256 - len(b)]) + b


def g_int(v):
    return g_uint((~v << 1) | 1 if v < 0 else v << 1)


def g_str(s):
    b = s.encode()
    return g_uint(len(b)) + b


def g_msg(payload):
    return g_uint(len(payload)) + payload


def struct_stream(fields):
    """A gob stream of one struct value: fields = [(name, value)], int or str
    values; zero values are omitted as Go's encoder omits them."""
    tid = 65
    wire = g_uint(3) + g_uint(1) + g_uint(2) + g_int(tid) + g_uint(0)   # StructT{CommonType{Id}}
    wire += g_uint(1) + g_uint(len(fields))
    for name, v in fields:
        wire += g_uint(1) + g_str(name) + g_uint(1) + g_int(2 if isinstance(v, int) else 6) + g_uint(0)
    wire += g_uint(0) + g_uint(0)
    out = g_msg(g_int(-tid) + wire)
    body, last = b"", -1
    for i, (name, v) in enumerate(fields):
        if v in (0, ""):
            continue
        body += g_uint(i - last) + (g_int(v) if isinstance(v, int) else g_str(v))
        last = i
    return out + g_msg(g_int(tid) + body + g_uint(0))


def loaded_fixture(*opts):
    idx = pqivf.New(8, pqivf.WithCoarseK(3), pqivf.WithNumSubquantizers(2), pqivf.WithPQK(4),
                    pqivf.WithKMeansIters(5), *opts)
    idx.Load(io.BytesIO(gob_file()))
    return idx


def test_encoder_helper_matches_gob_reader():
    vals = decode_stream(struct_stream([("Dimension", 6), ("CoarseK", 0), ("Name", "x"), ("Neg", -3)]))
    assert vals == [{"Dimension": 6, "Name": "x", "Neg": -3}]


def test_reference_file_loads_to_the_decoded_state(gob_fixture):
    """pqivf/golden_test.go:95-146: the file's index holds 29 entries; every
    centroid, codeword, list, code and vector equals the independent decode."""
    g = gob_fixture
    idx = loaded_fixture()
    assert idx.Stats().Count == 29
    assert idx.Stats().Dimension == 8
    k, m, p, d = idx.shape()
    assert (k, m, p, d) == (3, 2, 4, 8)
    assert idx.trained()
    C = pqivf.CoarseCentroids(idx)
    assert np.array_equal(C.view(np.uint32), np.asarray(g["CoarseCentroidsBits"], np.uint32))
    cb, cbl = pqivf.Codebooks(idx)
    want_cb = np.asarray(g["CodebooksBits"], np.uint32)
    assert list(cbl) == [want_cb.shape[1]] * 2
    assert np.array_equal(cb.view(np.uint32), want_cb)
    ids = pqivf.InvertedIDs(idx)
    codes = pqivf.InvertedCodes(idx)
    assert sorted(ids) == sorted(int(c) for c in g["InvertedLists"])
    for c, entries in g["InvertedLists"].items():
        assert ids[int(c)] == [e["ID"] for e in entries]
        assert codes[int(c)].tolist() == [e["Codes"] for e in entries]


def test_reference_file_reencodes_byte_for_byte():
    """Saving the loaded state in the file's layout (format 0: no
    FormatVersion, no NumCandidateClusters) reproduces Go's bytes."""
    raw = gob_file()
    idx = loaded_fixture()
    out = io.BytesIO()
    idx.Save(out, format_version=0)
    assert out.getvalue() == raw
    inner = decode_stream(raw)[0]
    assert idx.GobEncode(format_version=0) == inner


def test_current_layout_reads_back_with_the_independent_reader(gob_fixture):
    g = gob_fixture
    idx = loaded_fixture(pqivf.WithCandidateClusters(2))
    out = io.BytesIO()
    idx.Save(out)
    outer = decode_stream(out.getvalue())
    assert len(outer) == 1 and isinstance(outer[0], bytes)
    ser = decode_stream(outer[0])[0]
    assert ser["FormatVersion"] == 1
    # the legacy file carries no probe count: the configured one is kept
    assert ser["NumCandidateClusters"] == 2
    for key in ("Dimension", "CoarseK", "NumSubquantizers", "PqK", "KMeansIters", "Trained",
                "AllowBruteForceFallback"):
        assert ser[key] == g[key], key
    C = np.asarray(ser["CoarseCentroids"], np.float32)
    assert np.array_equal(C, bits_to_f32(g["CoarseCentroidsBits"]))
    assert ser["ClusterCounts"] == {int(c): len(e) for c, e in g["InvertedLists"].items()}
    for c, entries in g["InvertedLists"].items():
        got = ser["InvertedLists"][int(c)]
        assert [e.get("ID", 0) for e in got] == [e["ID"] for e in entries]
        assert [e.get("Codes", []) for e in got] == [e["Codes"] for e in entries]
        assert [e.get("Cluster", 0) for e in got] == [int(c)] * len(entries)
        assert np.array_equal(np.asarray([e["Vector"] for e in got], np.float32),
                              np.asarray([e["Vector"] for e in entries], np.float32))
    assert ser["PendingVectors"] == {}


def test_round_trip_after_delete_and_with_pending():
    """A mutated index saves and loads back to the same bytes; a deleted
    entry leaves its list and its count."""
    idx = loaded_fixture()
    idx.Delete(21)
    data = idx.GobEncode()
    ser = decode_stream(data)[0]
    assert 21 not in [e.get("ID", 0) for e in ser["InvertedLists"][0]]
    assert ser["ClusterCounts"][0] == 9
    other = pqivf.New(8)
    other.GobDecode(data)
    assert other.Stats().Count == 28
    assert other.GobEncode() == data
    # an untrained index: pending vectors only, written in ascending id order
    rng = np.random.default_rng(3)
    un = pqivf.New(6, pqivf.WithNumSubquantizers(3))
    vecs = rng.standard_normal((5, 6)).astype(np.float32)
    for i in (9, 2, 7, 0, 4):
        un.Add(i, vecs[i % 5])
    data = un.GobEncode()
    ser = decode_stream(data)[0]
    assert list(ser["PendingVectors"]) == [0, 2, 4, 7, 9]
    assert np.array_equal(np.asarray(ser["PendingVectors"][7], np.float32), vecs[2])
    assert "CoarseCentroids" not in ser and "Codebooks" not in ser and "Trained" not in ser
    back = pqivf.New(6)
    back.GobDecode(data)
    assert back.Stats().Count == 5
    assert back.GobEncode() == data


def test_decode_errors():
    """pqivf/index_test.go:741-764"""
    idx = pqivf.New(4, pqivf.WithCoarseK(2), pqivf.WithNumSubquantizers(2), pqivf.WithPQK(2),
                    pqivf.WithKMeansIters(5))
    with pytest.raises(HannError) as e:
        idx.GobDecode(b"not a gob stream")
    assert e.value.code == _lib.HANN_ERR_FORMAT
    with pytest.raises(HannError, match="format version"):
        idx.GobDecode(struct_stream([("FormatVersion", 999)]))
    # truncations of a real file fail cleanly
    raw = gob_file()
    for cut in (1, 10, 100, 1000, len(raw) - 1):
        with pytest.raises(HannError) as e:
            pqivf.New(8).Load(io.BytesIO(raw[:cut]))
        assert e.value.code == _lib.HANN_ERR_FORMAT


def test_decode_legacy_without_pending_vectors():
    """pqivf/index_test.go:766-793: a file from before PendingVectors loads
    into a usable index (unknown fields are skipped)."""
    idx = pqivf.New(4)
    idx.GobDecode(struct_stream([("Dimension", 4), ("CoarseK", 2), ("NumSubquantizers", 2), ("PqK", 4),
                                 ("KMeansIters", 5), ("Comment", "old file")]))
    idx.Add(1, [1, 2, 3, 4])
    assert idx.Stats().Count == 1


@pytest.mark.parametrize("fields,msg", [
    ([("Dimension", 0)], "invalid dimension 0"),
    ([("Dimension", 6), ("NumSubquantizers", 0)], "invalid subquantizer count 0 for dimension 6"),
    ([("Dimension", 6), ("NumSubquantizers", 4)], "invalid subquantizer count 4 for dimension 6"),
    ([("Dimension", 6), ("NumSubquantizers", 2), ("CoarseK", 0)], "invalid training parameters"),
])
def test_decode_rejects_invalid_parameters(fields, msg):
    """pqivf/index_test.go:944-978"""
    idx = pqivf.New(4)
    with pytest.raises(HannError, match=msg) as e:
        idx.GobDecode(struct_stream(fields))
    assert e.value.code == _lib.HANN_ERR_FORMAT
    assert idx.Stats().Dimension == 4   # the index is unchanged


def test_decode_rejects_what_the_device_index_cannot_hold():
    """A field of the wrong wire type, and a file whose vectors do not have
    the dimension, fail with a format error instead of loading."""
    idx = pqivf.New(4)
    with pytest.raises(HannError, match="type mismatch for field Dimension"):
        idx.GobDecode(struct_stream([("Dimension", "four")]))
    good = bytearray(loaded_fixture().GobEncode())
    # the last message is the value: type id 66 (ff 84), then field 0
    # (delta 1) Dimension 8 (zig-zag 16); declare dimension 4 instead
    r = gob2json.Reader(bytes(good))
    last = 0
    while r.p < len(good):
        n = r.uint()
        last = r.p
        r.p += n
    assert bytes(good[last:last + 4]) == bytes([0xff, 0x84, 0x01, 0x10])
    good[last + 3] = 0x08
    with pytest.raises(HannError, match="coarse centroid of 8 values in an index of dimension 4") as e:
        pqivf.New(8).GobDecode(bytes(good))
    assert e.value.code == _lib.HANN_ERR_FORMAT


def _load_raw(raw):
    idx = pqivf.New(8, pqivf.WithCoarseK(3), pqivf.WithNumSubquantizers(2), pqivf.WithPQK(4),
                    pqivf.WithKMeansIters(5))
    idx.Load(io.BytesIO(raw))
    return idx


def _codes_of(idx):
    return {c: v.tolist() for c, v in pqivf.InvertedCodes(idx).items()}


def test_uncodable_entry_loads_without_codes():
    """Search decodes an entry's codes only when each lies in its codebook
    (pqivf/index.go:906-928, decodePQCode 678-690); otherwise it ranks the
    entry by its vector. A file whose entry holds a code past its codebook
    loads, with that entry kept without codes and everything else as in the
    file; a negative code, which the reference indexes out of range with, is
    rejected."""
    from conftest import gob_code_byte
    raw = gob_file()
    p, c, r = gob_code_byte(raw, _load_raw, _codes_of)
    b = bytearray(raw)
    b[p] = 0x0A
    idx = _load_raw(bytes(b))
    assert idx.Stats().Count == 29
    assert pqivf.InvertedIDs(idx) == pqivf.InvertedIDs(_load_raw(raw))
    b[p] = 0x01   # gob's int -1
    with pytest.raises(HannError, match="negative code"):
        _load_raw(bytes(b))
