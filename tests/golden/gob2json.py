"""Decode the reference's pqivf/testdata/index_v1.gob into tests/golden/index_v1.json.

index_v1.gob is a PQIVF index trained and saved by the real reference
(pqivf/golden_test.go:39-71: 30 clustered vectors, d=8, coarseK 3, m 2,
pqK 4, 5 iterations, id 0 deleted, then trained again). Save writes the index
through GobEncode (pqivf/index.go:992-1017), so the file is one gob message
holding a byte slice, and that slice is itself a gob stream carrying
serializedPQIVF (pqivf/index.go:969-984).

This is a small reader for the parts of the gob wire format that file uses
(as documented by Go's encoding/gob package): length-prefixed messages, type
definitions through wireType, structs as (field delta, value) pairs ending in
a zero delta, slices and maps with a count prefix, ints zig-zag encoded,
floats as byte-reversed IEEE-754 float64. float32 values travel as float64
and are narrowed back exactly.

Run from the repository root:
    python tests/golden/gob2json.py [/root/reference/pqivf/testdata/index_v1.gob]
"""
import json
import os
import struct
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
DEFAULT_IN = "/root/reference/pqivf/testdata/index_v1.gob"
OUT = os.path.join(ROOT, "tests", "golden", "index_v1.json")

# predefined gob type ids
T_BOOL, T_INT, T_UINT, T_FLOAT, T_BYTES, T_STRING = 1, 2, 3, 4, 5, 6
# the bootstrap types that describe type definitions
T_WIRE, T_ARRAY, T_COMMON, T_SLICE, T_STRUCT, T_FIELD, T_FIELDS, T_MAP = 16, 17, 18, 19, 20, 21, 22, 23
BOOT = {
    T_WIRE: ("struct", ["ArrayT", "SliceT", "StructT", "MapT", "GobEncoderT", "BinaryMarshalerT",
                        "TextMarshalerT"], [T_ARRAY, T_SLICE, T_STRUCT, T_MAP, 24, 24, 24]),
    T_ARRAY: ("struct", ["CommonType", "Elem", "Len"], [T_COMMON, T_INT, T_INT]),
    T_COMMON: ("struct", ["Name", "Id"], [T_STRING, T_INT]),
    T_SLICE: ("struct", ["CommonType", "Elem"], [T_COMMON, T_INT]),
    T_STRUCT: ("struct", ["CommonType", "Field"], [T_COMMON, T_FIELDS]),
    T_FIELD: ("struct", ["Name", "Id"], [T_STRING, T_INT]),
    T_FIELDS: ("slice", T_FIELD),
    T_MAP: ("struct", ["CommonType", "Key", "Elem"], [T_COMMON, T_INT, T_INT]),
    24: ("struct", ["CommonType"], [T_COMMON]),  # gobEncoderType
}


class Reader:
    def __init__(self, data):
        self.b = data
        self.p = 0

    def byte(self):
        v = self.b[self.p]
        self.p += 1
        return v

    def uint(self):
        c = self.byte()
        if c < 128:
            return c
        n = 256 - c
        v = 0
        for _ in range(n):
            v = (v << 8) | self.byte()
        return v

    def int(self):
        u = self.uint()
        return ~(u >> 1) if u & 1 else u >> 1

    def float(self):
        u = self.uint()
        rev = int.from_bytes(u.to_bytes(8, "big")[::-1], "big")
        return struct.unpack(">d", rev.to_bytes(8, "big"))[0]

    def bytes(self):
        n = self.uint()
        v = self.b[self.p:self.p + n]
        self.p += n
        return bytes(v)


class Decoder:
    def __init__(self):
        self.types = dict(BOOT)

    def value(self, r, tid):
        if tid == T_BOOL:
            return r.uint() != 0
        if tid == T_INT:
            return r.int()
        if tid == T_UINT:
            return r.uint()
        if tid == T_FLOAT:
            return r.float()
        if tid == T_BYTES:
            return r.bytes()
        if tid == T_STRING:
            return r.bytes().decode()
        t = self.types[tid]
        kind = t[0]
        if kind == "struct":
            names, ids = t[1], t[2]
            out = {}
            f = -1
            while True:
                delta = r.uint()
                if delta == 0:
                    return out
                f += delta
                out[names[f]] = self.value(r, ids[f])
        if kind == "slice":
            return [self.value(r, t[1]) for _ in range(r.uint())]
        if kind == "array":
            n = r.uint()
            return [self.value(r, t[1]) for _ in range(n)]
        if kind == "map":
            out = {}
            for _ in range(r.uint()):
                k = self.value(r, t[1])
                out[k] = self.value(r, t[2])
            return out
        if kind == "gobencoder":
            return r.bytes()
        raise ValueError(f"unsupported gob type {tid} ({kind})")

    def define(self, tid, wire):
        if "StructT" in wire:
            fields = wire["StructT"].get("Field", [])
            self.types[tid] = ("struct", [f.get("Name", "") for f in fields], [f.get("Id", 0) for f in fields])
        elif "SliceT" in wire:
            self.types[tid] = ("slice", wire["SliceT"]["Elem"])
        elif "ArrayT" in wire:
            self.types[tid] = ("array", wire["ArrayT"]["Elem"])
        elif "MapT" in wire:
            self.types[tid] = ("map", wire["MapT"]["Key"], wire["MapT"]["Elem"])
        elif "GobEncoderT" in wire:
            self.types[tid] = ("gobencoder",)
        else:
            raise ValueError(f"unsupported wire type {wire}")

    def stream(self, data):
        """Decode every value message of one gob stream."""
        r = Reader(data)
        values = []
        while r.p < len(r.b):
            n = r.uint()
            end = r.p + n
            tid = r.int()
            if tid < 0:
                self.define(-tid, self.value(r, T_WIRE))
            else:
                if self.types.get(tid, ("struct",))[0] != "struct":
                    r.uint()  # a top-level non-struct value is framed as field 0
                values.append(self.value(r, tid))
            r.p = end
        return values


def f32(x):
    return float(np.float32(x))


def main():
    src = sys.argv[1] if len(sys.argv) > 1 else DEFAULT_IN
    with open(src, "rb") as fh:
        raw = fh.read()
    outer = Decoder().stream(raw)
    inner = Decoder().stream(outer[0])
    ser = inner[0]
    lists = {}
    for cluster, entries in sorted(ser.get("InvertedLists", {}).items()):
        lists[str(cluster)] = [
            {"ID": e.get("ID", 0), "Vector": [f32(v) for v in e.get("Vector", [])],
             "Codes": e.get("Codes", []), "Cluster": e.get("Cluster", 0)}
            for e in entries
        ]
    doc = {
        "source": "pqivf/testdata/index_v1.gob (habedi/hann), decoded by tests/golden/gob2json.py",
        "fields_present": sorted(ser.keys()),
        "Dimension": ser.get("Dimension", 0),
        "CoarseK": ser.get("CoarseK", 0),
        "NumSubquantizers": ser.get("NumSubquantizers", 0),
        "PqK": ser.get("PqK", 0),
        "KMeansIters": ser.get("KMeansIters", 0),
        "AllowBruteForceFallback": ser.get("AllowBruteForceFallback", False),
        "Trained": ser.get("Trained", False),
        "FormatVersion": ser.get("FormatVersion", 0),
        "NumCandidateClusters": ser.get("NumCandidateClusters", 0),
        "CoarseCentroids": [[f32(v) for v in c] for c in ser.get("CoarseCentroids", [])],
        "ClusterCounts": {str(k): v for k, v in sorted(ser.get("ClusterCounts", {}).items())},
        "Codebooks": [[[f32(v) for v in row] for row in cb] for cb in ser.get("Codebooks", [])],
        "InvertedLists": lists,
        "PendingVectors": {str(k): [f32(v) for v in vec]
                           for k, vec in sorted(ser.get("PendingVectors", {}).items())},
        # float32 bit patterns, so the fixture is exact whatever JSON float printing does
        "CoarseCentroidsBits": [[int(np.float32(v).view(np.uint32)) for v in c]
                                for c in ser.get("CoarseCentroids", [])],
        "CodebooksBits": [[[int(np.float32(v).view(np.uint32)) for v in row] for row in cb]
                          for cb in ser.get("Codebooks", [])],
        "VectorBits": {str(e["ID"]): [int(np.float32(v).view(np.uint32)) for v in e["Vector"]]
                       for es in lists.values() for e in es},
    }
    with open(OUT, "w") as fh:
        json.dump(doc, fh, indent=1)
    n = sum(len(v) for v in lists.values())
    print(f"wrote {OUT}: {n} entries in {len(lists)} lists, fields {doc['fields_present']}")


if __name__ == "__main__":
    main()
