import os
import sys

import numpy as np
import pytest

try:
    # torch first, as bench.py does: its HIP runtime is then the one the
    # library binds to, so that tests can hand torch device buffers and
    # streams to the library (one runtime per process)
    import torch  # noqa: F401
except ImportError:
    pass

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU and libhann_amd.so")


@pytest.fixture(scope="session")
def oracle():
    import oracle as o
    o.lib()
    return o


@pytest.fixture(scope="session")
def hann():
    """The product library, loaded once per test session on a GPU host."""
    import hann_amd
    from hann_amd import _lib
    lib = _lib.load()
    _lib.check(lib.hann_init(0))
    return hann_amd


@pytest.fixture(scope="session")
def distance_golden():
    return np.load(os.path.join(GOLDEN, "distance_golden.npz"))


@pytest.fixture(scope="session")
def gob_fixture():
    import json
    with open(os.path.join(GOLDEN, "index_v1.json")) as fh:
        return json.load(fh)


def bits_to_f32(a):
    return np.array(a, np.uint32).view(np.float32)


def gob_code_byte(raw, load, codes_of):
    """The offset of a byte of tests/golden/index_v1.gob that holds one
    entry's PQ code: codes there are 0..3, each one byte (gob writes int v >= 0
    as the unsigned 2v), so setting such a byte to 0x0A makes that code 5,
    past its 4-row codebook. Found by trial: the first offset whose change
    loads and leaves exactly one entry without codes and every other entry's
    codes as they were. load(bytes) -> index or raises; codes_of(index) ->
    {list: [[codes]...]} with rows of -1 for entries without codes."""
    base = codes_of(load(bytes(raw)))
    for p in range(len(raw)):
        if raw[p] not in (0, 2, 4, 6):
            continue
        b = bytearray(raw)
        b[p] = 0x0A
        try:
            got = codes_of(load(bytes(b)))
        except Exception:
            continue
        diff = [(c, r) for c in base for r in range(len(base[c])) if got[c][r] != base[c][r]]
        if len(diff) == 1:
            c, r = diff[0]
            if all(v == -1 for v in got[c][r]):
                return p, c, r
    raise AssertionError("no code byte found")
