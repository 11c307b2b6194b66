"""GPU parity of the batch-distance and batch top-k kernels (C-ABI calls).

Bar: bit-exact against the golden vectors written by the reference's own C
kernels and against the oracle; top-k ids and rank values identical to the
oracle's exact (rank, row) ordering."""
import numpy as np
import pytest

from test_oracle import split_pairs

pytestmark = pytest.mark.gpu

NAMES = [("squared_euclidean", 0), ("euclidean", 1), ("manhattan", 2), ("cosine_distance", 3)]


def test_version(hann):
    from hann_amd import _lib
    assert b"gfx950" in _lib.load().hann_version()


@pytest.mark.parametrize("name,mid", NAMES)
def test_drop_in_pair_symbols_bit_exact(hann, distance_golden, name, mid):
    from hann_amd import _lib
    lib = _lib.load()
    fn = {0: lib.simd_squared_euclidean, 1: lib.simd_euclidean, 2: lib.simd_manhattan,
          3: lib.simd_cosine_distance}[mid]
    want = distance_golden[f"pair_{name}_avx2"]
    for i, (d, a, b) in enumerate(split_pairs(distance_golden)):
        got = np.float32(fn(_lib.fptr(a), _lib.fptr(b), d))
        assert got.view(np.uint32) == want[i].view(np.uint32), (name, d, i, got, want[i])


@pytest.mark.parametrize("name,mid", NAMES)
def test_drop_in_batch_symbols_bit_exact(hann, distance_golden, name, mid):
    from hann_amd import _lib
    lib = _lib.load()
    fn = {0: lib.simd_squared_euclidean_batch, 1: lib.simd_euclidean_batch, 2: lib.simd_manhattan_batch,
          3: lib.simd_cosine_distance_batch}[mid]
    for si, (dim, n) in enumerate(distance_golden["batch_shapes"]):
        q = distance_golden[f"batch{si}_q"]
        flat = distance_golden[f"batch{si}_flat"]
        out = np.empty(int(n), np.float64)
        fn(_lib.fptr(q), _lib.fptr(flat), int(dim), int(n), _lib.dptr(out))
        want = distance_golden[f"batch{si}_{name}_avx2"]
        assert np.array_equal(out.view(np.uint64), want.view(np.uint64)), (name, dim, n)


def test_metric_api_semantics(hann):
    """core/metric_batch_test.go:76-100 and core/distance_test.go:125-153."""
    from hann_amd import core
    with pytest.raises(hann.HannError, match="same length"):
        core.Euclidean.Distance([1], [1, 2])
    assert core.Cosine.Distance([], []) == 0.0
    with pytest.raises(hann.HannError, match="query must not be empty"):
        core.Euclidean.DistanceBatch([], [1, 2], np.zeros(1))
    with pytest.raises(hann.HannError, match="not a multiple"):
        core.Euclidean.RankBatch([1, 2], [1, 2, 3], np.zeros(1))
    with pytest.raises(hann.HannError, match="does not match the candidate count"):
        core.Euclidean.DistanceBatch([1, 2], [1, 2, 3, 4], np.zeros(3))
    core.Euclidean.DistanceBatch([1, 2], [], np.zeros(0))
    assert core.Euclidean.Rank([0, 0], [3, 4]) == 25.0
    assert core.Euclidean.FromRank(25.0) == 5.0
    assert core.Cosine.Distance([0] * 8, list(range(1, 9))) == 1.0


@pytest.mark.parametrize("dim", [1, 7, 8, 24, 31, 32, 33, 64, 96, 128, 200])
@pytest.mark.parametrize("mid", [0, 1, 2, 3])
def test_distance_matrix_matches_oracle(hann, oracle, dim, mid):
    from hann_amd import core
    rng = np.random.default_rng(dim * 10 + mid)
    q = (rng.random((5, dim)) * 20 - 10).astype(np.float32)
    base = (rng.random((300, dim)) * 20 - 10).astype(np.float32)
    m = {0: core.SquaredEuclidean, 1: core.Euclidean, 2: core.Manhattan, 3: core.Cosine}[mid]
    got = m.distance_matrix(q, base)
    for i in range(5):
        want = oracle.batch(mid, q[i], base, dim)
        assert np.array_equal(got[i].view(np.uint64), want.view(np.uint64))


@pytest.mark.parametrize("nq,nb,dim,k", [(1, 1, 8, 1), (3, 50, 8, 10), (37, 5000, 128, 100),
                                          (17, 3000, 96, 64), (5, 2000, 100, 33), (20, 70, 16, 100),
                                          (64, 20000, 64, 10), (2, 4096, 784, 5)])
def test_batch_topk_exact_matches_oracle(hann, oracle, nq, nb, dim, k):
    from hann_amd import core
    rng = np.random.default_rng(nq * 1000 + nb)
    q = (rng.random((nq, dim)) * 20 - 10).astype(np.float32)
    base = (rng.random((nb, dim)) * 20 - 10).astype(np.float32)
    ids, dist = core.batch_topk(core.Euclidean, q, base, k)
    wid, wrank = oracle.batch_topk(0, q, base, k, nthreads=8)
    assert np.array_equal(ids, wid)
    kk = min(k, nb)
    assert np.array_equal(dist[:, :kk], np.sqrt(wrank[:, :kk].astype(np.float64)))
    if kk < k:
        assert np.all(ids[:, kk:] == -1) and np.all(np.isinf(dist[:, kk:]))


@pytest.mark.parametrize("nq,nb,dim,k,kind", [(500, 16384, 96, 128, "real"), (300, 20000, 128, 100, "int"),
                                              (64, 8192, 64, 512, "real"), (40, 5000, 64, 10, "dup"),
                                              (400, 65536, 96, 8, "real"), (33, 4000, 128, 1, "int"),
                                              (200, 150000, 128, 10, "int"), (24, 70000, 64, 10, "dup"),
                                              (120, 100000, 96, 100, "real")])
def test_batch_topk_rows_path_matches_oracle(hann, oracle, nq, nb, dim, k, kind):
    """The exact top-k's row path (d in {64, 96, 128}: a sample pass, each
    query's bound, a filter pass, the exact-record select and the rank
    kernel): ids and ranks identical to the oracle's, ties by row. With three
    distinct rows the records of every query overflow their room: up to
    65,536 rows the rank kernel ranks every row for them, past that the tiled
    kernel reruns them."""
    from hann_amd import core
    rng = np.random.default_rng(nq + nb + dim + k)
    if kind == "dup":
        rows = rng.standard_normal((3, dim)).astype(np.float32)
        base = rows[rng.integers(0, 3, nb)]
        q = rows[rng.integers(0, 3, nq)] + np.float32(0.01)
    elif kind == "int":
        lam = rng.uniform(10, 30, dim)
        base = np.minimum(rng.poisson(lam, (nb, dim)), 255).astype(np.float32)
        q = np.minimum(rng.poisson(lam, (nq, dim)), 255).astype(np.float32)
    else:
        cen = rng.standard_normal((64, dim)).astype(np.float32) * 3
        base = (cen[rng.integers(0, 64, nb)] + rng.standard_normal((nb, dim))).astype(np.float32)
        q = (cen[rng.integers(0, 64, nq)] + rng.standard_normal((nq, dim))).astype(np.float32)
    for metric, sq in ((core.SquaredEuclidean, True), (core.Euclidean, False)):
        ids, dist = core.batch_topk(metric, q, base, k)
        wid, wrank = oracle.batch_topk(0, q, base, k, nthreads=8)
        assert np.array_equal(ids, wid)
        want = wrank.astype(np.float64)
        assert np.array_equal(dist, want if sq else np.sqrt(want))


def test_batch_topk_integer_ties_ordered_by_row(hann, oracle):
    """Config-2 style integer data: huge tie groups, ordered by (rank, row)."""
    from hann_amd import core
    rng = np.random.default_rng(4)
    lam = rng.uniform(10, 30, 32)
    base = np.minimum(rng.poisson(lam, (20000, 32)), 255).astype(np.float32)
    q = np.minimum(rng.poisson(lam, (40, 32)), 255).astype(np.float32)
    ids, dist = core.batch_topk(core.SquaredEuclidean, q, base, 100)
    wid, wrank = oracle.batch_topk(0, q, base, 100, nthreads=8)
    assert np.array_equal(ids, wid)
    assert np.array_equal(dist, wrank.astype(np.float64))


@pytest.mark.parametrize("mid", [2, 3])
def test_batch_topk_other_metrics(hann, oracle, mid):
    from hann_amd import core
    rng = np.random.default_rng(mid)
    q = (rng.random((9, 40)) * 20 - 10).astype(np.float32)
    base = (rng.random((3000, 40)) * 20 - 10).astype(np.float32)
    m = core.Manhattan if mid == 2 else core.Cosine
    ids, dist = core.batch_topk(m, q, base, 20)
    wid, wrank = oracle.batch_topk(mid, q, base, 20, nthreads=8)
    assert np.array_equal(ids, wid)
    assert np.array_equal(dist, wrank.astype(np.float64))


@pytest.mark.parametrize("nq,nb,dim,k,kind", [(300, 20000, 128, 100, "int"), (130, 50000, 96, 10, "real"),
                                              (64, 4096, 100, 50, "real"), (5, 300, 32, 10, "real"),
                                              (200, 30000, 784, 20, "real"), (1000, 8000, 64, 1, "int"),
                                              (50, 5000, 64, 10, "dup"), (700, 16384, 96, 128, "real"),
                                              (200, 20001, 96, 50, "real"), (100, 70000, 64, 20, "int"),
                                              (7, 16, 32, 20, "real"), (9, 64, 96, 64, "dup")])
def test_batch_topk_bf16_certified_equals_exact(hann, oracle, nq, nb, dim, k, kind):
    """BF16 MFMA candidate pass + exact FP32 ranking: ids and distances are
    identical to the exact path (the filter threshold is a proven bound).
    Bases of at most 65,536 rows, a multiple of 8, and widths up to 128 take
    the dense form (every pair's key, each query's k-th smallest); the others
    (300 and 20,001 rows, width 784, 70,000 rows) the sampled form. Also k
    above the base's size (every row a candidate) and k equal to it on
    three distinct rows."""
    from hann_amd import _lib, core
    rng = np.random.default_rng(nq + nb + dim)
    if kind == "dup":
        # three distinct rows: every query ties with thousands of rows, the
        # candidate lists overflow and those queries take the exact path
        rows = rng.standard_normal((3, dim)).astype(np.float32)
        base = rows[rng.integers(0, 3, nb)]
        q = rows[rng.integers(0, 3, nq)] + np.float32(0.01)
    elif kind == "int":
        lam = rng.uniform(10, 30, dim)
        base = np.minimum(rng.poisson(lam, (nb, dim)), 255).astype(np.float32)
        q = np.minimum(rng.poisson(lam, (nq, dim)), 255).astype(np.float32)
    else:
        cen = rng.standard_normal((50, dim)).astype(np.float32) * 4
        base = (cen[rng.integers(0, 50, nb)] + rng.standard_normal((nb, dim))).astype(np.float32)
        q = (cen[rng.integers(0, 50, nq)] + rng.standard_normal((nq, dim))).astype(np.float32)
    ids, dist = core.batch_topk(core.Euclidean, q, base, k, precision=_lib.PREC_BF16)
    wid, wrank = oracle.batch_topk(0, q, base, k, nthreads=8)
    assert np.array_equal(ids, wid)
    assert np.array_equal(dist, np.sqrt(wrank.astype(np.float64)))


def test_normalize_bit_exact_vs_reference(hann):
    """simd_normalize, NormalizeVector and NormalizeBatch equal the
    reference's AVX2 bits (core/simd_kernels.inc.h:267-315) on lengths
    1..1,000 with every tail size, zero vectors (left as they are), small
    and large values."""
    import os
    from conftest import GOLDEN
    from hann_amd import core
    g = np.load(os.path.join(GOLDEN, "normalize_golden.npz"))
    off = np.concatenate([[0], np.cumsum(g["lens"])])
    vecs, wants = [], []
    for i in range(len(g["lens"])):
        v = g["inputs"][off[i]:off[i + 1]].copy()
        want = g["out_avx2"][off[i]:off[i + 1]]
        w = v.copy()
        core.NormalizeVector(w)
        assert np.array_equal(w.view(np.uint32), want.view(np.uint32)), v.size
        vecs.append(v)
        wants.append(want)
    core.NormalizeBatch(vecs)
    for v, want in zip(vecs, wants):
        assert np.array_equal(v.view(np.uint32), want.view(np.uint32)), v.size
    empty = np.zeros(0, np.float32)
    core.NormalizeVector(empty)


def test_normalize_batch_device_many_rows(hann, oracle):
    """The device batch form on 4,099 rows of 96 (one wave per row, rows past
    a workgroup's last): every row equals the oracle's."""
    import torch
    from hann_amd import _lib
    rng = np.random.default_rng(8)
    x = (rng.standard_normal((4099, 96)) * 5).astype(np.float32)
    x[17] = 0
    d = torch.from_numpy(x).cuda()
    torch.cuda.synchronize()
    _lib.check(_lib.load().hann_normalize_batch_dev(d.data_ptr(), x.shape[0], x.shape[1], None))
    got = d.cpu().numpy()
    want = np.stack([oracle.normalize(r) for r in x])
    assert np.array_equal(got.view(np.uint32), want.view(np.uint32))
