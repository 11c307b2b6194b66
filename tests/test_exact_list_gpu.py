"""The list-major exact scan of the plain path, on the GPU, through the C-ABI.

HANN_SCAN_EXACT with a rerank depth R > 0 ranks every probed entry by the
reference's decode-and-measure distance (pqivf/index.go:905-936), keeps the
k R best by (distance, storage position) and reranks them on the stored
vectors. Where the shape allows (d in {64, 96, 128}, d / m in {4, 8}, uint8
codes) the list-major exact scan does this; its results must equal the
oracle's or_search_rerank bit for bit (ids, distances, counts), like those
of the per-query exact scan it replaces. Also covered: the brute-force
fallback and its counter, queries whose records overflow their room (both
handed to the per-query exact scan), and a boundary bin of more than 1,024
records in the record select (its radix select over the record value), for
the exact records and the MFMA scan's 16-bit ones.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _index(x, coarse_k, m, pq_k=256, iters=6, seed=3):
    from hann_amd import pqivf as pq
    idx = pq.New(x.shape[1], pq.WithCoarseK(coarse_k), pq.WithNumSubquantizers(m), pq.WithPQK(pq_k),
                 pq.WithKMeansIters(iters), pq.WithSeed(seed))
    idx.bulk_add_arrays(np.arange(x.shape[0], dtype=np.int64), x)
    idx.Train()
    return idx


def _csr(oracle, idx):
    from hann_amd import pqivf
    off, ids, codes, vecs = idx.export_lists(with_vectors=True)
    cb, cbl = pqivf.Codebooks(idx)
    return oracle.CSRIndex(off, ids, codes, vecs, pqivf.CoarseCentroids(idx), cb, cbl)


def _mixture(n, d, comps, seed, spread=4.0):
    rng = np.random.default_rng(seed)
    c = rng.uniform(-spread, spread, (comps, d)).astype(np.float32)
    x = c[rng.integers(0, comps, n)] + rng.standard_normal((n, d)).astype(np.float32)
    return x.astype(np.float32), c


def _same(got, want):
    ids, dist, cnt = got
    wids, wdist, wcnt = want[:3]
    assert np.array_equal(cnt, wcnt)
    for i in range(cnt.size):
        c = int(cnt[i])
        assert np.array_equal(ids[i, :c], wids[i, :c]), i
        assert np.array_equal(dist[i, :c], wdist[i, :c]), i


@pytest.mark.parametrize("d,m", [(64, 8), (64, 16), (96, 24), (96, 12), (128, 16), (128, 32)])
def test_exact_list_scan_equals_oracle(hann, oracle, d, m):
    """Per-query k in {10, 50, 100} and nprobe in {1, 2, 8, 32}, R = 3 and
    R = 1: ids, distances and counts identical to or_search_rerank, and the
    list-major exact scan is what ran."""
    from hann_amd import _lib
    x, c = _mixture(12000, d, 48, seed=d + m)
    idx = _index(x, 64, m)
    ix = _csr(oracle, idx)
    rng = np.random.default_rng(7)
    qs = (c[rng.integers(0, 48, 96)] + rng.standard_normal((96, d))).astype(np.float32)
    ks = rng.choice(np.array([10, 50, 100], np.int32), 96)
    nps = rng.choice(np.array([1, 2, 8, 32], np.int32), 96)
    for R in (3, 1):
        got = idx.search_batch(qs, ks, nprobe=nps, scan_mode=_lib.SCAN_EXACT, rerank=R)
        assert idx.last_scan() == (_lib.SCAN_EXACT, 2)
        _same(got, oracle.search_rerank_many(ix, qs, ks, nps, R, nthreads=8))


@pytest.mark.parametrize("d,m", [(40, 8), (45, 9), (130, 10), (20, 4), (200, 8)])
def test_plain_path_rerank_any_width_equals_oracle(hann, oracle, d, m):
    """Widths the list-major exact scan does not take (the per-query exact
    scan ranks the candidates), through the rerank's reference-order
    distance: d = 40 exercises the 8-wide cleanup slots, 45 the cleanup and
    a 5-element scalar tail, 130 four blocks and a tail, 20 no full block,
    200 six blocks and a cleanup. Ids, distances and counts identical to
    or_search_rerank."""
    from hann_amd import _lib
    x, c = _mixture(6000, d, 24, seed=d)
    idx = _index(x, 32, m, pq_k=64, iters=4)
    ix = _csr(oracle, idx)
    rng = np.random.default_rng(d)
    qs = (c[rng.integers(0, 24, 64)] + rng.standard_normal((64, d))).astype(np.float32)
    ks = rng.choice(np.array([1, 10, 50], np.int32), 64)
    nps = rng.choice(np.array([1, 2, 8], np.int32), 64)
    for R in (3, 1):
        got = idx.search_batch(qs, ks, nprobe=nps, scan_mode=_lib.SCAN_EXACT, rerank=R)
        _same(got, oracle.search_rerank_many(ix, qs, ks, nps, R, nthreads=8))


def test_exact_list_scan_bf16_coarse_identical(hann, oracle):
    """The certified BF16 coarse step under the exact list scan: identical to
    the FP32 coarse step and to the oracle."""
    from hann_amd import _lib
    x, c = _mixture(20000, 96, 64, seed=11)
    idx = _index(x, 128, 24)
    ix = _csr(oracle, idx)
    rng = np.random.default_rng(8)
    qs = (c[rng.integers(0, 64, 200)] + rng.standard_normal((200, 96))).astype(np.float32)
    ks = rng.choice(np.array([10, 50, 100], np.int32), 200)
    nps = rng.choice(np.array([8, 16, 32, 64, 128], np.int32), 200)
    a = idx.search_batch(qs, ks, nprobe=nps, scan_mode=_lib.SCAN_EXACT, rerank=3, coarse_precision=_lib.PREC_FP32)
    b = idx.search_batch(qs, ks, nprobe=nps, scan_mode=_lib.SCAN_EXACT, rerank=3, coarse_precision=_lib.PREC_BF16)
    assert idx.last_scan() == (_lib.SCAN_EXACT, 2)
    for u, v in zip(a, b):
        assert np.array_equal(u, v)
    _same(a, oracle.search_rerank_many(ix, qs, ks, nps, 3, nthreads=8))


@pytest.mark.parametrize("path", ["exact", "table", "bf16"])
def test_fallback_counter_counts_each_search_once(hann, oracle, path):
    """A query whose probed lists hold fewer than k entries takes the
    reference's brute-force fallback (every list), through the per-query
    exact scan (plain path) or the table scan; the counter counts it once
    per search, also on a batch shape's first search, where the autotuner
    runs the scan once per candidate. The plain path's results equal the
    oracle's."""
    from hann_amd import _lib
    rng = np.random.default_rng(3)
    x = ((rng.standard_normal((4, 64)) * 20)[np.arange(400) % 4] + rng.standard_normal((400, 64))).astype(np.float32)
    idx = _index(x, 4, 16, pq_k=16, iters=5)
    ix = _csr(oracle, idx)
    qs = x[:6] + np.float32(0.01)
    ks = np.array([150, 10, 150, 10, 10, 10], np.int32)
    kw = {"exact": dict(scan_mode=_lib.SCAN_EXACT, rerank=2),
          "table": dict(scan_mode=_lib.SCAN_ADC_F32, rerank=2, fused=1),
          "bf16": dict(scan_mode=_lib.SCAN_LIST_BF16, rerank=2, fused=1)}[path]
    for rep in range(2):
        before = idx.Stats().FallbackSearches
        got = idx.search_batch(qs, ks, nprobe=1, **kw)
        assert idx.Stats().FallbackSearches == before + 2, rep
        if path == "exact":
            assert idx.last_scan() == (_lib.SCAN_EXACT, 2)
            _same(got, oracle.search_rerank_many(ix, qs, ks, np.ones(6, np.int32), 2, nthreads=8))


def test_exact_list_scan_record_overflow(hann, oracle):
    """Thousands of identical vectors share one distance: their records
    overflow the query's room, and the per-query exact scan ranks the query
    instead, with results identical to the oracle (ties by position, then
    the rerank's (distance, id))."""
    from hann_amd import _lib
    rng = np.random.default_rng(5)
    x = rng.standard_normal((3000, 64)).astype(np.float32)
    dup = np.repeat(rng.standard_normal((1, 64)).astype(np.float32), 6000, axis=0)
    data = np.concatenate([x, dup])
    idx = _index(data, 8, 16, pq_k=64, iters=3, seed=2)
    ix = _csr(oracle, idx)
    qs = dup[:8] + 0.01 * rng.standard_normal((8, 64)).astype(np.float32)
    got = idx.search_batch(qs, 10, nprobe=2, scan_mode=_lib.SCAN_EXACT, rerank=4)
    assert idx.last_scan() == (_lib.SCAN_EXACT, 2)
    _same(got, oracle.search_rerank_many(ix, qs, np.full(8, 10, np.int32), np.full(8, 2, np.int32), 4, nthreads=8))


@pytest.mark.parametrize("mode", ["exact", "bf16"])
def test_record_select_radix_boundary(hann, oracle, mode):
    """1,100 identical vectors in one list, k = 50, R = 3: all 1,100 records
    pass the bound and fit the room (8 k R = 1,200), and they fill one bin
    of the select's histogram beyond its 1,024-record room, so the select
    takes its radix select over the record value: the 150 smallest
    positions. Exact mode equals the oracle; the BF16 list scan equals the
    FP32 table scan (whose ties also go by position)."""
    from hann_amd import _lib
    rng = np.random.default_rng(9)
    far = (rng.standard_normal((5000, 64)) + 30.0).astype(np.float32)
    v = rng.standard_normal((1, 64)).astype(np.float32)
    data = np.concatenate([far[:2500], np.repeat(v, 1100, axis=0), far[2500:]])
    idx = _index(data, 8, 16, pq_k=64, iters=4, seed=4)
    qs = v + 0.01 * rng.standard_normal((4, 64)).astype(np.float32)
    if mode == "exact":
        got = idx.search_batch(qs, 50, nprobe=1, scan_mode=_lib.SCAN_EXACT, rerank=3)
        assert idx.last_scan() == (_lib.SCAN_EXACT, 2)
        ix = _csr(oracle, idx)
        _same(got, oracle.search_rerank_many(ix, qs, np.full(4, 50, np.int32), np.ones(4, np.int32), 3, nthreads=8))
        # the selection kept the smallest positions of the tied entries
        assert set(got[0][0].tolist()) <= set(range(2500, 3600))
    else:
        a = idx.search_batch(qs, 50, nprobe=1, scan_mode=_lib.SCAN_ADC_F32, rerank=3, fused=1)
        b = idx.search_batch(qs, 50, nprobe=1, scan_mode=_lib.SCAN_LIST_BF16, rerank=3, fused=1)
        assert idx.last_scan()[0] == _lib.SCAN_LIST_BF16
        for u, w in zip(a, b):
            assert np.array_equal(u, w)


def test_every_tuner_candidate_gives_identical_results(hann, oracle):
    """The autotuner may pick any of its candidates (work items of other
    chunk sizes and query groups, other partition counts): each gives the
    same results — the exact scan's equal the oracle's, and the BF16 scan's
    are identical across candidates. A search with list_config 0 tunes on
    its first batch of a shape and reports the choice."""
    from hann_amd import _lib
    x, c = _mixture(20000, 96, 64, seed=21)
    idx = _index(x, 128, 24)
    ix = _csr(oracle, idx)
    rng = np.random.default_rng(22)
    qs = (c[rng.integers(0, 64, 160)] + rng.standard_normal((160, 96))).astype(np.float32)
    ks = rng.choice(np.array([10, 50, 100], np.int32), 160)
    nps = rng.choice(np.array([8, 16, 32, 64], np.int32), 160)
    want = oracle.search_rerank_many(ix, qs, ks, nps, 3, nthreads=8)
    for cfg in range(1, 6):
        got = idx.search_batch(qs, ks, nprobe=nps, scan_mode=_lib.SCAN_EXACT, rerank=3, list_config=cfg)
        assert idx.last_scan() == (_lib.SCAN_EXACT, 2)
        _same(got, want)
    first = idx.search_batch(qs, ks, nprobe=nps, scan_mode=_lib.SCAN_LIST_BF16, rerank=3, fused=1, list_config=1)
    for cfg in range(2, 5):
        got = idx.search_batch(qs, ks, nprobe=nps, scan_mode=_lib.SCAN_LIST_BF16, rerank=3, fused=1, list_config=cfg)
        assert idx.last_scan()[0] == _lib.SCAN_LIST_BF16
        for u, v in zip(first, got):
            assert np.array_equal(u, v)
    tuned = idx.search_batch(qs, ks, nprobe=nps, scan_mode=_lib.SCAN_LIST_BF16, rerank=3, fused=1)
    assert "list-major MFMA scan tuned" in idx.last_tune()
    for u, v in zip(first, tuned):
        assert np.array_equal(u, v)
    with pytest.raises(Exception, match="list_config"):
        idx.search_batch(qs, ks, nprobe=nps, scan_mode=_lib.SCAN_LIST_BF16, rerank=3, fused=1, list_config=9)
