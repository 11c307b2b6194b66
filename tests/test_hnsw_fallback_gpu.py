"""The GPU batch top-k as the backend of HNSW's exhaustive fallback
(hnsw/index.go:1198-1319), through the C-ABI.

The fallback gathers the k nearest nodes by metric.Rank among the nodes that
are not already graph candidates (nodes taken in ascending id order), then
the caller merges them with the candidates and sorts by (rank, id). The
restatement below follows those lines on the host, with every rank from the
oracle's reference-order kernels; the backend's ids and ranks must equal it
for all four metrics: exactly where ranks are distinct, and by (rank, id)
where they tie (integer data), where the reference's heaps leave the choice
unspecified.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _restated(oracle, metric_rank, q, ids, vecs, k, cands):
    """hnsw/index.go:1212-1306: non-candidate nodes in ascending id order,
    each ranked, the k best kept; ties here by id"""
    keys = np.sort(ids)
    pos = {int(v): i for i, v in enumerate(ids)}
    nodes = [int(i) for i in keys if int(i) not in cands]
    if not nodes:
        return np.zeros(0, np.int64), np.zeros(0)
    flat = np.stack([vecs[pos[i]] for i in nodes])
    r = oracle.batch(metric_rank, q, flat, q.size)
    o = np.lexsort((np.array(nodes), r))[:k]
    return np.array(nodes)[o], r[o]


@pytest.mark.parametrize("metric_name", ["euclidean", "squared_euclidean", "manhattan", "cosine"])
@pytest.mark.parametrize("integer", [False, True])
def test_fallback_backend_equals_restatement(hann, oracle, metric_name, integer):
    from hann_amd import core
    metric, ok = core.MetricByName(metric_name)
    assert ok
    rank_id = {"euclidean": 0, "squared_euclidean": 0, "manhattan": 2, "cosine": 3}[metric_name]
    rng = np.random.default_rng(17 + integer)
    n, d = 3000, 40
    if integer:
        vecs = rng.integers(0, 4, (n, d)).astype(np.float32)
    else:
        vecs = rng.standard_normal((n, d)).astype(np.float32)
    ids = rng.permutation(10 * n)[:n].astype(np.int64)   # sparse, unordered node ids
    for t in range(4):
        q = (rng.integers(0, 4, d) if integer else rng.standard_normal(d)).astype(np.float32)
        cands = set(rng.choice(ids, 7 * t, replace=False).tolist())
        k = [1, 10, 25, 64][t]
        got_i, got_r = core.ExhaustiveFallback(metric, q, ids, vecs, k, cands)
        want_i, want_r = _restated(oracle, rank_id, q, ids, vecs, k, cands)
        assert np.array_equal(got_i, want_i), (metric_name, t)
        assert np.array_equal(got_r, want_r.astype(np.float64)), (metric_name, t)
        assert not (set(got_i.tolist()) & cands)


def test_fallback_backend_every_node_a_candidate(hann):
    """Every node already a candidate: nothing to gather (the reference skips
    the scan, hnsw/index.go:1226-1239)."""
    from hann_amd import core
    vecs = np.eye(5, dtype=np.float32)
    ids = np.arange(5)
    got_i, got_r = core.ExhaustiveFallback(core.Euclidean, vecs[0], ids, vecs, 3, set(range(5)))
    assert got_i.size == 0 and got_r.size == 0
