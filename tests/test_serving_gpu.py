"""The continuous-batching request loop (hann_server_*) on the GPU.

Bar: every request's result is identical whether it runs alone or mixed
into batches with other requests of other k and nprobe (the search of a
query does not depend on its batch); mutations submitted to the server are
seen by exactly the searches submitted after them."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def index(hann):
    from hann_amd import pqivf, synthetic
    x, _ = synthetic.config1(n=20000)
    idx = pqivf.New(64, pqivf.WithCoarseK(64), pqivf.WithPQK(256), pqivf.WithKMeansIters(5), pqivf.WithSeed(5))
    idx.bulk_add_arrays(np.arange(20000), x)
    idx.Train()
    return idx, x


@pytest.mark.parametrize("mode,rerank", [(2, 0), (1, 2), (0, 0), (3, 2), (4, 2)])
def test_alone_equals_mixed(hann, index, mode, rerank):
    from hann_amd import serving, synthetic
    idx, x = index
    _, qs = synthetic.config1(n=100, q=600)
    rng = np.random.default_rng(1)
    ks = rng.choice([1, 10, 50, 100], 600).astype(np.int32)
    nps = rng.choice([1, 4, 8, 32], 600).astype(np.int32)
    sizes = np.minimum(rng.lognormal(np.log(8), 1.0, 200).astype(int) + 1, 64)
    off = np.concatenate([[0], np.cumsum(sizes)])
    n_req = int(np.searchsorted(off, 600, side="right") - 1)
    srv = serving.Server(idx, max_batch=256, slots=3, scan_mode=mode, rerank=rerank, fused=1 if mode else 0)
    try:
        # alone: one request at a time
        alone = []
        for r in range(n_req):
            sl = slice(off[r], off[r + 1])
            t = srv.submit(qs[sl], ks[sl], nps[sl])
            alone.append(srv.result(t))
        # mixed: everything queued at once, batched together
        tickets = [srv.submit(qs[off[r]:off[r + 1]], ks[off[r]:off[r + 1]], nps[off[r]:off[r + 1]])
                   for r in range(n_req)]
        mixed = [srv.result(t) for t in tickets]
        st = srv.stats()
    finally:
        srv.close()
    for r, (a, b) in enumerate(zip(alone, mixed)):
        sl = slice(off[r], off[r + 1])
        info = f"request {r}: k {ks[sl].tolist()} nprobe {nps[sl].tolist()} counts alone {a[2].tolist()} mixed {b[2].tolist()}"
        assert np.array_equal(a[2], b[2]), info
        assert np.array_equal(a[0], b[0]), info
        assert np.array_equal(a[1], b[1]), info
    assert st["batches"] < 2 * n_req   # the mixed phase did batch requests together
    assert st["scratch_over_bound"] == 0, st   # each charge bounded its slot's device memory
    # and both equal the direct batched search
    ids, dist, cnt = idx.search_batch(qs[:off[n_req]], ks[:off[n_req]], nprobe=nps[:off[n_req]], scan_mode=mode,
                                      rerank=rerank, fused=1 if mode else 0)
    for r in range(n_req):
        kr = alone[r][0].shape[1]
        assert np.array_equal(alone[r][0], ids[off[r]:off[r + 1], :kr])


def test_mutations_are_ordered(hann, index):
    from hann_amd import serving
    idx, x = index
    srv = serving.Server(idx, max_batch=64, slots=2, scan_mode=1, rerank=2)
    try:
        probe = (x[:4] + 0.001).astype(np.float32)
        t0 = srv.submit(probe, 5, 8)
        td = srv.submit_bulk_delete(np.arange(4))
        t1 = srv.submit(probe, 5, 8)
        new = np.arange(100000, 100004)
        ta = srv.submit_bulk_add(new, probe)
        t2 = srv.submit(probe, 1, 8)
        r0, r1, r2 = srv.result(t0), srv.result(t1), srv.result(t2)
        srv.result(td)
        srv.result(ta)
    finally:
        srv.close()
    assert all(i in r0[0][j] for j, i in enumerate(range(4)))
    assert not np.isin(np.arange(4), r1[0]).any()
    assert r2[0][:, 0].tolist() == new.tolist()
    assert idx.Stats().Count == 20000


def test_replay_open_loop(hann, index):
    from hann_amd import serving, synthetic
    idx, x = index
    _, qs = synthetic.config1(n=100, q=2000)
    rng = np.random.default_rng(2)
    sizes = np.minimum(rng.lognormal(np.log(8), 1.0, 300).astype(int) + 1, 1024)
    off = np.concatenate([[0], np.cumsum(sizes)])
    n_req = int(np.searchsorted(off, 2000, side="right") - 1)
    at = np.cumsum(rng.exponential(1e6 / 20000, n_req))   # 20k requests/s
    ks = rng.choice([10, 50, 100], 2000).astype(np.int32)
    nps = rng.choice([8, 16, 32], 2000).astype(np.int32)
    srv = serving.Server(idx, max_batch=1024, slots=2, scan_mode=2, rerank=2)
    try:
        ids, dist, lat, wall = srv.replay(at, np.zeros(n_req), off[:n_req], sizes[:n_req], qs, ks, nps, kmax_out=10)
    finally:
        srv.close()
    want, _, _ = idx.search_batch(qs[:off[n_req]], ks[:off[n_req]], nprobe=nps[:off[n_req]], scan_mode=2, rerank=2,
                                  fused=1)
    assert np.array_equal(ids[:off[n_req]], want[:, :10])
    assert np.all(lat > 0) and wall > 0
