"""Parity at the full sizes of the benchmark configurations, on the GPU.

  C4  10M x 96 L2-normalised mixture, 16,384 lists, PQ 24 x 256, a batch of
      4,096 queries with k in {10, 50, 100} and nprobe in {8, ..., 128}:
      * the plain path (the reference's decode-and-measure distance, R = 3
        exact rerank) equals the oracle's or_search_rerank bit for bit;
      * the list-major BF16 / split-BF16 paths (R = 3): every returned
        distance is the reference-order FP32 one, counts are exact, ids equal
        the oracle's except for entries whose PQ distance lies inside the
        stated band around the query's candidate boundary, and recall@10
        against FP64 exact search is at least 0.95;
      * the certified BF16 coarse step gives the same results as FP32's.
  C5  the served list-major path (BF16 coarse, R = 3) on the same index: each
      request's results are identical alone and in the mix, with BulkAdd and
      BulkDelete requests of 1,000 to 10,000 vectors in between.
  C2  1M x 128 integer data, 10,000 queries, k = 100: certified BF16 top-k
      equals exact FP32 for every query, and both equal the oracle on a
      sample.
  C3  1M x 128, K = 4,096, 20 iterations, k-means++, PQ 16 x 256 on a
      65,536-residual sample: FP32 and BF16 assignment give bit-identical
      indexes; assignments and codes of a 10,000-row sample equal the
      oracle's on the final centroids and codebooks.
  A12 online inserts after Train (BulkAdd of 10,000 vectors): cluster, codes
      and list-tail order equal or_assign / or_encode.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

NCPU = 16


def _csr(oracle, idx, with_vectors=True):
    from hann_amd import pqivf
    off, ids, codes, vecs = idx.export_lists(with_vectors=with_vectors)
    cb, cbl = pqivf.Codebooks(idx)
    return oracle.CSRIndex(off, ids, codes, vecs if with_vectors else np.zeros((1, 1), np.float32),
                           pqivf.CoarseCentroids(idx), cb, cbl)


def _exact_dist_ok(oracle, q, x, ids, dist, cnt):
    """every returned distance is sqrt of the reference-order FP32 rank"""
    for qi in range(q.shape[0]):
        c = int(cnt[qi])
        want = np.sqrt(oracle.batch(0, q[qi], x[ids[qi, :c]], q.shape[1]))
        assert np.array_equal(dist[qi, :c], want), qi


def _fp64_truth(q, x, k=10):
    """FP64 exact top-k: exact FP32 top-(k + 22) on the GPU, re-ranked in FP64"""
    from hann_amd import core
    kk = k + 22
    ids, _ = core.batch_topk(core.SquaredEuclidean, q, x, kk)
    gt = np.empty((q.shape[0], k), np.int64)
    for i in range(q.shape[0]):
        d64 = ((x[ids[i]].astype(np.float64) - q[i].astype(np.float64)) ** 2).sum(1)
        gt[i] = ids[i][np.lexsort((ids[i], d64))[:k]]
    return gt


def _recall10(ids, gt):
    return np.mean([len(set(ids[i, :10].tolist()) & set(gt[i].tolist())) / 10 for i in range(gt.shape[0])])


# ------------------------------------------------------------------ C4 --
@pytest.fixture(scope="module")
def c4(hann):
    from hann_amd import pqivf, synthetic
    c = synthetic.CONFIG4
    x, centers = synthetic.config4()
    idx = pqivf.New(c["dim"], pqivf.WithCoarseK(c["nlist"]), pqivf.WithNumSubquantizers(c["m"]),
                    pqivf.WithPQK(c["pq_k"]), pqivf.WithKMeansIters(c["kmeans_iters"]), pqivf.WithSeed(c["seed"]),
                    pqivf.WithCandidateClusters(c["candidate_clusters"]), pqivf.WithTrainSample(c["train_sample"]),
                    pqivf.WithPQTrainSample(c["pq_train_sample"]))
    idx.bulk_add_arrays(np.arange(c["n"], dtype=np.int64), x)
    idx.Train()
    q, k, nprobe = synthetic.config4_batch(centers, 4096, seed=100)
    return {"idx": idx, "x": x, "centers": centers, "q": q, "k": k, "nprobe": nprobe}


@pytest.fixture(scope="module")
def c4_oracle(oracle, c4):
    """The exported CSR index, each entry's residual norm |y_e| and the FP64
    ground truth of the batch."""
    ix = _csr(oracle, c4["idx"])
    norms = np.zeros(ix.codes.shape[0], np.float64)
    for i in range(ix.m):
        tab = (ix.codebooks[i].astype(np.float64) ** 2).sum(1)
        norms += tab[ix.codes[:, i]]
    gt = _fp64_truth(c4["q"], c4["x"])
    return {"ix": ix, "ymax": float(np.sqrt(norms.max())), "gt": gt}


SAMPLE = np.r_[0:32, 4000:4032]   # 64 queries of the batch


def test_c4_plain_path_equals_oracle(hann, oracle, c4, c4_oracle):
    """The plain FP32 drop-in (decode-and-measure, R = 3 rerank) at full size:
    ids and distances identical to or_search_rerank on 64 queries with
    their own k and nprobe."""
    from hann_amd import _lib
    idx, q, k, nprobe = c4["idx"], c4["q"][SAMPLE], c4["k"][SAMPLE], c4["nprobe"][SAMPLE]
    ids, dist, cnt = idx.search_batch(q, k, nprobe=nprobe, scan_mode=_lib.SCAN_EXACT, rerank=3)
    wids, wdist, wcnt, _ = oracle.search_rerank_many(c4_oracle["ix"], q, k, nprobe, 3, nthreads=NCPU)
    assert np.array_equal(cnt, wcnt)
    for i in range(len(SAMPLE)):
        assert np.array_equal(ids[i, :cnt[i]], wids[i, :cnt[i]]), i
        assert np.array_equal(dist[i, :cnt[i]], wdist[i, :cnt[i]]), i


@pytest.mark.parametrize("mode", [3, 4])
def test_c4_list_scan_full_size(hann, oracle, c4, c4_oracle, mode):
    """The headline path (list-major BF16 or split-BF16 scan, certified BF16
    coarse step, R = 3) on the full batch. Bars:
      * counts equal k for every query (every query probes >= 8 lists);
      * every returned distance is the reference-order FP32 distance;
      * recall@10 against FP64 exact search >= 0.95;
      * on 64 queries, the ids equal or_search_rerank's except for entries
        whose PQ distance lies within the band of the query's candidate
        boundary B (the oracle's (k R)-th PQ distance), or ids that such an
        entry pushed out of a result: a candidate key differs from the PQ
        distance by at most 2^-6 |q| max|y| (BF16 operands; 2^-14 |q|
        max|y| for split BF16) plus 2^-10 of itself (the 16-bit key), so
        the two candidate sets can differ only on entries within twice that
        of B."""
    from hann_amd import _lib
    idx, x, q, k, nprobe = c4["idx"], c4["x"], c4["q"], c4["k"], c4["nprobe"]
    ids, dist, cnt = idx.search_batch(q, k, nprobe=nprobe, scan_mode=mode, rerank=3, fused=1,
                                      coarse_precision=_lib.PREC_BF16)
    assert idx.last_scan()[0] == mode
    assert np.array_equal(cnt, k)
    _exact_dist_ok(oracle, q, x, ids, dist, cnt)
    rec = _recall10(ids, c4_oracle["gt"])
    assert rec >= 0.95, rec
    ix = c4_oracle["ix"]
    qs, ks, ns = q[SAMPLE], k[SAMPLE], nprobe[SAMPLE]
    wids, wdist, wcnt, bound = oracle.search_rerank_many(ix, qs, ks, ns, 3, nthreads=NCPU)
    assert np.array_equal(wcnt, cnt[SAMPLE])
    pos = np.argsort(ix.ids)
    lists = np.searchsorted(ix.list_off, np.arange(ix.ids.size), side="right") - 1
    eps = 2.0 ** -6 if mode == 3 else 2.0 ** -14
    differ = 0

    def pq_of(qv, idl):
        e = pos[np.searchsorted(ix.ids[pos], idl)]
        r = np.concatenate([ix.codebooks[s][ix.codes[e, s]] for s in range(ix.m)], axis=1)
        r = (ix.coarse[lists[e]] + r).astype(np.float64)
        return ((r - qv.astype(np.float64)) ** 2).sum(1)

    for j, i in enumerate(SAMPLE):
        c = int(cnt[i])
        got, want = set(ids[i, :c].tolist()), set(wids[j, :c].tolist())
        if got == want:
            continue
        differ += 1
        B = float(bound[j])
        band = 2 * (eps * float(np.linalg.norm(qs[j])) * c4_oracle["ymax"] + 2.0 ** -10 * (B + 1e-6)) + 1e-5 * (1 + B)
        # an id in one result and not the other is either a candidate of one
        # side only (its PQ distance within the band of B) or was pushed
        # out by such a candidate (it ranks after the other side's last
        # result by (distance, id))
        g_last = (float(dist[i, c - 1]), int(ids[i, c - 1]))
        w_last = (float(wdist[j, c - 1]), int(wids[j, c - 1]))
        for r in range(c):
            for mine, other, d_, last in ((int(wids[j, r]), got, float(wdist[j, r]), g_last),
                                          (int(ids[i, r]), want, float(dist[i, r]), w_last)):
                if mine in other or (d_, mine) > last:
                    continue
                pqv = float(pq_of(qs[j], np.array([mine]))[0])
                assert abs(pqv - B) <= band, (i, mine, pqv, B, band)
    if mode == 4:
        assert differ <= len(SAMPLE) // 8, differ


def test_c4_bf16_coarse_equals_fp32(hann, c4):
    """The certified BF16 coarse step selects the same probes as FP32 at full
    size (16,384 centroids, nprobe up to 128): whole searches are identical."""
    from hann_amd import _lib
    idx, q, k, nprobe = c4["idx"], c4["q"], c4["k"], c4["nprobe"]
    a = idx.search_batch(q, k, nprobe=nprobe, scan_mode=3, rerank=3, fused=1, coarse_precision=_lib.PREC_FP32)
    b = idx.search_batch(q, k, nprobe=nprobe, scan_mode=3, rerank=3, fused=1, coarse_precision=_lib.PREC_BF16)
    for u, v in zip(a, b):
        assert np.array_equal(u, v)


# ------------------------------------------------------------------ C5 --
def test_c5_serving_alone_equals_mixed_with_mutations(hann, c4):
    """The served path of the headline (list-major BF16, certified BF16
    coarse step, R = 3) on the 10M index: lognormal request sizes (median 8,
    at most 1,024), per-query k and nprobe as C4, and BulkAdd / BulkDelete
    requests of 1,000 to 10,000 vectors in between (each added set deleted
    again later, so the index ends as it began). Every search request gets
    the same results alone (one request at a time) and mixed (everything
    queued at once and batched), and a search sees exactly the mutations
    submitted before it."""
    from hann_amd import _lib, serving, synthetic
    idx, centers = c4["idx"], c4["centers"]
    n0 = synthetic.CONFIG4["n"]
    rng = np.random.default_rng(55)
    q, k, nprobe = synthetic.config4_batch(centers, 3000, seed=300)
    sizes = np.minimum(rng.lognormal(np.log(8), 1.0, 400).astype(int) + 1, 1024)
    off = np.concatenate([[0], np.cumsum(sizes)])
    n_req = int(np.searchsorted(off, q.shape[0], side="right") - 1)
    # requests: searches, with an add of a fresh set after every 40th search
    # and the delete of that set 20 searches later; a probe request of four
    # of the set's vectors follows each add (it must find them) and each
    # delete (it must not)
    plan, nxt = [], n0
    adds = {}
    for r in range(n_req):
        plan.append(("search", r))
        if r % 40 == 10:
            cnt = int(rng.integers(1000, 10001))
            vecs, _ = synthetic.normalized_mixture(cnt, 96, 4096, 0.06, seed=1000 + r, centers=centers)
            ids = np.arange(nxt, nxt + cnt, dtype=np.int64)
            nxt += cnt
            adds[r] = (ids, vecs)
            plan += [("add", r), ("probe_in", r)]
        if r % 40 == 30 and (r - 20) in adds:
            plan += [("del", r - 20), ("probe_out", r - 20)]
    for r in adds:
        if ("del", r) not in plan:
            plan += [("del", r), ("probe_out", r)]
    srv = serving.Server(idx, max_batch=4096, slots=3, hbm_budget=8 << 30, scan_mode=_lib.SCAN_LIST_BF16, rerank=3,
                         fused=1, coarse_precision=_lib.PREC_BF16)
    try:
        def submit(kind, r):
            if kind == "search":
                sl = slice(off[r], off[r + 1])
                return srv.submit(q[sl], k[sl], nprobe[sl])
            if kind in ("probe_in", "probe_out"):
                return srv.submit(adds[r][1][:4], 1, 8)
            if kind == "add":
                return srv.submit_bulk_add(*adds[r])
            return srv.submit_bulk_delete(adds[r][0])

        def run(one_at_a_time):
            out = {}
            if one_at_a_time:
                for kind, r in plan:
                    out[(kind, r)] = srv.result(submit(kind, r))
            else:
                tickets = [(kind, r, submit(kind, r)) for kind, r in plan]
                for kind, r, t in tickets:
                    out[(kind, r)] = srv.result(t)
            return out

        alone = run(True)
        count_mid = idx.Stats().Count
        mixed = run(False)
        st = srv.stats()
    finally:
        srv.close()
    assert count_mid == n0 and idx.Stats().Count == n0
    for r in range(n_req):
        a, b = alone[("search", r)], mixed[("search", r)]
        sl = slice(off[r], off[r + 1])
        info = f"request {r}: k {k[sl].tolist()} nprobe {nprobe[sl].tolist()}"
        assert np.array_equal(a[2], b[2]), info
        assert np.array_equal(a[0], b[0]), info
        assert np.array_equal(a[1], b[1]), info
        assert np.array_equal(a[2], k[sl]), info
    for r in adds:
        for res in (alone, mixed):
            assert res[("probe_in", r)][0][:, 0].tolist() == adds[r][0][:4].tolist()
            if ("probe_out", r) in res:
                assert not np.isin(res[("probe_out", r)][0][:, 0], adds[r][0]).any()
    assert st["mutations"] == 2 * sum(1 for p in plan if p[0] in ("add", "del"))
    assert st["batches"] < 2 * n_req
    # every batch's charge bounded the device memory its slot's search took
    assert st["scratch_over_bound"] == 0, st
    assert 0 < st["budget_peak"] <= 8 << 30, st


# ------------------------------------------------------------------ C2 --
def test_c2_batch_topk_full_size(hann, oracle):
    """1,000,000 x 128 integer-valued base, 10,000 queries, squared
    Euclidean, k = 100: the certified BF16 pass returns exactly the FP32
    ids and distances for every query (integer data: both equal FP64), and
    both equal the oracle's exact top-k on 16 queries."""
    from hann_amd import _lib, core, synthetic
    base, qs = synthetic.config2()
    a_ids, a_d = core.batch_topk(core.SquaredEuclidean, qs, base, 100, precision=_lib.PREC_FP32)
    b_ids, b_d = core.batch_topk(core.SquaredEuclidean, qs, base, 100, precision=_lib.PREC_BF16)
    assert np.array_equal(a_ids, b_ids)
    assert np.array_equal(a_d, b_d)
    sample = np.r_[0:8, 9992:10000]
    wid, wrank = oracle.batch_topk(0, qs[sample], base, 100, nthreads=NCPU)
    assert np.array_equal(a_ids[sample], wid)
    assert np.array_equal(a_d[sample], wrank.astype(np.float64))
    # FP64: integer data, so every rank is exact
    d64 = np.stack([((base[a_ids[i]].astype(np.float64) - qs[i]) ** 2).sum(1) for i in sample])
    assert np.array_equal(d64, a_d[sample])


# ---------------------------------------------------------- C3 and A12 --
@pytest.fixture(scope="module")
def c3(hann):
    from hann_amd import _lib, pqivf, synthetic
    x, centers = synthetic.config3()
    out = {}
    for name, prec in (("fp32", _lib.PREC_FP32), ("bf16", _lib.PREC_BF16)):
        idx = pqivf.New(128, pqivf.WithCoarseK(4096), pqivf.WithNumSubquantizers(16), pqivf.WithPQK(256),
                        pqivf.WithKMeansIters(20), pqivf.WithSeed(3), pqivf.WithPQTrainSample(65536),
                        pqivf.WithTrainPrecision(prec), pqivf.WithKMeansInit(1))
        idx.bulk_add_arrays(np.arange(x.shape[0], dtype=np.int64), x)
        idx.Train()
        out[name] = idx
    out["x"], out["centers"] = x, centers
    return out


def test_c3_train_full_size(hann, oracle, c3):
    """Training at full size: FP32 and certified BF16 assignment give the
    same centroids (bits), codebooks, list membership and order, and codes;
    a 10,000-row sample's lists equal or_assign on the final centroids and
    its codes equal or_encode with the final codebooks."""
    from hann_amd import pqivf
    a, b, x = c3["fp32"], c3["bf16"], c3["x"]
    ca, cb_ = pqivf.CoarseCentroids(a), pqivf.CoarseCentroids(b)
    assert np.array_equal(ca.view(np.uint32), cb_.view(np.uint32))
    ka, la = pqivf.Codebooks(a)
    kb, lb = pqivf.Codebooks(b)
    assert np.array_equal(la, lb) and np.array_equal(ka.view(np.uint32), kb.view(np.uint32))
    ea = a.export_lists(with_vectors=False)
    eb = b.export_lists(with_vectors=False)
    for u, v in zip(ea[:3], eb[:3]):
        assert np.array_equal(u, v)
    off, ids, codes, _ = ea
    assert np.all(la == 256)
    rng = np.random.default_rng(17)
    rows = np.sort(rng.choice(x.shape[0], 10000, replace=False))
    where = np.empty(x.shape[0], np.int64)
    where[ids] = np.arange(ids.size)
    lists = np.searchsorted(off, where[rows], side="right") - 1
    want = oracle.assign(x[rows], ca, nthreads=NCPU)
    assert np.array_equal(lists, want)
    for r, e, c in zip(rows[:2000], where[rows[:2000]], want[:2000]):
        assert np.array_equal(codes[e], oracle.encode(x[r], ca[c], 16, ka, la)), r
    # lists hold ascending ids (pqivf/index.go:489-503)
    for c in range(0, 4096, 97):
        assert np.all(np.diff(ids[off[c]:off[c + 1]]) > 0)


def test_a12_online_insert_full_size(hann, oracle, c3):
    """BulkAdd of 10,000 vectors after Train (pqivf/index.go:245-297): each
    goes to the list or_assign picks on the trained centroids, carries the
    codes or_encode gives, and the new entries sit at their lists' tails in
    ascending id order."""
    from hann_amd import pqivf, synthetic
    idx = c3["fp32"]
    n0 = c3["x"].shape[0]
    rng = np.random.default_rng(23)
    centers = c3["centers"]
    new = (centers[rng.integers(0, centers.shape[0], 10000)] +
           rng.standard_normal((10000, 128), dtype=np.float32)).astype(np.float32)
    new_ids = n0 + rng.permutation(10000).astype(np.int64)
    off0, _, _, _ = idx.export_lists(with_vectors=False)
    idx.bulk_add_arrays(new_ids, new)
    off, ids, codes, _ = idx.export_lists(with_vectors=False)
    cents = pqivf.CoarseCentroids(idx)
    cbk, cbl = pqivf.Codebooks(idx)
    want = oracle.assign(new, cents, nthreads=NCPU)
    order = np.argsort(new_ids)
    for c in np.unique(want):
        tail = ids[off[c] + (off0[c + 1] - off0[c]):off[c + 1]]
        mine = np.sort(new_ids[want == c])
        assert tail.tolist() == mine.tolist(), c
    where = {int(v): j for j, v in enumerate(ids) if v >= n0}
    for j in order[:3000]:
        e = where[int(new_ids[j])]
        assert np.array_equal(codes[e], oracle.encode(new[j], cents[want[j]], 16, cbk, cbl)), j
    assert idx.Stats().Count == n0 + 10000
    del synthetic
