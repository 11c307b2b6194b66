"""Benchmark of the PQIVF search hot path on one MI355X.

Workload (the configuration the headline metric is quoted on): 10,000,000
x 96 float32 L2-normalised Gaussian-mixture vectors (4,096 components,
centers N(0, I) normalised, per-dimension std --sigma), indexed with 16,384
coarse lists and PQ 24 x 256; one step = one batch of --batch queries drawn
from the same mixture, each with its own k in {10, 50, 100} and nprobe in
{8, 16, 32, 64, 128}. Inputs are resident in HBM before the timed region.

Two search paths are measured on the same index and the same batches:
  plain   the FP32 drop-in: the reference's decode-and-measure PQ distance
          (bit-identical to pqivf/index.go:905-936) + an exact FP32 rerank
  chosen  the fastest at recall@10 >= 0.95 of three candidates, each with the
          certified BF16 MFMA coarse step (probes identical to FP32) and the
          same rerank: the persistent FP32 table scan (one table per query,
          per-entry list terms), or the list-major scan that decodes each
          probed list once and scores it against all its queries with BF16
          or split-BF16 MFMA
Recall: each path's rerank depth is the smallest of --rerank-grid whose
recall@10 against FP64 ground truth reaches 0.95 on a selection sample (the
first --gt-queries queries of a warm-up batch). The recall reported is then
measured on held-out queries (the first 1,024 of --heldout-batches timed
batches), with its standard error; a depth whose held-out recall falls below
0.95 is raised to the next one in the grid.

MALL: the 240 MB code plane fits in the 256 MB MALL, so the timed batches
rotate over --rotate identical index copies (built with the same seed), and
the one-stream pass that gives the stage and kernel times (the roofline)
also writes --flush-mb MB between batches.

`value` is the chosen path's queries per second with the timed batches
alternating over --streams streams (consecutive batches overlap on the
device; each stream has its own workspace in the library); p50/p99 are
per-batch latencies in that run.

python bench.py [--gpus 1] [--steps K] [--warmup W] [--n N] [--batch B]
"""
import argparse
import json
import os
import subprocess
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

METRIC = "QPS at recall@10≥0.95, 10M×96 PQIVF, chosen vs plain GPU; p50/p99 ms"
HBM_PEAK_GBS = 8000.0
# rocprofv3 counter summary of the list-major scan on the 10M x 96 workload:
# FETCH_SIZE and WRITE_SIZE per dispatch, scaled by what they read for a 1 GiB
# copy kernel
COUNTERS = os.path.join(ROOT, "profiles", "k_scan_list_pqivf_10m_x96_nlist16384_counters.txt")
# host threads the CPU baseline may use: one GPU's share of the host
CPU_SHARE = 16


def profiled_traffic(kernel):
    """HBM-side bytes per launch of `kernel` from the committed counter summary, or None."""
    try:
        with open(COUNTERS) as f:
            for line in f:
                if line.startswith(f"traffic_bytes_per_launch {kernel}:"):
                    return int(line.split(":")[1])
    except OSError:
        pass
    return None


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


def cpu_model():
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=10).stdout
        for line in out.splitlines():
            if line.startswith("Model name:"):
                return line.split(":", 1)[1].strip()
    except (OSError, subprocess.SubprocessError):
        pass
    return "unknown"


def make_batch(centers, sigma, nq, seed):
    from hann_amd import synthetic
    return synthetic.config4_batch(centers, nq, seed, sigma=sigma)


def exact_neighbors(torch, lib, base_dev, x, q, kk=32):
    """FP64 exact neighbours of each query: exact FP32 top-(kk + 22) on the
    GPU, re-ranked in FP64 and cut to kk, ordered by (distance, id).
    Returns (ids [nq][kk], FP64 squared distances [nq][kk])."""
    from hann_amd import _lib
    nq = q.shape[0]
    kw = min(kk + 22, base_dev.shape[0])
    qd = torch.from_numpy(np.ascontiguousarray(q)).cuda()
    ids = torch.empty((nq, kw), dtype=torch.int64, device="cuda")
    dist = torch.empty((nq, kw), dtype=torch.float64, device="cuda")
    _lib.check(lib.hann_batch_topk_dev(0, qd.data_ptr(), nq, base_dev.data_ptr(), base_dev.shape[0], q.shape[1], kw,
                                       0, 0, ids.data_ptr(), dist.data_ptr(), None))
    ids = ids.cpu().numpy()
    kk = min(kk, kw)
    out_i = np.empty((nq, kk), np.int64)
    out_d = np.empty((nq, kk), np.float64)
    for i in range(nq):
        d64 = ((x[ids[i]].astype(np.float64) - q[i].astype(np.float64)) ** 2).sum(1)
        o = np.lexsort((ids[i], d64))[:kk]
        out_i[i], out_d[i] = ids[i][o], d64[o]
    return out_i, out_d


def recall_hits(got_ids, gt):
    """recall@10 hits per query (0..10)"""
    return np.array([len(set(got_ids[i, :10].tolist()) & set(gt[i, :10].tolist())) for i in range(gt.shape[0])])


def recall_stats(hits):
    """recall@10 and its standard error over queries"""
    r = hits / 10.0
    return float(r.mean()), float(r.std(ddof=1) / np.sqrt(max(r.size, 1))) if r.size > 1 else 0.0


class DevBatch:
    def __init__(self, torch, q, k, nprobe):
        self.nq = q.shape[0]
        self.q = torch.from_numpy(q).cuda()
        self.k = torch.from_numpy(k).cuda()
        self.np = torch.from_numpy(nprobe).cuda()
        self.kmax = int(k.max())
        self.pmax = int(nprobe.max())
        self.ids = torch.empty((self.nq, self.kmax), dtype=torch.int64, device="cuda")
        self.dist = torch.empty((self.nq, self.kmax), dtype=torch.float32, device="cuda")
        self.cnt = torch.empty(self.nq, dtype=torch.int32, device="cuda")


def run_search(idx, b, stream, mode, rerank, fused, coarse):
    idx.search_batch_dev(b.q.data_ptr(), b.nq, b.k.data_ptr(), b.np.data_ptr(), b.kmax, b.pmax, b.ids.data_ptr(),
                         b.dist.data_ptr(), b.cnt.data_ptr(), stream, scan_mode=mode, rerank=rerank, fused=fused,
                         coarse_precision=coarse)


def probe_bytes(torch, lib, idx, b, list_sizes, m):
    """Algorithmic scan bytes of one batch: PQ codes of every distinct probed
    list read once (sum of n_l * m over the union), plus the per-pair total."""
    from hann_amd import _lib, pqivf
    cents = torch.from_numpy(pqivf.CoarseCentroids(idx)).cuda()
    ids = torch.empty((b.nq, b.pmax), dtype=torch.int64, device="cuda")
    dist = torch.empty((b.nq, b.pmax), dtype=torch.float64, device="cuda")
    _lib.check(lib.hann_batch_topk_dev(0, b.q.data_ptr(), b.nq, cents.data_ptr(), cents.shape[0], cents.shape[1],
                                       b.pmax, 0, 0, ids.data_ptr(), dist.data_ptr(), None))
    probes = ids.cpu().numpy()
    nps = b.np.cpu().numpy()
    mask = np.arange(b.pmax)[None, :] < nps[:, None]
    used = np.unique(probes[mask])
    distinct = int(list_sizes[used].sum()) * m
    pairs = int(list_sizes[probes[mask]].sum()) * m
    return distinct, pairs


def cpu_baseline(idx, q, k, nprobe, budget_s, threads_list):
    """The oracle (C restatement of pqivf/index.go Search) on the host cores,
    one query per thread, over a bounded sample of the batch, once per thread
    count (the first is the reported value)."""
    import oracle
    off, ids, codes, vecs = idx.export_lists(with_vectors=True)
    kc, m, p, d = idx.shape()
    from hann_amd import pqivf
    cb, cbl = pqivf.Codebooks(idx)

    class Ix:
        pass
    ix = Ix()
    ix.m = m
    ix.coarse = pqivf.CoarseCentroids(idx)
    ix.codebooks = np.ascontiguousarray(cb)
    ix.cb_len = cbl.astype(np.uint64)
    ix.list_off = off
    ix.ids = ids
    ix.vecs = vecs
    ix.codes = np.ascontiguousarray(codes)
    ix.has_codes = (codes[:, 0] >= 0).astype(np.uint8)
    per = budget_s / len(threads_list)
    by, samples = {}, []
    for threads in threads_list:
        done = 0
        t0 = time.perf_counter()
        chunk = max(threads * 4, 8)
        while done < q.shape[0] and time.perf_counter() - t0 < per:
            sl = slice(done, min(q.shape[0], done + chunk))
            # the chunk's queries grouped by their own (k, nprobe)
            for kk in np.unique(k[sl]):
                for pp in np.unique(nprobe[sl]):
                    sel = np.nonzero((k[sl] == kk) & (nprobe[sl] == pp))[0] + sl.start
                    if sel.size:
                        oracle.search_many(ix, q[sel], int(kk), int(pp), nthreads=threads)
            done = sl.stop
        el = time.perf_counter() - t0
        by[str(threads)] = round(done / el, 2)
        samples.append(f"{done} queries / {threads} threads / {el:.1f} s")
    t0 = threads_list[0]
    return {"value": by[str(t0)], "unit": "queries/s", "cores": t0, "kind": "port", "by_threads": by,
            "cpu_model": cpu_model(), "nproc": os.cpu_count(),
            "sample": "queries of the first batch in order (their own k and nprobe), oracle Search "
                      "(decode-and-measure, no rerank), one query per thread: " + "; ".join(samples)}


def bench_topk(a, torch, lib):
    """Config 2: 1,000,000 x 128 integer-valued float32 base, 10,000 queries,
    squared Euclidean, k = 100. Exact FP32 (reference order) vs the certified
    BF16 MFMA candidate pass + exact FP32 ranking (identical results)."""
    from hann_amd import _lib, synthetic
    import oracle
    base, qs = synthetic.config2(n=a.n_topk, q=a.q_topk)
    bd = torch.from_numpy(base).cuda()
    qd = torch.from_numpy(qs).cuda()
    k = 100
    out = {}
    stream = torch.cuda.Stream()
    for name, prec in (("fp32", _lib.PREC_FP32), ("bf16", _lib.PREC_BF16)):
        ids = torch.empty((qs.shape[0], k), dtype=torch.int64, device="cuda")
        dist = torch.empty((qs.shape[0], k), dtype=torch.float64, device="cuda")
        def run():
            _lib.check(lib.hann_batch_topk_dev(0, qd.data_ptr(), qs.shape[0], bd.data_ptr(), base.shape[0], 128, k,
                                               prec, 0, ids.data_ptr(), dist.data_ptr(), stream.cuda_stream))
        run()
        torch.cuda.synchronize()
        t = time.perf_counter()
        for _ in range(a.steps):
            run()
        torch.cuda.synchronize()
        el = (time.perf_counter() - t) / a.steps
        out[name] = {"s_per_run": round(el, 4), "queries_per_s": round(qs.shape[0] / el, 1),
                     "tflops": round(2.0 * qs.shape[0] * base.shape[0] * 128 / el / 1e12, 2),
                     "ids": ids.cpu().numpy(), "dist": dist.cpu().numpy()}
        log(f"topk {name}: {el * 1000:.1f} ms per run")
    same = bool(np.array_equal(out["fp32"]["ids"], out["bf16"]["ids"]) and
                np.array_equal(out["fp32"]["dist"], out["bf16"]["dist"]))
    # error of the returned distances against FP64 (integer data: exact)
    sq = out["bf16"]["dist"][:100]
    ids = out["bf16"]["ids"][:100]
    d64 = np.stack([((base[ids[i]].astype(np.float64) - qs[i]) ** 2).sum(1) for i in range(100)])
    err = float(np.max(np.abs(sq - d64) / np.maximum(d64, 1)))
    # CPU baseline: the oracle on a bounded sample of queries
    threads = min(os.cpu_count() or 1, CPU_SHARE)
    ns = max(threads, 16)
    t = time.perf_counter()
    wid, _ = oracle.batch_topk(0, qs[:ns], base, k, nthreads=threads)
    cpu_el = time.perf_counter() - t
    assert np.array_equal(wid, out["fp32"]["ids"][:ns])
    fl = out["bf16"]
    return {"metric": "batch top-k queries/s (1M x 128 int base, 10k queries, k=100)", "value": fl["queries_per_s"],
            "unit": "queries/s", "n_gpus": 1, "steps": a.steps, "warmup": 1, "ms_per_step": round(fl["s_per_run"] * 1e3, 3),
            "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "bf16+f32", "data": "synthetic",
            "config": {"workload": f"exact batch top-k {base.shape[0]}x128 integer data, {qs.shape[0]} queries, k=100"},
            "fp32": {k2: v for k2, v in out["fp32"].items() if k2 not in ("ids", "dist")},
            "bf16": {k2: v for k2, v in out["bf16"].items() if k2 not in ("ids", "dist")},
            "bf16_identical_to_fp32": same, "max_rel_err_vs_fp64": err,
            "roofline": {"bound": "mfma", "achieved": fl["tflops"], "peak": 2500.0, "unit": "TFLOP/s",
                         "frac": round(fl["tflops"] / 2500.0, 4), "traffic": None},
            "cpu_baseline": {"value": round(ns / cpu_el, 2), "unit": "queries/s", "cores": threads, "kind": "port",
                             "cpu_model": cpu_model(), "nproc": os.cpu_count(),
                             "sample": f"{ns} queries, oracle exact top-100 over the 1M base, {threads} threads"}}


def bench_train(a, torch, lib):
    """Config 3: 1,000,000 x 128 from a 1,024-component mixture (centers
    N(0, 25 I), std 1); coarse k-means K = 4,096, 20 iterations; PQ 16 x 256
    on a 65,536-residual sample; encode all. FP32 vs BF16 assignment."""
    from hann_amd import _lib, pqivf, synthetic
    import oracle
    n = a.n_train
    x, _ = synthetic.config3(n=n)
    res = {}
    lists = {}
    for name, prec in (("fp32", _lib.PREC_FP32), ("bf16", _lib.PREC_BF16)):
        idx = pqivf.New(128, pqivf.WithCoarseK(4096), pqivf.WithNumSubquantizers(16), pqivf.WithPQK(256),
                        pqivf.WithKMeansIters(a.train_iters), pqivf.WithSeed(3), pqivf.WithPQTrainSample(65536),
                        pqivf.WithTrainPrecision(prec), pqivf.WithKMeansInit(1))
        idx.bulk_add_arrays(np.arange(n), x)
        torch.cuda.synchronize()
        t = time.perf_counter()
        idx.Train()
        el = time.perf_counter() - t
        res[name] = round(el, 3)
        off, ids, codes, _ = idx.export_lists(with_vectors=False)
        lists[name] = (off, ids, codes, pqivf.CoarseCentroids(idx))
        log(f"train {name}: {el:.2f} s")
        del idx
    same = all(np.array_equal(lists["fp32"][i], lists["bf16"][i]) for i in range(3)) and \
        np.array_equal(lists["fp32"][3].view(np.uint32), lists["bf16"][3].view(np.uint32))
    # the assignment passes: one per Lloyd iteration plus the final one
    flops = 2.0 * n * 4096 * 128 * (a.train_iters + 1)
    best = min(res, key=res.get)   # both give the same index (checked below)
    # CPU baseline: the oracle's assignment (the reference's assignPoints, the
    # step that dominates Train) of a bounded sample against the final
    # centroids, extrapolated to the 21 passes over all points
    threads = min(os.cpu_count() or 1, CPU_SHARE)
    cents = lists[best][3]
    ns = 4096
    t = time.perf_counter()
    oracle.assign(x[:ns], cents, nthreads=threads)
    cpu_el = time.perf_counter() - t
    cpu_s = cpu_el * (n / ns) * (a.train_iters + 1)
    peak = 2500.0 if best == "bf16" else 157.3
    achieved = flops / res[best] / 1e12
    return {"metric": "PQIVF train seconds (1M x 128, K=4096, 20 iters, k-means++, PQ 16x256)",
            "value": res[best], "unit": "s", "n_gpus": 1, "steps": 1, "warmup": 0, "ms_per_step": res[best] * 1e3,
            "higher_is_better": False, "scaling": "weak", "vs_baseline": None,
            "dtype": "f32" if best == "fp32" else "bf16+f32", "data": "synthetic",
            "config": {"workload": f"train {n}x128 K=4096 iters={a.train_iters} k-means++ PQ16x256",
                       "assignment": best},
            "fp32_s": res["fp32"], "bf16_s": res["bf16"], "bf16_identical_to_fp32": bool(same),
            "assign_tflops_fp32": round(flops / res["fp32"] / 1e12, 2),
            "assign_tflops_bf16": round(flops / res["bf16"] / 1e12, 2),
            "roofline": {"bound": "mfma", "achieved": round(achieved, 2), "peak": peak, "unit": "TFLOP/s",
                         "frac": round(achieved / peak, 4), "traffic": None,
                         "note": "assignment flops (2 N K d per pass, 21 passes) over the whole Train time"},
            "cpu_baseline": {"value": round(cpu_s, 1), "unit": "s", "cores": threads, "kind": "port",
                             "cpu_model": cpu_model(), "nproc": os.cpu_count(),
                             "sample": f"oracle assignPoints of {ns} points against the 4,096 final centroids "
                                       f"({cpu_el:.2f} s, {threads} threads), scaled to {n} points x "
                                       f"{a.train_iters + 1} passes (the update, PQ training and encoding not "
                                       f"counted)"}}


def build_index(a, x):
    from hann_amd import pqivf
    idx = pqivf.New(a.dim, pqivf.WithCoarseK(a.nlist), pqivf.WithNumSubquantizers(a.m), pqivf.WithPQK(256),
                    pqivf.WithKMeansIters(a.iters), pqivf.WithSeed(1), pqivf.WithCandidateClusters(8),
                    pqivf.WithTrainSample(min(a.train_sample, a.n)), pqivf.WithPQTrainSample(min(65536, a.n)))
    idx.bulk_add_arrays(np.arange(a.n, dtype=np.int64), x)
    idx.Train()
    return idx


def serving_plan(rate, seconds, pool_n, rng, next_id):
    """An open-loop request schedule: Poisson arrivals of search requests
    (lognormal sizes, median 8, at most 1,024 queries) at `rate` offered
    queries/s, 1% of requests BulkAdd or BulkDelete of 1,000 to 10,000
    vectors. Adds insert fresh ids; deletes remove the oldest ids still
    present from earlier adds. Searches take pool queries in order."""
    mean_size = float(np.mean(np.minimum(rng.lognormal(np.log(8), 1.0, 100000).astype(int) + 1, 1024)))
    n_req = max(100, int(rate * seconds / mean_size))
    sizes = np.minimum(rng.lognormal(np.log(8), 1.0, n_req).astype(int) + 1, 1024).astype(np.int32)
    at = np.cumsum(rng.exponential(1e6 * mean_size / rate, n_req))
    kind = np.zeros(n_req, np.int32)
    mut = rng.random(n_req) < 0.01
    kind[mut] = rng.integers(1, 3, int(mut.sum()))
    sizes[mut] = rng.integers(1000, 10001, int(mut.sum()))
    q_off = np.zeros(n_req, np.int64)
    srch = kind == 0
    q_off[srch] = np.concatenate([[0], np.cumsum(sizes[srch])[:-1]])
    # mutations in arrival order: which ids each adds or deletes
    add_ids, del_ids = [], []
    present = []   # ids still present from adds, oldest first
    head = 0
    epoch = np.zeros(n_req, np.int64)   # mutations before each request
    born, died = {}, {}
    nmut = 0
    for r in range(n_req):
        epoch[r] = nmut
        if kind[r] == 1:
            ids = np.arange(next_id, next_id + sizes[r], dtype=np.int64)
            next_id += int(sizes[r])
            add_ids.append(ids)
            present.extend(ids.tolist())
            for i in ids.tolist():
                born[i] = nmut
            nmut += 1
        elif kind[r] == 2:
            take = present[head:head + int(sizes[r])]
            head += len(take)
            if not take:
                take = [next_id + 10 ** 9]   # an unknown id: BulkDelete skips it
            sizes[r] = len(take)
            del_ids.append(np.array(take, np.int64))
            for i in take:
                died[i] = nmut
            nmut += 1
    add_ids = np.concatenate(add_ids) if add_ids else np.zeros(0, np.int64)
    del_ids = np.concatenate(del_ids) if del_ids else np.zeros(0, np.int64)
    left = np.array(present[head:], np.int64)
    n_q = int(sizes[srch].sum())
    pool_idx = np.arange(n_q) % pool_n
    # per served query: the epoch of its request
    q_epoch = np.repeat(epoch[srch], sizes[srch])
    return {"at": at, "kind": kind, "sizes": sizes, "q_off": q_off, "n_req": n_req, "n_q": n_q, "pool_idx": pool_idx,
            "q_epoch": q_epoch, "add_ids": add_ids, "del_ids": del_ids, "left": left, "born": born, "died": died,
            "mutations": nmut, "next_id": next_id}


def served_recall(plan, ids_out, base_gt, added_nb, n_base):
    """recall@10 hits of every served query against FP64 exact search of the
    index as that query saw it (the base plus the added vectors present at
    its epoch). base_gt: (ids, d64) of the 32 nearest base vectors per pool
    query; added_nb: (ids, d64) of the nearest added vectors per pool query
    (at most 32, over every add of the plan). A query whose exact top 10 the
    two lists do not decide (fewer than 10 present added neighbours listed
    while the list is cut at 32 and its last is nearer than the merged 10th)
    is left out and counted."""
    bi, bd = base_gt
    ai, ad = added_nb
    p = plan["pool_idx"]
    e = plan["q_epoch"][:, None]
    cid, cd = bi[p], bd[p]
    if ai is not None:
        born = np.full(plan["next_id"] - n_base, 1 << 62, np.int64)
        died = np.full(plan["next_id"] - n_base, 1 << 62, np.int64)
        for k_, v in plan["born"].items():
            born[k_ - n_base] = v
        for k_, v in plan["died"].items():
            if k_ - n_base < died.size:
                died[k_ - n_base] = v
        aid = ai[p]
        live = (born[aid - n_base] < e) & (died[aid - n_base] >= e)
        cid = np.concatenate([cid, aid], axis=1)
        cd = np.concatenate([cd, np.where(live, ad[p], np.inf)], axis=1)
    o = np.argsort(cid, axis=1, kind="stable")
    cid, cd = np.take_along_axis(cid, o, 1), np.take_along_axis(cd, o, 1)
    o = np.argsort(cd, axis=1, kind="stable")[:, :10]
    top, top_d = np.take_along_axis(cid, o, 1), np.take_along_axis(cd, o, 1)
    keep = np.ones(p.size, bool)
    if ai is not None and ai.shape[1] == 32:
        keep = ~((live.sum(1) < 10) & (ad[p][:, -1] < top_d[:, -1]))
    hits = (ids_out[:, :10, None] == top[:, None, :]).any(2).sum(1)
    return hits[keep], int((~keep).sum())


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--workload", default="pqivf", choices=["pqivf", "topk", "train"])
    ap.add_argument("--n-topk", type=int, default=1_000_000)
    ap.add_argument("--q-topk", type=int, default=10_000)
    ap.add_argument("--n-train", type=int, default=1_000_000)
    ap.add_argument("--train-iters", type=int, default=20)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--n", type=int, default=10_000_000)
    ap.add_argument("--dim", type=int, default=96)
    ap.add_argument("--nlist", type=int, default=16384)
    ap.add_argument("--m", type=int, default=24)
    ap.add_argument("--components", type=int, default=4096)
    ap.add_argument("--sigma", type=float, default=0.06)
    ap.add_argument("--batch", type=int, default=4096)
    ap.add_argument("--iters", type=int, default=10)
    ap.add_argument("--train-sample", type=int, default=1_000_000)
    ap.add_argument("--gt-queries", type=int, default=2048, help="selection sample for the rerank depth")
    ap.add_argument("--heldout-batches", type=int, default=8, help="timed batches whose first 1,024 queries "
                                                                   "give the held-out recall")
    ap.add_argument("--rerank-grid", default="1,2,3,4,8")
    ap.add_argument("--rotate", type=int, default=2, help="identical index copies the timed batches rotate over")
    ap.add_argument("--flush-mb", type=int, default=512, help="MALL flush between batches of the one-stream pass")
    ap.add_argument("--cpu-budget", type=float, default=15.0)
    ap.add_argument("--cpu-threads", default=f"{CPU_SHARE},6,1")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--streams", type=int, default=4,
                    help="streams the timed batches alternate over (one workspace each)")
    ap.add_argument("--graphs", type=int, default=0,
                    help="1: each stream replays a captured search graph instead of direct calls")
    ap.add_argument("--serve-seconds", type=float, default=1.5, help="arrival window per offered rate")
    ap.add_argument("--serve-rates", default="5000,20000,50000,100000,200000", help="offered queries/s")
    ap.add_argument("--paths", default="", help="comma-separated subset of the chosen-path candidates")
    ap.add_argument("--copy-probe-mb", type=int, default=0,
                    help="run one device copy of this many MB first (a known-size kernel for counter scaling)")
    a = ap.parse_args()
    if a.gpus != 1:
        print(f"bench.py measures one GPU: this build runs on a single MI355X (asked for --gpus {a.gpus}).",
              file=sys.stderr)
        sys.exit(2)

    import torch
    from hann_amd import _lib, synthetic
    lib = _lib.load()
    _lib.check(lib.hann_init(0))
    torch.cuda.init()
    if a.copy_probe_mb > 0:
        src = torch.ones(a.copy_probe_mb << 18, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        log(f"copy probe: {a.copy_probe_mb} MB read, {a.copy_probe_mb} MB written")
        del src, dst
    if a.workload == "topk":
        print(json.dumps(bench_topk(a, torch, lib)), flush=True)
        return
    if a.workload == "train":
        print(json.dumps(bench_train(a, torch, lib)), flush=True)
        return

    t0 = time.time()
    x, centers = synthetic.normalized_mixture(a.n, a.dim, a.components, a.sigma, seed=7)
    log(f"data {x.shape} in {time.time() - t0:.1f}s")
    t0 = time.time()
    idx = build_index(a, x)
    train_s = time.time() - t0
    log(f"bulk add + train {train_s:.1f}s")
    idxs = [idx]
    for r in range(1, max(1, a.rotate)):
        t0 = time.time()
        idxs.append(build_index(a, x))
        log(f"index copy {r}: {time.time() - t0:.1f}s")
    off, _, _, _ = idx.export_lists(with_vectors=False)
    list_sizes = np.diff(off)
    if len(idxs) > 1:
        assert np.array_equal(np.diff(idxs[-1].export_lists(with_vectors=False)[0]), list_sizes)

    batches = [make_batch(centers, a.sigma, a.batch, seed=100 + i) for i in range(a.warmup + a.steps)]
    dev_batches = [DevBatch(torch, *b) for b in batches]
    base_dev = torch.from_numpy(x).cuda()
    # selection sample: the first gt_queries of warm-up batch 0 (and the
    # 32 nearest base vectors of all of batch 0: the serving pool)
    pool_gt = exact_neighbors(torch, lib, base_dev, x, batches[0][0])
    nsel = min(a.gt_queries, a.batch)
    n_ho = min(a.heldout_batches, a.steps)
    ho_q = np.concatenate([batches[a.warmup + i][0][:1024] for i in range(n_ho)])
    ho_k = np.concatenate([batches[a.warmup + i][1][:1024] for i in range(n_ho)])
    ho_p = np.concatenate([batches[a.warmup + i][2][:1024] for i in range(n_ho)])
    ho_gt = exact_neighbors(torch, lib, base_dev, x, ho_q)
    stream = torch.cuda.Stream()
    sh = stream.cuda_stream

    def selection_recall(mode, rerank, fused, coarse):
        b = dev_batches[0]
        run_search(idx, b, sh, mode, rerank, fused, coarse)
        torch.cuda.synchronize()
        return recall_stats(recall_hits(b.ids.cpu().numpy()[:nsel], pool_gt[0][:nsel]))

    def heldout(mode, rerank, fused, coarse):
        """recall@10 (and its standard error) of the held-out queries, and the
        returned distances' error against FP64 distances of the same ids"""
        ids, dist, cnt = idx.search_batch(ho_q, ho_k, nprobe=ho_p, scan_mode=mode, rerank=rerank, fused=fused,
                                          coarse_precision=coarse)
        rec, se = recall_stats(recall_hits(ids, ho_gt[0]))
        # every returned distance against the FP64 distance of its id; the
        # stated per-result bound is (d + 2) / 2 * 2^-24 relative
        worst = 0.0
        for i in range(0, ho_q.shape[0], 7):
            c = int(cnt[i])
            if c == 0:
                continue
            d64 = np.sqrt(((x[ids[i, :c]].astype(np.float64) - ho_q[i].astype(np.float64)) ** 2).sum(1))
            worst = max(worst, float(np.max(np.abs(dist[i, :c] - d64) / np.maximum(d64, 1e-30))))
        return rec, se, worst

    grid = [int(v) for v in a.rerank_grid.split(",")]
    # the plain FP32 drop-in, and the candidates for the chosen path (all
    # with the certified BF16 coarse step and an exact FP32 rerank)
    paths = {"plain": (_lib.SCAN_EXACT, 0, _lib.PREC_FP32),
             "table-f32": (_lib.SCAN_ADC_F32, 1, _lib.PREC_BF16),
             "list-bf16": (_lib.SCAN_LIST_BF16, 1, _lib.PREC_BF16),
             "list-bf16x3": (_lib.SCAN_LIST_BF16X3, 1, _lib.PREC_BF16),
             "list-fp8": (_lib.SCAN_LIST_FP8, 1, _lib.PREC_BF16)}
    if a.paths:
        keep = set(a.paths.split(",")) | {"plain"}
        paths = {k: v for k, v in paths.items() if k in keep}
    err_bound = (a.dim + 2) / 2 * 2.0 ** -24
    chosen_r, rec_sel, rec_ho = {}, {}, {}
    for name, (mode, fused, coarse) in paths.items():
        gi = 0
        while gi < len(grid):
            rs, _ = selection_recall(mode, grid[gi], fused, coarse)
            log(f"{name}: rerank {grid[gi]} selection recall@10 {rs:.4f}")
            if rs >= 0.95 or gi == len(grid) - 1:
                rh, se, worst = heldout(mode, grid[gi], fused, coarse)
                log(f"{name}: rerank {grid[gi]} held-out recall@10 {rh:.4f} +- {se:.4f} ({ho_q.shape[0]} queries), "
                    f"max rel err vs fp64 {worst:.3g} (bound {err_bound:.3g})")
                assert worst <= err_bound, (name, worst)
                chosen_r[name], rec_sel[name], rec_ho[name] = grid[gi], rs, (rh, se, worst)
                if rh >= 0.95 or gi == len(grid) - 1:
                    break
            gi += 1
    del base_dev
    torch.cuda.empty_cache()

    streams = [stream] + [torch.cuda.Stream() for _ in range(max(1, a.streams) - 1)]
    flush = torch.empty(max(1, a.flush_mb) << 18, dtype=torch.float32, device="cuda") if a.flush_mb > 0 else None
    kmax_all = max(b.kmax for b in dev_batches)
    pmax_all = max(b.pmax for b in dev_batches)

    def timed(mode, rerank, fused, coarse, profile, nstreams, graphs=False, rotate=None, flush_between=False):
        """Time a.steps batches, batch i on stream i mod nstreams (consecutive
        batches overlap on the device when nstreams > 1) and index copy
        i mod rotate; per-batch latency from events on the batch's own
        stream. flush_between: a write of the flush buffer before each batch
        (outside its events). With graphs, each stream owns one search graph
        bound to staging buffers: a batch is copied into its stream's staging
        buffers (inside the timed region) and the graph is replayed."""
        ss = streams[:nstreams]
        ix = idxs[:rotate or len(idxs)]
        gs = {}
        if graphs:
            for j, s_ in enumerate(ss):
                for r, ixr in enumerate(ix):
                    stage = DevBatch(torch, np.zeros((a.batch, a.dim), np.float32),
                                     np.full(a.batch, kmax_all, np.int32), np.full(a.batch, pmax_all, np.int32))
                    g = ixr.search_graph(stage.q.data_ptr(), stage.nq, stage.k.data_ptr(), stage.np.data_ptr(),
                                         kmax_all, pmax_all, stage.ids.data_ptr(), stage.dist.data_ptr(),
                                         stage.cnt.data_ptr(), s_.cuda_stream, scan_mode=mode, rerank=rerank,
                                         fused=fused, coarse_precision=coarse)
                    gs[(j, r)] = (stage, g)
            torch.cuda.synchronize()

        def step(b, i, j, s_):
            r = i % len(ix)
            if not graphs:
                run_search(ix[r], b, s_.cuda_stream, mode, rerank, fused, coarse)
                return
            stage, g = gs[(j, r)]
            with torch.cuda.stream(s_):
                stage.q.copy_(b.q, non_blocking=True)
                stage.k.copy_(b.k, non_blocking=True)
                stage.np.copy_(b.np, non_blocking=True)
            g.launch()

        for i in range(a.warmup):
            for j, s_ in enumerate(ss):
                step(dev_batches[i], i * len(ss) + j, j, s_)
        torch.cuda.synchronize()
        for ixr in ix:
            ixr.set_profile(profile)
            ixr.stage_times()
            ixr.scan_kernel_times()
        evs = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in range(a.steps)]
        torch.cuda.synchronize()
        t = time.perf_counter()
        for i in range(a.steps):
            j = i % nstreams
            s_ = ss[j]
            if flush_between and flush is not None:
                with torch.cuda.stream(s_):
                    flush.fill_(float(i))
            evs[i][0].record(s_)
            step(dev_batches[a.warmup + i], i, j, s_)
            evs[i][1].record(s_)
        torch.cuda.synchronize()
        el = time.perf_counter() - t
        st = np.zeros(4)
        kern = np.zeros(2)
        for ixr in ix:
            ixr.set_profile(False)
            st += np.array(ixr.stage_times())
            kern += np.array(ixr.scan_kernel_times())
        lat = np.array([s.elapsed_time(e) for s, e in evs])
        if graphs:
            log(f"graphs: {[g.counts() for _, g in gs.values()]}")
            for _, g in gs.values():
                g.close()
        return el, lat, st, kern

    res = {}
    for name, (mode, fused, coarse) in paths.items():
        # stage and scan-kernel times from one stream (no overlap between
        # batches, MALL flushed before each); throughput and latency with
        # a.streams streams
        el1, lat1, st, kern = timed(mode, chosen_r[name], fused, coarse, True, 1, flush_between=True)
        tuned = idx.last_tune()
        el, lat, _, _ = timed(mode, chosen_r[name], fused, coarse, False, a.streams, a.graphs)
        qps = a.steps * a.batch / el
        calls = max(st[3], 1)
        res[name] = {"qps": qps, "ms_per_step": 1000 * el / a.steps, "p50_ms": float(np.percentile(lat, 50)),
                     "p99_ms": float(np.percentile(lat, 99)), "stages": st, "rerank": chosen_r[name],
                     "recall": rec_ho[name][0], "recall_se": rec_ho[name][1], "recall_sel": rec_sel[name],
                     "err": rec_ho[name][2], "kernel_ms": kern[0] / calls, "sample_ms": kern[1] / calls,
                     "tuned": tuned,
                     "p50_ms_one_stream": float(np.percentile(lat1, 50)),
                     "p99_ms_one_stream": float(np.percentile(lat1, 99))}
        log(f"{name}: {qps:.0f} q/s ({a.streams} streams, {len(idxs)} index copies), {1000 * el / a.steps:.2f} "
            f"ms/step, stages {st[:3] / calls}, scan kernel {kern[0] / calls:.4f} ms, sample pass "
            f"{kern[1] / calls:.4f} ms")
    # chosen: the fastest candidate path at held-out recall@10 >= 0.95 (else the most accurate)
    cands = [n for n in res if n != "plain"]
    ok = [n for n in cands if res[n]["recall"] >= 0.95]
    chosen_name = max(ok, key=lambda n: res[n]["qps"]) if ok else max(cands, key=lambda n: res[n]["recall"])
    res["chosen"] = res[chosen_name]
    paths["chosen"] = paths[chosen_name]
    chosen_r["chosen"] = chosen_r[chosen_name]
    # the same path on one index copy (its code plane can stay in the MALL)
    mode, fused, coarse = paths["chosen"]
    el_w, lat_w, _, _ = timed(mode, chosen_r["chosen"], fused, coarse, False, a.streams, a.graphs, rotate=1)
    qps_warm = a.steps * a.batch / el_w
    log(f"chosen on one index copy: {qps_warm:.0f} q/s")
    path_desc = {"table-f32": "certified bf16 coarse + persistent fp32 table scan (one table per query, per-entry "
                              "list terms, autotuned waves/query) + fp32 rerank",
                 "list-bf16": "certified bf16 coarse + list-major decode + bf16 mfma: a sample pass over each "
                              "query's nearest lists bounds its k*R-th key, a filter pass keeps the records under "
                              "the bound, a per-query select of k*R + fp32 rerank",
                 "list-bf16x3": "certified bf16 coarse + list-major decode + split-bf16 mfma: a sample pass over "
                                "each query's nearest lists bounds its k*R-th key, a filter pass keeps the records "
                                "under the bound, a per-query select of k*R + fp32 rerank",
                 "list-fp8": "certified bf16 coarse + list-major decode + fp8 (e4m3) mfma: a sample pass over "
                             "each query's nearest lists bounds its k*R-th key, a filter pass keeps the records "
                             "under the bound, a per-query select of k*R + fp32 rerank"}
    kernel_name = {"table-f32": "k_scan_block", "list-bf16": "k_scan_list<bf16,filter>",
                   "list-bf16x3": "k_scan_list<bf16x3,filter>", "list-fp8": "k_scan_list<fp8,filter>"}

    # roofline of the dominant kernel: the chosen path's list scan (filter
    # pass; the sample pass beside it)
    dist_b, pair_b = 0, 0
    for i in range(a.steps):
        db, pb = probe_bytes(torch, lib, idx, dev_batches[a.warmup + i], list_sizes, a.m)
        dist_b += db
        pair_b += pb
    kern_ms = res["chosen"]["kernel_ms"]
    samp_ms = res["chosen"]["sample_ms"]
    bytes_per_launch = dist_b / a.steps
    achieved = bytes_per_launch / (kern_ms * 1e-3) / 1e9
    both = bytes_per_launch / ((kern_ms + samp_ms) * 1e-3) / 1e9
    # traffic is the counter-measured figure for the 10M x 96 workload only
    traffic = profiled_traffic(kernel_name[chosen_name]) if (a.n, a.dim, a.nlist) == (10_000_000, 96, 16384) else None
    st_c = res["chosen"]["stages"]
    roofline = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                "kernel": kernel_name[chosen_name], "kernel_ms_per_launch": round(kern_ms, 4),
                "sample_pass_ms_per_launch": round(samp_ms, 4),
                "frac_sample_plus_filter": round(both / HBM_PEAK_GBS, 4),
                "scan_stage_ms_per_launch": round(st_c[1] / max(st_c[3], 1), 4),
                "mall": f"flushed ({a.flush_mb} MB write) before each launch; {len(idxs)} index copies",
                "algorithmic_bytes_per_launch": int(bytes_per_launch),
                "pair_bytes_per_launch": int(pair_b / a.steps)}

    # serving: open-loop Poisson arrivals through the continuous-batching
    # loop, at each offered rate, for both paths
    serve = {}
    if a.serve_seconds > 0:
        from hann_amd import serving
        q0, k0, n0 = batches[0]
        rates = [float(v) for v in a.serve_rates.split(",") if v]
        add_pool = {}
        for rate in rates:
            for name in ("plain", "chosen"):
                rng = np.random.default_rng(9)
                plan = serving_plan(rate, a.serve_seconds, q0.shape[0], rng, a.n)
                n_add = plan["add_ids"].size
                if n_add not in add_pool:
                    av, _ = synthetic.normalized_mixture(max(n_add, 1), a.dim, a.components, a.sigma, seed=4242,
                                                         centers=centers)
                    add_pool[n_add] = av
                add_vecs = add_pool[n_add]
                pool = q0[plan["pool_idx"]]
                kk, pp = k0[plan["pool_idx"]], n0[plan["pool_idx"]]
                mode, fused, coarse = paths[name]
                srv = serving.Server(idx, max_batch=a.batch, slots=2, hbm_budget=2 << 30, scan_mode=mode,
                                     rerank=chosen_r[name], fused=fused, coarse_precision=coarse)
                try:
                    ids_out, _, lat, wall = srv.replay(plan["at"], plan["kind"], plan["q_off"], plan["sizes"], pool,
                                                       kk, pp, plan["add_ids"], add_vecs, plan["del_ids"],
                                                       kmax_out=10)
                    st = srv.stats()
                finally:
                    srv.close()
                # the adds still present are removed, so the next run sees the same index
                if plan["left"].size:
                    idx.BulkDelete(plan["left"].tolist())
                assert idx.Stats().Count == a.n
                # recall of every served query against the index it saw
                added_nb = (None, None)
                if n_add:
                    ad = torch.from_numpy(add_vecs[:n_add]).cuda()
                    ai, add_d = exact_neighbors(torch, lib, ad, add_vecs[:n_add], q0, kk=min(32, n_add))
                    added_nb = (plan["add_ids"][ai], add_d)
                    del ad
                hits, undecided = served_recall(plan, ids_out, pool_gt, added_nb, a.n)
                rec, se = recall_stats(hits)
                sl = lat[plan["kind"] == 0]
                ml = lat[plan["kind"] != 0]
                serve.setdefault(name, []).append({
                    "offered_qps": round(rate, 1), "achieved_qps": round(plan["n_q"] / wall, 1),
                    "p50_ms": round(float(np.percentile(sl, 50)), 3), "p99_ms": round(float(np.percentile(sl, 99)), 3),
                    "requests": int(plan["n_req"]), "queries": int(plan["n_q"]),
                    "mutations": int(plan["mutations"]), "vectors_added": int(n_add),
                    "vectors_deleted": int(plan["del_ids"].size), "batches": int(st["batches"]),
                    "mutation_p50_ms": round(float(np.percentile(ml, 50)), 3) if ml.size else None,
                    "budget_peak_bytes": int(st["budget_peak"]), "recall_at_10_all_served": round(rec, 4),
                    "recall_se": round(se, 4), "recall_undecided_queries": int(undecided)})
                log(f"serve {name} @ {rate:.0f}: {serve[name][-1]}")

    cpu = None
    if not a.no_cpu:
        q0, k0, n0 = batches[0]
        th = [int(v) for v in a.cpu_threads.split(",")]
        th = [min(t, os.cpu_count() or 1) for t in th]
        cpu = cpu_baseline(idx, q0, k0, n0, a.cpu_budget, th)
        log(f"cpu baseline {cpu}")

    ch, pl = res["chosen"], res["plain"]

    def stages(r):
        return [round(v / max(r["stages"][3], 1), 4) for v in r["stages"][:3]]
    out = {
        "metric": METRIC, "value": round(ch["qps"], 1), "unit": "queries/s", "n_gpus": 1, "steps": a.steps,
        "warmup": a.warmup, "ms_per_step": round(ch["ms_per_step"], 3), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "bf16+f32", "data": "synthetic",
        "config": {"workload": f"pqivf search {a.n}x{a.dim} nlist {a.nlist} PQ{a.m}x256, batch {a.batch}, "
                               f"k in {{10,50,100}}, nprobe in {{8..128}}",
                   "n": a.n, "dim": a.dim, "nlist": a.nlist, "m": a.m, "pq_k": 256, "batch": a.batch,
                   "sigma": a.sigma, "components": a.components, "train_sample": min(a.train_sample, a.n),
                   "kmeans_iters": a.iters, "streams": a.streams, "index_copies": len(idxs),
                   "graphs": bool(a.graphs)},
        "p50_ms": round(ch["p50_ms"], 3), "p99_ms": round(ch["p99_ms"], 3),
        "recall_at_10": round(ch["recall"], 4), "recall_at_10_se": round(ch["recall_se"], 4),
        "recall_queries": int(ho_q.shape[0]),
        "chosen": {"qps": round(ch["qps"], 1), "qps_one_index_copy": round(qps_warm, 1),
                   "recall_at_10": round(ch["recall"], 4), "recall_at_10_se": round(ch["recall_se"], 4),
                   "recall_at_10_selection": round(ch["recall_sel"], 4), "rerank": ch["rerank"],
                   "max_rel_err_vs_fp64": ch["err"], "err_bound_rel": err_bound,
                   "p50_ms": round(ch["p50_ms"], 3), "p99_ms": round(ch["p99_ms"], 3),
                   "p50_ms_one_stream": round(ch["p50_ms_one_stream"], 3),
                   "p99_ms_one_stream": round(ch["p99_ms_one_stream"], 3),
                   "path": path_desc[chosen_name], "name": chosen_name, "stage_ms_per_step": stages(ch)},
        "candidates": {n: {"qps": round(res[n]["qps"], 1), "recall_at_10": round(res[n]["recall"], 4),
                           "recall_at_10_se": round(res[n]["recall_se"], 4), "rerank": res[n]["rerank"],
                           "p99_ms": round(res[n]["p99_ms"], 3), "scan_kernel_ms": round(res[n]["kernel_ms"], 4),
                           "sample_pass_ms": round(res[n]["sample_ms"], 4), "autotuner": res[n]["tuned"],
                           "stage_ms_per_step": stages(res[n])} for n in cands},
        "plain": {"qps": round(pl["qps"], 1), "recall_at_10": round(pl["recall"], 4),
                  "recall_at_10_se": round(pl["recall_se"], 4), "rerank": pl["rerank"],
                  "max_rel_err_vs_fp64": pl["err"],
                  "p50_ms": round(pl["p50_ms"], 3), "p99_ms": round(pl["p99_ms"], 3),
                  "p50_ms_one_stream": round(pl["p50_ms_one_stream"], 3),
                  "p99_ms_one_stream": round(pl["p99_ms_one_stream"], 3),
                  "path": "reference-order fp32 decode + fp32 rerank", "stage_ms_per_step": stages(pl)},
        "chosen_vs_plain": round(ch["qps"] / pl["qps"], 3),
        "serving": serve,
        "train_s": round(train_s, 2),
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
