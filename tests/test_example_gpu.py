"""The example harness end to end on the GPU (example/load_data_test.go:75-109,
example/run_datasets_test.go): a dataset directory loads into a PQIVF index
that trains and answers; RunDataset runs in per-query, benchmark (worker
threads) and batched modes, and its recall is the one RecallAtK gives for
the results it returns."""
import os

import numpy as np
import pytest

from hann_amd import example, pqivf
from test_example_host import write, write_dataset

pytestmark = pytest.mark.gpu


def exact_factory():
    # six codewords for six vectors: every residual is its own codeword
    return pqivf.New(4, pqivf.WithCoarseK(2), pqivf.WithNumSubquantizers(2), pqivf.WithPQK(6), pqivf.WithSeed(42))


def test_load_dataset_then_search(hann, tmp_path):
    idx = exact_factory()
    test, nb, dist = example.LoadDataset(idx, write_dataset(str(tmp_path)))
    idx.Train()
    for q, want in zip(test, nb):
        got = idx.Search(q, 1)
        assert got[0].ID == want[0]


def test_run_dataset_modes(hann, tmp_path, monkeypatch):
    root = str(tmp_path)
    os.makedirs(os.path.join(root, "synthetic"))
    write_dataset(os.path.join(root, "synthetic"))
    lines = []
    r = example.RunDataset(exact_factory, "synthetic", root, 2, 2, 3, out=lines.append)
    assert r["recall"] == 1.0 and not r["benchmark_mode"] and r["queries"] == 2
    assert any(line.startswith("Query #2:") for line in lines)
    monkeypatch.setenv("HANN_BENCH_NTRD", "2")
    r = example.RunDataset(exact_factory, "synthetic", root, 2, -1, 3, out=lambda s: None)
    assert r["benchmark_mode"] and r["queries"] == 2 and r["recall"] == 1.0
    # a small PQ codebook, as the reference's trained-index case
    trained = lambda: pqivf.New(4, pqivf.WithCoarseK(2), pqivf.WithNumSubquantizers(2), pqivf.WithPQK(2))
    r = example.RunDataset(trained, "synthetic", root, 2, 2, 3, out=lambda s: None)
    assert 0.0 <= r["recall"] <= 1.0


def test_run_dataset_batched_equals_per_query(hann, tmp_path, monkeypatch):
    """A 4,000 x 32 clustered dataset with FP64 ground truth: the batched GPU
    mode returns the per-query mode's results, and the reported recall is
    RecallAtK over them."""
    rng = np.random.default_rng(3)
    centers = rng.uniform(-10, 10, (16, 32))
    x = (centers[rng.integers(0, 16, 4000)] + rng.standard_normal((4000, 32))).astype(np.float32)
    q = (centers[rng.integers(0, 16, 60)] + rng.standard_normal((60, 32))).astype(np.float32)
    d2 = ((q.astype(np.float64)[:, None, :] - x.astype(np.float64)[None, :, :]) ** 2).sum(-1)
    gt = np.lexsort((np.broadcast_to(np.arange(4000), d2.shape), d2), axis=1)[:, :10]
    gd = np.sqrt(np.take_along_axis(d2, gt, 1))
    d = os.path.join(str(tmp_path), "clustered")
    os.makedirs(d)
    write(d, "train.csv", [",".join(str(v) for v in row) for row in x])
    write(d, "test.csv", [",".join(str(v) for v in row) for row in q])
    write(d, "neighbors.csv", [",".join(str(v) for v in row) for row in gt])
    write(d, "distances.csv", [",".join(repr(float(v)) for v in row) for row in gd])
    factory = lambda: pqivf.New(32, pqivf.WithCoarseK(16), pqivf.WithNumSubquantizers(8), pqivf.WithPQK(64),
                                pqivf.WithCandidateClusters(4), pqivf.WithSeed(5))
    monkeypatch.setenv("HANN_BENCH_NTRD", "4")
    a = example.RunDataset(factory, "clustered", str(tmp_path), 10, -1, 3, out=lambda s: None)
    b = example.RunDataset(factory, "clustered", str(tmp_path), 10, -1, 3, batch=True, out=lambda s: None)
    for ra, rb in zip(a["results"], b["results"]):
        assert [n.ID for n in ra] == [n.ID for n in rb]
        assert [n.Distance for n in ra] == [n.Distance for n in rb]
    want = np.mean([example.RecallAtK(a["results"][i], gt[i].tolist(), 10) for i in range(60)])
    assert a["recall"] == b["recall"] == pytest.approx(want)
    # the level is the quantizer's (8 x 64 codes, no rerank), not the harness's
    assert 0.0 < a["recall"] <= 1.0
