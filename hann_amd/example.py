"""The reference's example harness (example/load_data.go, example/utils.go,
example/run_datasets.go) over this library's indexes.

A dataset directory holds train.csv (vectors to index, ids = row numbers),
test.csv (queries), neighbors.csv (ground-truth ids per query) and
distances.csv (ground-truth distances), all without headers, as the
ann-benchmarks exports the reference's download script writes.

Float files are parsed by numpy's C reader into float64 and narrowed to
float32: a value printed from a float32 (the exports' case) reads back to
that float32, as Go's strconv.ParseFloat(s, 32) gives. Integer files are
parsed strictly as decimal integers, as strconv.Atoi does. Errors carry the
reference's wording ("open", "read error", "parse error at col").

RunDataset adds one mode to the reference's: batch=True searches every test
query in one batched call on the GPU (search_batch) instead of one Search
per query from worker threads; recall is computed the same way.
"""
import io
import os
import queue
import threading
import time
import warnings

import numpy as np

from . import pqivf


class DatasetError(Exception):
    """A dataset file that cannot be opened, read or parsed."""


def _open(path):
    try:
        return open(path, "rb")
    except OSError as e:
        raise DatasetError(f"open {path}: {e.strerror}") from e


def _read_floats(path, skip_header=False):
    """readCSV[float32] (example/load_data.go:87-122): rows of float32."""
    with _open(path) as fh:
        data = fh.read()
    if skip_header:
        nl = data.find(b"\n")
        data = b"" if nl < 0 else data[nl + 1:]
    if not data.strip():
        return np.zeros((0, 0), np.float32)
    try:
        with warnings.catch_warnings():
            warnings.simplefilter("ignore")
            a = np.loadtxt(io.BytesIO(data), delimiter=",", dtype=np.float64, ndmin=2)
    except ValueError as e:
        msg = str(e)
        if "number of columns" in msg:
            raise DatasetError(f"read error in {path}: {msg}") from e
        col = msg.rsplit("column ", 1)[-1].rstrip(".") if "column " in msg else "?"
        raise DatasetError(f"parse error at col {col} in {path}: {msg}") from e
    return a.astype(np.float32)


def _read_ints(path, skip_header=False):
    """readCSV[int]: rows of decimal integers (no float or underscore forms)."""
    rows = []
    width = None
    with _open(path) as fh:
        for line_no, raw in enumerate(fh, 1):
            line = raw.decode().rstrip("\r\n")
            if skip_header:
                skip_header = False
                continue
            if line == "":
                continue
            fields = line.split(",")
            if width is None:
                width = len(fields)
            elif len(fields) != width:
                raise DatasetError(f"read error in {path}: record on line {line_no}: wrong number of fields")
            row = []
            for i, f in enumerate(fields):
                s = f.strip()
                body = s[1:] if s[:1] in "+-" else s
                if not body.isdigit() or not body.isascii():
                    raise DatasetError(f"parse error at col {i} in {path}: invalid syntax {s!r}")
                row.append(int(s))
            rows.append(row)
    return rows


def _read_doubles(path, skip_header=False):
    """readCSV[float64]."""
    with _open(path) as fh:
        data = fh.read()
    if skip_header:
        nl = data.find(b"\n")
        data = b"" if nl < 0 else data[nl + 1:]
    if not data.strip():
        return []
    try:
        with warnings.catch_warnings():
            warnings.simplefilter("ignore")
            a = np.loadtxt(io.BytesIO(data), delimiter=",", dtype=np.float64, ndmin=2)
    except ValueError as e:
        msg = str(e)
        if "number of columns" in msg:
            raise DatasetError(f"read error in {path}: {msg}") from e
        col = msg.rsplit("column ", 1)[-1].rstrip(".") if "column " in msg else "?"
        raise DatasetError(f"parse error at col {col} in {path}: {msg}") from e
    return a.tolist()


def LoadCSV(index, path, skip_header=False):
    """LoadCSV (example/load_data.go:71-85): add row i as id i."""
    vecs = _read_floats(path, skip_header)
    for i in range(vecs.shape[0]):
        try:
            index.Add(i, vecs[i])
        except Exception as e:
            raise DatasetError(f"failed to add vector {i}: {e}") from e


def LoadTrainingVectors(dir):
    """LoadTrainingVectors (example/load_data.go:141-157): {row: vector}
    as a float32 matrix (row i is id i)."""
    path = os.path.join(dir, "train.csv")
    try:
        return _read_floats(path)
    except DatasetError as e:
        raise DatasetError(f"failed to load train.csv: {e}") from e


def LoadTestDataset(dir):
    """LoadTestDataset (example/load_data.go:159-189): test vectors,
    ground-truth ids and distances."""
    out = []
    for name, reader in (("test.csv", _read_floats), ("neighbors.csv", _read_ints),
                         ("distances.csv", _read_doubles)):
        try:
            out.append(reader(os.path.join(dir, name)))
        except DatasetError as e:
            raise DatasetError(f"failed to load {name}: {e}") from e
    return tuple(out)


def LoadDataset(index, dir):
    """LoadDataset (example/load_data.go:16-69): train.csv into the index,
    then the test vectors and the ground truth."""
    try:
        LoadCSV(index, os.path.join(dir, "train.csv"), False)
    except DatasetError as e:
        raise DatasetError(f"failed to load train.csv: {e}") from e
    return LoadTestDataset(dir)


def FormatResults(results, max_results):
    """FormatResults (example/utils.go:10-22)."""
    return "".join(f"id={n.ID} (dist={n.Distance:.3f}) " for n in results[:max(max_results, 0)])


def FormatGroundTruth(neighbors, distances, k, max_results):
    """FormatGroundTruth (example/utils.go:24-36)."""
    lim = min(max(max_results, 0), len(neighbors))
    return "".join(f"id={neighbors[j]} (dist={distances[j]:.3f}) " for j in range(lim))


def RecallAtK(predicted, ground_truth, k):
    """RecallAtK (example/utils.go:38-66): the share of the first min(k,
    len(ground_truth)) true ids found among the first k predictions."""
    if k <= 0 or len(ground_truth) == 0:
        return 0.0
    pred = {int(n.ID if hasattr(n, "ID") else n) for n in list(predicted)[:k]}
    truth = list(ground_truth)[:min(k, len(ground_truth))]
    return sum(1 for i in truth if int(i) in pred) / len(truth)


def RunDataset(factory, dataset, root, k, num_queries, max_results, batch=False, out=print, **search_kw):
    """RunDataset (example/run_datasets.go:38-181): build the index from
    train.csv, train it if it trains, run kNN queries on test.csv, report
    Recall@k, per-query and overall times. num_queries < 0 or above the test
    set: every query, benchmark mode (no per-query lines). Worker threads
    come from HANN_BENCH_NTRD. batch=True: one batched GPU search of every
    query (search_kw passes its scan options). Returns the summary."""
    path = os.path.join(root, dataset)
    out(f"Loading dataset: {dataset}")
    t0 = time.perf_counter()
    index = factory()
    out(f"Created index: {type(index).__name__}")
    train = LoadTrainingVectors(path)
    if train.shape[0]:
        index.bulk_add_arrays(np.arange(train.shape[0], dtype=np.int64), train)
    if hasattr(index, "Train"):
        index.Train()
    test, gt_ids, gt_dist = LoadTestDataset(path)
    st = index.Stats()
    out(f"Indexed {st.Count} vectors ({st.Dimension} dimensions) in {time.perf_counter() - t0:.2f}s; "
        f"distance: {st.Distance}")
    bench = num_queries < 0 or num_queries > test.shape[0]
    if bench:
        num_queries = test.shape[0]
    threads = 1
    env = os.environ.get("HANN_BENCH_NTRD", "")
    if env.isdigit() and int(env) > 0:
        threads = int(env)
    out(f"Running kNN queries (k={k}) on {num_queries} test vectors using "
        f"{'one batched search' if batch else f'{threads} threads'}")
    results = [None] * num_queries
    durations = np.zeros(num_queries)
    if batch:
        tb = time.perf_counter()
        ids, dist, cnt = index.search_batch(test[:num_queries], k, **search_kw) if num_queries else ([], [], [])
        el = time.perf_counter() - tb
        for i in range(num_queries):
            results[i] = [pqivf.Neighbor(int(ids[i][j]), float(dist[i][j])) for j in range(int(cnt[i]))]
        durations[:] = el / max(num_queries, 1)
    else:
        tasks = queue.Queue()
        for i in range(num_queries):
            tasks.put(i)
        errors = []

        def worker():
            while True:
                try:
                    i = tasks.get_nowait()
                except queue.Empty:
                    return
                ts = time.perf_counter()
                try:
                    results[i] = index.Search(test[i], k)
                except Exception as e:   # the reference stops the program
                    errors.append((i, e))
                    return
                durations[i] = time.perf_counter() - ts

        ws = [threading.Thread(target=worker) for _ in range(threads)]
        for w in ws:
            w.start()
        for w in ws:
            w.join()
        if errors:
            i, e = errors[0]
            raise RuntimeError(f"Search error on query {i}: {e}") from e
    recalls = np.array([RecallAtK(results[i], gt_ids[i], k) for i in range(num_queries)])
    if not bench:
        for i in range(num_queries):
            out(f"Query #{i + 1}:")
            out(f" -> Predicted:     {FormatResults(results[i], max_results)}")
            out(f" -> Ground-truth:  {FormatGroundTruth(gt_ids[i], gt_dist[i], k, max_results)}")
            out(f" -> Recall@{k}:     {recalls[i]:.2f}, Response time: {durations[i] * 1e3:.3f}ms")
    avg_recall = float(recalls.mean()) if num_queries else 0.0
    avg_time = float(durations.mean()) if num_queries else 0.0
    runtime = time.perf_counter() - t0
    out(f"Average Recall@{k} over {num_queries} queries: {avg_recall:.2f}")
    out(f"Average query response time: {avg_time * 1e3:.3f}ms")
    out(f"Overall runtime: {runtime:.3f}s")
    return {"recall": avg_recall, "avg_response_s": avg_time, "runtime_s": runtime, "queries": num_queries,
            "benchmark_mode": bench, "results": results}
