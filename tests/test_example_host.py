"""The example harness's loaders, recall and formatting on the CPU
(example/load_data_test.go, example/utils_test.go). The datasets are the
reference tests' own tiny fixtures, written to a temporary directory; an
untrained PQIVF index stands in for the RPT index (it keeps added vectors on
the host, so nothing here touches a GPU)."""
import os

import numpy as np
import pytest

from hann_amd import example, pqivf

TRAIN = ["0,0,0,0", "1,0,0,0", "0,1,0,0", "10,10,10,10", "10,10,10,11", "5,5,5,5"]
TEST = ["0.1,0,0,0", "10,10,10,10.2"]
NEIGHBORS = ["0,1", "3,4"]
DISTANCES = ["0.1,0.9", "0.2,0.8"]


def write(dir, name, lines):
    path = os.path.join(dir, name)
    with open(path, "w") as fh:
        fh.write("\n".join(lines) + "\n")
    return path


def write_dataset(dir):
    write(dir, "train.csv", TRAIN)
    write(dir, "test.csv", TEST)
    write(dir, "neighbors.csv", NEIGHBORS)
    write(dir, "distances.csv", DISTANCES)
    return dir


def new_index():
    return pqivf.New(4)


def test_load_dataset(tmp_path):
    idx = new_index()
    test, nb, dist = example.LoadDataset(idx, write_dataset(str(tmp_path)))
    assert idx.Stats().Count == len(TRAIN)
    assert test.shape == (2, 4) and test.dtype == np.float32
    assert np.array_equal(test[1], np.array([10, 10, 10, 10.2], np.float32))
    assert nb == [[0, 1], [3, 4]]
    assert dist == [[0.1, 0.9], [0.2, 0.8]]


@pytest.mark.parametrize("missing", ["train.csv", "test.csv", "neighbors.csv", "distances.csv"])
def test_load_dataset_missing_file(tmp_path, missing):
    d = write_dataset(str(tmp_path))
    os.remove(os.path.join(d, missing))
    with pytest.raises(example.DatasetError, match=missing):
        example.LoadDataset(new_index(), d)


def test_load_csv(tmp_path):
    idx = new_index()
    example.LoadCSV(idx, write(str(tmp_path), "train.csv", TRAIN), False)
    assert idx.Stats().Count == len(TRAIN)
    idx = new_index()
    example.LoadCSV(idx, write(str(tmp_path), "h.csv", ["a,b,c,d"] + TRAIN), True)
    assert idx.Stats().Count == len(TRAIN)
    with pytest.raises(example.DatasetError, match="open"):
        example.LoadCSV(new_index(), str(tmp_path / "absent.csv"), False)
    with pytest.raises(example.DatasetError, match="parse error"):
        example.LoadCSV(new_index(), write(str(tmp_path), "p.csv", ["0,0,abc,0"]), False)
    with pytest.raises(example.DatasetError, match="read error"):
        example.LoadCSV(new_index(), write(str(tmp_path), "r.csv", ["0,0,0,0", "1,2"]), False)
    with pytest.raises(example.DatasetError, match="failed to add vector 0"):
        example.LoadCSV(new_index(), write(str(tmp_path), "d.csv", ["0,0,0"]), False)


def test_load_training_vectors(tmp_path):
    v = example.LoadTrainingVectors(write_dataset(str(tmp_path)))
    assert v.shape == (6, 4)
    assert np.array_equal(v[4], np.array([10, 10, 10, 11], np.float32))
    os.remove(os.path.join(str(tmp_path), "train.csv"))
    with pytest.raises(example.DatasetError, match="train.csv"):
        example.LoadTrainingVectors(str(tmp_path))


def test_load_test_dataset(tmp_path):
    d = write_dataset(str(tmp_path))
    test, nb, dist = example.LoadTestDataset(d)
    assert test.shape == (2, 4) and nb == [[0, 1], [3, 4]] and dist == [[0.1, 0.9], [0.2, 0.8]]
    write(d, "neighbors.csv", ["0,x"])
    with pytest.raises(example.DatasetError, match="parse error"):
        example.LoadTestDataset(d)
    write(d, "neighbors.csv", ["0,1.5"])
    with pytest.raises(example.DatasetError, match="parse error"):
        example.LoadTestDataset(d)
    write(d, "neighbors.csv", NEIGHBORS)
    write(d, "distances.csv", ["0.1,abc"])
    with pytest.raises(example.DatasetError, match="parse error"):
        example.LoadTestDataset(d)
    # whitespace around values is trimmed
    write(d, "distances.csv", ["0.1, 0.9", " 0.2 ,0.8"])
    assert example.LoadTestDataset(d)[2] == [[0.1, 0.9], [0.2, 0.8]]


def test_float32_values_read_back_exactly(tmp_path):
    """Values printed from float32 (shortest form) read back to the same
    float32, as strconv.ParseFloat(s, 32) gives."""
    rng = np.random.default_rng(1)
    x = (rng.standard_normal((200, 16)) * 10.0 ** rng.integers(-8, 8, (200, 16))).astype(np.float32)
    # str() of a numpy float32 is its shortest round-trip form
    write(str(tmp_path), "train.csv", [",".join(str(v) for v in row) for row in x])
    got = example.LoadTrainingVectors(str(tmp_path))
    assert np.array_equal(got.view(np.uint32), x.view(np.uint32))


def test_format_results():
    N = pqivf.Neighbor
    res = [N(1, 0.5), N(2, 1.25)]
    assert example.FormatResults(res, 2) == "id=1 (dist=0.500) id=2 (dist=1.250) "
    assert example.FormatResults(res, 1) == "id=1 (dist=0.500) "
    assert example.FormatResults(res, 10) == "id=1 (dist=0.500) id=2 (dist=1.250) "
    assert example.FormatResults([], 3) == ""


def test_format_ground_truth():
    assert example.FormatGroundTruth([7, 8], [0.5, 1.25], 2, 2) == "id=7 (dist=0.500) id=8 (dist=1.250) "
    assert example.FormatGroundTruth([7, 8], [0.5, 1.25], 2, 1) == "id=7 (dist=0.500) "
    assert example.FormatGroundTruth([7, 8], [0.5, 1.25], 2, 10) == "id=7 (dist=0.500) id=8 (dist=1.250) "
    assert example.FormatGroundTruth([], [], 2, 2) == ""


def test_recall_at_k():
    """example/utils_test.go:76-141"""
    N = pqivf.Neighbor
    gt = list(range(100))
    assert example.RecallAtK([N(i, 0) for i in range(10)], gt, 10) == 1.0
    assert example.RecallAtK([N(i, 0) for i in range(5)] + [N(1000 + i, 0) for i in range(5, 10)], gt, 10) == 0.5
    assert example.RecallAtK([N(i, 0) for i in range(10)], [0, 1, 2, 3], 10) == 1.0
    assert example.RecallAtK([N(i, 0) for i in range(100)], gt, 100) == 1.0
    assert example.RecallAtK([N(0, 0), N(1, 0)], gt, 10) == 0.2
    assert example.RecallAtK([], gt, 0) == 0.0
    assert example.RecallAtK([], [], 10) == 0.0
