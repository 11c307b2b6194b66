"""hann_batch_topk_noalloc_dev: the exact batch top-k without allocating.

Through the C-ABI, on device buffers: two calls in a row (the second with
fewer queries, rows and a smaller k) give exactly the ids and distances of
hann_batch_topk_dev with HANN_PREC_FP32; neither call changes the device's
free memory; a size beyond the scratch is refused with an error.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _run(lib, fn, metric, q, base, k, stream, *extra):
    import torch
    nq = q.shape[0]
    ids = torch.full((nq, k), -7, dtype=torch.int64, device="cuda")
    dist = torch.full((nq, k), -7.0, dtype=torch.float64, device="cuda")
    rc = getattr(lib, fn)(metric, q.data_ptr(), nq, base.data_ptr(), base.shape[0], q.shape[1], k, *extra,
                          ids.data_ptr(), dist.data_ptr(), stream)
    return rc, ids, dist


@pytest.mark.parametrize("metric", [0, 1, 2, 3])
def test_noalloc_topk_equals_allocating_entry(hann, metric):
    import torch
    from hann_amd import _lib
    lib = _lib.load()
    rng = np.random.default_rng(40 + metric)
    stream = torch.cuda.Stream()
    shapes = [(3000, 200_000, 96, 100), (700, 50_000, 40, 17)]   # second call: smaller
    data = []
    for nq, nb, d, k in shapes:
        q = torch.from_numpy(rng.integers(0, 64, (nq, d)).astype(np.float32) * np.float32(0.25)).cuda()
        base = torch.from_numpy(rng.integers(0, 64, (nb, d)).astype(np.float32) * np.float32(0.25)).cuda()
        rc, wi, wd = _run(lib, "hann_batch_topk_dev", metric, q, base, k, stream.cuda_stream, _lib.PREC_FP32, 0)
        _lib.check(rc)
        data.append((q, base, k, wi, wd))
    torch.cuda.synchronize()
    free0 = torch.cuda.mem_get_info()[0]
    outs = []
    for q, base, k, _, _ in data:
        rc, gi, gd = _run(lib, "hann_batch_topk_noalloc_dev", metric, q, base, k, stream.cuda_stream)
        _lib.check(rc)
        outs.append((gi, gd))
    torch.cuda.synchronize()
    assert torch.cuda.mem_get_info()[0] == free0
    for (q, base, k, wi, wd), (gi, gd) in zip(data, outs):
        assert torch.equal(gi, wi)
        assert np.array_equal(gd.cpu().numpy().view(np.uint64), wd.cpu().numpy().view(np.uint64))
    # the null stream: the call waits
    q, base, k, wi, wd = data[1]
    rc, gi, gd = _run(lib, "hann_batch_topk_noalloc_dev", metric, q, base, k, None)
    _lib.check(rc)
    assert torch.equal(gi, wi) and torch.equal(gd, wd)


def test_noalloc_topk_rejects_oversized(hann):
    import torch
    from hann_amd import _lib
    lib = _lib.load()
    nq, k = 30000, 100          # 36 MB of partial lists: above the 32 MiB scratch
    q = torch.zeros((nq, 8), dtype=torch.float32, device="cuda")
    base = torch.zeros((1000, 8), dtype=torch.float32, device="cuda")
    rc, _, _ = _run(lib, "hann_batch_topk_noalloc_dev", 0, q, base, k, None)
    assert rc == _lib.HANN_ERR_INVALID_ARG
    assert b"exceeds its scratch" in lib.hann_last_error()
    # at the limit it runs
    nq = (32 << 20) // (12 * k)
    q = torch.zeros((nq, 8), dtype=torch.float32, device="cuda")
    rc, ids, _ = _run(lib, "hann_batch_topk_noalloc_dev", 0, q, base, k, None)
    _lib.check(rc)
    assert torch.equal(ids[0].cpu(), torch.arange(100))
