"""Pins of the CPU oracle (oracle/oracle.c) -- CPU only.

Each check is chosen so that a plausible slip in the restatement (a wrong
accumulation order, an unfused tail, a transposed operand, a dropped term)
fails one of them.
"""
import math

import numpy as np
import pytest

from conftest import bits_to_f32

METRICS = [("squared_euclidean", 0), ("euclidean", 1), ("manhattan", 2), ("cosine_distance", 3)]


def split_pairs(g):
    dims = g["pair_dims"]
    off = np.concatenate([[0], np.cumsum(dims)])
    a, b = g["pair_a"], g["pair_b"]
    return [(int(d), a[off[i]:off[i + 1]], b[off[i]:off[i + 1]]) for i, d in enumerate(dims)]


@pytest.mark.parametrize("name,mid", METRICS)
def test_pairs_bit_exact_vs_reference_avx2(oracle, distance_golden, name, mid):
    """Golden bits from the reference's own C (AVX2 tier, AVX for Manhattan)."""
    want = distance_golden[f"pair_{name}_avx2"]
    for i, (d, a, b) in enumerate(split_pairs(distance_golden)):
        got = oracle.pair(mid, a, b)
        assert got.view(np.uint32) == want[i].view(np.uint32), (name, d, i, got, want[i])


def test_tiers_differ_so_the_pin_is_tier_specific(distance_golden):
    """The fallback tier rounds differently at some dims: the golden file
    really distinguishes summation orders."""
    a = distance_golden["pair_squared_euclidean_avx2"].view(np.uint32)
    f = distance_golden["pair_squared_euclidean_fallback"].view(np.uint32)
    assert (a != f).any()


@pytest.mark.parametrize("name,mid", METRICS)
def test_batches_bit_exact_vs_reference(oracle, distance_golden, name, mid):
    for si, (dim, n) in enumerate(distance_golden["batch_shapes"]):
        q = distance_golden[f"batch{si}_q"]
        flat = distance_golden[f"batch{si}_flat"]
        got = oracle.batch(mid, q, flat, int(dim))
        want = distance_golden[f"batch{si}_{name}_avx2"]
        assert np.array_equal(got.view(np.uint64), want.view(np.uint64)), (name, dim, n)


def test_closed_forms(oracle):
    """core/distance_test.go:14-123 (tolerance 1e-6 as there)."""
    cases = [
        ([1, 2, 3, 4, 5, 6], [1, 2, 3, 4, 5, 6], 0, 0, 0, 0),
        ([1, 2, 3, 4, 5, 6], [6, 5, 4, 3, 2, 1], math.sqrt(70), 70, 18, 1 - 56 / 91),
        ([1, 0, 0, 1, 0, 1], [0, 1, 1, 0, 1, 0], math.sqrt(6), 6, 6, 1),
        ([-1, -2, -3, -4, -5, -6], [-6, -5, -4, -3, -2, -1], math.sqrt(70), 70, 18, 1 - 56 / 91),
        ([0, 0, 0, 0, 0, 0], [1, 2, 3, 4, 5, 6], math.sqrt(91), 91, 21, 1),
        ([1, 2, 3, 4, 5, 6], [1, 2, 3, 4, 5, 6.000001], 1e-6, 1e-12, 1e-6, 0),
    ]
    for a, b, eu, sq, man, cos in cases:
        assert abs(oracle.pair(1, a, b) - eu) < 1e-6
        assert abs(oracle.pair(0, a, b) - sq) < 1e-6
        assert abs(oracle.pair(2, a, b) - man) < 1e-6
        assert abs(oracle.pair(3, a, b) - cos) < 1e-6


def test_fp64_differential(oracle):
    """core/distance_test.go:203-227: 1e-3 relative against float64."""
    rng = np.random.default_rng(7)
    for dim in [1, 2, 3, 4, 5, 6, 7, 8, 9, 15, 16, 17, 25, 31, 32, 33, 128, 200, 784]:
        for _ in range(5):
            a = (rng.random(dim) * 20 - 10).astype(np.float32)
            b = (rng.random(dim) * 20 - 10).astype(np.float32)
            a64, b64 = a.astype(np.float64), b.astype(np.float64)
            sq = float(((a64 - b64) ** 2).sum())
            ref = {0: sq, 1: math.sqrt(sq), 2: float(np.abs(a64 - b64).sum()),
                   3: 1 - max(-1, min(1, float(a64 @ b64) / (np.linalg.norm(a64) * np.linalg.norm(b64))))}
            for mid, want in ref.items():
                got = float(oracle.pair(mid, a, b))
                assert abs(got - want) <= 1e-3 * max(1, abs(want)), (mid, dim)


def test_integer_inputs_are_exact(oracle):
    """Integer-valued inputs: every partial sum is an integer below 2^24, so
    the reference order equals exact arithmetic."""
    rng = np.random.default_rng(3)
    for dim in [8, 64, 96, 128]:
        a = rng.integers(0, 256, dim).astype(np.float32)
        b = rng.integers(0, 256, dim).astype(np.float32)
        assert float(oracle.pair(0, a, b)) == float(((a.astype(np.int64) - b.astype(np.int64)) ** 2).sum())
        assert float(oracle.pair(2, a, b)) == float(np.abs(a.astype(np.int64) - b.astype(np.int64)).sum())


def fma32(a, b, c):
    """float32 fused multiply-add: the exact a*b+c rounded once to nearest-even."""
    from fractions import Fraction
    v = Fraction(float(a)) * Fraction(float(b)) + Fraction(float(c))
    f = np.float32(float(v))
    best = f
    for g in (np.nextafter(f, np.float32(np.inf)), np.nextafter(f, np.float32(-np.inf))):
        dg, db = abs(Fraction(float(g)) - v), abs(Fraction(float(best)) - v)
        if dg < db or (dg == db and int(g.view(np.uint32)) % 2 == 0):
            best = g
    return np.float32(best)


def test_s4_and_s8_special_orders(oracle):
    """s=4 is a plain fused chain; s=8 is one cleanup vector reduced by the
    horizontal-sum tree."""
    rng = np.random.default_rng(11)
    for _ in range(200):
        a = rng.standard_normal(4).astype(np.float32)
        b = rng.standard_normal(4).astype(np.float32)
        d = (a - b).astype(np.float32)
        want = np.float32(d[0] * d[0])
        for i in range(1, 4):
            want = fma32(d[i], d[i], want)
        assert oracle.pair(0, a, b) == want
        a8 = rng.standard_normal(8).astype(np.float32)
        b8 = rng.standard_normal(8).astype(np.float32)
        s = ((a8 - b8).astype(np.float32) ** 2).astype(np.float32)
        f = np.float32
        want8 = f(f(f(s[0] + s[4]) + f(s[1] + s[5])) + f(f(s[2] + s[6]) + f(s[3] + s[7])))
        assert oracle.pair(0, a8, b8) == want8


# ----------------------------------------------------------- gob fixture --
def fixture_arrays(g):
    C = bits_to_f32(g["CoarseCentroidsBits"])
    CB = bits_to_f32(g["CodebooksBits"])
    ents = [e for c, es in sorted(g["InvertedLists"].items(), key=lambda x: int(x[0])) for e in es]
    V = {e["ID"]: bits_to_f32(g["VectorBits"][str(e["ID"])]) for e in ents}
    return C, CB, ents, V


def test_fixture_assign_matches_reference(oracle, gob_fixture):
    """Stored Cluster of every entry = or_assign over the stored centroids."""
    C, _, ents, V = fixture_arrays(gob_fixture)
    data = np.stack([V[e["ID"]] for e in ents])
    got = oracle.assign(data, C)
    assert got.tolist() == [e["Cluster"] for e in ents]


def test_fixture_encode_matches_reference(oracle, gob_fixture):
    """Stored Codes of every entry = or_encode of its residual."""
    C, CB, ents, V = fixture_arrays(gob_fixture)
    m = gob_fixture["NumSubquantizers"]
    for e in ents:
        codes = oracle.encode(V[e["ID"]], C[e["Cluster"]], m, CB, [CB.shape[1]] * m)
        assert codes.tolist() == e["Codes"], e["ID"]


def test_fixture_codebooks_are_point_order_means(oracle, gob_fixture):
    """The stored codebooks are bit-for-bit the float32 point-order means the
    update step computes (training data in cluster-ascending list order,
    pqivf/index.go:520-541, assignment = the stored codes)."""
    C, CB, ents, V = fixture_arrays(gob_fixture)
    m = gob_fixture["NumSubquantizers"]
    d = gob_fixture["Dimension"]
    s = d // m
    for i in range(m):
        sub = np.stack([(V[e["ID"]] - C[e["Cluster"]]).astype(np.float32)[i * s:(i + 1) * s] for e in ents])
        assign = np.array([e["Codes"][i] for e in ents], np.int32)
        got, _ = oracle.kmeans_update(sub, assign, CB.shape[1], np.zeros_like(CB[i]), 1)
        assert np.array_equal(got.view(np.uint32), CB[i].view(np.uint32)), i


def test_fixture_coarse_centroids_are_member_means(oracle, gob_fixture):
    """The legacy file was trained on map-ordered data, so the stored coarse
    centroids equal the member means to rounding only (order unknown)."""
    C, _, ents, V = fixture_arrays(gob_fixture)
    data = np.stack([V[e["ID"]] for e in ents])
    assign = np.array([e["Cluster"] for e in ents], np.int32)
    got, _ = oracle.kmeans_update(data, assign, C.shape[0], np.zeros_like(C), 1)
    assert np.max(np.abs(got - C)) <= 2e-6 * max(1.0, float(np.abs(C).max()))


def test_update_is_sequential_float32(oracle):
    """Cluster sums are sequential float32 sums in point order: equal to the
    last element of numpy's float32 cumsum, and (on this data) different
    from a pairwise sum, so the test sees the order."""
    rng = np.random.default_rng(5)
    n, d, k = 3000, 16, 3
    data = (rng.standard_normal((n, d)) * 1000 + 0.1).astype(np.float32)
    assign = rng.integers(0, k, n).astype(np.int32)
    got, _ = oracle.kmeans_update(data, assign, k, np.zeros((k, d), np.float32), 1)
    differs = False
    for c in range(k):
        mem = data[assign == c]
        seq = np.cumsum(mem, axis=0, dtype=np.float32)[-1]
        want = (seq / np.float32(len(mem))).astype(np.float32)
        assert np.array_equal(got[c].view(np.uint32), want.view(np.uint32))
        # numpy reduces a contiguous axis pairwise
        pair = (np.ascontiguousarray(mem.T).sum(axis=1, dtype=np.float32) / np.float32(len(mem))).astype(np.float32)
        differs |= not np.array_equal(pair, want)
    assert differs


def test_empty_cluster_reseeds_from_stream(oracle):
    data = np.arange(40, dtype=np.float32).reshape(10, 4)
    assign = np.zeros(10, np.int32)
    state = 12345
    got, st2 = oracle.kmeans_update(data, assign, 3, np.zeros((3, 4), np.float32), state)
    s = np.uint64(state)
    import ctypes
    st = (ctypes.c_uint64 * 1)(state)
    i1 = oracle.lib().or_rng_bounded(st, 10)
    i2 = oracle.lib().or_rng_bounded(st, 10)
    assert np.array_equal(got[1], data[i1]) and np.array_equal(got[2], data[i2])
    assert st2 == st[0]
    assert np.array_equal(got[0], data.mean(axis=0, dtype=np.float32)) or np.allclose(got[0], data.mean(0))
    del s


def test_rng_perm_is_a_permutation_and_deterministic(oracle):
    p1, s1 = oracle.perm(99, 1000)
    p2, s2 = oracle.perm(99, 1000)
    assert sorted(p1.tolist()) == list(range(1000))
    assert np.array_equal(p1, p2) and s1 == s2
    p3, _ = oracle.perm(100, 1000)
    assert not np.array_equal(p1, p3)


def test_assign_tie_goes_to_lowest_index(oracle):
    """pqivf/index_test.go:1005-1066: duplicate centroids tie, lowest wins."""
    cents = np.array([[1, 1], [0, 0], [0, 0], [5, 5]], np.float32)
    data = np.array([[0, 0], [0.1, 0], [4, 4]], np.float32)
    assert oracle.assign(data, cents).tolist() == [1, 1, 3]


def test_assign_matches_fp64_brute_force_when_separated(oracle):
    rng = np.random.default_rng(8)
    cents = (rng.standard_normal((64, 32)) * 10).astype(np.float32)
    data = (cents[rng.integers(0, 64, 2000)] + rng.standard_normal((2000, 32)) * 0.5).astype(np.float32)
    d64 = ((data[:, None, :].astype(np.float64) - cents[None].astype(np.float64)) ** 2).sum(-1)
    srt = np.sort(d64, axis=1)
    ok = (srt[:, 1] - srt[:, 0]) > 1e-3 * srt[:, 1]
    got = oracle.assign(data, cents, nthreads=4)
    assert np.array_equal(got[ok], d64.argmin(1)[ok])


def test_batch_topk_matches_fp64_sort(oracle):
    rng = np.random.default_rng(9)
    base = rng.integers(0, 256, (3000, 24)).astype(np.float32)
    q = rng.integers(0, 256, (5, 24)).astype(np.float32)
    ids, rank = oracle.batch_topk(0, q, base, 17, nthreads=2)
    d = ((q[:, None, :].astype(np.int64) - base[None].astype(np.int64)) ** 2).sum(-1)
    for i in range(5):
        want = np.lexsort((np.arange(3000), d[i]))[:17]
        assert ids[i].tolist() == want.tolist()
        assert np.array_equal(rank[i], d[i][want].astype(np.float32))


def test_search_matches_fp64_on_reconstruction(oracle):
    """or_search ranks by the float32 reconstruction c + cb (decodePQCode +
    vectorAdd); its order matches FP64 distances to the same reconstruction
    except where they are within float32 rounding."""
    rng = np.random.default_rng(10)
    n, d = 2000, 16
    cent = (rng.random((8, d)) * 20 - 10).astype(np.float32)
    x = (cent[rng.integers(0, 8, n)] + rng.standard_normal((n, d))).astype(np.float32)
    ids = np.arange(n, dtype=np.int64) * 3 + 1
    ix = oracle.train(ids, x, 8, 4, 16, 5, seed=1)
    q = (x[:20] + rng.standard_normal((20, d)).astype(np.float32) * 0.1).astype(np.float32)
    got_ids, got_d, cnt = oracle.search_many(ix, q, 10, 3)
    s = d // 4
    recon = np.concatenate([ix.codebooks[i][ix.codes[:, i]] for i in range(4)], axis=1)
    cl = np.repeat(np.arange(len(ix.list_off) - 1), np.diff(ix.list_off))
    recon = (ix.coarse[cl] + recon).astype(np.float32)
    for qi in range(20):
        assert cnt[qi] == 10
        dq = np.sqrt(((recon.astype(np.float64) - q[qi].astype(np.float64)) ** 2).sum(1))
        pos = {int(i): j for j, i in enumerate(ix.ids)}
        for r in range(10):
            j = pos[int(got_ids[qi, r])]
            assert abs(dq[j] - got_d[qi, r]) <= 1e-5 * max(1, dq[j])
        assert np.all(np.diff(got_d[qi]) >= 0)
    del s


def test_train_lists_in_ascending_id_order(oracle):
    """pqivf/index_test.go:1072-1097"""
    rng = np.random.default_rng(12)
    ids = rng.permutation(500).astype(np.int64)
    x = rng.standard_normal((500, 8)).astype(np.float32)
    ix = oracle.train(ids, x, 4, 2, 8, 3, seed=2)
    for c in range(len(ix.list_off) - 1):
        seg = ix.ids[ix.list_off[c]:ix.list_off[c + 1]]
        assert np.all(np.diff(seg) > 0)


# ------------------------------------------- k-means++ (build extension) --
def _pp_draws(oracle, state, n, k):
    """The seeding's draws from one stream: the first row, then k-1 uniforms."""
    import ctypes
    st = (ctypes.c_uint64 * 1)(state)
    lib = oracle.lib()
    first = lib.or_rng_bounded(st, n)
    u = [(lib.or_rng_next(st) >> 11) * (1.0 / 9007199254740992.0) for _ in range(k - 1)]
    return first, u


def test_kmeanspp_seed_follows_d2_sampling(oracle):
    """Seeds are data rows: the first is the stream's bounded draw, each
    next one the row where the running D^2 sum passes u * total (the rows'
    D^2 values are spread far apart here, so rounding cannot move the pick)."""
    rng = np.random.default_rng(4)
    x = (rng.standard_normal((200, 3)) * np.array([100.0, 1.0, 0.01])).astype(np.float32)
    k, state = 6, 12345
    seeds, _ = oracle.kmeans_init(x, k, 0, state, init=1)
    first, u = _pp_draws(oracle, state, x.shape[0], k)
    assert np.array_equal(seeds[0], x[first])
    chosen = [x[first].astype(np.float64)]
    for j in range(1, k):
        d2 = np.min([((x.astype(np.float64) - c) ** 2).sum(1) for c in chosen], axis=0)
        cum = np.cumsum(d2)
        r = int(np.searchsorted(cum, u[j - 1] * cum[-1], side="right"))
        # the pick lies within one row of the exact crossing (float32 D^2)
        rows = {max(r - 1, 0), r, min(r + 1, len(x) - 1)}
        assert any(np.array_equal(seeds[j], x[i]) for i in rows), j
        chosen.append(seeds[j].astype(np.float64))


def test_kmeanspp_seeds_distinct_and_deterministic(oracle):
    rng = np.random.default_rng(5)
    cen = rng.uniform(-20, 20, (8, 16)).astype(np.float32)
    x = (cen[rng.integers(0, 8, 3000)] + rng.standard_normal((3000, 16))).astype(np.float32)
    a, _ = oracle.kmeans_init(x, 32, 0, 99, init=1)
    b, _ = oracle.kmeans_init(x, 32, 0, 99, init=1)
    assert np.array_equal(a, b)
    assert len({r.tobytes() for r in a}) == 32
    # every one of the 8 separated clusters gets a seed
    near = np.argmin(((a[:, None, :] - cen[None]) ** 2).sum(-1), axis=1)
    assert set(near.tolist()) == set(range(8))


def split_norm(g):
    off = np.concatenate([[0], np.cumsum(g["lens"])])
    return [(int(off[i]), int(off[i + 1])) for i in range(len(g["lens"]))]


def test_normalize_bit_exact_vs_reference_avx2(oracle):
    """core/simd_kernels.inc.h:267-315: golden bits from the reference's own
    C (normalize_avx2) for lengths 1..1,000, with tails of every size,
    zero vectors (left as they are) and very small and large values."""
    import os
    from conftest import GOLDEN
    g = np.load(os.path.join(GOLDEN, "normalize_golden.npz"))
    for a, b in split_norm(g):
        got = oracle.normalize(g["inputs"][a:b])
        assert np.array_equal(got.view(np.uint32), g["out_avx2"][a:b].view(np.uint32)), (b - a)
    # the scalar fallback tier sums in another order: the pin is tier-specific
    assert not np.array_equal(g["out_avx2"].view(np.uint32), g["out_fallback"].view(np.uint32))
    # zero vectors stay zero; others have unit norm to FP32 accuracy
    for a, b in split_norm(g):
        v = g["out_avx2"][a:b].astype(np.float64)
        if np.all(g["inputs"][a:b] == 0):
            assert np.all(v == 0)
        else:
            assert abs(np.sqrt((v * v).sum()) - 1.0) < 1e-5


# ------------------------------------------- build-owned random stream --
def test_splitmix64_known_answers(oracle):
    """The stream is splitmix64 (Steele, Lea and Flood, OOPSLA 2014): its
    published test sequences from state 0 and from state 1234567."""
    import ctypes
    lib = oracle.lib()
    st = (ctypes.c_uint64 * 1)(0)
    assert [lib.or_rng_next(st) for _ in range(3)] == [0xE220A8397B1DCDAF, 0x6E789E6AA1B965F4, 0x06C45D188009454F]
    st = (ctypes.c_uint64 * 1)(1234567)
    assert [lib.or_rng_next(st) for _ in range(5)] == [6457827717110365317, 3203168211198807973,
                                                       9817491932198370423, 4593380528125082431,
                                                       16408922859458223821]


def _chi2(counts, expected):
    return float((((counts - expected) ** 2) / expected).sum())


@pytest.mark.parametrize("n,draws", [(2, 40000), (3, 60000), (7, 70000), (10, 100000), (1000, 200000)])
def test_rng_bounded_uniform(oracle, n, draws):
    """Bounded draws stay in [0, n) and are uniform (chi-squared below the
    mean plus six standard deviations of its distribution)."""
    import ctypes
    lib = oracle.lib()
    st = (ctypes.c_uint64 * 1)(n * 7919)
    v = np.array([lib.or_rng_bounded(st, n) for _ in range(draws)], np.int64)
    assert v.min() >= 0 and v.max() < n
    c = np.bincount(v, minlength=n).astype(np.float64)
    dof = n - 1
    assert _chi2(c, draws / n) < dof + 6 * np.sqrt(2 * dof)


def test_rng_perm_uniform_over_all_permutations(oracle):
    """rand.Perm's inside-out shuffle (pqivf/index.go:805) gives every one of
    the 24 permutations of 4 items equally often. Drawing j from [0, i)
    instead of [0, i] (Sattolo's cycles) reaches only 6 of them, and a
    biased bound skews the counts: both fail the chi-squared bound."""
    import itertools
    perms = {p: i for i, p in enumerate(itertools.permutations(range(4)))}
    counts = np.zeros(24)
    state = 42
    for _ in range(48000):
        p, state = oracle.perm(state, 4)
        counts[perms[tuple(p.tolist())]] += 1
    assert np.count_nonzero(counts) == 24
    assert _chi2(counts, 2000.0) < 23 + 6 * np.sqrt(46)
    # position marginals at n = 10: P(perm[i] = v) = 1/10
    marg = np.zeros((10, 10))
    for _ in range(20000):
        p, state = oracle.perm(state, 10)
        marg[np.arange(10), p] += 1
    assert _chi2(marg.ravel(), 2000.0) < 99 + 6 * np.sqrt(198)


def test_kmeans_seeds_are_perm_rows(oracle):
    """runKMeans seeds centroid i with data[perm[i]] for the first k entries
    of one permutation of the rows (pqivf/index.go:804-810): with no Lloyd
    iteration the result is exactly those rows."""
    rng = np.random.default_rng(31)
    data = rng.standard_normal((500, 6)).astype(np.float32)
    state = oracle.stream_seed(9, 0, 0)
    got, st_after = oracle.kmeans(data, 37, 0, state)
    p, st_perm = oracle.perm(state, 500)
    assert np.array_equal(got, data[p[:37]])
    assert st_after == st_perm
    # k > n: every row, in permutation order
    got, _ = oracle.kmeans(data[:20], 37, 0, state)
    p, _ = oracle.perm(state, 20)
    assert np.array_equal(got, data[:20][p])


def test_train_stream_layout(oracle):
    """Train draws the coarse k-means from stream role 0 and subquantizer i's
    k-means from role 1 + i of (seed, epoch), over the residuals in list
    order (pqivf/index.go:520-551 builds dataPerSub that way)."""
    rng = np.random.default_rng(32)
    x = (rng.uniform(-10, 10, (6, 8))[rng.integers(0, 6, 600)] + rng.standard_normal((600, 8))).astype(np.float32)
    ids = np.arange(600, dtype=np.int64)
    ix = oracle.train(ids, x, 6, 2, 8, 4, seed=77, epoch=3)
    coarse, _ = oracle.kmeans(x, 6, 4, oracle.stream_seed(77, 3, 0))
    assert np.array_equal(ix.coarse, coarse)
    order = np.argsort(ix.assign, kind="stable")
    res = x[order] - coarse[ix.assign[order]]
    for i in range(2):
        cb, _ = oracle.kmeans(np.ascontiguousarray(res[:, 4 * i:4 * i + 4]), 8, 4, oracle.stream_seed(77, 3, 1 + i))
        assert np.array_equal(ix.codebooks[i, :len(cb)], cb)
    # distinct roles and epochs give distinct streams
    seeds = {oracle.stream_seed(77, e, r) for e in range(4) for r in range(40)}
    assert len(seeds) == 160


# ------------------------------------------------ search + exact rerank --
def _rerank_fixture(oracle, n=3000, d=16, K=12, m=4, P=32, seed=3):
    rng = np.random.default_rng(seed)
    cent = (rng.random((K, d)) * 20 - 10).astype(np.float32)
    x = (cent[rng.integers(0, K, n)] + rng.standard_normal((n, d))).astype(np.float32)
    ids = rng.permutation(10 * n)[:n].astype(np.int64)
    ix = oracle.train(ids, x, K, m, P, 5, seed=seed)
    q = (x[rng.integers(0, n, 40)] + rng.standard_normal((40, d)).astype(np.float32) * 0.3).astype(np.float32)
    return ix, q


def _probed(ix, q, nprobe):
    d64 = ((ix.coarse.astype(np.float64) - q.astype(np.float64)) ** 2).sum(1)
    return np.argsort(d64, kind="stable")[:nprobe]


def _pq_ranks64(ix, q, entries):
    """FP64 distance from q to the float32 reconstruction of each entry."""
    lists = np.repeat(np.arange(len(ix.list_off) - 1), np.diff(ix.list_off))
    m = ix.m
    s = q.size // m
    rec = np.concatenate([ix.codebooks[i][ix.codes[entries, i]] for i in range(m)], axis=1)
    rec = (ix.coarse[lists[entries]] + rec).astype(np.float32)
    return ((rec.astype(np.float64) - q.astype(np.float64)) ** 2).sum(1)


def test_search_rerank_all_candidates_is_exact_topk(oracle):
    """When rerank * k covers every probed entry the result is the exact
    top-k of the probed lists' stored vectors: FP64 brute force, ids equal
    except at near-ties, distances the reference-order FP32 ones."""
    ix, qs = _rerank_fixture(oracle)
    for qi, q in enumerate(qs):
        probed = _probed(ix, q, 2)
        ents = np.concatenate([np.arange(ix.list_off[c], ix.list_off[c + 1]) for c in probed])
        k = 10
        R = (len(ents) + k - 1) // k + 1
        ids, dist, cnt, bound = oracle.search_rerank_many(ix, q[None], k, 2, R)
        assert cnt[0] == k and np.isinf(bound[0])
        d64 = ((ix.vecs[ents].astype(np.float64) - q.astype(np.float64)) ** 2).sum(1)
        order = np.lexsort((ix.ids[ents], d64))[:k]
        want = ix.ids[ents][order]
        tie = np.sort(d64)
        for r in range(k):
            if ids[0, r] != want[r]:
                assert abs(tie[r] - tie[min(r + 1, len(tie) - 1)]) < 1e-5 * (1 + tie[r]), qi
        for r in range(k):
            e = ents[ix.ids[ents] == ids[0, r]][0]
            assert dist[0, r] == np.sqrt(np.float64(oracle.pair(0, q, ix.vecs[e])))


@pytest.mark.parametrize("k,R,nprobe", [(5, 1, 3), (10, 3, 2), (20, 2, 4), (7, 4, 12)])
def test_search_rerank_candidates_are_pq_topk(oracle, k, R, nprobe):
    """The candidates are the R*k best probed entries by (PQ rank, storage
    position) -- small codebooks make many code rows identical, so exact PQ
    ties are common and the position decides them -- and the result is the
    best k of those by (the stored vector's rank, id). PQ ranks are the
    reference-order FP32 distance to the float32 reconstruction, within
    rounding of the FP64 one; R = 1 returns the set Search returns."""
    ix, qs = _rerank_fixture(oracle, seed=4)
    ids, dist, cnt, bound = oracle.search_rerank_many(ix, qs, k, nprobe, R)
    pid, _, _ = oracle.search_many(ix, qs, k, nprobe)
    lists = np.repeat(np.arange(len(ix.list_off) - 1), np.diff(ix.list_off))
    for qi, q in enumerate(qs):
        probed = _probed(ix, q, nprobe)
        ents = np.concatenate([np.arange(ix.list_off[c], ix.list_off[c + 1]) for c in probed])
        rec = np.concatenate([ix.codebooks[i][ix.codes[ents, i]] for i in range(ix.m)], axis=1)
        rec = (ix.coarse[lists[ents]] + rec).astype(np.float32)
        pq32 = np.array([oracle.pair(0, q, r) for r in rec], np.float32)
        pq64 = ((rec.astype(np.float64) - q.astype(np.float64)) ** 2).sum(1)
        assert np.all(np.abs(pq32 - pq64) <= 1e-5 * (1 + pq64))
        sel = min(k * R, len(ents))
        cand = ents[np.lexsort((ents, pq32))[:sel]]
        if k * R < len(ents):
            assert bound[qi] == np.sort(pq32)[sel - 1]
        else:
            assert np.isinf(bound[qi])
        ex = np.array([oracle.pair(0, q, ix.vecs[e]) for e in cand], np.float32)
        best = np.lexsort((ix.ids[cand], ex))[:k]
        assert cnt[qi] == min(k, sel)
        assert ids[qi, :cnt[qi]].tolist() == ix.ids[cand][best].tolist()
        assert np.array_equal(dist[qi, :cnt[qi]], np.sqrt(ex[best].astype(np.float64)))
        if R == 1:
            assert set(ids[qi, :cnt[qi]].tolist()) == set(pid[qi, :cnt[qi]].tolist())


def test_search_rerank_fallback_rule(oracle):
    """Fewer than k probed entries: every list is searched
    (pqivf/index.go:892-899) and its rerank candidates come from all lists."""
    ix, qs = _rerank_fixture(oracle, n=400, K=12, seed=5)
    smallest = int(np.argmin(np.diff(ix.list_off)))
    q = ix.coarse[smallest][None] + np.float32(0.001)
    k = int(np.diff(ix.list_off)[smallest]) + 5
    ids, dist, cnt, bound = oracle.search_rerank_many(ix, q, k, 1, 2)
    assert cnt[0] == k
    want = np.lexsort((ix.ids, ((ix.vecs.astype(np.float64) - q[0].astype(np.float64)) ** 2).sum(1)))
    # the R * k candidates are drawn from every list, so the nearest stored
    # vectors overall are found
    assert set(ids[0, :5].tolist()) <= set(ix.ids[want[:2 * k]].tolist())
    ids2, _, cnt2, _ = oracle.search_rerank_many(ix, q, k, 1, 2, allow_fallback=False)
    assert cnt2[0] == k - 5
