"""Pins for the CPU oracle (no GPU needed).

Each check is independent of the oracle's own code: published known-answer
vectors, the reference's own known answers and checks, numpy routines
(cumsum, fancy indexing, stable argsort, bincount), the reference's compiled
OpenMP functions (oracle/_ref, when built) and brute force on tiny inputs.
"""
import itertools

import numpy as np
import pytest

# SplitMix64 seeded with 1234567: the published Rosetta Code known-answer vector
SPLITMIX_1234567 = [6457827717110365317, 3203168211198807973, 9817491932198370423,
                    4593380528125082431, 16408922859458223821]


def test_splitmix_known_answer(orc):
    assert [int(x) for x in orc.splitmix_seq(1234567, 5)] == SPLITMIX_1234567
    # the counter form used by every generator is the same stream
    L = orc.oracle()
    assert [L.oracle_splitmix_at(1234567, i) for i in range(5)] == SPLITMIX_1234567


def test_gen_values_follow_splitmix(orc):
    seq = orc.splitmix_seq(99, 1000)
    assert np.array_equal(orc.gen_i32(1000, 99).view(np.uint32), (seq >> np.uint64(32)).astype(np.uint32))
    assert np.array_equal(orc.gen_i32(1000, 99, 15), ((seq >> np.uint64(32)) & np.uint64(15)).astype(np.int32))
    assert np.array_equal(orc.gen_i64(1000, 99).view(np.uint64), seq)


@pytest.mark.parametrize("n", [1, 2, 3, 5, 17, 4095, 4096, 4097, 100_000, 1 << 18])
@pytest.mark.parametrize("kind", [0, 1])
def test_permutations_are_permutations(orc, n, kind):
    p = orc.gen_perm(n, 7, kind)
    assert np.array_equal(np.sort(p), np.arange(n, dtype=np.int32))
    if kind == 1:
        full = (n // 4096) * 4096
        blocks = p[:full].reshape(-1, 4096) if full else np.zeros((0, 4096), np.int32)
        # every output block reads exactly one source block, all of it
        for row in blocks:
            assert np.unique(row // 4096).size == 1
            assert np.array_equal(np.sort(row % 4096), np.arange(4096))
        assert np.all(p[full:] >= full)


def test_random_permutation_looks_random(orc):
    n = 1 << 16
    p = orc.gen_perm(n, 11, 0).astype(np.int64)
    # fixed points ~ Poisson(1); displacement spread like a uniform shuffle
    assert np.sum(p == np.arange(n)) < 10
    assert abs(np.mean(np.abs(p - np.arange(n))) / n - 1 / 3) < 0.01
    assert orc.gen_perm(n, 12, 0).tolist() != p.tolist()


def test_zipf_thresholds(orc):
    for B in [1, 2, 16, 256, 4096]:
        t = orc.zipf_thresholds(B).astype(np.float64)
        assert t[-1] == 2.0 ** 32 and np.all(np.diff(t) >= 0)
        h = sum(1.0 / r for r in range(1, B + 1))
        first = 2.0 ** 32 / h
        assert abs(t[0] - first) <= 1.0


@pytest.mark.parametrize("B", [2, 16, 256, 4096])
@pytest.mark.parametrize("dist", [0, 1])
def test_split_key_buckets(orc, B, dist):
    n = 1 << 18
    keys, vals = orc.gen_split_keys(n, B, dist, 5)
    b = keys & (B - 1)
    counts = np.bincount(b, minlength=B) / n
    if dist == 0:
        assert np.allclose(counts, 1.0 / B, atol=6 * np.sqrt(1.0 / B / n))
    else:
        h = sum(1.0 / r for r in range(1, B + 1))
        expect = 1.0 / np.arange(1, B + 1) / h
        assert np.allclose(counts, expect, atol=6 * np.sqrt(expect / n) + 1e-6)
    assert np.array_equal(vals, np.arange(n, dtype=np.int32))
    k64, _ = orc.gen_split_keys(n, B, dist, 5, wide=True)
    assert np.array_equal(k64 & (B - 1), b.astype(np.int64))


# ---------------------------------------------------------------- scan
def test_scan_all_ones_known_answer(orc):
    # cuda/test_scan.cu:53-58 and openmp/test_scan_CPU.cpp:174-177 + :39-49
    for n in [1, 1024, 1 << 20]:
        out, tot = orc.scan_exclusive(np.ones(n, np.int32))
        assert np.array_equal(out, np.arange(n, dtype=np.int32)) and tot == n


@pytest.mark.parametrize("dtype", [np.int32, np.int64])
def test_scan_matches_numpy_cumsum(orc, dtype):
    rng = np.random.default_rng(3)
    info = np.iinfo(dtype)
    x = rng.integers(info.min, info.max, size=100_003, dtype=dtype)  # wraps many times
    out, tot = orc.scan_exclusive(x)
    u = x.view(np.uint32 if dtype == np.int32 else np.uint64)
    with np.errstate(over="ignore"):
        inc = np.cumsum(u, dtype=u.dtype)
    assert np.array_equal(out.view(u.dtype)[1:], inc[:-1]) and out[0] == 0
    assert tot == inc[-1].view(dtype)


def test_scan_overflow_at_metric_size_is_wraparound(orc):
    # 2^28 values in [0,15] sum past 2^31: the total must wrap mod 2^32
    x = orc.gen_i32(1 << 22, 1, 15)
    _, tot = orc.scan_exclusive(np.tile(x, 64))
    assert int(tot) % (1 << 32) == (int(x.astype(np.int64).sum()) * 64) % (1 << 32)


def test_scan_linearity(orc):
    a = orc.gen_i32(5000, 1)
    b = orc.gen_i32(5000, 2)
    sa, _ = orc.scan_exclusive(a)
    sb, _ = orc.scan_exclusive(b)
    sab, _ = orc.scan_exclusive((a.view(np.uint32) + b.view(np.uint32)).view(np.int32))
    assert np.array_equal(sab.view(np.uint32), sa.view(np.uint32) + sb.view(np.uint32))


# ---------------------------------------------------------------- gather / scatter
@pytest.mark.parametrize("passes", [1, 2, 3, 7, 16])
def test_gather_scatter_known_answers(orc, passes):
    # opencl/test/test_gather.cpp:25,64-70 and test_scatter.cpp:63-69: in[i] = i
    n = 10_007
    idx = orc.gen_perm(n, 3, 0)
    iota = np.arange(n, dtype=np.int32)
    assert np.array_equal(orc.gather(iota, idx, passes), idx)
    inv = np.empty(n, np.int32)
    inv[idx] = iota
    assert np.array_equal(orc.scatter(iota, idx, passes), inv)


@pytest.mark.parametrize("wide", [False, True])
def test_gather_scatter_vs_numpy(orc, wide):
    n = 4099
    vals = orc.gen_i64(n, 8) if wide else orc.gen_i32(n, 8)
    for kind in [0, 1]:
        idx = orc.gen_perm(n, 4, kind)
        for p in [1, 2, 5, 16]:
            assert np.array_equal(orc.gather(vals, idx, p), vals[idx])
            ref = np.empty_like(vals)
            ref[idx] = vals
            assert np.array_equal(orc.scatter(vals, idx, p), ref)


def test_multipass_slices_brute_force(orc):
    # with a non-permutation index the pass structure decides which slots are
    # written: check the slice arithmetic of gatherImpl.cpp:37-45 by brute force
    n = 10
    idx = np.array([9, 0, 4, 4, 3, 8, 1, 5, 2, 7], np.int32)
    src = np.arange(100, 110, dtype=np.int32)
    for p in range(1, 12):
        length = -(-n // p)
        exp = np.full(n, -1, np.int32)
        for k in range(p):
            for i in range(n):
                if k * length <= idx[i] < (k + 1) * length:
                    exp[i] = src[idx[i]]
        assert np.array_equal(orc.gather(src, idx, p, out=np.full(n, -1, np.int32)), exp)


def test_reference_gather_scatter_agree(orc):
    ref = orc.reference()
    if ref is None:
        pytest.skip("reference OpenMP library not built (no /root/reference here)")
    n = 100_000
    src = orc.gen_i32(n, 21)
    idx = orc.gen_perm(n, 22, 0)
    out = np.zeros(n, np.int32)
    ref.edpp_ref_gather(orc._a(src), orc._a(out), orc._a(idx), n)
    assert np.array_equal(out, orc.gather(src, idx, 1))
    out2 = np.zeros(n, np.int32)
    ref.edpp_ref_scatter(orc._a(src), orc._a(out2), orc._a(idx), n)
    assert np.array_equal(out2, orc.scatter(src, idx, 1))


# ---------------------------------------------------------------- split
@pytest.mark.parametrize("B", [1, 2, 16, 256, 4096])
@pytest.mark.parametrize("dist", [0, 1])
def test_split_vs_numpy_stable_argsort(orc, B, dist):
    n = 50_001
    keys, vals = orc.gen_split_keys(n, B, dist, 9)
    ko, vo, start = orc.split(keys, B, vals)
    order = np.argsort(keys & (B - 1), kind="stable")
    assert np.array_equal(ko, keys[order]) and np.array_equal(vo, vals[order])
    cnt = np.bincount(keys & (B - 1), minlength=B)
    assert np.array_equal(start, np.concatenate([[0], np.cumsum(cnt)[:-1]]).astype(np.int32))


def test_split_reference_checks(orc):
    # opencl/test/test_split.cpp:164-203: bucket ids non-decreasing, sum kept;
    # cuda/test_split.cu:196-204: key == bucket -> output equals sorted input
    n, B = 30_000, 32
    keys = (orc.gen_i32(n, 4) & (B - 1)).astype(np.int32)
    ko, _, _ = orc.split(keys, B)
    assert np.all(np.diff(ko & (B - 1)) >= 0) and (ko & (B - 1)).sum() == (keys & (B - 1)).sum()
    assert np.array_equal(ko, np.sort(keys))


def test_split_brute_force_tiny(orc):
    for n in range(0, 7):
        for keys in itertools.product(range(4), repeat=min(n, 4)):
            k = np.array(keys, np.int32)
            ko, vo, start = orc.split(k, 4, np.arange(k.size, dtype=np.int32))
            exp = sorted(range(k.size), key=lambda i: (k[i], i))
            assert vo.tolist() == exp and ko.tolist() == [int(k[i]) for i in exp]
            assert start.tolist() == [int(np.sum(k < b)) for b in range(4)]


def test_split_int64_keys_use_low_bits(orc):
    n, B = 10_000, 256
    keys, vals = orc.gen_split_keys(n, B, 1, 2, wide=True)
    ko, vo, start = orc.split(keys, B, vals)
    order = np.argsort(keys & (B - 1), kind="stable")
    assert np.array_equal(ko, keys[order]) and np.array_equal(vo, vals[order])


# ---------------------------------------------------------------- STREAM
def test_stream_formulas(orc):
    a = orc.gen_i32(1000, 1)
    b = orc.gen_i32(1000, 2)
    ua, ub = a.view(np.uint32), b.view(np.uint32)
    assert np.array_equal(orc.stream(0, a), a)
    assert np.array_equal(orc.stream(1, a).view(np.uint32), ua * np.uint32(3))
    assert np.array_equal(orc.stream(2, a, b).view(np.uint32), ua + ub)
    assert np.array_equal(orc.stream(3, a, b).view(np.uint32), ua + np.uint32(3) * ub)


def test_reference_stream_agrees(orc):
    ref = orc.reference()
    if ref is None:
        pytest.skip("reference OpenMP library not built")
    a = orc.gen_i32(1 << 16, 3)
    out = np.zeros_like(a)
    ref.edpp_ref_scale(orc._a(a), orc._a(out), a.size)
    assert np.array_equal(out, orc.stream(1, a))


# ---------------------------------------------------------------- statistics
@pytest.mark.parametrize("x,expect", [
    ([1.0, 1.0, 1.0, 1.0, 100.0], 1.0),         # MAD = 0: only samples equal to the median
    ([1.0, 2.0, 3.0, 4.0, 5.0], 3.0),           # nothing rejected
    ([10.0, 10.5, 9.5, 10.2, 30.0], 10.05),     # median 10.2, MAD 0.5/0.6745: 30 rejected
    ([2.0, 4.0], 3.0),
])
def test_hampel_hand_computed(orc, x, expect):
    assert orc.hampel_mean(np.array(x)) == pytest.approx(expect)


def test_hampel_matches_reference(orc):
    ref = orc.reference()
    if ref is None:
        pytest.skip("reference OpenMP library not built")
    rng = np.random.default_rng(0)
    for _ in range(50):
        x = np.ascontiguousarray(rng.gamma(2.0, 1.0, size=int(rng.integers(1, 12))))
        if rng.random() < 0.3:
            x[0] = 50.0
        # the reference sums kept samples in order of deviation, the oracle in
        # input order: equal up to rounding of the final sum
        assert orc.hampel_mean(x) == pytest.approx(ref.edpp_ref_average_hampel(orc._a(x.copy()), x.size), rel=1e-12)


# ---------------------------------------------------------------- radix sort
def _twiddled_bits(k, begin, end, signed):
    u = k.view(np.uint32).astype(np.uint64)
    if signed:
        u ^= np.uint64(0x80000000)
    return (u >> np.uint64(begin)) & np.uint64((1 << (end - begin)) - 1)


@pytest.mark.parametrize("signed", [False, True])
@pytest.mark.parametrize("bits", [(0, 32), (0, 8), (4, 20), (8, 32), (31, 32), (5, 5)])
def test_sort_vs_numpy_stable_argsort(orc, signed, bits):
    rng = np.random.default_rng(7 + bits[0] + 40 * bits[1])
    n = 20_000
    keys = rng.integers(0, 2**32, n, dtype=np.uint64).astype(np.uint32)
    keys[: n // 10] = keys[n // 10: n // 5]  # duplicates make stability observable
    if signed:
        keys = keys.view(np.int32)
    vals = np.arange(n, dtype=np.int32)
    ko, vo = orc.sort(keys, vals, bits[0], bits[1], signed)
    order = np.argsort(_twiddled_bits(keys, bits[0], bits[1], signed), kind="stable")
    assert np.array_equal(vo, order)
    assert np.array_equal(ko, keys[order])


def test_sort_full_key_orders_like_std_sort(orc):
    # the reference checks its radix sort against std::sort of the keys
    # (obsolete/CUDA/src/testRadixSort.cpp:57-60)
    rng = np.random.default_rng(3)
    k = rng.integers(-2**31, 2**31, 50_000, dtype=np.int64).astype(np.int32)
    assert np.array_equal(orc.sort(k, signed=True)[0], np.sort(k))
    assert np.array_equal(orc.sort(k.view(np.uint32))[0], np.sort(k.view(np.uint32)))


def test_sort_brute_force_tiny(orc):
    rng = np.random.default_rng(11)
    for n in range(0, 9):
        k = rng.integers(0, 8, n).astype(np.uint32)
        v = np.arange(n, dtype=np.int32)
        ko, vo = orc.sort(k, v, 0, 3)
        ref = sorted(range(n), key=lambda i: (int(k[i]), i))  # stable by construction
        assert list(vo) == ref and list(ko) == [int(k[i]) for i in ref]


# ---------------------------------------------------------------- hash join
def _join_numpy(r, s):
    ur, cr = np.unique(r, return_counts=True)
    us, cs = np.unique(s, return_counts=True)
    common, ir, is_ = np.intersect1d(ur, us, assume_unique=True, return_indices=True)
    return int((cr[ir].astype(np.int64) * cs[is_].astype(np.int64)).sum())


@pytest.mark.parametrize("domain", [100, 10_000, 2**31])
def test_join_count_vs_numpy(orc, domain):
    rng = np.random.default_rng(domain % 97)
    r = rng.integers(-domain, domain, 30_000).astype(np.int32)
    s = rng.integers(-domain, domain, 20_000).astype(np.int32)
    assert orc.join_count(r, s) == _join_numpy(r, s)


def test_join_count_brute_force_and_special_keys(orc):
    rng = np.random.default_rng(5)
    for _ in range(20):
        r = rng.integers(-3, 3, rng.integers(0, 12)).astype(np.int32)
        s = rng.integers(-3, 3, rng.integers(0, 12)).astype(np.int32)
        assert orc.join_count(r, s) == sum(int(a == b) for a in r for b in s)
    # -1 and INT_MIN are ordinary keys; one key everywhere gives |R| * |S|
    r = np.full(3000, -1, np.int32)
    s = np.full(2000, -1, np.int32)
    assert orc.join_count(r, s) == 6_000_000
    r = np.array([-2**31, -1, 0, 2**31 - 1], np.int32)
    assert orc.join_count(r, r) == 4
    assert orc.join_count(r, np.empty(0, np.int32)) == 0
