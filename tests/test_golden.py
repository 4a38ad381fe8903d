"""The committed golden fixtures (tests/golden/*.npz, written by
tests/golden/make_golden.py) against closed forms that do not use the oracle,
and the oracle against the fixtures (CPU)."""
import os

import numpy as np
import pytest

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def load(name):
    return np.load(os.path.join(GOLDEN, f"{name}.npz"))  # plain arrays; pickles refused


def cases(z, suffix):
    return sorted(k[: -len(suffix)] for k in z.files if k.endswith(suffix))


def test_scan_fixtures(orc):
    z = load("scan")
    keys = cases(z, "_in")
    assert len(keys) >= 20
    for k in keys:
        x, y = z[k + "_in"], z[k + "_out"]
        cs = np.cumsum(x.astype(np.int64 if x.dtype == np.int32 else np.uint64))
        if x.dtype == np.int32:
            ref = np.concatenate([np.zeros(1, np.int64), cs[:-1]]).astype(np.uint32).view(np.int32)
        else:
            ref = np.concatenate([np.zeros(1, np.uint64), cs[:-1]]).view(np.int64)
        assert np.array_equal(y, ref), k
        got, tot = orc.scan_exclusive(x)
        assert np.array_equal(got, y), k
        if k + "_total" in z.files:
            assert int(z[k + "_total"][0]) == int(tot), k
    assert np.array_equal(z["ones_out"], np.arange(4096, dtype=np.int32))


def test_gather_scatter_fixtures(orc):
    z = load("gather_scatter")
    keys = cases(z, "_vals")
    assert len(keys) >= 8
    for k in keys:
        v, idx, g, s = z[k + "_vals"], z[k + "_idx"], z[k + "_gather"], z[k + "_scatter"]
        assert np.array_equal(np.sort(idx), np.arange(len(idx))), k  # a permutation
        assert np.array_equal(g, v[idx]), k
        assert np.array_equal(s[idx], v), k
        for p in (1, 2, 5, 16):
            assert np.array_equal(orc.gather(v, idx, p), g), (k, p)
            assert np.array_equal(orc.scatter(v, idx, p), s), (k, p)
    # iota known answers (opencl/test/test_gather.cpp:25,64-70; test_scatter.cpp:63-69)
    idx = z["iota_idx"]
    assert np.array_equal(z["iota_gather"], idx)
    assert np.array_equal(z["iota_scatter"][idx], np.arange(len(idx), dtype=np.int32))


def test_split_fixtures(orc):
    z = load("split")
    keys = cases(z, "_keys")
    assert len(keys) >= 20
    for k in keys:
        B = int(k.split("_B")[1].split("_")[0]) if "_B" in k else 256
        keys_in, vals = z[k + "_keys"], z[k + "_vals"]
        ko, vo, st = z[k + "_kout"], z[k + "_vout"], z[k + "_start"]
        b = (keys_in & (B - 1)).astype(np.int64)
        bo = (ko & (B - 1)).astype(np.int64)
        assert np.all(bo[1:] >= bo[:-1]), k                      # bucket ids non-decreasing
        assert np.array_equal(np.sort(vo), np.sort(vals)), k     # a permutation of the pairs
        cnt = np.bincount(b, minlength=B)
        assert np.array_equal(st, (np.cumsum(cnt) - cnt).astype(np.int32)), k
        for bb in np.unique(b):                                   # stable: input order inside a bucket
            assert np.array_equal(ko[bo == bb], keys_in[b == bb]), (k, bb)
            assert np.array_equal(vo[bo == bb], vals[b == bb]), (k, bb)
        ko2, vo2, st2 = orc.split(keys_in, B, vals)
        assert np.array_equal(ko2, ko) and np.array_equal(vo2, vo) and np.array_equal(st2, st), k


def test_sort_fixtures(orc):
    z = load("sort")
    keys = cases(z, "_keys")
    assert len(keys) >= 12
    for k in keys:
        signed = "_g1_" in k
        lo, hi = (int(x) for x in k.split("_b")[1].split("_"))
        kin, ko, vo = z[k + "_keys"], z[k + "_kout"], z[k + "_vout"]
        u = kin.view(np.uint32).astype(np.uint64) ^ np.uint64(0x80000000 if signed else 0)
        d = (u >> np.uint64(lo)) & np.uint64((1 << (hi - lo)) - 1)
        assert np.array_equal(np.sort(vo), np.arange(len(kin))), k          # a permutation
        assert np.array_equal(ko, kin[vo]), k                                # keys follow values
        dd = d[vo]
        assert np.all(dd[1:] >= dd[:-1]), k                                  # sorted by the key bits
        assert np.all((dd[1:] > dd[:-1]) | (vo[1:] > vo[:-1])), k           # ties in input order
        k2, v2 = orc.sort(kin if signed else kin.view(np.uint32), np.arange(len(kin), dtype=np.int32), lo, hi, signed)
        assert np.array_equal(k2.view(np.int32), ko) and np.array_equal(v2, vo), k


def test_stream_fixtures(orc):
    z = load("stream")
    for n in (1, 5, 1027):
        a, b = z[f"n{n}_a"].astype(np.int64), z[f"n{n}_b"].astype(np.int64)
        exp = [a, 3 * a, a + b, a + 3 * b]
        for op in range(4):
            assert np.array_equal(z[f"n{n}_op{op}"], exp[op].astype(np.int32)), (n, op)
            assert np.array_equal(orc.stream(op, z[f"n{n}_a"], z[f"n{n}_b"] if op >= 2 else None),
                                  z[f"n{n}_op{op}"])


def test_generator_fixtures(orc):
    z = load("generators")
    for seed in (1, 0xC0FFEE):
        n = len(z[f"s{seed}_i32"])
        assert np.array_equal(orc.gen_i32(n, seed), z[f"s{seed}_i32"])
        assert np.array_equal(orc.gen_i32(n, seed, 15), z[f"s{seed}_i32m15"])
        assert np.array_equal(orc.gen_i64(n, seed), z[f"s{seed}_i64"])
        for kind in (0, 1, 2):
            assert np.array_equal(orc.gen_perm(n, seed, kind), z[f"s{seed}_perm{kind}"])
        # block-local permutation: at n <= 4096 one block, shuffled inside
        p1 = z[f"s{seed}_perm1"]
        assert np.array_equal(p1 // 4096, np.arange(n) // 4096)
        assert np.array_equal(z[f"s{seed}_perm2"], np.arange(n))
        for B in (2, 256):
            for dist in (0, 1):
                k, v = orc.gen_split_keys(n, B, dist, seed)
                assert np.array_equal(k, z[f"s{seed}_keys_B{B}_d{dist}"]) and np.array_equal(
                    v, z[f"s{seed}_vals_B{B}_d{dist}"])


@pytest.mark.parametrize("name", ["scan", "gather_scatter", "split", "sort", "stream", "generators"])
def test_fixtures_are_plain_arrays(name):
    z = load(name)
    for k in z.files:
        assert z[k].dtype.kind in "iu", (name, k)
