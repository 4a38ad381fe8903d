"""Writes the golden fixtures of this directory: small seeded inputs and the
outputs of the sequential oracle (oracle/oracle.c), each output also checked
here against an independent closed form before it is written (numpy cumsum
for the scan, fancy indexing for gather/scatter, a stable argsort and
bincount for the split, a stable argsort of the key bits for the radix sort,
element-wise formulas for STREAM).  Run from the
repository root after ``python oracle/build.py``:

    python tests/golden/make_golden.py

Files are numpy .npz archives of plain integer arrays (no pickled objects).
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from oracle import pyoracle as orc  # noqa: E402

SEEDS = (1, 0xC0FFEE)
SIZES = (1, 63, 64, 65, 1000, 4096)
GEN_N = 1024


def _scan():
    out = {}
    for seed in SEEDS:
        for n in SIZES:
            for mask in (15, -1):
                x = orc.gen_i32(n, seed + n, mask)
                y, tot = orc.scan_exclusive(x)
                cs = np.cumsum(x.astype(np.int64)) & 0xFFFFFFFF
                ref = np.concatenate([np.zeros(1, np.int64), cs[:-1]]).astype(np.uint32).view(np.int32)
                assert np.array_equal(y, ref) and (int(tot) & 0xFFFFFFFF) == int(cs[-1])
                key = f"s{seed}_n{n}_m{mask & 0xFF}"
                out[key + "_in"], out[key + "_out"] = x, y
                out[key + "_total"] = np.array([tot], np.int32)
    ones = np.ones(4096, np.int32)
    y, tot = orc.scan_exclusive(ones)
    assert np.array_equal(y, np.arange(4096, dtype=np.int32)) and int(tot) == 4096  # all-ones known answer
    out["ones_in"], out["ones_out"] = ones, y
    x64 = orc.gen_i64(4096, 7)
    y64, t64 = orc.scan_exclusive(x64)
    cs = np.cumsum(x64.astype(np.uint64))
    assert np.array_equal(y64.view(np.uint64), np.concatenate([np.zeros(1, np.uint64), cs[:-1]]))
    out["i64_in"], out["i64_out"], out["i64_total"] = x64, y64, np.array([t64], np.int64)
    return out


def _gather_scatter():
    out = {}
    for seed in SEEDS:
        for n in (1, 65, 1000, 4096):
            for kind in (0, 1):
                v = orc.gen_i32(n, seed + 3 * n)
                idx = orc.gen_perm(n, seed + 5 * n, kind)
                g, s = orc.gather(v, idx, 1), orc.scatter(v, idx, 1)
                ref_s = np.empty_like(v)
                ref_s[idx] = v
                assert np.array_equal(g, v[idx]) and np.array_equal(s, ref_s)
                for p in (2, 3, 16):
                    assert np.array_equal(orc.gather(v, idx, p), g) and np.array_equal(orc.scatter(v, idx, p), s)
                key = f"s{seed}_n{n}_k{kind}"
                out[key + "_vals"], out[key + "_idx"], out[key + "_gather"], out[key + "_scatter"] = v, idx, g, s
    iota = np.arange(4096, dtype=np.int32)
    idx = orc.gen_perm(4096, 99, 0)
    inv = np.empty_like(idx)
    inv[idx] = iota
    assert np.array_equal(orc.gather(iota, idx, 4), idx) and np.array_equal(orc.scatter(iota, idx, 4), inv)
    out["iota_idx"], out["iota_gather"], out["iota_scatter"] = idx, idx, inv
    return out


def _split():
    out = {}
    for seed in SEEDS[:1]:
        for n in (1, 100, 4096):
            for B in (1, 2, 16, 256, 4096):
                for dist in (0, 1):
                    k, v = orc.gen_split_keys(n, B, dist, seed + n + B)
                    ko, vo, st = orc.split(k, B, v)
                    b = k & (B - 1)
                    order = np.argsort(b, kind="stable")
                    cnt = np.bincount(b, minlength=B)
                    assert np.array_equal(ko, k[order]) and np.array_equal(vo, v[order])
                    assert np.array_equal(st, (np.cumsum(cnt) - cnt).astype(np.int32))
                    key = f"s{seed}_n{n}_B{B}_d{dist}"
                    out[key + "_keys"], out[key + "_vals"] = k, v
                    out[key + "_kout"], out[key + "_vout"], out[key + "_start"] = ko, vo, st
    k, v = orc.gen_split_keys(4096, 256, 1, 11, wide=True)
    ko, vo, st = orc.split(k, 256, v)
    order = np.argsort(k & 255, kind="stable")
    assert np.array_equal(ko, k[order]) and np.array_equal(vo, v[order])
    out["i64_keys"], out["i64_vals"], out["i64_kout"], out["i64_vout"], out["i64_start"] = k, v, ko, vo, st
    return out


def _sort():
    out = {}
    for seed in SEEDS:
        for n in (1, 100, 4096):
            for signed in (0, 1):
                for lo, hi in ((0, 32), (4, 20)):
                    k = orc.gen_i32(n, seed + n + lo)
                    k[: n // 8] = k[n // 8: 2 * (n // 8)]  # duplicates
                    k = k if signed else k.view(np.uint32)
                    v = np.arange(n, dtype=np.int32)
                    ko, vo = orc.sort(k, v, lo, hi, bool(signed))
                    u = k.view(np.uint32).astype(np.uint64) ^ np.uint64(0x80000000 if signed else 0)
                    order = np.argsort((u >> np.uint64(lo)) & np.uint64((1 << (hi - lo)) - 1), kind="stable")
                    assert np.array_equal(vo, order) and np.array_equal(ko, k[order])
                    key = f"s{seed}_n{n}_g{signed}_b{lo}_{hi}"
                    out[key + "_keys"], out[key + "_kout"], out[key + "_vout"] = k.view(np.int32), ko.view(np.int32), vo
    return out


def _stream():
    out = {}
    for n in (1, 5, 1027):
        a, b = orc.gen_i32(n, 1 + n), orc.gen_i32(n, 2 + n)
        exp = [a, (3 * a.astype(np.int64)).astype(np.int32), (a.astype(np.int64) + b).astype(np.int32),
               (a.astype(np.int64) + 3 * b.astype(np.int64)).astype(np.int32)]
        for op in range(4):
            r = orc.stream(op, a, b if op >= 2 else None)
            assert np.array_equal(r, exp[op])
            out[f"n{n}_op{op}"] = r
        out[f"n{n}_a"], out[f"n{n}_b"] = a, b
    return out


def _generators():
    out = {}
    for seed in SEEDS:
        out[f"s{seed}_i32"] = orc.gen_i32(GEN_N, seed)
        out[f"s{seed}_i32m15"] = orc.gen_i32(GEN_N, seed, 15)
        out[f"s{seed}_i64"] = orc.gen_i64(GEN_N, seed)
        for kind in (0, 1, 2):
            p = orc.gen_perm(GEN_N, seed, kind)
            assert np.array_equal(np.sort(p), np.arange(GEN_N))
            out[f"s{seed}_perm{kind}"] = p
        for B in (2, 256):
            for dist in (0, 1):
                k, v = orc.gen_split_keys(GEN_N, B, dist, seed)
                out[f"s{seed}_keys_B{B}_d{dist}"], out[f"s{seed}_vals_B{B}_d{dist}"] = k, v
    return out


def main():
    for name, fn in (("scan", _scan), ("gather_scatter", _gather_scatter), ("split", _split),
                     ("sort", _sort), ("stream", _stream), ("generators", _generators)):
        path = os.path.join(HERE, f"{name}.npz")
        np.savez_compressed(path, **fn())
        print(path, os.path.getsize(path))


if __name__ == "__main__":
    main()
