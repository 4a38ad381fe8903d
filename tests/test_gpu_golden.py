"""The library (HIP kernels through the C ABI) against the committed golden
fixtures (tests/golden/*.npz), bit-exact."""
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def load(name):
    return np.load(os.path.join(GOLDEN, f"{name}.npz"))


def cases(z, suffix):
    return sorted(k[: -len(suffix)] for k in z.files if k.endswith(suffix))


def dev(torch, a):
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def host(t):
    return t.cpu().numpy()


def test_golden_scan(torch_gpu, ctx):
    torch = torch_gpu
    z = load("scan")
    for k in cases(z, "_in"):
        x, y = z[k + "_in"], z[k + "_out"]
        wide = x.dtype == np.int64
        for scheme in (1, 2, 3, 4):
            out = torch.empty(len(x), dtype=torch.int64 if wide else torch.int32, device="cuda")
            tot = torch.zeros(1, dtype=out.dtype, device="cuda")
            ctx.scan(dev(torch, x), out, len(x), scheme=scheme, total=tot, wide=wide)
            assert np.array_equal(host(out), y), (k, scheme)
            if k + "_total" in z.files:
                assert int(host(tot)[0]) == int(z[k + "_total"][0]), (k, scheme)


def test_golden_gather_scatter(torch_gpu, ctx):
    torch = torch_gpu
    z = load("gather_scatter")
    for k in cases(z, "_vals"):
        v, idx = dev(torch, z[k + "_vals"]), dev(torch, z[k + "_idx"])
        n = len(z[k + "_vals"])
        out = torch.empty(n, dtype=torch.int32, device="cuda")
        for p in (1, 2, 7, 16):
            out.fill_(0)
            ctx.gather(v, out, idx, n, p)
            assert np.array_equal(host(out), z[k + "_gather"]), (k, p)
            out.fill_(0)
            ctx.scatter(v, out, idx, n, p)
            assert np.array_equal(host(out), z[k + "_scatter"]), (k, p)


def test_golden_split(torch_gpu, ctx):
    torch = torch_gpu
    z = load("split")
    for k in cases(z, "_keys"):
        B = int(k.split("_B")[1].split("_")[0]) if "_B" in k else 256
        keys = z[k + "_keys"]
        wide = keys.dtype == np.int64
        n = len(keys)
        for strategy in (1, 2, 3):
            kout = torch.empty(n, dtype=torch.int64 if wide else torch.int32, device="cuda")
            vout = torch.empty(n, dtype=torch.int32, device="cuda")
            st = torch.full((B,), -1, dtype=torch.int32, device="cuda")
            ctx.split(dev(torch, keys), kout, B, dev(torch, z[k + "_vals"]), vout, st, n, strategy, wide=wide)
            assert np.array_equal(host(kout), z[k + "_kout"]), (k, strategy)
            assert np.array_equal(host(vout), z[k + "_vout"]), (k, strategy)
            assert np.array_equal(host(st), z[k + "_start"]), (k, strategy)


def test_golden_sort(torch_gpu, ctx):
    torch = torch_gpu
    z = load("sort")
    for k in cases(z, "_keys"):
        signed = "_g1_" in k
        lo, hi = (int(x) for x in k.split("_b")[1].split("_"))
        kin = z[k + "_keys"]
        n = len(kin)
        kout = torch.empty(n, dtype=torch.int32, device="cuda")
        vout = torch.empty(n, dtype=torch.int32, device="cuda")
        vals = torch.arange(n, dtype=torch.int32, device="cuda")
        ctx.sort(dev(torch, kin), kout, vals, vout, n, lo, hi, signed)
        assert np.array_equal(host(kout), z[k + "_kout"]), k
        assert np.array_equal(host(vout), z[k + "_vout"]), k


def test_golden_stream(torch_gpu, ctx):
    torch = torch_gpu
    z = load("stream")
    for n in (1, 5, 1027):
        a, b = dev(torch, z[f"n{n}_a"]), dev(torch, z[f"n{n}_b"])
        c = torch.empty(n, dtype=torch.int32, device="cuda")
        for op in range(4):
            ctx.stream_op(op, a, b, c, n)
            assert np.array_equal(host(c), z[f"n{n}_op{op}"]), (n, op)


def test_golden_generators(torch_gpu, ctx):
    torch = torch_gpu
    z = load("generators")
    for seed in (1, 0xC0FFEE):
        n = len(z[f"s{seed}_i32"])
        a = torch.empty(n, dtype=torch.int32, device="cuda")
        ctx.gen_i32(a, seed)
        assert np.array_equal(host(a), z[f"s{seed}_i32"])
        ctx.gen_i32(a, seed, 15)
        assert np.array_equal(host(a), z[f"s{seed}_i32m15"])
        b = torch.empty(n, dtype=torch.int64, device="cuda")
        ctx.gen_i64(b, seed)
        assert np.array_equal(host(b), z[f"s{seed}_i64"])
        for kind in (0, 1, 2):
            ctx.gen_perm(a, seed, kind)
            assert np.array_equal(host(a), z[f"s{seed}_perm{kind}"]), kind
        v = torch.empty(n, dtype=torch.int32, device="cuda")
        for B in (2, 256):
            for dist in (0, 1):
                ctx.gen_split_keys(a, v, B, dist, seed)
                assert np.array_equal(host(a), z[f"s{seed}_keys_B{B}_d{dist}"]), (B, dist)
                assert np.array_equal(host(v), z[f"s{seed}_vals_B{B}_d{dist}"]), (B, dist)
