"""GPU parity: every primitive through the C ABI against the CPU oracle on the
same seeded inputs, bit-exact, plus size-independent properties at the full
2^28 sizes."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

SIZES = [0, 1, 2, 3, 63, 64, 65, 1023, 1024, 1025, 4095, 4096, 4097, 65_537, 1 << 20, (1 << 20) + 17, 3_000_001]


def dev(torch, a):
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def host(t):
    return t.cpu().numpy()


def empty(torch, n, dtype):
    return torch.empty(max(n, 1), dtype=dtype, device="cuda")[:n]


# ---------------------------------------------------------------- generators
@pytest.mark.parametrize("n", [1, 1000, 4096 * 3 + 7, 1 << 20])
def test_generators_match_oracle(torch_gpu, ctx, orc, n):
    torch = torch_gpu
    seed = 0x1234 + n
    a = empty(torch, n, torch.int32)
    ctx.gen_i32(a, seed)
    assert np.array_equal(host(a), orc.gen_i32(n, seed))
    ctx.gen_i32(a, seed, 15)
    assert np.array_equal(host(a), orc.gen_i32(n, seed, 15))
    b = empty(torch, n, torch.int64)
    ctx.gen_i64(b, seed)
    assert np.array_equal(host(b), orc.gen_i64(n, seed))
    for kind in (0, 1, 2):
        ctx.gen_perm(a, seed, kind)
        assert np.array_equal(host(a), orc.gen_perm(n, seed, kind)), kind
    v = empty(torch, n, torch.int32)
    for B in (1, 2, 16, 256, 4096):
        for dist in (0, 1):
            ctx.gen_split_keys(a, v, B, dist, seed)
            k, vv = orc.gen_split_keys(n, B, dist, seed)
            assert np.array_equal(host(a), k) and np.array_equal(host(v), vv), (B, dist)
            ctx.gen_split_keys(b, v, B, dist, seed, wide=True)
            k, _ = orc.gen_split_keys(n, B, dist, seed, wide=True)
            assert np.array_equal(host(b), k), (B, dist)


# ---------------------------------------------------------------- STREAM
@pytest.mark.parametrize("n", [0, 1, 5, 1027, 1 << 20, (1 << 20) + 3])
def test_stream_ops(torch_gpu, ctx, orc, n):
    torch = torch_gpu
    a, b = orc.gen_i32(n, 1), orc.gen_i32(n, 2)
    da, db = dev(torch, a), dev(torch, b)
    c = empty(torch, n, torch.int32)
    for op in range(4):
        ctx.stream_op(op, da, db, c, n)
        assert np.array_equal(host(c), orc.stream(op, a, b if op >= 2 else None)), op
    # unaligned pointers take the scalar path
    if n > 8:
        ctx.stream_op(3, da[1:], db[1:], c[1:], n - 1)
        assert np.array_equal(host(c)[1:], orc.stream(3, a[1:], b[1:]))


# ---------------------------------------------------------------- scans
@pytest.mark.parametrize("scheme", [1, 2, 3, 4])
@pytest.mark.parametrize("n", SIZES)
def test_scan_i32(torch_gpu, ctx, orc, scheme, n):
    torch = torch_gpu
    for mask in (15, -1):
        x = orc.gen_i32(n, 100 + n, mask)
        ref, tot = orc.scan_exclusive(x)
        dx = dev(torch, x)
        out = empty(torch, n, torch.int32)
        total = torch.full((1,), 7, dtype=torch.int32, device="cuda")
        ctx.scan(dx, out, n, scheme=scheme, total=total)
        assert np.array_equal(host(out), ref), (scheme, n, mask)
        assert int(host(total)[0]) == int(tot)


@pytest.mark.parametrize("scheme", [1, 2, 3, 4])
@pytest.mark.parametrize("n", [1, 4097, 1 << 20, 3_000_001])
def test_scan_i64(torch_gpu, ctx, orc, scheme, n):
    torch = torch_gpu
    x = orc.gen_i64(n, 7 + n)
    ref, tot = orc.scan_exclusive(x)
    out = empty(torch, n, torch.int64)
    total = torch.zeros(1, dtype=torch.int64, device="cuda")
    ctx.scan(dev(torch, x), out, n, scheme=scheme, total=total, wide=True)
    assert np.array_equal(host(out), ref)
    assert int(host(total)[0]) == int(tot)


@pytest.mark.parametrize("block,vpt,scheme", [(256, 2, 1), (256, 4, 1), (256, 8, 1), (256, 16, 1), (512, 4, 1),
                                              (512, 8, 1), (512, 16, 1), (1024, 4, 1), (1024, 8, 1), (1024, 16, 1),
                                              (1024, 4, 4), (1024, 8, 4)])
def test_scan_lookback_shapes(torch_gpu, ctx, orc, block, vpt, scheme):
    # every look-back tile shape (lanes x 16-byte vectors per lane), one
    # workgroup per tile (scheme 1) or persistent workgroups (scheme 4),
    # 32- and 64-bit, whole and ragged last tiles, in place
    torch = torch_gpu
    ctx.set_tuning(scan_block=block, scan_vpt=vpt)
    try:
        for n in (1, 1000, (1 << 20) + 3, 3_000_001, 1 << 24):
            x = orc.gen_i32(n, 300 + n + vpt, -1)
            ref, tot = orc.scan_exclusive(x)
            out = empty(torch, n, torch.int32)
            total = torch.zeros(1, dtype=torch.int32, device="cuda")
            ctx.scan(dev(torch, x), out, n, scheme=scheme, total=total)
            assert np.array_equal(host(out), ref), (block, vpt, n)
            assert int(host(total)[0]) == int(tot)
            d = dev(torch, x)
            ctx.scan(d, d, n, scheme=scheme)
            assert np.array_equal(host(d), ref), ("in place", block, vpt, n)
        for n in (4097, 3_000_001):
            x = orc.gen_i64(n, 17 + n + vpt)
            ref, tot = orc.scan_exclusive(x)
            out = empty(torch, n, torch.int64)
            ctx.scan(dev(torch, x), out, n, scheme=scheme, wide=True)
            assert np.array_equal(host(out), ref), ("i64", block, vpt, n)
    finally:
        ctx.set_tuning()


@pytest.mark.parametrize("scheme", [1, 2, 3, 4])
def test_scan_in_place_and_unaligned(torch_gpu, ctx, orc, scheme):
    torch = torch_gpu
    n = 1_000_003
    x = orc.gen_i32(n, 5)
    ref, _ = orc.scan_exclusive(x)
    d = dev(torch, x)
    ctx.scan(d, d, n, scheme=scheme)  # in place, as split scans its histogram
    assert np.array_equal(host(d), ref)
    d = dev(torch, np.concatenate([np.zeros(1, np.int32), x]))
    out = empty(torch, n + 1, torch.int32)
    ctx.scan(d[1:], out[1:], n, scheme=scheme)  # 4-byte offset: not 16-byte aligned
    assert np.array_equal(host(out)[1:], ref)


def test_scan_c1_config(torch_gpu, ctx, orc):
    # 2^24 int32 values uniform in [0,15] from a fixed seed, every scheme
    torch = torch_gpu
    n = 1 << 24
    x = orc.gen_i32(n, 2024, 15)
    ref, tot = orc.scan_exclusive(x)
    d = empty(torch, n, torch.int32)
    ctx.gen_i32(d, 2024, 15)
    out = empty(torch, n, torch.int32)
    total = torch.zeros(1, dtype=torch.int32, device="cuda")
    for scheme in (1, 2, 3, 4):
        out.fill_(-1)
        ctx.scan(d, out, scheme=scheme, total=total)
        assert np.array_equal(host(out), ref), scheme
        assert int(host(total)[0]) == int(tot)


def test_scan_full_size_properties(torch_gpu, ctx):
    # 2^28 values in [0,15]: total wraps past 2^31; adjacent differences equal
    # the input and the total equals the (wrapped) sum
    torch = torch_gpu
    n = 1 << 28
    x = empty(torch, n, torch.int32)
    ctx.gen_i32(x, 77, 15)
    out = empty(torch, n, torch.int32)
    total = torch.zeros(1, dtype=torch.int32, device="cuda")
    for scheme in (1, 2, 3, 4):
        ctx.scan(x, out, scheme=scheme, total=total)
        assert int(out[0]) == 0
        assert torch.equal(out[1:] - out[:-1], x[:-1])  # int32 wrap-around subtraction
        s = int(x.to(torch.int64).sum().item())
        assert int(total.item()) % (1 << 32) == s % (1 << 32)
        assert (int(out[-1].item()) + int(x[-1].item())) % (1 << 32) == s % (1 << 32)
    del x, out


# ---------------------------------------------------------------- local (matrix) scans
@pytest.mark.parametrize("matrix", [0, 1, 2, 3])
@pytest.mark.parametrize("network", [0, 1, 2, 3, 4, 5, 6, 7, 8])
@pytest.mark.parametrize("tile,local", [(1, 512), (11, 512), (49, 512), (64, 64), (16, 1024), (3, 128), (2, 512),
                                        (64, 512), (38, 1024)])
def test_scan_local(torch_gpu, ctx, orc, matrix, network, tile, local):
    # even tiles take LDS rows padded to an odd stride, except (38, 1024), whose
    # padded rows would not fit in LDS
    torch = torch_gpu
    n = tile * local
    x = orc.gen_i32(n, tile * 1000 + local, 15)
    out = empty(torch, n, torch.int32)
    ctx.scan_local(dev(torch, x), out, tile, local, matrix, network)
    assert np.array_equal(host(out), orc.scan_exclusive(x)[0])


@pytest.mark.parametrize("matrix", [0, 2])
@pytest.mark.parametrize("tile", [11, 16])
def test_scan_local_unaligned(torch_gpu, ctx, orc, matrix, tile):
    # input and output one word past a 16-byte boundary: word-wise copies
    torch = torch_gpu
    local = 256
    n = tile * local
    x = orc.gen_i32(n, 4242 + tile, 15)
    xin = torch.zeros(n + 1, dtype=torch.int32, device="cuda")
    xin[1:] = dev(torch, x)
    out = torch.full((n + 1,), -7, dtype=torch.int32, device="cuda")
    ctx.scan_local(xin[1:], out[1:], tile, local, matrix, 3)
    assert int(out[0].item()) == -7
    assert np.array_equal(host(out[1:]), orc.scan_exclusive(x)[0])


# ---------------------------------------------------------------- gather / scatter
@pytest.mark.parametrize("n", [1, 3, 4, 5, 1000, 4096 * 5 + 3, 1 << 20, 3_000_001])
@pytest.mark.parametrize("kind", [0, 1])
def test_gather_scatter(torch_gpu, ctx, orc, n, kind):
    torch = torch_gpu
    vals = orc.gen_i32(n, 3 + n)
    idx = orc.gen_perm(n, 4 + n, kind)
    dv, di = dev(torch, vals), dev(torch, idx)
    out = empty(torch, n, torch.int32)
    gref = vals[idx]
    sref = np.empty_like(vals)
    sref[idx] = vals
    for passes in (1, 2, 3, 8, 16):
        out.fill_(0)
        ctx.gather(dv, out, di, n, passes)
        assert np.array_equal(host(out), gref), ("gather", passes)
        out.fill_(0)
        ctx.scatter(dv, out, di, n, passes)
        assert np.array_equal(host(out), sref), ("scatter", passes)


@pytest.mark.parametrize("unroll", [1, 2, 4, 8])
@pytest.mark.parametrize("kind", [0, 1])
def test_gather_scatter_window(torch_gpu, ctx, orc, unroll, kind):
    # block-local indices go through the LDS window: whole (4 vectors per
    # lane: one 4096-element block per workgroup), shared by two or four
    # workgroups (1, 2), cut by pass slices (3 and 5 passes), or too wide (8)
    torch = torch_gpu
    ctx.set_tuning(gs_unroll=unroll)
    try:
        for n in (4096 * 9, 4096 * 64 + 100):
            vals = orc.gen_i32(n, 31 + n)
            idx = orc.gen_perm(n, 32 + n, kind)
            dv, di = dev(torch, vals), dev(torch, idx)
            gref = vals[idx]
            sref = np.empty_like(vals)
            sref[idx] = vals
            out = empty(torch, n, torch.int32)
            for passes in (1, 3, 5):
                out.fill_(0)
                ctx.gather(dv, out, di, n, passes)
                assert np.array_equal(host(out), gref), ("gather", unroll, n, passes)
                out.fill_(0)
                ctx.scatter(dv, out, di, n, passes)
                assert np.array_equal(host(out), sref), ("scatter", unroll, n, passes)
            w = orc.gen_i64(n, 33 + n)
            dw = dev(torch, w)
            out64 = empty(torch, n, torch.int64)
            ctx.gather(dw, out64, di, n, 1, wide=True)
            assert np.array_equal(host(out64), w[idx])
            ctx.scatter(dw, out64, di, n, 2, wide=True)
            sw = np.empty_like(w)
            sw[idx] = w
            assert np.array_equal(host(out64), sw)
    finally:
        ctx.set_tuning()


@pytest.mark.parametrize("n", [7, 4096 * 3 + 1, 1 << 20])
def test_gather_scatter_i64_and_unaligned(torch_gpu, ctx, orc, n):
    torch = torch_gpu
    vals = orc.gen_i64(n, 9)
    idx = orc.gen_perm(n, 10, 0)
    dv, di = dev(torch, vals), dev(torch, idx)
    out = empty(torch, n, torch.int64)
    for passes in (1, 4):
        ctx.gather(dv, out, di, n, passes, wide=True)
        assert np.array_equal(host(out), vals[idx])
        ctx.scatter(dv, out, di, n, passes, wide=True)
        sref = np.empty_like(vals)
        sref[idx] = vals
        assert np.array_equal(host(out), sref)
    # unaligned int32 buffers use the scalar kernel
    v32 = orc.gen_i32(n + 1, 11)
    d32 = dev(torch, v32)
    i32 = dev(torch, np.concatenate([np.zeros(1, np.int32), idx]))
    o32 = empty(torch, n + 1, torch.int32)
    ctx.gather(d32[1:], o32[1:], i32[1:], n, 3)
    assert np.array_equal(host(o32)[1:], v32[1:][idx])


def test_gather_out_of_range_index_is_skipped(torch_gpu, ctx, orc):
    torch = torch_gpu
    n = 4096
    idx = orc.gen_perm(n, 1, 0)
    idx[5] = n + 100
    idx[9] = -3
    vals = orc.gen_i32(n, 2)
    out = torch.full((n,), -7, dtype=torch.int32, device="cuda")
    ctx.gather(dev(torch, vals), out, dev(torch, idx), n, 1)
    h = host(out)
    ok = (idx >= 0) & (idx < n)
    assert np.array_equal(h[ok], vals[idx[ok]]) and h[5] == -7 and h[9] == -7


@pytest.mark.parametrize("kind", [0, 1])
def test_gather_scatter_full_size_properties(torch_gpu, ctx, kind):
    # 2^28: gather of iota returns the index (test_gather.cpp:64-70), and
    # scatter undoes gather, for every pass count 1..16
    torch = torch_gpu
    n = 1 << 28
    idx = empty(torch, n, torch.int32)
    ctx.gen_perm(idx, 31, kind)
    iota = torch.arange(n, dtype=torch.int32, device="cuda")
    out = empty(torch, n, torch.int32)
    for passes in range(1, 17):
        ctx.gather(iota, out, idx, n, passes)
        assert torch.equal(out, idx)
        back = torch.zeros(n, dtype=torch.int32, device="cuda")
        ctx.scatter(out, back, idx, n, passes)
        assert torch.equal(back, iota)
        del back
    del idx, iota, out


# ---------------------------------------------------------------- split
SPLIT_SIZES = [0, 1, 63, 100, 4096 * 3 + 5, 100_003, 1 << 20]


@pytest.mark.parametrize("strategy", [1, 2, 3, 4])  # WG, WI, SINGLE, CHAINED
@pytest.mark.parametrize("B", [1, 2, 4, 8, 16, 32, 256, 4096])
@pytest.mark.parametrize("n", SPLIT_SIZES)
def test_split_i32(torch_gpu, ctx, orc, strategy, B, n):
    torch = torch_gpu
    for dist in (0, 1):
        keys, vals = orc.gen_split_keys(n, B, dist, 1000 + n + B)
        ko, vo, start = orc.split(keys, B, vals)
        dk, dvv = dev(torch, keys), dev(torch, vals)
        kout = empty(torch, n, torch.int32)
        vout = empty(torch, n, torch.int32)
        st = torch.full((B,), -1, dtype=torch.int32, device="cuda")
        ctx.split(dk, kout, B, dvv, vout, st, n, strategy)
        assert np.array_equal(host(kout), ko), (strategy, B, n, dist)
        assert np.array_equal(host(vout), vo), (strategy, B, n, dist)
        assert np.array_equal(host(st), start), (strategy, B, n, dist)
        # key-only
        kout2 = empty(torch, n, torch.int32)
        ctx.split(dk, kout2, B, n=n, strategy=strategy)
        assert np.array_equal(host(kout2), ko)


@pytest.mark.parametrize("strategy", [0, 1, 4])  # AUTO, WG, CHAINED
def test_split_starts_with_absent_buckets(torch_gpu, ctx, orc, strategy):
    # 4096 buckets take two passes, whose bucket starts come from a search of
    # the output: most buckets are absent here (the first and the last among
    # them), and one bucket holds half of the keys
    torch = torch_gpu
    B = 4096
    n = (1 << 22) + 3
    rng = np.random.default_rng(4096 + strategy)
    present = np.sort(rng.choice(np.arange(1, B - 1), 37, replace=False))
    bk = present[rng.integers(0, len(present), n)]
    bk[: n // 2] = present[5]
    keys = ((rng.integers(0, 1 << 19, n, dtype=np.int64) << 12) | bk).astype(np.int32)
    vals = np.arange(n, dtype=np.int32)
    ko, vo, start = orc.split(keys, B, vals)
    assert start[0] == 0 and start[-1] == n  # absent first and last buckets
    kout = empty(torch, n, torch.int32)
    vout = empty(torch, n, torch.int32)
    st = torch.full((B,), -1, dtype=torch.int32, device="cuda")
    ctx.split(dev(torch, keys), kout, B, dev(torch, vals), vout, st, n, strategy)
    assert np.array_equal(host(kout), ko), strategy
    assert np.array_equal(host(vout), vo), strategy
    assert np.array_equal(host(st), start), strategy


@pytest.mark.parametrize("reorder", [1, 2, 3])  # REORDER_NONE, _FIXED, _VARIED
@pytest.mark.parametrize("ipt,mult", [(8, 16), (16, 16), (16, 64), (8, 256)])
@pytest.mark.parametrize("B", [1, 2, 16, 32, 256, 4096])
def test_split_reorder_variants(torch_gpu, ctx, orc, reorder, ipt, mult, B):
    # the WG split's three ways of writing a tile (the reference's ReorderType):
    # tiles of one and of many sub-tiles, so the fixed variant's staging
    # carries elements across sub-tiles and flushes them at the tile end
    torch = torch_gpu
    ctx.set_tuning(split_ipt=ipt, split_tile_mult=mult, split_reorder=reorder)
    try:
        for n, dist in ((1, 0), (4096 * 7 + 19, 1), (300_007, 0), (1 << 20, 1)):
            keys, vals = orc.gen_split_keys(n, B, dist, 77 + n + B)
            ko, vo, start = orc.split(keys, B, vals)
            dk, dvv = dev(torch, keys), dev(torch, vals)
            kout = empty(torch, n, torch.int32)
            vout = empty(torch, n, torch.int32)
            st = torch.full((B,), -1, dtype=torch.int32, device="cuda")
            ctx.split(dk, kout, B, dvv, vout, st, n, 1)
            assert np.array_equal(host(kout), ko), (reorder, ipt, mult, B, n)
            assert np.array_equal(host(vout), vo), (reorder, ipt, mult, B, n)
            assert np.array_equal(host(st), start), (reorder, ipt, mult, B, n)
            kout2 = empty(torch, n, torch.int32)
            ctx.split(dk, kout2, B, n=n, strategy=1)  # key-only
            assert np.array_equal(host(kout2), ko), (reorder, B, n)
            k64, _ = orc.gen_split_keys(n, B, dist, 91 + n + B, wide=True)
            k64o, v64o, _ = orc.split(k64, B, vals)
            d64o = torch.empty(n, dtype=torch.int64, device="cuda")
            ctx.split(dev(torch, k64), d64o, B, dvv, vout, None, n, 1, wide=True)
            assert np.array_equal(host(d64o), k64o) and np.array_equal(host(vout), v64o), (reorder, B, n, "wide")
    finally:
        ctx.set_tuning()


@pytest.mark.parametrize("reorder", [1, 2])
def test_split_reorder_full_size(torch_gpu, ctx, reorder):
    # 2^28 pairs through the none and fixed variants, against the default one
    torch = torch_gpu
    n = 1 << 28
    keys = empty(torch, n, torch.int32)
    vals = empty(torch, n, torch.int32)
    ctx.gen_split_keys(keys, vals, 256, 0, 4343)
    ref_k, ref_v = empty(torch, n, torch.int32), empty(torch, n, torch.int32)
    ctx.split(keys, ref_k, 256, vals, ref_v, None, n, 1)
    ctx.set_tuning(split_reorder=reorder)
    try:
        ko, vo = empty(torch, n, torch.int32), empty(torch, n, torch.int32)
        ctx.split(keys, ko, 256, vals, vo, None, n, 1)
        assert torch.equal(ko, ref_k) and torch.equal(vo, ref_v)
    finally:
        ctx.set_tuning()
    del keys, vals, ref_k, ref_v, ko, vo


@pytest.mark.parametrize("ipt,mult,block", [(8, 16, 0), (8, 64, 0), (16, 16, 0), (16, 64, 0), (16, 256, 0),
                                            (16, 16, 256), (16, 16, 512), (16, 64, 512), (16, 96, 512)])
@pytest.mark.parametrize("B", [2, 8, 16, 128, 256, 4096])
def test_split_wg_shapes(torch_gpu, ctx, orc, ipt, mult, block, B):
    # sub-tile shapes of the WG strategy (256- and 512-lane scatter workgroups,
    # tiles of 1..several sub-tiles), sizes ragged against both the sub-tile
    # and the tile
    torch = torch_gpu
    ctx.set_tuning(split_ipt=ipt, split_tile_mult=mult, split_block=block)
    try:
        for n in (8192 * 3 + 4096 + 77, 200_003):
            keys, vals = orc.gen_split_keys(n, B, n % 2, 77 + B + ipt)
            ko, vo, start = orc.split(keys, B, vals)
            dk, dvv = dev(torch, keys), dev(torch, vals)
            kout = empty(torch, n, torch.int32)
            vout = empty(torch, n, torch.int32)
            st = empty(torch, B, torch.int32)
            ctx.split(dk, kout, B, dvv, vout, st, n, 1)
            assert np.array_equal(host(kout), ko), (ipt, mult, block, B, n)
            assert np.array_equal(host(vout), vo), (ipt, mult, block, B, n)
            assert np.array_equal(host(st), start), (ipt, mult, block, B, n)
            kout2 = empty(torch, n, torch.int32)
            ctx.split(dk, kout2, B, n=n, strategy=1)
            assert np.array_equal(host(kout2), ko), (ipt, mult, B, n)
    finally:
        ctx.set_tuning()


@pytest.mark.parametrize("strategy", [1, 2, 4])
@pytest.mark.parametrize("B", [2, 128, 256, 4096])
def test_split_i64(torch_gpu, ctx, orc, strategy, B):
    # from 128 buckets the WG strategy runs 8-byte keys on 512 lanes at 8 keys
    # per lane (values staged after the keys)
    torch = torch_gpu
    n = 300_007
    keys, vals = orc.gen_split_keys(n, B, 1, 5 + B, wide=True)
    ko, vo, start = orc.split(keys, B, vals)
    kout = empty(torch, n, torch.int64)
    vout = empty(torch, n, torch.int32)
    st = empty(torch, B, torch.int32)
    ctx.split(dev(torch, keys), kout, B, dev(torch, vals), vout, st, n, strategy, wide=True)
    assert np.array_equal(host(kout), ko) and np.array_equal(host(vout), vo) and np.array_equal(host(st), start)


@pytest.mark.parametrize("strategy", [1, 4])
@pytest.mark.parametrize("B", [2, 16, 256])
def test_split_unaligned_input(torch_gpu, ctx, orc, B, strategy):
    # keys and values one word past a 16-byte boundary
    torch = torch_gpu
    n = 100_003
    keys, vals = orc.gen_split_keys(n, B, 1, 900 + B)
    ko, vo, start = orc.split(keys, B, vals)
    dk = torch.zeros(n + 1, dtype=torch.int32, device="cuda")
    dvv = torch.zeros(n + 1, dtype=torch.int32, device="cuda")
    dk[1:] = dev(torch, keys)
    dvv[1:] = dev(torch, vals)
    kout = empty(torch, n, torch.int32)
    vout = empty(torch, n, torch.int32)
    st = empty(torch, B, torch.int32)
    ctx.split(dk[1:], kout, B, dvv[1:], vout, st, n, strategy)
    assert np.array_equal(host(kout), ko) and np.array_equal(host(vout), vo) and np.array_equal(host(st), start)


@pytest.mark.parametrize("strategy", [1, 4])
@pytest.mark.parametrize("dist", [0, 1])
@pytest.mark.parametrize("B", [2, 16, 256, 4096])
def test_split_full_size_properties(torch_gpu, ctx, B, dist, strategy):
    # 2^28 pairs, uniform or Zipf keys, vals = original index: bucket ids
    # non-decreasing, stable (vals increasing inside each bucket), a
    # permutation, keys follow their values, offsets equal the bucket counts
    torch = torch_gpu
    n = 1 << 28
    keys = empty(torch, n, torch.int32)
    vals = empty(torch, n, torch.int32)
    ctx.gen_split_keys(keys, vals, B, dist, 4242 + dist)
    kout = empty(torch, n, torch.int32)
    vout = empty(torch, n, torch.int32)
    st = empty(torch, B, torch.int32)
    ctx.split(keys, kout, B, vals, vout, st, n, strategy)
    b = kout & (B - 1)
    assert bool((b[1:] >= b[:-1]).all())
    same = b[1:] == b[:-1]
    assert bool((vout[1:][same] > vout[:-1][same]).all())
    assert torch.equal(keys[vout.long()], kout)
    cnt = torch.bincount((keys & (B - 1)).long(), minlength=B)
    exp = torch.cumsum(cnt, 0) - cnt
    assert torch.equal(st.long(), exp)
    del keys, vals, kout, vout, b, same


# ---------------------------------------------------------------- radix sort
SORT_STRATEGIES = [1, 4]  # per-digit WG splits, or one count of every digit and chained scatters


@pytest.fixture
def sort_ctx(ctx, request):
    ctx.set_tuning(sort_strategy=request.param)
    yield ctx
    ctx.set_tuning()


@pytest.mark.parametrize("sort_ctx", SORT_STRATEGIES, indirect=True)
@pytest.mark.parametrize("signed", [False, True])
@pytest.mark.parametrize("bits", [(0, 32), (0, 16), (4, 20), (8, 32), (0, 8), (3, 4), (7, 7)])
@pytest.mark.parametrize("n", [0, 1, 1000, 100_003, (1 << 20) - 1, (1 << 20) + 17, 3_000_001])
def test_sort(torch_gpu, sort_ctx, orc, signed, bits, n):
    torch = torch_gpu
    ctx = sort_ctx
    keys = orc.gen_i32(n, 500 + n + bits[0])
    if n > 10:
        keys[: n // 10] = keys[n // 10: 2 * (n // 10)]  # duplicates: stability observable
    keys = keys if signed else keys.view(np.uint32)
    vals = np.arange(n, dtype=np.int32)
    ko, vo = orc.sort(keys, vals, bits[0], bits[1], signed)
    dk = dev(torch, keys.view(np.int32))
    kout = empty(torch, n, torch.int32)
    vout = empty(torch, n, torch.int32)
    ctx.sort(dk, kout, dev(torch, vals), vout, n, bits[0], bits[1], signed)
    assert np.array_equal(host(kout), ko.view(np.int32)), (signed, bits, n)
    assert np.array_equal(host(vout), vo), (signed, bits, n)
    k2 = empty(torch, n, torch.int32)
    ctx.sort(dk, k2, n=n, begin_bit=bits[0], end_bit=bits[1], signed=signed)  # keys only
    assert np.array_equal(host(k2), ko.view(np.int32))
    # in place (keys and values)
    vin = dev(torch, vals)
    ctx.sort(dk, dk, vin, vin, n, bits[0], bits[1], signed)
    assert np.array_equal(host(dk), ko.view(np.int32)) and np.array_equal(host(vin), vo), ("in place", bits, n)


@pytest.mark.parametrize("sort_ctx", SORT_STRATEGIES, indirect=True)
@pytest.mark.parametrize("kind", ["equal", "two", "zipf"])
def test_sort_skewed_keys(torch_gpu, sort_ctx, orc, kind):
    # passes with every key of a tile in one bucket (equal), two values (0
    # and -1: the flipped top digit orders -1 first), Zipf low bytes
    torch = torch_gpu
    ctx = sort_ctx
    n = 3_000_001
    if kind == "equal":
        keys = np.full(n, 0x12345678, np.int32)
    elif kind == "two":
        keys = np.where(orc.gen_i32(n, 41) & 1, -1, 0).astype(np.int32)
    else:
        keys, _ = orc.gen_split_keys(n, 256, 1, 42)
    vals = np.arange(n, dtype=np.int32)
    for signed in (False, True):
        ko, vo = orc.sort(keys if signed else keys.view(np.uint32), vals, 0, 32, signed)
        kout = empty(torch, n, torch.int32)
        vout = empty(torch, n, torch.int32)
        ctx.sort(dev(torch, keys), kout, dev(torch, vals), vout, n, 0, 32, signed)
        assert np.array_equal(host(kout), ko.view(np.int32)), (kind, signed)
        assert np.array_equal(host(vout), vo), (kind, signed)


@pytest.mark.parametrize("sort_ctx", SORT_STRATEGIES, indirect=True)
def test_sort_full_size_properties(torch_gpu, sort_ctx):
    # 2^28 pairs, signed 32-bit keys, vals = original index: keys ascending,
    # ties in input order, a permutation, keys follow their values
    torch = torch_gpu
    ctx = sort_ctx
    n = 1 << 28
    keys = empty(torch, n, torch.int32)
    ctx.gen_i32(keys, 0x50F7)
    vals = torch.arange(n, dtype=torch.int32, device="cuda")
    kout = empty(torch, n, torch.int32)
    vout = empty(torch, n, torch.int32)
    ctx.sort(keys, kout, vals, vout, n, 0, 32, True)
    assert bool((kout[1:] >= kout[:-1]).all())
    same = kout[1:] == kout[:-1]
    assert bool((vout[1:][same] > vout[:-1][same]).all())
    assert torch.equal(keys[vout.long()], kout)
    assert int(torch.bincount(vout.long(), minlength=n).max()) == 1
    del keys, vals, kout, vout, same


# ---------------------------------------------------------------- hash join
@pytest.mark.parametrize("rn,sn", [(0, 10), (10, 0), (1, 1), (1000, 3000), (100_003, 70_001), (1 << 20, (1 << 20) + 5)])
@pytest.mark.parametrize("domain", [50, 100_000, 2**31])
def test_hash_join_count(torch_gpu, ctx, orc, rn, sn, domain):
    torch = torch_gpu
    r = (orc.gen_i64(rn, 900 + rn) % (2 * domain) - domain).astype(np.int32)
    s = (orc.gen_i64(sn, 901 + sn) % (2 * domain) - domain).astype(np.int32)
    cnt = torch.full((1,), 12345, dtype=torch.int64, device="cuda")
    ctx.join_count(dev(torch, r) if rn else empty(torch, 0, torch.int32),
                   dev(torch, s) if sn else empty(torch, 0, torch.int32), cnt, rn, sn)
    assert int(cnt.item()) == orc.join_count(r, s), (rn, sn, domain)


def test_hash_join_skew_and_special_keys(torch_gpu, ctx, orc):
    torch = torch_gpu
    cnt = torch.zeros(1, dtype=torch.int64, device="cuda")
    # one key everywhere: |R|*|S| > 2^32 pairs, and a partition far above a table fill
    r = np.full(100_000, -1, np.int32)
    s = np.full(70_000, -1, np.int32)
    ctx.join_count(dev(torch, r), dev(torch, s), cnt)
    assert int(cnt.item()) == 7_000_000_000
    # structured keys (multiples of 2^20) and the extremes of int32
    r = (np.arange(200_000, dtype=np.int64) << 20).astype(np.int32)
    s = np.concatenate([r[::3], np.array([-2**31, 2**31 - 1, -1, 0], np.int32)])
    ctx.join_count(dev(torch, r), dev(torch, s), cnt)
    assert int(cnt.item()) == orc.join_count(r, s)
    # heavy duplicates on both sides
    r = (orc.gen_i32(300_000, 3) & 63).astype(np.int32)
    s = (orc.gen_i32(200_000, 4) & 1023).astype(np.int32)
    ctx.join_count(dev(torch, r), dev(torch, s), cnt)
    assert int(cnt.item()) == orc.join_count(r, s)


def test_hash_join_heavy_partitions_among_light_ones(torch_gpu, ctx, orc):
    # 2048 partitions, 8 per workgroup: the 64 partitions of the heavy keys
    # (9 000 copies on each side) need chunked table fills, the others are
    # filled from registers, so workgroups switch between the two kinds
    torch = torch_gpu
    rng = np.random.default_rng(2048)
    heavy = rng.choice(1 << 30, 64, replace=False).astype(np.int64) + (1 << 30)
    r = np.concatenate([rng.permutation(1 << 23)[: 1 << 22], np.repeat(heavy, 9000)])
    s = np.concatenate([rng.permutation(1 << 23)[: 1 << 22], np.repeat(heavy, 9000)])
    r = rng.permutation(r).astype(np.int32)
    s = rng.permutation(s).astype(np.int32)
    cnt = torch.zeros(1, dtype=torch.int64, device="cuda")
    ctx.join_count(dev(torch, r), dev(torch, s), cnt)
    want = orc.join_count(r, s)
    assert want > 64 * 9000 * 9000
    assert int(cnt.item()) == want


@pytest.mark.parametrize("n", [(1 << 25) + 3, 1 << 26])
def test_hash_join_full_size_property(torch_gpu, ctx, n):
    # R = a random permutation of [0, n), S = R's first n//2 keys shifted by n
    # plus the rest of R: exactly n - n//2 matching pairs.  2^25+3 keys
    # partition by 13 bits (split passes of 6 and 7 bits), 2^26 by 14 (7 and 7)
    torch = torch_gpu
    r = empty(torch, n, torch.int32)
    ctx.gen_perm(r, 77, 0)
    s = r.clone()
    s[: n // 2] += n
    cnt = torch.zeros(1, dtype=torch.int64, device="cuda")
    ctx.join_count(r, s, cnt)
    assert int(cnt.item()) == n - n // 2
    del r, s


# ---------------------------------------------------------------- access probes
@pytest.mark.parametrize("pattern", [0, 1, 2])
@pytest.mark.parametrize("repeat", [1, 2, 3, 7, 20])
@pytest.mark.parametrize("k", [1, 37, 4096])
def test_access_scale(torch_gpu, ctx, orc, pattern, repeat, k):
    torch = torch_gpu
    n = 64 * repeat * k
    a = orc.gen_i32(n, 31 + n)
    out = torch.zeros(n, dtype=torch.int32, device="cuda")
    ctx.access_scale(pattern, dev(torch, a), out, repeat)
    assert np.array_equal(host(out), orc.stream(1, a)), (pattern, repeat, k)


def test_access_scale_rejects_ragged(torch_gpu, ctx, edpp):
    torch = torch_gpu
    a = torch.zeros(64 * 3 + 1, dtype=torch.int32, device="cuda")
    with pytest.raises(edpp.EdppError):
        ctx.access_scale(0, a, a.clone(), 3)


@pytest.mark.parametrize("n", [1, 1000, (1 << 20) + 3])
@pytest.mark.parametrize("chunk", [1, 256, 4096, 1 << 21])
def test_access_chunks(torch_gpu, ctx, orc, n, chunk):
    torch = torch_gpu
    a = orc.gen_i32(n, 71 + n)
    chunks = -(-n // chunk)
    cnt = torch.full((1,), -5, dtype=torch.int64, device="cuda")
    for order in (np.arange(chunks), np.arange(chunks)[::-1],
                  np.random.default_rng(chunks).permutation(chunks)):
        ctx.access_chunks(dev(torch, a), chunk, dev(torch, np.ascontiguousarray(order, np.int32)), cnt)
        assert int(cnt.item()) == int((a & 1).sum()), (n, chunk)


def _fmix32(x):
    # MurmurHash3's finaliser on uint32 numpy arrays (the probe's position hash)
    x = x.astype(np.uint32)
    x ^= x >> np.uint32(16)
    x *= np.uint32(0x85EBCA6B)
    x ^= x >> np.uint32(13)
    x *= np.uint32(0xC2B2AE35)
    x ^= x >> np.uint32(16)
    return x


@pytest.mark.parametrize("table_n,accesses", [(1, 5), (64, 1000), (1 << 12, 1 << 16), (1 << 20, 3_000_001),
                                              (1 << 16, 0)])
def test_access_random_probe(torch_gpu, ctx, orc, table_n, accesses):
    torch = torch_gpu
    seed = 0x9E37 + table_n
    tab = orc.gen_i32(table_n, 17 + table_n)
    with np.errstate(over="ignore"):
        pos = _fmix32(np.arange(accesses, dtype=np.uint64).astype(np.uint32) + np.uint32(seed & 0xFFFFFFFF))
    pos &= np.uint32(table_n - 1)
    want = int(tab.view(np.uint32)[pos].astype(np.uint64).sum()) & (2**64 - 1)
    dt = dev(torch, tab)
    tot = torch.full((1,), -1, dtype=torch.int64, device="cuda")
    ctx.access_random(dt, accesses, seed, False, tot)
    assert int(tot.item()) & (2**64 - 1) == want
    # writes: every touched position holds its own index, the others keep -1
    wt = torch.full((table_n,), -1, dtype=torch.int32, device="cuda")
    ctx.access_random(wt, accesses, seed, True)
    exp = np.full(table_n, -1, np.int32)
    exp[pos] = pos.astype(np.int32)
    assert np.array_equal(host(wt), exp)


def test_access_random_rejects(torch_gpu, ctx, edpp):
    torch = torch_gpu
    t = torch.zeros(100, dtype=torch.int32, device="cuda")
    tot = torch.zeros(1, dtype=torch.int64, device="cuda")
    with pytest.raises(edpp.EdppError):
        ctx.access_random(t, 10, 1, False, tot)  # 100 is not a power of two
    with pytest.raises(edpp.EdppError):
        ctx.access_random(t[:64], 1 << 32, 1, False, tot, table_n=64)
