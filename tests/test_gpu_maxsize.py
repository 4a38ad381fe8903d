"""The largest n the C ABI takes (2^31 - 1 elements, 8 GiB per int32 array),
checked through size-independent properties on the device: positions, tile
counts and histogram indices all exceed 2^30 here."""
import pytest

pytestmark = pytest.mark.gpu

N_MAX = (1 << 31) - 1


@pytest.fixture
def big(torch_gpu):
    yield torch_gpu
    torch_gpu.cuda.synchronize()
    torch_gpu.cuda.empty_cache()


def test_scan_max_n(big, ctx):
    torch = big
    n = N_MAX
    x = torch.empty(n, dtype=torch.int32, device="cuda")
    ctx.gen_i32(x, 11, 15)
    out = torch.empty(n, dtype=torch.int32, device="cuda")
    total = torch.zeros(1, dtype=torch.int32, device="cuda")
    for scheme in (1, 2):
        ctx.scan(x, out, n, scheme=scheme, total=total)
        assert int(out[0]) == 0
        assert bool((out[1:] - out[:-1] == x[:-1]).all())  # int32 wrap-around differences
        s = int(x.sum(dtype=torch.int64).item())
        assert int(total.item()) % (1 << 32) == s % (1 << 32)
    del x, out


def test_gather_scatter_max_n(big, ctx):
    torch = big
    n = N_MAX
    idx = torch.empty(n, dtype=torch.int32, device="cuda")
    ctx.gen_perm(idx, 12, 0)
    src = torch.empty(n, dtype=torch.int32, device="cuda")
    ctx.gen_i32(src, 13)
    mid = torch.empty(n, dtype=torch.int32, device="cuda")
    ctx.gather(src, mid, idx, n, 1)
    back = torch.empty(n, dtype=torch.int32, device="cuda")
    ctx.scatter(mid, back, idx, n, 4)
    assert torch.equal(back, src)  # scatter through idx undoes the gather through idx
    del idx, src, mid, back


def test_split_max_n(big, ctx):
    torch = big
    n = N_MAX
    keys = torch.empty(n, dtype=torch.int32, device="cuda")
    vals = torch.empty(n, dtype=torch.int32, device="cuda")
    ctx.gen_split_keys(keys, vals, 256, 0, 14)  # vals = original index
    kout = torch.empty(n, dtype=torch.int32, device="cuda")
    vout = torch.empty(n, dtype=torch.int32, device="cuda")
    st = torch.empty(256, dtype=torch.int32, device="cuda")
    ctx.split(keys, kout, 256, vals, vout, st, n, 0)
    b = kout & 255
    assert bool((b[1:] >= b[:-1]).all())
    same = b[1:] == b[:-1]
    assert bool((vout[1:][same] > vout[:-1][same]).all())
    del same, b
    assert torch.equal(keys[vout.long()], kout)
    cnt = torch.bincount((keys & 255).long(), minlength=256)
    assert torch.equal(st.long(), torch.cumsum(cnt, 0) - cnt)
    del keys, vals, kout, vout


def test_sort_max_n(big, ctx):
    torch = big
    n = N_MAX
    keys = torch.empty(n, dtype=torch.int32, device="cuda")
    ctx.gen_i32(keys, 15)
    vals = torch.arange(n, dtype=torch.int32, device="cuda")
    ctx.sort(keys, keys, vals, vals, n, 0, 32, True)  # in place
    assert bool((keys[1:] >= keys[:-1]).all())
    same = keys[1:] == keys[:-1]
    assert bool((vals[1:][same] > vals[:-1][same]).all())
    del same
    ref = torch.empty(n, dtype=torch.int32, device="cuda")
    ctx.gen_i32(ref, 15)
    assert torch.equal(ref[vals.long()], keys)
    del keys, vals, ref


@pytest.mark.parametrize("wide,n", [(False, (1 << 29) - 1), (False, 1 << 29), (True, (1 << 28) - 1), (True, 1 << 28)])
def test_split_at_the_buffer_store_limit(big, ctx, wide, n):
    """The WG scatter stores through buffer descriptors while the output's
    byte offsets fit 31 bits (n * key bytes < 2^31) and through global
    pointers from there on: both sides of that switch, 256 buckets."""
    torch = big
    dt = torch.int64 if wide else torch.int32
    keys = torch.randint(-2**31, 2**31 - 1, (n,), dtype=dt, device="cuda")
    vals = torch.arange(n, dtype=torch.int32, device="cuda")
    kout = torch.empty_like(keys)
    vout = torch.empty_like(vals)
    st = torch.empty(256, dtype=torch.int32, device="cuda")
    ctx.split(keys, kout, 256, vals, vout, st, n, 1, wide=wide)
    b = kout & 255
    assert bool((b[1:] >= b[:-1]).all())
    same = b[1:] == b[:-1]
    assert bool((vout[1:][same] > vout[:-1][same]).all())  # stable
    del same, b
    assert torch.equal(keys[vout.long()], kout)  # a permutation of the pairs
    cnt = torch.bincount((keys & 255).long(), minlength=256)
    assert torch.equal(st.long(), torch.cumsum(cnt, 0) - cnt)
    del keys, vals, kout, vout
