"""The exclusive scan that allocates nothing (edpp_scan_exclusive_i32_noalloc),
called through the C ABI, against the existing entry point
edpp_scan_exclusive_i32 on the same inputs: two calls in a row, the second
with a smaller input, for aligned and unaligned buffers, plus the size limit."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _pair(torch, ctx, orc, n, seed, offset):
    x = orc.gen_i32(n, seed, 15 if seed % 2 else -1)
    xin = torch.zeros(n + offset, dtype=torch.int32, device="cuda")
    xin[offset:] = torch.from_numpy(x).cuda()
    a = torch.full((n + offset,), -7, dtype=torch.int32, device="cuda")
    b = torch.full((n + offset,), -7, dtype=torch.int32, device="cuda")
    ta = torch.full((1,), -1, dtype=torch.int32, device="cuda")
    tb = torch.full((1,), -2, dtype=torch.int32, device="cuda")
    ctx.scan_noalloc(xin[offset:], a[offset:], n, ta)
    torch.cuda.synchronize()
    ctx.scan(xin[offset:], b[offset:], n, 1, tb)  # the existing entry point, look-back scheme
    assert torch.equal(a, b), (n, offset)  # including the untouched guard word below an unaligned output
    assert int(ta.item()) == int(tb.item()), (n, offset)
    return a[offset:].cpu().numpy(), x


@pytest.mark.parametrize("offset", [0, 1])
@pytest.mark.parametrize("sizes", [(1 << 24, 3_000_001), ((1 << 21) + 5, 1000), (100_003, 0), (1, 64)])
def test_noalloc_scan_matches_existing_entry_point(torch_gpu, ctx, orc, offset, sizes):
    torch = torch_gpu
    for k, n in enumerate(sizes):  # two calls in a row, the second smaller
        got, x = _pair(torch, ctx, orc, n, 1000 + n + k, offset)
        if n <= 1 << 22:
            assert np.array_equal(got, orc.scan_exclusive(x)[0]), (n, offset)


def test_noalloc_scan_full_size_and_limits(torch_gpu, ctx, edpp, orc):
    torch = torch_gpu
    n = 1 << 28
    x = torch.ones(n, dtype=torch.int32, device="cuda")
    y = torch.empty(n, dtype=torch.int32, device="cuda")
    t = torch.zeros(1, dtype=torch.int32, device="cuda")
    ctx.scan_noalloc(x, y, n, t)  # all ones: out[i] = i (cuda/test_scan.cu:53-58)
    assert torch.equal(y, torch.arange(n, dtype=torch.int32, device="cuda")) and int(t.item()) == n
    del y
    # unaligned buffers above 2^26 elements need more look-back words than the scratch holds
    m = (1 << 26) + 1
    with pytest.raises(edpp.EdppError):
        ctx.scan_noalloc(x[1:m + 1], x[1:m + 1], m)
    # the existing entry point is unaffected by a rejected call
    z = torch.empty(1000, dtype=torch.int32, device="cuda")
    ctx.scan_noalloc(x[:1000], z, 1000)
    assert torch.equal(z, torch.arange(1000, dtype=torch.int32, device="cuda"))
