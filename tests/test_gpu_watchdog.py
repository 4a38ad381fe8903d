"""Failure handling of the device-side waits, forced from the host: with a
watchdog of zero polls, a look-back that meets an unpublished predecessor gives
up at once.  Every kernel still finishes.  A call whose wait gave up reports
ETIMEOUT at the next sync, a call that is not flagged is exact, and the
context works normally once the watchdog is restored."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _outcome(ctx, edpp, run, exact):
    try:
        run()
        ctx.sync()
    except edpp.EdppError as e:
        assert e.status == edpp.ETIMEOUT, e
        return "timeout"
    return "exact" if exact() else "wrong"


def test_watchdog_flags_timeouts_and_the_context_recovers(torch_gpu, ctx, edpp, orc):
    torch = torch_gpu
    n = 1 << 24  # 256 to 512 look-back tiles
    a = orc.gen_i32(n, 99) & 15
    ref, _ = orc.scan_exclusive(a)
    da = torch.from_numpy(a).cuda()
    out = torch.empty(n, dtype=torch.int32, device="cuda")
    m = 1 << 22
    keys, vals = orc.gen_split_keys(m, 256, 0, 98)
    ko, vo, start = orc.split(keys, 256, vals)
    dk, dv = torch.from_numpy(keys).cuda(), torch.from_numpy(vals).cuda()
    kout, vout = torch.empty_like(dk), torch.empty_like(dv)
    st = torch.empty(256, dtype=torch.int32, device="cuda")

    def scan():
        ctx.scan(da, out, n, edpp.SCAN_LOOKBACK)

    def scan_persist():
        ctx.scan(da, out, n, edpp.SCAN_PERSIST)

    def scan_exact():
        return np.array_equal(out.cpu().numpy(), ref)

    def split():
        ctx.split(dk, kout, 256, dv, vout, st, m, edpp.SPLIT_WG)

    def split_chained():
        ctx.split(dk, kout, 256, dv, vout, st, m, edpp.SPLIT_CHAINED)

    def split_exact():
        return (np.array_equal(kout.cpu().numpy(), ko) and np.array_equal(vout.cpu().numpy(), vo)
                and np.array_equal(st.cpu().numpy(), start))

    outcomes = []
    ctx.set_watchdog(0)
    try:
        for _ in range(3):
            outcomes.append(("scan", _outcome(ctx, edpp, scan, scan_exact)))
            outcomes.append(("scan persistent", _outcome(ctx, edpp, scan_persist, scan_exact)))
            outcomes.append(("split", _outcome(ctx, edpp, split, split_exact)))
            outcomes.append(("split chained", _outcome(ctx, edpp, split_chained, split_exact)))
    finally:
        ctx.set_watchdog()
    print("zero-poll watchdog outcomes:", outcomes)
    assert all(o != "wrong" for _, o in outcomes), outcomes  # nothing wrong goes unflagged
    for run, exact in ((scan, scan_exact), (scan_persist, scan_exact), (split, split_exact),
                       (split_chained, split_exact)):
        run()
        ctx.sync()
        assert exact()
