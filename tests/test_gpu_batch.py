"""Batched executor: every job of a mixed batch bit-exact against the oracle,
for the plain per-call loop, one persistent launch and repeated hipGraph
replays (with eager work between the replays)."""
import importlib

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def check_job(orc, b, o, buckets=256):
    p = b["plan"]
    n = p["n"]
    src = b["in"].cpu().numpy()
    if p["kind"] == 0:
        ref, tot = orc.scan_exclusive(src)
        assert np.array_equal(b[f"out{o}"].cpu().numpy(), ref), p
        assert b[f"total{o}"].cpu().numpy()[0] == tot, p
    elif p["kind"] == 1:
        ko, vo, start = orc.split(src, buckets, b["vals"].cpu().numpy())
        assert np.array_equal(b[f"out{o}"].cpu().numpy(), ko), p
        assert np.array_equal(b[f"vout{o}"].cpu().numpy(), vo), p
        assert np.array_equal(b[f"start{o}"].cpu().numpy(), start), p
    else:
        idx = b["idx"].cpu().numpy()
        assert np.array_equal(b[f"out{o}"].cpu().numpy(), src[idx]), p
    assert n == src.size


@pytest.mark.parametrize("njobs,total,seed", [(60, 1 << 18, 1), (300, 1 << 22, 2)])
def test_batch_matches_oracle(torch_gpu, ctx, orc, njobs, total, seed):
    torch = torch_gpu
    edpp = importlib.import_module("efficient-dpp_amd")
    jm = importlib.import_module("efficient-dpp_amd.jobmix")
    plan, _ = jm.mix_plan(njobs, total, lo=1, hi=1 << 18, seed=seed)
    plan[0]["n"] = 0  # an empty job
    jobsets, bufs, arena = jm.build_mix(edpp, ctx, torch, plan, outputs=3)
    batch = edpp.Batch(ctx, jobsets[1])
    info = batch.info()
    assert info["tiles"] > 0
    edpp.Batch(ctx, jobsets[0]).run_plain()
    batch.run()
    for b in bufs:
        check_job(orc, b, 0)
        check_job(orc, b, 1)
    # graph replays: clobber the outputs with eager work in between
    g = edpp.Batch(ctx, jobsets[2])
    for rep in range(3):
        for b in bufs:
            b["out2"].fill_(-1)
        g.run(graph=True)
        torch.cuda.synchronize()
        for b in bufs:
            check_job(orc, b, 2)
    batch.close()
    g.close()


def test_batch_occupancy_and_rejects(torch_gpu, ctx, edpp):
    torch = torch_gpu
    n = 5000
    x = torch.ones(n, dtype=torch.int32, device="cuda")
    y = torch.empty(n, dtype=torch.int32, device="cuda")
    for occ in (1, 2, 8):
        b = edpp.Batch(ctx, [edpp.make_job(edpp.JOB_SCAN, n, x, y, 4)])
        b.set_occupancy(occ)
        b.run()
        assert torch.equal(y, torch.arange(n, dtype=torch.int32, device="cuda"))
        b.close()
    with pytest.raises(edpp.EdppError):
        edpp.Batch(ctx, [edpp.make_job(edpp.JOB_SPLIT, n, x, y, 4, buckets=512)])
    with pytest.raises(edpp.EdppError):
        edpp.Batch(ctx, [edpp.make_job(edpp.JOB_SCAN, n - 1, x[1:], y[1:], 4)])  # not 16-byte aligned


def test_batch_every_shape_matches_oracle(torch_gpu, ctx, orc, edpp):
    # every tile shape of the autotuner, at two occupancies, on one mixed batch
    torch = torch_gpu
    jm = importlib.import_module("efficient-dpp_amd.jobmix")
    shapes = edpp.batch_shapes()
    assert len(shapes) >= 4 and shapes[0]["split_tile"] == shapes[0]["block"] * 8
    assert len({s["block"] for s in shapes}) >= 2  # waves per workgroup vary
    plan, _ = jm.mix_plan(120, 1 << 21, lo=1, hi=1 << 17, seed=9)
    plan[3]["n"] = 0
    jobsets, bufs, arena = jm.build_mix(edpp, ctx, torch, plan, outputs=1)
    batch = edpp.Batch(ctx, jobsets[0])
    try:
        for k in range(len(shapes)):
            batch.set_shape(k)
            for occ in (2, 8):
                batch.set_occupancy(occ)
                order = (k + occ) % 3  # every queue order over the shapes
                batch.set_order(order)
                assert batch.get_shape() == (k, occ, order)
                for bb in bufs:
                    bb["out0"].fill_(-1)
                batch.run(graph=(occ == 8))
                torch.cuda.synchronize()
                for bb in bufs:
                    check_job(orc, bb, 0)
    finally:
        batch.close()


def test_batch_autotune_picks_a_timed_shape(torch_gpu, ctx, orc, edpp):
    torch = torch_gpu
    jm = importlib.import_module("efficient-dpp_amd.jobmix")
    plan, _ = jm.mix_plan(200, 1 << 22, lo=256, hi=1 << 18, seed=10)
    jobsets, bufs, arena = jm.build_mix(edpp, ctx, torch, plan, outputs=1)
    batch = edpp.Batch(ctx, jobsets[0])
    try:
        times = batch.autotune(reps=2)
        assert len(times) == len(edpp.BATCH_ORDERS) * len(edpp.batch_shapes()) * len(edpp.BATCH_OCCUPANCIES)
        assert all(t > 0 for t in times.values())
        chosen = batch.get_shape()
        assert times[chosen] == min(times.values())
        for bb in bufs:
            bb["out0"].fill_(-1)
        batch.run(graph=True)
        torch.cuda.synchronize()
        for bb in bufs:
            check_job(orc, bb, 0)
        # a job writing its own input cannot be autotuned
        x = torch.ones(1000, dtype=torch.int32, device="cuda")
        inplace = edpp.Batch(ctx, [edpp.make_job(edpp.JOB_SCAN, 1000, x, x, 4)])
        with pytest.raises(edpp.EdppError):
            inplace.autotune(1)
        inplace.close()
    finally:
        batch.close()
