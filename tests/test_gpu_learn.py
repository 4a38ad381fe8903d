"""Schedule candidates, autotuner, device features and the HIP cost model:
gradients against the float64 reference, bit-exact results for every
candidate and for model-chosen schedules, and what the model gains."""
import importlib

import numpy as np
import pytest

from oracle import mlp_ref

pytestmark = pytest.mark.gpu


def _rand_problem(rows, seed):
    rng = np.random.default_rng(seed)
    x = rng.normal(0, 1, (rows, 12)).astype(np.float32)
    mask = np.zeros((rows, 15), np.float32)
    for r in range(rows):
        lo = 5 * (r % 3)
        mask[r, lo:lo + 5] = 1
    t = (rng.random((rows, 15)) * mask).astype(np.float32)
    t /= t.sum(1, keepdims=True)
    return x, t, mask


def test_mlp_forward_and_gradient_match_reference(torch_gpu, ctx, edpp):
    torch = torch_gpu
    mlp = edpp.MLP(ctx, seed=3)
    p = mlp.get_params().astype(np.float64)
    # 16 and 17 rows: one whole chunk of the gradient kernel and one row past it;
    # 5000 rows: more chunks than workgroups, so workgroups take several
    for rows in (1, 16, 17, 37, 300, 5000):
        x, t, mask = _rand_problem(rows, rows)
        dx, dt, dm = (torch.from_numpy(a).cuda() for a in (x, t, mask))
        probs = torch.empty((rows, 15), dtype=torch.float32, device="cuda")
        choice = torch.empty(rows, dtype=torch.int32, device="cuda")
        mlp.forward(dx, dm, probs, choice)
        ref, _ = mlp_ref.forward(p, x, mask)
        assert np.allclose(probs.cpu().numpy(), ref, rtol=1e-4, atol=1e-6)  # float32 kernel vs float64
        assert np.array_equal(choice.cpu().numpy(), np.argmax(np.where(mask > 0, ref, -1), 1))
        loss = torch.empty(1, dtype=torch.float32, device="cuda")
        grad = torch.empty(mlp.nparams, dtype=torch.float32, device="cuda")
        mlp.loss_grad(dx, dt, dm, loss, grad)
        assert float(loss.item()) == pytest.approx(mlp_ref.loss(p, x, t, mask), rel=1e-4)
        g = mlp_ref.grad(p, x, t, mask)
        assert np.allclose(grad.cpu().numpy(), g, rtol=1e-3, atol=1e-5 * max(1.0, np.abs(g).max()))
    mlp.close()


def test_mlp_gradient_and_training_are_deterministic(torch_gpu, ctx, edpp):
    # partial gradients are summed in a fixed order: repeated calls agree bit for bit
    torch = torch_gpu
    x, t, mask = _rand_problem(5000, 11)
    dx, dt, dm = (torch.from_numpy(a).cuda() for a in (x, t, mask))
    runs = []
    for _ in range(2):
        mlp = edpp.MLP(ctx, seed=5)
        loss = torch.empty(1, dtype=torch.float32, device="cuda")
        grad = torch.empty(mlp.nparams, dtype=torch.float32, device="cuda")
        mlp.loss_grad(dx, dt, dm, loss, grad)
        losses = mlp.train(dx, dt, dm, epochs=50, lr=3e-3)
        runs.append((grad.cpu().numpy(), float(loss.item()), losses, mlp.get_params()))
        mlp.close()
    (g0, l0, h0, p0), (g1, l1, h1, p1) = runs
    assert np.array_equal(g0, g1) and l0 == l1
    assert np.array_equal(h0, h1) and np.array_equal(p0, p1)
    assert h0[0] == pytest.approx(l0, rel=1e-6)  # the first epoch's loss is the loss at the start


def test_mlp_training_learns_a_rule(torch_gpu, ctx, edpp):
    # best candidate determined by the sign of feature 3 within each family
    torch = torch_gpu
    rows = 512
    x, _, mask = _rand_problem(rows, 9)
    fam = np.arange(rows) % 3
    best = 5 * fam + np.where(x[:, 3] > 0, 1, 3)
    t = np.zeros((rows, 15), np.float32)
    t[np.arange(rows), best] = 1
    mlp = edpp.MLP(ctx, seed=1)
    dx, dt, dm = (torch.from_numpy(a).cuda() for a in (x, t, mask))
    losses = mlp.train(dx, dt, dm, epochs=400, lr=1e-2)
    assert losses[-1] < 0.25 * losses[0]
    choice = torch.empty(rows, dtype=torch.int32, device="cuda")
    mlp.forward(dx, dm, None, choice)
    assert np.mean(choice.cpu().numpy() == best) > 0.95
    mlp.close()


def _oracle_check(orc, ws, spec):
    n, w = spec["n"], spec["width"]
    torch = ws.torch
    dt = torch.int64 if w == 8 else torch.int32
    a = (ws.a.view(dt) if w == 8 else ws.a.view(torch.int32))[:n].cpu().numpy()
    b = (ws.b.view(dt) if w == 8 else ws.b.view(torch.int32))[:n].cpu().numpy()
    if spec["op"] == 0:
        return np.array_equal(b, orc.scan_exclusive(a)[0])
    if spec["op"] == 1:
        return np.array_equal(b, a[ws.idx[:n].cpu().numpy()])
    if spec["op"] == 2:
        ref = np.empty_like(a)
        ref[ws.idx[:n].cpu().numpy()] = a
        return np.array_equal(b, ref)
    ko, vo, _ = orc.split(a, spec["buckets"], ws.vin[:n].cpu().numpy())
    return np.array_equal(b, ko) and np.array_equal(ws.vout[:n].cpu().numpy(), vo)


def test_every_candidate_is_bit_exact(torch_gpu, ctx, edpp, orc):
    torch = torch_gpu
    tn = importlib.import_module("efficient-dpp_amd.tuner")
    ws = tn.Workspace(edpp, ctx, torch, 1 << 18)
    specs = tn.sample_specs(24, seed=5, lo=10, hi=18)
    cands = edpp.candidates()
    for spec in specs:
        call = ws.prepare(spec)
        for k, (fam, _) in enumerate(cands):
            if fam != tn.FAMILY_OF_OP[spec["op"]]:
                continue
            ws.b.fill_(-1)
            ctx.call_run(call, k)
            assert _oracle_check(orc, ws, spec), (spec, k)
        times = ctx.autotune(call, 1)
        fam = tn.FAMILY_OF_OP[spec["op"]]
        assert all((t > 0) == (cands[k][0] == fam) for k, t in enumerate(times))


def test_device_features(torch_gpu, ctx, edpp):
    torch = torch_gpu
    tn = importlib.import_module("efficient-dpp_amd.tuner")
    ws = tn.Workspace(edpp, ctx, torch, 1 << 20)
    n = 1 << 20
    f_uni = ws.features(ws.prepare({"op": 3, "n": n, "width": 4, "seed": 1, "buckets": 256, "dist": 0}))
    f_zipf = ws.features(ws.prepare({"op": 3, "n": n, "width": 4, "seed": 1, "buckets": 256, "dist": 1}))
    assert f_uni[2] == 1.0 and f_zipf[6] > f_uni[6] + 0.3  # Zipf(1) over 256 buckets: top bucket ~16% vs 0.4%
    f_rand = ws.features(ws.prepare({"op": 1, "n": n, "width": 4, "seed": 2, "perm": 0}))
    f_blk = ws.features(ws.prepare({"op": 2, "n": n, "width": 8, "seed": 2, "perm": 1}))
    assert f_rand[7] < 0.05 and f_blk[7] > 0.95 and f_blk[11] == 1.0 and f_blk[4] == 1.0
    assert abs(f_rand[3] - 20 / 28) < 1e-6


def test_model_gain_and_model_driven_calls(torch_gpu, ctx, edpp, orc):
    torch = torch_gpu
    tn = importlib.import_module("efficient-dpp_amd.tuner")
    ws = tn.Workspace(edpp, ctx, torch, 1 << 21)
    recs = tn.measure(ws, tn.sample_specs(90, seed=11, lo=14, hi=21), reps=3)
    mlp, losses = tn.train(edpp, ctx, torch, recs, epochs=800, lr=3e-3)
    assert losses[-1] < losses[0]
    choices = tn.predict(ctx, torch, mlp, recs)
    for r, c in zip(recs, choices):
        assert edpp.candidates()[c][0] == tn.FAMILY_OF_OP[r["spec"]["op"]]
    got = tn.regret(recs, choices)
    base = tn.regret(recs, tn.default_choices(recs))
    assert got["aggregate"] <= base["aggregate"] + 0.01, (got, base)
    # AUTO calls through the model stay bit-exact
    ctx.set_model(mlp)
    try:
        for spec in tn.sample_specs(12, seed=12, lo=12, hi=19):
            call = ws.prepare(spec)
            ws.b.fill_(-1)
            ctx.call_run(call, -1)
            assert _oracle_check(orc, ws, spec), spec
        n = 100_000
        x = torch.from_numpy(orc.gen_i32(n, 3, 15)).cuda()
        y = torch.empty(n, dtype=torch.int32, device="cuda")
        ctx.scan(x, y, n, edpp.SCAN_AUTO)
        assert np.array_equal(y.cpu().numpy(), orc.scan_exclusive(x.cpu().numpy())[0])
    finally:
        ctx.set_model(None)
    mlp.close()


def test_default_candidate_is_the_builtin_schedule(torch_gpu, ctx, edpp):
    torch = torch_gpu
    tn = importlib.import_module("efficient-dpp_amd.tuner")
    ws = tn.Workspace(edpp, ctx, torch, 1 << 26)
    names = [c[1] for c in edpp.candidates()]
    want = {
        (0, 1 << 16, 4, 256): "scan look-back 1024x4",
        (0, 1 << 22, 8, 256): "scan look-back 1024x8",
        (0, 1 << 24, 4, 256): "scan look-back 1024x16",
        (0, 1 << 24, 8, 256): "scan look-back 1024x16",  # runs 1024x8 for 64-bit values
        (0, 1 << 25, 4, 256): "scan look-back 1024x16",
        (0, 1 << 26, 4, 256): "scan persistent look-back 1024x8",
        (1, 1 << 20, 4, 256): "gather/scatter 1 pass, unroll 4",
        (2, 1 << 25, 4, 256): "gather/scatter 2 passes",
        (3, 1 << 20, 4, 256): "split WG 8/lane, tile 16B",
        (3, 1 << 25, 4, 256): "split WG 16/lane, tile 16B",
        (3, 1 << 24, 4, 16): "split WG 8/lane, tile 16B",
    }
    for (op, n, width, B), name in want.items():
        spec = {"op": op, "n": n, "width": width, "seed": 3, "buckets": B, "dist": 0, "perm": 0}
        assert names[ctx.default_candidate(ws.prepare(spec))] == name, (op, n, width, B)


def test_model_memo_returns_the_fresh_choice(torch_gpu, ctx, edpp, orc):
    # remembered choices equal fresh ones, follow parameter changes, survive
    # eviction (more signatures than slots), and AUTO calls stay bit-exact
    torch = torch_gpu
    tn = importlib.import_module("efficient-dpp_amd.tuner")
    ws = tn.Workspace(edpp, ctx, torch, 1 << 19)
    specs = tn.sample_specs(80, seed=31, lo=10, hi=19)
    mlp = edpp.MLP(ctx, seed=7)
    ctx.set_model(mlp)
    try:
        def choices():
            return [ctx.model_choice(ws.prepare(s)) for s in specs]
        mlp.memo(False)
        fresh = choices()
        mlp.memo(True)
        assert choices() == fresh  # filled
        assert choices() == fresh  # 80 signatures through 64 slots: evicted ones are recomputed
        p = mlp.get_params()
        q = (p + np.random.default_rng(5).normal(0, 0.5, p.shape)).astype(np.float32)
        mlp.set_params(q)
        mlp.memo(False)
        fresh_q = choices()
        assert fresh_q != fresh  # the perturbation changes some choice, so a stale memo would show
        mlp.memo(True)
        mlp.set_params(p)
        assert choices() == fresh  # remembered under p
        mlp.set_params(q)
        assert choices() == fresh_q  # the change of parameters dropped the choices made under p
        for spec in specs[:9]:
            call = ws.prepare(spec)
            for _ in range(2):  # the second call is served by the memo
                ws.b.fill_(-1)
                ctx.call_run(call, -1)
                assert _oracle_check(orc, ws, spec), spec
    finally:
        ctx.set_model(None)
    mlp.close()
