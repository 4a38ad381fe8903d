"""Exhaustive autotuner, its timing database, and training / evaluation of the
learned schedule model.

The timing of every candidate runs in the library (``edpp_autotune``: one
warm-up, ``reps`` timed runs, median); this module draws the calls, keeps the
database (JSON lines, one record per call) and drives the model's HIP
training kernels.  Calls follow the primitive configurations of the
benchmark: exclusive scans of values in [0,15], gathers and scatters through
a random or 4096-block-local permutation, and stable key/value splits into
B in {2,16,256,4096} buckets with uniform or Zipf(1) keys; n is log-uniform in
[2^16, 2^28] and the element (key) width is 4 or 8 bytes.
"""
import json
import math

import numpy as np

OPS = ("scan", "gather", "scatter", "split")
FAMILY_OF_OP = {0: 0, 1: 1, 2: 1, 3: 2}
# one fixed schedule per family: the method without the model (look-back 1024x8,
# one pass, WG split with 16 keys per lane and tiles of 64 elements per bucket)
DEFAULT_CANDIDATE = {0: 0, 1: 5, 2: 11}


def sample_specs(count, seed, lo=16, hi=28):
    rng = np.random.default_rng(seed)
    specs = []
    for i in range(count):
        op = int(rng.integers(0, 4))
        n = int(round(2.0 ** rng.uniform(lo, hi)))
        spec = {"op": op, "n": n, "width": int(rng.choice([4, 8])), "seed": int(seed * 100_003 + i)}
        if op == 3:
            spec["buckets"] = int(rng.choice([2, 16, 256, 4096]))
            spec["dist"] = int(rng.integers(0, 2))
        if op in (1, 2):
            spec["perm"] = int(rng.integers(0, 2))
        specs.append(spec)
    return specs


class Workspace:
    """Device buffers large enough for any sampled call (n <= 2^hi)."""

    def __init__(self, edpp, ctx, torch, nmax):
        self.edpp, self.ctx, self.torch = edpp, ctx, torch
        self.a = torch.empty(nmax, dtype=torch.int64, device="cuda")
        self.b = torch.empty(nmax, dtype=torch.int64, device="cuda")
        self.idx = torch.empty(nmax, dtype=torch.int32, device="cuda")
        self.vin = torch.empty(nmax, dtype=torch.int32, device="cuda")
        self.vout = torch.empty(nmax, dtype=torch.int32, device="cuda")
        self.feat = torch.empty(16, dtype=torch.float32, device="cuda")

    def prepare(self, spec):
        """Generate the call's inputs on the device and return its edpp_call."""
        edpp, ctx, torch = self.edpp, self.ctx, self.torch
        n, w, seed = spec["n"], spec["width"], spec["seed"]
        dt = torch.int64 if w == 8 else torch.int32
        a = self.a.view(dt)[:n] if w == 8 else self.a.view(torch.int32)[:n]
        b = self.b.view(dt)[:n] if w == 8 else self.b.view(torch.int32)[:n]
        op = spec["op"]
        if op == 0:
            (ctx.gen_i64 if w == 8 else ctx.gen_i32)(a, seed, 15)
            return edpp.make_call(edpp.OP_SCAN, n, a, b, w)
        if op in (1, 2):
            (ctx.gen_i64 if w == 8 else ctx.gen_i32)(a, seed)
            ctx.gen_perm(self.idx, seed, spec["perm"], n=n)
            return edpp.make_call(op, n, a, b, w, idx=self.idx)
        ctx.gen_split_keys(a, self.vin, spec["buckets"], spec["dist"], seed, n=n, wide=w == 8)
        return edpp.make_call(edpp.OP_SPLIT, n, a, b, w, buckets=spec["buckets"], vals_in=self.vin,
                              vals_out=self.vout)

    def features(self, call):
        self.ctx.call_features(call, self.feat)
        return self.feat[:12].cpu().numpy().astype(np.float64).tolist()


def call_bytes(spec):
    """Algorithmic bytes of a call (the per-element models of the benchmark)."""
    n, w = spec["n"], spec["width"]
    if spec["op"] == 0:
        return 2 * w * n
    if spec["op"] in (1, 2):
        return (4 + 2 * w) * n
    return 2 * (w + 4) * n


def measure(ws, specs, reps=3, log=None):
    """Exhaustive autotune of every spec: list of records {spec, times, features}."""
    out = []
    for i, spec in enumerate(specs):
        call = ws.prepare(spec)
        times = ws.ctx.autotune(call, reps)
        out.append({"spec": spec, "times": times, "features": ws.features(call),
                    "auto": ws.ctx.default_candidate(call)})
        if log and (i + 1) % 100 == 0:
            log(f"autotuned {i + 1}/{len(specs)} calls")
    return out


def save_db(path, records):
    with open(path, "w") as f:
        for r in records:
            f.write(json.dumps(r) + "\n")


def load_db(path):
    with open(path) as f:
        return [json.loads(line) for line in f if line.strip()]


def arrays(records, tau=0.02, k=15):
    """Features X, soft targets T (candidates within a few % of the best share
    the mass) and masks M, all float32."""
    X = np.array([r["features"] for r in records], np.float32)
    T = np.zeros((len(records), k), np.float32)
    M = np.zeros((len(records), k), np.float32)
    for i, r in enumerate(records):
        t = np.array(r["times"], np.float64)
        valid = t > 0
        M[i, valid] = 1.0
        best = t[valid].min()
        s = np.zeros(k)
        s[valid] = np.exp(-(t[valid] / best - 1.0) / tau)
        T[i] = s / s.sum()
    return X, T, M


def train(edpp, ctx, torch, records, epochs=1500, lr=3e-3, seed=7):
    X, T, M = arrays(records)
    dx, dt, dm = (torch.from_numpy(a).cuda() for a in (X, T, M))
    mlp = edpp.MLP(ctx, seed)
    losses = mlp.train(dx, dt, dm, epochs, lr)
    return mlp, losses


def predict(ctx, torch, mlp, records):
    X, _, M = arrays(records)
    dx, dm = torch.from_numpy(X).cuda(), torch.from_numpy(M).cuda()
    choice = torch.empty(len(records), dtype=torch.int32, device="cuda")
    mlp.forward(dx, dm, None, choice)
    return choice.cpu().numpy()


def regret(records, choices):
    """Aggregate and mean per-call regret of `choices` against the exhaustive best."""
    chosen = np.array([r["times"][c] for r, c in zip(records, choices)])
    best = np.array([min(t for t in r["times"] if t > 0) for r in records])
    return {"aggregate": float(chosen.sum() / best.sum() - 1.0), "mean": float(np.mean(chosen / best - 1.0)),
            "worst": float(np.max(chosen / best - 1.0)), "ms_chosen": float(chosen.sum()),
            "ms_best": float(best.sum())}


def default_choices(records):
    return [DEFAULT_CANDIDATE[FAMILY_OF_OP[r["spec"]["op"]]] for r in records]


def auto_choices(records):
    """The built-in schedule's candidate for every record (the method without
    the model); a record whose built-in schedule is no candidate falls back to
    the family's fixed default."""
    return [r["auto"] if r.get("auto", -1) >= 0 else DEFAULT_CANDIDATE[FAMILY_OF_OP[r["spec"]["op"]]]
            for r in records]


def end_to_end(ws, ctx, specs, use_model, reps=3):
    """Sum of the measured time of the calls run with cand = -1 (the context's
    schedule choice, including the model's feature and inference kernel when
    the model's decision memo does not already hold the call)."""
    total_ms, total_bytes = 0.0, 0.0
    for spec in specs:
        call = ws.prepare(spec)
        ctx.call_run(call, -1)
        ts = sorted(ctx.call_run(call, -1, timed=True) for _ in range(reps))
        total_ms += ts[reps // 2]
        total_bytes += call_bytes(spec)
    return {"ms": total_ms, "gbs": total_bytes / (total_ms * 1e-3) / 1e9, "model": bool(use_model)}


def summarize_times(records):
    fam = {0: [], 1: [], 2: []}
    for r in records:
        t = np.array(r["times"])
        fam[FAMILY_OF_OP[r["spec"]["op"]]].append(int(np.argmin(np.where(t > 0, t, math.inf))))
    return {k: np.bincount(v, minlength=15).tolist() for k, v in fam.items()}
