"""Synthetic mixes of many independent jobs for the batched executor.

A mix of ``njobs`` jobs draws, from one seed: a length log-uniform in
[lo, hi], a kind uniform over {exclusive scan, 256-bucket stable split,
random-permutation gather} and an element width uniform over {4, 8} bytes.
Log-uniform lengths in [256, 2^22] average about 4.3e5, so 10,000 of them sum
to about 4.3e9, not 2^28; the lengths are therefore rescaled by one common
factor s (found by bisection) so that sum(clamp(round(s * L_j), lo, hi)) is
within 0.1% of ``total``.  The shape of the distribution (log-uniform,
clamped to [lo, hi]) is kept; only its scale changes.
"""
import math

import numpy as np

KIND_NAMES = ("scan", "split256", "gather")


def mix_lengths(njobs, total, lo=256, hi=1 << 22, seed=0):
    rng = np.random.default_rng(seed)
    raw = np.exp(rng.uniform(math.log(lo), math.log(hi), size=njobs))

    def total_for(s):
        return int(np.clip(np.rint(s * raw), lo, hi).sum())

    a, b = 1e-6, 1e3
    for _ in range(200):
        m = math.sqrt(a * b)
        if total_for(m) < total:
            a = m
        else:
            b = m
        if abs(total_for(m) - total) <= total * 1e-3:
            break
    s = math.sqrt(a * b)
    return np.clip(np.rint(s * raw), lo, hi).astype(np.int64), s


def mix_plan(njobs, total, lo=256, hi=1 << 22, seed=0):
    """List of dicts {kind, width, n, seed} plus the scale factor used."""
    lengths, s = mix_lengths(njobs, total, lo, hi, seed)
    rng = np.random.default_rng(seed + 1)
    kinds = rng.integers(0, 3, size=njobs)
    widths = np.where(rng.integers(0, 2, size=njobs) == 0, 4, 8)
    plan = [{"kind": int(k), "width": int(w), "n": int(n), "seed": int(seed * 1_000_003 + j)}
            for j, (k, w, n) in enumerate(zip(kinds, widths, lengths))]
    return plan, s


class Arena:
    """One torch allocation carved into 256-byte aligned slices."""

    def __init__(self, torch, nbytes):
        self.buf = torch.empty(max(nbytes, 256), dtype=torch.uint8, device="cuda")
        self.off = 0

    @staticmethod
    def need(nbytes):
        return (nbytes + 255) // 256 * 256

    def take(self, torch, n, dtype):
        nbytes = n * torch.empty((), dtype=dtype).element_size()
        t = self.buf[self.off:self.off + max(nbytes, 1)].view(torch.uint8)
        self.off += self.need(nbytes)
        return t[:nbytes].view(dtype) if nbytes else torch.empty(0, dtype=dtype, device="cuda")


def build_mix(edpp, ctx, torch, plan, buckets=256, outputs=1):
    """Allocate and fill every job's inputs on the device; returns
    (list of job-descriptor lists, one per output set, list of buffer dicts)."""
    def sizes(p):
        n, w = p["n"], p["width"]
        if p["kind"] == 0:
            return [n * w] + [n * w, w] * outputs
        if p["kind"] == 1:
            return [n * w, n * 4] + [n * w, n * 4, buckets * 4] * outputs
        return [n * w, n * 4] + [n * w] * outputs

    total = sum(Arena.need(x) for p in plan for x in sizes(p))
    arena = Arena(torch, total)
    jobsets = [[] for _ in range(outputs)]
    bufs = []
    for p in plan:
        n, w, seed = p["n"], p["width"], p["seed"]
        dt = torch.int64 if w == 8 else torch.int32
        b = {"plan": p}
        if p["kind"] == 0:
            b["in"] = arena.take(torch, n, dt)
            (ctx.gen_i64 if w == 8 else ctx.gen_i32)(b["in"], seed, 15)
            for o in range(outputs):
                b[f"out{o}"] = arena.take(torch, n, dt)
                b[f"total{o}"] = arena.take(torch, 1, dt)
                jobsets[o].append(edpp.make_job(edpp.JOB_SCAN, n, b["in"], b[f"out{o}"], w, aux=b[f"total{o}"]))
        elif p["kind"] == 1:
            b["in"] = arena.take(torch, n, dt)
            b["vals"] = arena.take(torch, n, torch.int32)
            ctx.gen_split_keys(b["in"], b["vals"], buckets, edpp.KEYS_ZIPF if seed % 2 else edpp.KEYS_UNIFORM, seed,
                               n=n, wide=w == 8)
            for o in range(outputs):
                b[f"out{o}"] = arena.take(torch, n, dt)
                b[f"vout{o}"] = arena.take(torch, n, torch.int32)
                b[f"start{o}"] = arena.take(torch, buckets, torch.int32)
                jobsets[o].append(edpp.make_job(edpp.JOB_SPLIT, n, b["in"], b[f"out{o}"], w, buckets=buckets,
                                                vals_in=b["vals"], vals_out=b[f"vout{o}"], aux=b[f"start{o}"]))
        else:
            b["in"] = arena.take(torch, n, dt)
            (ctx.gen_i64 if w == 8 else ctx.gen_i32)(b["in"], seed)
            b["idx"] = arena.take(torch, n, torch.int32)
            ctx.gen_perm(b["idx"], seed, edpp.PERM_RANDOM, n=n)
            for o in range(outputs):
                b[f"out{o}"] = arena.take(torch, n, dt)
                jobsets[o].append(edpp.make_job(edpp.JOB_GATHER, n, b["in"], b[f"out{o}"], w, idx=b["idx"]))
        bufs.append(b)
    return jobsets, bufs, arena
