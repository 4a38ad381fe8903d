"""Benchmark of the data-parallel primitives on one MI355X.

Default line (the driver's contract): one step = one multi-pass gather plus
one multi-pass scatter of 2^28 int32 values through a full random
permutation (12 algorithmic bytes per element each: index, source and
destination once), inputs resident in HBM, pass count from the library's
schedule.  ``value`` is GB/s (1e9 B/s) of algorithmic bytes.

``--suite`` adds every primitive at its full size (STREAM, scans, gather /
scatter over pass counts and index kinds, stable split over bucket counts and
key skews) under the key "suite" of the same JSON line.

Timing: HIP events recorded on the library context's stream, which is torch's
current stream, around exactly K steps after W warm-up steps; the dominant
kernel's launch time comes from events around its calls inside those steps.
``roofline.traffic`` (HBM bytes per launch) comes from two rocprofv3 --pmc
child runs (FETCH_SIZE, WRITE_SIZE) made before the parent touches the GPU.
"""
import argparse
import json
import os
import platform
import sys
import time

import numpy as np

REPO = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, REPO)

METRIC = "GB/s & % HBM peak at 2^28 int32; batched+learned vs plain per-call, bit-exact"
HBM_PEAK_GBS = 8000.0
N_FULL = 1 << 28


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=10)
    p.add_argument("--warmup", type=int, default=3)
    p.add_argument("--suite", action="store_true", help="measure every primitive too")
    p.add_argument("--batched", action="store_true", help="10,000-job mix: one persistent launch vs per-call")
    p.add_argument("--learned", action="store_true", help="train the schedule model on autotuner timings")
    p.add_argument("--access", action="store_true",
                   help="access-pattern (row/column/mixed x repeat) and workgroup-order (chunk x order x cache) sweeps")
    p.add_argument("--train-calls", type=int, default=2000)
    p.add_argument("--test-calls", type=int, default=500)
    p.add_argument("--db-out", default=None, help="write the autotuner database (JSON lines) here")
    p.add_argument("--headline-only", action="store_true",
                   help="only the gather/scatter step: skip the per-configuration entries (scan, split, batched, "
                        "learned) and the random-access probe")
    p.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    p.add_argument("--no-traffic", action="store_true",
                   help="skip the two rocprofv3 --pmc child runs that measure HBM bytes per launch")
    p.add_argument("--cpu-sample", type=int, default=1 << 28,
                   help="elements of the CPU baseline's gather and scatter (default: the full workload)")
    p.add_argument("--passes", type=int, default=0, help="pass count of the default step (0 = library schedule)")
    p.add_argument("--out", default=None, help="also write the JSON line to this file")
    p.add_argument("--buckets", type=int, default=256, help="bucket count of --profile split")
    p.add_argument("--split-strategy", type=int, default=0,
                   help="strategy of --profile split (0 auto, 1 WG, 2 WI, 3 single, 4 chained) and, when 1 or 4, "
                        "the digit passes of --profile sort")
    p.add_argument("--profile", default=None,
                   choices=["stream", "gather", "scatter", "gather_local", "scatter_local", "scan", "split", "sort",
                            "join", "batched", "mlp", "probe"],
                   help="only run one workload at full size (for rocprofv3), plus one STREAM copy")
    return p.parse_args()


def hampel(x):
    """average_Hampel (opencl/util/utility.cpp:38-83) restated for the report."""
    x = np.asarray(x, np.float64)
    med = np.median(x)
    d = np.abs(x - med)
    mad = np.median(d) / 0.6745
    keep = (d == 0) if mad == 0 else (d / mad <= 3.5)
    return float(x[keep].mean())


class Timer:
    """HIP events on torch's current stream (the library context's stream)."""

    def __init__(self, torch):
        self.torch = torch

    def time(self, fn, reps, warm=3, flush=None):
        """`warm` untimed runs, then `reps` runs each bracketed by events.
        flush, when given, runs before each timed run outside the events (a
        1 GiB scratch write that evicts the 256 MiB Infinity Cache: cold runs)."""
        torch = self.torch
        for _ in range(warm):
            fn()
        evs = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in range(reps)]
        torch.cuda.synchronize()
        for a, b in evs:
            if flush is not None:
                flush()
            a.record()
            fn()
            b.record()
        torch.cuda.synchronize()
        ms = [a.elapsed_time(b) for a, b in evs]
        return {"median_ms": float(np.median(ms)), "hampel_ms": hampel(ms), "min_ms": float(np.min(ms))}


def gbs(nbytes, ms):
    return nbytes / (ms * 1e-3) / 1e9


def cpu_baseline(args, torch_unused=None):
    """The reference's OpenMP gather/scatter (oracle/_ref) or the port, on host cores."""
    from oracle import pyoracle as orc
    n = args.cpu_sample
    idx = orc.gen_perm(n, 0xC0FFEE, 0)
    src = orc.gen_i32(n, 0xBEEF)
    out = np.empty(n, np.int32)
    ref = orc.reference()
    b = orc.baseline()
    threads = b.baseline_threads()
    if ref is not None:
        kind = "reference"
        g = lambda: ref.edpp_ref_gather(orc._a(src), orc._a(out), orc._a(idx), n)
        s = lambda: ref.edpp_ref_scatter(orc._a(src), orc._a(out), orc._a(idx), n)
    else:
        kind = "port"
        g = lambda: b.baseline_gather(orc._a(src), orc._a(out), orc._a(idx), n)
        s = lambda: b.baseline_scatter(orc._a(src), orc._a(out), orc._a(idx), n)
    g(), s()  # first touch, discarded (the reference discards run 0 in its bandwidth tests)
    times = []
    t_end = time.time() + 20.0
    while len(times) < 5 or (time.time() < t_end and len(times) < 40):
        times.append(g() + s())
        if len(times) >= 5 and time.time() > t_end:
            break
    ms = hampel(times)
    cpu = platform.processor() or platform.machine()
    sockets = set()
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name") and cpu == (platform.processor() or platform.machine()):
                    cpu = line.split(":", 1)[1].strip()
                elif line.startswith("physical id"):
                    sockets.add(line.split(":", 1)[1].strip())
    except OSError:
        pass
    return {
        "value": round(gbs(24.0 * n, ms), 3),
        "unit": "GB/s",
        "cores": int(threads),
        "kind": kind,
        "sample": f"one OpenMP gather + one scatter of 2^{int(np.log2(n))} int32 through a random permutation, "
                  f"Hampel mean of {len(times)} runs ({ms:.1f} ms); host {cpu}, nproc={os.cpu_count()}, "
                  f"sockets={len(sockets) or 'unknown'}, OMP_NUM_THREADS={os.environ.get('OMP_NUM_THREADS', 'unset')}, "
                  f"OMP_PROC_BIND={os.environ.get('OMP_PROC_BIND', 'unset')}",
        "primitives": cpu_primitives(orc, ref, b, 1 << 27, src, idx),
    }


def cpu_primitives(orc, ref, b, n, gsrc, gidx):
    """The other host primitives of the CPU baseline: the scans (scan-scan-add
    and reduce-then-scan, openmp/test_scan_CPU.cpp:52-163, restated: that file
    needs oneTBB), copy and scale (openmp/test_bandwidth_CPU.cpp:39-83, the
    reference's own when built) on n int32, and gather and scatter apart on
    the gather/scatter sample (gsrc, gidx).  Run 0 is first touch and
    discarded, then 5 timed runs; GB/s from the Hampel mean."""
    def aligned(count):  # 32-byte aligned int32 array: the streaming-store kernels use aligned AVX2 moves
        raw = np.empty(count + 8, np.int32)
        off = (-raw.ctypes.data % 32) // 4
        return raw[off:off + count]

    src = aligned(n)
    src[:] = orc.gen_i32(n, 0x5CA7) & 15
    out = aligned(n)
    m = gsrc.size
    gout = np.empty(m, np.int32)
    A = orc._a
    if ref is not None:
        copy = lambda: ref.edpp_ref_copy(A(src), A(out), n)
        scale = lambda: ref.edpp_ref_scale(A(src), A(out), n)
        gather = lambda: ref.edpp_ref_gather(A(gsrc), A(gout), A(gidx), m)
        scatter = lambda: ref.edpp_ref_scatter(A(gsrc), A(gout), A(gidx), m)
    else:
        copy = lambda: b.baseline_copy(A(src), A(out), n)
        scale = lambda: b.baseline_scale(A(src), A(out), n)
        gather = lambda: b.baseline_gather(A(gsrc), A(gout), A(gidx), m)
        scatter = lambda: b.baseline_scatter(A(gsrc), A(gout), A(gidx), m)
    res = {}
    legs = [("scan_ssa", lambda: b.baseline_scan_ssa(A(src), A(out), n), 8, n, "port"),
            ("scan_rts", lambda: b.baseline_scan_rts(A(src), A(out), n), 8, n, "port"),
            ("copy", copy, 8, n, "reference" if ref is not None else "port"),
            ("scale", scale, 8, n, "reference" if ref is not None else "port"),
            ("gather", gather, 12, m, "reference" if ref is not None else "port"),
            ("scatter", scatter, 12, m, "reference" if ref is not None else "port")]
    if ref is not None:  # streaming-store copy and scale (openmp/test_bandwidth_CPU.cpp:58-83)
        legs += [("copy_ss", lambda: ref.edpp_ref_copy_ss(A(src), A(out), n), 8, n, "reference"),
                 ("scale_ss", lambda: ref.edpp_ref_scale_ss(A(src), A(out), n), 8, n, "reference")]
    else:
        res["copy_ss"] = res["scale_ss"] = "unavailable: the reference's OpenMP build (oracle/_ref) is absent"
    for name, fn, nb, cnt, kind in legs:
        fn()
        ts = [fn() for _ in range(5)]
        res[name] = {"n": cnt, "gbs": round(gbs(nb * cnt, hampel(ts)), 3),
                     "median_ms": round(float(np.median(ts)), 3), "kind": kind}
        if name.startswith("scan"):  # exclusive prefix, spot-checked at both ends
            head = np.concatenate(([0], np.cumsum(src[:4095], dtype=np.int64))).astype(np.int32)
            tail = int(np.sum(src[:-1], dtype=np.int64) & 0xFFFFFFFF)
            res[name]["checked"] = bool(np.array_equal(out[:4096], head)) and int(out[-1]) & 0xFFFFFFFF == tail
        elif name in ("copy_ss", "scale_ss"):  # whole output against the formula
            res[name]["checked"] = bool(np.array_equal(out, src * (3 if name == "scale_ss" else 1)))
    if os.path.exists("/usr/include/tbb/task_scheduler_init.h"):
        res["tbb_scan"] = "not built: the TBB headers are present but this baseline does not link TBB"
    elif os.path.isdir("/usr/include/tbb") or os.path.isdir("/usr/include/oneapi/tbb"):
        res["tbb_scan"] = "unavailable: the reference's scan_tbb needs tbb/task_scheduler_init.h, absent from oneTBB"
    else:
        res["tbb_scan"] = "unavailable: TBB is not installed (the reference's scan_tbb)"
    return res


def pmc_traffic(kernel_prefix, profile, passes, timeout_s=150):
    """HBM bytes per launch of one kernel from two rocprofv3 --pmc child runs
    (FETCH_SIZE, then WRITE_SIZE, KiB per dispatch) of ``bench.py --profile``.
    Each counter is scaled by what it reads for the STREAM copy kernel of 2^28
    int32 in the same run (1 GiB read, 1 GiB written).  Returns a dict, or
    None with the reason when the profiler is absent or fails."""
    import csv
    import shutil
    import subprocess
    import tempfile
    prof = shutil.which("rocprofv3") or ("/opt/rocm/bin/rocprofv3" if os.path.exists("/opt/rocm/bin/rocprofv3")
                                         else None)
    if prof is None:
        return None, "rocprofv3 not found"
    known = float(N_FULL * 4) / 1024.0  # KiB read (FETCH) or written (WRITE) by the copy
    out = {}
    env = dict(os.environ, TMPDIR="/tmp")
    for ctr in ("FETCH_SIZE", "WRITE_SIZE"):
        d = tempfile.mkdtemp(prefix="edpp_pmc_", dir="/tmp")
        cmd = ["timeout", "-s", "KILL", str(timeout_s), prof, "--pmc", ctr, "--output-format", "csv", "-d", d,
               "-o", "run", "--", sys.executable, os.path.join(REPO, "bench.py"), "--profile", profile,
               "--steps", "2", "--passes", str(passes)]
        r = subprocess.run(cmd, cwd="/tmp", env=env, capture_output=True, text=True)
        rows = []
        for root, _, files in os.walk(d):
            for fn in files:
                if fn.endswith("counter_collection.csv"):
                    with open(os.path.join(root, fn), newline="") as f:
                        rows += [row for row in csv.DictReader(f) if row.get("Counter_Name") == ctr]
        shutil.rmtree(d, ignore_errors=True)
        kern = [float(rw["Counter_Value"]) for rw in rows if rw["Kernel_Name"].startswith(kernel_prefix)]
        copy = [float(rw["Counter_Value"]) for rw in rows if rw["Kernel_Name"].startswith("void edpp::k_stream<0,")]
        if r.returncode != 0 or not kern or not copy:
            return None, f"{ctr} run rc={r.returncode}, {len(kern)} kernel / {len(copy)} copy dispatches"
        factor = known / float(np.mean(copy))
        out[ctr] = {"kib_per_dispatch": float(np.mean(kern)), "copy_kib": float(np.mean(copy)),
                    "factor": round(factor, 4), "bytes": float(np.mean(kern)) * factor * 1024.0}
    return out, None


def profile_only(args, ctx, edpp, torch):
    """One workload at full size, `steps` times, then one STREAM copy of 2^28
    int32 (the calibration kernel for the byte counters)."""
    n = N_FULL
    i32 = torch.int32
    a = torch.empty(n, dtype=i32, device="cuda")
    b = torch.empty(n, dtype=i32, device="cuda")
    ctx.gen_i32(a, 0xDA7A)
    reps = max(1, args.steps)
    if args.profile in ("gather", "scatter", "gather_local", "scatter_local"):
        idx = torch.empty(n, dtype=i32, device="cuda")
        local = args.profile.endswith("_local")
        ctx.gen_perm(idx, 0x5EED, edpp.PERM_BLOCK_LOCAL if local else edpp.PERM_RANDOM)
        f = ctx.gather if args.profile.startswith("gather") else ctx.scatter
        for _ in range(reps):
            f(a, b, idx, n, args.passes)
    elif args.profile == "scan":
        ctx.gen_i32(a, 2024, 15)
        for _ in range(reps):
            ctx.scan(a, b, n, edpp.SCAN_AUTO)
    elif args.profile == "split":
        v = torch.empty(n, dtype=i32, device="cuda")
        vo = torch.empty(n, dtype=i32, device="cuda")
        ctx.gen_split_keys(a, v, args.buckets, edpp.KEYS_UNIFORM, 0x5917)
        for _ in range(reps):
            ctx.split(a, b, args.buckets, v, vo, None, n, args.split_strategy)
    elif args.profile == "sort":
        v = torch.arange(n, dtype=i32, device="cuda")
        vo = torch.empty(n, dtype=i32, device="cuda")
        if args.split_strategy in (edpp.SPLIT_WG, edpp.SPLIT_CHAINED):
            ctx.set_tuning(sort_strategy=args.split_strategy)
        for _ in range(reps):
            ctx.sort(a, b, v, vo, n, 0, 32)
        ctx.set_tuning()
    elif args.profile == "join":
        m = 1 << 24
        ctx.gen_perm(a[:m], 0x1A, edpp.PERM_RANDOM, n=m)
        ctx.gen_perm(b[:m], 0x1B, edpp.PERM_RANDOM, n=m)
        cnt = torch.zeros(1, dtype=torch.int64, device="cuda")
        for _ in range(reps):
            ctx.join_count(a[:m], b[:m], cnt, m, m)
    elif args.profile == "mlp":
        # the cost model's HIP kernels at the learned run's shapes: training on
        # 2000 rows x 1500 epochs, inference on 500 rows, and the per-call
        # selection (features + forward) of AUTO calls with the memo off
        rng = np.random.default_rng(5)
        x = torch.from_numpy(rng.normal(size=(2000, 12)).astype(np.float32)).cuda()
        t = torch.from_numpy(rng.dirichlet(np.ones(15), size=2000).astype(np.float32)).cuda()
        msk = torch.ones(2000, 15, dtype=torch.float32, device="cuda")
        mlp = edpp.MLP(ctx, 7)
        for _ in range(reps):
            mlp.train(x, t, msk, 1500, 3e-3)
        choice = torch.empty(500, dtype=torch.int32, device="cuda")
        for _ in range(100):
            mlp.forward(x[:500], msk[:500], None, choice)
        ctx.set_model(mlp)
        mlp.memo(False)
        m = 1 << 20
        for _ in range(100):
            ctx.scan(a[:m], b[:m], m, edpp.SCAN_AUTO)
        ctx.set_model(None)
        mlp.close()
    elif args.profile == "probe":
        # random 4-byte reads, then writes, over 1 GiB / 128 MiB / 2 MiB tables
        tot = torch.zeros(1, dtype=torch.int64, device="cuda")
        for lg in (28, 25, 19):
            for write in (False, True):
                for _ in range(reps):
                    ctx.access_random(b if write else a, n, 0xA11CE, write, tot, table_n=1 << lg)
    elif args.profile == "batched":
        del a, b
        batched(ctx, edpp, torch)
        a = torch.empty(n, dtype=i32, device="cuda")
        b = torch.empty(n, dtype=i32, device="cuda")
    else:
        for _ in range(reps):
            ctx.stream_op(edpp.STREAM_COPY, a, None, b, n)
    ctx.stream_op(edpp.STREAM_COPY, a, None, b, n)
    torch.cuda.synchronize()
    print(json.dumps({"profile": args.profile, "n": n, "reps": reps}), flush=True)


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {args.gpus}); multi-GPU sharding is design-only.",
              file=sys.stderr)
        sys.exit(2)
    if args.profile:
        import importlib
        import torch
        edpp = importlib.import_module("efficient-dpp_amd")
        torch.cuda.set_device(0)
        ctx = edpp.Context(0, stream=torch.cuda.current_stream().cuda_stream)
        profile_only(args, ctx, edpp, torch)
        ctx.close()
        return
    # HBM bytes per launch of the dominant kernel, from profiler child runs made
    # before this process touches the GPU
    traffic, traffic_note = (None, "skipped (--no-traffic)")
    if not args.no_traffic:
        traffic, traffic_note = pmc_traffic("void edpp::k_gather<", "gather", args.passes)
    import importlib
    import torch
    edpp = importlib.import_module("efficient-dpp_amd")
    torch.cuda.set_device(0)
    ctx = edpp.Context(0, stream=torch.cuda.current_stream().cuda_stream)
    T = Timer(torch)
    n = N_FULL
    i32 = torch.int32

    # ---- inputs of the step, resident in HBM before timing
    idx = torch.empty(n, dtype=i32, device="cuda")
    ctx.gen_perm(idx, 0x5EED, edpp.PERM_RANDOM)
    src = torch.empty(n, dtype=i32, device="cuda")
    ctx.gen_i32(src, 0xDA7A)
    dst = torch.empty(n, dtype=i32, device="cuda")
    dst2 = torch.empty(n, dtype=i32, device="cuda")
    passes = args.passes

    # events on the context's stream: around each step and around each primitive
    # inside it, so the dominant kernel's launch time comes from the timed region
    ev = [[torch.cuda.Event(enable_timing=True) for _ in range(3)] for _ in range(args.steps)]

    def step(k=None):
        if k is not None:
            ev[k][0].record()
        ctx.gather(src, dst, idx, n, passes)
        if k is not None:
            ev[k][1].record()
        ctx.scatter(dst, dst2, idx, n, passes)
        if k is not None:
            ev[k][2].record()

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for k in range(args.steps):
        step(k)
    e1.record()
    torch.cuda.synchronize()
    total_ms = e0.elapsed_time(e1)
    ms_step = total_ms / args.steps
    value = gbs(24.0 * n, ms_step)

    # bit-exact spot check of the step's result (scatter undoes gather)
    assert torch.equal(dst2, src), "gather/scatter round trip mismatch"

    # dominant kernel: the gather launch(es) of the timed steps
    g_ms = float(np.mean([e[0].elapsed_time(e[1]) for e in ev]))
    s_ms = float(np.mean([e[1].elapsed_time(e[2]) for e in ev]))
    g_passes = passes if passes > 0 else ctx.schedule_passes(n, scatter=False)
    s_passes = passes if passes > 0 else ctx.schedule_passes(n, scatter=True)
    achieved = gbs(12.0 * n, g_ms)
    copy = T.time(lambda: ctx.stream_op(edpp.STREAM_COPY, src, None, dst, n), reps=20)
    copy_gbs = gbs(8.0 * n, copy["hampel_ms"])

    line = {
        "metric": METRIC,
        "value": round(value, 3),
        "unit": "GB/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "int32",
        "data": "synthetic (SplitMix64 values, keyed-Feistel random permutation; generated on device)",
        "config": {"workload": "multi-pass gather + scatter, 2^28 int32, full random permutation",
                   "n": n, "gather_passes": g_passes, "scatter_passes": s_passes,
                   "passes_from": "argument" if passes else "library schedule", "index": "random permutation",
                   "bytes_per_element_per_primitive": 12},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": round(traffic["FETCH_SIZE"]["bytes"] + traffic["WRITE_SIZE"]["bytes"]) if traffic
                     else None,
                     "kernel": "k_gather (one launch per pass)",
                     "algorithmic_bytes_per_launch": 12 * n // g_passes,
                     "avg_launch_ms": round(g_ms / g_passes, 4),
                     "vs_stream_copy": round(achieved / copy_gbs, 4), "stream_copy_gbs": round(copy_gbs, 2),
                     "traffic_counters": traffic if traffic else traffic_note},
        "per_primitive": {"gather_gbs": round(achieved, 2), "scatter_gbs": round(gbs(12.0 * n, s_ms), 2),
                          "gather_ms": round(g_ms, 4), "scatter_ms": round(s_ms, 4)},
    }
    del dst2
    if not args.headline_only:
        probe = random_probe(ctx, edpp, torch, src)
        ceil = probe["table_1GiB"]["read"]["gather_equivalent_gbs"]
        line["roofline"]["random_read_ceiling_gbs"] = ceil
        line["roofline"]["frac_of_random_read_ceiling"] = round(achieved / ceil, 4)
        line["random_access_probe"] = probe
        line["configs"] = configs(ctx, edpp, torch, T, src, dst, idx, copy_gbs, line)
    if args.suite:
        line["suite"] = suite(ctx, edpp, torch, T, src, dst, idx, copy_gbs)
    big = args.batched or args.learned or args.access or not args.headline_only
    if big:
        del src, dst, idx
        torch.cuda.empty_cache()
    if args.access:
        line["access"] = access(ctx, edpp, torch, T)
    if args.batched or not args.headline_only:
        bt = batched(ctx, edpp, torch)
        if args.batched:
            line["batched"] = bt
        if not args.headline_only:
            bt["frac_peak"] = round(bt["batched_gbs"] / HBM_PEAK_GBS, 4)
            bt["frac_copy"] = round(bt["batched_gbs"] / copy_gbs, 4)
            line["configs"]["C4_batched_10k_mix"] = bt
    if args.learned or not args.headline_only:
        ln = learned(ctx, edpp, torch, args)
        if args.learned:
            line["learned"] = ln
        if not args.headline_only:
            line["configs"]["C5_learned_schedule"] = ln
    if not args.no_cpu:
        line["cpu_baseline"] = cpu_baseline(args)
    s = json.dumps(line)
    print(s, flush=True)
    if args.out:
        with open(args.out, "w") as f:
            f.write(s + "\n")
    ctx.close()


def random_probe(ctx, edpp, torch, table):
    """Random 4-byte accesses with positions hashed in registers (no index
    stream): 2^28 reads, then 2^28 writes, over tables of 1 GiB (HBM), 128 MiB
    (Infinity Cache) and 2 MiB (each XCD's L2).  `gather_equivalent_gbs` is
    12 B x accesses / time: the rate of a 12 B/element gather (or scatter)
    whose only cost were its random accesses, the ceiling k_gather is graded
    against."""
    n = N_FULL
    tot = torch.zeros(1, dtype=torch.int64, device="cuda")
    scratch = table.clone()  # the probe's writes go to a copy of the step's source
    res = {}
    for lg, name in ((28, "table_1GiB"), (25, "table_128MiB"), (19, "table_2MiB")):
        row = {}
        for write in (False, True):
            tab = scratch if write else table
            ms = []
            ctx.access_random(tab, n, 0xA11CE, write, tot, table_n=1 << lg)
            for _ in range(7):
                ms.append(ctx.access_random(tab, n, 0xA11CE, write, tot, table_n=1 << lg, timed=True))
            t = hampel(ms)
            row["write" if write else "read"] = {"ms": round(t, 4), "g_accesses_s": round(n / (t * 1e-3) / 1e9, 2),
                                                 "gb_s_4byte": round(gbs(4.0 * n, t), 1),
                                                 "gb_s_128byte_lines": round(gbs(128.0 * n, t), 1),
                                                 "gather_equivalent_gbs": round(gbs(12.0 * n, t), 1)}
        res[name] = row
        print("random probe", name, row, file=sys.stderr, flush=True)
    del scratch
    return res


def configs(ctx, edpp, torch, T, src, dst, idx, copy_gbs, line):
    """One entry per benchmark configuration that fits this process's buffers:
    C1 (exclusive scan of 2^24 values in [0,15], warm and cold, and 2^28), C2
    (gather / scatter at 2^28: the default step, every pass count's best, and
    the 4096-block-local permutation) and C3 (stable key/value split of 2^28
    pairs, B in {2,16,256,4096}, uniform and Zipf(1) keys), each in GB/s of
    algorithmic bytes, as a fraction of the 8 TB/s peak and of the measured
    STREAM copy.  C4 and C5 are added by the caller."""
    n = N_FULL
    i32 = torch.int32
    out = {"stream_copy_gbs": round(copy_gbs, 2)}

    def entry(nbytes, t, cold=None):
        g = gbs(nbytes, t["hampel_ms"])
        e = {"gbs": round(g, 2), "ms": round(t["hampel_ms"], 4), "median_ms": round(t["median_ms"], 4),
             "frac_peak": round(g / HBM_PEAK_GBS, 4), "frac_copy": round(g / copy_gbs, 4)}
        if cold is not None:
            gc = gbs(nbytes, cold["hampel_ms"])
            e.update({"cold_gbs": round(gc, 2), "cold_ms": round(cold["hampel_ms"], 4),
                      "cold_frac_peak": round(gc / HBM_PEAK_GBS, 4), "cold_frac_copy": round(gc / copy_gbs, 4)})
        return e

    b = torch.empty(n, dtype=i32, device="cuda")
    # C1: look-back scan (the library's AUTO schedule); cold runs evict the
    # 256 MiB Infinity Cache with a 1 GiB write before each run
    scratch = torch.empty(1 << 28, dtype=i32, device="cuda")
    flush = lambda: scratch.fill_(0x5A5A)
    c1 = {}
    for lg in (24, 28):
        m = 1 << lg
        x = b[:m]
        ctx.gen_i32(x, 2024, 15)
        t = T.time(lambda: ctx.scan(x, dst, m, edpp.SCAN_AUTO), reps=20)
        tc = T.time(lambda: ctx.scan(x, dst, m, edpp.SCAN_AUTO), reps=20, flush=flush)
        c1[f"2^{lg}"] = entry(8.0 * m, t, tc)
    del scratch
    out["C1_scan"] = c1
    print("C1", c1, file=sys.stderr, flush=True)
    # C2: the default step is the headline; add pass counts and block-local
    c2 = {"default_step": {"gbs": line["value"], "gather_gbs": line["per_primitive"]["gather_gbs"],
                           "scatter_gbs": line["per_primitive"]["scatter_gbs"],
                           "frac_copy": round(line["value"] / copy_gbs, 4)}}
    for kind, kname in ((edpp.PERM_RANDOM, "random"), (edpp.PERM_BLOCK_LOCAL, "block4096")):
        ctx.gen_perm(idx, 0x5EED, kind)
        for p in (1, 2, 4, 8, 16):
            tg = T.time(lambda: ctx.gather(src, dst, idx, n, p), reps=5, warm=1)
            ts = T.time(lambda: ctx.scatter(src, dst, idx, n, p), reps=5, warm=1)
            c2[f"{kname}_p{p}"] = {"gather": entry(12.0 * n, tg), "scatter": entry(12.0 * n, ts)}
    ctx.gen_perm(idx, 0x5EED, edpp.PERM_RANDOM)
    out["C2_gather_scatter"] = c2
    print("C2", {k: (v.get("gather", {}).get("gbs"), v.get("scatter", {}).get("gbs")) for k, v in c2.items()},
          file=sys.stderr, flush=True)
    # C3: stable split with bucket offsets, AUTO schedule
    ko = torch.empty(n, dtype=i32, device="cuda")
    vo = torch.empty(n, dtype=i32, device="cuda")
    start = torch.empty(4096, dtype=i32, device="cuda")
    c3 = {}
    for B in (2, 16, 256, 4096):
        for dist, dname in ((0, "uniform"), (1, "zipf")):
            ctx.gen_split_keys(src, b, B, dist, 0x5917 + B)
            t = T.time(lambda: ctx.split(src, ko, B, b, vo, start, n, edpp.SPLIT_AUTO), reps=20)
            c3[f"B{B}_{dname}"] = entry(16.0 * n, t)
    out["C3_split"] = c3
    print("C3", {k: v["frac_copy"] for k, v in c3.items()}, file=sys.stderr, flush=True)
    del ko, vo, b
    ctx.gen_i32(src, 0xDA7A)
    return out


def suite(ctx, edpp, torch, T, src, dst, idx, copy_gbs):
    n = N_FULL
    res = {}
    i32 = torch.int32
    # STREAM (opencl/test/test_bandwidth.cpp: 2^28 int32 per array)
    b = torch.empty(n, dtype=i32, device="cuda")
    ctx.gen_i32(b, 0x11)
    st, st_ms = {}, {}
    for name, op, nb in [("copy", 0, 8), ("scale", 1, 8), ("add", 2, 12), ("triad", 3, 12)]:
        t = T.time(lambda: ctx.stream_op(op, src, b, dst, n), reps=20)
        st[name] = round(gbs(nb * n, t["hampel_ms"]), 2)
        st_ms[name] = {k: round(v, 4) for k, v in t.items()}
    res["stream_gbs"] = st
    res["stream_ms"] = st_ms
    print("stream", st, file=sys.stderr, flush=True)
    # scans: 2^24 (values in [0,15]) and 2^28; warm (back to back) and cold
    # (the 256 MiB Infinity Cache evicted by a 1 GiB write before each run):
    # 2^24 in + out is 128 MiB, so only its cold runs price HBM
    sc = {}
    scratch = torch.empty(1 << 28, dtype=i32, device="cuda")
    flush = lambda: scratch.fill_(0x5A5A)
    for lg in (24, 28):
        m = 1 << lg
        x = b[:m]
        ctx.gen_i32(x, 2024, 15)
        for name, scheme in [("lookback", 1), ("rts", 2), ("ssa", 3), ("persist", 4)]:
            t = T.time(lambda: ctx.scan(x, dst, m, scheme), reps=20)
            tc = T.time(lambda: ctx.scan(x, dst, m, scheme), reps=20, flush=flush)
            g = gbs(8.0 * m, t["hampel_ms"])
            gc = gbs(8.0 * m, tc["hampel_ms"])
            sc[f"2^{lg}_{name}"] = {"gbs": round(g, 2), "ms": round(t["hampel_ms"], 4),
                                    "median_ms": round(t["median_ms"], 4), "min_ms": round(t["min_ms"], 4),
                                    "frac_peak": round(g / HBM_PEAK_GBS, 4), "frac_copy": round(g / copy_gbs, 4),
                                    "cold_gbs": round(gc, 2), "cold_ms": round(tc["hampel_ms"], 4),
                                    "cold_frac_copy": round(gc / copy_gbs, 4)}
    del scratch
    res["scan"] = sc
    print("scan", sc, file=sys.stderr, flush=True)
    # gather / scatter over pass counts and index kinds
    gs = {}
    for kind, kname in [(edpp.PERM_RANDOM, "random"), (edpp.PERM_BLOCK_LOCAL, "block4096")]:
        ctx.gen_perm(idx, 0x5EED, kind)
        for p in (1, 2, 4, 8, 16):
            tg = T.time(lambda: ctx.gather(src, dst, idx, n, p), reps=10)
            ts = T.time(lambda: ctx.scatter(src, dst, idx, n, p), reps=10)
            gs[f"{kname}_p{p}"] = {"gather_gbs": round(gbs(12.0 * n, tg["hampel_ms"]), 2),
                                   "scatter_gbs": round(gbs(12.0 * n, ts["hampel_ms"]), 2),
                                   "gather_ms": {k: round(v, 4) for k, v in tg.items()},
                                   "scatter_ms": {k: round(v, 4) for k, v in ts.items()}}
    ctx.gen_perm(idx, 0x5EED, edpp.PERM_RANDOM)
    res["gather_scatter"] = gs
    print("gather_scatter", gs, file=sys.stderr, flush=True)
    # stable split, 2^28 (key, value) pairs
    sp = {}
    ko = torch.empty(n, dtype=i32, device="cuda")
    vo = torch.empty(n, dtype=i32, device="cuda")
    start = torch.empty(4096, dtype=i32, device="cuda")
    for B in (2, 16, 256, 4096):
        for dist, dname in [(0, "uniform"), (1, "zipf")]:
            ctx.gen_split_keys(src, b, B, dist, 0x5917 + B)
            t = T.time(lambda: ctx.split(src, ko, B, b, vo, start, n, edpp.SPLIT_AUTO), reps=20)
            g = gbs(16.0 * n, t["hampel_ms"])
            sp[f"B{B}_{dname}"] = {"gbs": round(g, 2), "ms": round(t["hampel_ms"], 3),
                                   "median_ms": round(t["median_ms"], 3), "min_ms": round(t["min_ms"], 3),
                                   "frac_copy": round(g / copy_gbs, 4)}
            print("split", B, dname, sp[f"B{B}_{dname}"], file=sys.stderr, flush=True)
    res["split_kv"] = sp
    # the three split strategies (WG, WI, single-lane chunks), B=256 uniform
    ctx.gen_split_keys(src, b, 256, 0, 0x5917 + 256)
    ss = {}
    for sname, strat in [("wg", edpp.SPLIT_WG), ("wi", edpp.SPLIT_WI), ("single", edpp.SPLIT_SINGLE)]:
        t = T.time(lambda: ctx.split(src, ko, 256, b, vo, start, n, strat), reps=3, warm=1)
        g = gbs(16.0 * n, t["hampel_ms"])
        ss[sname] = {"gbs": round(g, 2), "ms": round(t["hampel_ms"], 3), "frac_copy": round(g / copy_gbs, 4)}
    res["split_strategies_B256"] = ss
    print("split strategies", ss, file=sys.stderr, flush=True)
    # local (matrix) scans: one 512-lane workgroup scans 512*TILE int32, 1000
    # launches back to back, as opencl/test/test_scan_local.cpp:6-137 repeats them
    ls = {}
    for tile in (1, 11, 49):
        m = 512 * tile
        x = b[:m]
        ctx.gen_i32(x, 0x10CA1 + tile, 15)
        for mname, mt in [("lm", edpp.MATRIX_LM), ("reg", edpp.MATRIX_REG), ("lm_reg", edpp.MATRIX_LM_REG),
                          ("lm_serial", edpp.MATRIX_LM_SERIAL)]:
            def launches():
                for _ in range(1000):
                    ctx.scan_local(x, dst, tile, 512, mt, edpp.NET_WAVE)
            t = T.time(launches, reps=3, warm=1)
            us = t["hampel_ms"]  # ms per 1000 launches = us per launch
            ls[f"tile{tile}_{mname}"] = {"us_per_launch": round(us, 2), "mkeys_s": round(m / us, 1)}
    res["scan_local_512"] = ls
    print("scan_local", ls, file=sys.stderr, flush=True)
    # LSD radix sort on the split: 2^28 (int32 key, int32 value) pairs, all 32 key bits
    ctx.gen_i32(src, 0x50F7)
    b.copy_(torch.arange(n, dtype=i32, device="cuda"))
    srt = {}
    for name, fn in [("pairs_u32", lambda: ctx.sort(src, ko, b, vo, n, 0, 32)),
                     ("pairs_i32", lambda: ctx.sort(src, ko, b, vo, n, 0, 32, True)),
                     ("keys_u32", lambda: ctx.sort(src, ko, None, None, n, 0, 32))]:
        t = T.time(fn, reps=10)
        per = 16.0 if name.startswith("pairs") else 8.0  # one read and one write of each pair
        srt[name] = {"ms": round(t["hampel_ms"], 3), "median_ms": round(t["median_ms"], 3),
                     "min_ms": round(t["min_ms"], 3), "gkeys_s": round(n / (t["hampel_ms"] * 1e-3) / 1e9, 2),
                     "gbs_one_pass_bytes": round(gbs(per * n, t["hampel_ms"]), 1)}
    res["sort"] = srt
    print("sort", srt, file=sys.stderr, flush=True)
    # partitioned hash join count: two relations of 2^24 and of 2^26 distinct keys
    jn = {}
    cnt = torch.zeros(1, dtype=torch.int64, device="cuda")
    for lg in (24, 26):
        m = 1 << lg
        ctx.gen_perm(ko[:m], 0x1A, edpp.PERM_RANDOM, n=m)
        ctx.gen_perm(vo[:m], 0x1B, edpp.PERM_RANDOM, n=m)
        t = T.time(lambda: ctx.join_count(ko[:m], vo[:m], cnt, m, m), reps=5)
        assert int(cnt.item()) == m, "join count mismatch"
        jn[f"2^{lg}x2^{lg}"] = {"ms": round(t["hampel_ms"], 3), "median_ms": round(t["median_ms"], 3),
                                "min_ms": round(t["min_ms"], 3),
                                "g_tuples_s": round(2 * m / (t["hampel_ms"] * 1e-3) / 1e9, 2)}
    res["hash_join"] = jn
    print("hash_join", jn, file=sys.stderr, flush=True)
    # 64-bit elements (int64 keys and values, int32 index and split values):
    # 2^27 elements, so each 64-bit array is 1 GiB as at 2^28 int32
    m = 1 << 27
    i64 = torch.int64
    wa = src.view(i64)[:m]
    wb = dst.view(i64)[:m]
    ctx.gen_i64(wa, 0x6464, 15)
    wide = {}
    t = T.time(lambda: ctx.scan(wa, wb, m, edpp.SCAN_LOOKBACK, wide=True), reps=20)
    wide["scan_lookback"] = {"gbs": round(gbs(16.0 * m, t["hampel_ms"]), 2), "ms": round(t["hampel_ms"], 4)}
    t = T.time(lambda: ctx.scan(wa, wb, m, edpp.SCAN_AUTO, wide=True), reps=20)
    wide["scan_auto"] = {"gbs": round(gbs(16.0 * m, t["hampel_ms"]), 2), "ms": round(t["hampel_ms"], 4)}
    widx = idx[:m]
    ctx.gen_perm(widx, 0x5EED, edpp.PERM_RANDOM, n=m)
    for p in (1, 4):
        tg = T.time(lambda: ctx.gather(wa, wb, widx, m, p, wide=True), reps=10)
        ts = T.time(lambda: ctx.scatter(wa, wb, widx, m, p, wide=True), reps=10)
        wide[f"random_p{p}"] = {"gather_gbs": round(gbs(20.0 * m, tg["hampel_ms"]), 2),
                                "scatter_gbs": round(gbs(20.0 * m, ts["hampel_ms"]), 2)}
    ctx.gen_perm(idx, 0x5EED, edpp.PERM_RANDOM)
    kv = b[:m]
    kvo = vo[:m]
    for B in (16, 256):
        ctx.gen_split_keys(wa, kv, B, 0, 0x5917 + B, wide=True)
        t = T.time(lambda: ctx.split(wa, wb, B, kv, kvo, start, m, edpp.SPLIT_AUTO, wide=True), reps=10)
        wide[f"split_B{B}"] = {"gbs": round(gbs(24.0 * m, t["hampel_ms"]), 2), "ms": round(t["hampel_ms"], 3)}
    res["int64_2p27"] = wide
    print("int64", wide, file=sys.stderr, flush=True)
    del ko, vo, b
    ctx.gen_i32(src, 0xDA7A)
    return res


def access(ctx, edpp, torch, T):
    """The reference's access studies (opencl/test/test_access.cpp,
    opencl/test/test_access_wg.cpp) on this GPU: out = 3*in through the
    column / row / mixed lane mappings for repeat 1..20 over 512*32768 lanes,
    and chunked reads of 2^28 int32 per workgroup chunk size, in interleaved or
    random workgroup order, with the data cold (a 512 MiB write flushed the
    caches) or loaded (just written)."""
    i32 = torch.int32
    lanes = 512 * 32768
    big = torch.empty(lanes * 20, dtype=i32, device="cuda")
    ctx.gen_i32(big, 0xACC)
    out = torch.empty_like(big)
    res = {"scale": {}}
    for pname, pat in (("column", edpp.ACCESS_COLUMN), ("row", edpp.ACCESS_ROW), ("mixed", edpp.ACCESS_MIXED)):
        row = {}
        for rep in (1, 2, 4, 8, 16, 20):
            n = lanes * rep
            t = T.time(lambda: ctx.access_scale(pat, big, out, rep, n=n), reps=5)
            row[rep] = round(gbs(8.0 * n, t["hampel_ms"]), 1)
        res["scale"][pname] = row
        print("access scale", pname, row, file=sys.stderr, flush=True)
    del big, out
    n = N_FULL
    data = torch.empty(n, dtype=i32, device="cuda")
    ctx.gen_i32(data, 0xACD)
    flush = torch.empty(1 << 27, dtype=i32, device="cuda")  # 512 MiB: twice the Infinity Cache
    cnt = torch.zeros(1, dtype=torch.int64, device="cuda")
    expect = None
    res["chunks"] = {}
    for kb in (8, 64, 512, 4096, 8192):
        chunk = kb * 256  # int32 elements in kb KiB
        chunks = -(-n // chunk)
        orders = {"interleaved": torch.arange(chunks, dtype=i32, device="cuda"),
                  "random": torch.randperm(chunks, device="cuda").to(i32)}
        for oname, order in orders.items():
            for cname in ("cold", "loaded"):
                ms = []
                for _ in range(5):
                    if cname == "cold":
                        flush.fill_(1)
                    else:
                        ctx.stream_op(edpp.STREAM_COPY, data, None, data, n)  # rewrite: the tail is cached
                    ms.append(ctx.access_chunks(data, chunk, order, cnt, timed=True))
                c = int(cnt.item())
                expect = c if expect is None else expect
                assert c == expect, "chunk count changed with the order"
                res["chunks"][f"{kb}KiB_{oname}_{cname}"] = round(gbs(4.0 * n, hampel(ms)), 1)
        print("access chunks", kb, {k: v for k, v in res["chunks"].items() if k.startswith(f"{kb}KiB")},
              file=sys.stderr, flush=True)
    return res


def batched(ctx, edpp, torch):
    """C4: 10,000 jobs (scan / 256-bucket split / gather, int32 or int64,
    log-uniform lengths rescaled to 2^28 elements in total): one persistent
    launch (and its hipGraph replay) against one library call per job."""
    import importlib
    jm = importlib.import_module("efficient-dpp_amd.jobmix")
    plan, scale = jm.mix_plan(10000, 1 << 28, seed=4)
    jobsets, bufs, arena = jm.build_mix(edpp, ctx, torch, plan, outputs=2)
    b = edpp.Batch(ctx, jobsets[0])
    plain = edpp.Batch(ctx, jobsets[1])
    fixed = b.get_shape()  # the default shape, occupancy and queue order the launch uses without tuning

    def med(f, k):
        f()
        return float(np.median([f() for _ in range(k)]))

    t_run = med(lambda: b.run(timed=True), 5)
    t_graph = med(lambda: b.run(timed=True, graph=True), 5)
    t_plain = med(lambda: plain.run_plain(timed=True), 3)
    # the batch's own autotuner: every tile shape x occupancy, median of 3
    tune_t0 = time.time()
    tune = b.autotune(reps=3)
    tune_s = time.time() - tune_t0
    shape, occ, order = b.get_shape()
    t_tuned = med(lambda: b.run(timed=True, graph=True), 5)
    # cold: a 1 GiB write evicts the 256 MiB Infinity Cache before each replay
    scratch = torch.empty(1 << 28, dtype=torch.int32, device="cuda")

    def cold_graph():
        scratch.fill_(0x5A5A)
        return b.run(timed=True, graph=True)

    t_graph_cold = med(cold_graph, 5)
    del scratch
    bad = 0
    for bb in bufs:
        same = torch.equal(bb["out0"], bb["out1"])
        if bb["plan"]["kind"] == 1:
            same = same and torch.equal(bb["vout0"], bb["vout1"]) and torch.equal(bb["start0"], bb["start1"])
        if bb["plan"]["kind"] == 0:
            same = same and torch.equal(bb["total0"], bb["total1"])
        bad += 0 if same else 1
    elems = sum(p["n"] for p in plan)
    nbytes = 0
    for p in plan:
        w = p["width"]
        nbytes += p["n"] * (2 * w if p["kind"] == 0 else (2 * (w + 4) if p["kind"] == 1 else 4 + 2 * w))
    best = min(t_tuned, t_graph)
    out = {"jobs": len(plan), "elements": elems, "length_scale": scale, "tiles": b.info()["tiles"],
           "batched_ms": round(t_run, 4), "graph_ms": round(t_graph, 4), "graph_cold_ms": round(t_graph_cold, 4),
           "autotuned_graph_ms": round(t_tuned, 4),
           "autotuned": {"shape": shape, "wgs_per_cu": occ, "queue_order": edpp.BATCH_ORDERS[order],
                         "tile_shape": edpp.batch_shapes()[shape], "search_s": round(tune_s, 2),
                         "fixed_default_ms": round(tune.get(fixed, -1.0), 4),
                         "all_ms": {f"{edpp.BATCH_ORDERS[k[2]]}_s{k[0]}x{k[1]}": round(v, 4)
                                    for k, v in tune.items()}},
           "plain_per_call_ms": round(t_plain, 3),
           "speedup": round(t_plain / best, 2), "speedup_fixed_default_shape": round(t_plain / t_graph, 2),
           "batched_gbs": round(nbytes / (best * 1e-3) / 1e9, 1),
           "jobs_differing_from_per_call": bad}
    print("batched", out, file=sys.stderr, flush=True)
    b.close()
    plain.close()
    del bufs, arena, jobsets
    torch.cuda.empty_cache()
    return out


def learned(ctx, edpp, torch, args):
    """C5: exhaustive autotune of seeded calls, HIP training of the cost
    model, held-out regret against exhaustive search and end-to-end rate."""
    import importlib
    tn = importlib.import_module("efficient-dpp_amd.tuner")
    t0 = time.time()
    ws = tn.Workspace(edpp, ctx, torch, N_FULL)
    log = lambda m: print(m, file=sys.stderr, flush=True)
    train_recs = tn.measure(ws, tn.sample_specs(args.train_calls, seed=2024), reps=3, log=log)
    test_recs = tn.measure(ws, tn.sample_specs(args.test_calls, seed=777), reps=3, log=log)
    if args.db_out:
        tn.save_db(args.db_out, train_recs + test_recs)
    t1 = time.time()
    mlp, losses = tn.train(edpp, ctx, torch, train_recs, epochs=1500, lr=3e-3)
    t2 = time.time()
    pred = tn.predict(ctx, torch, mlp, test_recs)
    model = tn.regret(test_recs, pred)
    default = tn.regret(test_recs, tn.default_choices(test_recs))
    train_fit = tn.regret(train_recs, tn.predict(ctx, torch, mlp, train_recs))
    specs = [r["spec"] for r in test_recs]
    ctx.set_model(None)
    e2e_default = tn.end_to_end(ws, ctx, specs, False)
    ctx.set_model(mlp)
    mlp.memo(False)  # every call runs the selection kernel and waits for it
    e2e_model_cold = tn.end_to_end(ws, ctx, specs, True)
    mlp.memo(True)  # repeated calls reuse the model's decision
    e2e_model = tn.end_to_end(ws, ctx, specs, True)
    ctx.set_model(None)
    auto = tn.regret(test_recs, tn.auto_choices(test_recs))
    best_bytes = sum(tn.call_bytes(s) for s in specs)
    out = {"train_calls": len(train_recs), "test_calls": len(test_recs), "autotune_s": round(t1 - t0, 1),
           "train_s": round(t2 - t1, 2), "loss_first": float(losses[0]), "loss_last": float(losses[-1]),
           "regret_model": model, "regret_default": default, "regret_builtin_schedule": auto,
           "regret_model_on_train": train_fit,
           "exhaustive_gbs": round(best_bytes / (model["ms_best"] * 1e-3) / 1e9, 1),
           "end_to_end_model": e2e_model, "end_to_end_model_every_call_selects": e2e_model_cold,
           "end_to_end_default": e2e_default,
           "best_candidate_histogram": tn.summarize_times(train_recs)}
    print("learned", json.dumps(out), file=sys.stderr, flush=True)
    mlp.close()
    return out


if __name__ == "__main__":
    main()
