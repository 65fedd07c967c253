"""Benchmark: plain FP64 single-RHS SpTRSV on one MI355X.

Workload (default): forward solve with the lower part of the 5-point
Laplacian on a 1500x1500 grid in natural ordering (n = 2,250,000,
nnz = 6,747,000, 2999 level sets), b ~ U(-1,1) seeded, lv from the auto rule
(4).  A "step" is one solve of one right-hand side with the matrix and b
already resident in HBM.

Metric: GB/s of algorithmic traffic, B = 12*nnz + 4*(m+1) + 16*m bytes per
solve (value + index per entry, the row pointer, reading b and writing x
once), and that as a fraction of the 8 TB/s HBM peak.

Prints ONE JSON line.  --gpus above 1 is refused: this build runs on one GPU.
"""
import argparse
import json
import os
import sys
import time

REPO = os.path.dirname(os.path.abspath(__file__))
sys.path[:0] = [REPO, os.path.join(REPO, "recblock-sptrsv_amd")]

HBM_PEAK_GBS = 8000.0
METRIC = ("SpTRSV GB/s & % HBM roofline; solves/s + p99 latency "
                   "(batched FP32+IR vs FP64)")
TRAFFIC_FILE = os.path.join(REPO, "profiles", "traffic.json")


def measured_traffic(workload, schedule):
    """HBM bytes per launch of the dominant kernel from the committed
    rocprofv3 counter runs (FETCH_SIZE + WRITE_SIZE, each scaled by the
    same counter's reading for a copy kernel of known size), or None when
    no run covers this workload and schedule."""
    try:
        with open(TRAFFIC_FILE) as fh:
            tab = json.load(fh)
    except (OSError, ValueError):
        return None
    e = tab.get(workload)
    if not e or e.get("schedule") != schedule:
        return None
    return e.get("bytes_per_launch")


def algorithmic_bytes(m, nnz, nrhs=1):
    return 12 * nnz + 4 * (m + 1) + 16 * m * nrhs


def make_workload(name):
    from rbsptrsv import gen
    if name == "laplacian":
        return gen.laplacian(1500), "laplacian-1500x1500-lower-forward"
    if name == "laplacian-backward":
        return gen.laplacian(1500, lower=False), "laplacian-1500x1500-upper-backward"
    if name == "powerlaw":
        return gen.powerlaw(4_000_000, seed=0), "powerlaw-4M-zipf2-lower-forward"
    if name == "banded":
        return gen.banded(100_000, seed=0), "banded-100k-w2000-lower-forward"
    raise SystemExit("unknown workload %s" % name)


def host_cpu():
    """(nproc, usable cores, CPU model) of this host."""
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    try:
        usable = len(os.sched_getaffinity(0))
    except AttributeError:
        usable = os.cpu_count()
    return os.cpu_count(), usable, model


def _median_time(fn, budget_s, min_reps=20, warm=3):
    """Median seconds per call after `warm` warm-up calls, over at least
    min_reps timed calls (fewer only if budget_s runs out first)."""
    import numpy as np
    for _ in range(warm):
        fn()
    ts = []
    t_start = time.perf_counter()
    while len(ts) < min_reps or (time.perf_counter() - t_start < 1.0 and
                                 len(ts) < 200):
        t0 = time.perf_counter()
        fn()
        ts.append(time.perf_counter() - t0)
        if time.perf_counter() - t_start > budget_s and len(ts) >= 3:
            break
    return float(np.median(ts)), len(ts)


def f32ir_leg(args, dcp, dri, dv, f, db, K, fp64_ms):
    """FP32 solves refined in FP64 to 1e-12 (rb_solve_f32ir) on the same
    factor and right-hand sides: a plan built with mixed=True under the
    same schedule rule, timed over args.steps calls with their report (one
    host synchronisation per call, after the second correction).  Beside
    the plain FP64 line, not instead of it."""
    import numpy as np
    import torch
    from rbsptrsv import Plan
    q = Plan(dcp, dri, dv, f.substitution, -1, schedule=args.schedule,
             fused_waves=args.fused_waves, mixed=True)
    x, it, rel = q.solve_f32ir(db, tol=1e-12)
    for _ in range(args.warmup):
        q.solve_f32ir(db, x=x, tol=1e-12)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        x, it, rel = q.solve_f32ir(db, x=x, tol=1e-12)
    torch.cuda.synchronize()
    ms = (time.perf_counter() - t0) * 1e3 / args.steps
    q.check()
    sched = {1: "blocked", 2: "fused"}.get(q.info().schedule, q.info().schedule)
    q.close()
    X = x.t().cpu().numpy() if K > 1 else x.cpu().numpy()[None, :]
    B = db.t().cpu().numpy() if K > 1 else db.cpu().numpy()[None, :]
    res = max(float(np.linalg.norm(B[j] - f.matvec(X[j])) /
                    np.linalg.norm(B[j])) for j in range(K))
    return {"ms_per_call": round(ms, 4), "solves_per_s": round(K / (ms * 1e-3), 1),
            "vs_fp64_solve": round(fp64_ms / ms, 3), "corrections": it,
            "relres_max": rel, "residual": res, "tol": 1e-12,
            "schedule": sched, "dtype": "f32 solves, f64 residual and update"}


def cpu_baseline(f, lv, budget_s):
    """The reference's algorithm on the host cores: the oracle's level
    reorder, partition and block order, triangular blocks level by level
    over OpenMP threads (serially when their levels are narrow), square
    blocks as row-parallel SpMV (oracle/rb_cpu.c); timed on all threads
    OpenMP is given and on one.  Beside it the plain serial CSR
    substitution of the original matrix.  Medians after 3 warm-ups."""
    import oracle
    from rbsptrsv import gen
    B = algorithmic_bytes(f.n, f.nnz)
    # timed as the device build is: the first call, then the median of two
    # more
    pre = []
    for _ in range(3):
        tp = time.perf_counter()
        oracle.full_pipeline(f.colptr, f.rowidx, f.val, f.n, f.substitution, lv)
        pre.append((time.perf_counter() - tp) * 1e3)
    pre_ms, pre_cold_ms = (pre[1] + pre[2]) / 2, pre[0]
    solver = oracle.CpuSolver(f, lv)
    b = gen.rhs(f.n, seed=1)
    nthr = oracle.cpu_threads()
    t_all, n_all = _median_time(lambda: solver.solve(b, nthr), budget_s / 3)
    t_one, n_one = _median_time(lambda: solver.solve(b, 1), budget_s / 3)
    solver.close()
    rp, ci, v = f.csr()
    t_csr, n_csr = _median_time(
        lambda: oracle.sptrsv_serial(rp, ci, v, f.n, f.substitution, b),
        budget_s / 3)
    nproc, usable, model = host_cpu()
    gbs = lambda t: round(B / t / 1e9, 4)
    return dict(value=gbs(t_all), unit="GB/s", cores=nthr, kind="port",
                sample="median of %d block-ordered solves of the same "
                       "workload on %d OpenMP threads (%.2f ms each)"
                       % (n_all, nthr, t_all * 1e3),
                ms_per_solve=round(t_all * 1e3, 3),
                one_core={"value": gbs(t_one), "ms_per_solve":
                          round(t_one * 1e3, 3), "reps": n_one},
                serial_csr={"value": gbs(t_csr), "ms_per_solve":
                            round(t_csr * 1e3, 3), "reps": n_csr,
                            "what": "textbook row-oriented substitution of "
                                    "the original CSR, one thread"},
                nproc=nproc, usable_cores=usable, cpu_model=model,
                # the reference's preprocessing restated on the CPU (level
                # analysis, reorder, partition, block extraction and
                # selector statistics), one thread: the mean of the 2nd
                # and 3rd calls, and the first call
                preprocess_ms=round(pre_ms, 1),
                preprocess_cold_ms=round(pre_cold_ms, 1))


def bench_batch(args):
    """64 uneven factors (banded recipe, n log-uniform in [1e3, 5e5], seeds
    per factor), 1..32 right-hand sides each, all solved per call: batched
    FP64 (one union plan), batched FP32 + FP64 refinement to 1e-12, and the
    plain path (one plan per factor, one right-hand side per call)."""
    import numpy as np
    import torch
    from rbsptrsv import Batch, Plan, gen
    def say(*a):
        print("[batch64]", *a, file=sys.stderr, flush=True)

    facs, nrhs = gen.batch_factors(64, seed=3)
    K = int(sum(nrhs))
    say("factors generated: n_total", sum(f.n for f in facs), "rhs", K)
    rng = np.random.default_rng(5)
    Bh = [rng.uniform(-1, 1, (k, f.n)) for f, k in zip(facs, nrhs)]
    bs = [torch.as_tensor(B, device="cuda").t() for B in Bh]
    B_alg = (sum(12 * f.nnz + 4 * (f.n + 1) for f in facs) +
             sum(16 * f.n * k for f, k in zip(facs, nrhs)))

    def timed(fn, steps):
        fn()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            out = fn()
        torch.cuda.synchronize()
        return (time.perf_counter() - t0) / steps * 1e3, out

    def worst_residual(xs):
        w = 0.0
        for f, B, x in zip(facs, Bh, xs):
            X = x.t().cpu().numpy()
            for j in range(B.shape[0]):
                w = max(w, float(np.linalg.norm(B[j] - f.matvec(X[j])) /
                                 np.linalg.norm(B[j])))
        return w

    t0 = time.perf_counter()
    bt = Batch(facs, lv=-1, schedule=args.schedule, tune_rhs=max(nrhs))
    pre_ms = (time.perf_counter() - t0) * 1e3
    info = bt.info()
    say("union plan built: %.0f ms, schedule %d (warm solve ms: blocked "
        "%.3f, fused %.3f)" % (pre_ms, info.schedule, info.tune_ms[0],
                               info.tune_ms[1]))
    ms64, xs = timed(lambda: bt.solve(bs), args.steps)
    say("batched fp64: %.3f ms per batch" % ms64)
    res64 = worst_residual(xs)
    bt.close()
    bm = Batch(facs, lv=-1, schedule=args.schedule, mixed=True,
               tune_rhs=max(nrhs))
    say("mixed union plan built")
    ms32, (xs32, iters, rel) = timed(lambda: bm.solve_f32ir(bs, tol=1e-12),
                                     args.steps)
    say("batched fp32+ir: %.3f ms per batch" % ms32)
    res32 = worst_residual(xs32)
    bm.close()
    plain = None
    if not args.no_plain:
        plans = []
        for j, f in enumerate(facs):
            plans.append(Plan.from_factor(f, lv=-1))
            if j % 16 == 15:
                say("plain plans", j + 1)
        b1 = [b[:, 0].contiguous() for b in bs]
        msp, _ = timed(lambda: [p.solve(b) for p, b in zip(plans, b1)], 3)
        say("plain: %.3f ms per 64 solves" % msp)
        for p in plans:
            p.close()
        plain = {"solves_per_s": round(len(facs) / (msp * 1e-3), 1),
                 "ms_per_64_solves": round(msp, 3),
                 "how": "one plan per factor (auto schedule), one "
                        "right-hand side per call, 64 calls"}
    # the workload's own form (FP32 solves refined in FP64 to 1e-12) is the
    # line's value; batched FP64 and the plain per-factor path beside it
    sps32 = K / (ms32 * 1e-3)
    sps64 = K / (ms64 * 1e-3)
    out = {
        "metric": METRIC,
        "value": round(sps32, 1),
        "unit": "solves/s",
        "n_gpus": 1, "steps": args.steps, "warmup": 1,
        "ms_per_step": round(ms32, 4),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32 solves, f64 residual and update",
        "data": "synthetic (seeded generator)",
        "config": {"workload": "batch-64-banded-1e3..5e5-rhs1..32",
                   "factors": len(facs), "rhs_total": K,
                   "n_total": int(sum(f.n for f in facs)),
                   "nnz_total": int(sum(f.nnz for f in facs)),
                   "schedule": {1: "blocked", 2: "fused"}.get(info.schedule),
                   "levels": info.nlevel, "preprocess_ms": round(pre_ms, 1),
                   "tune_ms": [round(info.tune_ms[0], 3),
                               round(info.tune_ms[1], 3)],
                   "tol": 1e-12},
        "f32ir": {"solves_per_s": round(sps32, 1),
                  "ms_per_batch": round(ms32, 3), "iterations": iters,
                  "relres_max_inf": rel, "residual": res32,
                  "vs_plain_fp64": (round(sps32 / plain["solves_per_s"], 2)
                                    if plain else None)},
        "batched_fp64": {"solves_per_s": round(sps64, 1),
                         "ms_per_batch": round(ms64, 4), "residual": res64,
                         "gbs_algorithmic": round(B_alg / (ms64 * 1e-3) / 1e9, 3),
                         "vs_plain_fp64": (round(sps64 / plain["solves_per_s"], 2)
                                           if plain else None)},
        "plain_fp64": plain,
    }
    print(json.dumps(out), flush=True)


def serve_factors():
    """16 resident factors, n log-spaced in [1e4, 4e6], recipes cycling
    banded / Laplacian lower / power-law."""
    from rbsptrsv import gen
    return gen.serve_factors(16, 1e4, 4e6, seed=100)


def bench_serve(args):
    """Poisson request stream over 16 resident factors (Zipf(1.1) choice),
    served with continuous cross-request batching; the same stream served
    one request per solve for comparison, and FP32 solves refined to 1e-12.
    Each server also serves streams at the other offered rates
    (--serve-rates): throughput, p50 and p99 per rate."""
    import numpy as np
    from rbsptrsv.serve import Server, poisson_requests

    def say(*a):
        print("[serve16]", *a, file=sys.stderr, flush=True)

    facs = serve_factors()
    say("factors", [f.n for f in facs])
    sizes = [f.n for f in facs]
    # the headline rate first, then the other offered rates of the
    # latency-vs-load curve (a stream of at most ~4 s at each)
    rates = [args.rate] + [float(r) for r in args.serve_rates.split(",")
                           if r and float(r) != args.rate]
    streams = {}
    for rate in rates:
        count = args.requests if rate == args.rate else \
            max(200, min(args.requests, int(4 * rate)))
        streams[rate] = poisson_requests(16, sizes, count, rate=rate, seed=9)
    res, curve = {}, {}
    modes = (("batched_fp64", "fp64", 64), ("unbatched_fp64", "fp64", 1),
             ("batched_f32ir", "f32ir", 64))
    if args.serve_modes:
        modes = [m for m in modes if m[0] in args.serve_modes.split(",")]
    for name, mode, mb in modes:
        t0 = time.perf_counter()
        srv = Server(facs, mode=mode, max_batch=mb, streams=args.serve_streams,
                     mem_budget=8 << 30,
                     log=lambda msg, name=name: say(name, msg))
        build_s = time.perf_counter() - t0
        say(name, "server ready in %.1f s" % build_s)
        curve[name] = []
        for rate in rates:
            reqs = streams[rate]
            stats, out = srv.run(reqs)
            for p in srv.plans:
                p.check()
            rels = []
            for r in out[:: max(1, len(out) // 200)]:
                f = facs[r.factor]
                b = np.asarray(reqs[r.rid][1])
                x = r.x.cpu().numpy()
                rels.append(float(np.linalg.norm(b - f.matvec(x)) /
                                  np.linalg.norm(b)))
            rels = np.array(rels)
            stats["residual_sampled_max"] = float(rels.max())  # NaN if any is
            stats["residual_sampled_nonfinite"] = int((~np.isfinite(rels)).sum())
            stats["offered_rate_per_s"] = rate
            stats = {k: (round(v, 4) if isinstance(v, float) and
                         not k.startswith("residual") else v)
                     for k, v in stats.items()}
            say(name, stats)
            curve[name].append({k: stats[k] for k in (
                "offered_rate_per_s", "requests", "solves_per_s", "p50_ms",
                "p99_ms", "mean_batch")})
            if rate == args.rate:
                stats["server_build_s"] = round(build_s, 1)
                res[name] = stats
        srv.close()
    # the workload's own form (batched, FP32 solves refined in FP64) is the
    # line's value when it ran; the others beside it
    key = "batched_f32ir" if "batched_f32ir" in res else next(iter(res))
    b = res[key]
    plain = res.get("unbatched_fp64")
    ratios = {k: round(v["solves_per_s"] / plain["solves_per_s"], 2)
              for k, v in res.items()} if plain else None
    out = {
        "metric": METRIC,
        "value": round(b["solves_per_s"], 1),
        "unit": "solves/s",
        "n_gpus": 1, "steps": args.requests, "warmup": 1,
        "ms_per_step": round(1e3 / b["solves_per_s"], 4),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": ("f32 solves, f64 residual and update" if key == "batched_f32ir"
                  else "f64"),
        "data": "synthetic (seeded generator)",
        "config": {"workload": "serve-16-factors-1e4..4e6-poisson",
                   "server": key, "offered_rate_per_s": args.rate,
                   "requests": args.requests, "zipf": 1.1, "tol": 1e-12},
        "p50_ms": b["p50_ms"], "p99_ms": b["p99_ms"],
        "vs_unbatched_fp64": ratios,
        "servers": res,
        "latency_vs_load": curve,
    }
    print(json.dumps(out), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--workload", default="laplacian")
    ap.add_argument("--schedule", type=int, default=0,
                    help="0 auto, 1 blocked, 2 fused")
    ap.add_argument("--fused-waves", type=int, default=0,
                    help="cap on the fused schedule's consumer waves (0 auto)")
    ap.add_argument("--rhs", type=int, default=1,
                    help="right-hand sides per solve call (column-major)")
    ap.add_argument("--cpu-budget", type=float, default=10.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--f32ir", action="store_true",
                    help="single-factor workloads: also time FP32 solves "
                         "refined in FP64 to 1e-12 on the same factor")
    ap.add_argument("--no-plain", action="store_true",
                    help="batch workload: skip the one-plan-per-factor path")
    ap.add_argument("--rate", type=float, default=16000.0,
                    help="serve workload: offered requests per second (the "
                         "default is above every server's capacity, so the "
                         "line reports capacity)")
    ap.add_argument("--requests", type=int, default=3000,
                    help="serve workload: requests in the stream")
    ap.add_argument("--serve-modes", default="",
                    help="serve workload: comma list of servers to run")
    ap.add_argument("--serve-streams", type=int, default=4,
                    help="serve workload: streams the factors are spread over")
    ap.add_argument("--serve-rates", default="250,1000,4000,32000",
                    help="serve workload: further offered rates (comma "
                         "list) for the latency-vs-load curve")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py runs on one GPU only (this build has no multi-GPU "
              "path); got --gpus %d" % args.gpus, file=sys.stderr)
        sys.exit(2)

    if args.workload == "batch64":
        return bench_batch(args)
    if args.workload == "serve16":
        return bench_serve(args)

    import torch
    from rbsptrsv import Plan, gen

    f, wname = make_workload(args.workload)
    K = max(1, args.rhs)
    b = gen.rhs(f.n, seed=1)
    stream = torch.cuda.current_stream()
    # plan build from device arrays (the reference's preprocess_time also
    # starts from the factor in device memory)
    dcp = torch.as_tensor(f.colptr, device="cuda")
    dri = torch.as_tensor(f.rowidx, device="cuda")
    dv = torch.as_tensor(f.val, device="cuda")
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    plan = Plan(dcp, dri, dv, f.substitution, -1, schedule=args.schedule,
                fused_waves=args.fused_waves)
    pre_cold_ms = (time.perf_counter() - t0) * 1e3
    # two more builds: the first one in a process also loads the kernels'
    # code objects and maps fresh memory
    info = plan.info()
    pre_warm, info_w = [], None
    for _ in range(2):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        q = Plan(dcp, dri, dv, f.substitution, -1, schedule=args.schedule,
                 fused_waves=args.fused_waves)
        pre_warm.append((time.perf_counter() - t0) * 1e3)
        info_w = q.info()
        q.close()
    pre_ms = sum(pre_warm) / len(pre_warm)
    if K == 1:
        db = torch.as_tensor(b, device="cuda")
        dx = torch.empty_like(db)
    else:   # column j = b scaled by (1 + j): distinct, known solutions
        import numpy as np
        Bh = np.stack([b * (1.0 + j) for j in range(K)])
        db = torch.as_tensor(Bh, device="cuda").t()
        dx = torch.empty((K, f.n), dtype=torch.float64, device="cuda").t()
    for _ in range(args.warmup):
        plan.solve(db, dx)
    plan.check()
    torch.cuda.synchronize()
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    e0.record(stream)
    for _ in range(args.steps):
        plan.solve(db, dx)
    e1.record(stream)
    torch.cuda.synchronize()
    plan.check()
    ms = e0.elapsed_time(e1) / args.steps
    # per-launch duration of the dominant kernel (the fused kernel, or the
    # captured block walk): HIP events that the library records on the
    # solve stream around that launch alone
    plan.set_timing(True)
    kms = []
    for _ in range(args.steps):
        plan.solve(db, dx)
        kms.append(plan.kernel_ms())
    launch_ms = sum(kms) / len(kms)
    # the same launch with a cold Infinity Cache (MALL, 256 MB): before each
    # solve a 1 GiB buffer is written on the same stream, which evicts the
    # factor's data from the MALL and the L2s
    flush = torch.empty(1 << 27, dtype=torch.float64, device="cuda")
    cold = []
    for _ in range(max(3, min(args.steps, 20))):
        flush.fill_(1.0)
        plan.solve(db, dx)
        cold.append(plan.kernel_ms())
    del flush
    cold_ms = sum(cold) / len(cold)
    plan.set_timing(False)
    plan.check()

    # correctness of what was timed: relative residual of the last solve
    import numpy as np
    X = dx.t().cpu().numpy() if K > 1 else dx.cpu().numpy()[None, :]
    res = max(float(np.linalg.norm(b * (1.0 + j) - f.matvec(X[j])) /
                    np.linalg.norm(b * (1.0 + j))) for j in range(K))

    B = algorithmic_bytes(f.n, f.nnz, K)
    gbs = B / (ms * 1e-3) / 1e9
    sched = {1: "blocked", 2: "fused"}.get(info.schedule, info.schedule)
    kern_gbs = B / (launch_ms * 1e-3) / 1e9
    out = {
        "metric": METRIC,
        "value": round(gbs, 3),
        "unit": "GB/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms, 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded generator)",
        "solves_per_s": round(K / (ms * 1e-3), 1),
        "config": {"workload": wname, "n": f.n, "nnz": f.nnz, "lv": info.lv,
                   "levels": info.nlevel, "rhs": K,
                   "schedule": sched,
                   "preprocess_ms": round(pre_ms, 1),
                   "residual": res},
        "roofline": {"bound": "hbm", "achieved": round(kern_gbs, 3),
                     "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(kern_gbs / HBM_PEAK_GBS, 5),
                     # counter runs cover one right-hand side
                     "traffic": measured_traffic(wname, sched)
                                if args.rhs == 1 else None,
                     "kernel": "k_fused" if sched == "fused" else
                               "blocked block walk (graph of block kernels)",
                     "launch_ms": round(launch_ms, 5),
                     "algorithmic_bytes_per_launch": B,
                     "cache": "warm: repeated solves, the factor's data "
                              "may sit in the 256 MB Infinity Cache",
                     "cold": {"launch_ms": round(cold_ms, 5),
                              "achieved": round(B / (cold_ms * 1e-3) / 1e9, 3),
                              "frac": round(B / (cold_ms * 1e-3) / 1e9 /
                                            HBM_PEAK_GBS, 5),
                              "how": "1 GiB written on the solve stream "
                                     "before each timed launch"}},
    }
    if K > 1:
        out["metric"] = METRIC
        out["note"] = ("%d right-hand sides per call; bytes count the matrix "
                       "once and b/x per column" % K)
    # plan build throughput (preprocessing as a metric of its own): rows
    # of the factor per second of plan build, GPU build against the CPU
    # restatement of the reference's preprocessing when that is timed
    out["preprocess"] = {"ms": round(pre_ms, 1),
                         "cold_ms": round(pre_cold_ms, 1),
                         "how": "ms: mean of the 2nd and 3rd plan builds in "
                                "the process; cold_ms: the first (it also "
                                "loads the kernels); the CPU port alike",
                         "rows_per_s": round(f.n / (pre_ms * 1e-3), 1),
                         "includes": "level analysis, reorder, partition, "
                                     "blocks, selector, schedule build and "
                                     "tuning",
                         "phases_ms": {k: round(info_w.phase_ms[i], 1) for i, k in
                                       enumerate(("levels", "blocks",
                                                  "fused_schedule", "tuning"))}}
    if args.f32ir:
        out["f32ir"] = f32ir_leg(args, dcp, dri, dv, f, db, K, ms)
    plan.close()    # explicitly, before the interpreter (and any profiler
                    # attached to the HIP runtime) shuts down
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(f, info.lv, args.cpu_budget)
        out["preprocess"]["cpu_port_ms"] = out["cpu_baseline"]["preprocess_ms"]
        out["preprocess"]["cpu_port_cold_ms"] = \
            out["cpu_baseline"]["preprocess_cold_ms"]
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
