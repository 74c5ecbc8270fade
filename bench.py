#!/usr/bin/env python3
"""Benchmark of the periodic Poisson solve on one MI355X.

Workload (the plain-method configuration): FP64 solve of the benchmark's
analytic f at 1024^3 points, 1x1x1 tasks.  A step is one full solve (five
passes) of the same HBM-resident f; the plan is created with
FF_PRESERVE_INPUT so back-to-back solves see identical input, exactly like the
reference's loop which regenerates the same f before every solve
(FishFry.hpp:173-180).

Prints one JSON line on stdout (progress goes to stderr):
  value        whole-job points/s = N * steps / wall time of the timed steps
  p50/p99      per-solve latency from HIP events around each solve
  roofline     dominant pass kernel: algorithmic bytes / average duration
               (HIP events on the launch stream over a timed pass loop)
  passes       the same for all five passes
  fp32_refine  FP32 passes + FP64 residual, 0..3 refinement steps on the
               analytic f plus 16 random modes: points/s and Linf beside the
               FP64 solve of the same f
  fd_symbols   the same Pareto with the 3-point and 5-point finite-difference
               symbols (FP64 stencil residual), beside FP64 solves of the
               same symbol
  cpu_baseline the CPU oracle's FP64 solve on the host cores (bounded sample)
  extras       the other configurations, measured beside the headline: FP64
               solves at 512^3, 768^3 and 1000^3; the 64-grid batching
               workload (FP64 and FP32 + 2 steps); the Poisson serving
               workload (p50/p99, points/s) with the default policy and,
               at the same arrivals, with large requests nearly alone;
               same-shape bursts through the server
"""
import argparse
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E, 8 TB/s
METRIC = "points/s at 1024³ (FP64 vs FP32+refine, same L∞); p50/p99 solve latency"


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def pct(sorted_vals, q):
    if not sorted_vals:
        return float("nan")
    pos = q * (len(sorted_vals) - 1)
    lo = int(pos)
    hi = min(lo + 1, len(sorted_vals) - 1)
    return sorted_vals[lo] + (pos - lo) * (sorted_vals[hi] - sorted_vals[lo])


def pass_bytes(n, prec_bytes=8):
    """Algorithmic HBM bytes per launch of each pass (read + write once)."""
    N = n[0] * n[1] * n[2]
    Nh = n[0] * n[1] * (n[2] // 2 + 1)
    c = 2 * prec_bytes  # complex element
    return [
        ("z_forward_r2c", 8 * N + c * Nh),
        ("y_forward", 2 * c * Nh),
        ("x_forward_filter_inverse", 2 * c * Nh),
        ("y_inverse", 2 * c * Nh),
        ("z_inverse_c2r_scale", c * Nh + 8 * N),
    ]


def random_modes(seed, count=16, kmax=12):
    """count distinct integer wave vectors with 1 <= |k| <= kmax (+-k counted
    once) and N(0,1) cosine / sine amplitudes."""
    import numpy as np

    rng = np.random.default_rng(seed)
    out, seen = [], set()
    while len(out) < count:
        k = tuple(int(v) for v in rng.integers(-kmax, kmax + 1, size=3))
        if k == (0, 0, 0) or not (1 <= math.sqrt(sum(v * v for v in k)) <= kmax):
            continue
        if k in seen or tuple(-v for v in k) in seen:
            continue
        seen.add(k)
        out.append((k, float(rng.standard_normal()), float(rng.standard_normal())))
    return out


def host_cpus():
    """CPUs this process may run on: the affinity mask, capped by a cgroup
    CPU quota where one is set; plus nproc and the CPU model for the record."""
    nproc = os.cpu_count() or 1
    try:
        aff = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        aff = nproc
    quota = None
    try:
        with open("/sys/fs/cgroup/cpu.max") as fh:
            q, per = fh.read().split()[:2]
            if q != "max":
                quota = max(1, int(math.ceil(int(q) / int(per))))
    except (OSError, ValueError):
        pass
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    use = min(aff, quota) if quota else aff
    return use, nproc, aff, quota, model


def cpu_baseline(size, budget_s=30.0, runs=10):
    """The CPU oracle's FP64 solve of the benchmark f at size^3 on every CPU
    this process may use: 1 warmup, then up to `runs` timed solves within
    budget_s seconds (the sample is bounded so the benchmark stays short)."""
    import oracle  # the CPU oracle: timed here as the reported baseline only

    threads, nproc, aff, quota, model = host_cpus()
    n = (size, size, size)
    f = oracle.init_field(n)
    times = []
    oracle.solve(f, nthreads=threads)  # warmup
    t_start = time.perf_counter()
    while len(times) < runs and (not times or time.perf_counter() - t_start < budget_s):
        t0 = time.perf_counter()
        phi = oracle.solve(f, nthreads=threads)
        times.append(time.perf_counter() - t0)
    l1, l2, linf = oracle.norms(phi)
    del f, phi
    ts = sorted(times)
    t = pct(ts, 0.5)
    return {
        "value": float(n[0] * n[1] * n[2] / t),
        "unit": "points/s",
        "cores": threads,
        "kind": "port",
        "sample": f"CPU oracle FP64 solve (same algorithm: R2C z, C2C y, C2C x, filter, inverses) of the "
                  f"benchmark f at {size}^3, OpenMP over lines on {threads} threads (nproc {nproc}, affinity {aff}, "
                  f"cgroup CPU quota {quota if quota else 'none'}; {model}), 1 warmup + {len(times)} timed "
                  f"(at most {runs} within {budget_s:.0f} s), p50 {t:.3f} s, p99 {pct(ts, 0.99):.3f} s, "
                  f"Linf {linf:.2e}",
    }


def bench_sizes(ff, torch, sizes, K):
    """FP64 solves of the analytic f at other sizes of the plain-method
    configuration: ms per solve, points/s and Linf."""
    out = []
    for n in sizes:
        lo, hi, _ = ff.benchmark_grid(n)
        plan = ff.ParisPencil(n, lo, hi, flags=ff.FF_PRESERVE_INPUT)
        nb = plan.bytes()
        f = torch.empty(nb // 8, dtype=torch.float64, device="cuda")
        phi = torch.empty(nb // 8, dtype=torch.float64, device="cuda")
        field = ff.field_default(n)
        ff.field_init(field, f)
        plan.solve_async(nb, f, phi)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(K):
            plan.solve_async(nb, f, phi)
        torch.cuda.synchronize()
        ms = 1e3 * (time.perf_counter() - t0) / K
        linf = ff.field_norms(field, phi)[2]
        N = n[0] * n[1] * n[2]
        out.append({"n": list(n), "ms": round(ms, 4), "points_per_s": N / (ms * 1e-3), "linf": linf,
                    "hbm_frac": round(sum(b for _, b in pass_bytes(n)) / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)})
        log(f"FP64 {n}: {ms:.3f} ms, {N / (ms * 1e-3):.4e} points/s, Linf {linf:.3e}")
        plan.close()
        del f, phi
        torch.cuda.empty_cache()
    return out


def bench_batching(ff, torch, K):
    """The batching configuration: 64 grids drawn (seeded) from {32^3, 48^3,
    64^3, 96^3, 128^3, 160^3, 128x64x96}, f_b(x) = s_b f(x + delta_b), one
    HIP-graph batch per precision, autotuned, then the same batch grouped (one
    launch per pass for each shape) and step by step (one multi-shape launch
    per pass step over all 64 grids); aggregate points/s, launches per batch
    and the worst Linf against phi_b = s_b phi(x + delta_b)."""
    import numpy as np

    sizes = [(32, 32, 32), (48, 48, 48), (64, 64, 64), (96, 96, 96), (128, 128, 128), (160, 160, 160),
             (128, 64, 96)]
    rng = np.random.default_rng(20261202)
    shapes = [sizes[i] for i in rng.integers(0, len(sizes), 64)]
    fields = []
    for n in shapes:
        fd = ff.field_default(n)
        for a in range(3):
            fd.shift[a] = float(rng.uniform(0, 2 * math.pi))
        fd.scale = float(rng.uniform(0.5, 2.0))
        fields.append(fd)
    points = sum(n[0] * n[1] * n[2] for n in shapes)
    out = []
    stream = torch.cuda.Stream()  # graph capture needs a stream of its own
    for precision, steps in ((ff.FF_FP64, 0), (ff.FF_FP32_REFINE, 2)):
        items = [{"n": n, "precision": precision, "refine_steps": steps} for n in shapes]
        b = ff.Batch(items, ff.FF_BATCH_GRAPH)
        d = [torch.zeros(b.bytes(i) // 8, dtype=torch.float64, device="cuda") for i in range(64)]
        p = [torch.zeros(b.bytes(i) // 8, dtype=torch.float64, device="cuda") for i in range(64)]
        torch.cuda.synchronize()
        tuned, before, after = b.autotune(d, p, reps=2, stream=stream)
        stream.synchronize()
        for k in range(64):
            ff.field_init(fields[k], d[k])
        torch.cuda.synchronize()
        b.solve(d, p, stream)
        b.wait(stream)
        stream.synchronize()
        worst = max(ff.field_norms(fields[k], p[k])[2] for k in range(64))
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(K):
            b.solve(d, p, stream)
        b.wait(stream)
        stream.synchronize()
        ms = 1e3 * (time.perf_counter() - t0) / K
        b.close()
        # the same batch grouped: one launch per pass for each shape (tuned configurations stay in force)
        g = ff.Batch(items, ff.FF_BATCH_GRAPH | ff.FF_BATCH_GROUPED)
        for k in range(64):
            ff.field_init(fields[k], d[k])
        torch.cuda.synchronize()
        g.solve(d, p, stream)
        g.wait(stream)
        stream.synchronize()
        gworst = max(ff.field_norms(fields[k], p[k])[2] for k in range(64))
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(K):
            g.solve(d, p, stream)
        g.wait(stream)
        stream.synchronize()
        gms = 1e3 * (time.perf_counter() - t0) / K
        gdesc = json.loads(g.describe())
        ngroups = len(gdesc["groups"])
        glaunches = sum(x["launches"] for x in gdesc["groups"])
        g.close()
        # the same batch step by step: one multi-shape launch per pass step
        # over all 64 instances (descriptor table), captured in a graph.  The
        # autotuner's choices target the single-solve kernels (some move a
        # pass off the multi-shape variants), so this batch plans with the
        # default configurations.
        ff.tune_reset()
        sb = ff.Batch(items, ff.FF_BATCH_GRAPH | ff.FF_BATCH_STEPS)
        for k in range(64):
            ff.field_init(fields[k], d[k])
        torch.cuda.synchronize()
        sb.solve(d, p, stream)
        sb.wait(stream)
        stream.synchronize()
        sworst = max(ff.field_norms(fields[k], p[k])[2] for k in range(64))
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(K):
            sb.solve(d, p, stream)
        sb.wait(stream)
        stream.synchronize()
        sms = 1e3 * (time.perf_counter() - t0) / K
        fams = json.loads(sb.describe())["steps"]
        slaunches = sum(x["launches"] for x in fams)
        sb.close()
        out.append({"precision": "fp64" if precision == ff.FF_FP64 else "fp32+2", "grids": 64, "points": points,
                    "ms_per_batch": round(ms, 4), "points_per_s": points / (ms * 1e-3), "worst_linf": worst,
                    "launches": len(shapes) * (5 if precision == ff.FF_FP64 else 5 + 10 * steps),
                    "autotuned_passes": tuned, "ms_before_tuning": round(before, 4), "ms_after_tuning": round(after, 4),
                    "grouped": {"groups": ngroups, "launches": glaunches, "ms_per_batch": round(gms, 4),
                                "points_per_s": points / (gms * 1e-3), "worst_linf": gworst},
                    "steps": {"launches": slaunches, "ms_per_batch": round(sms, 4),
                              "points_per_s": points / (sms * 1e-3), "worst_linf": sworst}})
        log(f"batch {out[-1]['precision']}: {ms:.3f} ms per 64-grid batch, {points / (ms * 1e-3):.4e} points/s, "
            f"worst Linf {worst:.3e}, {tuned} passes tuned ({before:.3f} -> {after:.3f} ms); grouped ({ngroups} "
            f"groups, {glaunches} launches): {gms:.3f} ms, {points / (gms * 1e-3):.4e} points/s, worst Linf "
            f"{gworst:.3e}; steps ({slaunches} launches): {sms:.3f} ms, {points / (sms * 1e-3):.4e} points/s, "
            f"worst Linf {sworst:.3e}")
        del d, p
        torch.cuda.empty_cache()
    return out


def bench_serving_grouped(ff, nreq=128):
    """Same-shape bursts through the server: nreq requests of one small grid
    submitted at once, with dispatch groups of one (every request its own
    launches) and of eight (each pass step one grouped launch for the
    group).  Reports requests/s and p50 for each."""
    out = []
    for n, prec, steps in (((48, 48, 48), ff.FF_FP64, 0), ((64, 64, 64), ff.FF_FP32_REFINE, 2)):
        row = {"n": list(n), "precision": "fp64" if prec == ff.FF_FP64 else "fp32+%d" % steps}
        for mb in (1, 8):
            srv = ff.Server(hbm_budget=8 << 30, streams=4, max_batch=mb, batch_window_us=200.0)
            fld = ff.field_default(n)
            srv.wait(srv.submit(n, prec, steps, field=fld))  # plan cache
            best = None
            for _ in range(3):
                t0 = time.perf_counter()
                ids = [srv.submit(n, prec, steps, field=fld) for _ in range(nreq)]
                for rid in ids:
                    srv.wait(rid)
                dt = time.perf_counter() - t0
                best = dt if best is None else min(best, dt)
            st = srv.stats()
            srv.close()
            row["groups_of_%d" % mb] = {"requests_per_s": round(nreq / best, 1), "p50_ms": round(1e3 * st["p50_s"], 4),
                                        "grouped_solves": st["grouped_solves"]}
        log(f"serving bursts {row['precision']} {n}: {nreq} requests, groups of 1 "
            f"{row['groups_of_1']['requests_per_s']:.0f}/s (p50 {row['groups_of_1']['p50_ms']:.2f} ms), groups of 8 "
            f"{row['groups_of_8']['requests_per_s']:.0f}/s (p50 {row['groups_of_8']['p50_ms']:.2f} ms)")
        out.append(row)
    return out


def bench_serving(ff, nreq=96, seed=20261202, flags=0, rate=None):
    """The serving configuration: Poisson arrivals, edges log-uniform in
    [32, 512] (any integer; half cubic), half FP64 and half FP32 + 2
    refinement steps, f = the analytic field plus 16 random modes; the arrival
    rate (unless given) is 75 % of the throughput a closed-loop burst of the same mix
    reaches.  Reports p50/p99 latency, points/s and the worst Linf."""
    import numpy as np

    rng = np.random.default_rng(seed)

    def draw():
        if rng.random() < 0.5:
            e = int(round(math.exp(rng.uniform(math.log(32), math.log(512)))))
            return (e, e, e)
        return tuple(int(round(math.exp(rng.uniform(math.log(32), math.log(512))))) for _ in range(3))

    def request(k):
        n = draw()
        prec = ff.FF_FP64 if k % 2 == 0 else ff.FF_FP32_REFINE
        modes = random_modes(int(rng.integers(1 << 30)))
        return n, prec, ff.field_set_modes(ff.field_default(n), modes)

    # size-class groups of at most two requests, gathered for 100 us: each
    # request runs its own solve, so larger groups only queue their later
    # members behind the earlier ones on one stream.  flags 0: the default
    # policy; FF_SERVE_LARGE_ALONE is measured at the default policy's rate,
    # so that both see the same arrivals
    cfg = dict(hbm_budget=48 << 30, streams=4, max_batch=2, batch_window_us=100.0, flags=flags)
    srv = ff.Server(**cfg)
    # closed-loop burst: capacity in requests/s for this mix
    burst = [request(k) for k in range(32)]
    t0 = time.perf_counter()
    ids = [srv.submit(n, prec, 2, field=fld) for n, prec, fld in burst]
    for rid in ids:
        srv.wait(rid, timeout=300.0)
    cap = len(burst) / (time.perf_counter() - t0)
    srv.close()
    rate = rate or 0.75 * cap
    srv = ff.Server(**cfg)
    reqs = [request(k) for k in range(nreq)]
    arrivals = np.cumsum(rng.exponential(1.0 / rate, nreq))
    t0 = time.perf_counter()
    ids = []
    for (n, prec, fld), ta in zip(reqs, arrivals):
        dt = ta - (time.perf_counter() - t0)
        if dt > 0:
            time.sleep(dt)
        ids.append((prec, srv.submit(n, prec, 2, field=fld, want_norms=True)))
    worst = {ff.FF_FP64: 0.0, ff.FF_FP32_REFINE: 0.0}
    for prec, rid in ids:
        worst[prec] = max(worst[prec], srv.wait(rid, timeout=300.0)["norms"][2])
    st = srv.stats()
    srv.close()
    res = {"requests": nreq, "policy": "large_alone" if flags else "default", "arrival_rate_per_s": round(rate, 2),
           "burst_capacity_per_s": round(cap, 2),
           "p50_ms": round(1e3 * st["p50_s"], 4), "p99_ms": round(1e3 * st["p99_s"], 4),
           "points_per_s": st["points_per_s"], "groups": st["groups"], "arena_peak_bytes": st["arena_peak_bytes"],
           "worst_linf_fp64": worst[ff.FF_FP64], "worst_linf_fp32_refine2": worst[ff.FF_FP32_REFINE]}
    log(f"serving ({res['policy']}): {nreq} requests at {rate:.1f}/s (burst capacity {cap:.1f}/s): "
        f"p50 {res['p50_ms']:.2f} ms, "
        f"p99 {res['p99_ms']:.2f} ms, {st['points_per_s']:.4e} points/s, worst Linf {worst}")
    return res


def timed_solves(torch, plan, nb, f, phi, K):
    """K back-to-back solves after one warmup: (points/s, sorted latencies ms)."""
    stream = torch.cuda.current_stream()
    plan.solve_async(nb, f, phi)
    torch.cuda.synchronize()
    ev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in range(K)]
    ta = time.perf_counter()
    for k in range(K):
        ev[k][0].record(stream)
        plan.solve_async(nb, f, phi)
        ev[k][1].record(stream)
    torch.cuda.synchronize()
    tb = time.perf_counter()
    N = plan.n[0] * plan.n[1] * plan.n[2]
    return N * K / (tb - ta), sorted(a.elapsed_time(b) for a, b in ev)


def precision_pareto(ff, torch, n, K, rfield, flags, value64):
    """FP32 passes with 0..3 FP64-residual refinement steps beside the FP64
    solve of the same f and symbol (flags: 0 spectral, or FF_SYMBOL_3PT /
    FF_SYMBOL_5PT): points/s and Linf against the exact potential.  value64:
    the FP64 rate to compare with (None: measured here)."""
    lo, hi, _ = ff.benchmark_grid(n)
    p64 = ff.ParisPencil(n, lo, hi, flags=flags | ff.FF_PRESERVE_INPUT)
    nb = p64.bytes()
    f = torch.empty(nb // 8, dtype=torch.float64, device="cuda")
    phi = torch.empty(nb // 8, dtype=torch.float64, device="cuda")
    ff.field_init(rfield, f)
    v64, _ = timed_solves(torch, p64, nb, f, phi, K)
    linf64 = ff.field_norms(rfield, phi)[2]
    p64.close()
    if value64 is None:
        value64 = v64
    tag = "spectral" if not flags else ("3-point" if flags == ff.FF_SYMBOL_3PT else "5-point")
    log(f"{tag} symbol, random-mode f: FP64 {v64:.4e} points/s, Linf {linf64:.3e}")
    levels = []
    for steps in (0, 1, 2, 3):
        p32 = ff.ParisPencil(n, lo, hi, precision=ff.FF_FP32_REFINE, refine_steps=steps, flags=flags)
        pts, l32 = timed_solves(torch, p32, nb, f, phi, K)
        r = ff.field_norms(rfield, phi)
        levels.append({"refine_steps": steps, "points_per_s": pts, "p50_ms": round(pct(l32, .5), 4),
                       "p99_ms": round(pct(l32, .99), 4), "l1": r[0], "l2": r[1], "linf": r[2],
                       "speed_vs_fp64": round(pts / value64, 4), "linf_vs_fp64": round(r[2] / linf64, 4)})
        log(f"  FP32+refine steps={steps}: {pts:.4e} points/s ({pts / value64:.3f}x FP64), Linf {r[2]:.3e} "
            f"({r[2] / linf64:.3g}x FP64)")
        p32.close()
    del f, phi
    torch.cuda.empty_cache()
    return {"fp64_points_per_s": v64, "fp64_linf": linf64, "levels": levels}


def main():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--size", type=int, nargs=3, default=[1024, 1024, 1024])
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU oracle baseline")
    ap.add_argument("--no-fp32", action="store_true", help="skip the FP32+refinement Pareto")
    ap.add_argument("--cpu-size", type=int, default=1024)
    ap.add_argument("--no-extras", action="store_true",
                    help="skip the other plain-method sizes, the batching and the serving workloads")
    args = ap.parse_args()

    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py: this build runs on exactly one GPU (no collectives, no sharding); "
              "run it with --gpus 1", file=sys.stderr)
        return 2

    import torch

    import fishfry_amd as ff

    torch.cuda.set_device(0)
    stream = torch.cuda.current_stream()
    n = tuple(args.size)
    N = n[0] * n[1] * n[2]
    lo, hi, _ = ff.benchmark_grid(n)
    K, W = args.steps, args.warmup

    # ---- headline: FP64 solve, input resident and preserved -----------------
    plan = ff.ParisPencil(n, lo, hi, flags=ff.FF_PRESERVE_INPUT)
    nb = plan.bytes()
    log(f"plan {plan.describe()}")
    f = torch.empty(nb // 8, dtype=torch.float64, device="cuda")
    phi = torch.empty(nb // 8, dtype=torch.float64, device="cuda")
    field = ff.field_default(n)
    ff.field_init(field, f)
    for _ in range(W):
        plan.solve_async(nb, f, phi)
    torch.cuda.synchronize()
    evs = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in range(K)]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for k in range(K):
        evs[k][0].record(stream)
        plan.solve_async(nb, f, phi)
        evs[k][1].record(stream)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    wall = t1 - t0
    lat = sorted(a.elapsed_time(b) for a, b in evs)
    value = N * K / wall
    l1, l2, linf = ff.field_norms(field, phi)
    log(f"FP64 {n}: {value:.4e} points/s, {1e3 * wall / K:.3f} ms/solve, p50 {pct(lat, .5):.3f} ms, "
        f"p99 {pct(lat, .99):.3f} ms, L1 {l1:.3e} L2 {l2:.3e} Linf {linf:.3e}")

    # ---- per-pass timing: the same five kernels launched one by one -----------
    work = torch.empty(nb // 8, dtype=torch.float64, device="cuda")
    bufs = [(f, phi), (phi, work), (work, phi), (phi, work), (work, phi)]
    for p, (s, d) in enumerate(bufs):
        plan.solve_pass(p, s, d)
    torch.cuda.synchronize()
    pev = [[torch.cuda.Event(enable_timing=True) for _ in range(6)] for _ in range(K)]
    for k in range(K):
        pev[k][0].record(stream)
        for p, (s, d) in enumerate(bufs):
            plan.solve_pass(p, s, d)
            pev[k][p + 1].record(stream)
    torch.cuda.synchronize()
    passes = []
    for p, (name, nbytes) in enumerate(pass_bytes(n)):
        ms = sum(pev[k][p].elapsed_time(pev[k][p + 1]) for k in range(K)) / K
        gbs = nbytes / (ms * 1e-3) / 1e9
        passes.append({"kernel": name, "ms": round(ms, 4), "bytes": nbytes, "GB/s": round(gbs, 1),
                       "frac": round(gbs / HBM_PEAK_GBS, 4)})
        log(f"  pass {p} {name:28s} {ms:8.3f} ms  {gbs:8.1f} GB/s  {100 * gbs / HBM_PEAK_GBS:5.1f}% of HBM peak")
    del work
    dom = max(passes, key=lambda r: r["ms"])
    traffic = None
    tpath = os.path.join(ROOT, "profiles", "traffic_1024.json")
    if tuple(n) == (1024, 1024, 1024) and os.path.exists(tpath):
        with open(tpath) as fh:
            traffic = json.load(fh).get(dom["kernel"])
    roofline = {"bound": "hbm", "achieved": dom["GB/s"], "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": dom["frac"], "traffic": traffic, "kernel": dom["kernel"],
                "algorithmic_bytes_per_launch": dom["bytes"], "avg_launch_ms": dom["ms"]}
    solve_bytes = sum(b for _, b in pass_bytes(n))
    plan.close()
    del f, phi
    torch.cuda.empty_cache()

    # ---- precision: FP32 passes + FP64 residual refinement ----------------------
    # f = the analytic field plus 16 random Fourier modes (1 <= |k| <= 12,
    # +-k deduplicated, N(0,1) amplitudes, seeded); the FP64 solve of the same
    # f gives the Linf each refinement level is compared with.  The spectral
    # symbol (the reference's filter) is compared with the headline FP64 rate;
    # the finite-difference symbols with an FP64 solve of the same symbol.
    fp32 = []
    fd = []
    if not args.no_fp32:
        rfield = ff.field_set_modes(ff.field_default(n), random_modes(20261202))
        fp32 = precision_pareto(ff, torch, n, K, rfield, 0, value)
        fp32["field"] = "analytic f + 16 random modes, 1 <= |k| <= 12, N(0,1) amplitudes, seed 20261202"
        for name, flag in (("3-point", ff.FF_SYMBOL_3PT), ("5-point", ff.FF_SYMBOL_5PT)):
            r = precision_pareto(ff, torch, n, K, rfield, flag, None)
            r["symbol"] = name
            fd.append(r)

    extras = None
    if not args.no_extras:
        extras = {
            "sizes_fp64": bench_sizes(ff, torch, [(512, 512, 512), (768, 768, 768), (1000, 1000, 1000)], 5),
            "batching": bench_batching(ff, torch, 5),
            "serving": bench_serving(ff),
            "serving_same_shape": bench_serving_grouped(ff),
        }
        extras["serving_large_alone"] = bench_serving(ff, flags=ff.FF_SERVE_LARGE_ALONE,
                                                      rate=extras["serving"]["arrival_rate_per_s"])

    cpu = None
    if not args.no_cpu:
        log(f"CPU baseline (oracle) at {args.cpu_size}^3 ...")
        cpu = cpu_baseline(args.cpu_size)
        log(f"  {cpu['value']:.4e} points/s on {cpu['cores']} threads")

    out = {
        "metric": METRIC,
        "value": value,
        "unit": "points/s",
        "n_gpus": 1,
        "steps": K,
        "warmup": W,
        "ms_per_step": 1e3 * wall / K,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"FP64 periodic Poisson solve {n[0]}x{n[1]}x{n[2]}, 1x1x1 tasks, FishFry analytic f",
                   "n": list(n), "precision": "fp64", "input": "HBM-resident, preserved across steps"},
        "p50_ms": round(pct(lat, .5), 4),
        "p99_ms": round(pct(lat, .99), 4),
        "l1": l1, "l2": l2, "linf": linf,
        "solve_bytes": solve_bytes,
        "solve_hbm_frac": round(solve_bytes / (wall / K) / 1e9 / HBM_PEAK_GBS, 4),
        "roofline": roofline,
        "passes": passes,
        "fp32_refine": fp32,
        "fd_symbols": fd,
        "cpu_baseline": cpu,
        "extras": extras,
    }
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
