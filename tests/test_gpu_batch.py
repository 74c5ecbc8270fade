"""GPU tests of the batched solve (many uneven grids launched together on
parallel streams, optionally captured as one HIP graph): per-instance parity
with the single-grid solve and the CPU oracle, bit-identical results alone and
in a mix, graph replays, the autotuner, and the benchmark's batching workload
with phase-shifted and scaled fields (phi_b = s_b * phi(x + delta_b))."""
import json
import math
import sys

import numpy as np
import pytest

import fishfry_amd as ff
import oracle

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")

TOL = 1e-13


def say(*a):
    print(*a, flush=True)


def rel(a, b):
    return np.linalg.norm(np.ravel(a) - np.ravel(b)) / np.linalg.norm(np.ravel(b))


def alloc(batch):
    d = [torch.zeros(batch.bytes(i) // 8, dtype=torch.float64, device="cuda") for i in range(batch.count)]
    p = [torch.zeros(batch.bytes(i) // 8, dtype=torch.float64, device="cuda") for i in range(batch.count)]
    return d, p


def put(buf, arr):
    flat = np.ascontiguousarray(arr).ravel()
    buf[: flat.size].copy_(torch.from_numpy(flat))


def get(buf, n):
    return buf[: n[0] * n[1] * n[2]].cpu().numpy().reshape(n)


MIX = [(32, 32, 32), (48, 40, 36), (12, 10, 9), (64, 32, 16), (9, 8, 12), (128, 64, 96), (30, 25, 27),
       (16, 24, 20), (5, 6, 4), (20, 18, 15)]


def run_batch(shapes, fields, flags=0, precision=ff.FF_FP64, refine=0, stream=None):
    items = [{"n": n, "precision": precision, "refine_steps": refine} for n in shapes]
    b = ff.Batch(items, flags)
    d, p = alloc(b)
    for k, f in enumerate(fields):
        put(d[k], f)
    b.solve(d, p, stream)
    b.wait(stream)
    out = [get(p[k], shapes[k]) for k in range(len(shapes))]
    desc = b.describe()
    b.close()
    return out, desc


def test_batch_fp64_matches_oracle_and_single():
    rng = np.random.default_rng(1)
    fields = [rng.standard_normal(n) for n in MIX]
    out, desc = run_batch(MIX, fields)
    say("batch", desc)
    for n, f, phi in zip(MIX, fields, out):
        assert rel(phi, oracle.solve(f)) < TOL, n
        lo, hi, _ = ff.benchmark_grid(n)
        plan = ff.ParisPencil(n, lo, hi)
        nb = plan.bytes()
        a = torch.zeros(nb // 8, dtype=torch.float64, device="cuda")
        c = torch.zeros(nb // 8, dtype=torch.float64, device="cuda")
        put(a, f)
        plan.solve(nb, a, c)
        assert rel(phi, get(c, n)) < TOL, n
        plan.close()


def test_batch_alone_equals_in_mix_bitwise():
    rng = np.random.default_rng(2)
    fields = [rng.standard_normal(n) for n in MIX]
    mix, _ = run_batch(MIX, fields)
    for k in (0, 5, 6):
        alone, _ = run_batch([MIX[k]], [fields[k]])
        assert np.array_equal(alone[0], mix[k]), MIX[k]


def smooth_field(n, rng, modes=8):
    """The benchmark's f plus a few random low Fourier modes (the precision
    workload's kind of input)."""
    f = oracle.init_field(n)
    lo, hi, dx = oracle.grid(n)
    x = [lo[a] + np.arange(n[a]) * dx[a] for a in range(3)]
    X, Y, Z = np.meshgrid(*x, indexing="ij")
    for _ in range(modes):
        k = rng.integers(-4, 5, size=3)
        f = f + rng.standard_normal() * np.cos(k[0] * X + k[1] * Y + k[2] * Z)
    return f


def test_batch_fp32_refine():
    rng = np.random.default_rng(3)
    shapes = [(32, 32, 32), (48, 40, 36), (30, 25, 27)]
    fields = [smooth_field(n, rng) for n in shapes]
    for steps, bar in ((0, 1e-6), (2, 1e-14)):
        out, _ = run_batch(shapes, fields, precision=ff.FF_FP32_REFINE, refine=steps)
        for n, f, phi in zip(shapes, fields, out):
            e = rel(phi, oracle.solve(f))
            say("fp32 batch", n, steps, e)
            assert e < bar, (n, steps, e)
    # broadband input: refinement still converges monotonically, more slowly
    noisy = [rng.standard_normal(n) for n in shapes[:1]]
    errs = []
    for steps in (0, 1, 2, 3):
        out, _ = run_batch(shapes[:1], noisy, precision=ff.FF_FP32_REFINE, refine=steps)
        errs.append(rel(out[0], oracle.solve(noisy[0])))
    say("fp32 white noise", errs)
    assert errs[0] > errs[1] > errs[2] and errs[3] < 1e-10


def test_batch_mixed_precisions_in_one_launch():
    rng = np.random.default_rng(4)
    shapes = [(32, 32, 32), (24, 20, 18), (16, 16, 16), (40, 30, 20)]
    fields = [rng.standard_normal(n) for n in shapes]
    items = [{"n": n, "precision": (ff.FF_FP64 if k % 2 == 0 else ff.FF_FP32_REFINE), "refine_steps": 2}
             for k, n in enumerate(shapes)]
    b = ff.Batch(items)
    d, p = alloc(b)
    for k, f in enumerate(fields):
        put(d[k], f)
    b.solve(d, p)
    b.wait()
    for k, n in enumerate(shapes):
        assert rel(get(p[k], n), oracle.solve(fields[k])) < (TOL if k % 2 == 0 else 1e-14), n
    b.close()


def test_batch_graph_replays():
    rng = np.random.default_rng(5)
    fields = [rng.standard_normal(n) for n in MIX[:6]]
    eager, _ = run_batch(MIX[:6], fields)
    s = torch.cuda.Stream()
    items = [{"n": n} for n in MIX[:6]]
    b = ff.Batch(items, ff.FF_BATCH_GRAPH)
    d, p = alloc(b)
    for rep in range(4):
        with torch.cuda.stream(s):
            for k, f in enumerate(fields):
                put(d[k], f)  # FP64 clobbers density: reload it eagerly between replays
                p[k].zero_()
        s.synchronize()
        b.solve(d, p, s)
        b.wait(s)
        torch.cuda.synchronize()
        for k, n in enumerate(MIX[:6]):
            assert np.array_equal(get(p[k], n), eager[k]), (rep, n)
    say("graph", json.loads(b.describe())["graph"])
    b.close()


@pytest.mark.parametrize("flags", [0, ff.FF_BATCH_GROUPED])
def test_batch_autotune_keeps_results(flags):
    # grouped: the autotuner re-plans, which must rebuild the recorded groups
    rng = np.random.default_rng(6)
    shapes = MIX[:5] + [MIX[0], MIX[1]]
    fields = [rng.standard_normal(n) for n in shapes]
    items = [{"n": n} for n in shapes]
    b = ff.Batch(items, flags)
    d, p = alloc(b)
    try:
        tuned, before, after = b.autotune(d, p, reps=2)
        say("autotune", tuned, "passes", before, "->", after, "ms")
        assert tuned >= 5 and before > 0 and after > 0
        for k, f in enumerate(fields):
            put(d[k], f)
        b.solve(d, p)
        b.wait()
        for k, n in enumerate(shapes):
            assert rel(get(p[k], n), oracle.solve(fields[k])) < TOL
    finally:
        ff.tune_reset()
        b.close()


def test_batching_workload_analytic():
    """64 grids drawn from the batching configuration's size set, f_b(x) =
    s_b * f(x + delta_b), checked against phi_b = s_b * phi(x + delta_b)."""
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
    for precision, refine, bar in ((ff.FF_FP64, 0, 4e-15), (ff.FF_FP32_REFINE, 2, 4e-15)):
        b = ff.Batch([{"n": n, "precision": precision, "refine_steps": refine} for n in shapes])
        d, p = alloc(b)
        for k in range(64):
            ff.field_init(fields[k], d[k])
        b.solve(d, p)
        b.wait()
        worst = 0.0
        for k in range(64):
            worst = max(worst, ff.field_norms(fields[k], p[k])[2])
        say("batching workload", precision, refine, "worst Linf", worst, len(json.loads(b.describe())["plans"]))
        assert worst < bar
        b.close()
        del d, p
        torch.cuda.empty_cache()


GROUP_SHAPES = [(32, 32, 32), (48, 40, 36), (32, 32, 32), (12, 10, 9), (48, 40, 36), (32, 32, 32), (7, 5, 11),
                (64, 32, 16), (7, 5, 11), (12, 10, 9)]


def grouped_items(precision, refine):
    items = []
    for k, n in enumerate(GROUP_SHAPES):
        it = {"n": n, "precision": precision, "refine_steps": refine}
        if k % 3 == 1:  # some instances of a shape on another domain (other filter constants)
            it["lo"] = [0.0, -1.0, 0.5]
            it["hi"] = [2.0 + 0.25 * k, 3.0, 4.5]
        items.append(it)
    return items


def batch_outputs(items, fields, flags, stream=None):
    b = ff.Batch(items, flags)
    d, p = alloc(b)
    for k, f in enumerate(fields):
        put(d[k], f)
    b.solve(d, p, stream)
    b.wait(stream)
    out = [get(p[k], it["n"]) for k, it in enumerate(items)]
    desc = json.loads(b.describe())
    b.close()
    return out, desc


@pytest.mark.parametrize("precision,refine", [(ff.FF_FP64, 0), (ff.FF_FP32_REFINE, 2)])
def test_grouped_batch_bitwise(precision, refine):
    """FF_BATCH_GROUPED runs each pass of a group of same-shape instances as
    one launch with a per-instance table: the same kernels and arithmetic as
    the stream-parallel batch, so the results are bit-identical, with
    instances of one shape on different domains in the same launch."""
    rng = np.random.default_rng(21)
    items = grouped_items(precision, refine)
    fields = [rng.standard_normal(it["n"]) for it in items]
    say("grouped", precision, refine)
    ref, _ = batch_outputs(items, fields, 0)
    out, desc = batch_outputs(items, fields, ff.FF_BATCH_GROUPED)
    groups = desc["groups"]
    say("  groups", groups)
    assert sum(g["instances"] for g in groups) == len(items)
    assert len(groups) == len(set(GROUP_SHAPES))  # one group per shape: domains differ only in the table
    for k, it in enumerate(items):
        assert np.array_equal(out[k], ref[k]), (k, it["n"])
        if "lo" not in it:
            assert rel(out[k], oracle.solve(fields[k])) < (TOL if precision == ff.FF_FP64 else 1e-12)


def test_grouped_batch_graph_replays_and_rebinding():
    """Grouped launches captured in a graph replay bit-identically (eager
    reloads between replays), and new buffers rebuild the per-instance table."""
    rng = np.random.default_rng(22)
    items = grouped_items(ff.FF_FP64, 0)
    fields = [rng.standard_normal(it["n"]) for it in items]
    eager, _ = batch_outputs(items, fields, 0)
    s = torch.cuda.Stream()
    b = ff.Batch(items, ff.FF_BATCH_GROUPED | ff.FF_BATCH_GRAPH)
    for binding in range(2):
        d, p = alloc(b)  # fresh buffers: the table is rebuilt for them
        for rep in range(3):
            with torch.cuda.stream(s):
                for k, f in enumerate(fields):
                    put(d[k], f)  # FP64 clobbers density: reload it eagerly between replays
                    p[k].zero_()
            s.synchronize()
            b.solve(d, p, s)
            b.wait(s)
            torch.cuda.synchronize()
            for k, it in enumerate(items):
                assert np.array_equal(get(p[k], it["n"]), eager[k]), (binding, rep, it["n"])
        say("grouped graph", binding, json.loads(b.describe())["graph"])
        del d, p
    b.close()
