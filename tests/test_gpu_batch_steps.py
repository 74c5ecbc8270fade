"""GPU tests of the step-batched solve (FF_BATCH_STEPS): pass step s of every
instance of one precision and refinement count is one launch of a
multi-shape kernel driven by per-instance descriptors.  Results are compared
bitwise with the stream-parallel batch (each instance's single-solve
kernels), with the CPU oracle, and with the batching workload's analytic
potentials; the launch counts come from the batch's own description."""
import json
import math

import numpy as np
import pytest

import fishfry_amd as ff
import oracle

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")

TOL = 1e-13
# the batching configuration's size set
SIZES = [(32, 32, 32), (48, 48, 48), (64, 64, 64), (96, 96, 96), (128, 128, 128), (160, 160, 160), (128, 64, 96)]


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


def items_of(shapes, precision, refine, domains=True):
    items = []
    for k, n in enumerate(shapes):
        it = {"n": n, "precision": precision, "refine_steps": refine}
        if domains and k % 3 == 1:  # some instances on another domain (other filter constants)
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


def expect_bitwise(out, ref, items):
    bad = []
    for k, it in enumerate(items):
        if not np.array_equal(out[k], ref[k]):
            bad.append((k, it["n"], float(np.max(np.abs(out[k] - ref[k]))), rel(out[k], ref[k])))
    for b in bad:
        say("  differs:", b)
    assert not bad


@pytest.mark.parametrize("precision,refine", [(ff.FF_FP64, 0), (ff.FF_FP32_REFINE, 2)])
def test_step_batch_bitwise(precision, refine):
    """Every shape of the batching configuration, several on other domains:
    one launch per pass step (5 for FP64, 5 + 10 per refinement step for
    FP32), results bit-identical to the stream-parallel batch."""
    rng = np.random.default_rng(31)
    shapes = SIZES + SIZES[::-1]
    items = items_of(shapes, precision, refine)
    fields = [rng.standard_normal(it["n"]) for it in items]
    say("steps", precision, refine)
    ref, _ = batch_outputs(items, fields, 0)
    out, desc = batch_outputs(items, fields, ff.FF_BATCH_STEPS)
    fams = desc["steps"]
    say("  families", fams)
    steps = 5 + 10 * refine if precision == ff.FF_FP32_REFINE else 5
    assert len(fams) == 1 and fams[0]["instances"] == len(items)
    assert fams[0]["steps"] == steps and fams[0]["launches"] == steps and fams[0]["multi_launches"] == steps
    expect_bitwise(out, ref, items)
    for k in (0, 3, 6):  # oracle spot checks (the oracle is slow at 160^3)
        want = oracle.solve(fields[k], items[k].get("lo"), items[k].get("hi"))
        assert rel(out[k], want) < (TOL if precision == ff.FF_FP64 else 1e-12)


def test_step_batch_with_other_lengths():
    """Shapes without a multi-shape variant (Bluestein lengths, odd nz, short
    lines) run their steps as grouped launches beside the multi-shape one;
    every result is still bit-identical to the stream-parallel batch and
    matches the oracle."""
    rng = np.random.default_rng(32)
    shapes = [(32, 32, 32), (7, 5, 11), (48, 40, 36), (12, 10, 9), (64, 32, 16), (32, 32, 32), (37, 48, 64),
              (30, 25, 27)]
    items = items_of(shapes, ff.FF_FP64, 0)
    fields = [rng.standard_normal(it["n"]) for it in items]
    ref, _ = batch_outputs(items, fields, 0)
    out, desc = batch_outputs(items, fields, ff.FF_BATCH_STEPS)
    fams = desc["steps"]
    say("  families", fams)
    assert len(fams) == 1 and fams[0]["launches"] > fams[0]["steps"]
    assert fams[0]["multi_launches"] == 5  # every step has some instance with a variant
    expect_bitwise(out, ref, items)
    for k, it in enumerate(items):
        assert rel(out[k], oracle.solve(fields[k], it.get("lo"), it.get("hi"))) < TOL, it["n"]


def test_step_batch_mixed_precisions():
    """FP64 and FP32 + 2 instances in one batch form two families, each with
    its own launch sequence; each result equals its stream-parallel one."""
    rng = np.random.default_rng(33)
    shapes = [(32, 32, 32), (64, 64, 64), (128, 64, 96), (48, 48, 48)]
    items = []
    for k, n in enumerate(shapes * 2):
        items.append({"n": n, "precision": ff.FF_FP64 if k % 2 == 0 else ff.FF_FP32_REFINE,
                      "refine_steps": 0 if k % 2 == 0 else 2})
    fields = [rng.standard_normal(it["n"]) for it in items]
    ref, _ = batch_outputs(items, fields, 0)
    out, desc = batch_outputs(items, fields, ff.FF_BATCH_STEPS)
    fams = desc["steps"]
    say("  families", fams)
    assert sorted(f["launches"] for f in fams) == [5, 25]
    expect_bitwise(out, ref, items)


def test_step_batch_graph_replays_and_rebinding():
    """Step launches captured in a graph replay bit-identically (eager reloads
    between replays), and new buffers rebuild the descriptor tables."""
    rng = np.random.default_rng(34)
    items = items_of(SIZES, ff.FF_FP64, 0)
    fields = [rng.standard_normal(it["n"]) for it in items]
    eager, _ = batch_outputs(items, fields, 0)
    s = torch.cuda.Stream()
    b = ff.Batch(items, ff.FF_BATCH_STEPS | ff.FF_BATCH_GRAPH)
    for binding in range(2):
        d, p = alloc(b)
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
        say("steps graph", binding, json.loads(b.describe())["graph"])
        del d, p
    b.close()


def test_step_batching_workload_analytic():
    """The batching workload (64 grids drawn from the size set, f_b(x) =
    s_b * f(x + delta_b)) in 5 launches per FP64 batch and 25 per FP32 + 2
    batch, checked against phi_b = s_b * phi(x + delta_b)."""
    rng = np.random.default_rng(20261202)
    shapes = [SIZES[i] for i in rng.integers(0, len(SIZES), 64)]
    fields = []
    for n in shapes:
        fd = ff.field_default(n)
        for a in range(3):
            fd.shift[a] = float(rng.uniform(0, 2 * math.pi))
        fd.scale = float(rng.uniform(0.5, 2.0))
        fields.append(fd)
    for precision, refine, launches in ((ff.FF_FP64, 0, 5), (ff.FF_FP32_REFINE, 2, 25)):
        b = ff.Batch([{"n": n, "precision": precision, "refine_steps": refine} for n in shapes], ff.FF_BATCH_STEPS)
        d, p = alloc(b)
        for k in range(64):
            ff.field_init(fields[k], d[k])
        b.solve(d, p)
        b.wait()
        worst = max(ff.field_norms(fields[k], p[k])[2] for k in range(64))
        fams = json.loads(b.describe())["steps"]
        say("step batching workload", precision, refine, "worst Linf", worst, fams)
        assert worst < 4e-15
        assert len(fams) == 1 and fams[0]["launches"] == launches
        b.close()
        del d, p
        torch.cuda.empty_cache()
