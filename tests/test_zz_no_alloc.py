"""GPU test of ff_solve_noalloc: the allocation-free variant of
ff_solve_async, whose work area is the library's static device array.  It is
called through the C-ABI and checked bitwise against ff_solve_async on the
same plan (the same kernels and arithmetic), on two calls in a row, the
second with a smaller grid, for the plans that use a work area (FP32 +
refinement, FP64 with FF_PRESERVE_INPUT)."""
import numpy as np
import pytest

import fishfry_amd as ff
import oracle

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")


def say(*a):
    print(*a, flush=True)


def run(n, precision, refine, flags, f, noalloc):
    lo, hi, _ = ff.benchmark_grid(n)
    plan = ff.ParisPencil(n, lo, hi, precision=precision, refine_steps=refine, flags=flags)
    nb = plan.bytes()
    d = torch.zeros(nb // 8, dtype=torch.float64, device="cuda")
    p = torch.full((nb // 8,), float("nan"), dtype=torch.float64, device="cuda")
    d[: f.size].copy_(torch.from_numpy(f.ravel()))
    torch.cuda.synchronize()
    if noalloc:
        plan.solve_noalloc(nb, d, p)
    else:
        plan.solve_async(nb, d, p)
    torch.cuda.synchronize()
    out = p[: f.size].cpu().numpy().reshape(n)
    kept = d[: f.size].cpu().numpy().reshape(n)
    plan.close()
    return out, kept


@pytest.mark.parametrize("precision,refine,flags", [(ff.FF_FP32_REFINE, 2, 0), (ff.FF_FP32_REFINE, 0, 0),
                                                    (ff.FF_FP64, 0, ff.FF_PRESERVE_INPUT)])
def test_noalloc_matches_solve_async(precision, refine, flags):
    rng = np.random.default_rng(41)
    for n in ((96, 80, 72), (48, 40, 36)):  # the second call is smaller
        say("noalloc", n, precision, refine, flags)
        f = rng.standard_normal(n)
        want, _ = run(n, precision, refine, flags, f, False)
        got, kept = run(n, precision, refine, flags, f, True)
        assert np.array_equal(got, want), n
        assert np.array_equal(kept, f)  # these plans preserve density
        if precision == ff.FF_FP64 or refine:
            rel = np.linalg.norm(got - oracle.solve(f)) / np.linalg.norm(oracle.solve(f))
            assert rel < 1e-12, rel


def test_noalloc_rejects_large_work_area():
    """A plan whose work area exceeds the static array is refused before any
    GPU work (256^3 FP32: about 337 MB > 64 MiB)."""
    n = (256, 256, 256)
    lo, hi, _ = ff.benchmark_grid(n)
    plan = ff.ParisPencil(n, lo, hi, precision=ff.FF_FP32_REFINE, refine_steps=1)
    nb = plan.bytes()
    d = torch.zeros(16, dtype=torch.float64, device="cuda")
    p = torch.zeros(16, dtype=torch.float64, device="cuda")
    with pytest.raises(ff.FishFryError) as e:
        plan.solve_noalloc(nb, d, p)  # rejected on size before the buffers are touched
    assert e.value.status == ff.FF_ERR_INVALID and "static work area" in str(e.value)
    plan.close()
