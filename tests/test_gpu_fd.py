"""GPU tests of the finite-difference symbols (FF_SYMBOL_3PT / FF_SYMBOL_5PT):
the FP64 solve against the CPU oracle's finite-difference solve, the
real-space stencil against the oracle's, FP32 refinement with the stencil
residual converging to the FP64 finite-difference solution, and the
benchmark's analytic norms (discretization error of order h^2 / h^4, which
the FP32 solve already reaches without or with one refinement step)."""
import math
import sys

import numpy as np
import pytest

import fishfry_amd as ff
import oracle

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")

TOL = 1e-13
FLAG = {3: ff.FF_SYMBOL_3PT, 5: ff.FF_SYMBOL_5PT}


def rel(a, b):
    return np.linalg.norm(np.ravel(a) - np.ravel(b)) / np.linalg.norm(np.ravel(b))


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def solve(f, order, lo=None, hi=None, precision=ff.FF_FP64, refine=0):
    n = f.shape
    if lo is None:
        lo, hi, _ = ff.benchmark_grid(n)
    plan = ff.ParisPencil(n, lo, hi, precision=precision, refine_steps=refine, flags=FLAG[order])
    nb = plan.bytes()
    a = torch.zeros(nb // 8, dtype=torch.float64, device="cuda")
    b = torch.zeros(nb // 8, dtype=torch.float64, device="cuda")
    a[: f.size].copy_(torch.from_numpy(np.ascontiguousarray(f).ravel()))
    plan.solve(nb, a, b)
    torch.cuda.synchronize()
    out = b[: f.size].cpu().numpy().reshape(n)
    plan.close()
    return out


@pytest.mark.parametrize("order", [3, 5])
@pytest.mark.parametrize("n", [(64, 48, 40), (30, 25, 27), (37, 16, 64), (1024, 8, 6), (96, 96, 96), (12, 10, 9),
                               (5, 6, 4)])
def test_fd_solve_matches_oracle(n, order):
    """FP64 solve with the finite-difference symbol against the oracle's, on
    the benchmark grid and on another domain (compiled-in 1024 x kernel,
    Bluestein lengths, odd nz included)."""
    rng = np.random.default_rng(81)
    f = rng.standard_normal(n)
    say("fd", n, order)
    got = solve(f, order)
    assert rel(got, oracle.solve(f, symbol=order)) < TOL
    lo, hi = [0.0, -1.0, 0.5], [2.0, 3.0, 4.5]
    got = solve(f, order, lo, hi)
    assert rel(got, oracle.solve(f, lo, hi, symbol=order)) < TOL


@pytest.mark.parametrize("order", [3, 5])
@pytest.mark.parametrize("n", [(1024, 8, 6), (1024, 12, 32)])
def test_fd_x_factors_from_lds_bitwise(n, order):
    """The compiled-in 1024 x kernels with LDS twiddles also read the x
    axis's finite-difference factors from an LDS copy: the solves are
    bit-identical to the generic x kernel's (which reads them from memory),
    FP64 and FP32 + 0 and + 1 steps, plain and pencil layouts."""
    rng = np.random.default_rng(82)
    f = rng.standard_normal(n)
    try:
        for prec, steps in ((ff.FF_FP64, 0), (ff.FF_FP32_REFINE, 0), (ff.FF_FP32_REFINE, 1)):
            for lay in (ff.FF_PLAIN_LAYOUT, ff.FF_PENCIL_LAYOUT):
                outs = []
                for generic in (False, True):
                    ff.tune_reset()
                    if generic:
                        for p in (0, 1):
                            cur = ff.tune_query(2, 1024, p, n[2] // 2 + 1)
                            ff.tune_set(2, 1024, p, cur.E, cur.T,
                                        ff.FF_TUNE_GENERIC | (cur.flags & ff.FF_TUNE_SPLIT_LDS))
                    else:
                        assert ff.tune_query(2, 1024, 0, n[2] // 2 + 1).flags & ff.FF_TUNE_LDS_TWIDDLES
                    lo, hi, _ = ff.benchmark_grid(n)
                    plan = ff.ParisPencil(n, lo, hi, precision=prec, refine_steps=steps, flags=FLAG[order] | lay)
                    nb = plan.bytes()
                    a = torch.zeros(nb // 8, dtype=torch.float64, device="cuda")
                    b = torch.zeros(nb // 8, dtype=torch.float64, device="cuda")
                    a[: f.size].copy_(torch.from_numpy(np.ascontiguousarray(f).ravel()))
                    say("fd lds", n, order, prec, steps, lay, generic)
                    plan.solve(nb, a, b)
                    torch.cuda.synchronize()
                    outs.append(b[: f.size].cpu().numpy().reshape(n))
                    plan.close()
                assert np.array_equal(outs[0], outs[1])
    finally:
        ff.tune_reset()
    assert rel(outs[0], oracle.solve(f, symbol=order)) < 1e-9


@pytest.mark.parametrize("order", [3, 5])
@pytest.mark.parametrize("n", [(16, 24, 20), (7, 5, 11), (2, 3, 300), (64, 64, 64)])
def test_stencil_apply_matches_oracle(n, order):
    """ff_stencil_apply (the real-space operator) against the oracle's stencil,
    and L_h(solve(f)) = f - mean(f) on the GPU."""
    rng = np.random.default_rng(82)
    lo, hi = [0.0, -1.0, 0.5], [2.0, 3.0, 4.5]
    plan = ff.ParisPencil(n, lo, hi, flags=FLAG[order])
    N = n[0] * n[1] * n[2]
    g = rng.standard_normal(n)
    src = torch.from_numpy(g.ravel()).cuda()
    out = torch.zeros(N, dtype=torch.float64, device="cuda")
    plan.stencil_apply(src, out)
    torch.cuda.synchronize()
    want = oracle.stencil(g, order, lo, hi)
    assert np.max(np.abs(out.cpu().numpy().reshape(n) - want)) < 1e-13 * np.max(np.abs(want))
    plan.close()
    f = rng.standard_normal(n)
    phi = solve(f, order, lo, hi)
    plan = ff.ParisPencil(n, lo, hi, flags=FLAG[order])
    plan.stencil_apply(torch.from_numpy(phi.ravel()).cuda(), out)
    torch.cuda.synchronize()
    assert np.max(np.abs(out.cpu().numpy().reshape(n) - (f - f.mean()))) < 1e-11 * np.max(np.abs(f))
    plan.close()


def test_stencil_apply_rejects_spectral_plans():
    n = (8, 8, 8)
    lo, hi, _ = ff.benchmark_grid(n)
    plan = ff.ParisPencil(n, lo, hi)
    buf = torch.zeros(512, dtype=torch.float64, device="cuda")
    out = torch.zeros(512, dtype=torch.float64, device="cuda")
    with pytest.raises(ff.FishFryError) as e:
        plan.stencil_apply(buf, out)
    assert e.value.status == ff.FF_ERR_INVALID
    plan.close()


@pytest.mark.parametrize("order", [3, 5])
@pytest.mark.parametrize("n", [(64, 48, 40), (96, 96, 96), (30, 25, 27)])
def test_fd_refinement_converges(n, order):
    """FP32 passes with the FP64 stencil residual and FP64 mean removal:
    each step gains FP32 precision, two steps reach the FP64
    finite-difference solution to round-off."""
    rng = np.random.default_rng(83)
    f = rng.standard_normal(n) + 3.0  # a mean that must be removed
    ref = solve(f, order)
    errs = [rel(solve(f, order, precision=ff.FF_FP32_REFINE, refine=s), ref) for s in range(3)]
    say("fd refine", n, order, errs)
    assert errs[0] < 1e-5
    assert errs[1] < 1e-9 and errs[1] < errs[0] / 100
    assert errs[2] < 1e-13


@pytest.mark.parametrize("order,steps", [(3, 0), (5, 1)])
def test_fd_fp32_reaches_fp64_linf(order, steps):
    """Against the benchmark's analytic potential the finite-difference solve
    carries a discretization error (h^2 for 3 points, h^4 for 5) far above
    FP32 round-off: the FP32 solve with `steps` refinement steps has an Linf
    within 1.1 times the FP64 solve's, on the benchmark field at 128^3."""
    n = (128, 128, 128)
    fd = ff.field_default(n)
    lo, hi, _ = ff.benchmark_grid(n)
    out = []
    for precision, s in ((ff.FF_FP64, 0), (ff.FF_FP32_REFINE, steps)):
        plan = ff.ParisPencil(n, lo, hi, precision=precision, refine_steps=s, flags=FLAG[order])
        nb = plan.bytes()
        a = torch.zeros(nb // 8, dtype=torch.float64, device="cuda")
        b = torch.zeros(nb // 8, dtype=torch.float64, device="cuda")
        ff.field_init(fd, a)
        plan.solve(nb, a, b)
        out.append(ff.field_norms(fd, b)[2])
        plan.close()
    say("fd linf", order, steps, out)
    # 3-point: ~h^2 (1e-3 at 128^3); 5-point: ~h^4 (1e-6)
    assert out[0] > (1e-4 if order == 3 else 1e-8)
    assert out[1] <= 1.1 * out[0]
