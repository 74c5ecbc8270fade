"""GPU parity of the HIP solve against the CPU oracle (oracle/), the committed
golden vectors (tests/golden/) and the benchmark's analytic potential.

Bars: FP64 passes and solves agree with the oracle to a relative L2 of 1e-13
(the acceptance bar for parity with the reference's FP64 solve is 1e-12); the
analytic-phi norms stay at FP64 round-off.
"""
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


def rel(a, b):
    d = np.linalg.norm(np.ravel(b))
    return np.linalg.norm(np.ravel(a) - np.ravel(b)) / (d if d else 1.0)


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def make(n, **kw):
    lo, hi, _ = ff.benchmark_grid(n)
    plan = ff.ParisPencil(n, lo, hi, **kw)
    nb = plan.bytes()
    a = torch.zeros(nb // 8, dtype=torch.float64, device="cuda")
    b = torch.zeros(nb // 8, dtype=torch.float64, device="cuda")
    return plan, nb, a, b


def put(buf, arr):
    flat = np.ascontiguousarray(arr).view(np.float64).ravel()
    buf[: flat.size].copy_(torch.from_numpy(flat))


def get_real(buf, n):
    N = n[0] * n[1] * n[2]
    return buf[:N].cpu().numpy().reshape(n)


def kt_of(plan):
    return json.loads(plan.describe())["kt"]


PLAIN, YB, XB = 0, 1, 2


def to_layout(spec, lay, KT):
    """Flat complex array of spec[i][j][kh] in a spectrum layout: plain
    [i][j][kh], y-blocked [kb][i][j][kt] or x-blocked [kb][j][i][kt]."""
    if lay == PLAIN:
        return np.ascontiguousarray(spec).ravel()
    parts = []
    for k0 in range(0, spec.shape[2], KT):
        blk = spec[:, :, k0:k0 + KT]
        parts.append(np.ascontiguousarray(blk if lay == YB else blk.transpose(1, 0, 2)).ravel())
    return np.concatenate(parts)


def from_layout(flat, n, lay, KT):
    nh = n[2] // 2 + 1
    if lay == PLAIN:
        return flat.reshape(n[0], n[1], nh)
    out = np.empty((n[0], n[1], nh), dtype=flat.dtype)
    off = 0
    for k0 in range(0, nh, KT):
        W = min(KT, nh - k0)
        size = n[0] * n[1] * W
        blk = flat[off:off + size]
        out[:, :, k0:k0 + W] = blk.reshape(n[0], n[1], W) if lay == YB else blk.reshape(n[1], n[0], W).transpose(1, 0, 2)
        off += size
    return out


def get_half(buf, n, lay=PLAIN, KT=0):
    M = n[0] * n[1] * (n[2] // 2 + 1)
    return from_layout(buf[: 2 * M].cpu().numpy().view(np.complex128), n, lay, KT)


def put_half(buf, spec, lay=PLAIN, KT=0):
    put(buf, to_layout(spec, lay, KT))


def gpu_solve(f, **kw):
    kw.setdefault("flags", 0)
    n = f.shape
    plan, nb, a, b = make(n, **kw)
    put(a, f)
    plan.solve(nb, a, b)
    torch.cuda.synchronize()
    out = get_real(b, n)
    plan.close()
    return out


PASS_SHAPES = [(16, 24, 20), (32, 32, 32), (12, 10, 9), (64, 48, 40), (5, 6, 4), (9, 8, 12), (2, 8, 16),
               (30, 25, 27), (32, 1024, 512), (24, 600, 1000), (1024, 8, 6), (512, 512, 10), (8, 6, 1024),
               # prime factors above 5 (Bluestein lines): odd and even nz, nz/2 prime
               (7, 11, 13), (14, 22, 26), (37, 20, 18), (12, 10, 14), (13, 509, 8), (4, 6, 1018)]


@pytest.mark.parametrize("flags", [0, ff.FF_BLOCKED_LAYOUT])
@pytest.mark.parametrize("n", PASS_SHAPES)
def test_each_pass_matches_oracle(n, flags):
    rng = np.random.default_rng(sum(n))
    plan, nb, a, b = make(n, flags=flags)
    KT = kt_of(plan)
    assert KT == (8 if flags else 0)
    # layouts read / written by passes 1..4 (plain throughout unless blocked)
    lay = {1: (PLAIN, YB), 2: (YB, XB), 3: (XB, YB), 4: (YB, PLAIN)} if KT else \
        {p: (PLAIN, PLAIN) for p in (1, 2, 3, 4)}
    say("pass parity", n, "kt", KT)
    nh = n[2] // 2 + 1
    lo, hi, _ = oracle.grid(n)
    # pass 0: z forward (D2Z)
    f = rng.standard_normal(n)
    put(a, f)
    plan.solve_pass(0, a, b)
    torch.cuda.synchronize()
    assert rel(get_half(b, n), oracle.rfft(f)) < TOL
    spec = rng.standard_normal((n[0], n[1], nh)) + 1j * rng.standard_normal((n[0], n[1], nh))
    # pass 1: y forward
    put_half(a, spec, lay[1][0], KT)
    plan.solve_pass(1, a, b)
    torch.cuda.synchronize()
    assert rel(get_half(b, n, lay[1][1], KT), oracle.fft(spec, -1, axis=1)) < TOL
    # pass 2: x forward * filter * inverse
    put_half(a, spec, lay[2][0], KT)
    plan.solve_pass(2, a, b)
    torch.cuda.synchronize()
    want = oracle.fft(oracle.filter_half(oracle.fft(spec, -1, axis=0), n[2], lo, hi), +1, axis=0)
    assert rel(get_half(b, n, lay[2][1], KT), want) < TOL
    # pass 3: y inverse
    put_half(a, spec, lay[3][0], KT)
    plan.solve_pass(3, a, b)
    torch.cuda.synchronize()
    assert rel(get_half(b, n, lay[3][1], KT), oracle.fft(spec, +1, axis=1)) < TOL
    # pass 4: z inverse (Z2D) and the 1/N scale; Im of DC / Nyquist ignored
    put_half(a, spec, lay[4][0], KT)
    plan.solve_pass(4, a, b)
    torch.cuda.synchronize()
    want = oracle.irfft(spec, n[2]) * (1.0 / (float(n[0]) * float(n[1]) * float(n[2])))
    assert rel(get_real(b, n), want) < TOL
    plan.close()


@pytest.mark.parametrize("n", [(1024, 4, 8), (1024, 8, 10), (1024, 12, 2), (1024, 6, 7), (64, 48, 40)])
def test_x_pass_tile_shift_bitwise(n):
    """The x pass with its line plan compiled in (length 1024) shifts each
    plane's column tiles onto the 64-byte grid of the destination rows when
    planes start off that grid (ny*nh a multiple of 4, nh not).  Writing one
    spectrum to destinations 0..3 complex elements apart moves the tile
    boundaries, never the values.  (1024, 6, 7): rows off the grid, no
    shift; (64, 48, 40): generic kernel, no shift."""
    rng = np.random.default_rng(11)
    plan, nb, a, _ = make(n)
    nh = n[2] // 2 + 1
    spec = rng.standard_normal((n[0], n[1], nh)) + 1j * rng.standard_normal((n[0], n[1], nh))
    put_half(a, spec)
    lo, hi, _ = oracle.grid(n)
    want = oracle.fft(oracle.filter_half(oracle.fft(spec, -1, axis=0), n[2], lo, hi), +1, axis=0)
    big = torch.zeros(nb // 8 + 8, dtype=torch.float64, device="cuda")
    outs = []
    for k in range(4):
        say("x pass tile shift", n, "dst offset", k)
        big.zero_()
        dst = big[2 * k:]
        plan.solve_pass(2, a, dst)
        torch.cuda.synchronize()
        outs.append(get_half(dst, n))
    plan.close()
    assert all(np.array_equal(outs[0], o) for o in outs[1:])
    assert rel(outs[0], want) < TOL


@pytest.mark.parametrize("n", [(512, 512, 512), (2048, 1024, 256), (512, 256, 514)])
def test_pencil_default_matches_plain(n):
    """FP64 plans of 2^25 spectrum points and more take the pencil layout by
    default; FF_PLAIN_LAYOUT keeps the reference's [i][j][kh].  The two give
    bit-identical potentials, and the default one meets the analytic norms."""
    field = ff.field_default(n)
    lo, hi, _ = ff.benchmark_grid(n)
    outs = []
    for flags in (0, ff.FF_PLAIN_LAYOUT):
        plan = ff.ParisPencil(n, lo, hi, flags=flags)
        info = json.loads(plan.describe())
        say("pencil default", n, flags, "pencil", info["pencil"])
        assert info["pencil"] == (0 if flags else 1)
        nb = plan.bytes()
        a = torch.empty(nb // 8, dtype=torch.float64, device="cuda")
        b = torch.empty(nb // 8, dtype=torch.float64, device="cuda")
        ff.field_init(field, a)
        plan.solve(nb, a, b)
        torch.cuda.synchronize()
        N = n[0] * n[1] * n[2]
        outs.append(b[:N].clone())
        if not flags:
            assert ff.field_norms(field, b)[2] < 4e-15
        plan.close()
        del a, b
    assert torch.equal(outs[0], outs[1])
    del outs
    torch.cuda.empty_cache()


@pytest.mark.parametrize("n", [(48, 40, 36), (30, 25, 27), (9, 7, 12), (1024, 24, 64), (36, 1000, 20)])
@pytest.mark.parametrize("flags", [0, ff.FF_SYMBOL_3PT, ff.FF_SYMBOL_5PT])
def test_pencil_layout_fp32_and_symbols_bitwise(n, flags):
    """FP32 refinement plans run their FP32 passes in the pencil layout
    between the two halves of the work area, and the finite-difference
    symbols index the pencil x pass by (i, j, kh) as the plain one does: the
    potentials are bit-identical to the plain layout's (FP64, FP32 + 0 and +
    2 steps), and match the oracle."""
    rng = np.random.default_rng(7)
    f = rng.standard_normal(n)
    f -= f.mean()
    for prec, steps in ((ff.FF_FP64, 0), (ff.FF_FP32_REFINE, 0), (ff.FF_FP32_REFINE, 2)):
        outs = []
        for lay in (ff.FF_PLAIN_LAYOUT, ff.FF_PENCIL_LAYOUT):
            say("pencil", n, flags, prec, steps, lay)
            outs.append(gpu_solve(f, precision=prec, refine_steps=steps, flags=flags | lay))
        assert np.array_equal(outs[0], outs[1])
        want = oracle.solve(f, symbol={0: 0, ff.FF_SYMBOL_3PT: 3, ff.FF_SYMBOL_5PT: 5}[flags])
        # FP32 + 2 steps: the residual floor (1.03e-13 for the 5-point symbol
        # at 1024x24x64); FP32 without refinement: FP32 rounding
        tol = TOL if prec == ff.FF_FP64 else (1e-12 if steps else 1e-5)
        assert rel(outs[1], want) < tol


@pytest.mark.parametrize("n", [(32, 1024, 512), (48, 40, 36), (30, 25, 27), (64, 64, 2), (96, 160, 48),
                               (9, 7, 12), (5, 3, 9), (1024, 24, 64), (24, 1024, 30), (36, 1000, 20)])
def test_layouts_agree_bitwise(n):
    """The blocked and pencil spectrum layouts change addresses only, not
    arithmetic.  The pencil shapes include ny below the z inverse's rows per
    workgroup (rows of one workgroup in several x-planes), odd nz, Bluestein
    and 1024-point columns."""
    rng = np.random.default_rng(4)
    f = rng.standard_normal(n)
    outs = []
    P = ff.FF_PENCIL_LAYOUT
    for flags in (0, ff.FF_BLOCKED_LAYOUT, ff.FF_BLOCKED_LAYOUT | ff.FF_PRESERVE_INPUT, P, P | ff.FF_PRESERVE_INPUT):
        plan, nb, a, b = make(n, flags=flags)
        say("layout", n, flags, "kt", kt_of(plan))
        put(a, f)
        plan.solve(nb, a, b)
        torch.cuda.synchronize()
        outs.append(get_real(b, n))
        plan.close()
    for k in range(1, len(outs)):
        assert np.array_equal(outs[0], outs[k]), k
    assert rel(outs[0], oracle.solve(f)) < TOL


@pytest.mark.parametrize("variant", [ff.FF_TUNE_SPLIT_LDS, ff.FF_TUNE_TWO_GROUPS,
                                     ff.FF_TUNE_SPLIT_LDS | ff.FF_TUNE_TWO_GROUPS, ff.FF_TUNE_GENERIC,
                                     ff.FF_TUNE_LDS_TWIDDLES, ff.FF_TUNE_SPLIT_LDS | ff.FF_TUNE_LDS_TWIDDLES])
@pytest.mark.parametrize("n,E", [((64, 48, 40), 0), ((30, 25, 27), 0), ((96, 1000, 64), 10), ((1024, 16, 8), 16),
                                 ((768, 12, 6), 12), ((128, 64, 96), 8), ((256, 1024, 20), 16),
                                 ((512, 512, 12), 0), ((1024, 1024, 4), 0), ((64, 1000, 10), 20),
                                 ((1000, 16, 10), 20), ((768, 8, 6), 12), ((128, 16, 64), 0), ((64, 32, 64), 0),
                                 ((96, 16, 128), 0), ((48, 8, 64), 0), ((160, 8, 64), 0), ((32, 8, 40), 0)])
def test_column_variants_agree(n, E, variant):
    """Column-pass variants -- the split real/imaginary LDS exchange, two
    column groups per workgroup, the generic kernel in place of the one
    with the line plan compiled in, and twiddles read from a (half-table,
    swizzled) LDS copy instead of memory (length 1024) -- move the same
    values as the kernel with flags 0.  The pass kernels are compiled without
    floating-point contraction and write their multiply-adds out, so every
    variant performs the same rounded operations: results are bit-identical,
    and all match the oracle.  Two groups apply where E is 8 or 16.
    Compiled-in plans exist for 1024 (E 16), 32, 48, 64, 96, 128, 160 and 256,
    and for the x passes of 512, 768 and 1000."""
    rng = np.random.default_rng(5)
    f = rng.standard_normal(n)
    try:
        for prec, steps in ((ff.FF_FP64, 0), (ff.FF_FP32_REFINE, 1)):
            outs = []
            pi = 0 if prec == ff.FF_FP64 else 1
            for flags in (0, variant):
                ff.tune_reset()
                for kind, length in ((1, n[1]), (2, n[0])):
                    for p in sorted({0, pi}):  # FP32 refinement also runs FP64 residual passes
                        cur = ff.tune_query(kind, length, p, n[2] // 2 + 1)
                        # drop what this E / length cannot do
                        for fl in (flags, flags & ff.FF_TUNE_SPLIT_LDS, 0):
                            try:
                                ff.tune_set(kind, length, p, E or cur.E, 0 if E else cur.T, fl)
                                break
                            except ff.FishFryError:  # E does not divide this length
                                try:
                                    ff.tune_set(kind, length, p, cur.E, cur.T, fl)
                                    break
                                except ff.FishFryError:
                                    continue
                say("variant", n, E, prec, flags)
                outs.append(gpu_solve(f, precision=prec, refine_steps=steps))
            d = rel(outs[1], outs[0])
            say("  variant vs default kernel: rel L2", d)
            assert np.array_equal(outs[1], outs[0])
            want = oracle.solve(f)
            tol = TOL if prec == ff.FF_FP64 else 1e-9  # FP32 + 1 step: ~1e-10 at 768-long lines
            assert rel(outs[0], want) < tol and rel(outs[1], want) < tol
    finally:
        ff.tune_reset()


@pytest.mark.parametrize("layout", [ff.FF_PLAIN_LAYOUT, ff.FF_PENCIL_LAYOUT])
@pytest.mark.parametrize("n,E,T", [((6, 5, 1024), 8, 2), ((5, 7, 1024), 8, 8), ((6, 5, 1024), 8, 1),
                                   ((8, 6, 512), 16, 0), ((5, 4, 768), 12, 4), ((5, 4, 768), 24, 0),
                                   ((6, 6, 1000), 10, 0), ((7, 9, 160), 20, 0), ((12, 10, 32), 16, 0),
                                   ((10, 9, 48), 12, 0), ((9, 11, 128), 8, 3)])
def test_row_variants_agree(n, E, T, layout):
    """The z passes (both directions) with twiddles read from a half-table LDS
    copy, and with the generic kernel, move the same values as the kernel
    with the row plan compiled in and twiddles from memory: bit-identical
    solves (FP64, and FP32 + one refinement step, whose residual passes run
    FP64 rows too) that match the oracle, in the plain and the pencil spectrum
    layout (the pencil z inverse stages its rows from a strided spectrum).
    T = 0 keeps the default rows per workgroup."""
    rng = np.random.default_rng(7)
    f = rng.standard_normal(n)
    M = n[2] // 2
    try:
        for prec, steps in ((ff.FF_FP64, 0), (ff.FF_FP32_REFINE, 1)):
            pi = 0 if prec == ff.FF_FP64 else 1
            outs = []
            for flags in (0, ff.FF_TUNE_LDS_TWIDDLES, ff.FF_TUNE_GENERIC):
                ff.tune_reset()
                for kind in (0, 3):
                    for p in sorted({0, pi}):
                        cur = ff.tune_query(kind, M, p, 1)
                        ff.tune_set(kind, M, p, E, T or cur.T, flags)
                        got = ff.tune_query(kind, M, p, 1)
                        assert (got.E, got.flags) == (E, flags), (kind, p, got.E, got.flags)
                say("row variant", n, E, T, layout, prec, flags)
                outs.append(gpu_solve(f, precision=prec, refine_steps=steps, flags=layout))
            assert np.array_equal(outs[1], outs[0]) and np.array_equal(outs[2], outs[0])
            want = oracle.solve(f)
            assert rel(outs[0], want) < (TOL if prec == ff.FF_FP64 else 1e-9)
    finally:
        ff.tune_reset()


@pytest.mark.parametrize("n,slab", [((10, 12, 16), 3), ((64, 48, 40), 5), ((30, 25, 27), 7), ((9, 8, 12), 1),
                                    ((96, 160, 48), 0), ((12, 10, 9), 12)])
def test_slab_pipeline_bitwise(n, slab):
    """FF_SLAB_PIPELINE runs the same kernels over plane ranges: the solve is
    bit-identical to the whole-pass one, with and without FF_PRESERVE_INPUT
    (without it the slab order keeps the aliased work array from overwriting
    real planes not yet read)."""
    rng = np.random.default_rng(6)
    f = rng.standard_normal(n)
    outs = []
    for flags in (0, ff.FF_SLAB_PIPELINE, ff.FF_SLAB_PIPELINE | ff.FF_PRESERVE_INPUT):
        lo, hi, _ = ff.benchmark_grid(n)
        plan = ff.ParisPencil(n, lo, hi, flags=flags, slab_planes=slab)
        info = json.loads(plan.describe())
        say("slab", n, slab, flags, info["slab"])
        if flags and 0 < slab < n[0]:
            assert info["slab"] == slab
        nb = plan.bytes()
        a = torch.zeros(nb // 8, dtype=torch.float64, device="cuda")
        b = torch.zeros(nb // 8, dtype=torch.float64, device="cuda")
        put(a, f)
        stamps = []
        plan.solve(nb, a, b, stamps)
        torch.cuda.synchronize()
        assert len(stamps) == 14
        outs.append(get_real(b, n))
        if flags & ff.FF_PRESERVE_INPUT:
            assert np.array_equal(get_real(a, n), f)
        plan.close()
    assert np.array_equal(outs[0], outs[1]) and np.array_equal(outs[0], outs[2])
    assert rel(outs[0], oracle.solve(f)) < TOL


SOLVE_SHAPES = [(8, 8, 8), (12, 10, 9), (16, 24, 20), (32, 32, 32), (64, 64, 64), (6, 9, 15), (5, 6, 4),
                (9, 8, 12), (48, 40, 36), (128, 64, 96), (2, 2, 2), (2, 3, 4), (2, 8, 16), (16, 2, 9),
                (30, 25, 27), (100, 96, 80), (64, 64, 2), (4, 4, 2), (3, 3, 3), (250, 2, 6), (96, 160, 48),
                (48, 600, 1000), (40, 1024, 300),
                # prime factors above 5 (Bluestein lines)
                (7, 5, 11), (11, 13, 17), (101, 8, 9), (37, 41, 43), (509, 4, 6), (14, 49, 22), (2, 7, 7),
                (97, 96, 194)]


@pytest.mark.parametrize("n", SOLVE_SHAPES)
def test_solve_matches_oracle(n):
    say("solve parity", n)
    rng = np.random.default_rng(n[0] * 10007 + n[1] * 101 + n[2])
    f = rng.standard_normal(n)
    got = gpu_solve(f)
    want = oracle.solve(f)
    assert rel(got, want) < TOL, rel(got, want)


def test_solve_matches_golden(golden):
    """Every committed golden solve, including extents with prime factors
    above 5 (7x5x11), against the GPU solve."""
    tags = sorted({k[len("solve_rand_in_"):] for k in golden if k.startswith("solve_rand_in_")})
    checked = 0
    for kind in ("rand", "field"):
        for tag in tags:
            n = tuple(int(v) for v in tag.split("x"))
            say("golden", kind, tag)
            got = gpu_solve(golden[f"solve_{kind}_in_{tag}"])
            assert rel(got, golden[f"solve_{kind}_out_{tag}"]) < TOL, (kind, tag)
            checked += 1
    assert checked == 2 * len(tags) and len(tags) >= 8


@pytest.mark.parametrize("n", [(32, 32, 32), (64, 64, 64), (48, 40, 36), (128, 96, 64)])
def test_benchmark_field_and_norms(n):
    say("analytic", n)
    plan, nb, a, b = make(n)
    field = ff.field_default(n)
    ff.field_init(field, a)
    torch.cuda.synchronize()
    f = get_real(a, n)
    assert rel(f, oracle.init_field(n)) < 1e-14
    plan.solve(nb, a, b)
    torch.cuda.synchronize()
    phi = get_real(b, n)
    assert rel(phi, oracle.solve(f)) < TOL
    l1, l2, linf = ff.field_norms(field, b)
    assert linf < 2e-15 and l2 < 2e-15 and l1 < 2e-15
    # the device error sums agree with the oracle's on a phi whose error is far
    # above the 1-ulp differences between the two libms' exact potentials
    rng = np.random.default_rng(3)
    pert = phi + 1e-6 * rng.standard_normal(n)
    put(b, pert)
    got = np.array(ff.field_diff(field, b))
    want = oracle.diff(pert)
    assert np.all(np.abs(got - want) <= 1e-9 * np.abs(want)), (got, want)
    # phi = None evaluates the exact potential against zero (FishFry.hpp:146-159)
    s0 = np.array(ff.field_diff(field, None))
    w0 = oracle.diff(np.zeros(n))
    assert np.all(np.abs(s0 - w0) <= 1e-13 * np.abs(w0))
    plan.close()


def test_preserve_input_flag():
    n = (48, 40, 36)
    rng = np.random.default_rng(9)
    f = rng.standard_normal(n)
    plan, nb, a, b = make(n, flags=ff.FF_PRESERVE_INPUT)
    put(a, f)
    plan.solve(nb, a, b)
    torch.cuda.synchronize()
    assert np.array_equal(get_real(a, n), f)
    first = get_real(b, n)
    assert rel(first, oracle.solve(f)) < TOL
    plan.solve(nb, a, b)
    torch.cuda.synchronize()
    assert np.array_equal(get_real(b, n), first)  # deterministic
    plan.close()


def test_stamps_and_contract():
    n = (32, 24, 20)
    plan, nb, a, b = make(n)
    ff.field_init(ff.field_default(n), a)
    stamps = []
    plan.solve(nb, a, b, stamps)
    names = [s[0] for s in stamps]
    assert names == ["Total", "Reorder 3D block into sub-pencils", "Redistribute into Z pencils",
                     "Real-to-complex FFT in Z", "Redistribute for Y pencils", "Forward FFT in Y",
                     "Redistribute for X pencils", "X FFTs and filter", "Redistribute for Y pencils",
                     "Backward FFT in Y", "Redistribute in Z pencils", "Complex-to-real FFT in Z",
                     "Redistribute for 3D blocks", "Rearrange into 3D blocks and normalize"]
    t = [s[1] for s in stamps]
    assert all(t[i] <= t[i + 1] for i in range(len(t) - 1))
    assert t[-1] > t[0]
    # bytes check (ParisPencil.cpp:101)
    with pytest.raises(ff.FishFryError) as e:
        plan.solve(nb - 8, a, b)
    assert e.value.status == ff.FF_ERR_BYTES
    with pytest.raises(ff.FishFryError) as e:
        plan.solve(nb, a, a)
    assert e.value.status == ff.FF_ERR_INVALID
    plan.close()
    # a prime x extent above the Bluestein limit: no whole-line plan and no
    # four-step split
    with pytest.raises(ff.FishFryError) as e:
        ff.ParisPencil((5003, 4, 4), *ff.benchmark_grid((5003, 4, 4))[:2])
    assert e.value.status == ff.FF_ERR_UNSUPPORTED
    assert "5003" in str(e.value)


# x or y extents no workgroup transforms whole: four-step lines n = n1*n2
# (4100 = 41*100: a Bluestein sub-line; 4097 = 17*241: two Bluestein
# sub-lines; 8448 = 2^8*3*11; 70000 = 2^4*5^4*7, above the former 65535 bound;
# 4400 x 4200: both axes; 9000: 5-smooth, but 300 threads of 30 elements
# exceed a workgroup's register budget)
FOUR_STEP_SHAPES = [(8, 4100, 6), (4100, 6, 8), (4097, 5, 4), (6, 8448, 5), (70000, 3, 4), (4400, 4200, 2),
                    (6, 9000, 5)]


@pytest.mark.parametrize("n", FOUR_STEP_SHAPES)
def test_four_step_lines_match_oracle(n):
    """The reference plans any line length (cufftPlanMany,
    ParisPencil.cpp:78-81).  Extents too long for one workgroup's line run
    as four-step sub-line passes; the x filter reads y frequencies in
    four-step order.  FP64 solves match the oracle; FP32 + 2 steps converge
    to it; the 3-point symbol indexes its tables by the true frequencies."""
    rng = np.random.default_rng(n[0] + 7 * n[1] + 13 * n[2])
    f = rng.standard_normal(n)
    f -= f.mean()
    lo, hi, _ = ff.benchmark_grid(n)
    plan = ff.ParisPencil(n, lo, hi)
    info = json.loads(plan.describe())
    say("four-step", n, info["four_step"])
    fs = info["four_step"]
    if n[1] > 8192 or n[1] in (4100, 4200):
        assert fs["y"][0] * fs["y"][1] == n[1] and fs["y"][0] > 1
    if n[0] > 4096:
        assert fs["x"][0] * fs["x"][1] == n[0] and fs["x"][0] > 1
    plan.close()
    want = oracle.solve(f)
    got = gpu_solve(f)
    assert rel(got, want) < TOL, rel(got, want)
    got32 = gpu_solve(f, precision=ff.FF_FP32_REFINE, refine_steps=2)
    assert rel(got32, want) < 1e-12, rel(got32, want)
    if n[0] * n[1] * n[2] <= 200000:
        fd = gpu_solve(f, flags=ff.FF_SYMBOL_3PT)
        assert rel(fd, oracle.solve(f, symbol=3)) < TOL


def test_four_step_lines_in_a_batch():
    """A four-step instance in a step-batched batch falls back to its own
    launches and matches its single-plan solve bit for bit."""
    shapes = [(8, 4100, 6), (16, 16, 16), (4097, 5, 4)]
    items = []
    fs, outs = [], []
    rng = np.random.default_rng(11)
    for n in shapes:
        lo, hi, _ = ff.benchmark_grid(n)
        items.append({"n": n, "lo": lo, "hi": hi})
        f = rng.standard_normal(n)
        fs.append(f)
        outs.append(gpu_solve(f))
    b = ff.Batch(items, ff.FF_BATCH_STEPS)
    dens, pots = [], []
    for n, f in zip(shapes, fs):
        nb = b.bytes(len(dens))
        a = torch.zeros(nb // 8, dtype=torch.float64, device="cuda")
        put(a, f)
        dens.append(a)
        pots.append(torch.zeros(nb // 8, dtype=torch.float64, device="cuda"))
    b.solve(dens, pots)
    torch.cuda.synchronize()
    for k, n in enumerate(shapes):
        assert np.array_equal(get_real(pots[k], n), outs[k]), n
    b.close()


# even nz whose packed rows (nz/2 complex values) no workgroup transforms
# whole: 5-smooth (10000, 16384) and with Bluestein sub-lines (8500 = 85 *
# 100, 8195 = 5 * 11 * 149, 10001 = 73 * 137); odd nz above a workgroup's
# line (two real rows per complex row): 16875 = 3^3 5^4, 10125 = 3^4 5^3,
# 9801 = 99^2 (Bluestein sub-lines), 16385 = 5 * 29 * 113, and with nx*ny odd
# (the last row alone, through the work array's tail)
LONG_Z_SHAPES = [(3, 2, 20000), (2, 2, 32768), (4, 3, 17000), (2, 5, 16390), (5, 2, 20002),
                 (2, 2, 16875), (6, 5, 10125), (4, 2, 9801), (2, 3, 16385), (3, 3, 16875), (3, 5, 9801)]


@pytest.mark.parametrize("n", LONG_Z_SHAPES)
def test_four_step_z_rows_match_oracle(n):
    """The reference transforms z rows of any length (D2Z / Z2D plans,
    ParisPencil.cpp:80-81, 165, 380).  Packed rows too long for one workgroup run as
    four-step sub-line passes through the plan's work array, with the
    real-to-complex split and the complex-to-real merge as row-local steps:
    the solve, the z forward pass (pass 0) and the z inverse pass (pass 4)
    match the oracle; the density is left unchanged; the allocation-free
    entry point gives the same bits; FP32 + 2 steps (FP32 passes and FP64
    residual passes through the same work array) converges to the oracle."""
    rng = np.random.default_rng(n[2] + n[1])
    f = rng.standard_normal(n)
    lo, hi, _ = ff.benchmark_grid(n)
    plan, nb, a, b = make(n)
    fs = json.loads(plan.describe())["four_step"]
    say("four-step z", n, fs)
    assert fs["z"][0] * fs["z"][1] == (n[2] if n[2] % 2 else n[2] // 2) and fs["z"][0] > 1
    put(a, f)
    plan.solve(nb, a, b)
    torch.cuda.synchronize()
    got = get_real(b, n)
    assert rel(got, oracle.solve(f)) < TOL, rel(got, oracle.solve(f))
    assert np.array_equal(get_real(a, n), f)  # the rows of M live in the plan's work array
    c = torch.zeros_like(b)
    plan.solve_noalloc(nb, a, c)
    torch.cuda.synchronize()
    assert np.array_equal(get_real(c, n), got)
    # pass 0: z forward (D2Z)
    plan.solve_pass(0, a, b)
    torch.cuda.synchronize()
    assert rel(get_half(b, n), oracle.rfft(f)) < TOL
    # pass 4: z inverse (Z2D) and the 1/N scale (Im of DC / Nyquist ignored)
    nh = n[2] // 2 + 1
    spec = rng.standard_normal((n[0], n[1], nh)) + 1j * rng.standard_normal((n[0], n[1], nh))
    put_half(a, spec)
    plan.solve_pass(4, a, b)
    torch.cuda.synchronize()
    want = oracle.irfft(spec, n[2]) * (1.0 / (float(n[0]) * float(n[1]) * float(n[2])))
    assert rel(get_real(b, n), want) < TOL
    assert np.array_equal(get_half(a, n), spec)  # the source half spectrum is read only
    with pytest.raises(ff.FishFryError) as e:
        plan.solve_pass(4, a, a)  # the merge cannot run in place
    assert e.value.status == ff.FF_ERR_INVALID
    plan.close()
    got32 = gpu_solve(f, precision=ff.FF_FP32_REFINE, refine_steps=2)
    assert rel(got32, oracle.solve(f)) < 1e-12, rel(got32, oracle.solve(f))


@pytest.mark.parametrize("n", [(2, 4100, 16400), (4100, 2, 16400)])
def test_four_step_z_with_four_step_columns(n):
    """Four-step z rows with four-step y (or x) lines in one plan (1.3e8
    points; largest prime factor 41, so the oracle finishes in seconds), FP64
    and FP32 + 2 steps."""
    rng = np.random.default_rng(n[0])
    f = rng.standard_normal(n)
    plan, nb, a, b = make(n)
    fs = json.loads(plan.describe())["four_step"]
    say("four-step z + columns", n, fs)
    assert fs["z"][0] > 1 and (fs["y"][0] > 1 or fs["x"][0] > 1)
    plan.close()
    got = gpu_solve(f)
    want = oracle.solve(f)
    assert rel(got, want) < TOL, rel(got, want)
    got32 = gpu_solve(f, precision=ff.FF_FP32_REFINE, refine_steps=2)
    assert rel(got32, want) < 1e-12, rel(got32, want)


def test_four_step_z_limits():
    """Four-step z rows run in the plain layout: the pencil, blocked and slab
    layouts are refused with a message.  Every batch mode runs them (the
    grouped and step modes solve such a grid on its own), an odd nz with
    nx*ny odd included, and matches the single-plan solve bit for bit, also
    on a second solve and under graph capture."""
    n = (3, 2, 20000)
    lo, hi, _ = ff.benchmark_grid(n)
    for kw in ({"flags": ff.FF_PENCIL_LAYOUT}, {"flags": ff.FF_BLOCKED_LAYOUT}, {"flags": ff.FF_SLAB_PIPELINE},
               {"precision": ff.FF_FP32_REFINE, "refine_steps": 1, "flags": ff.FF_PENCIL_LAYOUT}):
        with pytest.raises(ff.FishFryError) as e:
            ff.ParisPencil(n, lo, hi, **kw)
        assert e.value.status == ff.FF_ERR_UNSUPPORTED, kw
        assert "four steps" in str(e.value)
    n3 = (3, 3, 16875)
    lo3, hi3, _ = ff.benchmark_grid(n3)
    items = [{"n": n, "lo": lo, "hi": hi}, {"n": n3, "lo": lo3, "hi": hi3}, {"n": (8, 8, 8)}, {"n": (8, 8, 8)}]
    rng = np.random.default_rng(12)
    fs = [rng.standard_normal(n), rng.standard_normal(n3), rng.standard_normal((8, 8, 8)),
          rng.standard_normal((8, 8, 8))]
    outs = [gpu_solve(f) for f in fs]
    # every batch mode runs them: the grouped and step modes solve such a
    # grid on its own, beside their grouped or multi-shape launches
    for flags in (0, ff.FF_BATCH_GROUPED, ff.FF_BATCH_STEPS, ff.FF_BATCH_STEPS | ff.FF_BATCH_GRAPH):
        bt = ff.Batch(items, flags)
        dens, pots = [], []
        for k, f in enumerate(fs):
            nb = bt.bytes(k)
            d = torch.zeros(nb // 8, dtype=torch.float64, device="cuda")
            put(d, f)
            dens.append(d)
            pots.append(torch.zeros(nb // 8, dtype=torch.float64, device="cuda"))
        s = torch.cuda.Stream()
        for _ in range(2):
            torch.cuda.synchronize()
            bt.solve(dens, pots, s)
            bt.wait(s)
            torch.cuda.synchronize()
            for k, f in enumerate(fs):
                assert np.array_equal(get_real(pots[k], f.shape), outs[k]), (flags, k)
            for k, f in enumerate(fs):
                put(dens[k], f)  # the solve may use density as work space
        bt.close()


def test_four_step_z_lone_row_only():
    """nx = ny = 1 with an odd nz above a workgroup's line: no row pairs at all,
    only the lone row's launches run (the pair launches have no rows).  The
    solve (on a domain of its own: the benchmark grid has hi = lo at n = 1),
    the z passes and FP32 + 2 steps match the oracle."""
    n = (1, 1, 16875)
    lo, hi = [0.0, 0.0, -math.pi], [1.0, 1.0, math.pi - 2.0 * math.pi / n[2]]
    rng = np.random.default_rng(31)
    f = rng.standard_normal(n)
    want = oracle.solve(f, lo, hi)
    for kw, tol in (({}, TOL), ({"precision": ff.FF_FP32_REFINE, "refine_steps": 2}, 1e-12)):
        plan = ff.ParisPencil(n, lo, hi, **kw)
        nb = plan.bytes()
        a = torch.zeros(nb // 8, dtype=torch.float64, device="cuda")
        b = torch.zeros(nb // 8, dtype=torch.float64, device="cuda")
        put(a, f)
        plan.solve(nb, a, b)
        torch.cuda.synchronize()
        assert rel(get_real(b, n), want) < tol, (kw, rel(get_real(b, n), want))
        if not kw:
            plan.solve_pass(0, a, b)
            torch.cuda.synchronize()
            assert rel(get_half(b, n), oracle.rfft(f)) < TOL
            spec = rng.standard_normal((1, 1, n[2] // 2 + 1)) + 1j * rng.standard_normal((1, 1, n[2] // 2 + 1))
            put_half(a, spec)
            plan.solve_pass(4, a, b)
            torch.cuda.synchronize()
            assert rel(get_real(b, n), oracle.irfft(spec, n[2]) / n[2]) < TOL
        plan.close()


@pytest.mark.parametrize("n", [(512, 512, 512), (768, 768, 768), (1000, 1000, 1000), (1024, 1024, 1024)])
def test_large_benchmark_norms(n):
    """Full-size FP64 solves: the analytic norms stay at round-off, and the
    linearity / mean-removal properties hold (size-independent checks)."""
    say("large", n)
    plan, nb, a, b = make(n)
    field = ff.field_default(n)
    ff.field_init(field, a)
    plan.solve(nb, a, b)
    l1, l2, linf = ff.field_norms(field, b)
    say("  norms", l1, l2, linf)
    assert linf < 4e-15 and l2 < 4e-15 and l1 < 4e-15
    # scale + constant shift: solve(2 f + 5) = 2 phi
    N = n[0] * n[1] * n[2]
    phi1 = b[:N].clone()
    ff.field_init(field, a)
    a[:N].mul_(2.0).add_(5.0)
    plan.solve(nb, a, b)
    d = (b[:N] - 2.0 * phi1).abs().max().item()
    assert d < 1e-13 * phi1.abs().max().item() * 2
    del phi1
    plan.close()
    torch.cuda.empty_cache()


def test_solve_beyond_2_31_points():
    """1296^3 = 2,176,782,336 points, more than 2^31 (the reference indexes in
    32-bit int and asserts its arrays stay below INT_MAX, ParisPencil.cpp:74):
    every pass indexes with 64-bit offsets.  Checked on the whole array by the
    analytic norms (the last planes included), and bitwise by exact scaling:
    doubling f doubles every operation of the solve exactly, so
    solve(2 f) = 2 solve(f) in every element."""
    n = (1296, 1296, 1296)
    N = n[0] * n[1] * n[2]
    assert N > 2 ** 31
    say("beyond 2^31", n)
    plan, nb, a, b = make(n)
    field = ff.field_default(n)
    ff.field_init(field, a)
    plan.solve(nb, a, b)
    l1, l2, linf = ff.field_norms(field, b)
    say("  norms", l1, l2, linf)
    assert linf < 4e-15 and l2 < 4e-15 and l1 < 4e-15
    phi1 = b[:N].clone()
    ff.field_init(field, a)  # the solve clobbered f
    a[:N].mul_(2.0)
    plan.solve(nb, a, b)
    assert torch.equal(b[:N], 2.0 * phi1)
    # the last four planes lie past element 2^31 and are not zero
    tail = b[N - 4 * n[1] * n[2]:N]
    assert tail.abs().max().item() > 0.1
    del phi1, tail
    plan.close()
    torch.cuda.empty_cache()


@pytest.mark.parametrize("n", [(64, 64, 64), (48, 40, 36), (22, 14, 26)])
def test_fp32_refinement_converges(n):
    say("refine", n)
    rng = np.random.default_rng(7)
    f = oracle.init_field(n)
    lo, hi, dx = oracle.grid(n)
    x = [lo[a] + np.arange(n[a]) * dx[a] for a in range(3)]
    X, Y, Z = np.meshgrid(*x, indexing="ij")
    for _ in range(8):
        k = rng.integers(-6, 7, size=3)
        f = f + rng.standard_normal() * np.cos(k[0] * X + k[1] * Y + k[2] * Z)
    want = oracle.solve(f)
    errs = []
    for steps in (0, 1, 2, 3):
        got = gpu_solve(f, precision=ff.FF_FP32_REFINE, refine_steps=steps)
        errs.append(rel(got, want))
        say("  steps", steps, "rel err vs FP64 oracle", errs[-1])
    assert errs[0] < 1e-6
    assert errs[1] < 1e-11
    assert errs[2] < 1e-14
    assert errs[3] < 1e-14
    # the input is preserved in FP32 mode (refinement re-reads f)
    plan, nb, a, b = make(n, precision=ff.FF_FP32_REFINE, refine_steps=1)
    put(a, f)
    plan.solve(nb, a, b)
    torch.cuda.synchronize()
    assert np.array_equal(get_real(a, n), f)
    plan.close()


def test_random_shapes_match_oracle():
    """60 seeded random shapes, 2..96 points per axis (odd and even nz, prime
    and prime-half lengths through Bluestein): FP64 solves against the oracle,
    and every fourth shape also in FP32 + 2 refinement steps."""
    rng = np.random.default_rng(20261203)
    worst64, worst32 = 0.0, 0.0
    for k in range(60):
        n = tuple(int(v) for v in rng.integers(2, 97, size=3))
        f = rng.standard_normal(n)
        want = oracle.solve(f)
        say("random shape", k, n)
        e = rel(gpu_solve(f), want)
        worst64 = max(worst64, e)
        assert e < TOL, (n, e)
        if k % 4 == 0:
            e32 = rel(gpu_solve(f, precision=ff.FF_FP32_REFINE, refine_steps=2), want)
            worst32 = max(worst32, e32)
            assert e32 < 1e-12, (n, e32)
    say("random shapes: worst rel L2 FP64", worst64, "FP32+2", worst32)
