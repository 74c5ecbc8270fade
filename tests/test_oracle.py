"""Pins the CPU oracle (oracle/) before anything is checked against it.

Each check is chosen so a plausible mistake in the oracle fails it: a wrong
exponent sign or twiddle index fails the brute-force DFT, a dropped scale fails
the golden solves, a wrong filter constant fails the single-mode closed form, a
missing mean removal fails the constant-shift test, a transposed axis fails the
non-cubic golden solves and the shift-equivariance test.
"""
import math

import numpy as np
import pytest

import oracle


def rel(a, b):
    a = np.asarray(a)
    b = np.asarray(b)
    den = np.linalg.norm(b.ravel())
    return np.linalg.norm((a - b).ravel()) / (den if den else 1.0)


PI_L = 4 * np.arctan(np.longdouble(1))  # pi in long double (np.pi is only a double)


def brute_dft(x, sign):
    n = len(x)
    L = np.longdouble
    k = np.arange(n)
    ang = (2 * PI_L) * (np.outer(k, k) % n).astype(L) / L(n)
    m = np.cos(ang) + sign * 1j * np.sin(ang)
    return m.astype(np.clongdouble) @ x.astype(np.clongdouble)


@pytest.mark.parametrize("n", [1, 2, 3, 4, 5, 6, 7, 8, 9, 11, 12, 13, 16, 17, 25, 27, 31, 32, 37, 48, 60, 64])
def test_fft_brute_force(n):
    rng = np.random.default_rng(n)
    x = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    for sign in (-1, +1):
        assert rel(oracle.fft(x, sign=sign), brute_dft(x, sign)) < 3e-16 * max(1.0, math.log2(n)) * (1 if n in (2, 3, 4, 5) or max(f for f in range(1, n + 1) if n % f == 0 and all(f % q for q in range(2, f))) <= 5 else math.sqrt(n))


def test_fft_golden(golden):
    ns = sorted(int(k[len("fft_in_"):]) for k in golden if k.startswith("fft_in_"))
    assert len(ns) > 50
    for n in ns:
        x = golden[f"fft_in_{n}"]
        tol = 4e-16 * max(1.0, math.log2(n)) if n != 509 else 3e-15
        assert rel(oracle.fft(x, -1), golden[f"fft_fwd_{n}"]) < tol, n
        assert rel(oracle.fft(x, +1), golden[f"fft_inv_{n}"]) < tol, n


def test_fft_known_answers():
    n = 48
    # impulse at t0 -> pure tone exp(-2 pi i k t0 / n)
    x = np.zeros(n, complex)
    x[5] = 1.0
    y = oracle.fft(x)
    k = np.arange(n)
    assert np.abs(y - np.exp(-2j * PI_L * ((k * 5) % n) / n)).max() < 3e-16
    # tone -> impulse of height n
    t = np.exp(2j * PI_L * ((7 * k) % n) / n).astype(complex)
    y = oracle.fft(t)
    e = np.zeros(n, complex)
    e[7] = n
    assert np.abs(y - e).max() < 1e-13
    # Parseval
    rng = np.random.default_rng(3)
    x = rng.standard_normal(1000) + 1j * rng.standard_normal(1000)
    y = oracle.fft(x)
    assert abs(np.vdot(y, y).real / 1000 - np.vdot(x, x).real) < 1e-11 * np.vdot(x, x).real
    # inverse(forward(x)) = n x
    assert rel(oracle.fft(oracle.fft(x), +1) / 1000, x) < 1e-15


def test_twiddles():
    for n in (4, 8, 12, 1000, 1024, 768, 509):
        for t in range(0, n, max(1, n // 97)):
            w = oracle.twiddle(t, n)
            ref = np.exp(-2j * PI_L * np.longdouble(t) / np.longdouble(n))
            assert abs(w - complex(ref)) < 2.3e-16
        # exact special points and symmetry
        if n % 4 == 0:
            assert oracle.twiddle(n // 4, n) == complex(0, -1)
            assert oracle.twiddle(n // 2, n) == complex(-1, 0)
            assert oracle.twiddle(3 * n // 4, n) == complex(0, 1)
        for t in range(1, n):
            a = oracle.twiddle(t, n)
            b = oracle.twiddle(n - t, n)
            assert a == b.conjugate()


def test_rfft_golden(golden):
    ns = sorted(int(k[len("rfft_in_"):]) for k in golden if k.startswith("rfft_in_"))
    for n in ns:
        x = golden[f"rfft_in_{n}"]
        tol = 5e-16 * max(1.0, math.log2(n))
        assert rel(oracle.rfft(x), golden[f"rfft_out_{n}"]) < tol, n
        h = golden[f"irfft_in_{n}"]
        assert rel(oracle.irfft(h, n), golden[f"irfft_out_{n}"]) < tol, n
        # round trip
        assert rel(oracle.irfft(oracle.rfft(x), n) / n, x) < tol, n


@pytest.mark.parametrize("kind", ["rand", "field"])
def test_solve_golden(golden, kind):
    shapes = sorted({k[len(f"solve_{kind}_in_"):] for k in golden if k.startswith(f"solve_{kind}_in_")})
    assert len(shapes) >= 8
    for tag in shapes:
        d = golden[f"solve_{kind}_in_{tag}"]
        expect = golden[f"solve_{kind}_out_{tag}"]
        got = oracle.solve(d.copy())
        assert rel(got, expect) < 5e-15, tag


@pytest.mark.parametrize("shape", [(16, 24, 20), (12, 10, 9), (32, 32, 32), (9, 7, 13)])
def test_single_mode_closed_form(shape):
    lo, hi, dx = oracle.grid(shape)
    n = shape
    dd = [oracle.dd(n[a], lo[a], hi[a]) for a in range(3)]
    x = [lo[a] + np.arange(n[a]) * dx[a] for a in range(3)]
    X, Y, Z = np.meshgrid(*x, indexing="ij")
    a = (2, 3, 1)
    arg = a[0] * X + a[1] * Y + a[2] * Z
    f = np.cos(arg) + 0.5 * np.sin(arg)
    lam = (a[0] * dd[0]) ** 2 + (a[1] * dd[1]) ** 2 + (a[2] * dd[2]) ** 2
    phi = oracle.solve(f.copy())
    assert rel(phi, -f / lam) < 1e-14


def test_mean_removal_linearity_shift():
    shape = (16, 12, 10)
    rng = np.random.default_rng(11)
    f = rng.standard_normal(shape)
    g = rng.standard_normal(shape)
    pf = oracle.solve(f)
    pg = oracle.solve(g)
    # the mean mode is zeroed (ParisPencil.cpp:275,282): adding a constant changes nothing
    assert rel(oracle.solve(f + 3.25), pf) < 1e-14
    assert abs(pf.mean()) < 1e-15 * np.abs(pf).max()
    # linearity
    assert rel(oracle.solve(2.0 * f - 0.5 * g), 2.0 * pf - 0.5 * pg) < 1e-14
    # circular-shift equivariance on each axis
    for ax, s in ((0, 3), (1, 5), (2, 7)):
        assert rel(oracle.solve(np.roll(f, s, axis=ax)), np.roll(pf, s, axis=ax)) < 1e-14


def test_laplacian_of_solution():
    """L_h phi = f - mean(f) with L_h applied by an explicit long-double DFT."""
    shape = (6, 5, 4)
    rng = np.random.default_rng(5)
    f = rng.standard_normal(shape)
    phi = oracle.solve(f)
    lo, hi, _ = oracle.grid(shape)
    L = np.longdouble

    def dftm(n, sign):
        k = np.arange(n)
        ang = (2 * PI_L) * (np.outer(k, k) % n).astype(L) / L(n)
        return (np.cos(ang) + sign * 1j * np.sin(ang)).astype(np.clongdouble)

    def sym(n, d):
        k = np.arange(n)
        return (np.minimum(k, n - k).astype(np.float64) * d) ** 2

    p = phi.astype(np.clongdouble)
    for ax in range(3):
        p = np.moveaxis(np.tensordot(dftm(shape[ax], -1), np.moveaxis(p, ax, 0), axes=1), 0, ax)
    dd = [oracle.dd(shape[a], lo[a], hi[a]) for a in range(3)]
    s = sym(shape[0], dd[0])[:, None, None] + sym(shape[1], dd[1])[None, :, None] + sym(shape[2], dd[2])[None, None, :]
    p = -p * s.astype(L)
    for ax in range(3):
        p = np.moveaxis(np.tensordot(dftm(shape[ax], +1), np.moveaxis(p, ax, 0), axes=1), 0, ax)
    lap = (p / L(f.size)).real.astype(np.float64)
    assert rel(lap, f - f.mean()) < 1e-14


def test_dd_constants():
    for n in (32, 48, 64, 96, 128, 160):
        lo, hi, _ = oracle.grid((n, n, n))
        assert oracle.dd(n, lo[0], hi[0]) == 1.0
    lo, hi, _ = oracle.grid((512,) * 3)
    assert oracle.dd(512, lo[0], hi[0]) == 1.0 - 2.0 ** -53
    lo, hi, _ = oracle.grid((1024,) * 3)
    assert oracle.dd(1024, lo[0], hi[0]) == 1.0 + 2.0 ** -52


def test_reference_bytes():
    # at 1x1x1 the largest work array is the complex half spectrum, 2*nx*ny*nh doubles
    for n in ((64, 64, 64), (1024, 1024, 1024), (12, 10, 9), (128, 64, 96), (7, 5, 11)):
        nh = n[2] // 2 + 1
        assert oracle.ref_bytes(n) == 16 * n[0] * n[1] * nh
    assert oracle.ref_bytes((64, 64, 64)) == 2162688
    assert oracle.ref_bytes((1024, 1024, 1024)) == 8606711808


def test_analytic_norms():
    # the benchmark's own self-check (FishFry.hpp:189-202): FP64 solve is exact to round-off
    for n in ((32, 32, 32), (64, 64, 64), (48, 40, 36)):
        f = oracle.init_field(n)
        l1, l2, linf = oracle.norms(oracle.solve(f))
        assert linf < 2e-15 and l2 < 2e-15 and l1 < 2e-15, (n, l1, l2, linf)
    # a wrong answer is detected: phi = 0 has relative norms exactly 1
    l1, l2, linf = oracle.norms(np.zeros((8, 8, 8)))
    assert (l1, l2, linf) == (1.0, 1.0, 1.0)


def test_init_field_long_double(golden):
    for tag in ("8x8x8", "12x10x9", "16x24x20"):
        shp = tuple(int(v) for v in tag.split("x"))
        f = oracle.init_field(shp)
        ref = golden[f"solve_field_in_{tag}"]
        assert np.abs(f - ref).max() < 1e-14 * np.abs(ref).max()


def test_diff_direct_long_double():
    """ora_diff against sums formed here in numpy long double from the
    analytic potential phi = e^{sin x} cos x * e^{cos y} sin y * e^{cos z} cos z
    (FishFry.hpp:57-101) on an arbitrary field: a wrong axis function, a
    missing square or a wrong maximum fails it, and phi = 0 gives the scales."""
    rng = np.random.default_rng(71)
    for n in ((8, 12, 10), (16, 16, 16), (7, 5, 11)):
        lo, _, dx = oracle.grid(n)
        ld = np.longdouble
        x = ld(lo[0]) + np.arange(n[0], dtype=ld) * ld(dx[0])
        y = ld(lo[1]) + np.arange(n[1], dtype=ld) * ld(dx[1])
        z = ld(lo[2]) + np.arange(n[2], dtype=ld) * ld(dx[2])
        px = np.exp(np.sin(x)) * np.cos(x)
        py = np.exp(np.cos(y)) * np.sin(y)
        pz = np.exp(np.cos(z)) * np.cos(z)
        exact = px[:, None, None] * py[None, :, None] * pz[None, None, :]
        for phi in (rng.standard_normal(n), np.zeros(n), np.asarray(exact, dtype=np.float64) * 0.5):
            e = np.abs(exact - phi.astype(ld))
            want = (np.sum(e), np.sum(e * e), np.max(e))
            got = oracle.diff(phi)
            for a, b in zip(got, want):
                assert abs(ld(a) - b) <= 1e-13 * abs(b), (n, got, want)


def numpy_stencil(phi, order, lo, hi):
    """L_h phi written independently with np.roll (periodic)."""
    out = np.zeros_like(phi)
    for ax in range(3):
        n = phi.shape[ax]
        h = (hi[ax] - lo[ax]) / (n - 1)
        p1, m1 = np.roll(phi, -1, ax), np.roll(phi, 1, ax)
        if order == 3:
            d2 = p1 + m1 - 2.0 * phi
        else:
            p2, m2 = np.roll(phi, -2, ax), np.roll(phi, 2, ax)
            d2 = (16.0 * (p1 + m1) - (p2 + m2) - 30.0 * phi) / 12.0
        out += d2 / (h * h)
    return out


@pytest.mark.parametrize("order", [3, 5])
@pytest.mark.parametrize("n", [(8, 10, 12), (9, 7, 11), (16, 16, 16), (5, 6, 4), (12, 3, 2)])
def test_fd_solve_inverts_the_stencil(order, n):
    """The finite-difference solve inverts the stencil written here with
    np.roll: L_h phi = f - mean f, phi has no mean; and the oracle's own
    real-space stencil agrees with the numpy one on a random field.  A wrong
    symbol (sin for sin^2, a dropped 1/h^2, a wrong 5-point coefficient)
    fails the first check."""
    rng = np.random.default_rng(72)
    f = rng.standard_normal(n)
    lo, hi = [0.0, -1.0, 0.5], [2.0, 3.0, 4.5]
    phi = oracle.solve(f, lo, hi, symbol=order)
    lap = numpy_stencil(phi, order, lo, hi)
    assert np.max(np.abs(lap - (f - f.mean()))) < 1e-12 * np.max(np.abs(f))
    assert abs(phi.mean()) < 1e-14 * np.max(np.abs(phi))
    g = rng.standard_normal(n)
    assert np.max(np.abs(oracle.stencil(g, order, lo, hi) - numpy_stencil(g, order, lo, hi))) < 1e-12 * np.max(
        np.abs(numpy_stencil(g, order, lo, hi)))


def test_fd_convergence_orders():
    """Against the benchmark's analytic potential the 3-point solve's error
    falls as h^2 and the 5-point one's as h^4 (error ratios near 4 and 16 when
    n doubles), while the spectral solve is exact to round-off."""
    errs = {3: [], 5: []}
    for m in (32, 64):
        n = (m, m, m)
        f = oracle.init_field(n)
        for order in (3, 5):
            errs[order].append(oracle.norms(oracle.solve(f, symbol=order))[2])
    r3 = errs[3][0] / errs[3][1]
    r5 = errs[5][0] / errs[5][1]
    assert 3.6 < r3 < 4.4, errs
    assert 14.0 < r5 < 18.0, errs
    assert errs[5][1] < errs[3][1] / 50


def test_fd_factor_limits():
    """Small angles: with theta = 2 pi k / n the 3-point factor is theta^2 -
    theta^4/12 + ..., the 5-point one theta^2 - theta^6/90 + ... (the Taylor
    series of the two stencils' eigenvalues); symmetric in k -> n - k; zero at
    k = 0; at the Nyquist mode 4 and 16/3."""
    n = 4096
    for k in (1, 2, 5, 9):
        t = 2 * math.pi * k / n
        g3, g5 = oracle.fd_factor(k, n, 3), oracle.fd_factor(k, n, 5)
        assert abs((t * t - g3) / t ** 4 - 1 / 12) < 1e-3
        assert abs((t * t - g5) / t ** 6 - 1 / 90) < 2e-2 or abs(t * t - g5) < 1e-15 * t * t
        assert oracle.fd_factor(n - k, n, 3) == g3 and oracle.fd_factor(n - k, n, 5) == g5
    assert oracle.fd_factor(0, 64, 3) == 0.0 and oracle.fd_factor(0, 64, 5) == 0.0
    assert oracle.fd_factor(32, 64, 3) == 4.0
    assert abs(oracle.fd_factor(32, 64, 5) - 16.0 / 3.0) < 1e-15


@pytest.mark.parametrize("order", [3, 5])
def test_fd_single_mode(order):
    """One Fourier mode: the discrete solution is -f divided by the stencil's
    eigenvalue, formed here from cosines (2 - 2 cos(theta), 5-point
    (30 - 32 cos(theta) + 2 cos(2 theta))/12), a different route to the
    same number than the oracle's sin^2 factor."""
    n = (12, 10, 16)
    lo, hi, dx = oracle.grid(n)
    k = (2, 3, 5)
    idx = np.meshgrid(*[np.arange(m) for m in n], indexing="ij")
    f = np.cos(sum(2 * math.pi * k[a] * idx[a] / n[a] for a in range(3)))
    lam = 0.0
    for a in range(3):
        th = 2 * math.pi * k[a] / n[a]
        h = (hi[a] - lo[a]) / (n[a] - 1)
        e = 2 - 2 * math.cos(th) if order == 3 else (30 - 32 * math.cos(th) + 2 * math.cos(2 * th)) / 12
        lam += e / (h * h)
    phi = oracle.solve(f, symbol=order)
    # (the cosine forms cancel for small angles: a few ulp of the eigenvalue)
    assert np.max(np.abs(phi + f / lam)) < 1e-13 * np.max(np.abs(f / lam))
