"""Write tests/golden/golden.npz: known-answer vectors for the FFT pieces and
the full periodic Poisson solve of fishfry (ParisPencil.cpp:95-425 at 1x1x1
tasks).

The reference ships no tests or fixtures and its FFT lives in cuFFT/hipFFT
(not vendored, not linkable here), so the expected outputs come from an
independent implementation: scipy.fft (pocketfft) evaluated in long double
(80-bit) on float64 inputs, then rounded to float64.  The solve follows the
reference's sequence: D2Z along z (ParisPencil.cpp:165), forward C2C along y
(:212) and x (:259), the spectral filter (:263-285) with the reference's dd
(:9-11), inverse C2C along x (:288) and y (:334), C2R along z (:380), scale by
1/(nx*ny*nz) (:406-421).

Run from the repo root:  python tests/golden/make_golden.py
"""
import math
import os

import numpy as np
import scipy.fft as sf

SEED = 20261202
HERE = os.path.dirname(os.path.abspath(__file__))

FFT_LENGTHS = [1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 11, 12, 13, 15, 16, 17, 20, 24, 25, 27, 30, 31, 32,
               36, 37, 40, 45, 48, 49, 50, 60, 64, 72, 80, 81, 96, 97, 100, 101, 120, 125, 128,
               160, 192, 243, 250, 256, 384, 500, 509, 512, 625, 640, 729, 768, 1000, 1024]
RFFT_LENGTHS = [1, 2, 3, 4, 5, 6, 8, 9, 10, 15, 16, 17, 27, 32, 37, 48, 64, 96, 100, 101,
                128, 160, 250, 384, 500, 512, 768, 1000, 1024]
SOLVE_SHAPES = [(8, 8, 8), (12, 10, 9), (16, 24, 20), (32, 32, 32), (7, 5, 11), (6, 9, 15),
                (5, 6, 4), (9, 8, 12)]


def grid(n):
    dx = [2.0 * math.pi / n[a] for a in range(3)]
    lo = [-math.pi] * 3
    hi = [math.pi - dx[a] for a in range(3)]
    return lo, hi, dx


def dd(n, lo, hi):
    return 2.0 * math.pi * float(n - 1) / (float(n) * (hi - lo))


def field(n):
    """f of FishFry.hpp:14-45, evaluated in long double."""
    lo, _, dx = grid(n)
    L = np.longdouble
    x = L(lo[0]) + np.arange(n[0], dtype=L) * L(dx[0])
    y = L(lo[1]) + np.arange(n[1], dtype=L) * L(dx[1])
    z = L(lo[2]) + np.arange(n[2], dtype=L) * L(dx[2])
    cx, sx = np.cos(x), np.sin(x)
    ex = np.exp(sx)
    px, dpx = cx * ex, -ex * sx * cx * (sx + 3)
    cy, sy = np.cos(y), np.sin(y)
    ey = np.exp(cy)
    py, dpy = sy * ey, sy * ey * (sy * sy - 3 * cy - 1)
    cz, sz = np.cos(z), np.sin(z)
    ez = np.exp(cz)
    pz, dpz = cz * ez, ez * (3 * sz * sz - cz * cz * cz - 1)
    f = (px[:, None, None] * py[None, :, None] * dpz[None, None, :]
         + px[:, None, None] * dpy[None, :, None] * pz[None, None, :]
         + dpx[:, None, None] * py[None, :, None] * pz[None, None, :])
    return f


def solve_ld(density):
    """Long-double restatement of the reference's solve sequence."""
    n = density.shape
    lo, hi, _ = grid(n)
    ddi, ddj, ddk = (dd(n[a], lo[a], hi[a]) for a in range(3))
    a = sf.rfft(density.astype(np.longdouble), axis=2)
    a = sf.fft(a, axis=1)
    a = sf.fft(a, axis=0)
    nh = n[2] // 2 + 1
    L = np.longdouble
    i = np.arange(n[0])
    j = np.arange(n[1])
    k = np.arange(nh)
    i2 = (np.minimum(i, n[0] - i).astype(np.float64) * ddi) ** 2
    j2 = (np.minimum(j, n[1] - j).astype(np.float64) * ddj) ** 2
    k2 = (k.astype(np.float64) * ddk) ** 2
    s = i2[:, None, None].astype(L) + j2[None, :, None].astype(L) + k2[None, None, :].astype(L)
    s[0, 0, 0] = 1
    d = -1 / s
    d[0, 0, 0] = 0
    a = a * d
    a = sf.ifft(a, axis=0, norm="forward")
    a = sf.ifft(a, axis=1, norm="forward")
    phi = sf.irfft(a, n=n[2], axis=2, norm="forward")
    divN = 1.0 / (float(n[0]) * float(n[1]) * float(n[2]))
    return (phi * L(divN)).astype(np.float64)


def main():
    rng = np.random.default_rng(SEED)
    out = {}
    for n in FFT_LENGTHS:
        x = rng.standard_normal(n) + 1j * rng.standard_normal(n)
        xl = x.astype(np.clongdouble)
        out[f"fft_in_{n}"] = x
        out[f"fft_fwd_{n}"] = sf.fft(xl).astype(np.complex128)
        out[f"fft_inv_{n}"] = sf.ifft(xl, norm="forward").astype(np.complex128)
    for n in RFFT_LENGTHS:
        x = rng.standard_normal((3, n))
        out[f"rfft_in_{n}"] = x
        out[f"rfft_out_{n}"] = sf.rfft(x.astype(np.longdouble), axis=1).astype(np.complex128)
        nh = n // 2 + 1
        # a half spectrum with non-zero Im on DC (and Nyquist for even n): the
        # C2R must read it as Hermitian, i.e. drop those imaginary parts
        h = rng.standard_normal((3, nh)) + 1j * rng.standard_normal((3, nh))
        out[f"irfft_in_{n}"] = h
        out[f"irfft_out_{n}"] = sf.irfft(h.astype(np.clongdouble), n=n, axis=1,
                                         norm="forward").astype(np.float64)
    for shp in SOLVE_SHAPES:
        tag = "x".join(map(str, shp))
        r = rng.standard_normal(shp)
        out[f"solve_rand_in_{tag}"] = r
        out[f"solve_rand_out_{tag}"] = solve_ld(r)
        f = field(shp)
        out[f"solve_field_in_{tag}"] = f.astype(np.float64)
        out[f"solve_field_out_{tag}"] = solve_ld(f.astype(np.float64))
    out["meta_seed"] = np.array([SEED])
    np.savez_compressed(os.path.join(HERE, "golden.npz"), **out)
    print("wrote", len(out), "arrays")


if __name__ == "__main__":
    main()
