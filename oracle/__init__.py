"""CPU oracle for fishfry's periodic Poisson solve -- TEST INFRASTRUCTURE ONLY.

This package is the checker: only tests/, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` may import it.  The product
(``fishfry_amd``) never imports it and never falls back to it.

It wraps ``liboracle.so`` (built from ``oracle.c`` by ``make -C oracle``) with
numpy-friendly helpers.  Each function names the reference lines it restates;
``oracle.h`` lists how every routine is pinned.
"""
from __future__ import annotations

import ctypes
import math
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")
_lib = None

_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)


def build(force: bool = False) -> str:
    """Compile liboracle.so with the Makefile beside this file."""
    args = ["make", "-C", _HERE]
    if force:
        args.insert(1, "-B")
    subprocess.run(args, check=True, stdout=subprocess.DEVNULL)
    return _LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        L.ora_fft.argtypes = [ctypes.c_int, _dp, ctypes.c_long, ctypes.c_long, ctypes.c_long, ctypes.c_int]
        L.ora_rfft.argtypes = [ctypes.c_int, _dp, _dp, ctypes.c_long]
        L.ora_irfft.argtypes = [ctypes.c_int, _dp, _dp, ctypes.c_long]
        L.ora_dd.argtypes = [ctypes.c_int, ctypes.c_double, ctypes.c_double]
        L.ora_dd.restype = ctypes.c_double
        L.ora_bytes.argtypes = [_ip, _ip]
        L.ora_bytes.restype = ctypes.c_size_t
        L.ora_solve.argtypes = [_ip, _dp, _dp, _dp, _dp, _dp, ctypes.c_int]
        L.ora_filter.argtypes = [_ip, _dp, _dp, _dp]
        L.ora_filter.restype = None
        L.ora_init_field.argtypes = [ctypes.c_int] * 3 + [ctypes.c_double] * 6 + [_dp]
        L.ora_diff.argtypes = [ctypes.c_int] * 3 + [ctypes.c_double] * 6 + [_dp, _dp]
        L.ora_twiddle.argtypes = [ctypes.c_long, ctypes.c_long, _dp, _dp]
        L.ora_fd_factor.argtypes = [ctypes.c_long, ctypes.c_long, ctypes.c_int]
        L.ora_fd_factor.restype = ctypes.c_double
        L.ora_filter_sym.argtypes = [_ip, _dp, _dp, ctypes.c_int, _dp]
        L.ora_filter_sym.restype = None
        L.ora_solve_sym.argtypes = [_ip, _dp, _dp, ctypes.c_int, _dp, _dp, _dp, ctypes.c_int]
        L.ora_stencil.argtypes = [_ip, _dp, _dp, ctypes.c_int, _dp, _dp]
        L.ora_stencil.restype = None
        for f in ("ora_fft", "ora_rfft", "ora_irfft", "ora_solve", "ora_solve_sym"):
            getattr(L, f).restype = ctypes.c_int
        _lib = L
    return _lib


def _ptr(a: np.ndarray):
    return a.ctypes.data_as(_dp)


def _ints(v):
    return (ctypes.c_int * 3)(*[int(x) for x in v])


def _dbls(v):
    return (ctypes.c_double * 3)(*[float(x) for x in v])


def fft(x: np.ndarray, sign: int = -1, axis: int = -1) -> np.ndarray:
    """Unnormalized complex FFT along `axis` (exponent `sign`)."""
    a = np.ascontiguousarray(np.moveaxis(np.asarray(x, dtype=np.complex128), axis, -1)).copy()
    n = a.shape[-1]
    how = a.size // n if n else 0
    rc = lib().ora_fft(n, _ptr(a.view(np.float64)), how, 1, n, sign)
    if rc:
        raise RuntimeError("ora_fft failed")
    return np.moveaxis(a, -1, axis)


def rfft(x: np.ndarray) -> np.ndarray:
    """R2C along the last (contiguous) axis, ParisPencil.cpp:81,165."""
    a = np.ascontiguousarray(x, dtype=np.float64)
    n = a.shape[-1]
    out = np.empty(a.shape[:-1] + (n // 2 + 1,), dtype=np.complex128)
    if lib().ora_rfft(n, _ptr(a), _ptr(out.view(np.float64)), a.size // n):
        raise RuntimeError("ora_rfft failed")
    return out


def irfft(x: np.ndarray, n: int) -> np.ndarray:
    """C2R (unnormalized) along the last axis, ParisPencil.cpp:80,380."""
    a = np.ascontiguousarray(x, dtype=np.complex128)
    assert a.shape[-1] == n // 2 + 1
    out = np.empty(a.shape[:-1] + (n,), dtype=np.float64)
    if lib().ora_irfft(n, _ptr(a.view(np.float64)), _ptr(out), a.size // a.shape[-1]):
        raise RuntimeError("ora_irfft failed")
    return out


def dd(n: int, lo: float, hi: float) -> float:
    return lib().ora_dd(n, lo, hi)


def ref_bytes(n, m=(1, 1, 1)) -> int:
    return int(lib().ora_bytes(_ints(n), _ints(m)))


def grid(n):
    """Grid of FishFry.hpp:108-139 at 1x1x1 tasks: lo, hi, dx per axis."""
    dx = [2.0 * math.pi / n[a] for a in range(3)]
    lo = [-math.pi] * 3
    hi = [math.pi - dx[a] for a in range(3)]
    return lo, hi, dx


def solve(density: np.ndarray, lo=None, hi=None, nthreads: int = 0, symbol: int = 0) -> np.ndarray:
    """Full periodic Poisson solve, ParisPencil.cpp:95-425 at 1x1x1 tasks.
    symbol 3 or 5: the 3-point / 5-point finite-difference symbol in place of
    the spectral filter (the options make-frontier.sh:9-10 names)."""
    d = np.ascontiguousarray(density, dtype=np.float64)
    n = d.shape
    if lo is None or hi is None:
        lo, hi, _ = grid(n)
    nh = n[2] // 2 + 1
    work = np.empty(2 * n[0] * n[1] * nh, dtype=np.float64)
    out = np.empty_like(d)
    if lib().ora_solve_sym(_ints(n), _dbls(lo), _dbls(hi), int(symbol), _ptr(d), _ptr(out), _ptr(work),
                           int(nthreads)):
        raise RuntimeError("ora_solve failed")
    return out


def fd_factor(k: int, n: int, order: int) -> float:
    """4 s^2 (3-point) or 4 s^2 (1 + s^2/3) (5-point), s = sin(pi k / n)."""
    return lib().ora_fd_factor(int(k), int(n), int(order))


def stencil(phi: np.ndarray, order: int, lo=None, hi=None) -> np.ndarray:
    """L_h phi with the 3- or 5-point stencil on every axis (periodic)."""
    p = np.ascontiguousarray(phi, dtype=np.float64)
    n = p.shape
    if lo is None or hi is None:
        lo, hi, _ = grid(n)
    out = np.empty_like(p)
    lib().ora_stencil(_ints(n), _dbls(lo), _dbls(hi), int(order), _ptr(p), _ptr(out))
    return out


def filter_half(spec: np.ndarray, nz: int, lo=None, hi=None, symbol: int = 0) -> np.ndarray:
    """Spectral filter of ParisPencil.cpp:263-285 (or, symbol 3 / 5, the
    finite-difference one) applied to a copy of the half spectrum
    spec[nx][ny][nz/2+1]."""
    a = np.ascontiguousarray(spec, dtype=np.complex128).copy()
    n = (a.shape[0], a.shape[1], nz)
    assert a.shape[2] == nz // 2 + 1
    if lo is None or hi is None:
        lo, hi, _ = grid(n)
    lib().ora_filter_sym(_ints(n), _dbls(lo), _dbls(hi), int(symbol), _ptr(a.view(np.float64)))
    return a


def init_field(n) -> np.ndarray:
    """f of FishFry.hpp:14-45 on the 1x1x1 grid."""
    lo, _, dx = grid(n)
    f = np.empty(tuple(n), dtype=np.float64)
    lib().ora_init_field(n[0], n[1], n[2], lo[0], lo[1], lo[2], dx[0], dx[1], dx[2], _ptr(f))
    return f


def diff(phi: np.ndarray):
    """(sum|e|, sum e^2, max|e|) against analytic phi, FishFry.hpp:57-101."""
    p = np.ascontiguousarray(phi, dtype=np.float64)
    n = p.shape
    lo, _, dx = grid(n)
    out = np.zeros(3, dtype=np.float64)
    lib().ora_diff(n[0], n[1], n[2], lo[0], lo[1], lo[2], dx[0], dx[1], dx[2], _ptr(p), _ptr(out))
    return out


def norms(phi: np.ndarray):
    """Relative L1, L2, Linf of FishFry.hpp:146-159,198."""
    s = diff(np.zeros_like(phi))
    e = diff(phi)
    return e[0] / s[0], math.sqrt(e[1] / s[1]), e[2] / s[2]


def twiddle(t: int, n: int):
    c = ctypes.c_double()
    s = ctypes.c_double()
    lib().ora_twiddle(t, n, ctypes.byref(c), ctypes.byref(s))
    return complex(c.value, s.value)
