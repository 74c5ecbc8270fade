"""CPU checks of the index algebra and formulas of the row-local steps of
four-step z rows (fishfry_amd/csrc/ff_zsplit.hip), restated here in numpy:

* even nz: the real-to-complex split of the packed transform Z (item k <= M/2
  reads Z[k], Z[M-k] and writes X[k], X[M-k], and X[M] at k = 0) and the
  complex-to-real merge, against numpy's rfft / irfft;
* odd nz (row pairs): the in-place separation of a pair's transform into two
  spectrum rows -- each item (k <= k' = nh-1-k) reads and writes the same four
  places, the items' places are disjoint and cover both rows -- and the merge
  of two half spectra into the pair's complex spectrum; with nx*ny odd, the
  lone last row as a pair with a zero second row.

A plausible slip (a conjugate on the wrong term, a sign of the 1/(2i), an
index k against n-k, an overlap between items) fails one of these."""
import numpy as np
import pytest


def split_even(Z, nz):
    """k_zsplit_fwd, item by item, in place on a row of nh = M+1 values."""
    M = nz // 2
    row = np.zeros(M + 1, dtype=complex)
    row[:M] = Z
    twp = np.exp(-2j * np.pi * np.arange(M) / nz)

    def half(zk, zo, k):
        zm = np.conj(zo)
        ev = 0.5 * (zk + zm)
        od = complex(0.5 * (zk.imag - zm.imag), -0.5 * (zk.real - zm.real))
        return ev + od * twp[k]

    written = set()
    for k in range(M // 2 + 1):
        km = M - k if k else 0
        zk, zo = row[k], row[km]
        out = {k: half(zk, zo, k)}
        if km != k:
            out[km] = half(zo, zk, km)
        if k == 0:
            out[M] = complex(zk.real - zk.imag, 0.0)
        assert not (set(out) & written)  # no place is written twice
        written |= set(out)
        for i, v in out.items():
            row[i] = v
    assert written == set(range(M + 1))
    return row


def merge_even(X, nz):
    """k_zsplit_inv: the half spectrum (Hermitian) to the packed spectrum."""
    M = nz // 2
    twp = np.exp(-2j * np.pi * np.arange(M) / nz)
    Z = np.empty(M, dtype=complex)
    for k in range(M):
        xk, xm = X[k], X[M - k]
        if k == 0:
            xk, xm = complex(xk.real, 0), complex(xm.real, 0)
        cm = np.conj(xm)
        A = xk + cm
        B = (xk - cm) * np.conj(twp[k])
        Z[k] = complex(A.real - B.imag, A.imag + B.real)
    return Z


@pytest.mark.parametrize("nz", [2, 4, 6, 10, 16, 30, 64, 100, 1000])
def test_even_split_and_merge(nz):
    rng = np.random.default_rng(nz)
    x = rng.standard_normal(nz)
    Z = np.fft.fft(x[0::2] + 1j * x[1::2])
    X = split_even(Z, nz)
    assert np.allclose(X, np.fft.rfft(x), rtol=0, atol=1e-12 * nz)
    # merge then the M-point inverse gives the packed real pairs (times M)
    z = np.fft.ifft(merge_even(X, nz)) * (nz // 2)
    back = np.empty(nz)
    back[0::2], back[1::2] = z.real, z.imag
    assert np.allclose(back / (nz // 2) / 2, x, rtol=0, atol=1e-12)


def separate_pair(Z, n):
    """k_zpair_split, item by item, in place on a pair region of n+1 values."""
    nh = (n + 1) // 2
    R = np.zeros(n + 1, dtype=complex)
    R[:n] = Z

    def x1(u, v):
        return complex(0.5 * (u.real + v.real), 0.5 * (u.imag - v.imag))

    def x2(u, v):
        return complex(0.5 * (u.imag + v.imag), -0.5 * (u.real - v.real))

    written = set()
    for k in range((nh + 1) // 2):
        kp = nh - 1 - k
        reads = {k, n - k if k else 0, kp, n - kp}
        zk, zmk, zkp, zmkp = R[k], R[n - k if k else 0], R[kp], R[n - kp]
        out = {k: x1(zk, zmk), nh + k: x2(zk, zmk)}
        if kp != k:
            out[kp] = x1(zkp, zmkp)
            out[nh + kp] = x2(zkp, zmkp)
        # the item reads only places it writes itself, and no place twice
        assert reads <= set(out), (n, k)
        assert not (set(out) & written), (n, k)
        written |= set(out)
        for i, v in out.items():
            R[i] = v
    assert written == set(range(2 * nh))  # both spectrum rows, the spare slot n included
    return R[:nh], R[nh:2 * nh]


def merge_pair(X1, X2, n):
    """k_zpair_merge: two half spectra to the pair's complex spectrum."""
    nh = (n + 1) // 2
    Z = np.empty(n, dtype=complex)
    for k in range(n):
        if k < nh:
            u, v = X1[k], X2[k]
            if k == 0:
                u, v = complex(u.real, 0), complex(v.real, 0)
        else:
            u, v = np.conj(X1[n - k]), np.conj(X2[n - k])
        Z[k] = complex(u.real - v.imag, u.imag + v.real)
    return Z


@pytest.mark.parametrize("n", [1, 3, 5, 7, 9, 15, 101, 243, 1001])
def test_odd_pair_separation_and_merge(n):
    rng = np.random.default_rng(n)
    a, b = rng.standard_normal(n), rng.standard_normal(n)
    X1, X2 = separate_pair(np.fft.fft(a + 1j * b), n)
    assert np.allclose(X1, np.fft.rfft(a), rtol=0, atol=1e-12 * n)
    assert np.allclose(X2, np.fft.rfft(b), rtol=0, atol=1e-12 * n)
    z = np.fft.ifft(merge_pair(X1, X2, n))
    assert np.allclose(z.real, a, rtol=0, atol=1e-12) and np.allclose(z.imag, b, rtol=0, atol=1e-12)


@pytest.mark.parametrize("n", [1, 9, 243, 1001])
def test_odd_lone_row(n):
    """nx*ny odd: the last row runs alone, its pair's second row zero on the
    way in (k_zpair_pack with lone) and its second spectrum taken as zero on
    the way back (k_zpair_merge with lone): the separation then gives the
    row's spectrum and zero, and the merge's inverse gives the row back with a
    zero imaginary part."""
    rng = np.random.default_rng(n + 7)
    a = rng.standard_normal(n)
    X1, X2 = separate_pair(np.fft.fft(a + 0j), n)
    assert np.allclose(X1, np.fft.rfft(a), rtol=0, atol=1e-12 * n)
    assert np.allclose(X2, 0, rtol=0, atol=1e-12 * n)
    z = np.fft.ifft(merge_pair(X1, np.zeros_like(X1), n))
    assert np.allclose(z.real, a, rtol=0, atol=1e-12) and np.allclose(z.imag, 0, rtol=0, atol=1e-12)


def test_planted_slips_are_caught():
    """The checks above fail for a conjugate dropped from the odd separation
    and for a sign flipped in the even split's odd spectrum."""
    n = 9
    rng = np.random.default_rng(1)
    a, b = rng.standard_normal(n), rng.standard_normal(n)
    Z = np.fft.fft(a + 1j * b)
    nh = (n + 1) // 2
    # separation with conj dropped: X1 = (Z[k] + Z[n-k]) / 2
    bad = np.array([0.5 * (Z[k] + Z[(n - k) % n]) for k in range(nh)])
    assert not np.allclose(bad, np.fft.rfft(a), atol=1e-9)
    # even split with 1/(2i) replaced by 1/(-2i)
    nz = 10
    x = rng.standard_normal(nz)
    Zp = np.fft.fft(x[0::2] + 1j * x[1::2])
    M = nz // 2
    twp = np.exp(-2j * np.pi * np.arange(M) / nz)
    wrong = []
    for k in range(M):
        zk, zm = Zp[k], np.conj(Zp[(M - k) % M])
        wrong.append(0.5 * (zk + zm) - (zk - zm) / 2j * twp[k])
    assert not np.allclose(wrong, np.fft.rfft(x)[:M], atol=1e-9)
