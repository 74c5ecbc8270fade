/*
 * oracle.h -- CPU restatement of fishfry's periodic Poisson solve (FP64).
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product (fishfry_amd/, include/,
 * the C-ABI library) may include, link or call this code.  It is used by
 * tests/ (as the checker), by __graft_entry__.smoke() (as the checker) and by
 * bench.py's cpu_baseline leg (as the timed CPU baseline).
 *
 * Every routine cites the reference file:line whose behaviour it restates.
 * How each routine is pinned (so that a plausible mistake fails a check):
 *   ora_fft        brute-force long-double DFT (tests/test_oracle.py), scipy
 *                  long-double golden vectors, Parseval, impulse/single-tone
 *   ora_rfft/irfft scipy long-double rfft/irfft golden vectors, round trip,
 *                  Hermitian handling of DC/Nyquist imaginary parts
 *   ora_solve      scipy long-double golden solves (8^3, 12x10x9, 16x24x20,
 *                  32^3), single-mode closed form phi = -f/|k dd|^2, mean
 *                  removal, linearity, analytic-phi norms
 *   ora_init_field closed form evaluated in long double by numpy
 *   ora_diff       numpy long-double sums
 *   ora_dd         closed-form values for n in {32..1024} (1 ulp cases)
 *   ora_solve_sym  the stencil (numpy, np.roll) applied to its phi gives
 *                  f - mean f; errors against the analytic phi fall as h^2
 *                  (3-point) and h^4 (5-point); single-mode closed forms
 *   ora_stencil    numpy stencil on random fields
 *   ora_fd_factor  small-angle limits and the 3-point / 5-point relation
 */
#ifndef FISHFRY_ORACLE_H
#define FISHFRY_ORACLE_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Batched unnormalized 1D complex FFT, interleaved (re,im) doubles.
 * sign = -1 forward (exponent -1), +1 inverse; element (b,r) lives at
 * data[2*(b*dist + r*stride)].  Follows the cuFFT/hipFFT Z2Z convention used at
 * ParisPencil.cpp:78-79 (plans) and :212,259,288,334 (execs). */
int ora_fft(int n, double *data, long howmany, long stride, long dist, int sign);

/* Batched R2C along contiguous rows: in[b*n + k] -> out complex[b*(n/2+1) + kh].
 * ParisPencil.cpp:81 (D2Z plan), :165 (exec). */
int ora_rfft(int n, const double *in, double *out, long howmany);

/* Batched C2R (unnormalized) along contiguous rows: complex[b*(n/2+1)+kh] ->
 * out[b*n + k].  The imaginary parts of the DC term and (even n) the Nyquist
 * term are ignored, i.e. the input is read as a Hermitian half spectrum.
 * ParisPencil.cpp:80 (Z2D plan), :380 (exec). */
int ora_irfft(int n, const double *in, double *out, long howmany);

/* Filter constants dd = 2*pi*(n-1)/(n*(hi-lo)), ParisPencil.cpp:9-11. */
double ora_dd(int n, double lo, double hi);

/* Bytes the reference asks callers to allocate per work array,
 * ParisPencil.cpp:49-75 (general m[3]). */
size_t ora_bytes(const int n[3], const int m[3]);

/* The spectral filter alone, in place on a half spectrum [ni][nj][nk/2+1]
 * (complex interleaved): ParisPencil.cpp:263-285 at 1x1x1 tasks. */
void ora_filter(const int n[3], const double lo[3], const double hi[3], double *spec);

/* Full periodic Poisson solve at 1x1x1 tasks, ParisPencil.cpp:95-425:
 * D2Z(z), Z2Z fwd (y), Z2Z fwd (x), spectral filter (:263-285), Z2Z inv (x),
 * Z2Z inv (y), Z2D (z), scale by 1/(nx*ny*nz) (:406-421).
 * density [nx][ny][nz] (z fastest) is read, potential receives phi in the
 * same layout.  work must hold 2*nx*ny*(nz/2+1) doubles.  nthreads <= 0 uses
 * the OpenMP default. */
int ora_solve(const int n[3], const double lo[3], const double hi[3],
              const double *density, double *potential, double *work, int nthreads);

/* Finite-difference symbols (order 3: the 3-point Laplacian, 5: the 5-point
 * one, on every axis; any other order: the spectral filter above), named but
 * not implemented by the reference (make-frontier.sh:9-10, make-summit.sh:7-8):
 * for mode k of an n-point axis with spacing h = (hi - lo)/(n - 1) and
 * s = sin(pi k/n), the axis term is (4/h^2) s^2 (3-point) or
 * (4/h^2) s^2 (1 + s^2/3) (5-point); the filter is -1/(sum of the axis terms),
 * the (0,0,0) mode set to 0.  ora_fd_factor: 4 s^2 [(1 + s^2/3)].
 * ora_stencil applies the matching stencil in real space (periodic). */
double ora_fd_factor(long k, long n, int order);
void ora_filter_sym(const int n[3], const double lo[3], const double hi[3], int order, double *spec);
int ora_solve_sym(const int n[3], const double lo[3], const double hi[3], int order,
                  const double *density, double *potential, double *work, int nthreads);
void ora_stencil(const int n[3], const double lo[3], const double hi[3], int order, const double *phi, double *out);

/* The benchmark's right-hand side f = lap(phi), FishFry.hpp:14-45, at
 * x_i = xlo + i*dx (and likewise y, z).  Layout [ni][nj][nk]. */
void ora_init_field(int ni, int nj, int nk, double xlo, double ylo, double zlo,
                    double dx, double dy, double dz, double *f);

/* Error sums against the analytic phi, FishFry.hpp:57-101:
 * out[0] = sum|e|, out[1] = sum e^2, out[2] = max|e|, e = phi_exact - phi.
 * Accumulated in long double (the reference uses double atomics). */
void ora_diff(int ni, int nj, int nk, double xlo, double ylo, double zlo,
              double dx, double dy, double dz, const double *phi, double out[3]);

/* Twiddle exp(-2*pi*i*t/n) with exact integer reduction (quadrant and
 * octant symmetry), evaluated in long double and rounded. */
void ora_twiddle(long t, long n, double *c, double *s);

#ifdef __cplusplus
}
#endif
#endif
