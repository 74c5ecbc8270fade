/*
 * oracle.c -- CPU restatement of fishfry's periodic Poisson solve (FP64).
 *
 * TEST INFRASTRUCTURE ONLY: the checker for tests/, __graft_entry__.smoke()
 * and the timed CPU baseline of bench.py.  The product never links it.
 *
 * The FFT is a textbook mixed-radix Stockham autosort transform (radix 4, 2,
 * 3, 5 specialised, any other prime factor by a direct DFT of that length),
 * with twiddles from ora_twiddle().  The transforms are unnormalized with
 * exponent -1 forward and +1 inverse, the cuFFT/hipFFT convention that
 * ParisPencil.cpp relies on (forward at :165,212,259, inverse at :288,334,380,
 * single 1/N scale at :408).
 */
#define _GNU_SOURCE
#include "oracle.h"

#include <complex.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

typedef double _Complex cplx;

/* ------------------------------------------------------------------------ */
/* twiddles                                                                   */

void ora_twiddle(long t, long n, double *c, double *s)
{
  /* exp(-2 pi i t/n) = cos(2 pi t/n) - i sin(2 pi t/n).
   * 4t = q n + r, 0 <= r < n: angle = q*pi/2 + (pi/2)*(r/n).
   * For r/n > 1/2 use the complementary angle so the long-double argument
   * never exceeds pi/4 and symmetric points get identical values. */
  t %= n;
  if (t < 0) t += n;
  const long v = 4 * t;
  const long q = v / n;
  const long r = v - q * n;
  long double cb, sb;
  if (2 * r <= n) {
    const long double a = (M_PIl / 2.0L) * ((long double)r / (long double)n);
    cb = cosl(a);
    sb = sinl(a);
  } else {
    const long double a = (M_PIl / 2.0L) * ((long double)(n - r) / (long double)n);
    cb = sinl(a);
    sb = cosl(a);
  }
  long double cc, ss; /* cos, sin of the full angle */
  switch (q & 3) {
    case 0: cc = cb; ss = sb; break;
    case 1: cc = -sb; ss = cb; break;
    case 2: cc = -cb; ss = -sb; break;
    default: cc = sb; ss = -cb; break;
  }
  *c = (double)cc;
  *s = (double)(-ss);
}

/* ------------------------------------------------------------------------ */
/* 1D plan: factorization + twiddle table W[t] = exp(-2 pi i t / n)          */

typedef struct {
  int n;
  int nf;
  int fac[64];
  cplx *w; /* n entries */
} plan1d;

static int plan_init(plan1d *p, int n)
{
  if (n < 1) return -1;
  p->n = n;
  p->nf = 0;
  int m = n;
  while (m % 4 == 0) { p->fac[p->nf++] = 4; m /= 4; }
  while (m % 2 == 0) { p->fac[p->nf++] = 2; m /= 2; }
  for (int f = 3; m > 1; f += 2) {
    while (m % f == 0) { p->fac[p->nf++] = f; m /= f; }
    if ((long)f * f > m && m > 1) { p->fac[p->nf++] = m; m = 1; }
  }
  p->w = (cplx *)malloc(sizeof(cplx) * (size_t)n);
  if (!p->w) return -1;
  for (int t = 0; t < n; t++) {
    double c, s;
    ora_twiddle(t, n, &c, &s);
    p->w[t] = c + I * s;
  }
  return 0;
}

static void plan_free(plan1d *p) { free(p->w); p->w = NULL; }

/* W^e for the active direction; e is reduced modulo n. */
static inline cplx tw(const plan1d *p, long e, int sign)
{
  const cplx w = p->w[e % p->n];
  return sign < 0 ? w : conj(w);
}

/* One length-n transform of contiguous x (in place), scratch y (n entries). */
static void fft_contig(const plan1d *p, cplx *x, cplx *y, int sign)
{
  const int n = p->n;
  cplx *src = x, *dst = y;
  long ns = 1;
  cplx a[64], b[64];
  cplx *ga = NULL, *gb = NULL;
  for (int f = 0; f < p->nf; f++) {
    const int R = p->fac[f];
    const long nr = n / R;
    const long span = ns * R;           /* length of the sub-transforms after this pass */
    const long wstep = n / span;        /* W_span = W_n^(n/span) */
    cplx *va = a, *vb = b;
    if (R > 64) {
      ga = (cplx *)realloc(ga, sizeof(cplx) * R);
      gb = (cplx *)realloc(gb, sizeof(cplx) * R);
      va = ga; vb = gb;
    }
    for (long j = 0; j < nr; j++) {
      const long k = j % ns;
      for (int m = 0; m < R; m++) {
        cplx v = src[j + m * nr];
        if (k && m) v *= tw(p, (long)m * k * wstep, sign);
        va[m] = v;
      }
      /* DFT of length R */
      switch (R) {
        case 2:
          vb[0] = va[0] + va[1];
          vb[1] = va[0] - va[1];
          break;
        case 4: {
          const cplx t0 = va[0] + va[2], t1 = va[0] - va[2];
          const cplx t2 = va[1] + va[3], t3 = va[1] - va[3];
          /* forward: -i * t3 ; inverse: +i * t3 */
          const cplx rt3 = sign < 0 ? -I * t3 : I * t3;
          vb[0] = t0 + t2;
          vb[2] = t0 - t2;
          vb[1] = t1 + rt3;
          vb[3] = t1 - rt3;
          break;
        }
        default: {
          const long st = n / R; /* W_R = W_n^(n/R) */
          for (int q = 0; q < R; q++) {
            cplx acc = 0;
            for (int m = 0; m < R; m++) acc += va[m] * tw(p, (long)((q * (long)m) % R) * st, sign);
            vb[q] = acc;
          }
        }
      }
      const long base = (j / ns) * span + k;
      for (int q = 0; q < R; q++) dst[base + q * ns] = vb[q];
    }
    cplx *tmp = src; src = dst; dst = tmp;
    ns = span;
  }
  if (src != x) memcpy(x, src, sizeof(cplx) * (size_t)n);
  free(ga);
  free(gb);
}

/* Transform `howmany` lines of plan p: line b, element r at
 * d[b*dist + r*stride].  Lines are processed in blocks of up to CB
 * neighbouring lines (dist == 1 for the column passes) so strided gathers read
 * whole cache lines. */
#define CB 8
static int fft_lines(const plan1d *p, cplx *d, long howmany, long stride, long dist, int sign)
{
  const int n = p->n;
  int err = 0;
#pragma omp parallel
  {
    cplx *x = (cplx *)malloc(sizeof(cplx) * (size_t)n * CB);
    cplx *y = (cplx *)malloc(sizeof(cplx) * (size_t)n);
    if (!x || !y) {
#pragma omp atomic write
      err = -1;
    } else {
      const long nblk = (howmany + CB - 1) / CB;
#pragma omp for schedule(static)
      for (long blk = 0; blk < nblk; blk++) {
        const long b0 = blk * CB;
        const int nb = (int)(howmany - b0 < CB ? howmany - b0 : CB);
        for (int r = 0; r < n; r++)
          for (int c = 0; c < nb; c++) x[(long)c * n + r] = d[(b0 + c) * dist + (long)r * stride];
        for (int c = 0; c < nb; c++) fft_contig(p, x + (long)c * n, y, sign);
        for (int r = 0; r < n; r++)
          for (int c = 0; c < nb; c++) d[(b0 + c) * dist + (long)r * stride] = x[(long)c * n + r];
      }
    }
    free(x);
    free(y);
  }
  return err;
}

int ora_fft(int n, double *data, long howmany, long stride, long dist, int sign)
{
  if (n < 1 || howmany < 0) return -1;
  plan1d p;
  if (plan_init(&p, n)) return -1;
  const int err = fft_lines(&p, (cplx *)data, howmany, stride, dist, sign);
  plan_free(&p);
  return err;
}

/* ------------------------------------------------------------------------ */
/* real transforms along contiguous rows (the z axis of the reference)       */

int ora_rfft(int n, const double *in, double *out, long howmany)
{
  if (n < 1 || howmany < 0) return -1;
  const int nh = n / 2 + 1;
  plan1d p;
  if (plan_init(&p, n)) return -1;
#pragma omp parallel
  {
    cplx *x = (cplx *)malloc(sizeof(cplx) * (size_t)n);
    cplx *y = (cplx *)malloc(sizeof(cplx) * (size_t)n);
#pragma omp for schedule(static)
    for (long b = 0; b < howmany; b++) {
      for (int k = 0; k < n; k++) x[k] = in[b * n + k];
      fft_contig(&p, x, y, -1);
      cplx *o = (cplx *)out + b * nh;
      for (int k = 0; k < nh; k++) o[k] = x[k];
    }
    free(x);
    free(y);
  }
  plan_free(&p);
  return 0;
}

int ora_irfft(int n, const double *in, double *out, long howmany)
{
  if (n < 1 || howmany < 0) return -1;
  const int nh = n / 2 + 1;
  plan1d p;
  if (plan_init(&p, n)) return -1;
#pragma omp parallel
  {
    cplx *x = (cplx *)malloc(sizeof(cplx) * (size_t)n);
    cplx *y = (cplx *)malloc(sizeof(cplx) * (size_t)n);
#pragma omp for schedule(static)
    for (long b = 0; b < howmany; b++) {
      const cplx *h = (const cplx *)in + b * nh;
      /* Hermitian extension; Im of DC and of the even-n Nyquist term dropped */
      x[0] = creal(h[0]);
      for (int k = 1; k < nh; k++) x[k] = h[k];
      if (n % 2 == 0) x[n / 2] = creal(h[n / 2]);
      for (int k = nh; k < n; k++) x[k] = conj(h[n - k]);
      fft_contig(&p, x, y, +1);
      for (int k = 0; k < n; k++) out[b * n + k] = creal(x[k]);
    }
    free(x);
    free(y);
  }
  plan_free(&p);
  return 0;
}

/* ------------------------------------------------------------------------ */
/* constants                                                                  */

double ora_dd(int n, double lo, double hi)
{
  /* ParisPencil.cpp:9-11, same operand order */
  return 2.0 * M_PI * (double)(n - 1) / ((double)n * (hi - lo));
}

static long lmax(long a, long b) { return a > b ? a : b; }

size_t ora_bytes(const int n[3], const int m[3])
{
  /* ParisPencil.cpp:18,24-26,51-75 */
  const int nh = n[2] / 2 + 1;
  const int mi = m[0], mj = m[1], mk = m[2];
  int mq = (int)lround(sqrt((double)mk));
  while (mk % mq) mq--;
  const int mp = mk / mq;
  const long dh = (nh + mk - 1) / mk;
  const long di = (n[0] + mi - 1) / mi;
  const long dj = (n[1] + mj - 1) / mj;
  const long dk = (n[2] + mk - 1) / mk;
  const long dip = (di + mp - 1) / mp;
  const long djq = (dj + mq - 1) / mq;
  const long mjq = (long)mj * mq;
  const long dhq = (nh + mjq - 1) / mjq;
  const long mip = (long)mi * mp;
  const long djp = (n[1] + mip - 1) / mip;
  long nmax = di * dj * dk;
  nmax = lmax(nmax, (long)mp * mq * dip * djq * dk);
  nmax = lmax(nmax, 2L * dip * djq * mk * dh);
  nmax = lmax(nmax, 2L * dip * mp * djq * mq * dh);
  nmax = lmax(nmax, 2L * dip * djq * mjq * dhq);
  nmax = lmax(nmax, 2L * dip * dhq * mip * djp);
  nmax = lmax(nmax, 2L * djp * dhq * mip * dip);
  return (size_t)nmax * sizeof(double);
}

/* ------------------------------------------------------------------------ */
/* the solve                                                                  */

static inline double sqr(double x) { return x * x; }

/* Finite-difference symbols, per axis, for mode k of an n-point periodic
 * axis with spacing h = (hi - lo)/(n - 1) (the reference's hi is the last
 * cell, FishFry.hpp:133-134).  With s = sin(pi k / n):
 *   3-point  (u[j+1] - 2u[j] + u[j-1]) / h^2                     -> -(4/h^2) s^2
 *   5-point  (-u[j+2] + 16u[j+1] - 30u[j] + 16u[j-1] - u[j-2]) / (12 h^2)
 *                                                                 -> -(4/h^2) s^2 (1 + s^2/3)
 * (the eigenvalues of the stencils on exp(2 pi i k j / n)).  The reference's
 * build scripts name these symbols (PARIS_3PT / PARIS_5PT, make-frontier.sh:9-10,
 * make-summit.sh:7-8) but its source implements only the spectral one.
 * ora_fd_factor returns 4 s^2 [(1 + s^2/3)], in long double from sinl of
 * pi min(k, n - k) / n, rounded to double. */
double ora_fd_factor(long k, long n, int order)
{
  long kk = k % n;
  if (kk < 0) kk += n;
  if (n - kk < kk) kk = n - kk;
  const long double sv = sinl(3.14159265358979323846264338327950288L * (long double)kk / (long double)n);
  const long double s2 = sv * sv;
  const long double g = order == 5 ? 4.0L * s2 * (1.0L + s2 / 3.0L) : 4.0L * s2;
  return (double)g;
}

/* 1/h^2 of an axis, h = (hi - lo)/(n - 1). */
static double fd_scale(int n, double lo, double hi)
{
  const double ih = (double)(n - 1) / (hi - lo);
  return ih * ih;
}

void ora_filter_sym(const int n[3], const double lo[3], const double hi[3], int order, double *spec)
{
  if (order != 3 && order != 5) {
    ora_filter(n, lo, hi, spec);
    return;
  }
  const int ni = n[0], nj = n[1], nk = n[2];
  const int nh = nk / 2 + 1;
  const double ci = fd_scale(ni, lo[0], hi[0]);
  const double cj = fd_scale(nj, lo[1], hi[1]);
  const double ck = fd_scale(nk, lo[2], hi[2]);
  cplx *w = (cplx *)spec;
#pragma omp parallel for schedule(static)
  for (int i = 0; i < ni; i++) {
    const double ti = ora_fd_factor(i, ni, order) * ci;
    for (int j = 0; j < nj; j++) {
      const double tj = ora_fd_factor(j, nj, order) * cj;
      for (int k = 0; k < nh; k++) {
        const long iab = ((long)i * nj + j) * nh + k;
        if (i || j || k) {
          const double tk = ora_fd_factor(k, nk, order) * ck;
          const double d = -1.0 / ((ti + tj) + tk);
          w[iab] = d * creal(w[iab]) + I * (d * cimag(w[iab]));
        } else {
          w[iab] = 0;
        }
      }
    }
  }
}

/* L_h phi with the 3- or 5-point stencil on every axis, periodic, in real
 * space: the operator whose symbol ora_filter_sym inverts. */
void ora_stencil(const int n[3], const double lo[3], const double hi[3], int order, const double *phi, double *out)
{
  const int ni = n[0], nj = n[1], nk = n[2];
  const double c[3] = {fd_scale(ni, lo[0], hi[0]), fd_scale(nj, lo[1], hi[1]), fd_scale(nk, lo[2], hi[2])};
#pragma omp parallel for schedule(static)
  for (int i = 0; i < ni; i++) {
    for (int j = 0; j < nj; j++) {
      for (int k = 0; k < nk; k++) {
        const int idx[3] = {i, j, k}, ext[3] = {ni, nj, nk};
        const double center = phi[((long)i * nj + j) * nk + k];
        double lap = 0.0;
        for (int a = 0; a < 3; a++) {
          double v[5];
          for (int o = -2; o <= 2; o++) {
            int q[3] = {idx[0], idx[1], idx[2]};
            q[a] = ((q[a] + o) % ext[a] + ext[a]) % ext[a];
            v[o + 2] = phi[((long)q[0] * nj + q[1]) * nk + q[2]];
          }
          double d2;
          if (order == 5) d2 = (16.0 * (v[3] + v[1]) - (v[4] + v[0]) - 30.0 * center) / 12.0;
          else d2 = (v[3] + v[1]) - 2.0 * center;
          lap += c[a] * d2;
        }
        out[((long)i * nj + j) * nk + k] = lap;
      }
    }
  }
}

void ora_filter(const int n[3], const double lo[3], const double hi[3], double *spec)
{
  const int ni = n[0], nj = n[1], nk = n[2];
  const int nh = nk / 2 + 1;
  const double ddi = ora_dd(ni, lo[0], hi[0]);
  const double ddj = ora_dd(nj, lo[1], hi[1]);
  const double ddk = ora_dd(nk, lo[2], hi[2]);
  cplx *w = (cplx *)spec;
  /* spectral filter  (ParisPencil.cpp:263-285), same expression order */
#pragma omp parallel for schedule(static)
  for (int i = 0; i < ni; i++) {
    for (int j = 0; j < nj; j++) {
      for (int k = 0; k < nh; k++) {
        const long iab = ((long)i * nj + j) * nh + k;
        if (i || j || k) {
          const double i2 = sqr((double)(i < ni - i ? i : ni - i) * ddi);
          const double j2 = sqr((double)(j < nj - j ? j : nj - j) * ddj);
          const double k2 = sqr((double)k * ddk);
          const double d = -1.0 / (i2 + j2 + k2);
          w[iab] = d * creal(w[iab]) + I * (d * cimag(w[iab]));
        } else {
          w[iab] = 0;
        }
      }
    }
  }

}

int ora_solve(const int n[3], const double lo[3], const double hi[3],
              const double *density, double *potential, double *work, int nthreads)
{
  return ora_solve_sym(n, lo, hi, 0, density, potential, work, nthreads);
}

int ora_solve_sym(const int n[3], const double lo[3], const double hi[3], int order,
                  const double *density, double *potential, double *work, int nthreads)
{
  const int ni = n[0], nj = n[1], nk = n[2];
  if (ni < 1 || nj < 1 || nk < 1) return -1;
  const int nh = nk / 2 + 1;
#ifdef _OPENMP
  if (nthreads > 0) omp_set_num_threads(nthreads);
#else
  (void)nthreads;
#endif

  plan1d pi_, pj_;
  if (plan_init(&pi_, ni)) return -1;
  if (plan_init(&pj_, nj)) { plan_free(&pi_); return -1; }
  cplx *w = (cplx *)work;
  int err = 0;

  /* D2Z along z: [i][j][k] -> [i][j][kh]  (ParisPencil.cpp:165) */
  err |= ora_rfft(nk, density, work, (long)ni * nj);
  /* Z2Z forward along y: column stride nh  (ParisPencil.cpp:212) */
  for (int i = 0; i < ni && !err; i++) err |= fft_lines(&pj_, w + (long)i * nj * nh, nh, nh, 1, -1);
  /* Z2Z forward along x: column stride nj*nh  (ParisPencil.cpp:259) */
  if (!err) err |= fft_lines(&pi_, w, (long)nj * nh, (long)nj * nh, 1, -1);
  if (err) { plan_free(&pi_); plan_free(&pj_); return -1; }

  ora_filter_sym(n, lo, hi, order, work);

  /* Z2Z inverse along x, then y  (ParisPencil.cpp:288,334) */
  err |= fft_lines(&pi_, w, (long)nj * nh, (long)nj * nh, 1, +1);
  for (int i = 0; i < ni && !err; i++) err |= fft_lines(&pj_, w + (long)i * nj * nh, nh, nh, 1, +1);
  plan_free(&pi_);
  plan_free(&pj_);
  if (err) return -1;
  /* Z2D along z  (ParisPencil.cpp:380) */
  if (ora_irfft(nk, work, potential, (long)ni * nj)) return -1;
  /* normalize  (ParisPencil.cpp:408,419) */
  const double divN = 1.0 / ((double)ni * (double)nj * (double)nk);
  const long N = (long)ni * nj * nk;
#pragma omp parallel for schedule(static)
  for (long t = 0; t < N; t++) potential[t] = divN * potential[t];
  return 0;
}

/* ------------------------------------------------------------------------ */
/* driver kernels                                                             */

void ora_init_field(int ni, int nj, int nk, double xlo, double ylo, double zlo,
                    double dx, double dy, double dz, double *f)
{
  /* FishFry.hpp:14-45 */
#pragma omp parallel for schedule(static)
  for (int i = 0; i < ni; i++) {
    const double x = xlo + i * dx;
    const double cx = cos(x), sx = sin(x), ex = exp(sx);
    const double px = cx * ex;
    const double dpx = -ex * sx * cx * (sx + 3.0);
    for (int j = 0; j < nj; j++) {
      const double y = ylo + j * dy;
      const double cy = cos(y), sy = sin(y), ey = exp(cy);
      const double py = sy * ey;
      const double dpy = sy * ey * (sy * sy - 3.0 * cy - 1.0);
      for (int k = 0; k < nk; k++) {
        const double z = zlo + k * dz;
        const double cz = cos(z), sz = sin(z), ez = exp(cz);
        const double pz = cz * ez;
        const double dpz = ez * (3.0 * sz * sz - cz * cz * cz - 1.0);
        f[((long)i * nj + j) * nk + k] = px * py * dpz + px * dpy * pz + dpx * py * pz;
      }
    }
  }
}

void ora_diff(int ni, int nj, int nk, double xlo, double ylo, double zlo,
              double dx, double dy, double dz, const double *phi, double out[3])
{
  /* FishFry.hpp:57-101 */
  long double s1 = 0, s2 = 0;
  double mx = 0;
  for (int i = 0; i < ni; i++) {
    const double x = xlo + i * dx;
    const double px = cos(x) * exp(sin(x));
    for (int j = 0; j < nj; j++) {
      const double y = ylo + j * dy;
      const double py = sin(y) * exp(cos(y));
      for (int k = 0; k < nk; k++) {
        const double z = zlo + k * dz;
        const double cz = cos(z);
        const double pz = cz * exp(cz);
        const double e = fabs(px * py * pz - phi[((long)i * nj + j) * nk + k]);
        s1 += e;
        s2 += (long double)e * e;
        if (e > mx) mx = e;
      }
    }
  }
  out[0] = (double)s1;
  out[1] = (double)s2;
  out[2] = mx;
}
