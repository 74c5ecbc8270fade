// Device FFT engine for the Poisson passes (gfx950, wave64).
//
// A "line" is one 1D transform of length n.  Each line is owned by L = n/E
// threads; thread r0 of a line holds the E elements r0 + i*L (i = 0..E-1) in
// registers.  A transform is a sequence of Stockham autosort passes with radix
// R (every R divides E, the product of the radices is n).  In every pass a
// thread performs E/R butterflies j = r0 + q*L, q = 0..E/R-1, whose inputs are
// the elements j + m*n/R = r0 + (q + m*E/R)*L -- registers q + m*E/R -- so the
// input side of a pass never needs data from another thread.  The outputs go
// to their autosorted positions ((j - j%ns)*R + j%ns + m*ns) in LDS, and the
// next pass reads them back in natural order.  After the last pass the
// registers hold the result in natural order again, so a kernel can chain a
// forward transform, a pointwise multiply and an inverse transform without an
// extra exchange.
//
// Twiddles come from a per-length table W[t] = exp(-2*pi*i*t/n), t < n,
// generated on the host in long double with exact integer argument reduction.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "ff_consts.hpp"

#define FF_MAXP 16
#define FF_MAX_THREADS 1024

namespace ff {
// Workgroup-size bound per elements-per-thread, chosen so the register file
// holds the kernel without spills: 1024 threads leave 128 VGPRs per lane
// (E <= 8), 512 leave 256 (E <= 16), 256 leave 512 (E <= 30).
__host__ __device__ constexpr int max_threads_for(int E) { return E <= 8 ? 1024 : (E <= 16 ? 512 : 256); }
}  // namespace ff

namespace ff {

template <typename F>
struct alignas(2 * sizeof(F)) cx {
  F x, y;
};

template <typename F>
__device__ __forceinline__ cx<F> operator+(cx<F> a, cx<F> b) { return {a.x + b.x, a.y + b.y}; }
template <typename F>
__device__ __forceinline__ cx<F> operator-(cx<F> a, cx<F> b) { return {a.x - b.x, a.y - b.y}; }
template <typename F>
__device__ __forceinline__ cx<F> scale(cx<F> a, F s) { return {a.x * s, a.y * s}; }
template <typename F>
__device__ __forceinline__ cx<F> conj(cx<F> a) { return {a.x, -a.y}; }

// Fused multiply-add with one rounding.  The pass kernels are compiled with
// floating-point contraction off, and every multiply-add of the engine is
// written out with fmadd: each kernel variant (runtime or compiled-in plan,
// any tile width, one solve or many instances per launch) then performs the
// same rounded operations in the same order, so their results are
// bit-identical.
__device__ __forceinline__ double fmadd(double a, double b, double c) { return __builtin_fma(a, b, c); }
__device__ __forceinline__ float fmadd(float a, float b, float c) { return __builtin_fmaf(a, b, c); }

// a * w for the forward direction (DIR < 0), a * conj(w) for the inverse.
template <int DIR, typename F>
__device__ __forceinline__ cx<F> twmul(cx<F> a, cx<F> w) {
  if (DIR < 0) return {fmadd(a.x, w.x, -(a.y * w.y)), fmadd(a.x, w.y, a.y * w.x)};
  return {fmadd(a.x, w.x, a.y * w.y), fmadd(a.y, w.x, -(a.x * w.y))};
}

// a * (-i) forward, a * (+i) inverse.
template <int DIR, typename F>
__device__ __forceinline__ cx<F> mul_mi(cx<F> a) {
  if (DIR < 0) return {a.y, -a.x};
  return {-a.y, a.x};
}

// a * exp(DIR * 2*pi*i*t/R) with t, R compile-time.
template <int R, int t_, int DIR, typename F>
__device__ __forceinline__ cx<F> rot(cx<F> a) {
  constexpr int t = ((t_ % R) + R) % R;
  if constexpr (t == 0) {
    return a;
  } else if constexpr ((4 * t) % R == 0) {
    constexpr int q = (4 * t) / R;
    if constexpr (q == 2) return {-a.x, -a.y};
    else if constexpr (q == 1) return mul_mi<DIR>(a);
    else return mul_mi<-DIR>(a);
  } else if constexpr ((8 * t) % R == 0) {
    constexpr F h = F(0.7071067811865475244008444);
    constexpr int o = (8 * t) / R;  // odd multiple of pi/4
    // exp(DIR*i*o*pi/4) = (cx, DIR*sx)
    constexpr F cxv = (o == 1 || o == 7) ? h : -h;
    constexpr F sxv = (o == 1 || o == 3) ? h : -h;
    constexpr F s = DIR < 0 ? -sxv : sxv;
    return {fmadd(cxv, a.x, -(s * a.y)), fmadd(cxv, a.y, s * a.x)};
  } else {
    constexpr F c = F(kcos(R, t));
    constexpr F s0 = F(ksin(R, t));
    constexpr F s = DIR < 0 ? -s0 : s0;
    return {fmadd(c, a.x, -(s * a.y)), fmadd(c, a.y, s * a.x)};
  }
}

// ---------------------------------------------------------------------------
// In-register DFTs of length R, in place on a[0..R-1].

template <int R, int DIR, typename F>
struct Dft;

template <int DIR, typename F>
struct Dft<1, DIR, F> {
  static __device__ __forceinline__ void run(cx<F>*) {}
};

template <int DIR, typename F>
struct Dft<2, DIR, F> {
  static __device__ __forceinline__ void run(cx<F>* a) {
    const cx<F> t = a[0];
    a[0] = t + a[1];
    a[1] = t - a[1];
  }
};

template <int DIR, typename F>
struct Dft<3, DIR, F> {
  static __device__ __forceinline__ void run(cx<F>* a) {
    constexpr F h = F(0.8660254037844386467637232);
    const cx<F> t1 = a[1] + a[2];
    const cx<F> t2 = {fmadd(F(-0.5), t1.x, a[0].x), fmadd(F(-0.5), t1.y, a[0].y)};
    const cx<F> t3 = mul_mi<DIR>(scale(a[1] - a[2], h));
    a[0] = a[0] + t1;
    a[1] = t2 + t3;
    a[2] = t2 - t3;
  }
};

template <int DIR, typename F>
struct Dft<4, DIR, F> {
  static __device__ __forceinline__ void run(cx<F>* a) {
    const cx<F> t0 = a[0] + a[2], t1 = a[0] - a[2];
    const cx<F> t2 = a[1] + a[3], t3 = mul_mi<DIR>(a[1] - a[3]);
    a[0] = t0 + t2;
    a[2] = t0 - t2;
    a[1] = t1 + t3;
    a[3] = t1 - t3;
  }
};

template <int DIR, typename F>
struct Dft<5, DIR, F> {
  static __device__ __forceinline__ void run(cx<F>* a) {
    constexpr F c1 = F(0.3090169943749474241022934);  // cos(2pi/5)
    constexpr F c2 = F(0.8090169943749474241022934);  // -cos(4pi/5)
    constexpr F s1 = F(0.9510565162951535721164393);  // sin(2pi/5)
    constexpr F s2 = F(0.5877852522924731291687060);  // sin(4pi/5)
    const cx<F> t1 = a[1] + a[4], t2 = a[2] + a[3];
    const cx<F> t3 = a[1] - a[4], t4 = a[2] - a[3];
    const cx<F> a0 = a[0];
    const cx<F> u1 = {fmadd(-c2, t2.x, fmadd(c1, t1.x, a0.x)), fmadd(-c2, t2.y, fmadd(c1, t1.y, a0.y))};
    const cx<F> u2 = {fmadd(c1, t2.x, fmadd(-c2, t1.x, a0.x)), fmadd(c1, t2.y, fmadd(-c2, t1.y, a0.y))};
    const cx<F> v1 = mul_mi<DIR>(cx<F>{fmadd(s1, t3.x, s2 * t4.x), fmadd(s1, t3.y, s2 * t4.y)});
    const cx<F> v2 = mul_mi<DIR>(cx<F>{fmadd(s2, t3.x, -(s1 * t4.x)), fmadd(s2, t3.y, -(s1 * t4.y))});
    a[0] = a0 + t1 + t2;
    a[1] = u1 + v1;
    a[4] = u1 - v1;
    a[2] = u2 + v2;
    a[3] = u2 - v2;
  }
};

template <int R, int R1, int IDX, int DIR, typename F>
__device__ __forceinline__ void apply_rot(cx<F>* c) {
  if constexpr (IDX < R) {
    c[IDX] = rot<R, (IDX / R1) * (IDX % R1), DIR>(c[IDX]);
    apply_rot<R, R1, IDX + 1, DIR, F>(c);
  }
}

// Cooley-Tukey split R = R1*R2 inside registers: input m = R2*p + q,
// output k = k1 + R1*k2.
template <int R1, int R2, int DIR, typename F>
__device__ __forceinline__ void dft_ct(cx<F>* a) {
  constexpr int R = R1 * R2;
  cx<F> c[R];
#pragma unroll
  for (int q = 0; q < R2; ++q) {
    cx<F> s[R1];
#pragma unroll
    for (int p = 0; p < R1; ++p) s[p] = a[R2 * p + q];
    Dft<R1, DIR, F>::run(s);
#pragma unroll
    for (int k1 = 0; k1 < R1; ++k1) c[q * R1 + k1] = s[k1];
  }
  // twiddles exp(DIR*2*pi*i*q*k1/R) on c[q*R1 + k1]
  apply_rot<R, R1, 0, DIR, F>(c);
#pragma unroll
  for (int k1 = 0; k1 < R1; ++k1) {
    cx<F> s[R2];
#pragma unroll
    for (int q = 0; q < R2; ++q) s[q] = c[q * R1 + k1];
    Dft<R2, DIR, F>::run(s);
#pragma unroll
    for (int k2 = 0; k2 < R2; ++k2) a[k1 + R1 * k2] = s[k2];
  }
}

template <int DIR, typename F>
struct Dft<6, DIR, F> {
  static __device__ __forceinline__ void run(cx<F>* a) { dft_ct<3, 2, DIR, F>(a); }
};
template <int DIR, typename F>
struct Dft<8, DIR, F> {
  static __device__ __forceinline__ void run(cx<F>* a) { dft_ct<4, 2, DIR, F>(a); }
};
template <int DIR, typename F>
struct Dft<10, DIR, F> {
  static __device__ __forceinline__ void run(cx<F>* a) { dft_ct<5, 2, DIR, F>(a); }
};
template <int DIR, typename F>
struct Dft<12, DIR, F> {
  static __device__ __forceinline__ void run(cx<F>* a) { dft_ct<4, 3, DIR, F>(a); }
};
template <int DIR, typename F>
struct Dft<16, DIR, F> {
  static __device__ __forceinline__ void run(cx<F>* a) { dft_ct<4, 4, DIR, F>(a); }
};

// ---------------------------------------------------------------------------
// Plan of one line transform (host-built, passed by value in kernel args).

struct LinePlan {
  int n;                    // transform length
  int L;                    // threads per line = n / E
  int np;                   // number of passes
  int radix[FF_MAXP];       // radix of each pass
  int ns[FF_MAXP];          // product of the radices before the pass
  uint32_t nsinv[FF_MAXP];  // ceil(2^32 / ns) for ns > 1 (exact j/ns for j, ns < 2^16)
  int twstep[FF_MAXP];      // n / (ns * radix): twiddle exponent step
};

__device__ __forceinline__ int fast_mod(int j, int ns, uint32_t inv) {
  const int qt = (int)__umulhi((uint32_t)j, inv);
  return j - qt * ns;
}

// Radices supported by a given elements-per-thread E.
template <int E, int R>
constexpr bool radix_ok() {
  return R <= E && E % R == 0 && (R == 2 || R == 3 || R == 4 || R == 5 || R == 6 || R == 8 ||
                                  R == 10 || R == 12 || R == 16);
}

// Twiddle sources for the passes: tw_at(tw, t) = W[t].
// A table in global memory:
template <typename F>
__device__ __forceinline__ cx<F> tw_at(const cx<F>* __restrict__ tw, int t) {
  return tw[t];
}
// The first half of a table of even length 2H, held in LDS by the kernel:
// W[t + H] = -W[t] exactly in the host's tables (a half turn adds two quarter
// turns to the same reduced angle, so only the signs differ), so the negated
// entry is bit-identical to the stored one.  Entry t is stored at
// tw_swizzle(t): a permutation within blocks of 16 entries that spreads the
// strided indices of a wave's lanes (t = m*k*twstep) over the LDS banks.
__device__ __forceinline__ int tw_swizzle(int t) { return t ^ (((t >> 4) ^ (t >> 8)) & 15); }
template <typename F, int H>
struct TwHalf {
  const cx<F>* p;
};
template <typename T>
struct is_tw_half {
  static constexpr bool value = false;
};
template <typename F, int H>
struct is_tw_half<TwHalf<F, H>> {
  static constexpr bool value = true;
};
template <typename F, int H>
__device__ __forceinline__ cx<F> tw_at(const TwHalf<F, H>& tw, int t) {
  const bool hi = t >= H;
  const cx<F> w = tw.p[tw_swizzle(hi ? t - H : t)];
  return hi ? cx<F>{-w.x, -w.y} : w;
}

// Row tiles: line-major with a padded pitch.
struct RowIdx {
  int base, ps;
  __device__ __forceinline__ int operator()(int r) const { return base + r + (r >> ps); }
};
// A row held whole by the lanes of one wavefront (its length / E divides 64):
// the exchanges between its passes touch only that wavefront's lanes and LDS
// row, so they need a wavefront-level barrier, not a workgroup one, and the
// workgroup's rows proceed independently.
struct RowIdxW {
  int base, ps;
  __device__ __forceinline__ int operator()(int r) const { return base + r + (r >> ps); }
};
// The barrier around an exchange through LDS laid out by Idx.
template <typename Idx>
__device__ __forceinline__ void exchange_sync() {
  __syncthreads();
}
template <>
__device__ __forceinline__ void exchange_sync<RowIdxW>() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Butterflies of one pass on registers (twiddles applied to the inputs).
template <int E, int R, int DIR, typename F, typename TW>
__device__ __forceinline__ void pass_regs(cx<F> (&v)[E], int r0, int L, int ns, uint32_t nsinv, int twstep,
                                          const TW& tw) {
  constexpr int Q = E / R;
#pragma unroll
  for (int q = 0; q < Q; ++q) {
    cx<F> a[R];
#pragma unroll
    for (int m = 0; m < R; ++m) a[m] = v[q + m * Q];
    if (ns > 1) {
      const int j = r0 + q * L;
      const int e = fast_mod(j, ns, nsinv) * twstep;
#pragma unroll
      for (int m = 1; m < R; ++m) a[m] = twmul<DIR>(a[m], tw_at(tw, m * e));
    }
    Dft<R, DIR, F>::run(a);
#pragma unroll
    for (int m = 0; m < R; ++m) v[q + m * Q] = a[m];
  }
}

// Outputs of a pass to their autosorted positions in LDS.
template <int E, int R, typename F, typename Idx>
__device__ __forceinline__ void scatter(const cx<F> (&v)[E], cx<F>* lds, const Idx& idx, int r0, int L, int ns,
                                        uint32_t nsinv) {
  constexpr int Q = E / R;
#pragma unroll
  for (int q = 0; q < Q; ++q) {
    const int j = r0 + q * L;
    const int k = ns > 1 ? fast_mod(j, ns, nsinv) : 0;
    const int base = (j - k) * R + k;
#pragma unroll
    for (int m = 0; m < R; ++m) lds[idx(base + m * ns)] = v[q + m * Q];
  }
}

template <int E, typename F, typename Idx>
__device__ __forceinline__ void gather(cx<F> (&v)[E], const cx<F>* lds, const Idx& idx, int r0, int L) {
#pragma unroll
  for (int i = 0; i < E; ++i) v[i] = lds[idx(r0 + i * L)];
}

template <int E, typename F, typename Idx>
__device__ __forceinline__ void put_natural(const cx<F> (&v)[E], cx<F>* lds, const Idx& idx, int r0, int L) {
#pragma unroll
  for (int i = 0; i < E; ++i) lds[idx(r0 + i * L)] = v[i];
}

// Split exchange: one component (PART 0 real, 1 imaginary) at a time through
// an LDS array of F, half the LDS of a complex exchange.  After the real
// gather v[].x is in the new order while v[].y still holds the old order,
// which the imaginary scatter then reads.
template <int E, int R, int PART, typename F, typename Idx>
__device__ __forceinline__ void scatter_part(const cx<F> (&v)[E], F* lds, const Idx& idx, int r0, int L, int ns,
                                             uint32_t nsinv) {
  constexpr int Q = E / R;
#pragma unroll
  for (int q = 0; q < Q; ++q) {
    const int j = r0 + q * L;
    const int k = ns > 1 ? fast_mod(j, ns, nsinv) : 0;
    const int base = (j - k) * R + k;
#pragma unroll
    for (int m = 0; m < R; ++m) lds[idx(base + m * ns)] = PART ? v[q + m * Q].y : v[q + m * Q].x;
  }
}

template <int E, int PART, typename F, typename Idx>
__device__ __forceinline__ void gather_part(cx<F> (&v)[E], const F* lds, const Idx& idx, int r0, int L) {
#pragma unroll
  for (int i = 0; i < E; ++i) {
    if constexpr (PART) v[i].y = lds[idx(r0 + i * L)];
    else v[i].x = lds[idx(r0 + i * L)];
  }
}

// One pass with compile-time radix R: butterflies, then (unless last) the LDS
// exchange to natural order for the next pass.
template <int E, int R, int DIR, typename F, bool SPLIT, typename Idx, typename TW>
__device__ __forceinline__ void run_pass(cx<F> (&v)[E], const LinePlan& lp, int p, const TW& tw,
                                         cx<F>* lds, const Idx& idx, int r0, bool last) {
  if constexpr (radix_ok<E, R>()) {
    pass_regs<E, R, DIR, F>(v, r0, lp.L, lp.ns[p], lp.nsinv[p], lp.twstep[p], tw);
    if (!last) {
      if constexpr (!SPLIT) {
        exchange_sync<Idx>();
        scatter<E, R, F>(v, lds, idx, r0, lp.L, lp.ns[p], lp.nsinv[p]);
        exchange_sync<Idx>();
        gather<E, F>(v, lds, idx, r0, lp.L);
      } else {
        F* l = reinterpret_cast<F*>(lds);
        exchange_sync<Idx>();
        scatter_part<E, R, 0, F>(v, l, idx, r0, lp.L, lp.ns[p], lp.nsinv[p]);
        exchange_sync<Idx>();
        gather_part<E, 0, F>(v, l, idx, r0, lp.L);
        exchange_sync<Idx>();
        scatter_part<E, R, 1, F>(v, l, idx, r0, lp.L, lp.ns[p], lp.nsinv[p]);
        exchange_sync<Idx>();
        gather_part<E, 1, F>(v, l, idx, r0, lp.L);
      }
    }
  }
}

// Full line transform.  All threads of the workgroup must call it (it contains
// workgroup barriers); threads of padding lines simply transform garbage.
// SPLIT: exchange real and imaginary parts separately (LDS holds one
// component per slot).
template <int E, int DIR, typename F, typename Idx, bool SPLIT = false>
__device__ __forceinline__ void fft_line(cx<F> (&v)[E], const LinePlan& lp, const cx<F>* __restrict__ tw,
                                         cx<F>* lds, const Idx& idx, int r0) {
  for (int p = 0; p < lp.np; ++p) {
    const bool last = (p + 1 == lp.np);
    switch (lp.radix[p]) {
      case 2: run_pass<E, 2, DIR, F, SPLIT>(v, lp, p, tw, lds, idx, r0, last); break;
      case 3: run_pass<E, 3, DIR, F, SPLIT>(v, lp, p, tw, lds, idx, r0, last); break;
      case 4: run_pass<E, 4, DIR, F, SPLIT>(v, lp, p, tw, lds, idx, r0, last); break;
      case 5: run_pass<E, 5, DIR, F, SPLIT>(v, lp, p, tw, lds, idx, r0, last); break;
      case 6: run_pass<E, 6, DIR, F, SPLIT>(v, lp, p, tw, lds, idx, r0, last); break;
      case 8: run_pass<E, 8, DIR, F, SPLIT>(v, lp, p, tw, lds, idx, r0, last); break;
      case 10: run_pass<E, 10, DIR, F, SPLIT>(v, lp, p, tw, lds, idx, r0, last); break;
      case 12: run_pass<E, 12, DIR, F, SPLIT>(v, lp, p, tw, lds, idx, r0, last); break;
      case 16: run_pass<E, 16, DIR, F, SPLIT>(v, lp, p, tw, lds, idx, r0, last); break;
      default: break;
    }
  }
}

// ---------------------------------------------------------------------------
// Compile-time line plans.  For the lengths that dominate the solve, kernels
// are also instantiated with the plan fixed at compile time, so every LDS
// index, modulus and twiddle stride folds to a constant and the pass loop
// unrolls.  The plan is the one the host planner builds for (n, E): fewest
// passes with radices dividing E (first radix of 16, 12, 10, 8, 6, 5, 4, 3, 2
// that reaches the minimum), largest radix first, so the arithmetic is that
// of the generic kernel.

constexpr int kPlanRadices[9] = {16, 12, 10, 8, 6, 5, 4, 3, 2};

constexpr int static_min_passes(int m, int E) {
  if (m == 1) return 0;
  int best = 1 << 20;
  for (int t = 0; t < 9; ++t) {
    const int R = kPlanRadices[t];
    if (R > E || E % R || m % R) continue;
    const int p = static_min_passes(m / R, E);
    if (p + 1 < best) best = p + 1;
  }
  return best;
}

constexpr int static_best_radix(int m, int E) {
  int best = 1 << 20, bestR = 0;
  for (int t = 0; t < 9; ++t) {
    const int R = kPlanRadices[t];
    if (R > E || E % R || m % R) continue;
    const int p = static_min_passes(m / R, E);
    if (p + 1 < best) {
      best = p + 1;
      bestR = R;
    }
  }
  return bestR;
}

template <int N, int E>
constexpr LinePlan static_plan() {
  LinePlan lp{};
  int seq[FF_MAXP] = {};
  int np = 0;
  for (int m = N; m > 1 && np < FF_MAXP;) {
    const int R = static_best_radix(m, E);
    if (!R) break;
    seq[np++] = R;
    m /= R;
  }
  for (int a = 1; a < np; ++a)  // descending
    for (int b = a; b > 0 && seq[b] > seq[b - 1]; --b) {
      const int t = seq[b];
      seq[b] = seq[b - 1];
      seq[b - 1] = t;
    }
  lp.n = N;
  lp.L = N / E;
  lp.np = np;
  int ns = 1;
  for (int p = 0; p < np; ++p) {
    lp.radix[p] = seq[p];
    lp.ns[p] = ns;
    lp.nsinv[p] = ns > 1 ? (uint32_t)(((1ull << 32) + (uint64_t)ns - 1) / (uint64_t)ns) : 0u;
    lp.twstep[p] = N / (ns * seq[p]);
    ns *= seq[p];
  }
  return lp;
}

// fft_line with the plan of (N, E) fixed at compile time; TW: a twiddle
// source (tw_at).
template <int N, int E, int DIR, typename F, typename Idx, bool SPLIT, int P = 0, typename TW>
__device__ __forceinline__ void fft_line_static(cx<F> (&v)[E], const TW& tw, cx<F>* lds, const Idx& idx, int r0) {
  constexpr LinePlan lp = static_plan<N, E>();
  static_assert(lp.np > 0 && N % E == 0, "no compile-time plan for this length");
  if constexpr (P < lp.np) {
    run_pass<E, lp.radix[P], DIR, F, SPLIT>(v, lp, P, tw, lds, idx, r0, P + 1 == lp.np);
    fft_line_static<N, E, DIR, F, Idx, SPLIT, P + 1>(v, tw, lds, idx, r0);
  }
}

// LDS layouts.  Column tiles: element (line c, position r) at r*T + c plus one
// pad slot every 2^PS positions (row tiles: RowIdx above).
// The column pad shift of every column kernel:
#ifndef FF_COL_PS
#define FF_COL_PS 4
#endif
constexpr int kColPadShift = FF_COL_PS;
struct ColIdx {
  int T, c, ps;
  __device__ __forceinline__ int operator()(int r) const { return r * T + c + (r >> ps); }
};

}  // namespace ff
