// Row-local steps of the z passes whose packed rows run as four-step
// sub-line passes (even nz, M = nz/2 too long for one workgroup's line
// transform): the real-to-complex split after the forward M-point transform,
// the complex-to-real merge before the inverse one, and the inverse's
// epilogue.  The whole-row kernels (k_r2c, k_c2r) do the same arithmetic on
// rows held in LDS (ParisPencil.cpp:165 and 380: D2Z and Z2D along z).
#include <algorithm>
#include <type_traits>

#include "ff_internal.hpp"
#include "ff_kernels.hpp"

namespace ff {

namespace {

constexpr int kZThreads = 256;

// Workgroups of a grid-stride launch over total items.
unsigned int zsplit_blocks(long long total) {
  const long long b = (total + kZThreads - 1) / kZThreads;
  return (unsigned int)std::max(1LL, std::min(b, 256LL * 64));
}

// Forward split, in place on rows of nh complex values (row r at r*nh) that
// hold the packed transform Z[k], k < M, in natural order: X[k] = (Z[k] +
// conj Z[M-k])/2 - i W_nz^k (Z[k] - conj Z[M-k])/2 for k < M and X[M] =
// Re Z[0] - Im Z[0].  Item (r, k), k <= M/2, reads Z[k] and Z[M-k] and writes
// X[k] and X[M-k] (and X[M] for k = 0), so the pairs never overlap.
template <typename F>
__global__ __launch_bounds__(kZThreads) void k_zsplit_fwd(const RowArgs a) {
  const PassInst io = inst_of(a);
  const int M = a.nline;
  const long long P = M / 2 + 1;
  const long long total = a.nrows * P;
  FF_GLB cx<F>* z = glb<cx<F>>(io.dst);
  const FF_GLB cx<F>* twp = glb<const cx<F>>(a.twpost);
  auto half = [&](cx<F> zk, cx<F> zo, int k) {
    const cx<F> zm = conj(zo);
    // even spectrum (zk + zm)/2, odd spectrum (zk - zm)/(2i)
    const cx<F> ev = {F(0.5) * (zk.x + zm.x), F(0.5) * (zk.y + zm.y)};
    const cx<F> od = {F(0.5) * (zk.y - zm.y), F(-0.5) * (zk.x - zm.x)};
    return ev + twmul<-1>(od, twp[k]);
  };
  for (long long it = (long long)blockIdx.x * kZThreads + threadIdx.x; it < total;
       it += (long long)gridDim.x * kZThreads) {
    const long long r = it / P;
    const int k = (int)(it - r * P);
    FF_GLB cx<F>* row = z + r * a.nh;
    const int km = k ? M - k : 0;
    const cx<F> zk = row[k];
    const cx<F> zo = row[km];
    row[k] = half(zk, zo, k);
    if (km != k) row[km] = half(zo, zk, km);
    if (k == 0) row[M] = cx<F>{zk.x - zk.y, F(0)};
  }
}

// Inverse merge, src -> dst, both rows of nh complex values: the half
// spectrum X (Hermitian: the imaginary parts of X[0] and X[M] are ignored)
// to the packed spectrum Z'[k] = A + i B, A = X[k] + conj X[M-k],
// B = W_nz^-k (X[k] - conj X[M-k]), k < M.
template <typename F>
__global__ __launch_bounds__(kZThreads) void k_zsplit_inv(const RowArgs a) {
  const PassInst io = inst_of(a);
  const int M = a.nline;
  const long long total = a.nrows * M;
  const FF_GLB cx<F>* x = glb<const cx<F>>(io.src);
  FF_GLB cx<F>* z = glb<cx<F>>(io.dst);
  const FF_GLB cx<F>* twp = glb<const cx<F>>(a.twpost);
  for (long long it = (long long)blockIdx.x * kZThreads + threadIdx.x; it < total;
       it += (long long)gridDim.x * kZThreads) {
    const long long r = it / M;
    const int k = (int)(it - r * M);
    const FF_GLB cx<F>* row = x + r * a.nh;
    cx<F> xk = row[k];
    cx<F> xm = row[M - k];
    if (k == 0) {
      xk.y = F(0);
      xm.y = F(0);
    }
    const cx<F> cm = conj(xm);
    const cx<F> A = xk + cm;
    const cx<F> B = twmul<+1>(xk - cm, twp[k]);
    z[r * a.nh + k] = cx<F>{A.x - B.y, A.y + B.x};
  }
}

// Four-step twiddle, in place on packed rows of M = m1*m2 complex values
// (rows contiguous): element t1 + m1*k2 of every row times W_M^(t1*k2)
// (DIR < 0: between the forward sub-line passes) or its conjugate (DIR > 0:
// between the inverse ones).  a.tw: W_M, t < M.
template <typename F, int DIR>
__global__ __launch_bounds__(kZThreads) void k_ztwiddle(const RowArgs a) {
  const PassInst io = inst_of(a);
  const int M = a.nline, m1 = a.m1;
  const long long total = a.nrows * M;
  FF_GLB cx<F>* z = glb<cx<F>>(io.dst);
  const FF_GLB cx<F>* w = glb<const cx<F>>(a.tw);
  for (long long it = (long long)blockIdx.x * kZThreads + threadIdx.x; it < total;
       it += (long long)gridDim.x * kZThreads) {
    const int e = (int)(it % M);
    const int k2 = e / m1;
    const int t1 = e - k2 * m1;
    z[it] = twmul<DIR>(z[it], w[(long long)t1 * k2]);
  }
}

// FP64 real rows -> FP32 (the packed rows of an FP32 pass whose input is
// FP64): N values, src -> dst.
__global__ __launch_bounds__(kZThreads) void k_zpack(const RowArgs a) {
  const PassInst io = inst_of(a);
  const long long total = 2 * a.nrows * a.nline;
  const FF_GLB double* x = glb<const double>(io.src);
  FF_GLB float* y = glb<float>(io.dst);
  for (long long it = (long long)blockIdx.x * kZThreads + threadIdx.x; it < total;
       it += (long long)gridDim.x * kZThreads)
    y[it] = float(x[it]);
}

// Inverse epilogue: the packed rows z (M complex values per row, rows
// contiguous, so z is nrows*M pairs of real values in order) to the real
// output, stored, as the residual or accumulated (c2r_out2).
template <typename F, typename OUT, int EPI>
__global__ __launch_bounds__(kZThreads) void k_zepi(const RowArgs a) {
  const PassInst io = inst_of(a);
  const long long total = a.nrows * a.nline;
  const FF_GLB cx<F>* z = glb<const cx<F>>(io.src);
  FF_GLB OUT* out = glb<OUT>(io.dst);
  const FF_GLB double* f = glb<const double>(io.aux);
  const double shift = io.shiftp ? a.shift * *glb<const double>(io.shiftp) : a.shift;
  for (long long it = (long long)blockIdx.x * kZThreads + threadIdx.x; it < total;
       it += (long long)gridDim.x * kZThreads) {
    const cx<F> v = z[it];
    c2r_out2<EPI, OUT>(out, f, it, to_d(v.x), to_d(v.y), a.scale, shift);
  }
}

// ---- odd nz: two real rows per complex row ------------------------------------
// Pair p holds rows (2p, 2p+1): z[t] = x[2p][t] + i x[2p+1][t], t < n = nz
// (packed rows of n, pair stride n).  Its transform Z has the two spectra
// X1[k] = (Z[k] + conj Z[n-k])/2 and X2[k] = (Z[k] - conj Z[n-k])/(2i).
// With nx*ny odd the last row is alone: a launch of its own (a.lone, one
// pair) takes its second row as zero and never reads or writes that row.

// Packing: src real rows (IN) -> dst packed pairs (F).
template <typename F, typename IN>
__global__ __launch_bounds__(kZThreads) void k_zpair_pack(const RowArgs a) {
  const PassInst io = inst_of(a);
  const long long n = a.nline;
  const long long total = a.nrows * n;
  const FF_GLB IN* x = glb<const IN>(io.src);
  FF_GLB cx<F>* z = glb<cx<F>>(io.dst);
  for (long long it = (long long)blockIdx.x * kZThreads + threadIdx.x; it < total;
       it += (long long)gridDim.x * kZThreads) {
    const long long p = it / n, t = it - p * n;
    z[it] = cx<F>{F(x[2 * p * n + t]), a.lone ? F(0) : F(x[(2 * p + 1) * n + t])};
  }
}

// Separation in place: pair p's transform, in natural order at p*(n+1)
// (B's output), becomes the spectrum rows 2p (at p*(n+1)) and 2p+1 (at
// p*(n+1) + nh).  Item (p, k), k <= k' = nh-1-k, reads Z[k], Z[n-k], Z[k'] and
// Z[n-k'] = Z[nh+k] and writes X1[k], X2[k], X1[k'], X2[k'] at k, nh+k, k'
// and nh+k' = n-k: the same four places (k = 0: Z[0] twice, X2[nh-1] into the
// pair's spare slot n), so the items never overlap.
template <typename F>
__global__ __launch_bounds__(kZThreads) void k_zpair_split(const RowArgs a) {
  const PassInst io = inst_of(a);
  const int n = a.nline, nh = a.nh;
  const long long K = (nh + 1) / 2;
  const long long total = a.nrows * K;
  FF_GLB cx<F>* z = glb<cx<F>>(io.dst);
  auto x1 = [](cx<F> u, cx<F> v) { return cx<F>{F(0.5) * (u.x + v.x), F(0.5) * (u.y - v.y)}; };
  auto x2 = [](cx<F> u, cx<F> v) { return cx<F>{F(0.5) * (u.y + v.y), F(-0.5) * (u.x - v.x)}; };
  for (long long it = (long long)blockIdx.x * kZThreads + threadIdx.x; it < total;
       it += (long long)gridDim.x * kZThreads) {
    const long long p = it / K;
    const int k = (int)(it - p * K);
    const int kp = nh - 1 - k;
    FF_GLB cx<F>* Z = z + p * (long long)(n + 1);
    const cx<F> zk = Z[k];
    const cx<F> zmk = Z[k ? n - k : 0];
    const cx<F> zkp = Z[kp];
    const cx<F> zmkp = Z[n - kp];
    Z[k] = x1(zk, zmk);
    Z[nh + k] = x2(zk, zmk);
    if (kp != k) {
      Z[kp] = x1(zkp, zmkp);
      Z[nh + kp] = x2(zkp, zmkp);
    }
  }
}

// Merge, src -> dst: the spectrum rows 2p and 2p+1 (rows of nh; Hermitian,
// Im X[0] ignored) into pair p's complex spectrum Z[k] = X1[k] + i X2[k]
// (X[n-k] = conj X[k] above nh), natural order at p*(n+1).
template <typename F>
__global__ __launch_bounds__(kZThreads) void k_zpair_merge(const RowArgs a) {
  const PassInst io = inst_of(a);
  const int n = a.nline, nh = a.nh;
  const long long total = a.nrows * n;
  const FF_GLB cx<F>* x = glb<const cx<F>>(io.src);
  FF_GLB cx<F>* z = glb<cx<F>>(io.dst);
  for (long long it = (long long)blockIdx.x * kZThreads + threadIdx.x; it < total;
       it += (long long)gridDim.x * kZThreads) {
    const long long p = it / n;
    const int k = (int)(it - p * n);
    const FF_GLB cx<F>* r1 = x + 2 * p * nh;
    const FF_GLB cx<F>* r2 = r1 + nh;
    cx<F> u, v{F(0), F(0)};
    if (k < nh) {
      u = r1[k];
      if (!a.lone) v = r2[k];
      if (k == 0) {
        u.y = F(0);
        v.y = F(0);
      }
    } else {
      u = conj(r1[n - k]);
      if (!a.lone) v = conj(r2[n - k]);
    }
    z[p * (long long)(n + 1) + k] = cx<F>{u.x - v.y, u.y + v.x};
  }
}

// Unpacking epilogue: pair p's packed row (src, pair stride n) to the real
// rows 2p (real parts) and 2p+1 (imaginary parts), stored, as the residual or
// accumulated (c2r_out1).
template <typename F, typename OUT, int EPI>
__global__ __launch_bounds__(kZThreads) void k_zpair_epi(const RowArgs a) {
  const PassInst io = inst_of(a);
  const long long n = a.nline;
  const long long total = a.nrows * n;
  const FF_GLB cx<F>* z = glb<const cx<F>>(io.src);
  FF_GLB OUT* out = glb<OUT>(io.dst);
  const FF_GLB double* f = glb<const double>(io.aux);
  const double shift = io.shiftp ? a.shift * *glb<const double>(io.shiftp) : a.shift;
  for (long long it = (long long)blockIdx.x * kZThreads + threadIdx.x; it < total;
       it += (long long)gridDim.x * kZThreads) {
    const long long p = it / n, t = it - p * n;
    const cx<F> v = z[it];
    c2r_out1<EPI, OUT>(out, f, 2 * p * n + t, to_d(v.x), a.scale, shift);
    if (!a.lone) c2r_out1<EPI, OUT>(out, f, (2 * p + 1) * n + t, to_d(v.y), a.scale, shift);
  }
}

template <typename F>
const void* zsplit_fn(int step, int out_kind, int epi) {
  switch (step) {
    case ZSPLIT_FWD: return (const void*)&k_zsplit_fwd<F>;
    case ZSPLIT_INV: return (const void*)&k_zsplit_inv<F>;
    case ZSPLIT_TWF: return (const void*)&k_ztwiddle<F, -1>;
    case ZSPLIT_TWI: return (const void*)&k_ztwiddle<F, +1>;
    case ZSPLIT_PACK: return std::is_same<F, float>::value ? (const void*)&k_zpack : nullptr;
    case ZSPLIT_PPACK:  // out_kind: FP32 input rows
      return out_kind ? (std::is_same<F, float>::value ? (const void*)&k_zpair_pack<float, float> : nullptr)
                      : (const void*)&k_zpair_pack<F, double>;
    case ZSPLIT_PFWD: return (const void*)&k_zpair_split<F>;
    case ZSPLIT_PINV: return (const void*)&k_zpair_merge<F>;
    case ZSPLIT_PEPI:
      if constexpr (std::is_same<F, double>::value) {
        if (epi == EPI_STORE && out_kind == 0) return (const void*)&k_zpair_epi<double, double, EPI_STORE>;
        if (epi == EPI_RESIDUAL && out_kind == 1) return (const void*)&k_zpair_epi<double, float, EPI_RESIDUAL>;
        if (epi == EPI_RESIDUAL && out_kind == 0) return (const void*)&k_zpair_epi<double, double, EPI_RESIDUAL>;
      } else {
        if (epi == EPI_STORE && out_kind == 0) return (const void*)&k_zpair_epi<float, double, EPI_STORE>;
        if (epi == EPI_ACCUM && out_kind == 0) return (const void*)&k_zpair_epi<float, double, EPI_ACCUM>;
      }
      return nullptr;
    case ZSPLIT_EPI:
      // the epilogues the passes use: FP64 stored or residual (FP64 or FP32
      // output), FP32 stored or accumulated into FP64 phi
      if constexpr (std::is_same<F, double>::value) {
        if (epi == EPI_STORE && out_kind == 0) return (const void*)&k_zepi<double, double, EPI_STORE>;
        if (epi == EPI_RESIDUAL && out_kind == 1) return (const void*)&k_zepi<double, float, EPI_RESIDUAL>;
        if (epi == EPI_RESIDUAL && out_kind == 0) return (const void*)&k_zepi<double, double, EPI_RESIDUAL>;
      } else {
        if (epi == EPI_STORE && out_kind == 0) return (const void*)&k_zepi<float, double, EPI_STORE>;
        if (epi == EPI_ACCUM && out_kind == 0) return (const void*)&k_zepi<float, double, EPI_ACCUM>;
      }
      return nullptr;
    default: return nullptr;
  }
}

}  // namespace

hipError_t launch_zsplit(int step, int f32, int out_kind, int epi, const RowArgs& a, hipStream_t s) {
  if (launch_recording()) return hipErrorInvalidValue;  // never recorded
  const void* fn = f32 ? zsplit_fn<float>(step, out_kind, epi) : zsplit_fn<double>(step, out_kind, epi);
  if (!fn) return hipErrorInvalidValue;
  // items: the split's pairs of k, else the rows' complex values (real values
  // for the pack)
  long long total = (step == ZSPLIT_PACK ? 2 : 1) * a.nrows * a.nline;
  if (step == ZSPLIT_FWD) total = a.nrows * (a.nline / 2 + 1);
  if (step == ZSPLIT_PFWD) total = a.nrows * ((a.nh + 1) / 2);
  void* args[] = {const_cast<RowArgs*>(&a)};
  return hipLaunchKernel(fn, dim3(zsplit_blocks(total)), dim3(kZThreads), args, 0, s);
}

}  // namespace ff
