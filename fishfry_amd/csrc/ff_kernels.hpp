// Pass kernels of the periodic Poisson solve.
//
// Data layout in HBM (all passes): the half spectrum is complex [i][j][kh],
// kh fastest, row pitch nh = nz/2+1 -- the layout the reference has after its
// D2Z (ParisPencil.cpp:165) -- so no pass ever transposes.  The z passes work
// on contiguous rows; the y and x passes work on strided columns, a workgroup
// owning a tile of T neighbouring kh columns so every row of the tile is one
// contiguous T*sizeof(complex) segment in HBM.
//
//   k_r2c   z forward, real rows -> half-spectrum rows       (ParisPencil.cpp:165)
//   k_col   y forward / inverse, or x forward*filter*inverse (:212, :259-288, :334)
//   k_c2r   z inverse + the 1/(nx*ny*nz) scale                (:380, :406-421)
#pragma once

#include "ff_fft.hpp"

namespace ff {

// Workgroups are dispatched round-robin over the 8 XCDs; remap so that each
// XCD walks a contiguous range of tiles (neighbouring tiles share L2 lines).
__device__ __forceinline__ int xcd_remap(int b, int nb) {
  const int q = nb >> 3, r = nb & 7;
  const int x = b & 7, s = b >> 3;
  return x * q + (x < r ? x : r) + s;
}

__device__ __forceinline__ double sqr(double x) { return x * x; }

// Spectrum layout: element (i, j, kh) of the half spectrum lives at complex
// offset ((i >> lgB)*ny + j)*(B*nh) + (i & (B-1))*nh + kh, i.e. x is split
// into blocks of B planes with y between the blocks.  B = 1 is the
// reference's [i][j][kh]; B > 1 keeps both column passes within a few dozen
// pages per tile instead of one page per row for x.
// Two-level strides address it: position p along an axis maps to
// (p >> lg)*hi + (p & ((1 << lg) - 1))*lo.
struct Stride2 {
  long long hi, lo;
  int lg;
  __device__ __forceinline__ long long operator()(int p) const {
    return (long long)(p >> lg) * hi + (long long)(p & ((1 << lg) - 1)) * lo;
  }
};

struct ColArgs {
  const void* src;
  void* dst;
  Stride2 so;         // plane index -> offset
  Stride2 sr;         // transform position -> offset
  int ncols;          // columns per plane (nh)
  int nouter;         // planes
  int T, lgT;         // columns per workgroup (power of two)
  int tiles;          // tiles per plane
  int nblocks;        // nouter * tiles
  int ps;             // LDS pad shift
  int nx, ny;         // extents for the spectral symbol (x pass)
  double ddi, ddj, ddk;
  double* dc_out;  // FILTER / SYMBOL: receives the forward (0,0,0) term (sum of the input)
  long long rstep; // sr(p + L) - sr(p) for every p when that is constant, else 0
  const void* tw;
  LinePlan lp;
};

// MODE: 0 forward, 1 inverse, 2 forward * filter * inverse (the solve),
//       3 forward * symbol * inverse (applies L_h for a residual).
enum { COL_FWD = 0, COL_INV = 1, COL_FILTER = 2, COL_SYMBOL = 3 };

// SPLIT: exchange real and imaginary parts separately through half the LDS.
template <typename F, int E, int MODE, bool SPLIT>
__global__ __launch_bounds__(max_threads_for(E)) void k_col(const ColArgs a) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  cx<F>* lds = reinterpret_cast<cx<F>*>(smem_raw);
  const int tid = threadIdx.x;
  const int c = tid & (a.T - 1);
  const int r0 = tid >> a.lgT;
  const int L = a.lp.L;
  const int b = xcd_remap(blockIdx.x, a.nblocks);
  const int outer = b / a.tiles;
  const int tile = b - outer * a.tiles;
  const int col = tile * a.T + c;
  const bool valid = col < a.ncols;
  // Threads past the last column load the last column (in bounds, never
  // stored), so the loads need no predicate.
  const long long base = a.so(outer) + (valid ? col : a.ncols - 1);
  // src may equal dst (in-place pass): every thread stores exactly the
  // elements it loaded, after all of its loads.
  const cx<F>* src = reinterpret_cast<const cx<F>*>(a.src) + base;
  cx<F>* dst = reinterpret_cast<cx<F>*>(a.dst) + base;
  const cx<F>* __restrict__ tw = reinterpret_cast<const cx<F>*>(a.tw);

  cx<F> v[E];
  if (a.rstep) {
    const cx<F>* s = src + a.sr(r0);
#pragma unroll
    for (int i = 0; i < E; ++i) v[i] = s[(long long)i * a.rstep];
  } else {
#pragma unroll
    for (int i = 0; i < E; ++i) v[i] = src[a.sr(r0 + i * L)];
  }
  const ColIdx idx{a.T, c, a.ps};
  if constexpr (MODE == COL_FWD) {
    fft_line<E, -1, F, ColIdx, SPLIT>(v, a.lp, tw, lds, idx, r0);
  } else if constexpr (MODE == COL_INV) {
    fft_line<E, +1, F, ColIdx, SPLIT>(v, a.lp, tw, lds, idx, r0);
  } else {
    fft_line<E, -1, F, ColIdx, SPLIT>(v, a.lp, tw, lds, idx, r0);
    // spectral filter of ParisPencil.cpp:269-284 at (i, j, kh) = (r, outer, col)
    const int jj = outer, kk = col;
    const double j2 = sqr(double(jj < a.ny - jj ? jj : a.ny - jj) * a.ddj);
    const double k2 = sqr(double(kk) * a.ddk);
    // the (0,0,0) term before the multiply is the sum of the input field
    if (a.dc_out && valid && r0 == 0 && jj == 0 && kk == 0) *a.dc_out = double(v[0].x);
#pragma unroll
    for (int i = 0; i < E; ++i) {
      const int ii = r0 + i * L;
      const double i2 = sqr(double(ii < a.nx - ii ? ii : a.nx - ii) * a.ddi);
      double d;
      if constexpr (MODE == COL_FILTER) d = (ii | jj | kk) ? -1.0 / (i2 + j2 + k2) : 0.0;
      else d = -(i2 + j2 + k2);
      const F df = F(d);
      v[i] = cx<F>{df * v[i].x, df * v[i].y};
    }
    fft_line<E, +1, F, ColIdx, SPLIT>(v, a.lp, tw, lds, idx, r0);
  }
  if (valid) {
    if (a.rstep) {
      cx<F>* d = dst + a.sr(r0);
#pragma unroll
      for (int i = 0; i < E; ++i) d[(long long)i * a.rstep] = v[i];
    } else {
#pragma unroll
      for (int i = 0; i < E; ++i) dst[a.sr(r0 + i * L)] = v[i];
    }
  }
}

struct RowArgs {
  const void* src;
  void* dst;
  const void* aux;  // f for the residual epilogue
  long long nrows;
  int nz, nh;
  int ny, lgB;    // spectrum layout (see Stride2)
  int T;          // rows per workgroup
  uint32_t Linv;  // ceil(2^32 / L)
  int nblocks;
  int ps, pitch;  // LDS pad shift and row pitch (elements)
  double scale;   // 1/(nx*ny*nz)
  double shift;   // subtracted by the accumulate epilogue (times *shiftp if set)
  const double* shiftp;
  const void* tw;      // W_n of the transform length
  const void* twpost;  // exp(-2 pi i k / nz), k < nz/2 (even nz)
  LinePlan lp;
};

template <typename T>
__device__ __forceinline__ double to_d(T x) { return double(x); }

// Complex offset of spectrum row (i, j), row = i*ny + j.
__device__ __forceinline__ long long spec_row(const RowArgs& a, long long row) {
  const int i = (int)(row / a.ny);
  const int j = (int)(row - (long long)i * a.ny);
  const int B = 1 << a.lgB;
  return ((long long)(i >> a.lgB) * a.ny + j) * ((long long)B * a.nh) + (long long)(i & (B - 1)) * a.nh;
}

// z forward: real rows (IN precision) -> half-spectrum rows (F precision).
// Even nz: pack (x[2r], x[2r+1]) as one complex, an nz/2-point transform, then
// split the even/odd spectra.  Odd nz: a full nz-point transform of (x, 0).
template <typename F, typename IN, int E, bool ODD>
__global__ __launch_bounds__(max_threads_for(E)) void k_r2c(const RowArgs a) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  cx<F>* lds = reinterpret_cast<cx<F>*>(smem_raw);
  const int tid = threadIdx.x;
  const int L = a.lp.L;
  const int line = L == 1 ? tid : (int)__umulhi((uint32_t)tid, a.Linv);
  const int r0 = tid - line * L;
  const long long row = (long long)xcd_remap(blockIdx.x, a.nblocks) * a.T + line;
  const bool valid = row < a.nrows;
  const long long rowc = valid ? row : a.nrows - 1;  // loads stay in bounds; stores are predicated
  const int M = a.lp.n;
  const cx<F>* __restrict__ tw = reinterpret_cast<const cx<F>*>(a.tw);
  const RowIdx idx{line * a.pitch, a.ps};
  cx<F> v[E];
  if constexpr (!ODD) {
    const IN* __restrict__ src = reinterpret_cast<const IN*>(a.src) + rowc * a.nz;
#pragma unroll
    for (int i = 0; i < E; ++i) {
      const int k = r0 + i * L;
      if constexpr (sizeof(IN) == 8) {
        const cx<double> p = reinterpret_cast<const cx<double>*>(src)[k];
        v[i] = cx<F>{F(p.x), F(p.y)};
      } else {
        const cx<float> p = reinterpret_cast<const cx<float>*>(src)[k];
        v[i] = cx<F>{F(p.x), F(p.y)};
      }
    }
    fft_line<E, -1, F>(v, a.lp, tw, lds, idx, r0);
    __syncthreads();
    put_natural<E, F>(v, lds, idx, r0, L);
    __syncthreads();
    const cx<F>* __restrict__ twp = reinterpret_cast<const cx<F>*>(a.twpost);
    cx<F>* __restrict__ dst = reinterpret_cast<cx<F>*>(a.dst) + (valid ? spec_row(a, row) : 0);
    if (valid) {
#pragma unroll
      for (int i = 0; i < E; ++i) {
        const int k = r0 + i * L;
        const cx<F> zk = lds[idx(k)];
        const cx<F> zm = conj(lds[idx(k ? M - k : 0)]);
        // even spectrum (zk + zm)/2, odd spectrum (zk - zm)/(2i)
        const cx<F> ev = {F(0.5) * (zk.x + zm.x), F(0.5) * (zk.y + zm.y)};
        const cx<F> od = {F(0.5) * (zk.y - zm.y), F(-0.5) * (zk.x - zm.x)};
        dst[k] = ev + twmul<-1>(od, twp[k]);
      }
      if (r0 == 0) {
        const cx<F> z0 = lds[idx(0)];
        dst[M] = cx<F>{z0.x - z0.y, F(0)};
      }
    }
  } else {
    const IN* __restrict__ src = reinterpret_cast<const IN*>(a.src) + rowc * a.nz;
#pragma unroll
    for (int i = 0; i < E; ++i) v[i] = cx<F>{F(src[r0 + i * L]), F(0)};
    fft_line<E, -1, F>(v, a.lp, tw, lds, idx, r0);
    cx<F>* __restrict__ dst = reinterpret_cast<cx<F>*>(a.dst) + (valid ? spec_row(a, row) : 0);
    if (valid) {
#pragma unroll
      for (int i = 0; i < E; ++i) {
        const int k = r0 + i * L;
        if (k < a.nh) dst[k] = v[i];
      }
    }
  }
}

// Epilogues of the z inverse: store scale*x, store the residual
// f - shift - scale*x, or accumulate out += scale*x - shift.
enum { EPI_STORE = 0, EPI_RESIDUAL = 1, EPI_ACCUM = 2 };

template <int EPI, typename OUT>
__device__ __forceinline__ void c2r_out1(OUT* out, const double* f, long long off, double val, double scale,
                                         double shift) {
  if constexpr (EPI == EPI_STORE) {
    out[off] = OUT(scale * val);
  } else if constexpr (EPI == EPI_RESIDUAL) {
    out[off] = OUT((f[off] - shift) - scale * val);
  } else {
    out[off] = OUT(double(out[off]) + (scale * val - shift));
  }
}

// Two neighbouring outputs (2r, 2r+1) with one vector access.
template <int EPI, typename OUT>
__device__ __forceinline__ void c2r_out2(OUT* out, const double* f, long long r, double v0, double v1,
                                         double scale, double shift) {
  cx<OUT>* o = reinterpret_cast<cx<OUT>*>(out) + r;
  if constexpr (EPI == EPI_STORE) {
    *o = cx<OUT>{OUT(scale * v0), OUT(scale * v1)};
  } else if constexpr (EPI == EPI_RESIDUAL) {
    const cx<double> fp = reinterpret_cast<const cx<double>*>(f)[r];
    *o = cx<OUT>{OUT((fp.x - shift) - scale * v0), OUT((fp.y - shift) - scale * v1)};
  } else {
    const cx<OUT> p = *o;
    *o = cx<OUT>{OUT(double(p.x) + (scale * v0 - shift)), OUT(double(p.y) + (scale * v1 - shift))};
  }
}

// z inverse: half-spectrum rows (F) -> real rows (OUT) with an epilogue.
// The imaginary parts of the DC and (even nz) Nyquist terms are ignored,
// i.e. the rows are read as Hermitian half spectra.
template <typename F, typename OUT, int E, bool ODD, int EPI>
__global__ __launch_bounds__(max_threads_for(E)) void k_c2r(const RowArgs a) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  cx<F>* lds = reinterpret_cast<cx<F>*>(smem_raw);
  const int tid = threadIdx.x;
  const int L = a.lp.L;
  const int line = L == 1 ? tid : (int)__umulhi((uint32_t)tid, a.Linv);
  const int r0 = tid - line * L;
  const long long row = (long long)xcd_remap(blockIdx.x, a.nblocks) * a.T + line;
  const bool valid = row < a.nrows;
  const int M = a.lp.n;
  const cx<F>* __restrict__ tw = reinterpret_cast<const cx<F>*>(a.tw);
  const RowIdx idx{line * a.pitch, a.ps};
  // rows past the end load the last row (in bounds); their stores are predicated
  const cx<F>* __restrict__ src = reinterpret_cast<const cx<F>*>(a.src) + spec_row(a, valid ? row : a.nrows - 1);
  // stage the half spectrum in LDS
#pragma unroll
  for (int i = 0; i < E; ++i) {
    const int k = r0 + i * L;
    if (!ODD || k < a.nh) lds[idx(k)] = src[k];
  }
  if (!ODD && r0 == 0) lds[idx(M)] = src[M];
  __syncthreads();
  cx<F> v[E];
  if constexpr (!ODD) {
    const cx<F>* __restrict__ twp = reinterpret_cast<const cx<F>*>(a.twpost);
#pragma unroll
    for (int i = 0; i < E; ++i) {
      const int k = r0 + i * L;
      cx<F> xk = lds[idx(k)];
      cx<F> xm = lds[idx(M - k)];
      if (k == 0) {
        xk.y = F(0);
        xm.y = F(0);
      }
      const cx<F> cm = conj(xm);
      const cx<F> A = xk + cm;
      const cx<F> B = twmul<+1>(xk - cm, twp[k]);
      v[i] = cx<F>{A.x - B.y, A.y + B.x};
    }
  } else {
#pragma unroll
    for (int i = 0; i < E; ++i) {
      const int k = r0 + i * L;
      cx<F> y = k < a.nh ? lds[idx(k)] : conj(lds[idx(a.nz - k)]);
      if (k == 0) y.y = F(0);
      v[i] = y;
    }
  }
  fft_line<E, +1, F>(v, a.lp, tw, lds, idx, r0);
  if (valid) {
    const double shift = a.shiftp ? a.shift * *a.shiftp : a.shift;
    OUT* out = reinterpret_cast<OUT*>(a.dst) + row * a.nz;
    const double* f = reinterpret_cast<const double*>(a.aux) + (EPI == EPI_RESIDUAL ? row * a.nz : 0);
#pragma unroll
    for (int i = 0; i < E; ++i) {
      const int r = r0 + i * L;
      if constexpr (!ODD) {
        c2r_out2<EPI, OUT>(out, f, r, to_d(v[i].x), to_d(v[i].y), a.scale, shift);
      } else {
        c2r_out1<EPI, OUT>(out, f, r, to_d(v[i].x), a.scale, shift);
      }
    }
  }
}

// ---------------------------------------------------------------------------
// Launchers: one explicit instantiation unit per E (ff_inst_E*.hip).

struct Launch {
  dim3 grid, block;
  size_t lds;
  hipStream_t stream;
};

template <int E>
struct PassLaunchers {
  static hipError_t col(int f32, int mode, int split, const Launch& l, const ColArgs& a);
  static hipError_t r2c(int f32, int in_f32, int odd, const Launch& l, const RowArgs& a);
  static hipError_t c2r(int f32, int out_kind, int odd, int epi, const Launch& l, const RowArgs& a);
};

}  // namespace ff
