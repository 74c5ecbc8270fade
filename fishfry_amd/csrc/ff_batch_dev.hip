// Device side of the batched solve: the LDS-resident runtime-radix Stockham
// engine, the three tile handlers (z forward, column, z inverse) and the
// persistent kernel that dequeues (op, tile) items.
#include "ff_batch.hpp"
#include "ff_kernels.hpp"

namespace ff {

// floor(u / d) for u, d < 2^16 with inv = ceil(2^32 / d), d > 1.
__device__ __forceinline__ int qdiv(int u, int d, uint32_t inv) {
  return d == 1 ? u : (int)__umulhi((uint32_t)u, inv);
}
__device__ __forceinline__ uint32_t qinv(int d) {
  return d > 1 ? (uint32_t)(((1ull << 32) + (unsigned long long)d - 1) / (unsigned long long)d) : 0u;
}

template <int R, int DIR, typename F>
__device__ __forceinline__ void bfly_lds(const cx<F>* src, cx<F>* dst, int j, int nr, int ns, uint32_t nsinv,
                                         int twstep, const cx<F>* __restrict__ tw) {
  cx<F> a[R];
#pragma unroll
  for (int m = 0; m < R; ++m) a[m] = src[j + m * nr];
  const int k = ns > 1 ? fast_mod(j, ns, nsinv) : 0;
  if (ns > 1) {
    const int e = k * twstep;
#pragma unroll
    for (int m = 1; m < R; ++m) a[m] = twmul<DIR>(a[m], tw[m * e]);
  }
  Dft<R, DIR, F>::run(a);
  const int base = (j - k) * R + k;
#pragma unroll
  for (int m = 0; m < R; ++m) dst[base + m * ns] = a[m];
}

// Transforms T lines (line l at A + l*pitch) in LDS, ping-ponging with B.
// Returns the buffer that holds the result.  Contains barriers: every
// thread of the workgroup must call it.
template <typename F, int DIR>
__device__ cx<F>* lds_fft(cx<F>* A, cx<F>* B, int T, int pitch, const BLine& lp, const cx<F>* __restrict__ tw) {
  for (int p = 0; p < lp.np; ++p) {
    const int R = lp.radix[p];
    const int nr = lp.nr[p];
    const int total = T * nr;
    for (int u = threadIdx.x; u < total; u += blockDim.x) {
      const int line = qdiv(u, nr, lp.nrinv[p]);
      const int j = u - line * nr;
      const cx<F>* s = A + line * pitch;
      cx<F>* d = B + line * pitch;
      switch (R) {
        case 2: bfly_lds<2, DIR, F>(s, d, j, nr, lp.ns[p], lp.nsinv[p], lp.twstep[p], tw); break;
        case 3: bfly_lds<3, DIR, F>(s, d, j, nr, lp.ns[p], lp.nsinv[p], lp.twstep[p], tw); break;
        case 4: bfly_lds<4, DIR, F>(s, d, j, nr, lp.ns[p], lp.nsinv[p], lp.twstep[p], tw); break;
        case 5: bfly_lds<5, DIR, F>(s, d, j, nr, lp.ns[p], lp.nsinv[p], lp.twstep[p], tw); break;
        case 6: bfly_lds<6, DIR, F>(s, d, j, nr, lp.ns[p], lp.nsinv[p], lp.twstep[p], tw); break;
        case 8: bfly_lds<8, DIR, F>(s, d, j, nr, lp.ns[p], lp.nsinv[p], lp.twstep[p], tw); break;
        case 10: bfly_lds<10, DIR, F>(s, d, j, nr, lp.ns[p], lp.nsinv[p], lp.twstep[p], tw); break;
        case 12: bfly_lds<12, DIR, F>(s, d, j, nr, lp.ns[p], lp.nsinv[p], lp.twstep[p], tw); break;
        case 16: bfly_lds<16, DIR, F>(s, d, j, nr, lp.ns[p], lp.nsinv[p], lp.twstep[p], tw); break;
        default: break;
      }
    }
    __syncthreads();
    cx<F>* t = A;
    A = B;
    B = t;
  }
  return A;
}

// ---- column op: y forward / inverse, x forward * filter * inverse ----------
template <typename F>
__device__ void item_col(const BOp& o, long long tile, unsigned char* smem) {
  cx<F>* A = reinterpret_cast<cx<F>*>(smem);
  cx<F>* B = A + (size_t)o.T * o.pitch;
  const cx<F>* __restrict__ tw = reinterpret_cast<const cx<F>*>(o.tw);
  const int outer = (int)(tile / o.tiles_pp);
  const int c0 = (int)(tile - (long long)outer * o.tiles_pp) << o.lgT;
  const int n = o.line.n;
  const int T = o.T;
  const int tot = T * n;
  const long long base = (long long)outer * o.s_outer + c0;
  const cx<F>* src = reinterpret_cast<const cx<F>*>(o.src) + base;
  for (int u = threadIdx.x; u < tot; u += blockDim.x) {
    const int r = u >> o.lgT;
    const int c = u & (T - 1);
    cx<F> v = {F(0), F(0)};
    if (c0 + c < o.nh) v = src[(long long)r * o.s_r + c];
    A[c * o.pitch + r] = v;
  }
  __syncthreads();
  cx<F>* X;
  if (o.mode == COL_INV) X = lds_fft<F, +1>(A, B, T, o.pitch, o.line, tw);
  else X = lds_fft<F, -1>(A, B, T, o.pitch, o.line, tw);
  if (o.mode == COL_FILTER || o.mode == COL_SYMBOL) {
    // spectral filter of ParisPencil.cpp:269-284 at (i, j, kh) = (r, outer, c0 + c)
    const uint32_t ninv = qinv(n);
    const int jj = outer;
    const double j2 = sqr(double(jj < o.ny - jj ? jj : o.ny - jj) * o.ddj);
    for (int u = threadIdx.x; u < tot; u += blockDim.x) {
      const int c = qdiv(u, n, ninv);
      const int ii = u - c * n;
      const int kk = c0 + c;
      const double k2 = sqr(double(kk) * o.ddk);
      const double i2 = sqr(double(ii < o.nx - ii ? ii : o.nx - ii) * o.ddi);
      cx<F>& e = X[c * o.pitch + ii];
      double d;
      if (o.mode == COL_FILTER) {
        d = (ii | jj | kk) ? -1.0 / (i2 + j2 + k2) : 0.0;
      } else {
        if (o.dc_out && (ii | jj | kk) == 0) *o.dc_out = double(e.x);
        d = -(i2 + j2 + k2);
      }
      e = cx<F>{F(d) * e.x, F(d) * e.y};
    }
    __syncthreads();
    X = lds_fft<F, +1>(X, X == A ? B : A, T, o.pitch, o.line, tw);
  }
  cx<F>* dst = reinterpret_cast<cx<F>*>(o.dst) + base;
  for (int u = threadIdx.x; u < tot; u += blockDim.x) {
    const int r = u >> o.lgT;
    const int c = u & (T - 1);
    if (c0 + c < o.nh) dst[(long long)r * o.s_r + c] = X[c * o.pitch + r];
  }
}

// ---- z forward: real rows -> half-spectrum rows ------------------------------
template <typename F, typename IN>
__device__ void item_zf(const BOp& o, long long tile, unsigned char* smem) {
  cx<F>* A = reinterpret_cast<cx<F>*>(smem);
  cx<F>* B = A + (size_t)o.T * o.pitch;
  const cx<F>* __restrict__ tw = reinterpret_cast<const cx<F>*>(o.tw);
  const long long row0 = tile * o.T;
  const int T = o.T;
  const int M = o.line.n;
  const IN* src = reinterpret_cast<const IN*>(o.src);
  cx<F>* dst = reinterpret_cast<cx<F>*>(o.dst);
  const uint32_t minv = qinv(M);
  for (int u = threadIdx.x; u < T * M; u += blockDim.x) {
    const int line = qdiv(u, M, minv);
    const int r = u - line * M;
    const long long row = row0 + line;
    cx<F> v = {F(0), F(0)};
    if (row < o.nlines) {
      if (!o.odd) v = cx<F>{F(src[row * o.nz + 2 * r]), F(src[row * o.nz + 2 * r + 1])};
      else v = cx<F>{F(src[row * o.nz + r]), F(0)};
    }
    A[line * o.pitch + r] = v;
  }
  __syncthreads();
  const cx<F>* X = lds_fft<F, -1>(A, B, T, o.pitch, o.line, tw);
  if (!o.odd) {
    const cx<F>* __restrict__ twp = reinterpret_cast<const cx<F>*>(o.twpost);
    const int M1 = M + 1;
    const uint32_t m1inv = qinv(M1);
    for (int u = threadIdx.x; u < T * M1; u += blockDim.x) {
      const int line = qdiv(u, M1, m1inv);
      const int k = u - line * M1;
      const long long row = row0 + line;
      if (row >= o.nlines) continue;
      const cx<F>* x = X + line * o.pitch;
      cx<F> out;
      if (k == M) {
        out = cx<F>{x[0].x - x[0].y, F(0)};
      } else {
        const cx<F> zk = x[k];
        const cx<F> zm = conj(x[k ? M - k : 0]);
        const cx<F> ev = {F(0.5) * (zk.x + zm.x), F(0.5) * (zk.y + zm.y)};
        const cx<F> od = {F(0.5) * (zk.y - zm.y), F(-0.5) * (zk.x - zm.x)};
        out = ev + twmul<-1>(od, twp[k]);
      }
      dst[row * o.nh + k] = out;
    }
  } else {
    const uint32_t hinv = qinv(o.nh);
    for (int u = threadIdx.x; u < T * o.nh; u += blockDim.x) {
      const int line = qdiv(u, o.nh, hinv);
      const int k = u - line * o.nh;
      const long long row = row0 + line;
      if (row < o.nlines) dst[row * o.nh + k] = X[line * o.pitch + k];
    }
  }
}

// ---- z inverse: half-spectrum rows -> real rows with an epilogue -------------
template <typename OUT>
__device__ __forceinline__ void zi_store(const BOp& o, OUT* out, long long off, double val) {
  if (o.mode == EPI_STORE) out[off] = OUT(o.scale * val);
  else if (o.mode == EPI_RESIDUAL) out[off] = OUT(o.aux[off] - o.scale * val);
  else out[off] = OUT(double(out[off]) + (o.scale * val - o.scale * *o.shiftp));
}

template <typename F>
__device__ void item_zi(const BOp& o, long long tile, unsigned char* smem) {
  cx<F>* A = reinterpret_cast<cx<F>*>(smem);
  cx<F>* B = A + (size_t)o.T * o.pitch;
  const cx<F>* __restrict__ tw = reinterpret_cast<const cx<F>*>(o.tw);
  const long long row0 = tile * o.T;
  const int T = o.T;
  const int M = o.line.n;
  const cx<F>* src = reinterpret_cast<const cx<F>*>(o.src);
  const int nload = o.odd ? o.nh : M + 1;
  const uint32_t linv = qinv(nload);
  for (int u = threadIdx.x; u < T * nload; u += blockDim.x) {
    const int line = qdiv(u, nload, linv);
    const int k = u - line * nload;
    const long long row = row0 + line;
    A[line * o.pitch + k] = row < o.nlines ? src[row * o.nh + k] : cx<F>{F(0), F(0)};
  }
  __syncthreads();
  const uint32_t minv = qinv(M);
  if (!o.odd) {
    const cx<F>* __restrict__ twp = reinterpret_cast<const cx<F>*>(o.twpost);
    for (int u = threadIdx.x; u < T * M; u += blockDim.x) {
      const int line = qdiv(u, M, minv);
      const int k = u - line * M;
      const cx<F>* x = A + line * o.pitch;
      cx<F> xk = x[k];
      cx<F> xm = x[M - k];
      if (k == 0) {
        xk.y = F(0);
        xm.y = F(0);
      }
      const cx<F> cm = conj(xm);
      const cx<F> a = xk + cm;
      const cx<F> b = twmul<+1>(xk - cm, twp[k]);
      B[line * o.pitch + k] = cx<F>{a.x - b.y, a.y + b.x};
    }
  } else {
    for (int u = threadIdx.x; u < T * M; u += blockDim.x) {
      const int line = qdiv(u, M, minv);
      const int k = u - line * M;
      const cx<F>* x = A + line * o.pitch;
      cx<F> y = k < o.nh ? x[k] : conj(x[o.nz - k]);
      if (k == 0) y.y = F(0);
      B[line * o.pitch + k] = y;
    }
  }
  __syncthreads();
  const cx<F>* X = lds_fft<F, +1>(B, A, T, o.pitch, o.line, tw);
  for (int u = threadIdx.x; u < T * M; u += blockDim.x) {
    const int line = qdiv(u, M, minv);
    const int r = u - line * M;
    const long long row = row0 + line;
    if (row >= o.nlines) continue;
    const cx<F> v = X[line * o.pitch + r];
    if (o.io32) {
      float* out = reinterpret_cast<float*>(o.dst) + row * o.nz;
      if (!o.odd) {
        zi_store<float>(o, out, 2 * r, double(v.x));
        zi_store<float>(o, out, 2 * r + 1, double(v.y));
      } else {
        zi_store<float>(o, out, r, double(v.x));
      }
    } else {
      double* out = reinterpret_cast<double*>(o.dst) + row * o.nz;
      if (!o.odd) {
        zi_store<double>(o, out, 2 * r, double(v.x));
        zi_store<double>(o, out, 2 * r + 1, double(v.y));
      } else {
        zi_store<double>(o, out, r, double(v.x));
      }
    }
  }
}

// ---- the persistent kernel ---------------------------------------------------

__global__ __launch_bounds__(256) void k_batch_reset(unsigned int* done, int nops, unsigned long long* head, int nheads,
                                                     unsigned int* err) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < nops; i += gridDim.x * blockDim.x) done[i] = 0u;
  if (blockIdx.x == 0) {
    for (int i = threadIdx.x; i < nheads; i += blockDim.x) head[i] = 0ull;
    if (threadIdx.x == 0) *err = 0u;
  }
}

__global__ __launch_bounds__(256) void k_batch(const BatchDev d) {
  extern __shared__ __align__(16) unsigned char smem[];
  __shared__ long long s_item;
  for (;;) {
    if (threadIdx.x == 0) s_item = d.item_begin + (long long)atomicAdd(d.head, 1ull);
    __syncthreads();
    const long long it = s_item;
    if (it >= d.item_end) break;  // uniform: every thread read the same s_item
    // op of this item: item_start[op] <= it < item_start[op + 1]
    int lo = 0, hi = d.nops;
    while (hi - lo > 1) {
      const int mid = (lo + hi) >> 1;
      if (d.item_start[mid] <= it) lo = mid;
      else hi = mid;
    }
    lo = __builtin_amdgcn_readfirstlane(lo);
    const BOp& o = d.ops[lo];
    const long long tile = it - d.item_start[lo];
    if (o.dep >= 0) {
      if (threadIdx.x == 0) {
        // Poll with relaxed loads (an acquire load would invalidate this
        // XCD's L2 on every poll) and back off; the wait is bounded in wall
        // time by the 100 MHz real-time counter, and once any wait has
        // failed every later wait gives up at once, so the grid always
        // drains quickly.
        const unsigned need = (unsigned)d.ops[o.dep].ntiles;
        const unsigned long long t0 = __builtin_amdgcn_s_memrealtime();
        while (__hip_atomic_load(&d.done[o.dep], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < need) {
          if (__builtin_amdgcn_s_memrealtime() - t0 > (unsigned long long)d.spin_limit ||
              __hip_atomic_load(d.err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) {
            atomicOr(d.err, 1u);
            break;
          }
          __builtin_amdgcn_s_sleep(32);
        }
      }
      __syncthreads();
      // one acquire per item: makes the producers' data visible to every wave
      __atomic_thread_fence(__ATOMIC_ACQUIRE);
    }
    if (o.kind == BOP_COL) {
      if (o.prec) item_col<float>(o, tile, smem);
      else item_col<double>(o, tile, smem);
    } else if (o.kind == BOP_ZF) {
      if (!o.prec) item_zf<double, double>(o, tile, smem);
      else if (o.io32) item_zf<float, float>(o, tile, smem);
      else item_zf<float, double>(o, tile, smem);
    } else {
      if (o.prec) item_zi<float>(o, tile, smem);
      else item_zi<double>(o, tile, smem);
    }
    // all of this workgroup's stores are complete before the release
    __builtin_amdgcn_s_waitcnt(0);
    __syncthreads();
    if (threadIdx.x == 0) __hip_atomic_fetch_add(&d.done[lo], 1u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
  }
}

// Allows lds_bytes of dynamic LDS; called when a batch binds its buffers,
// outside any stream capture.
hipError_t batch_prepare(int lds_bytes) {
  if (lds_bytes <= 65536) return hipSuccess;
  return hipFuncSetAttribute((const void*)&k_batch, hipFuncAttributeMaxDynamicSharedMemorySize, lds_bytes);
}

hipError_t batch_launch(const BatchDev& d, int grid, hipStream_t s) {
  hipLaunchKernelGGL(k_batch, dim3(grid), dim3(256), d.lds_bytes, s, d);
  return hipGetLastError();
}

hipError_t batch_reset(unsigned int* done, int nops, unsigned long long* head, int nheads, unsigned int* err,
                       hipStream_t s) {
  const int blocks = (nops + 255) / 256 > 0 ? ((nops + 255) / 256 < 64 ? (nops + 255) / 256 : 64) : 1;
  hipLaunchKernelGGL(k_batch_reset, dim3(blocks), dim3(256), 0, s, done, nops, head, nheads, err);
  return hipGetLastError();
}

hipError_t batch_occupancy(int lds_bytes, int* blocks_per_cu) {
  return hipOccupancyMaxActiveBlocksPerMultiprocessor(blocks_per_cu, (const void*)&k_batch, 256, lds_bytes);
}

}  // namespace ff
