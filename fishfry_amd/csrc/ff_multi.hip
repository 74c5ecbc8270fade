// Multi-shape pass kernels (see ff_multi.hpp).  The build compiles this file
// once per FF_MULTI_PART so the kernel families build in parallel:
//   0  column passes, FP64, forward and inverse (y)
//   1  column passes, FP64, forward*filter*inverse and forward*symbol*inverse (x)
//   2  column passes, FP32, all three
//   3  z forward (R2C) rows
//   4  z inverse (C2R) rows with their epilogues
#include <climits>

#include "ff_multi.hpp"

#ifndef FF_MULTI_PART
#error "FF_MULTI_PART must be defined"
#endif

namespace ff {

// Descriptor of logical tile b: the number of first tiles <= b, minus one.
// Lane l of the wave compares descriptor base + l; the count of a ballot is
// wave-uniform, so the result lives in scalar registers.
__device__ __forceinline__ int multi_find(const int* block0, int ndesc, int b) {
  const int lane = threadIdx.x & 63;
  int base = 0;
  for (;;) {
    const int i = base + lane;
    const int v = i < ndesc ? glb<const int>(block0)[i] : INT_MAX;
    const int c = __popcll(__ballot(v <= b));
    if (c < 64 || base + 64 >= ndesc) return base + c - 1;
    base += 64;
  }
}

template <typename F, int MODE, int K>
__device__ __forceinline__ void mcol_run(int v, const ColArgs& a, const PassInst& io, int b) {
  if constexpr (K < kNumMultiCols) {
    constexpr MultiVar mv = kMultiCols[K];
    if (v == K) {
      col_tile<F, mv.E, MODE, false, 1, false, mv.n, multi_tile(mv)>(a, io, b);
      return;
    }
    mcol_run<F, MODE, K + 1>(v, a, io, b);
  }
}

// Logical tile of this workgroup, or -1 past its XCD chunk's end.
__device__ __forceinline__ int multi_tile_of(const MultiArgs& m) {
  const int x = blockIdx.x % kXcds, s = blockIdx.x / kXcds;
  return s < m.chunkn[x] ? m.chunk0[x] + s : -1;
}

template <typename F, int MODE>
__global__ __attribute__((amdgpu_flat_work_group_size(1, kMultiThreads), amdgpu_waves_per_eu(2))) void k_mcol(
    const MultiArgs m) {
  const int gb = multi_tile_of(m);
  if (gb < 0) return;
  const int d = __builtin_amdgcn_readfirstlane(multi_find(m.block0, m.ndesc, gb));
  const FF_GLB MultiColDesc* D = glb<const MultiColDesc>(m.desc) + d;
  const ColArgs a = D->a;
  const int v = __builtin_amdgcn_readfirstlane(D->variant);
  const int b = __builtin_amdgcn_readfirstlane(gb - glb<const int>(m.block0)[d]);
  const PassInst io = uniform_inst(PassInst{a.src, a.dst, nullptr, a.dc_out, nullptr, a.ddi, a.ddj, a.ddk});
  mcol_run<F, MODE, 0>(v, a, io, b);
}

template <typename F, typename IN, int K>
__device__ __forceinline__ void mr2c_run(int v, const RowArgs& a, const PassInst& io, int b) {
  if constexpr (K < kNumMultiRows) {
    constexpr MultiVar mv = kMultiRows[K];
    if (v == K) {
      r2c_tile<F, IN, mv.E, false, false, mv.n>(a, io, b);
      return;
    }
    mr2c_run<F, IN, K + 1>(v, a, io, b);
  }
}

template <typename F, typename OUT, int EPI, int K>
__device__ __forceinline__ void mc2r_run(int v, const RowArgs& a, const PassInst& io, int b) {
  if constexpr (K < kNumMultiRows) {
    constexpr MultiVar mv = kMultiRows[K];
    if (v == K) {
      c2r_tile<F, OUT, mv.E, false, EPI, false, mv.n>(a, io, b);
      return;
    }
    mc2r_run<F, OUT, EPI, K + 1>(v, a, io, b);
  }
}

// The row descriptor of logical tile gb: (descriptor arguments, its
// per-instance fields, its row block).
__device__ __forceinline__ void multi_row_desc(const MultiArgs& m, int gb, RowArgs& a, PassInst& io, int& v,
                                               int& b) {
  const int d = __builtin_amdgcn_readfirstlane(multi_find(m.block0, m.ndesc, gb));
  const FF_GLB MultiRowDesc* D = glb<const MultiRowDesc>(m.desc) + d;
  a = D->a;
  v = __builtin_amdgcn_readfirstlane(D->variant);
  b = __builtin_amdgcn_readfirstlane(gb - glb<const int>(m.block0)[d]);
  io = uniform_inst(PassInst{a.src, a.dst, a.aux, nullptr, a.shiftp, 0.0, 0.0, 0.0});
}

template <typename F, typename IN>
__global__ __attribute__((amdgpu_flat_work_group_size(1, kMultiThreads), amdgpu_waves_per_eu(2))) void k_mr2c(
    const MultiArgs m) {
  const int gb = multi_tile_of(m);
  if (gb < 0) return;
  RowArgs a;
  PassInst io;
  int v, b;
  multi_row_desc(m, gb, a, io, v, b);
  mr2c_run<F, IN, 0>(v, a, io, b);
}

template <typename F, typename OUT, int EPI>
__global__ __attribute__((amdgpu_flat_work_group_size(1, kMultiThreads), amdgpu_waves_per_eu(2))) void k_mc2r(
    const MultiArgs m) {
  const int gb = multi_tile_of(m);
  if (gb < 0) return;
  RowArgs a;
  PassInst io;
  int v, b;
  multi_row_desc(m, gb, a, io, v, b);
  mc2r_run<F, OUT, EPI, 0>(v, a, io, b);
}

// The kernel of a role in this part, or null.
template <>
const void* multi_fn<FF_MULTI_PART>(int kind, int f32, int sel, int epi) {
#if FF_MULTI_PART == 0
  if (kind == MULTI_COL && !f32 && sel == COL_FWD) return (const void*)&k_mcol<double, COL_FWD>;
  if (kind == MULTI_COL && !f32 && sel == COL_INV) return (const void*)&k_mcol<double, COL_INV>;
#elif FF_MULTI_PART == 1
  if (kind == MULTI_COL && !f32 && sel == COL_FILTER) return (const void*)&k_mcol<double, COL_FILTER>;
  if (kind == MULTI_COL && !f32 && sel == COL_SYMBOL) return (const void*)&k_mcol<double, COL_SYMBOL>;
#elif FF_MULTI_PART == 2
  if (kind == MULTI_COL && f32 && sel == COL_FWD) return (const void*)&k_mcol<float, COL_FWD>;
  if (kind == MULTI_COL && f32 && sel == COL_INV) return (const void*)&k_mcol<float, COL_INV>;
  if (kind == MULTI_COL && f32 && sel == COL_FILTER) return (const void*)&k_mcol<float, COL_FILTER>;
#elif FF_MULTI_PART == 3
  if (kind == MULTI_R2C && !f32 && !sel) return (const void*)&k_mr2c<double, double>;
  if (kind == MULTI_R2C && f32 && !sel) return (const void*)&k_mr2c<float, double>;
  if (kind == MULTI_R2C && f32 && sel) return (const void*)&k_mr2c<float, float>;
#elif FF_MULTI_PART == 4
  if (kind == MULTI_C2R && !f32 && !sel && epi == EPI_STORE) return (const void*)&k_mc2r<double, double, EPI_STORE>;
  if (kind == MULTI_C2R && f32 && !sel && epi == EPI_STORE) return (const void*)&k_mc2r<float, double, EPI_STORE>;
  if (kind == MULTI_C2R && !f32 && sel && epi == EPI_RESIDUAL)
    return (const void*)&k_mc2r<double, float, EPI_RESIDUAL>;
  if (kind == MULTI_C2R && f32 && !sel && epi == EPI_ACCUM) return (const void*)&k_mc2r<float, double, EPI_ACCUM>;
#endif
  (void)kind;
  (void)f32;
  (void)sel;
  (void)epi;
  return nullptr;
}

}  // namespace ff
