// Explicit instantiation of the pass kernels for one elements-per-thread
// value, FF_E (the build compiles this file once per E so the units build in
// parallel).
#include "ff_kernels.hpp"

#ifndef FF_E
#error "FF_E must be defined"
#endif

namespace ff {

static hipError_t launch(const void* fn, const Launch& l, const void* args, size_t) {
  if (l.lds > 65536) {
    hipError_t e = hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)l.lds);
    if (e != hipSuccess) return e;
  }
  void* argv[] = {const_cast<void*>(args)};
  return hipLaunchKernel(fn, l.grid, l.block, argv, l.lds, l.stream);
}

template <int E, bool S, int G>
static const void* col_fn(int f32, int mode) {
  if (f32) {
    switch (mode) {
      case COL_FWD: return (const void*)&k_col<float, E, COL_FWD, S, G>;
      case COL_INV: return (const void*)&k_col<float, E, COL_INV, S, G>;
      case COL_FILTER: return (const void*)&k_col<float, E, COL_FILTER, S, G>;
      default: return nullptr;
    }
  }
  switch (mode) {
    case COL_FWD: return (const void*)&k_col<double, E, COL_FWD, S, G>;
    case COL_INV: return (const void*)&k_col<double, E, COL_INV, S, G>;
    case COL_FILTER: return (const void*)&k_col<double, E, COL_FILTER, S, G>;
    case COL_SYMBOL: return (const void*)&k_col<double, E, COL_SYMBOL, S, G>;
    default: return nullptr;
  }
}

// Two column groups per workgroup are compiled for the elements-per-thread
// values that long column transforms use.
constexpr bool kTwoGroups = FF_E == 8 || FF_E == 16;

template <>
hipError_t PassLaunchers<FF_E>::col(int f32, int mode, int split, int groups, const Launch& l, const ColArgs& a) {
  const void* fn = nullptr;
  if (groups == 1) {
    fn = split ? col_fn<FF_E, true, 1>(f32, mode) : col_fn<FF_E, false, 1>(f32, mode);
  } else if (groups == 2) {
    if constexpr (kTwoGroups) fn = split ? col_fn<FF_E, true, 2>(f32, mode) : col_fn<FF_E, false, 2>(f32, mode);
  }
  if (!fn) return hipErrorInvalidValue;
  return launch(fn, l, &a, sizeof(a));
}

template <>
hipError_t PassLaunchers<FF_E>::r2c(int f32, int in_f32, int odd, const Launch& l, const RowArgs& a) {
  constexpr int E = FF_E;
  const void* fn = nullptr;
  if (!f32) {
    if (in_f32) return hipErrorInvalidValue;
    fn = odd ? (const void*)&k_r2c<double, double, E, true> : (const void*)&k_r2c<double, double, E, false>;
  } else if (!in_f32) {
    fn = odd ? (const void*)&k_r2c<float, double, E, true> : (const void*)&k_r2c<float, double, E, false>;
  } else {
    fn = odd ? (const void*)&k_r2c<float, float, E, true> : (const void*)&k_r2c<float, float, E, false>;
  }
  return launch(fn, l, &a, sizeof(a));
}

// out_kind: 0 = double output, 1 = float output.
template <>
hipError_t PassLaunchers<FF_E>::c2r(int f32, int out_kind, int odd, int epi, const Launch& l, const RowArgs& a) {
  constexpr int E = FF_E;
  const void* fn = nullptr;
#define FF_PICK(F, OUT, EPI) \
  fn = odd ? (const void*)&k_c2r<F, OUT, E, true, EPI> : (const void*)&k_c2r<F, OUT, E, false, EPI>
  if (!f32) {
    if (epi == EPI_STORE && out_kind == 0) FF_PICK(double, double, EPI_STORE);
    else if (epi == EPI_RESIDUAL && out_kind == 1) FF_PICK(double, float, EPI_RESIDUAL);
    else if (epi == EPI_RESIDUAL && out_kind == 0) FF_PICK(double, double, EPI_RESIDUAL);
    else return hipErrorInvalidValue;
  } else {
    if (epi == EPI_STORE && out_kind == 0) FF_PICK(float, double, EPI_STORE);
    else if (epi == EPI_ACCUM && out_kind == 0) FF_PICK(float, double, EPI_ACCUM);
    else return hipErrorInvalidValue;
  }
#undef FF_PICK
  return launch(fn, l, &a, sizeof(a));
}

}  // namespace ff
