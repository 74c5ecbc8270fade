// Explicit instantiation of the pass kernels for one elements-per-thread
// value, FF_E (the build compiles this file once per E so the units build in
// parallel).
#include <type_traits>

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

template <int E, bool S, int G, bool BL>
static const void* col_fn(int f32, int mode) {
  if (f32) {
    switch (mode) {
      case COL_FWD: return (const void*)&k_col<float, E, COL_FWD, S, G, BL>;
      case COL_INV: return (const void*)&k_col<float, E, COL_INV, S, G, BL>;
      case COL_FILTER: return (const void*)&k_col<float, E, COL_FILTER, S, G, BL>;
      case COL_FILTER_FD: return (const void*)&k_col<float, E, COL_FILTER_FD, S, G, BL>;
      default: return nullptr;
    }
  }
  switch (mode) {
    case COL_FWD: return (const void*)&k_col<double, E, COL_FWD, S, G, BL>;
    case COL_INV: return (const void*)&k_col<double, E, COL_INV, S, G, BL>;
    case COL_FILTER: return (const void*)&k_col<double, E, COL_FILTER, S, G, BL>;
    case COL_SYMBOL: return (const void*)&k_col<double, E, COL_SYMBOL, S, G, BL>;
    case COL_FILTER_FD: return (const void*)&k_col<double, E, COL_FILTER_FD, S, G, BL>;
    default: return nullptr;
  }
}

// Two column groups per workgroup are compiled for the elements-per-thread
// values that long column transforms use.
constexpr bool kTwoGroups = FF_E == 8 || FF_E == 16;

// Kernels with the line plan compiled in: the kStaticCols entries of this E,
// with twiddles from the table in memory or (TWL) from LDS.
template <typename F, int MODE, bool S, int N, int T, int W, bool TWL>
static const void* col_static_kernel() {
  return (const void*)&k_col<F, FF_E, MODE, S, 1, false, N, T, W, TWL>;
}

template <typename F, bool S, int N, int T, int W, int KINDS, bool TWL>
static const void* col_static_modes(int mode) {
  constexpr bool y = (KINDS & 2) != 0, x = (KINDS & 4) != 0;
  switch (mode) {
    case COL_FWD:
      if constexpr (y) return col_static_kernel<F, COL_FWD, S, N, T, W, TWL>();
      return nullptr;
    case COL_INV:
      if constexpr (y) return col_static_kernel<F, COL_INV, S, N, T, W, TWL>();
      return nullptr;
    case COL_FILTER:
      if constexpr (x) return col_static_kernel<F, COL_FILTER, S, N, T, W, TWL>();
      return nullptr;
    case COL_FILTER_FD:
      if constexpr (x) return col_static_kernel<F, COL_FILTER_FD, S, N, T, W, TWL>();
      return nullptr;
    case COL_SYMBOL:
      if constexpr (x && std::is_same<F, double>::value) return col_static_kernel<F, COL_SYMBOL, S, N, T, W, TWL>();
      return nullptr;
    default: return nullptr;
  }
}

// stat: 1 twiddles from the table in memory, 2 from LDS.
template <int K>
static const void* col_static_fn(int f32, int split, int mode, int n, int T, int stat) {
  if constexpr (K < kNumStaticCols) {
    constexpr StaticCol sc = kStaticCols[K];
    if constexpr (sc.E == FF_E) {
      if (sc.n == n && sc.T == T && sc.f32 == f32 && sc.split == split) {
        // entries may serve different passes of one shape: keep looking when
        // this one has no kernel for the mode
        using F = typename std::conditional<sc.f32 != 0, float, double>::type;
        const void* fn = nullptr;
        if (stat == 1) {
          fn = col_static_modes<F, sc.split != 0, sc.n, sc.T, sc.waves, sc.kinds, false>(mode);
        } else if constexpr (sc.twl_waves > 0) {
          fn = col_static_modes<F, sc.split != 0, sc.n, sc.T, sc.twl_waves, sc.kinds, true>(mode);
        }
        if (fn) return fn;
      }
    }
    return col_static_fn<K + 1>(f32, split, mode, n, T, stat);
  } else {
    return nullptr;
  }
}

// Bluestein lines use the default kernel shape (one group, complex exchange).
template <>
hipError_t PassLaunchers<FF_E>::col(int f32, int mode, int split, int groups, int blue, int stat, const Launch& l,
                                    const ColArgs& a) {
  const void* fn = nullptr;
  if (stat) {
    fn = col_static_fn<0>(f32, split, mode, a.lp.n, a.T, stat);
  } else if (blue) {
    if (groups == 1 && !split) fn = col_fn<FF_E, false, 1, true>(f32, mode);
  } else if (groups == 1) {
    fn = split ? col_fn<FF_E, true, 1, false>(f32, mode) : col_fn<FF_E, false, 1, false>(f32, mode);
  } else if (groups == 2) {
    if constexpr (kTwoGroups)
      fn = split ? col_fn<FF_E, true, 2, false>(f32, mode) : col_fn<FF_E, false, 2, false>(f32, mode);
  }
  if (!fn) return hipErrorInvalidValue;
  return launch(fn, l, &a, sizeof(a));
}

template <bool BL, int SN = 0, bool TWL = false>
static const void* r2c_fn(int f32, int in_f32, int odd) {
  constexpr int E = FF_E;
  if constexpr (SN > 0) {  // compiled-in plans: even nz only
    if (odd) return nullptr;
    if (!f32) return in_f32 ? nullptr : (const void*)&k_r2c<double, double, E, false, BL, SN, TWL>;
    return in_f32 ? (const void*)&k_r2c<float, float, E, false, BL, SN, TWL>
                  : (const void*)&k_r2c<float, double, E, false, BL, SN, TWL>;
  }
  if (!f32) {
    if (in_f32) return nullptr;
    return odd ? (const void*)&k_r2c<double, double, E, true, BL> : (const void*)&k_r2c<double, double, E, false, BL>;
  }
  if (!in_f32)
    return odd ? (const void*)&k_r2c<float, double, E, true, BL> : (const void*)&k_r2c<float, double, E, false, BL>;
  return odd ? (const void*)&k_r2c<float, float, E, true, BL> : (const void*)&k_r2c<float, float, E, false, BL>;
}

// Row kernels with the plan of a kStaticRows entry of this E compiled in
// (stat 2: twiddles from LDS).
template <int K>
static const void* r2c_static_fn(int f32, int in_f32, int odd, int M, int stat) {
  if constexpr (K < kNumStaticRows) {
    constexpr StaticRow sr = kStaticRows[K];
    if constexpr (sr.E == FF_E) {
      if (sr.M == M)
        return stat == 2 ? r2c_fn<false, sr.M, true>(f32, in_f32, odd) : r2c_fn<false, sr.M>(f32, in_f32, odd);
    }
    return r2c_static_fn<K + 1>(f32, in_f32, odd, M, stat);
  } else {
    return nullptr;
  }
}

template <>
hipError_t PassLaunchers<FF_E>::r2c(int f32, int in_f32, int odd, int blue, int stat, const Launch& l,
                                    const RowArgs& a) {
  const void* fn = stat ? r2c_static_fn<0>(f32, in_f32, odd, a.lp.n, stat)
                        : (blue ? r2c_fn<true>(f32, in_f32, odd) : r2c_fn<false>(f32, in_f32, odd));
  if (!fn) return hipErrorInvalidValue;
  return launch(fn, l, &a, sizeof(a));
}

// out_kind: 0 = double output, 1 = float output.
template <bool BL, int SN = 0, bool TWL = false>
static const void* c2r_fn(int f32, int out_kind, int odd, int epi) {
  constexpr int E = FF_E;
  // (compiled-in plans: even nz only)
#define FF_PICK(F, OUT, EPI)                                                                           \
  {                                                                                                    \
    if constexpr (SN > 0)                                                                              \
      return odd ? nullptr : (const void*)&k_c2r<F, OUT, E, false, EPI, BL, SN, TWL>;                  \
    else                                                                                               \
      return odd ? (const void*)&k_c2r<F, OUT, E, true, EPI, BL>                                       \
                 : (const void*)&k_c2r<F, OUT, E, false, EPI, BL>;                                     \
  }
  if (!f32) {
    if (epi == EPI_STORE && out_kind == 0) FF_PICK(double, double, EPI_STORE)
    if (epi == EPI_RESIDUAL && out_kind == 1) FF_PICK(double, float, EPI_RESIDUAL)
    if (epi == EPI_RESIDUAL && out_kind == 0) FF_PICK(double, double, EPI_RESIDUAL)
    return nullptr;
  }
  if (epi == EPI_STORE && out_kind == 0) FF_PICK(float, double, EPI_STORE)
  if (epi == EPI_ACCUM && out_kind == 0) FF_PICK(float, double, EPI_ACCUM)
  return nullptr;
#undef FF_PICK
}

template <int K>
static const void* c2r_static_fn(int f32, int out_kind, int odd, int epi, int M, int stat) {
  if constexpr (K < kNumStaticRows) {
    constexpr StaticRow sr = kStaticRows[K];
    if constexpr (sr.E == FF_E) {
      if (sr.M == M)
        return stat == 2 ? c2r_fn<false, sr.M, true>(f32, out_kind, odd, epi)
                         : c2r_fn<false, sr.M>(f32, out_kind, odd, epi);
    }
    return c2r_static_fn<K + 1>(f32, out_kind, odd, epi, M, stat);
  } else {
    return nullptr;
  }
}

template <>
hipError_t PassLaunchers<FF_E>::c2r(int f32, int out_kind, int odd, int epi, int blue, int stat, const Launch& l,
                                    const RowArgs& a) {
  const void* fn = stat ? c2r_static_fn<0>(f32, out_kind, odd, epi, a.lp.n, stat)
                        : (blue ? c2r_fn<true>(f32, out_kind, odd, epi) : c2r_fn<false>(f32, out_kind, odd, epi));
  if (!fn) return hipErrorInvalidValue;
  return launch(fn, l, &a, sizeof(a));
}

}  // namespace ff
