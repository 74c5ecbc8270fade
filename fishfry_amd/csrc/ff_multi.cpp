// Host side of the multi-shape pass launches (see ff_multi.hpp): variant
// lookup and the launch itself.
#include "ff_multi.hpp"

namespace ff {

namespace {

bool same_plan(const LinePlan& s, const LinePlan& lp) {
  if (s.n != lp.n || s.L != lp.L || s.np != lp.np) return false;
  for (int p = 0; p < s.np; ++p)
    if (s.radix[p] != lp.radix[p] || s.ns[p] != lp.ns[p] || s.nsinv[p] != lp.nsinv[p] ||
        s.twstep[p] != lp.twstep[p])
      return false;
  return true;
}

template <int K>
int col_variant(int n, int E, const LinePlan& lp) {
  if constexpr (K < kNumMultiCols) {
    constexpr MultiVar mv = kMultiCols[K];
    if (mv.n == n && mv.E == E) {
      constexpr LinePlan s = static_plan<mv.n, mv.E>();
      return same_plan(s, lp) ? K : -1;
    }
    return col_variant<K + 1>(n, E, lp);
  } else {
    return -1;
  }
}

template <int K>
int row_variant(int M, int E, const LinePlan& lp) {
  if constexpr (K < kNumMultiRows) {
    constexpr MultiVar mv = kMultiRows[K];
    if (mv.n == M && mv.E == E) {
      constexpr LinePlan s = static_plan<mv.n, mv.E>();
      return same_plan(s, lp) ? K : -1;
    }
    return row_variant<K + 1>(M, E, lp);
  } else {
    return -1;
  }
}

const void* find_fn(int kind, int f32, int sel, int epi) {
  const void* fn = multi_fn<0>(kind, f32, sel, epi);
  if (!fn) fn = multi_fn<1>(kind, f32, sel, epi);
  if (!fn) fn = multi_fn<2>(kind, f32, sel, epi);
  if (!fn) fn = multi_fn<3>(kind, f32, sel, epi);
  if (!fn) fn = multi_fn<4>(kind, f32, sel, epi);
  return fn;
}

}  // namespace

int multi_col_variant(int n, int E, const LinePlan& lp) { return col_variant<0>(n, E, lp); }
int multi_row_variant(int M, int E, const LinePlan& lp) { return row_variant<0>(M, E, lp); }

bool multi_kernel_exists(int kind, int f32, int sel, int epi) { return find_fn(kind, f32, sel, epi) != nullptr; }

hipError_t launch_multi(int kind, int f32, int sel, int epi, const MultiArgs& m, size_t lds, hipStream_t s) {
  const void* fn = find_fn(kind, f32, sel, epi);
  if (!fn || m.nblocks < 1 || m.ndesc < 1) return hipErrorInvalidValue;
  if (lds > 65536) {
    hipError_t e = hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    if (e != hipSuccess) return e;
  }
  void* argv[] = {const_cast<MultiArgs*>(&m)};
  return hipLaunchKernel(fn, dim3(m.nblocks), dim3(kMultiThreads), argv, lds, s);
}

}  // namespace ff
