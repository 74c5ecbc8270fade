// The library's static device work area, for ff_solve_noalloc: a solve whose
// kernels need a work area beyond density and potential (FP32 + refinement:
// spectrum, residual and two sums; FP64 with FF_PRESERVE_INPUT: the third
// spectrum array) takes it from this fixed array instead of memory the plan
// allocated, so the call itself allocates nothing.
//
// The array exists once per GPU the library runs on and is held until the
// process ends.  Its device address differs per GPU and is looked up once
// per device (hipGetSymbolAddress, with that device current) and kept; host
// code uses only that address, never the array's name.  Every call on one
// GPU shares the array: calls that overlap on a GPU race.  No kernel needs
// it cleared (every byte a solve reads was written earlier in that solve).
#include <hip/hip_runtime.h>

#include <mutex>

#include "ff_internal.hpp"

static __device__ __align__(256) unsigned char scratch[ff::kStaticWorkBytes];

namespace ff {

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_static_mutex;
void* g_static_ptr[kMaxDevices] = {};
}  // namespace

ff_status static_work_area(int device, void** out) {
  *out = nullptr;
  if (device < 0 || device >= kMaxDevices) return set_error(FF_ERR_INVALID, "static work area: device out of range");
  std::lock_guard<std::mutex> g(g_static_mutex);
  if (!g_static_ptr[device]) {
    int cur = -1;
    ff_status st = hip_check(hipGetDevice(&cur), "hipGetDevice");
    if (st != FF_OK) return st;
    if (cur != device)
      return set_error(FF_ERR_INVALID, "static work area: the plan's device must be the current device");
    void* p = nullptr;
    st = hip_check(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)), "hipGetSymbolAddress(static work area)");
    if (st != FF_OK) return st;
    if (!p) return set_error(FF_ERR_HIP, "static work area: null symbol address");
    g_static_ptr[device] = p;
  }
  *out = g_static_ptr[device];
  return FF_OK;
}

}  // namespace ff
