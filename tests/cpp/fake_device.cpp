// A CPU stand-in for the device, linked in place of the HIP runtime and the
// solver so the serving scheduler (fishfry_amd/csrc/ff_server.cpp) runs under
// ThreadSanitizer without a GPU.
//
// "Device" memory is host memory.  Every operation enqueued on a stream takes
// effect at once, and the stream is busy until a simulated completion time
// (a fixed cost plus one per element), so events complete asynchronously as
// the scheduler's polling expects.  The solve stand-in writes
// phi = 0.5 * f + 1; the field stand-ins write f(i) = i and error sums whose
// relative norms are (0.5, 1, 0.5).
#include <hip/hip_runtime.h>

#include <chrono>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <string>
#include <thread>

#include "ff_internal.hpp"
#include "ff_solver.hpp"

namespace {

double now_s() {
  return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

struct FakeStream {
  std::mutex mu;
  double ready_at = 0;  // completion time of the work enqueued so far
  // the stream is busy for cost seconds more (returns the completion time)
  double enqueue(double cost) {
    std::lock_guard<std::mutex> g(mu);
    ready_at = std::max(ready_at, now_s()) + cost;
    return ready_at;
  }
  double ready() {
    std::lock_guard<std::mutex> g(mu);
    return ready_at;
  }
};

struct FakeEvent {
  std::mutex mu;
  double t = 0;
};

FakeStream* fs(hipStream_t s) { return reinterpret_cast<FakeStream*>(s); }
FakeEvent* fe(hipEvent_t e) { return reinterpret_cast<FakeEvent*>(e); }

constexpr double kOpCost = 50e-6;      // seconds per operation
constexpr double kElemCost = 2e-9;     // seconds per element moved or computed

thread_local std::string g_error;

}  // namespace

extern "C" {

hipError_t hipGetDevice(int* d) {
  *d = 0;
  return hipSuccess;
}
hipError_t hipSetDevice(int) { return hipSuccess; }
hipError_t hipMalloc(void** p, size_t n) {
  *p = std::calloc(1, n ? n : 1);
  return *p ? hipSuccess : hipErrorOutOfMemory;
}
hipError_t hipFree(void* p) {
  std::free(p);
  return hipSuccess;
}
hipError_t hipHostMalloc(void** p, size_t n, unsigned int) { return hipMalloc(p, n); }
hipError_t hipHostFree(void* p) { return hipFree(p); }
hipError_t hipStreamCreateWithFlags(hipStream_t* s, unsigned int) {
  *s = reinterpret_cast<hipStream_t>(new FakeStream());
  return hipSuccess;
}
hipError_t hipStreamDestroy(hipStream_t s) {
  delete fs(s);
  return hipSuccess;
}
hipError_t hipStreamSynchronize(hipStream_t s) {
  const double t = fs(s)->ready();
  while (now_s() < t) std::this_thread::sleep_for(std::chrono::microseconds(10));
  return hipSuccess;
}
hipError_t hipEventCreateWithFlags(hipEvent_t* e, unsigned int) {
  *e = reinterpret_cast<hipEvent_t>(new FakeEvent());
  return hipSuccess;
}
hipError_t hipEventDestroy(hipEvent_t e) {
  delete fe(e);
  return hipSuccess;
}
hipError_t hipEventRecord(hipEvent_t e, hipStream_t s) {
  const double t = fs(s)->ready();
  std::lock_guard<std::mutex> g(fe(e)->mu);
  fe(e)->t = t;
  return hipSuccess;
}
hipError_t hipEventQuery(hipEvent_t e) {
  std::lock_guard<std::mutex> g(fe(e)->mu);
  return now_s() >= fe(e)->t ? hipSuccess : hipErrorNotReady;
}
hipError_t hipMemcpy(void* d, const void* s, size_t n, hipMemcpyKind) {
  std::memmove(d, s, n);
  return hipSuccess;
}
hipError_t hipMemcpyAsync(void* d, const void* s, size_t n, hipMemcpyKind, hipStream_t st) {
  std::memmove(d, s, n);
  fs(st)->enqueue(kOpCost + kElemCost * double(n / 8));
  return hipSuccess;
}

}  // extern "C"

namespace ff {

ff_status set_error(ff_status st, const std::string& msg) {
  g_error = msg;
  return st;
}
ff_status hip_check(hipError_t e, const char* what) {
  return e == hipSuccess ? FF_OK : set_error(FF_ERR_HIP, what);
}
void clear_error() { g_error.clear(); }

size_t reference_bytes(const int n[3], const int[3]) {
  const size_t N = (size_t)n[0] * n[1] * n[2];
  const size_t Nh = (size_t)n[0] * n[1] * (n[2] / 2 + 1);
  return std::max(N * 8, Nh * 16);
}

ff_status plan_create(const int n[3], const double lo[3], const double hi[3], const ff_options* opts, ff_plan** out,
                      bool) {
  ff_plan* p = new ff_plan();
  for (int a = 0; a < 3; ++a) {
    p->n[a] = n[a];
    p->lo[a] = lo[a];
    p->hi[a] = hi[a];
  }
  p->nh = n[2] / 2 + 1;
  p->N = (long long)n[0] * n[1] * n[2];
  p->Nh = (long long)n[0] * n[1] * p->nh;
  if (opts) p->opt = *opts;
  *out = p;
  return FF_OK;
}
void plan_destroy(ff_plan* p) { delete p; }
bool z_rows_four_step(const int[3], int) { return false; }
size_t z_work_bytes(const int[3]) { return 0; }
void plan_bind_scratch(ff_plan* p, void* base) { p->scratch = base; }

// Launch recording: the stand-in solve records nothing, so the server's
// grouped solves fall back to solving requests one by one.
thread_local std::vector<PassLaunch>* t_rec = nullptr;
void set_launch_recorder(std::vector<PassLaunch>* rec) { t_rec = rec; }
bool launch_recording() { return t_rec != nullptr; }
hipError_t replay_launch(const PassLaunch&) { return hipErrorInvalidValue; }
PassInst launch_inst_fields(const PassLaunch&) { return PassInst{}; }
bool launch_same(const PassLaunch&, const PassLaunch&) { return false; }

ff_status solve_enqueue(const ff_plan* p, double* density, double* potential, hipStream_t s, hipEvent_t*) {
  if (t_rec) return FF_OK;  // recording: nothing runs, nothing is recorded
  for (long long i = 0; i < p->N; ++i) potential[i] = 0.5 * density[i] + 1.0;
  fs(s)->enqueue(kOpCost + kElemCost * double(p->N) * 5);
  return FF_OK;
}

ff_status field_check(const ff_field* f) {
  if (!f) return set_error(FF_ERR_INVALID, "ff_field: null descriptor");
  for (int a = 0; a < 3; ++a)
    if (f->n[a] < 1) return set_error(FF_ERR_INVALID, "ff_field: extents must be >= 1");
  return FF_OK;
}
ff_status field_init(const ff_field* f, double* out, hipStream_t s) {
  const long long N = (long long)f->n[0] * f->n[1] * f->n[2];
  for (long long i = 0; i < N; ++i) out[i] = double(i);
  fs(s)->enqueue(kOpCost + kElemCost * double(N));
  return FF_OK;
}
ff_status field_diff_enqueue(const ff_field* f, const double* phi, double* work, hipStream_t s) {
  // sums of |e|, e^2, max|e| (phi given) or of the exact potential (null)
  double* r = work + kDiffResult;
  r[0] = phi ? 1.0 : 2.0;
  r[1] = phi ? 4.0 : 4.0;
  r[2] = phi ? 2.0 : 4.0;
  const long long N = (long long)f->n[0] * f->n[1] * f->n[2];
  fs(s)->enqueue(kOpCost + kElemCost * double(N));
  return FF_OK;
}

}  // namespace ff
