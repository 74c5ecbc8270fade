// Host-side internals shared by the library's translation units.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/fishfry.h"
#include "ff_kernels.hpp"

namespace ff {

// ---- errors ---------------------------------------------------------------
ff_status set_error(ff_status st, const std::string& msg);
ff_status hip_check(hipError_t e, const char* what);
void clear_error();

// Library-owned device memory for small reductions (one block per device,
// grown on demand, kept until process exit).  Callers hold scratch_mutex()
// while the memory is in use.
ff_status device_scratch(size_t ndoubles, double** out);
std::mutex& scratch_mutex();

// The static device work area of ff_solve_noalloc (ff_static.hip): its size,
// and its address on a device (looked up once per device, which must be the
// current one on that first call).
constexpr size_t kStaticWorkBytes = size_t(64) << 20;
ff_status static_work_area(int device, void** out);

// ---- line transforms --------------------------------------------------------
// Supported elements-per-thread values (one kernel instantiation unit each).
constexpr int kEValues[] = {1, 2, 3, 4, 5, 6, 8, 10, 12, 15, 16, 20, 24, 30};

struct LineCfg {
  int n = 0;      // transform length
  int E = 0;
  LinePlan lp{};  // engine transform: length n, or for Bluestein the 2^a 3^b 5^c length m >= 2n - 1
  bool bluestein() const { return lp.n != n; }
};

// Chooses E and a radix sequence for a length-n transform.  A length with a
// prime factor above 5 is transformed by Bluestein's algorithm over the
// 2^a 3^b 5^c length m >= 2n - 1 with a plan and the least work m * (passes +
// 0.04 E) (forceE then applies to m).  False when no plan fits one workgroup
// per line.
bool plan_line(int n, LineCfg& out, int forceE = 0);

// Bluestein tables of a length-n transform computed over length m, in
// precision prec (0 FP64, 1 FP32), appended to buf at 256-byte aligned
// offsets: the chirp w[t] = exp(-pi i t^2/n) (t < n) and, for each direction,
// the length-m spectrum of the convolution kernel divided by m (forward:
// b[t] = conj(w[|t|]); inverse: b[t] = w[|t|]; |t| < n, indices mod m).
void push_bluestein(std::vector<unsigned char>& buf, int n, int m, int prec, size_t& chirp, size_t& bsf, size_t& bsi);

// Device-resident tables shared by every plan, batch and server on the
// current device (built on first use, kept until the process ends; null /
// false when the allocation fails): W_n[t] (t < n); the even-nz post-twiddles
// exp(-2 pi i k/nz) (k < nz/2); the Bluestein tables of length n over m
// (chirp, forward and inverse kernel spectra).
const void* table_twiddle(int n, int prec);
const void* table_post(int nz, int prec);
bool table_bluestein(int n, int m, int prec, const void* out[3]);
// Finite-difference factor table g[k] = 4 sin^2(pi k/n) (order 3) or
// 4 sin^2(pi k/n) (1 + sin^2(pi k/n)/3) (order 5), k < n, FP64.
const double* table_symbol(int n, int order);
long double fd_factor_ld(long long k, long long n, int order);

// exp(-2 pi i t / n) in long double with exact integer argument reduction.
void twiddle_ld(long long t, long long n, long double& c, long double& s);

struct ColCfg {
  LineCfg line;
  int T = 1, lgT = 0, ps = 30;
  int split = 0;   // FF_TUNE_SPLIT_LDS: one component per LDS slot
  int groups = 1;  // FF_TUNE_TWO_GROUPS: column groups of T per workgroup
  int stat = 0;    // a kStaticCols kernel runs this configuration (FF_TUNE_GENERIC clears it):
                   // 1 twiddles from the table in memory, 2 from LDS (FF_TUNE_LDS_TWIDDLES)
  size_t lds = 0;
  int threads = 0;
};

struct RowCfg {
  LineCfg line;  // transform length nz/2 (even nz) or nz (odd nz)
  int T = 1, ps = 30, pitch = 0;
  size_t lds = 0;
  int threads = 0;
  uint32_t Linv = 0;
  int stat = 0;  // a kStaticRows kernel runs this configuration (FF_TUNE_GENERIC clears it);
                 // 2: the one that reads its twiddles from LDS (FF_TUNE_LDS_TWIDDLES)
};

// kind: 0 z forward rows, 1 y columns, 2 x columns, 3 z inverse rows;
// prec: 0 FP64, 1 FP32.  Tuning
// overrides (ff_tune_set) take precedence over the built-in heuristic.
// pencil: the plan runs the pencil spectrum layout (the FP64 column passes
// transpose as they transform), which has its own measured defaults
bool make_col_cfg(int n, int ncols, int elem_bytes, ColCfg& out, int kind, int prec, bool pencil = false);
bool make_row_cfg(int nz, int elem_bytes, RowCfg& out, int kind, int prec);
bool col_cfg_with(int n, int ncols, int elem_bytes, ColCfg& out, int E, int T, int flags, int kind);
bool row_cfg_with(int nz, int elem_bytes, RowCfg& out, int E, int T, int flags);
bool tune_lookup(int kind, int n, int prec, int& E, int& T, int& flags);
ff_status tune_set(const ff_tune* t);
void tune_reset();

hipError_t launch_col(int E, int f32, int mode, int split, int groups, int blue, int stat, const Launch& l,
                      const ColArgs& a);

// Launch recording.  While a recorder is installed on the calling thread, the
// three pass launchers append their launches to it instead of launching; a
// recorded launch can be replayed later, e.g. with a per-instance table and
// one grid row per instance (the batch's grouped mode).
struct PassLaunch {
  int kind;  // 0 column pass, 1 z forward, 2 z inverse
  int E, f32, sel, split, groups, blue, stat, odd, epi;  // sel: mode / in_f32 / out_kind
  Launch l;
  ColArgs col;
  RowArgs row;
};
void set_launch_recorder(std::vector<PassLaunch>* rec);
hipError_t replay_launch(const PassLaunch& r);
// The per-instance fields of a recorded launch, and whether two recorded
// launches differ only in those fields and their stream (so that one launch
// with a per-instance table runs both).
PassInst launch_inst_fields(const PassLaunch& r);
bool launch_same(const PassLaunch& x, const PassLaunch& y);
hipError_t launch_r2c(int E, int f32, int in_f32, int odd, int blue, int stat, const Launch& l, const RowArgs& a);
hipError_t launch_c2r(int E, int f32, int out_kind, int odd, int epi, int blue, int stat, const Launch& l,
                      const RowArgs& a);
// whether a recorder is installed on the calling thread
bool launch_recording();

// Row-local steps of z passes over four-step rows (ff_zsplit.hip; f32:
// FP32 data): the forward split (in place on io.dst), the inverse merge
// (src -> dst), the inverse epilogue (packed rows src -> real dst, out_kind
// and epi as for launch_c2r), the four-step twiddle between the sub-line
// passes (forward or conjugate, in place on io.dst; a.tw = W_M, a.m1) and
// the rounding of FP64 real rows to FP32 (src -> dst).  a.nline = M = nz/2.
// Odd nz (row pairs, a.nrows pairs, a.nline = nz): the packing of real rows
// (2p, 2p+1) into one complex row (src -> dst; out_kind: FP32 input), the
// separation of the two spectra (in place on io.dst, pair stride 2*nh), the
// merge of two half spectra into the complex row of a pair (src -> dst) and
// the unpacking epilogue.  Never recorded: refused while a recorder is
// installed.
enum {
  ZSPLIT_FWD = 0,
  ZSPLIT_INV = 1,
  ZSPLIT_EPI = 2,
  ZSPLIT_TWF = 3,
  ZSPLIT_TWI = 4,
  ZSPLIT_PACK = 5,
  ZSPLIT_PPACK = 6,
  ZSPLIT_PFWD = 7,
  ZSPLIT_PINV = 8,
  ZSPLIT_PEPI = 9
};
hipError_t launch_zsplit(int step, int f32, int out_kind, int epi, const RowArgs& a, hipStream_t s);

// ---- driver kernels ---------------------------------------------------------
ff_status field_check(const ff_field* f);
ff_status field_init(const ff_field* f, double* out, hipStream_t s);
ff_status field_diff(const ff_field* f, const double* phi, double out[3], hipStream_t s);
// Enqueues the same reduction on s over a caller-provided device work area of
// kDiffDoubles doubles; the three sums land at work + kDiffResult.
constexpr int kDiffDoubles = 3 * 1024 + 3;
constexpr int kDiffResult = 3 * 1024;
ff_status field_diff_enqueue(const ff_field* f, const double* phi, double* work, hipStream_t s);

}  // namespace ff
