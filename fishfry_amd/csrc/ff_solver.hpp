// The plan object behind the C-ABI handle and the pass sequencer.
#pragma once

#include "ff_internal.hpp"

struct ff_plan {
  int n[3];
  double lo[3], hi[3];
  int nh;
  int lgKT;      // FP64 spectrum between the y passes: 0 plain, else kh blocks of 2^lgKT
  int pencil;    // FF_PENCIL_LAYOUT (FP64 passes)
  int slab;      // FF_SLAB_PIPELINE: x-planes per slab (0: whole passes)
  long long N;   // nx*ny*nz
  long long Nh;  // nx*ny*nh
  double ddi, ddj, ddk;
  // finite-difference symbol (FF_SYMBOL_3PT / _5PT): order 3 or 5, 0 for the
  // spectral filter; 1/h^2 per axis and the per-axis factor tables
  // 4 sin^2(pi k/n) [(1 + sin^2(pi k/n)/3)] (resident cache, FP64)
  int sym;
  double fd[3];
  const double* symtab[3];
  size_t bytes;
  ff_options opt;
  int device;
  // pass configurations per precision (0 = FP64, 1 = FP32)
  ff::RowCfg zr[2];  // z forward rows
  ff::RowCfg zi[2];  // z inverse rows
  ff::ColCfg cy[2], cx[2];
  const void* tw_z[2];   // z forward engine twiddles
  const void* tw_zi[2];  // z inverse engine twiddles (its E, hence m, may differ)
  const void* tw_zpost[2];
  const void* tw_y[2];
  const void* tw_x[2];
  // Bluestein tables (chirp, forward and inverse kernel spectra) of the z, y
  // and x line transforms; null when the line length has a direct plan
  // (all tables live in the resident per-device cache; the plan owns none)
  const void* bl_z[2][3];
  const void* bl_zi[2][3];
  const void* bl_y[2][3];
  const void* bl_x[2][3];
  // FP32 + refinement work area: FP64 spectrum (also hosts the FP32 one),
  // the FP32 residual and one double for the mean of phi.
  void* scratch;
  size_t scratch_bytes;
  void* spec;   // 16*Nh bytes
  float* res;   // N floats
  double* dc;   // sum of phi (refinement mean removal)
  double* dcf;  // sum of f (removed from the residual before it is rounded to FP32)
  // FP64 with FF_PRESERVE_INPUT or four-step z rows: third work array
  // (16*Nh bytes)
  void* work;
  // Four-step lines: an x or y extent n that no workgroup transforms whole
  // runs as n1-point and n2-point sub-line passes, n = n1*n2; so do the
  // packed z rows (n = nz/2, even nz, FP64 plans) when nz/2 is too long.
  struct FourStep {
    int n, n1, n2;              // n = 0: the axis is transformed whole
    ff::ColCfg c1[2], c2[2];    // n1-point and n2-point sub-lines (generic kernels)
    const void* tw1[2];         // engine twiddles of the sub-lines
    const void* tw2[2];
    const void* twn[2];         // W_n, t < n
    const void* bl1[2][3];      // Bluestein tables of the sub-lines
    const void* bl2[2][3];
  } fy, fx, fz;
  // four-step z rows of odd nz: the complex row of two real rows (2p, 2p+1),
  // nz long, is transformed (fz.n = nz) and the two spectra separated
  int zodd;
  // odd nz with nx*ny odd: the last row has no partner and runs alone, with
  // a zero second row, through the work array's tail (z_work_bytes)
  int zlone;
};

namespace ff {

// Pass kernels over caller-provided buffers.  prec: 0 FP64, 1 FP32.
// Row passes over rows [row0, row0 + nrows) of i*ny + j (nrows < 0: all rows),
// column passes over planes [outer0, outer0 + nouter) (nouter < 0: all);
// ranges apply to the plain spectrum layout.
ff_status run_zfwd(const ff_plan* p, int prec, int in_f32, const void* src, void* dst, hipStream_t s,
                   long long row0 = 0, long long nrows = -1);
// Half-spectrum layouts (see SpecSide): plain [i][j][kh], y-blocked
// [kb][i][j][kt], x-blocked [kb][j][i][kt].
// spectrum layouts: plain [i][j][kh]; blocked y [kb][i][j][kt] and x
// [kb][j][i][kt]; pencil A [i][kh][j], B [kh][j][i] and D [kh][i][j]
enum { LAY_PLAIN = 0, LAY_Y = 1, LAY_X = 2, LAY_A = 3, LAY_B = 4, LAY_D = 5 };
ff_status run_col(const ff_plan* p, int prec, int axis, int mode, const void* src, void* dst, hipStream_t s,
                  double* dc_out = nullptr, int lay_in = LAY_PLAIN, int lay_out = LAY_PLAIN, int outer0 = 0,
                  int nouter = -1);
ff_status run_zinv(const ff_plan* p, int prec, int out_kind, int epi, const void* src, void* dst, const double* f,
                   const double* shiftp, hipStream_t s, int lay_in = LAY_PLAIN, long long row0 = 0,
                   long long nrows = -1);
// Layouts the FP64 solve's column passes read and write: y forward, x, y inverse
// (and the z inverse reads what the y inverse wrote).
void pass_layouts(const ff_plan* p, int pass, int& lay_in, int& lay_out);

// The whole solve; ev (optional) receives one record after each of the five
// FP64 passes.
ff_status solve_enqueue(const ff_plan* p, double* density, double* potential, hipStream_t s, hipEvent_t* ev);

// own_scratch false: an FP32 plan gets no refinement work area of its own and
// must be bound to one (plan_bind_scratch) before it is enqueued, and an FP64
// plan with four-step z rows no work array (the caller sets work); callers
// that sub-allocate device memory themselves (the server's arena) use this.
ff_status plan_create(const int n[3], const double lo[3], const double hi[3], const ff_options* opts,
                      ff_plan** out, bool own_scratch = true);
void plan_destroy(ff_plan* p);
// Whether an FP64 grid of extents n runs its z rows in four steps, which
// needs the work array (z_work_bytes; the plan's own unless created with
// own_scratch false).
bool z_rows_four_step(const int n[3], int precision);
// The work array of four-step z rows: the rows of M (16*Nh bytes) and, for
// odd nz with nx*ny odd, a tail for the lone last row (its packed row, 16*nz
// bytes, and its pair region, 32*nh bytes), each part 256-byte aligned.
size_t z_work_bytes(const int n[3]);
// The size of a plan's work array (0: none).
size_t plan_work_bytes(const ff_plan* p);
// FP32 refinement work area: its size, and pointing a plan (or a copy of one)
// at a work area of that size.
size_t plan_scratch_bytes(const ff_plan* p);
void plan_bind_scratch(ff_plan* p, void* base);
size_t reference_bytes(const int n[3], const int m[3]);
// The FP64 stencil residual of a finite-difference plan: r = (f - sum_f/N) -
// L_h phi stored as float, and sum(phi) into *sum_phi (deterministic two-stage
// reduction through `partial`, at least nx*ny doubles).  With r_f64 set, out
// receives L_h phi in FP64 instead (f, sum_f unused).
ff_status run_stencil(const ff_plan* p, const double* phi, const double* f, void* out, bool r_f64,
                      const double* sum_f, double* partial, double* sum_phi, hipStream_t s);

}  // namespace ff
