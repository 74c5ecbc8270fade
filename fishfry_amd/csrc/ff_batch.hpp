// Batched solves: many independent grids (uneven shapes, FP64 or FP32 +
// refinement each) in one persistent launch.
//
// Every instance's solve is a chain of pass ops (the same five passes as the
// single-grid solve, plus ten per refinement step).  The ops of all instances
// are listed step-major -- step 0 of every instance, then step 1, ... -- and
// each op is cut into tiles of a few lines.  Workgroups of the persistent
// kernel dequeue (op, tile) items from one atomic counter in that order; an
// item of step s waits until every tile of its instance's step s-1 is done.
// Because items are handed out in step-major order, every item an item waits
// on was dequeued earlier by a running workgroup that never waits on a later
// item, so the waits always end; a spin limit turns a broken invariant into an
// error flag instead of a hang.
//
// Tiles are transformed by an LDS-resident Stockham engine with a runtime radix
// sequence, so one kernel handles every transform length.
#pragma once

#include "ff_fft.hpp"

namespace ff {

enum { BOP_ZF = 0, BOP_COL = 1, BOP_ZI = 2 };

// BatchCore flag: one launch per pass step instead of one persistent launch.
constexpr int kBatchPerStep = 2;

struct BLine {
  int n, np;
  int radix[FF_MAXP];
  int nr[FF_MAXP];          // n / radix: butterflies per line
  int ns[FF_MAXP];
  uint32_t nsinv[FF_MAXP];
  int twstep[FF_MAXP];
  uint32_t nrinv[FF_MAXP];  // ceil(2^32 / (n / radix)): butterfly -> line
};

struct BOp {
  int kind;       // BOP_*
  int mode;       // BOP_COL: COL_FWD/INV/FILTER/SYMBOL; BOP_ZI: EPI_*
  int prec;       // 0 FP64, 1 FP32 (storage of the half spectrum and arithmetic)
  int io32;       // BOP_ZF: real input is float; BOP_ZI: real output is float
  int odd;        // nz odd (BOP_ZF / BOP_ZI)
  int T;          // lines per tile
  int lgT;        // BOP_COL: T = 1 << lgT
  int ntiles;     // tiles of this op
  int dep;        // op index this op waits for (-1: none)
  int tiles_pp;   // BOP_COL: tiles per plane
  int nx, ny, nz, nh;
  int pitch;      // LDS line pitch (elements)
  long long nlines;     // lines of the op
  long long s_outer;    // BOP_COL: plane stride (complex elements)
  long long s_r;        // BOP_COL: stride along the transform
  const void* src;
  void* dst;
  const double* aux;    // f (residual epilogue)
  double* dc_out;       // SYMBOL: sum of the input
  const double* shiftp; // ACCUM: sum of phi
  double scale;         // 1/(nx ny nz)
  double ddi, ddj, ddk;
  const void* tw;       // W_n of the line transform
  const void* twpost;   // exp(-2 pi i k/nz), k < nz/2
  BLine line;
};

struct BatchDev {
  const BOp* ops;
  const long long* item_start;  // nops + 1 prefix sums of ntiles
  int nops;
  unsigned int* done;            // per op completed tiles
  unsigned long long* head;      // next item of this launch (relative to item_begin)
  unsigned int* err;             // set when a wait exceeds the spin limit
  long long item_begin, item_end;
  long long spin_limit;  // longest dependency wait, in 100 MHz real-time ticks
  int lds_bytes;
};

}  // namespace ff

#include <vector>

#include "../../include/fishfry.h"

namespace ff {

// Resident twiddle tables shared by every batch and the server (per device,
// created on first use, kept until process exit).
const void* twiddle_table(int n, int prec);  // W_n[t], t < n
const void* twiddle_post(int nz, int prec);  // exp(-2 pi i k/nz), k < nz/2

size_t batch_item_bytes(const ff_batch_item& it);    // density / potential bytes
size_t batch_scratch_bytes(const ff_batch_item& it); // FP32 refinement work area

// A batch's host-side state: validated items, the op list in step-major
// order and the image of its device descriptor block.  Device memory (the
// descriptor block and the refinement work areas) is provided by the owner.
struct BatchCore {
  int count = 0;
  int flags = 0;
  int budget_kib = 32;
  int strategy = 0;
  std::vector<ff_batch_item> items;
  std::vector<long long> N, Nh;
  std::vector<size_t> bytes, scratch_off;
  size_t scratch_bytes = 0;
  std::vector<BOp> ops;
  std::vector<long long> item_start, step_start;
  int lds_bytes = 0, grid = 0, nsteps = 0;
  std::vector<unsigned char> image;  // ops | item_start, uploaded to dev
  size_t o_is = 0, o_done = 0, o_head = 0, o_err = 0, desc_bytes = 0;
  void* dev = nullptr;

  ff_status init(int count, const ff_batch_item* items, int flags);
  // Builds the op chains over the given buffers and work area (host only).
  ff_status build(double* const* density, double* const* potential, void* scratch);
  // Copies the descriptor image to dev (desc_bytes bytes) on stream s.
  ff_status upload(void* dev, hipStream_t s);
  BatchDev view() const;
  ff_status enqueue(hipStream_t s) const;
};

}  // namespace ff
