// Multi-shape pass launches: one kernel launch runs the same pass step of many
// independent solves whose grids differ in shape.
//
// Each solve's pass is described by a descriptor: the pass arguments it would
// be launched with alone (ColArgs / RowArgs, including its buffers and domain
// constants), the compiled-in variant that transforms its lines and the first
// workgroup of its range.  The descriptors of one launch are ordered so their
// first workgroups ascend; a workgroup finds its descriptor by counting the
// first workgroups at or below its own index (one wave-wide compare and
// ballot per 64 descriptors), then runs that solve's tile with the variant's
// line plan and tile width compiled in.  No workgroup waits on another: the
// launch boundary between steps orders the passes of each solve.
//
// A variant is a line length n with elements per thread E (the plan of
// (n, E) compiled in, as for the single-solve kernels); its tile is
// kMultiThreads / (n / E) columns or rows, so every variant runs the same
// workgroup size.
#pragma once

#include "ff_kernels.hpp"

namespace ff {

constexpr int kMultiThreads = 128;

struct MultiVar {
  int n, E;
};
// Column passes (y and x): line lengths of the batching configurations.
constexpr MultiVar kMultiCols[] = {{32, 8}, {48, 12}, {64, 8}, {96, 24}, {128, 16}, {160, 20}, {256, 16}};
constexpr int kNumMultiCols = sizeof(kMultiCols) / sizeof(kMultiCols[0]);
// Row passes (z, even nz): transform length M = nz / 2.
constexpr MultiVar kMultiRows[] = {{16, 16}, {24, 12}, {32, 8}, {48, 12}, {64, 8}, {80, 20}, {128, 16}};
constexpr int kNumMultiRows = sizeof(kMultiRows) / sizeof(kMultiRows[0]);

// Columns (or rows) per workgroup of a variant.
__host__ __device__ constexpr int multi_tile(const MultiVar& v) { return kMultiThreads / (v.n / v.E); }

struct MultiColDesc {
  ColArgs a;    // the pass as launched alone, with the variant's tile (T, tiles, nblocks)
  int variant;  // index into kMultiCols
  int pad;
};
struct MultiRowDesc {
  RowArgs a;
  int variant;  // index into kMultiRows
  int pad;
};

// Workgroups are dispatched round-robin over the 8 XCDs (workgroup b runs on
// XCD b mod 8).  The launch's logical tiles are cut into 8 contiguous chunks
// of about equal work (tiles of different variants differ in size), chunk x
// running on XCD x in order, so neighbouring tiles of one solve share an L2
// and no XCD idles while another still works; the grid is 8 times the largest
// chunk, and workgroups past their chunk's end return at once.
constexpr int kXcds = 8;

struct MultiArgs {
  const void* desc;    // MultiColDesc[] or MultiRowDesc[]
  const int* block0;   // first logical tile of each descriptor, ascending (block0[0] = 0)
  int ndesc;
  int nblocks;         // workgroups of the launch (kXcds * the largest chunk)
  int chunk0[kXcds];   // first logical tile of each XCD's chunk
  int chunkn[kXcds];   // tiles in each XCD's chunk
};

// The kernels are compiled in kMultiParts units (ff_multi.hip); unit PART
// returns its kernel for a role, or null.
constexpr int kMultiParts = 5;
template <int PART>
const void* multi_fn(int kind, int f32, int sel, int epi);

// Kernel roles, as the single-solve launchers select them.
enum { MULTI_COL = 0, MULTI_R2C = 1, MULTI_C2R = 2 };

// Host side.  Variant of a line (n, E) whose runtime plan equals the
// compiled-in plan of that variant, or -1.
int multi_col_variant(int n, int E, const LinePlan& lp);
int multi_row_variant(int M, int E, const LinePlan& lp);
// Whether a multi-shape kernel exists for the role: kind MULTI_COL (sel =
// mode), MULTI_R2C (sel = real input is float), MULTI_C2R (sel = real output
// is float, epi = epilogue).
bool multi_kernel_exists(int kind, int f32, int sel, int epi);
hipError_t launch_multi(int kind, int f32, int sel, int epi, const MultiArgs& m, size_t lds, hipStream_t s);

}  // namespace ff
