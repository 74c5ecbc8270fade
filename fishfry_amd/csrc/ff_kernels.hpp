// Pass kernels of the periodic Poisson solve.
//
// Data layout in HBM: the half spectrum leaves the z forward pass as complex
// [i][j][kh], kh fastest, row pitch nh = nz/2+1 -- the layout the reference
// has after its D2Z (ParisPencil.cpp:165).  In the plain layout every pass
// keeps it; in the pencil layout each column pass reads strided columns and
// writes its lines contiguously in a transposed order ([kh][i][j], [kh][j][i],
// [i][kh][j]; SpecSide describes each).  The z passes work on rows; the y and
// x passes on strided columns, a workgroup owning a tile of T neighbouring
// columns so every row of the tile is one contiguous T*sizeof(complex)
// segment in HBM.
//
//   k_r2c   z forward, real rows -> half-spectrum rows       (ParisPencil.cpp:165)
//   k_col   y forward / inverse, or x forward*filter*inverse (:212, :259-288, :334)
//   k_c2r   z inverse + the 1/(nx*ny*nz) scale                (:380, :406-421)
#pragma once

#include <type_traits>

#include <climits>

#include "ff_fft.hpp"

namespace ff {

// Workgroups are dispatched round-robin over the 8 XCDs; remap so that each
// XCD walks a contiguous range of tiles (neighbouring tiles share L2 lines).
__device__ __forceinline__ int xcd_remap(int b, int nb) {
  const int q = nb >> 3, r = nb & 7;
  const int x = b & 7, s = b >> 3;
  return x * q + (x < r ? x : r) + s;
}

__device__ __forceinline__ double sqr(double x) { return x * x; }

// Half-spectrum layouts.  Element (i, j, kh) lives at complex offset
//   plain      [i][j][kh]:      (i*ny + j)*nh + kh     (the reference's, after its D2Z)
//   y-blocked  [kb][i][j][kt]:  kb*KS + (i*ny + j)*W + kt
//   x-blocked  [kb][j][i][kt]:  kb*KS + (j*nx + i)*W + kt
// with kh = kb*KT + kt, KS = nx*ny*KT and W = KT except for the last block,
// whose width is nh - (nb-1)*KT (so every layout holds exactly nx*ny*nh).
// The blocked layouts let a column pass write each tile of KT columns as one
// contiguous region; the strided side of a pass is then always a read.
//
// A SpecSide addresses one layout as seen by a column pass with planes
// `outer` and transform positions r:
//   offset(outer, r, col) = kb*KS + (outer*fo + r*fr)*W + kt,
//   kb = col >> lgKT, kt = col & (KT-1).
// The plain layout is the single-block case (KT >= nh, W = nh).  With
// lgKT = 0 (blocks of one column, W = 1) it is any strided layout,
//   offset = col*KS + outer*fo + r*fr,
// which the pencil layouts use (FF_PENCIL_LAYOUT): [i][kh][j] after the y
// passes and [kh][j][i] after the x pass, so every column pass writes its
// lines contiguously.
// Four-step sub-lines (generic kernels only): with nlo > 0 a plane index is
// the pair (outer % nlo, outer / nlo) with strides fo and fo2.
struct SpecSide {
  long long KS;
  long long fo, fr;
  int lgKT;
  int nb, Wlast;
  int nlo;
  long long fo2;
};

// Offset of (outer, r0, col) and the offset step between positions r and r+L.
template <bool FS = false>
__device__ __forceinline__ void side_addr(const SpecSide& s, int outer, int col, int r0, int L, long long& base,
                                          long long& step) {
  const int kb = col >> s.lgKT;
  const int kt = col & ((1 << s.lgKT) - 1);
  const long long W = kb == s.nb - 1 ? (long long)s.Wlast : (1LL << s.lgKT);
  long long ot = (long long)outer * s.fo;
  if constexpr (FS)
    if (s.nlo) ot = (long long)(outer % s.nlo) * s.fo + (long long)(outer / s.nlo) * s.fo2;
  base = (long long)kb * s.KS + (ot + (long long)r0 * s.fr) * W + kt;
  step = (long long)L * s.fr * W;
}

// Per-instance fields of a grouped launch: grid row blockIdx.y runs instance
// blockIdx.y of a group of solves of one shape, precision and refinement
// count, whose other pass arguments are all equal (their buffers and domain
// constants differ).
struct PassInst {
  const void* src;
  void* dst;
  const void* aux;       // rows: f of the residual epilogue
  double* dc_out;        // columns: sum of the input
  const double* shiftp;  // rows: sum of phi
  double ddi, ddj, ddk;  // columns: filter constants
};

struct ColArgs {
  const void* src;
  void* dst;
  SpecSide in, out;   // layouts read and written
  int ncols;          // columns per plane (nh)
  int nouter;         // planes
  int outer_fast;     // workgroup order: 1 = consecutive workgroups take consecutive planes
  int T, lgT;         // columns per group (power of two); a tile is G groups
  int tiles;          // tiles per plane
  int nblocks;        // nouter * tiles
  int ps;             // LDS pad shift
  int nx, ny;         // extents for the spectral symbol (x pass)
  int cols_j;         // x pass of the pencil layout: columns are j and planes kh (else planes j, columns kh)
  // Four-step lines (generic kernels; the planes are (lo, hi) pairs, in.nlo > 0):
  // xtw: W_n, n = line length; a forward transform multiplies position r of
  // plane lo by W_n[lo*r] after the transform, an inverse one by its
  // conjugate before (0: no twiddle).  The x filter's i index is lo +
  // imap_n2*r where imap_n2 > 0, and its j index p (the plane, or the hi
  // part of it) is p / jmap_n1 + jmap_n2 * (p % jmap_n1) where jmap_n1 > 0
  // (y lines held in four-step order).
  const void* xtw;
  int imap_n2, jmap_n1, jmap_n2;
  // Alignment of the written row segments (kernels with a compiled-in plan).
  // When all rows of a plane start at one offset within a 64-byte block but
  // planes start at different ones (plain x pass, odd nh), a plane's tiles
  // are shifted left by sh = (element offset of its column 0) & cshift_mask
  // columns: every row segment a workgroup writes then starts on a segment
  // boundary and no 32-byte sector is written in parts by two workgroups.
  // cshift_mask = 0: no shift.
  int cshift_mul, cshift_mask;
  double ddi, ddj, ddk;  // filter constants (finite-difference symbol: 1/h^2 per axis)
  // finite-difference symbol: per-axis factor tables g[k] (null: the
  // spectral filter); the axis term is g[k] * dd
  const double *symx, *symy, *symz;
  double* dc_out;  // FILTER / SYMBOL: receives the forward (0,0,0) term (sum of the input)
  const void* tw;
  LinePlan lp;
  int nline;                     // transform length (lp.n: engine length; larger for Bluestein)
  const void *chirp, *bsf, *bsi;  // Bluestein tables (see line_dft)
  const PassInst* inst;           // grouped launch: per-instance fields (null: one solve)
};

// Buffers in global memory.  The per-instance fields come either from the
// kernel arguments or from a table in memory, so the compiler cannot infer
// their address space (it would emit flat accesses): the kernels address
// them through these global-space views.
// (The host pass of the HIP compilation parses these functions too; address
// spaces exist only in the device pass.)
#if defined(__HIP_DEVICE_COMPILE__)
#define FF_GLB __attribute__((address_space(1)))
#else
#define FF_GLB
#endif
template <typename T>
__device__ __forceinline__ FF_GLB T* glb(const void* p) {
  return (FF_GLB T*)(p);
}

// The per-instance fields are the same for the whole workgroup: keeping them
// in scalar registers keeps the address arithmetic scalar as well.
__device__ __forceinline__ unsigned long long uniform64(unsigned long long v) {
  const unsigned lo = __builtin_amdgcn_readfirstlane((unsigned)v);
  const unsigned hi = __builtin_amdgcn_readfirstlane((unsigned)(v >> 32));
  return ((unsigned long long)hi << 32) | lo;
}
template <typename P>
__device__ __forceinline__ P uniform_ptr(P p) {
  return reinterpret_cast<P>(uniform64(reinterpret_cast<unsigned long long>(p)));
}
__device__ __forceinline__ double uniform_dbl(double x) {
  return __longlong_as_double((long long)uniform64((unsigned long long)__double_as_longlong(x)));
}
__device__ __forceinline__ PassInst uniform_inst(const PassInst& t) {
  return PassInst{uniform_ptr(t.src),    uniform_ptr(t.dst),    uniform_ptr(t.aux), uniform_ptr(t.dc_out),
                  uniform_ptr(t.shiftp), uniform_dbl(t.ddi), uniform_dbl(t.ddj), uniform_dbl(t.ddk)};
}

// The per-instance fields of this workgroup's solve.
__device__ __forceinline__ PassInst inst_of(const ColArgs& a) {
  if (a.inst) return uniform_inst(glb<const PassInst>(a.inst)[blockIdx.y]);
  return PassInst{a.src, a.dst, nullptr, a.dc_out, nullptr, a.ddi, a.ddj, a.ddk};
}

// MODE: 0 forward, 1 inverse, 2 forward * filter * inverse (the solve),
//       3 forward * symbol * inverse (applies L_h for a residual).
//       4 forward * finite-difference filter * inverse (symbol from the
//         per-axis tables symx, symy, symz; a separate instantiation so the
//         spectral kernels carry none of its registers).
enum { COL_FWD = 0, COL_INV = 1, COL_FILTER = 2, COL_SYMBOL = 3, COL_FILTER_FD = 4 };

// Unnormalized DFT of direction DIR (-1 forward, +1 inverse) of a line of
// length nline held at positions r0 + i*L.  When nline equals the engine
// length lp.n this is the Stockham transform itself.  Otherwise (a length
// with a prime factor above 5) it is Bluestein's algorithm over the engine
// length m = lp.n >= 2*nline - 1: with w[t] = exp(-pi i t^2/n),
//   X[k] = w[k]^s * sum_j (x[j] w[j]^s) conj(w[k-j])^s,   s = 1 (forward) or conj (inverse),
// the sum being a length-m circular convolution done as forward transform,
// product with the kernel spectrum (bsf / bsi, already divided by m) and
// inverse transform.  Positions >= nline are zero on entry to the
// convolution and undefined on exit.
// SN > 0: the plan of length SN is fixed at compile time (lp is not read).
// TW: the twiddle source (tw_at); a table in global memory except where SN > 0.
template <int E, int DIR, typename F, typename Idx, bool SPLIT, bool BL, int SN = 0, typename TW>
__device__ __forceinline__ void line_dft(cx<F> (&v)[E], const LinePlan& lp, int nline, const TW& tw,
                                         const void* chirp_p, const void* bsf_p, const void* bsi_p, cx<F>* lds,
                                         const Idx& idx, int r0) {
  static_assert(SN == 0 || !BL, "compile-time plans are for direct lengths");
  if constexpr (SN > 0) {
    fft_line_static<SN, E, DIR, F, Idx, SPLIT>(v, tw, lds, idx, r0);
  } else if constexpr (!BL) {
    fft_line<E, DIR, F, Idx, SPLIT>(v, lp, tw, lds, idx, r0);
  } else {
    const cx<F>* __restrict__ chirp = reinterpret_cast<const cx<F>*>(chirp_p);
    const cx<F>* __restrict__ bs = reinterpret_cast<const cx<F>*>(DIR < 0 ? bsf_p : bsi_p);
    const int L = lp.L;
#pragma unroll
    for (int i = 0; i < E; ++i) {
      const int r = r0 + i * L;
      if (r < nline) {
        const cx<F> w = chirp[r];
        v[i] = twmul<DIR>(v[i], w);
      } else {
        v[i] = cx<F>{F(0), F(0)};
      }
    }
    fft_line<E, -1, F, Idx, SPLIT>(v, lp, tw, lds, idx, r0);
#pragma unroll
    for (int i = 0; i < E; ++i) v[i] = twmul<-1>(v[i], bs[r0 + i * L]);
    fft_line<E, +1, F, Idx, SPLIT>(v, lp, tw, lds, idx, r0);
#pragma unroll
    for (int i = 0; i < E; ++i) {
      const int r = r0 + i * L;
      if (r < nline) {
        const cx<F> w = chirp[r];
        v[i] = twmul<DIR>(v[i], w);
      }
    }
  }
}

// One column line held in registers through the transform, filter or symbol
// and the inverse transform of MODE.
template <typename F, int E, int MODE, bool SPLIT, bool BL, int SN, typename TW>
__device__ __forceinline__ void col_line(cx<F> (&v)[E], const ColArgs& a, const PassInst& io, const TW& tw,
                                         cx<F>* lds, const ColIdx& idx, int r0, int outer, int col, bool valid) {
  const int L = SN > 0 ? SN / E : a.lp.L;
  // four-step lines (generic kernels): plane (lo, hi), the twiddle table
  int plo = 0, phi = outer;
  if constexpr (SN == 0) {
    if (a.in.nlo) {
      plo = outer % a.in.nlo;
      phi = outer / a.in.nlo;
    }
  }
  // (the four-step twiddles of forward and inverse sub-lines are applied by
  // col_store and col_load, where the transform's registers are free)
  if constexpr (MODE == COL_FWD) {
    line_dft<E, -1, F, ColIdx, SPLIT, BL, SN>(v, a.lp, a.nline, tw, a.chirp, a.bsf, a.bsi, lds, idx, r0);
  } else if constexpr (MODE == COL_INV) {
    line_dft<E, +1, F, ColIdx, SPLIT, BL, SN>(v, a.lp, a.nline, tw, a.chirp, a.bsf, a.bsi, lds, idx, r0);
  } else {
    line_dft<E, -1, F, ColIdx, SPLIT, BL, SN>(v, a.lp, a.nline, tw, a.chirp, a.bsf, a.bsi, lds, idx, r0);
    // spectral filter of ParisPencil.cpp:269-284 at (i, j, kh) = (r, outer, col)
    // (pencil layout: (r, col, outer)), or the finite-difference symbol from
    // the per-axis tables.  Four-step lines: i = lo + imap_n2*r, and j
    // planes in four-step order map to their frequency.
    int jj = a.cols_j ? col : phi;
    const int kk = a.cols_j ? outer : col;
    int istep = 1;
    if constexpr (SN == 0) {
      if (a.jmap_n1) jj = jj / a.jmap_n1 + a.jmap_n2 * (jj % a.jmap_n1);
      if (a.imap_n2) istep = a.imap_n2;
    }
    const bool jv = a.cols_j ? valid : true, kv = a.cols_j ? true : valid;  // index in range
    // the (0,0,0) term before the multiply is the sum of the input field
    if (io.dc_out && valid && r0 == 0 && plo == 0 && jj == 0 && kk == 0) *glb<double>(io.dc_out) = double(v[0].x);
    if constexpr (MODE == COL_FILTER_FD) {
      // columns outside the plane and Bluestein positions past the line read
      // entry 0 (their values are never stored)
      const double tj = glb<const double>(a.symy)[jv ? jj : 0] * io.ddj;
      const double gk = glb<const double>(a.symz)[kv ? kk : 0];
#pragma unroll
      for (int i = 0; i < E; ++i) {
        const int ii = plo + istep * (r0 + i * L);
        double gi;
        if constexpr (is_tw_half<TW>::value)
          gi = reinterpret_cast<const double*>(tw.p + SN / 2)[ii];  // LDS copy (compiled-in: ii < SN = nx)
        else
          gi = glb<const double>(a.symx)[ii < a.nx ? ii : 0];
        const double s2 = fmadd(gk, io.ddk, fmadd(gi, io.ddi, tj));  // (i + j) + k
        const double d = (ii | jj | kk) ? -1.0 / s2 : 0.0;
        const F df = F(d);
        v[i] = cx<F>{df * v[i].x, df * v[i].y};
      }
    } else {
      const double j2 = sqr(double(jj < a.ny - jj ? jj : a.ny - jj) * io.ddj);
      const double dk = double(kk) * io.ddk;
#pragma unroll
      for (int i = 0; i < E; ++i) {
        const int ii = plo + istep * (r0 + i * L);
        const double di = double(ii < a.nx - ii ? ii : a.nx - ii) * io.ddi;
        const double s2 = fmadd(dk, dk, fmadd(di, di, j2));  // (i^2 + j^2) + k^2, as the reference sums
        double d;
        if constexpr (MODE == COL_FILTER) d = (ii | jj | kk) ? -1.0 / s2 : 0.0;
        else d = -s2;
        const F df = F(d);
        v[i] = cx<F>{df * v[i].x, df * v[i].y};
      }
    }
    if constexpr (SN > 0) {
      // With the plan compiled in, the inverse transform's twiddle loads and
      // LDS addresses are the forward's; left visible, the compiler keeps them
      // live across the filter and spills.  An opaque copy of the thread's
      // position (from which both derive) makes it recompute them.
      int r0b = r0;
      ColIdx idx2 = idx;
      asm volatile("" : "+v"(r0b), "+v"(idx2.c));
      line_dft<E, +1, F, ColIdx, SPLIT, BL, SN>(v, a.lp, a.nline, tw, a.chirp, a.bsf, a.bsi, lds, idx2, r0b);
    } else {
      line_dft<E, +1, F, ColIdx, SPLIT, BL, SN>(v, a.lp, a.nline, tw, a.chirp, a.bsf, a.bsi, lds, idx, r0);
    }
  }
}

// A workgroup owns a tile of G*T neighbouring columns of one plane: thread
// (c, r0) holds positions r0 + i*L of columns c + g*T (g < G).  The G groups
// are transformed one after the other through the same LDS (T columns of
// it), but all loads are issued first and all stores last, so each row of the
// tile -- G*T*sizeof(complex) contiguous bytes -- is read and written whole.
// SPLIT: exchange real and imaginary parts separately through half the LDS.
// BL: Bluestein lines (nline < lp.n); a separate instantiation so the common
// kernels carry none of its registers.
// SN, ST > 0: line length SN and ST columns per group fixed at compile time
// (LDS pad shift 4), for the lengths that dominate the solve; the workgroup is
// then exactly (SN/E)*ST threads and SW the waves per SIMD the register
// allocation must allow (the LDS-limited occupancy of the tile).
//
// A tile's work is split into load, transform and store (col_load,
// col_transform, col_store); col_tile does one tile (k_col, and the
// multi-shape kernels of ff_multi.hip).

// Tile b of a pass (b < a.nblocks): its plane, its index within the plane
// and the plane's column shift.
struct ColTile {
  int outer, tile, sh;
};

template <typename F, int SN>
__device__ __forceinline__ ColTile col_tile_at(const ColArgs& a, const PassInst& io, int b) {
  ColTile t;
  if (a.outer_fast) {
    t.tile = b / a.nouter;
    t.outer = b - t.tile * a.nouter;
  } else {
    t.outer = b / a.tiles;
    t.tile = b - t.outer * a.tiles;
  }
  // (the generic kernels never shift: the extra live column index costs
  // them an occupancy step)
  t.sh = 0;
  if constexpr (SN > 0)
    t.sh = (int)(((unsigned long long)(uintptr_t)io.dst / sizeof(cx<F>) + (unsigned long long)t.outer * a.cshift_mul) &
                 (unsigned long long)a.cshift_mask);
  return t;
}

// This thread's column within a group and first position.
template <int E, int SN, int ST>
__device__ __forceinline__ void col_thread(const ColArgs& a, int& c, int& r0) {
  const int T = SN > 0 ? ST : a.T;
  c = threadIdx.x & (T - 1);
  r0 = SN > 0 ? (int)threadIdx.x / ST : (int)threadIdx.x >> a.lgT;
}

// Columns below 0 exist only when shifted.
template <int SN>
constexpr int col_min() {
  return SN > 0 ? 0 : INT_MIN;
}

// Threads outside the columns load the nearest column (in bounds, never
// stored), so the loads need no predicate.  src may equal dst only when both
// sides use the same layout (in-place pass): every thread stores exactly the
// elements it loaded, after all of its loads.
// Four-step sub-lines (generic kernels): position r of plane lo times W_n[lo*r]
// (DIR < 0, after a forward transform) or its conjugate (before an inverse).
template <int DIR, typename F, int E, int G>
__device__ __forceinline__ void four_step_twiddle(cx<F> (&v)[G][E], const ColArgs& a, int outer, int r0, int L) {
  if (!a.xtw) return;
  const int lo = outer % a.in.nlo;
  const FF_GLB cx<F>* w = glb<const cx<F>>(a.xtw);
#pragma unroll
  for (int g = 0; g < G; ++g)
#pragma unroll
    for (int i = 0; i < E; ++i) {
      const int r = r0 + i * L;
      if (r < a.nline) v[g][i] = twmul<DIR>(v[g][i], cx<F>(w[lo * r]));
    }
}

template <typename F, int E, int G, bool BL, int SN, int ST, int MODE = COL_FWD>
__device__ __forceinline__ void col_load(cx<F> (&v)[G][E], const ColArgs& a, const PassInst& io, const ColTile& t) {
  constexpr int kMinCol = col_min<SN>();
  const int T = SN > 0 ? ST : a.T;
  const int L = SN > 0 ? SN / E : a.lp.L;
  int c, r0;
  col_thread<E, SN, ST>(a, c, r0);
#pragma unroll
  for (int g = 0; g < G; ++g) {
    const int col = (t.tile * G + g) * T + c - t.sh;
    long long ib, is;
    side_addr<SN == 0>(a.in, t.outer, col < kMinCol ? kMinCol : (col < a.ncols ? col : a.ncols - 1), r0, L, ib, is);
    const FF_GLB cx<F>* src = glb<const cx<F>>(io.src) + ib;
    if constexpr (!BL) {
#pragma unroll
      for (int i = 0; i < E; ++i) v[g][i] = src[(long long)i * is];
    } else {  // Bluestein: the engine line is longer than the data line
#pragma unroll
      for (int i = 0; i < E; ++i) v[g][i] = r0 + i * L < a.nline ? cx<F>(src[(long long)i * is]) : cx<F>{F(0), F(0)};
    }
  }
  if constexpr (SN == 0 && MODE == COL_INV) four_step_twiddle<+1>(v, a, t.outer, r0, L);
}

// LDS bytes of the exchange of a compiled-in column tile (SN positions, ST
// columns, a pad slot every 16 positions), rounded up to 16.
template <typename F, bool SPLIT, int SN, int ST>
constexpr int col_exchange_bytes() {
  return (int)(((SN * ST + ((SN - 1) >> kColPadShift) + 1) * (SPLIT ? sizeof(F) : sizeof(cx<F>)) + 15) / 16 * 16);
}

// TWL: the twiddles come from the first half of the table, which the kernel
// has copied to LDS after the exchange area (col_twiddles_to_lds).
template <typename F, int E, int MODE, bool SPLIT, int G, bool BL, int SN, int ST, bool TWL = false>
__device__ __forceinline__ void col_transform(cx<F> (&v)[G][E], const ColArgs& a, const PassInst& io,
                                              const ColTile& t) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  cx<F>* lds = reinterpret_cast<cx<F>*>(smem_raw);
  static_assert(SN == 0 || (ST >= 4 && (ST & (ST - 1)) == 0), "static column tile: 4, 8, ... columns");
  constexpr int kMinCol = col_min<SN>();
  const int T = SN > 0 ? ST : a.T;
  int c, r0;
  col_thread<E, SN, ST>(a, c, r0);
  if constexpr (TWL) {
    static_assert(SN > 0 && SN % 32 == 0, "LDS twiddles: compiled-in plans, length a multiple of 32");
    const ColIdx idx{T, c, kColPadShift};
    const TwHalf<F, SN / 2> tw{reinterpret_cast<const cx<F>*>(smem_raw + col_exchange_bytes<F, SPLIT, SN, ST>())};
#pragma unroll
    for (int g = 0; g < G; ++g) {
      const int col = (t.tile * G + g) * T + c - t.sh;
      col_line<F, E, MODE, SPLIT, BL, SN>(v[g], a, io, tw, lds, idx, r0, t.outer, col, col >= kMinCol && col < a.ncols);
    }
    return;
  }
  const ColIdx idx{T, c, SN > 0 ? kColPadShift : a.ps};
  const cx<F>* __restrict__ tw = reinterpret_cast<const cx<F>*>(a.tw);
#pragma unroll
  for (int g = 0; g < G; ++g) {
    const int col = (t.tile * G + g) * T + c - t.sh;
    col_line<F, E, MODE, SPLIT, BL, SN>(v[g], a, io, tw, lds, idx, r0, t.outer, col, col >= kMinCol && col < a.ncols);
  }
}

template <typename F, int E, int G, bool BL, int SN, int ST, int MODE = COL_FWD>
__device__ __forceinline__ void col_store(cx<F> (&v)[G][E], const ColArgs& a, const PassInst& io, const ColTile& t) {
  constexpr int kMinCol = col_min<SN>();
  const int T = SN > 0 ? ST : a.T;
  const int L = SN > 0 ? SN / E : a.lp.L;
  int c, r0;
  col_thread<E, SN, ST>(a, c, r0);
  if constexpr (SN == 0 && MODE == COL_FWD) four_step_twiddle<-1>(v, a, t.outer, r0, L);
#pragma unroll
  for (int g = 0; g < G; ++g) {
    const int col = (t.tile * G + g) * T + c - t.sh;
    if (col >= kMinCol && col < a.ncols) {
      long long ob, os;
      side_addr<SN == 0>(a.out, t.outer, col, r0, L, ob, os);
      FF_GLB cx<F>* dst = glb<cx<F>>(io.dst) + ob;
#pragma unroll
      for (int i = 0; i < E; ++i)
        if (!BL || r0 + i * L < a.nline) dst[(long long)i * os] = v[g][i];
    }
  }
}

template <typename F, int E, int MODE, bool SPLIT, int G, bool BL, int SN, int ST, bool TWL = false>
__device__ __forceinline__ void col_tile(const ColArgs& a, const PassInst& io, int b) {
  const ColTile t = col_tile_at<F, SN>(a, io, b);
  cx<F> v[G][E];
  col_load<F, E, G, BL, SN, ST, MODE>(v, a, io, t);
  col_transform<F, E, MODE, SPLIT, G, BL, SN, ST, TWL>(v, a, io, t);
  col_store<F, E, G, BL, SN, ST, MODE>(v, a, io, t);
}

// Copies the first half of the twiddle table (SN/2 entries, swizzled) to LDS
// after the exchange area, for col_transform<..., TWL = true>.  The tile's
// loads are issued first, so the copy overlaps them.  The twiddle loads of
// the transforms otherwise go through the vector memory path that the tile's
// own loads and stores keep busy: from LDS the FP64 x pass at 1024 takes 7 %
// less time on MI355X.  Launch LDS: col_exchange_bytes + (SN/2) complex values.
template <typename F, bool SPLIT, int SN, int ST, int THREADS, int MODE = COL_FWD>
__device__ __forceinline__ void col_twiddles_to_lds(const ColArgs& a) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  cx<F>* twl = reinterpret_cast<cx<F>*>(smem_raw + col_exchange_bytes<F, SPLIT, SN, ST>());
  const FF_GLB cx<F>* twg = glb<const cx<F>>(a.tw);
  // every load is issued before the first LDS store, so the copy costs one
  // memory round trip: a loop that stores each value as it arrives waits
  // for each load in turn, behind the tile's loads
  // (indexes clamped into the table: every load is unconditional, which
  // keeps the values in registers)
  constexpr int NT = (SN / 2 + THREADS - 1) / THREADS;
  cx<F> tv[NT];
#pragma unroll
  for (int k = 0; k < NT; ++k) {
    const int i = min((int)threadIdx.x + k * THREADS, SN / 2 - 1);
    tv[k] = cx<F>{twg[i].x, twg[i].y};
  }
  // (finite-difference x pass: the x axis's factors, SN doubles, follow the
  // twiddles, and the filter reads them from LDS instead of through the
  // vector memory path; launch LDS: SN * 8 bytes more)
  constexpr int NS = MODE == COL_FILTER_FD ? (SN + THREADS - 1) / THREADS : 0;
  double sv[NS > 0 ? NS : 1];
  if constexpr (NS > 0) {
    const FF_GLB double* symg = glb<const double>(a.symx);
#pragma unroll
    for (int k = 0; k < NS; ++k) sv[k] = symg[min((int)threadIdx.x + k * THREADS, SN - 1)];
  }
#pragma unroll
  for (int k = 0; k < NT; ++k) {
    const int i = threadIdx.x + k * THREADS;
    if (i < SN / 2) twl[tw_swizzle(i)] = tv[k];
  }
  if constexpr (NS > 0) {
    double* sym = reinterpret_cast<double*>(twl + SN / 2);
#pragma unroll
    for (int k = 0; k < NS; ++k) {
      const int i = threadIdx.x + k * THREADS;
      if (i < SN) sym[i] = sv[k];
    }
  }
  __syncthreads();
}

// TWL (compiled-in plans): twiddles from LDS.
template <typename F, int E, int MODE, bool SPLIT, int G, bool BL, int SN = 0, int ST = 0, int SW = 1,
          bool TWL = false>
__global__ __attribute__((amdgpu_flat_work_group_size(1, SN > 0 ? SN / E * ST : max_threads_for(E)),
                          amdgpu_waves_per_eu(SN > 0 ? SW : 1))) void k_col(const ColArgs a) {
  const PassInst io = inst_of(a);
  const int b = xcd_remap(blockIdx.x, a.nblocks);
  if constexpr (TWL) {
    static_assert(SN > 0 && G == 1 && !BL, "LDS twiddles: compiled-in plans");
    const ColTile t = col_tile_at<F, SN>(a, io, b);
    cx<F> v[1][E];
    col_load<F, E, 1, false, SN, ST>(v, a, io, t);
    col_twiddles_to_lds<F, SPLIT, SN, ST, SN / E * ST, MODE>(a);
    col_transform<F, E, MODE, SPLIT, 1, false, SN, ST, true>(v, a, io, t);
    col_store<F, E, 1, false, SN, ST>(v, a, io, t);
  } else {
    col_tile<F, E, MODE, SPLIT, G, BL, SN, ST>(a, io, b);
  }
}

struct RowArgs {
  const void* src;
  void* dst;
  const void* aux;  // f for the residual epilogue
  long long nrows;
  int nz, nh;
  SpecSide spec;  // k_c2r: layout of the half spectrum read (lgKT >= 30: plain rows; lgKT = 0:
                  // strided, element (i, j, kh) at i*fo + j*fr + kh*KS, row = i*ny + j)
  int ny;
  int T;          // rows per workgroup
  uint32_t Linv;  // ceil(2^32 / L)
  int nblocks;
  int ps, pitch;  // LDS pad shift and row pitch (elements)
  double scale;   // 1/(nx*ny*nz)
  double shift;   // subtracted by the accumulate epilogue (times *shiftp if set)
  const double* shiftp;
  const void* tw;      // W_n of the engine transform length
  const void* twpost;  // exp(-2 pi i k / nz), k < nz/2 (even nz)
  LinePlan lp;
  int nline;                     // transform length M (lp.n: engine length; larger for Bluestein)
  const void *chirp, *bsf, *bsi;  // Bluestein tables (see line_dft)
  const PassInst* inst;           // grouped launch: per-instance fields (null: one solve)
  int m1;                         // four-step z rows: the stride-1 sub-line length (ff_zsplit.hip)
  int lone;                       // odd nz, nx*ny odd: the launch's one pair is the last row alone
                                  // (no second row: zero on the way in, not written on the way out)
};

__device__ __forceinline__ PassInst inst_of(const RowArgs& a) {
  if (a.inst) return uniform_inst(glb<const PassInst>(a.inst)[blockIdx.y]);
  return PassInst{a.src, a.dst, a.aux, nullptr, a.shiftp, 0.0, 0.0, 0.0};
}

template <typename T>
__device__ __forceinline__ double to_d(T x) { return double(x); }

// Complex offset of plain spectrum row (i, j), row = i*ny + j.
__device__ __forceinline__ long long spec_row(const RowArgs& a, long long row) { return row * a.nh; }

// z forward: real rows (IN precision) -> half-spectrum rows (F precision).
// Even nz: pack (x[2r], x[2r+1]) as one complex, an nz/2-point transform, then
// split the even/odd spectra.  Odd nz: a full nz-point transform of (x, 0).
// SN > 0: the plan of the row transform, length SN, is compiled in (even nz).
// r2c_tile: the work of one workgroup, row block blk (< a.nblocks), with the
// per-instance fields io (k_r2c: one pass; ff_multi.hip: many instances).
// TWL (compiled-in plans): the first half of the W_SN table is copied to LDS
// after the T row tiles and the passes read their twiddles there.
template <typename F, int SN>
__device__ __forceinline__ TwHalf<F, SN / 2> row_twiddles_to_lds(const RowArgs& a, cx<F>* lds) {
  cx<F>* twl = lds + (long long)a.T * a.pitch;
  const cx<F>* __restrict__ twg = reinterpret_cast<const cx<F>*>(a.tw);
  for (int i = threadIdx.x; i < SN / 2; i += blockDim.x) twl[tw_swizzle(i)] = twg[i];
  return TwHalf<F, SN / 2>{twl};
}

template <typename F, typename IN, int E, bool ODD, bool BL, int SN, bool TWL = false>
__device__ __forceinline__ void r2c_tile(const RowArgs& a, const PassInst& io, int blk) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  cx<F>* lds = reinterpret_cast<cx<F>*>(smem_raw);
  const int tid = threadIdx.x;
  const int L = SN > 0 ? SN / E : a.lp.L;
  const int line = SN > 0 ? tid / L : (L == 1 ? tid : (int)__umulhi((uint32_t)tid, a.Linv));
  const int r0 = tid - line * L;
  const long long row = (long long)blk * a.T + line;
  const bool valid = row < a.nrows;
  const long long rowc = valid ? row : a.nrows - 1;  // loads stay in bounds; stores are predicated
  const int M = a.nline;                  // transform length (the engine's lp.n is longer for Bluestein)
  constexpr bool blue = BL;               // positions >= M exist only for Bluestein
  const cx<F>* __restrict__ tw = reinterpret_cast<const cx<F>*>(a.tw);
  // (compiled-in plans whose rows fit a wavefront: wavefront-level exchanges)
  using RI = std::conditional_t<(SN > 0 && 64 % (SN > 0 ? SN / E : 64) == 0), RowIdxW, RowIdx>;
  const RI idx{line * a.pitch, a.ps};
  cx<F> v[E];
  if constexpr (!ODD) {
    const FF_GLB IN* __restrict__ src = glb<const IN>(io.src) + rowc * a.nz;
#pragma unroll
    for (int i = 0; i < E; ++i) {
      const int k = r0 + i * L;
      if (!blue || k < M) {
        if constexpr (sizeof(IN) == 8) {
          const cx<double> p = ((const FF_GLB cx<double>*)src)[k];
          v[i] = cx<F>{F(p.x), F(p.y)};
        } else {
          const cx<float> p = ((const FF_GLB cx<float>*)src)[k];
          v[i] = cx<F>{F(p.x), F(p.y)};
        }
      } else {
        v[i] = cx<F>{F(0), F(0)};
      }
    }
    if constexpr (TWL) {
      static_assert(SN > 0 && !BL, "LDS twiddles: compiled-in plans");
      const TwHalf<F, SN / 2> twl = row_twiddles_to_lds<F, SN>(a, lds);
      __syncthreads();
      line_dft<E, -1, F, RI, false, BL, SN>(v, a.lp, M, twl, a.chirp, a.bsf, a.bsi, lds, idx, r0);
    } else {
      line_dft<E, -1, F, RI, false, BL, SN>(v, a.lp, M, tw, a.chirp, a.bsf, a.bsi, lds, idx, r0);
    }
    exchange_sync<RI>();
    put_natural<E, F>(v, lds, idx, r0, L);
    exchange_sync<RI>();
    const cx<F>* __restrict__ twp = reinterpret_cast<const cx<F>*>(a.twpost);
    FF_GLB cx<F>* __restrict__ dst = glb<cx<F>>(io.dst) + (valid ? spec_row(a, row) : 0);
    if (valid) {
#pragma unroll
      for (int i = 0; i < E; ++i) {
        const int k = r0 + i * L;
        if (blue && k >= M) continue;
        const cx<F> zk = lds[idx(k)];
        const cx<F> zm = conj(lds[idx(k ? M - k : 0)]);
        // even spectrum (zk + zm)/2, odd spectrum (zk - zm)/(2i)
        const cx<F> ev = {F(0.5) * (zk.x + zm.x), F(0.5) * (zk.y + zm.y)};
        const cx<F> od = {F(0.5) * (zk.y - zm.y), F(-0.5) * (zk.x - zm.x)};
        dst[k] = ev + twmul<-1>(od, twp[k]);
      }
      if (r0 == 0) {
        const cx<F> z0 = lds[idx(0)];
        dst[M] = cx<F>{z0.x - z0.y, F(0)};
      }
    }
  } else {
    const FF_GLB IN* __restrict__ src = glb<const IN>(io.src) + rowc * a.nz;
#pragma unroll
    for (int i = 0; i < E; ++i) {
      const int k = r0 + i * L;
      v[i] = cx<F>{(!blue || k < M) ? F(src[k]) : F(0), F(0)};
    }
    line_dft<E, -1, F, RI, false, BL, SN>(v, a.lp, M, tw, a.chirp, a.bsf, a.bsi, lds, idx, r0);
    FF_GLB cx<F>* __restrict__ dst = glb<cx<F>>(io.dst) + (valid ? spec_row(a, row) : 0);
    if (valid) {
#pragma unroll
      for (int i = 0; i < E; ++i) {
        const int k = r0 + i * L;
        if (k < a.nh) dst[k] = v[i];
      }
    }
  }
}

template <typename F, typename IN, int E, bool ODD, bool BL, int SN = 0, bool TWL = false>
__global__ __launch_bounds__(max_threads_for(E)) void k_r2c(const RowArgs a) {
  r2c_tile<F, IN, E, ODD, BL, SN, TWL>(a, inst_of(a), xcd_remap(blockIdx.x, a.nblocks));
}

// Epilogues of the z inverse: store scale*x, store the residual
// f - shift - scale*x, or accumulate out += scale*x - shift.
enum { EPI_STORE = 0, EPI_RESIDUAL = 1, EPI_ACCUM = 2 };

template <int EPI, typename OUT>
__device__ __forceinline__ void c2r_out1(FF_GLB OUT* out, const FF_GLB double* f, long long off, double val,
                                         double scale, double shift) {
  if constexpr (EPI == EPI_STORE) {
    out[off] = OUT(scale * val);
  } else if constexpr (EPI == EPI_RESIDUAL) {
    out[off] = OUT(fmadd(-scale, val, f[off] - shift));
  } else {
    out[off] = OUT(double(out[off]) + fmadd(scale, val, -shift));
  }
}

// Two neighbouring outputs (2r, 2r+1) with one vector access.
template <int EPI, typename OUT>
__device__ __forceinline__ void c2r_out2(FF_GLB OUT* out, const FF_GLB double* f, long long r, double v0,
                                         double v1, double scale, double shift) {
  FF_GLB cx<OUT>* o = (FF_GLB cx<OUT>*)out + r;
  if constexpr (EPI == EPI_STORE) {
    *o = cx<OUT>{OUT(scale * v0), OUT(scale * v1)};
  } else if constexpr (EPI == EPI_RESIDUAL) {
    const cx<double> fp = ((const FF_GLB cx<double>*)f)[r];
    *o = cx<OUT>{OUT(fmadd(-scale, v0, fp.x - shift)), OUT(fmadd(-scale, v1, fp.y - shift))};
  } else {
    const cx<OUT> p = *o;
    *o = cx<OUT>{OUT(double(p.x) + fmadd(scale, v0, -shift)), OUT(double(p.y) + fmadd(scale, v1, -shift))};
  }
}

// z inverse: half-spectrum rows (F) -> real rows (OUT) with an epilogue.
// The imaginary parts of the DC and (even nz) Nyquist terms are ignored,
// i.e. the rows are read as Hermitian half spectra.
// SN > 0: the plan of the row transform, length SN, is compiled in (even nz).
// c2r_tile: the work of one workgroup, row block blk (< a.nblocks), with the
// per-instance fields io (k_c2r: one pass; ff_multi.hip: many instances).
template <typename F, typename OUT, int E, bool ODD, int EPI, bool BL, int SN, bool TWL = false>
__device__ __forceinline__ void c2r_tile(const RowArgs& a, const PassInst& io, int blk) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  cx<F>* lds = reinterpret_cast<cx<F>*>(smem_raw);
  const int tid = threadIdx.x;
  const int L = SN > 0 ? SN / E : a.lp.L;
  const int line = SN > 0 ? tid / L : (L == 1 ? tid : (int)__umulhi((uint32_t)tid, a.Linv));
  const int r0 = tid - line * L;
  const long long row = (long long)blk * a.T + line;
  const bool valid = row < a.nrows;
  const int M = a.nline;                // transform length (the engine's lp.n is longer for Bluestein)
  constexpr bool blue = BL;             // positions >= M exist only for Bluestein
  const cx<F>* __restrict__ tw = reinterpret_cast<const cx<F>*>(a.tw);
  // (compiled-in plans whose rows fit a wavefront: wavefront-level exchanges
  // after the staging)
  using RI = std::conditional_t<(SN > 0 && 64 % (SN > 0 ? SN / E : 64) == 0), RowIdxW, RowIdx>;
  const RI idx{line * a.pitch, a.ps};
  // stage the half spectrum of the workgroup's rows in LDS; rows past the end
  // load the last row (in bounds), their stores are predicated
  if (a.spec.lgKT >= 30) {
    const FF_GLB cx<F>* __restrict__ src = glb<const cx<F>>(io.src) + spec_row(a, valid ? row : a.nrows - 1);
#pragma unroll
    for (int i = 0; i < E; ++i) {
      const int k = r0 + i * L;
      // even nz without Bluestein: k < M < nh always
      if ((!ODD && !BL) || k < a.nh) lds[idx(k)] = src[k];
    }
    if (!ODD && r0 == 0) lds[idx(M)] = src[M];
  } else if (a.spec.lgKT == 0) {
    // strided spectrum (pencil layout [i][kh][j]): the workgroup's T rows are
    // T consecutive j, contiguous for each kh.  Thread t copies row t % T at
    // kh = t / T + i*L (i < E, all loads issued first) and the last entry kh
    // = M: a wave reads 64/T segments of T elements.  Rows past the end load
    // the last row (in bounds, never stored).
    const int rr = tid % a.T;
    const int m = tid / a.T;
    long long r = (long long)blk * a.T + rr;
    if (r >= a.nrows) r = a.nrows - 1;
    const long long ii = r / a.ny;
    const long long jj = r - ii * a.ny;
    const FF_GLB cx<F>* __restrict__ src = glb<const cx<F>>(io.src) + (ii * a.spec.fo + jj * a.spec.fr);
    const RowIdx sidx{rr * a.pitch, a.ps};
    cx<F> t[E];
#pragma unroll
    for (int i = 0; i < E; ++i) {
      const int k = m + i * L;
      // even nz without Bluestein: k < M < nh always
      if ((!ODD && !BL) || k < a.nh) t[i] = src[(long long)k * a.spec.KS];
    }
    cx<F> last{};
    if (!ODD && m == 0) last = src[(long long)M * a.spec.KS];
#pragma unroll
    for (int i = 0; i < E; ++i) {
      const int k = m + i * L;
      if ((!ODD && !BL) || k < a.nh) lds[sidx(k)] = t[i];
    }
    if (!ODD && m == 0) lds[sidx(M)] = last;
  } else {
    // y-blocked spectrum: the workgroup's T rows are contiguous inside each
    // kh block (T*KT elements), so all threads copy element e of the tile in
    // (block, row, kt) order
    const FF_GLB cx<F>* __restrict__ src = glb<const cx<F>>(io.src);
    const long long row0 = (long long)blk * a.T;
    const int KT = 1 << a.spec.lgKT;
    const int TK = a.T * KT;
    const int total = a.T * a.nh;
    const int nthreads = a.T * L;
    for (int e = tid; e < total; e += nthreads) {
      int kb = e / TK;
      int W = KT;
      if (kb >= a.spec.nb - 1) {
        kb = a.spec.nb - 1;
        W = a.spec.Wlast;
      }
      const int rem = e - kb * TK;
      const int rr = rem / W;
      const int kt = rem - rr * W;
      long long r = row0 + rr;
      if (r >= a.nrows) r = a.nrows - 1;
      lds[rr * a.pitch + ((kb * KT + kt) + ((kb * KT + kt) >> a.ps))] =
          src[(long long)kb * a.spec.KS + r * W + kt];
    }
  }
  TwHalf<F, (SN > 0 ? SN / 2 : 1)> twl{nullptr};
  if constexpr (TWL) {
    static_assert(SN > 0 && !BL, "LDS twiddles: compiled-in plans");
    twl = row_twiddles_to_lds<F, SN>(a, lds);
  }
  __syncthreads();
  cx<F> v[E];
  if constexpr (!ODD) {
    const cx<F>* __restrict__ twp = reinterpret_cast<const cx<F>*>(a.twpost);
#pragma unroll
    for (int i = 0; i < E; ++i) {
      const int k = r0 + i * L;
      if (blue && k >= M) {
        v[i] = cx<F>{F(0), F(0)};
        continue;
      }
      cx<F> xk = lds[idx(k)];
      cx<F> xm = lds[idx(M - k)];
      if (k == 0) {
        xk.y = F(0);
        xm.y = F(0);
      }
      const cx<F> cm = conj(xm);
      const cx<F> A = xk + cm;
      const cx<F> B = twmul<+1>(xk - cm, twp[k]);
      v[i] = cx<F>{A.x - B.y, A.y + B.x};
    }
  } else {
#pragma unroll
    for (int i = 0; i < E; ++i) {
      const int k = r0 + i * L;
      if (blue && k >= M) {
        v[i] = cx<F>{F(0), F(0)};
        continue;
      }
      cx<F> y = k < a.nh ? lds[idx(k)] : conj(lds[idx(a.nz - k)]);
      if (k == 0) y.y = F(0);
      v[i] = y;
    }
  }
  if constexpr (TWL) {
    line_dft<E, +1, F, RI, false, BL, SN>(v, a.lp, M, twl, a.chirp, a.bsf, a.bsi, lds, idx, r0);
  } else {
    line_dft<E, +1, F, RI, false, BL, SN>(v, a.lp, M, tw, a.chirp, a.bsf, a.bsi, lds, idx, r0);
  }
  if (valid) {
    const double shift = io.shiftp ? a.shift * *glb<const double>(io.shiftp) : a.shift;
    FF_GLB OUT* out = glb<OUT>(io.dst) + row * a.nz;
    const FF_GLB double* f = glb<const double>(io.aux) + (EPI == EPI_RESIDUAL ? row * a.nz : 0);
#pragma unroll
    for (int i = 0; i < E; ++i) {
      const int r = r0 + i * L;
      if (blue && r >= M) continue;
      if constexpr (!ODD) {
        c2r_out2<EPI, OUT>(out, f, r, to_d(v[i].x), to_d(v[i].y), a.scale, shift);
      } else {
        c2r_out1<EPI, OUT>(out, f, r, to_d(v[i].x), a.scale, shift);
      }
    }
  }
}

template <typename F, typename OUT, int E, bool ODD, int EPI, bool BL, int SN = 0, bool TWL = false>
__global__ __launch_bounds__(max_threads_for(E)) void k_c2r(const RowArgs a) {
  c2r_tile<F, OUT, E, ODD, EPI, BL, SN, TWL>(a, inst_of(a), xcd_remap(blockIdx.x, a.nblocks));
}

// ---------------------------------------------------------------------------
// Launchers: one explicit instantiation unit per E (ff_inst_E*.hip).

struct Launch {
  dim3 grid, block;
  size_t lds;
  hipStream_t stream;
};

// Column kernels compiled with the line plan of (n, E) and the tile width T
// fixed (LDS pad shift 4, one group).  The FP64 x pass at 1024 runs the split
// exchange (real and imaginary parts through half the LDS, 33 KB): at 166
// VGPRs three workgroups then fit a CU instead of two, and the pass takes
// 4.84 ms instead of 5.14 ms on MI355X (the y passes gain nothing from it).  At 1024 (E = 16, the tile of the
// default configuration) the folded indices save 15-20 % of the x pass and a
// few percent of the y passes on MI355X.  Measured and not kept: 512 (E = 8,
// three waves per SIMD: the solve as fast as with the generic kernels), other
// tiles at 1024 (E = 8, or T = 2: slower) and the split exchange at three
// waves per SIMD (slower).
struct StaticCol {
  int n, E, T, f32;
  int split;     // FF_TUNE_SPLIT_LDS exchange
  int waves;     // waves per SIMD the registers must allow
  int kinds;     // passes it serves, as 1 << kind: 2 the y pass (forward, inverse), 4 the x pass (filter, symbol)
  int twl_waves;  // > 0: the kernel with LDS twiddles (FF_TUNE_LDS_TWIDDLES) is compiled too, for this
                  // many waves per SIMD
};
constexpr StaticCol kStaticCols[] = {{1024, 16, 4, 0, 0, 2, 6, 2}, {1024, 16, 8, 1, 0, 2, 6, 2}, {1024, 16, 4, 0, 1, 3, 6, 3},
                                     // x passes only at 512, 768 and 1000 (the y passes are as fast or
                                     // faster generic there); both column passes at 256 and below
                                     {768, 12, 4, 0, 0, 3, 4}, {1000, 20, 4, 0, 0, 2, 4},
                                     {512, 8, 4, 0, 0, 4, 4}, {256, 16, 16, 0, 0, 2, 6}, {128, 16, 32, 0, 0, 2, 6},
                                     // small grids (FP64 and FP32)
                                     {160, 20, 32, 0, 0, 2, 6}, {96, 24, 64, 0, 0, 1, 6}, {64, 8, 32, 0, 0, 4, 6},
                                     {48, 12, 32, 0, 0, 3, 6}, {32, 8, 32, 0, 0, 4, 6}, {128, 16, 32, 1, 0, 2, 6},
                                     {160, 20, 32, 1, 0, 2, 6}, {96, 24, 64, 1, 0, 2, 6}, {64, 8, 32, 1, 0, 4, 6},
                                     {48, 12, 32, 1, 0, 4, 6}, {32, 8, 32, 1, 0, 4, 6}, {256, 16, 16, 1, 0, 2, 6},
                                     {512, 8, 8, 1, 0, 4, 4}};
constexpr int kNumStaticCols = sizeof(kStaticCols) / sizeof(kStaticCols[0]);

// Row transforms (z passes at even nz, length M = nz/2) with the plan of
// (M, E) compiled in: the default configurations of nz = 32 ... 1024
// (768 has E = 12 in FP64 and E = 24 in FP32).
struct StaticRow {
  int M, E;
};
constexpr StaticRow kStaticRows[] = {{512, 8}, {500, 10}, {384, 12}, {384, 24}, {256, 16}, {128, 16}, {80, 20},
                                     {64, 8},  {48, 12},  {32, 8},   {24, 12},  {16, 16}};
constexpr int kNumStaticRows = sizeof(kStaticRows) / sizeof(kStaticRows[0]);

template <int E>
struct PassLaunchers {
  static hipError_t col(int f32, int mode, int split, int groups, int blue, int stat, const Launch& l,
                        const ColArgs& a);
  static hipError_t r2c(int f32, int in_f32, int odd, int blue, int stat, const Launch& l, const RowArgs& a);
  static hipError_t c2r(int f32, int out_kind, int odd, int epi, int blue, int stat, const Launch& l,
                        const RowArgs& a);
};

}  // namespace ff
