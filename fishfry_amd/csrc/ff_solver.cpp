// Plan construction and the pass sequence of the periodic Poisson solve.
//
// At 1x1x1 tasks the reference's 25 steps (ParisPencil.cpp:95-425) are 6 FFT
// executions, one filter, one scale and 17 reorders / self all-to-alls.  Here
// they are five kernels, each reading and writing the whole array once:
//   z forward (D2Z, :165), y forward (:212), x forward*filter*inverse
//   (:259,:263-285,:288), y inverse (:334), z inverse (Z2D, :380) with the
//   1/(nx*ny*nz) scale of :466-481 in its epilogue.
#include <algorithm>
#include <initializer_list>
#include <climits>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <new>

#include "ff_solver.hpp"

namespace ff {

size_t reference_bytes(const int n[3], const int m[3]) {
  // ParisPencil.cpp:18 (nh), :24-26 (mp x mq split of mk), :51-75
  const long nh = n[2] / 2 + 1;
  const long mi = m[0], mj = m[1], mk = m[2];
  long mq = std::lround(std::sqrt((double)mk));
  while (mk % mq) mq--;
  const long mp = mk / mq;
  const long dh = (nh + mk - 1) / mk;
  const long di = (n[0] + mi - 1) / mi;
  const long dj = (n[1] + mj - 1) / mj;
  const long dk = (n[2] + mk - 1) / mk;
  const long dip = (di + mp - 1) / mp;
  const long djq = (dj + mq - 1) / mq;
  const long mjq = mj * mq;
  const long dhq = (nh + mjq - 1) / mjq;
  const long mip = mi * mp;
  const long djp = (n[1] + mip - 1) / mip;
  const long nmax = std::max({di * dj * dk, mp * mq * dip * djq * dk, 2 * dip * djq * mk * dh,
                              2 * dip * mp * djq * mq * dh, 2 * dip * djq * mjq * dhq, 2 * dip * dhq * mip * djp,
                              2 * djp * dhq * mip * dip});
  return (size_t)nmax * sizeof(double);
}

static double dd_const(int n, double lo, double hi) {
  // ParisPencil.cpp:9-11, same operand order
  return 2.0 * M_PI * double(n - 1) / (double(n) * (hi - lo));
}

// Four-step split of an x or y extent n that no workgroup transforms whole:
// n = n1*n2 (n1 <= n2), both sub-line lengths planned (generic kernels,
// Bluestein allowed), the split with the least work n2*cost(n1) +
// n1*cost(n2), cost = engine length * (passes + 0.04 E).  False when n has
// no such split (a prime above the Bluestein limit).  ncols: the columns of
// the sub-line passes (x, y: nh); z rows (zrows): the n1-point lines run over
// n2 columns and the n2-point ones over n1.
static bool four_step_cfg(int n, int ncols, int eb, int kind, int prec, ff_plan::FourStep& fs,
                          bool zrows = false) {
  if (fs.n == n) {
    // another precision of the plan chose the split: the same one here
    const int n1 = fs.n1, n2 = fs.n2;
    return col_cfg_with(n1, zrows ? n2 : ncols, eb, fs.c1[prec], 0, 0, FF_TUNE_GENERIC, kind) &&
           col_cfg_with(n2, zrows ? n1 : ncols, eb, fs.c2[prec], 0, 0, FF_TUNE_GENERIC, kind);
  }
  auto cost = [](const ColCfg& c) { return double(c.line.lp.n) * (double(c.line.lp.np) + 0.04 * c.line.E); };
  double best = 1e300;
  int b1 = 0;
  ColCfg bc1, bc2;
  for (int n1 = 2; (long long)n1 * n1 <= n; ++n1) {
    if (n % n1) continue;
    const int n2 = n / n1;
    ColCfg c1, c2;
    if (!col_cfg_with(n1, zrows ? n2 : ncols, eb, c1, 0, 0, FF_TUNE_GENERIC, kind) ||
        !col_cfg_with(n2, zrows ? n1 : ncols, eb, c2, 0, 0, FF_TUNE_GENERIC, kind))
      continue;
    const double w = double(n2) * cost(c1) + double(n1) * cost(c2);
    if (w < best) {
      best = w;
      b1 = n1;
      bc1 = c1;
      bc2 = c2;
    }
  }
  if (!b1) return false;
  fs.n = n;
  fs.n1 = b1;
  fs.n2 = n / b1;
  fs.c1[prec] = bc1;
  fs.c2[prec] = bc2;
  return true;
}

ff_status plan_create(const int n[3], const double lo[3], const double hi[3], const ff_options* opts,
                      ff_plan** out, bool own_scratch) {
  *out = nullptr;
  for (int a = 0; a < 3; ++a) {
    if (n[a] < 1) return set_error(FF_ERR_INVALID, "grid extents must be >= 1");
    if (!(hi[a] > lo[a])) return set_error(FF_ERR_INVALID, "hi must exceed lo on every axis");
  }
  ff_options o{FF_FP64, 0, 0};
  if (opts) o = *opts;
  if (o.precision != FF_FP64 && o.precision != FF_FP32_REFINE)
    return set_error(FF_ERR_INVALID, "options.precision must be FF_FP64 or FF_FP32_REFINE");
  if (o.refine_steps < 0 || o.refine_steps > 8) return set_error(FF_ERR_INVALID, "refine_steps must be 0..8");
  if (o.flags & ~(FF_PRESERVE_INPUT | FF_BLOCKED_LAYOUT | FF_SLAB_PIPELINE | FF_SYMBOL_3PT | FF_SYMBOL_5PT |
                  FF_PENCIL_LAYOUT | FF_PLAIN_LAYOUT))
    return set_error(FF_ERR_INVALID, "unknown bits in options.flags");
  if ((o.flags & FF_PENCIL_LAYOUT) && (o.flags & (FF_BLOCKED_LAYOUT | FF_SLAB_PIPELINE | FF_PLAIN_LAYOUT)))
    return set_error(FF_ERR_INVALID,
                     "options.flags: FF_PENCIL_LAYOUT excludes FF_BLOCKED_LAYOUT, FF_SLAB_PIPELINE and FF_PLAIN_LAYOUT");
  if ((o.flags & FF_SYMBOL_3PT) && (o.flags & FF_SYMBOL_5PT))
    return set_error(FF_ERR_INVALID, "options.flags: at most one of FF_SYMBOL_3PT and FF_SYMBOL_5PT");

  ff_plan* p = new (std::nothrow) ff_plan();
  if (!p) return set_error(FF_ERR_NOMEM, "host allocation failed");
  memcpy(p->n, n, sizeof p->n);
  memcpy(p->lo, lo, sizeof p->lo);
  memcpy(p->hi, hi, sizeof p->hi);
  p->opt = o;
  p->nh = n[2] / 2 + 1;
  p->N = (long long)n[0] * n[1] * n[2];
  p->Nh = (long long)n[0] * n[1] * p->nh;
  p->ddi = dd_const(n[0], lo[0], hi[0]);
  p->ddj = dd_const(n[1], lo[1], hi[1]);
  p->ddk = dd_const(n[2], lo[2], hi[2]);
  p->sym = (o.flags & FF_SYMBOL_3PT) ? 3 : ((o.flags & FF_SYMBOL_5PT) ? 5 : 0);
  for (int a = 0; a < 3; ++a) {
    const double ih = double(n[a] - 1) / (hi[a] - lo[a]);  // 1/h, h = (hi - lo)/(n - 1)
    p->fd[a] = ih * ih;
    p->symtab[a] = nullptr;
  }
  const int ones[3] = {1, 1, 1};
  p->bytes = reference_bytes(n, ones);
  // kh blocks of 8 complex doubles (one 128-byte line per block row)
  p->lgKT = (o.flags & FF_BLOCKED_LAYOUT) ? 3 : 0;
  // pencil layout: on request, or by default for FP64 grids of 2^25 spectrum
  // points and more with nz <= 1024 (measured on MI355X: 512^3 and 768^3
  // 12-13 % faster, 1024^3 8 %, 2048x1024x256 36 %, 1000^3 even; 128^3 and
  // 256^3 slower, and 1296^3, whose z inverse rows do not fit eight to a
  // workgroup in half the LDS)
  // FP32 plans run their FP32 passes in the pencil layout on request only,
  // between the two halves of the refinement work area (at 1024^3 their
  // FP32 passes are 4 % slower than in place); their FP64 residual passes
  // stay in place in the plain layout.
  const bool other_layout = (o.flags & (FF_PLAIN_LAYOUT | FF_BLOCKED_LAYOUT | FF_SLAB_PIPELINE)) != 0;
  p->pencil = (o.flags & FF_PENCIL_LAYOUT) ||
                      (o.precision == FF_FP64 && !other_layout && p->Nh >= (1LL << 25) && n[2] <= 1024)
                  ? 1
                  : 0;
  const int mainp = o.precision == FF_FP64 ? 0 : 1;  // precision of the passes the layout applies to
  p->slab = 0;
  if ((o.flags & FF_SLAB_PIPELINE) && !p->lgKT) {
    const long long plane = (long long)n[1] * p->nh * 16;  // bytes of one x-plane of the spectrum
    p->slab = o.slab_planes > 0 ? o.slab_planes : (int)std::max(1LL, (64LL << 20) / plane);
    if (p->slab >= n[0]) p->slab = 0;  // one slab: plain passes
  }

  ff_status st = hip_check(hipGetDevice(&p->device), "hipGetDevice");
  if (st != FF_OK) {
    delete p;
    return st;
  }
  const int nprec = (o.precision == FF_FP64) ? 1 : 2;
  // x or y extents that no workgroup transforms whole: four-step lines
  p->fy.n = p->fx.n = p->fz.n = 0;
  for (int prec = 0; prec < nprec; ++prec) {
    const int eb = prec ? 8 : 16;
    ColCfg c;
    if (!make_col_cfg(n[1], p->nh, eb, c, 1, prec)) four_step_cfg(n[1], p->nh, eb, 1, prec, p->fy);
    if (!make_col_cfg(n[0], p->nh, eb, c, 2, prec)) four_step_cfg(n[0], p->nh, eb, 2, prec, p->fx);
  }
  // an axis in four steps in one precision runs so in all of the plan's
  for (int prec = 0; prec < nprec; ++prec) {
    const int eb = prec ? 8 : 16;
    if (p->fy.n && !p->fy.c1[prec].threads) four_step_cfg(n[1], p->nh, eb, 1, prec, p->fy);
    if (p->fx.n && !p->fx.c1[prec].threads) four_step_cfg(n[0], p->nh, eb, 2, prec, p->fx);
  }
  // even nz whose packed rows (nz/2 complex values) no workgroup transforms
  // whole, in any of the plan's precisions: four-step rows in all of them
  // (odd nz: the rows are transformed whole as complex rows, two real rows
  // packed per complex row; with nx*ny odd the last row runs alone)
  bool zfour = false;
  for (int prec = 0; prec < nprec; ++prec) {
    RowCfg r0, r1;
    const int eb = prec ? 8 : 16;
    if (!make_row_cfg(n[2], eb, r0, 0, prec) || !make_row_cfg(n[2], eb, r1, 3, prec)) zfour = true;
  }
  p->zodd = zfour && n[2] % 2 ? 1 : 0;
  p->zlone = p->zodd && ((long long)n[0] * n[1]) % 2 ? 1 : 0;
  for (int prec = 0; prec < nprec && zfour; ++prec)
    if (!four_step_cfg(p->zodd ? n[2] : n[2] / 2, 0, prec ? 8 : 16, 1, prec, p->fz, true)) p->fz.c1[prec].threads = 0;
  const bool four = p->fy.n || p->fx.n || p->fz.n;
  if (four) {
    if (o.flags & (FF_PENCIL_LAYOUT | FF_BLOCKED_LAYOUT | FF_SLAB_PIPELINE)) {
      delete p;
      return set_error(FF_ERR_UNSUPPORTED, "extents transformed in four steps take the plain spectrum layout only");
    }
    p->pencil = 0;
  }
  for (int prec = 0; prec < nprec; ++prec) {
    const int eb = prec ? 8 : 16;
    // with four-step y lines the x pass maps its j planes to frequencies: a
    // generic kernel (compiled-in ones take the plain order)
    auto x_cfg = [&](ColCfg& c) -> bool {
      if (p->fx.n) return true;  // four-step x lines: p->fx
      if (!make_col_cfg(n[0], p->nh, eb, c, 2, prec, p->pencil && prec == mainp)) return false;
      if (!p->fy.n || !c.stat) return true;
      const int fl = FF_TUNE_GENERIC | (c.split ? FF_TUNE_SPLIT_LDS : 0) | (c.groups == 2 ? FF_TUNE_TWO_GROUPS : 0);
      return col_cfg_with(n[0], p->nh, eb, c, c.line.E, c.T, fl, 2);
    };
    const bool z_ok = p->fz.n ? p->fz.c1[prec].threads > 0
                              : make_row_cfg(n[2], eb, p->zr[prec], 0, prec) && make_row_cfg(n[2], eb, p->zi[prec], 3, prec);
    if (!z_ok || !(p->fy.n || make_col_cfg(n[1], p->nh, eb, p->cy[prec], 1, prec, p->pencil && prec == mainp)) ||
        !x_cfg(p->cx[prec]) || (p->fy.n && !p->fy.c1[prec].threads) || (p->fx.n && !p->fx.c1[prec].threads)) {
      delete p;
      char buf[256];
      snprintf(buf, sizeof buf,
               "no transform plan for %dx%dx%d: an extent (nz/2 for even nz) is too long for one "
               "workgroup's line transform and has no four-step split",
               n[0], n[1], n[2]);
      return set_error(FF_ERR_UNSUPPORTED, buf);
    }
  }

  // pencil layout: the z inverse reads its rows from [i][kh][j], T
  // consecutive j per kh, so it takes 128 bytes of rows per workgroup (8
  // FP64 or 16 FP32 rows) where two such workgroups fit a CU's LDS, else 64
  // (unless tuned: ff_tune_set takes precedence)
  int tE, tT, tF;
  if (p->pencil && !tune_lookup(3, n[2] % 2 ? n[2] : n[2] / 2, mainp, tE, tT, tF)) {
    RowCfg& zi = p->zi[mainp];
    const int eb = mainp ? 8 : 16;
    for (int T : {128 / eb, 64 / eb}) {
      RowCfg c;
      const int fl = zi.stat == 2 ? FF_TUNE_LDS_TWIDDLES : (zi.stat ? 0 : FF_TUNE_GENERIC);
      if (row_cfg_with(n[2], eb, c, zi.line.E, T, fl) &&
          (T == 64 / eb || c.lds <= 80 * 1024)) {
        zi = c;
        break;
      }
    }
  }

  // one workgroup per tile: every pass's grid must fit a 32-bit grid size
  for (int prec = 0; prec < nprec; ++prec) {
    const long long rows = (long long)n[0] * n[1];
    // (four-step z rows: n2-point lines over n1 columns, n1-point ones over n2)
    auto tiles = [](long long cols, int T) { return (cols + T - 1) / T; };
    const long long rb =
        p->fz.n ? rows * std::max(tiles(p->fz.n1, p->fz.c2[prec].T), tiles(p->fz.n2, p->fz.c1[prec].T))
                : std::max((rows + p->zr[prec].T - 1) / p->zr[prec].T, (rows + p->zi[prec].T - 1) / p->zi[prec].T);
    auto col_blocks = [&](const ColCfg& c, int nouter) {
      const long long tiles = (p->nh + 7 + (long long)c.groups * c.T - 1) / ((long long)c.groups * c.T);
      return (long long)nouter * tiles;
    };
    // (pencil layout: the x and y inverse passes run nh planes of ny and nx columns)
    auto pencil_blocks = [&](const ColCfg& c, int ncols) {
      return p->pencil ? (long long)p->nh * ((ncols + c.T - 1) / c.T) : 0LL;
    };
    // (four-step lines: sub-line passes over n1 and n2 times as many planes)
    auto axis_blocks = [&](const ff_plan::FourStep& fs, const ColCfg& c, int nouter) {
      if (!fs.n) return col_blocks(c, nouter);
      return std::max(col_blocks(fs.c1[prec], nouter) * fs.n2, col_blocks(fs.c2[prec], nouter) * fs.n1);
    };
    if (rb > INT_MAX || axis_blocks(p->fy, p->cy[prec], n[0]) > INT_MAX ||
        axis_blocks(p->fx, p->cx[prec], n[1]) > INT_MAX || pencil_blocks(p->cx[prec], n[1]) > INT_MAX ||
        pencil_blocks(p->cy[prec], n[0]) > INT_MAX) {
      delete p;
      return set_error(FF_ERR_UNSUPPORTED, "grid too large: a pass would need more than 2^31-1 workgroups");
    }
  }

  // twiddle tables of the engine transforms (length m for Bluestein lines),
  // the even-nz post-twiddles and the Bluestein tables, from the resident
  // per-device cache (shared by every plan with the same line lengths)
  const int nz = n[2];
  bool ok = true;
  auto blue = [&](const LineCfg& c, int prec, const void* (&o)[3]) {
    o[0] = o[1] = o[2] = nullptr;
    if (c.bluestein()) ok = ok && table_bluestein(c.n, c.lp.n, prec, o);
  };
  for (int prec = 0; prec < nprec && ok; ++prec) {
    p->tw_zpost[prec] = table_post(nz, prec);
    ok = p->tw_zpost[prec] != nullptr;
    if (!p->fz.n) {
      p->tw_z[prec] = table_twiddle(p->zr[prec].line.lp.n, prec);
      p->tw_zi[prec] = table_twiddle(p->zi[prec].line.lp.n, prec);
      ok = ok && p->tw_z[prec] && p->tw_zi[prec];
      blue(p->zr[prec].line, prec, p->bl_z[prec]);
      blue(p->zi[prec].line, prec, p->bl_zi[prec]);
    }
    // column axes and four-step z rows: whole lines, or the two sub-line
    // lengths and W_n
    for (int axis = 0; axis < 3; ++axis) {
      ff_plan::FourStep& fs = axis == 2 ? p->fz : axis ? p->fy : p->fx;
      if (axis == 2 && !fs.n) continue;
      const ColCfg& c = axis ? p->cy[prec] : p->cx[prec];
      if (fs.n) {
        fs.tw1[prec] = table_twiddle(fs.c1[prec].line.lp.n, prec);
        fs.tw2[prec] = table_twiddle(fs.c2[prec].line.lp.n, prec);
        fs.twn[prec] = table_twiddle(fs.n, prec);
        ok = ok && fs.tw1[prec] && fs.tw2[prec] && fs.twn[prec];
        blue(fs.c1[prec].line, prec, fs.bl1[prec]);
        blue(fs.c2[prec].line, prec, fs.bl2[prec]);
      } else {
        const void*& tw = axis ? p->tw_y[prec] : p->tw_x[prec];
        tw = table_twiddle(c.line.lp.n, prec);
        ok = ok && tw;
        blue(c.line, prec, axis ? p->bl_y[prec] : p->bl_x[prec]);
      }
    }
  }
  for (int a = 0; a < 3 && ok && p->sym; ++a) ok = (p->symtab[a] = table_symbol(n[a], p->sym)) != nullptr;
  if (!ok) {
    plan_destroy(p);
    return set_error(FF_ERR_NOMEM, "twiddle table allocation failed");
  }
  if (o.precision == FF_FP32_REFINE && own_scratch) {
    const size_t spec = (size_t)p->Nh * 16;
    const size_t res = ((size_t)p->N * 4 + 255) / 256 * 256;
    p->scratch_bytes = spec + res + 256;
    st = hip_check(hipMalloc(&p->scratch, p->scratch_bytes), "refinement work area hipMalloc");
    if (st != FF_OK) {
      plan_destroy(p);
      return st;
    }
    plan_bind_scratch(p, p->scratch);
  }
  // (four-step z rows: own_scratch false leaves the work array to the caller,
  // who points the plan, or a copy of it, at z_work_bytes of its own)
  if ((o.precision == FF_FP64 && (o.flags & FF_PRESERVE_INPUT)) || (p->fz.n && own_scratch)) {
    st = hip_check(hipMalloc(&p->work, p->fz.n ? z_work_bytes(p->n) : (size_t)p->Nh * 16), "work array hipMalloc");
    if (st != FF_OK) {
      plan_destroy(p);
      return st;
    }
  }
  *out = p;
  return FF_OK;
}

static size_t align256(size_t b) { return (b + 255) / 256 * 256; }

size_t z_work_bytes(const int n[3]) {
  const size_t nh = (size_t)n[2] / 2 + 1;
  size_t b = align256((size_t)n[0] * n[1] * nh * 16);
  if (n[2] % 2 && ((long long)n[0] * n[1]) % 2) b += align256(16 * (size_t)n[2]) + align256(32 * nh);
  return b;
}

size_t plan_work_bytes(const ff_plan* p) {
  if (!p->work) return 0;
  return p->fz.n ? z_work_bytes(p->n) : (size_t)p->Nh * 16;
}

bool z_rows_four_step(const int n[3], int precision) {
  for (int prec = 0; prec < (precision == FF_FP64 ? 1 : 2); ++prec) {
    RowCfg r0, r1;
    const int eb = prec ? 8 : 16;
    if (!make_row_cfg(n[2], eb, r0, 0, prec) || !make_row_cfg(n[2], eb, r1, 3, prec)) return true;
  }
  return false;
}

size_t plan_scratch_bytes(const ff_plan* p) {
  if (p->opt.precision != FF_FP32_REFINE) return 0;
  return (size_t)p->Nh * 16 + ((size_t)p->N * 4 + 255) / 256 * 256 + 256;
}

void plan_bind_scratch(ff_plan* p, void* base) {
  unsigned char* b = static_cast<unsigned char*>(base);
  const size_t spec = (size_t)p->Nh * 16;
  const size_t res = ((size_t)p->N * 4 + 255) / 256 * 256;
  p->spec = b;
  p->res = reinterpret_cast<float*>(b + spec);
  p->dc = reinterpret_cast<double*>(b + spec + res);
  p->dcf = p->dc + 1;
}

void plan_destroy(ff_plan* p) {
  if (!p) return;
  if (p->scratch) (void)hipFree(p->scratch);
  if (p->work) (void)hipFree(p->work);
  delete p;
}

// ---- passes -------------------------------------------------------------------

// A layout as seen by a column pass along axis (1 = y: planes i, positions j;
// 0 = x: planes j, positions i), or by the row passes (axis 1).  The pencil
// layouts address (col, outer, r) by strides, with the roles of the pass:
// y forward (col kh, outer i, r j), x (col j, outer kh, r i), y inverse
// (col i, outer kh, r j); the z inverse reads (row i*ny + j, kh).
enum { ROLE_Y = 1, ROLE_X = 0, ROLE_XT = 2, ROLE_YT = 3 };
static SpecSide spec_side(const ff_plan* p, int axis, int lay) {
  SpecSide s;
  memset(&s, 0, sizeof s);
  const long long nx = p->n[0], ny = p->n[1], nh = p->nh;
  if (lay == LAY_A || lay == LAY_B || lay == LAY_D) {
    // element strides of (i, j, kh)
    long long si, sj, sk;
    if (lay == LAY_A) si = nh * ny, sj = 1, sk = ny;        // [i][kh][j]
    else if (lay == LAY_B) si = 1, sj = nx, sk = ny * nx;   // [kh][j][i]
    else si = ny, sj = 1, sk = nx * ny;                     // [kh][i][j]
    s.lgKT = 0;  // blocks of one column: offset = col*KS + outer*fo + r*fr
    s.Wlast = 1;
    switch (axis) {
      case ROLE_Y: s.KS = sk, s.fo = si, s.fr = sj, s.nb = (int)nh; break;
      case ROLE_XT: s.KS = sj, s.fo = sk, s.fr = si, s.nb = (int)ny; break;
      case ROLE_YT: s.KS = si, s.fo = sk, s.fr = sj, s.nb = (int)nx; break;
      default: break;
    }
    return s;
  }
  if (lay == LAY_PLAIN) {
    s.lgKT = 30;  // one block of width nh
    s.nb = 1;
    s.Wlast = p->nh;
  } else {
    const int KT = 1 << p->lgKT;
    s.lgKT = p->lgKT;
    s.nb = (p->nh + KT - 1) / KT;
    s.Wlast = p->nh - (s.nb - 1) * KT;
    s.KS = nx * ny * KT;
  }
  if (lay == LAY_X) {  // [kb][j][i][kt]
    s.fo = axis == 1 ? 1 : nx;
    s.fr = axis == 1 ? nx : 1;
  } else {             // [i][j] inside each block
    s.fo = axis == 1 ? ny : 1;
    s.fr = axis == 1 ? 1 : ny;
  }
  return s;
}

void pass_layouts(const ff_plan* p, int pass, int& lay_in, int& lay_out) {
  lay_in = lay_out = LAY_PLAIN;
  if (p->pencil) {
    switch (pass) {
      case 1: lay_out = LAY_D; break;                  // y forward: [i][j][kh] -> [kh][i][j]
      case 2: lay_in = LAY_D; lay_out = LAY_B; break;  // x: [kh][i][j] -> [kh][j][i]
      case 3: lay_in = LAY_B; lay_out = LAY_A; break;  // y inverse: [kh][j][i] -> [i][kh][j]
      case 4: lay_in = LAY_A; break;                   // z inverse reads [i][kh][j]
      default: break;
    }
    return;
  }
  if (!p->lgKT) return;
  switch (pass) {
    case 1: lay_out = LAY_Y; break;                  // y forward: plain -> y-blocked
    case 2: lay_in = LAY_Y; lay_out = LAY_X; break;  // x: y-blocked -> x-blocked
    case 3: lay_in = LAY_X; lay_out = LAY_Y; break;  // y inverse: x-blocked -> y-blocked
    case 4: lay_in = LAY_Y; break;                   // z inverse reads y-blocked
    default: break;
  }
}

static RowArgs row_args(const ff_plan* p, int prec, bool inverse) {
  const RowCfg& c = inverse ? p->zi[prec] : p->zr[prec];
  RowArgs a;
  memset(&a, 0, sizeof a);
  a.nrows = (long long)p->n[0] * p->n[1];
  a.nz = p->n[2];
  a.nh = p->nh;
  a.ny = p->n[1];
  a.spec = spec_side(p, 1, LAY_PLAIN);
  a.T = c.T;
  a.Linv = c.Linv;
  a.nblocks = (int)((a.nrows + c.T - 1) / c.T);
  a.ps = c.ps;
  a.pitch = c.pitch;
  a.scale = 1.0 / (double(p->n[0]) * double(p->n[1]) * double(p->n[2]));  // ParisPencil.cpp:408
  a.shift = 0.0;
  a.shiftp = nullptr;
  a.tw = inverse ? p->tw_zi[prec] : p->tw_z[prec];
  a.twpost = p->tw_zpost[prec];
  a.lp = c.line.lp;
  a.nline = c.line.n;
  const void* const* bl = inverse ? p->bl_zi[prec] : p->bl_z[prec];
  a.chirp = bl[0];
  a.bsf = bl[1];
  a.bsi = bl[2];
  return a;
}

// z rows in four steps (ParisPencil.cpp:80-81, 165 and 380 transform z
// rows of any length): the packed row z[t] = x[2t] + i x[2t+1], t = t1 +
// m1*t2 < M = nz/2 = m1*m2, in rows of M complex values, is transformed as
//   A: for each t1, the m2-point DFT over t2 (stride m1), in place at t1 +
//      m1*k2 (column t1, stride 1), then times W_M^(t1*k2) (a row-local
//      kernel: the column kernels take their four-step twiddle by plane);
//   B: for each k2, the m1-point DFT over t1 (stride 1), written in natural
//      order, frequency k2 + m2*k1, into rows of nh;
// then the real-to-complex split in place (ff_zsplit.hip).  The inverse
// merges the half spectrum into packed rows of nh, runs B^-1 into rows of M,
// the conjugate twiddle and A^-1 in place, and the epilogue.  Whole passes,
// plain layout; the plan's work array (16*Nh bytes) holds the rows of M or
// the merged rows.  prec: 0 FP64, 1 FP32.
static ff_status run_zsub(const ff_plan* p, int prec, int mode, bool stepA, const void* src, void* dst,
                          hipStream_t s, int nrows = -1) {
  const ff_plan::FourStep& fs = p->fz;
  const long long m1 = fs.n1, m2 = fs.n2, M = fs.n;
  const ColCfg& c = stepA ? fs.c2[prec] : fs.c1[prec];
  // (odd nz: rows are the row pairs, and the natural order is written with a
  // pair stride of nz + 1 = 2*nh, the pair's two spectrum rows; nrows: the
  // lone last row's one pair)
  const int rows = nrows >= 0 ? nrows : p->zodd ? p->n[0] * p->n[1] / 2 : p->n[0] * p->n[1];
  if (!rows) return FF_OK;
  const long long nat = p->zodd ? 2LL * p->nh : p->nh;
  auto side = [](long long KS, long long fo, long long fr) {
    SpecSide sd;
    memset(&sd, 0, sizeof sd);
    sd.KS = KS;  // (lgKT 0: element (col, outer, r) at col*KS + outer*fo + r*fr)
    sd.fo = fo;
    sd.fr = fr;
    sd.lgKT = 0;
    sd.nb = INT_MAX;
    sd.Wlast = 1;
    return sd;
  };
  ColArgs a;
  memset(&a, 0, sizeof a);
  a.src = src;
  a.dst = dst;
  if (stepA) {
    a.in = a.out = side(1, M, m1);  // column t1, position t2 (k2)
    a.ncols = (int)m1;
  } else {
    const SpecSide packed = side(m1, M, 1);          // column k2, position t1
    const SpecSide natural = side(1, nat, m2);       // column k2, position k1: k2 + m2*k1
    a.in = mode == COL_FWD ? packed : natural;
    a.out = mode == COL_FWD ? natural : packed;
    a.ncols = (int)m2;
  }
  a.nouter = rows;
  a.T = c.T;
  a.lgT = c.lgT;
  a.tiles = (a.ncols + c.T - 1) / c.T;
  a.nblocks = a.nouter * a.tiles;  // bounded at plan creation
  a.ps = c.ps;
  a.nx = p->n[0];
  a.ny = p->n[1];
  a.tw = stepA ? fs.tw2[prec] : fs.tw1[prec];
  a.lp = c.line.lp;
  a.nline = c.line.n;
  const void* const* bl = stepA ? fs.bl2[prec] : fs.bl1[prec];
  a.chirp = bl[0];
  a.bsf = bl[1];
  a.bsi = bl[2];
  const Launch l{dim3(a.nblocks), dim3(c.threads), c.lds, s};
  return hip_check(launch_col(c.line.E, prec, mode, c.split, c.groups, c.line.bluestein(), c.stat, l, a),
                   "z four-step launch");
}

// Arguments of the row-local steps of four-step z rows.
static RowArgs zsplit_args(const ff_plan* p, int prec, const void* src, void* dst) {
  RowArgs a;
  memset(&a, 0, sizeof a);
  a.src = src;
  a.dst = dst;
  a.nrows = (long long)p->n[0] * p->n[1] / (p->zodd ? 2 : 1);  // (odd nz: row pairs)
  a.nz = p->n[2];
  a.nh = p->nh;
  a.ny = p->n[1];
  a.nline = p->fz.n;
  a.m1 = p->fz.n1;
  a.scale = 1.0 / (double(p->n[0]) * double(p->n[1]) * double(p->n[2]));  // ParisPencil.cpp:408
  a.twpost = p->tw_zpost[prec];
  a.tw = p->fz.twn[prec];
  return a;
}

static ff_status zsplit(int step, int prec, int out_kind, int epi, const RowArgs& a, hipStream_t s) {
  if (!a.nrows) return FF_OK;  // (odd nz, one row: only the lone row's launches)
  return hip_check(launch_zsplit(step, prec, out_kind, epi, a, s), "z four-step row-local launch");
}

// The lone last row of odd nz with nx*ny odd, in the work array's tail: its
// packed row (nz complex values) and its pair region (2*nh), one pair.
static unsigned char* zlone_packed(const ff_plan* p) {
  return static_cast<unsigned char*>(p->work) + align256((size_t)p->Nh * 16);
}
static unsigned char* zlone_natural(const ff_plan* p) { return zlone_packed(p) + align256(16 * (size_t)p->n[2]); }
static RowArgs zlone_args(const ff_plan* p, int prec, const void* src, void* dst) {
  RowArgs a = zsplit_args(p, prec, src, dst);
  a.nrows = 1;
  a.lone = 1;
  return a;
}

// The buffers of a four-step z pass must differ, except where noted.
static ff_status check_zfour(const ff_plan* p, long long nrows, std::initializer_list<const void*> distinct) {
  if (nrows >= 0) return set_error(FF_ERR_INVALID, "z rows in four steps: whole passes only");
  if (!p->work) return set_error(FF_ERR_INVALID, "z rows in four steps need the plan's work array");
  for (auto x = distinct.begin(); x != distinct.end(); ++x)
    for (auto y = x + 1; y != distinct.end(); ++y)
      if (*x == *y)
        return set_error(FF_ERR_INVALID, "z rows in four steps: source, destination and work array must differ");
  return FF_OK;
}

// Forward: A from src (rows of M; FP64 input of an FP32 pass is first
// rounded into the work array) into the work array, the twiddle, B into dst,
// the split in dst.  src is read only.
// Odd nz: the rows (2p, 2p+1) are packed as one complex row first, and the
// spectra separated in place at the end.
static ff_status zfwd_four(const ff_plan* p, int prec, int in_f32, const void* src, void* dst, hipStream_t s) {
  void* w = p->work;
  ff_status st;
  const void* a_in = src;
  if (p->zodd) {
    if ((st = zsplit(ZSPLIT_PPACK, prec, in_f32, 0, zsplit_args(p, prec, src, w), s))) return st;
    a_in = w;
  } else if (prec && !in_f32) {
    if ((st = zsplit(ZSPLIT_PACK, 1, 0, 0, zsplit_args(p, 1, src, w), s))) return st;
    a_in = w;
  }
  if ((st = run_zsub(p, prec, COL_FWD, true, a_in, w, s))) return st;
  if ((st = zsplit(ZSPLIT_TWF, prec, 0, 0, zsplit_args(p, prec, w, w), s))) return st;
  if ((st = run_zsub(p, prec, COL_FWD, false, w, dst, s))) return st;
  if ((st = zsplit(p->zodd ? ZSPLIT_PFWD : ZSPLIT_FWD, prec, 0, 0, zsplit_args(p, prec, dst, dst), s))) return st;
  if (!p->zlone) return FF_OK;
  // the lone last row: the same steps on its own pair in the work array's
  // tail, then its spectrum row (the pair's first) into dst
  const long long R = (long long)p->n[0] * p->n[1];
  const size_t csz = prec ? 8 : 16;
  unsigned char* tp = zlone_packed(p);
  unsigned char* tn = zlone_natural(p);
  const void* row = static_cast<const unsigned char*>(src) + (R - 1) * p->n[2] * (in_f32 ? 4 : 8);
  if ((st = zsplit(ZSPLIT_PPACK, prec, in_f32, 0, zlone_args(p, prec, row, tp), s))) return st;
  if ((st = run_zsub(p, prec, COL_FWD, true, tp, tp, s, 1))) return st;
  if ((st = zsplit(ZSPLIT_TWF, prec, 0, 0, zlone_args(p, prec, tp, tp), s))) return st;
  if ((st = run_zsub(p, prec, COL_FWD, false, tp, tn, s, 1))) return st;
  if ((st = zsplit(ZSPLIT_PFWD, prec, 0, 0, zlone_args(p, prec, tn, tn), s))) return st;
  return hip_check(hipMemcpyAsync(static_cast<unsigned char*>(dst) + (R - 1) * p->nh * csz, tn, p->nh * csz,
                                  hipMemcpyDeviceToDevice, s),
                   "z four-step lone row copy");
}

// Inverse.  The stored FP64 pass keeps src: the merge into dst, B^-1 into the
// work array, the twiddle and A^-1 there, the epilogue into dst (src may be
// the work array itself).  The other epilogues, whose dst holds phi or a
// residual too small for the merged rows, use src as their rows of M: the
// merge into the work array, B^-1 back into src, the twiddle and A^-1 there,
// the epilogue from src.  Odd nz: the merge forms the complex rows of the row
// pairs (natural order, pair stride 2*nh) and the epilogue unpacks them.
static ff_status zinv_four(const ff_plan* p, int prec, int out_kind, int epi, const void* src, void* dst,
                           const double* f, const double* shiftp, hipStream_t s) {
  void* w = p->work;
  ff_status st;
  RowArgs e;
  void* rows;
  const int merge = p->zodd ? ZSPLIT_PINV : ZSPLIT_INV;
  // the lone last row (odd nz, nx*ny odd) runs through the work array's
  // tail: its merge first, while src is whole (src may be the work array)
  const long long R = (long long)p->n[0] * p->n[1];
  const size_t csz = prec ? 8 : 16;
  unsigned char* tp = p->zlone ? zlone_packed(p) : nullptr;
  unsigned char* tn = p->zlone ? zlone_natural(p) : nullptr;
  if (p->zlone) {
    const void* row = static_cast<const unsigned char*>(src) + (R - 1) * p->nh * csz;
    if ((st = zsplit(ZSPLIT_PINV, prec, 0, 0, zlone_args(p, prec, row, tn), s))) return st;
  }
  if (!prec && !out_kind && epi == EPI_STORE && !f && !shiftp) {
    if ((st = zsplit(merge, 0, 0, 0, zsplit_args(p, 0, src, dst), s))) return st;
    if ((st = run_zsub(p, 0, COL_INV, false, dst, w, s))) return st;
    rows = w;
  } else {
    void* srcw = const_cast<void*>(src);
    if ((st = zsplit(merge, prec, 0, 0, zsplit_args(p, prec, src, w), s))) return st;
    if ((st = run_zsub(p, prec, COL_INV, false, w, srcw, s))) return st;
    rows = srcw;
  }
  if ((st = zsplit(ZSPLIT_TWI, prec, 0, 0, zsplit_args(p, prec, rows, rows), s))) return st;
  if ((st = run_zsub(p, prec, COL_INV, true, rows, rows, s))) return st;
  e = zsplit_args(p, prec, rows, dst);
  e.aux = f;
  if (shiftp) {
    e.shiftp = shiftp;
    e.shift = e.scale;  // shift = (sum of phi) / N
  }
  if ((st = zsplit(p->zodd ? ZSPLIT_PEPI : ZSPLIT_EPI, prec, out_kind, epi, e, s))) return st;
  if (!p->zlone) return FF_OK;
  if ((st = run_zsub(p, prec, COL_INV, false, tn, tp, s, 1))) return st;
  if ((st = zsplit(ZSPLIT_TWI, prec, 0, 0, zlone_args(p, prec, tp, tp), s))) return st;
  if ((st = run_zsub(p, prec, COL_INV, true, tp, tp, s, 1))) return st;
  // its epilogue writes (and reads f at) the last real row only
  const long long off = (R - 1) * p->n[2];
  RowArgs el = zlone_args(p, prec, tp, static_cast<unsigned char*>(dst) + off * (out_kind ? 4 : 8));
  el.aux = f ? f + off : nullptr;
  el.shiftp = e.shiftp;
  el.shift = e.shift;
  return zsplit(ZSPLIT_PEPI, prec, out_kind, epi, el, s);
}

ff_status run_zfwd(const ff_plan* p, int prec, int in_f32, const void* src, void* dst, hipStream_t s,
                   long long row0, long long nrows) {
  if (p->fz.n) {
    ff_status st = check_zfour(p, nrows, {src, dst, p->work});
    if (st) return st;
    return zfwd_four(p, prec, in_f32, src, dst, s);
  }
  RowArgs a = row_args(p, prec, false);
  const RowCfg& c = p->zr[prec];
  if (nrows >= 0) {
    a.nrows = nrows;
    a.nblocks = (int)((nrows + c.T - 1) / c.T);
    if (!nrows) return FF_OK;
  }
  a.src = static_cast<const unsigned char*>(src) + row0 * p->n[2] * (in_f32 ? 4 : 8);
  a.dst = static_cast<unsigned char*>(dst) + row0 * p->nh * (prec ? 8 : 16);
  const Launch l{dim3(a.nblocks), dim3(c.threads), c.lds, s};
  return hip_check(launch_r2c(c.line.E, prec, in_f32, p->n[2] % 2, c.line.bluestein(), c.stat, l, a), "z forward launch");
}

ff_status run_zinv(const ff_plan* p, int prec, int out_kind, int epi, const void* src, void* dst, const double* f,
                   const double* shiftp, hipStream_t s, int lay_in, long long row0, long long nrows) {
  if (p->fz.n) {
    if (lay_in != LAY_PLAIN) return set_error(FF_ERR_INVALID, "z rows in four steps: plain layout only");
    // (the stored FP64 pass may read the work array itself)
    const bool keep_src = !prec && !out_kind && epi == EPI_STORE && !f && !shiftp;
    ff_status st = keep_src ? check_zfour(p, nrows, {src, dst}) : check_zfour(p, nrows, {src, dst, p->work});
    if (!st && keep_src && dst == p->work) st = check_zfour(p, nrows, {dst, p->work});
    if (st) return st;
    return zinv_four(p, prec, out_kind, epi, src, dst, f, shiftp, s);
  }
  RowArgs a = row_args(p, prec, true);
  a.spec = spec_side(p, 1, lay_in);
  if (nrows >= 0) {
    if (lay_in != LAY_PLAIN) return set_error(FF_ERR_INVALID, "row ranges need the plain spectrum layout");
    a.nrows = nrows;
    a.nblocks = (int)((nrows + p->zi[prec].T - 1) / p->zi[prec].T);
    if (!nrows) return FF_OK;
  }
  a.src = static_cast<const unsigned char*>(src) + row0 * p->nh * (prec ? 8 : 16);
  a.dst = static_cast<unsigned char*>(dst) + row0 * p->n[2] * (out_kind ? 4 : 8);
  a.aux = f ? f + row0 * p->n[2] : nullptr;
  if (shiftp) {
    a.shiftp = shiftp;
    a.shift = a.scale;  // shift = (sum of phi) / N
  }
  const RowCfg& c = p->zi[prec];
  const Launch l{dim3(a.nblocks), dim3(c.threads), c.lds, s};
  return hip_check(launch_c2r(c.line.E, prec, out_kind, p->n[2] % 2, epi, c.line.bluestein(), c.stat, l, a),
                   "z inverse launch");
}

// Plane and position geometry of a four-step sub-line pass (plain layout, in
// rows of nh): planes (lo, hi) with lo < nlo, strides fo (lo), fo2 (hi) and
// fr (positions); the W_n twiddle and the x filter's i map.
struct SubLines {
  int nouter, nlo;
  long long fo, fo2, fr;
  const void* xtw;
  int imap_n2;
};

// One column pass launch with line configuration c (engine twiddles tw,
// Bluestein tables bl); sub: four-step geometry (null: whole lines).
static ff_status run_col_cfg(const ff_plan* p, int prec, int axis, int mode, const void* src, void* dst,
                             hipStream_t s, double* dc_out, int lay_in, int lay_out, int outer0, int nouter,
                             const ColCfg& c, const void* tw, const void* const* bl, const SubLines* sub) {
  if (src == dst && lay_in != lay_out) return set_error(FF_ERR_INVALID, "in-place column pass must keep its layout");
  if (mode == COL_FILTER && p->sym) mode = COL_FILTER_FD;
  if (mode == COL_SYMBOL && p->sym)
    return set_error(FF_ERR_INVALID, "finite-difference plans apply L_h with the stencil, not the x pass");
  ColArgs a;
  memset(&a, 0, sizeof a);
  a.src = src;
  a.dst = dst;
  a.ncols = p->nh;
  // y: planes i, positions j (ParisPencil.cpp:79 plan, batch nx*nh);
  // x: planes j, positions i (ParisPencil.cpp:78 plan, batch ny*nh)
  a.nouter = axis == 1 ? p->n[0] : p->n[1];
  // pencil layout: the x pass reads [i][kh][j] (columns j, planes kh) and the
  // y inverse reads [kh][j][i] (columns i, planes kh)
  int role = axis;
  if (axis == 0 && (lay_in == LAY_A || lay_in == LAY_D)) role = ROLE_XT;
  if (axis == 1 && lay_in == LAY_B) role = ROLE_YT;
  auto pencil = [](int l) { return l == LAY_A || l == LAY_B || l == LAY_D; };
  if ((pencil(lay_in) || pencil(lay_out)) &&
      !((role == ROLE_Y && lay_in == LAY_PLAIN && (lay_out == LAY_A || lay_out == LAY_D)) ||
        (role == ROLE_XT && lay_out == LAY_B) || (role == ROLE_YT && lay_out == LAY_A)))
    return set_error(FF_ERR_INVALID, "column pass: unsupported pencil layout pair");
  if (role == ROLE_XT) {
    a.ncols = p->n[1];
    a.nouter = p->nh;
    a.cols_j = 1;
  } else if (role == ROLE_YT) {
    a.ncols = p->n[0];
    a.nouter = p->nh;
  }
  a.in = spec_side(p, role, lay_in);
  a.out = spec_side(p, role, lay_out);
  if (sub) {
    a.nouter = sub->nouter;
    for (SpecSide* sd : {&a.in, &a.out}) {
      sd->nlo = sub->nlo;
      sd->fo = sub->fo;
      sd->fo2 = sub->fo2;
      sd->fr = sub->fr;
    }
    a.xtw = sub->xtw;
    a.imap_n2 = sub->imap_n2;
  }
  // four-step y lines leave the y frequencies of the x pass's j planes in
  // four-step order
  if (axis == 0 && p->fy.n) {
    a.jmap_n1 = p->fy.n1;
    a.jmap_n2 = p->fy.n2;
  }
  if (nouter >= 0) {
    // planes [outer0, outer0 + nouter) of a y pass over the plain layout
    if (axis != 1 || lay_in != LAY_PLAIN || lay_out != LAY_PLAIN)
      return set_error(FF_ERR_INVALID, "plane ranges need a y pass over the plain spectrum layout");
    if (!nouter) return FF_OK;
    const long long off = (long long)outer0 * p->n[1] * p->nh * (prec ? 8 : 16);
    a.src = static_cast<const unsigned char*>(src) + off;
    a.dst = static_cast<unsigned char*>(dst) + off;
    a.nouter = nouter;
  }
  a.outer_fast = 0;
  a.T = c.T;
  a.lgT = c.lgT;
  // Kernels with a compiled-in plan, plain layout on both sides, rows that
  // start 64-byte aligned within a plane but planes that do not (the x pass at
  // odd nh): shift each plane's tiles onto the 64-byte grid of its rows.  The
  // tile count covers the shift.
  a.cshift_mask = a.cshift_mul = 0;
  if (c.stat) {
    const long long mask = prec ? 7 : 3;  // complex elements per 64 bytes, minus one
    const long long W = a.out.Wlast;
    if (a.in.nb == 1 && a.out.nb == 1 && ((a.out.fr * W) & mask) == 0 && ((a.out.fo * W) & mask) != 0) {
      a.cshift_mask = (int)mask;
      a.cshift_mul = (int)((a.out.fo * W) & mask);
    }
  }
  a.tiles = (a.ncols + a.cshift_mask + c.groups * c.T - 1) / (c.groups * c.T);
  const long long nb = (long long)a.nouter * a.tiles;
  if (nb > INT_MAX) return set_error(FF_ERR_INVALID, "grid too large");
  a.nblocks = (int)nb;
  a.ps = c.ps;
  a.nx = p->n[0];
  a.ny = p->n[1];
  if (p->sym) {
    a.ddi = p->fd[0];
    a.ddj = p->fd[1];
    a.ddk = p->fd[2];
    a.symx = p->symtab[0];
    a.symy = p->symtab[1];
    a.symz = p->symtab[2];
  } else {
    a.ddi = p->ddi;
    a.ddj = p->ddj;
    a.ddk = p->ddk;
  }
  a.dc_out = dc_out;
  a.tw = tw;
  a.lp = c.line.lp;
  a.nline = c.line.n;
  a.chirp = bl[0];
  a.bsf = bl[1];
  a.bsi = bl[2];
  // kernels with LDS twiddles also copy the x axis's finite-difference
  // factors to LDS, after the twiddles
  const size_t lds = c.lds + (mode == COL_FILTER_FD && c.stat == 2 ? (size_t)c.line.lp.n * sizeof(double) : 0);
  const Launch l{dim3(a.nblocks), dim3(c.threads), lds, s};
  return hip_check(launch_col(c.line.E, prec, mode, c.split, c.groups, c.line.bluestein(), c.stat, l, a),
                   "column pass launch");
}

// Four-step lines (ParisPencil.cpp:78-79 plans any length through
// cufftPlanMany): a line x[j], j = j1 + n1*j2, of length n = n1*n2 is
//   A: for each j1, the n2-point DFT over j2 (stride n1), times W_n^(j1*k2),
//      in place at j1 + n1*k2;
//   B: for each k2, the n1-point DFT over j1 (stride 1), in place: frequency
//      k2 + n2*k1 lands at k1 + n1*k2.
// The inverse runs B^-1, then the conjugate twiddle and A^-1, and restores
// natural order.  The x pass between the y passes reads its j planes in that
// order (jmap); an x pass of four-step lines runs A, B with the filter at
// frequency i = k2 + n2*k1, then A^-1.
ff_status run_col(const ff_plan* p, int prec, int axis, int mode, const void* src, void* dst, hipStream_t s,
                  double* dc_out, int lay_in, int lay_out, int outer0, int nouter) {
  const ff_plan::FourStep& fs = axis == 1 ? p->fy : p->fx;
  if (!fs.n) {
    const ColCfg& c = axis == 1 ? p->cy[prec] : p->cx[prec];
    return run_col_cfg(p, prec, axis, mode, src, dst, s, dc_out, lay_in, lay_out, outer0, nouter, c,
                       axis == 1 ? p->tw_y[prec] : p->tw_x[prec], axis == 1 ? p->bl_y[prec] : p->bl_x[prec],
                       nullptr);
  }
  if (lay_in != LAY_PLAIN || lay_out != LAY_PLAIN || nouter >= 0)
    return set_error(FF_ERR_INVALID, "four-step column lines run over whole passes in the plain layout");
  const long long n1 = fs.n1, n2 = fs.n2, nx = p->n[0], ny = p->n[1];
  // geometry in rows: y lines (planes i, row j), x lines (planes j, row i*ny)
  SubLines ga, gb;
  if (axis == 1) {
    ga = SubLines{(int)(nx * n1), (int)n1, 1, ny, n1, fs.twn[prec], 0};
    gb = SubLines{(int)(nx * n2), (int)n2, n1, ny, 1, nullptr, 0};
  } else {
    ga = SubLines{(int)(ny * n1), (int)n1, ny, 1, n1 * ny, fs.twn[prec], 0};
    gb = SubLines{(int)(ny * n2), (int)n2, n1 * ny, 1, ny, nullptr, (int)n2};
  }
  auto A = [&](int m, const void* in, void* out) {
    return run_col_cfg(p, prec, axis, m, in, out, s, nullptr, LAY_PLAIN, LAY_PLAIN, 0, -1, fs.c2[prec], fs.tw2[prec],
                       fs.bl2[prec], &ga);
  };
  auto B = [&](int m, const void* in, void* out, double* dc) {
    return run_col_cfg(p, prec, axis, m, in, out, s, dc, LAY_PLAIN, LAY_PLAIN, 0, -1, fs.c1[prec], fs.tw1[prec],
                       fs.bl1[prec], &gb);
  };
  ff_status st;
  if (mode == COL_FWD) {
    if ((st = A(COL_FWD, src, dst))) return st;
    return B(COL_FWD, dst, dst, nullptr);
  }
  if (mode == COL_INV) {
    if ((st = B(COL_INV, src, dst, nullptr))) return st;
    return A(COL_INV, dst, dst);
  }
  // the x pass: forward, filter / symbol at the four-step frequencies, inverse
  if ((st = A(COL_FWD, src, dst))) return st;
  if ((st = B(mode, dst, dst, dc_out))) return st;
  return A(COL_INV, dst, dst);
}

// The FP32 transform round trip of the solve and of each correction, from
// src (FP64 f or the FP32 residual r) to phi (stored or accumulated): in
// place in the work area (plain layout), or, with the pencil layout, between
// its two halves (each holds one FP32 spectrum).
static ff_status fp32_round_trip(const ff_plan* p, int in_f32, const void* src, double* phi, int epi, double* dc_col,
                                 const double* shiftp, hipStream_t s) {
  void* a = p->spec;
  void* b = p->pencil ? static_cast<void*>(static_cast<char*>(p->spec) + p->Nh * 8) : p->spec;
  ff_status st;
  if ((st = run_zfwd(p, 1, in_f32, src, a, s))) return st;
  if (p->pencil) {
    if ((st = run_col(p, 1, 1, COL_FWD, a, b, s, nullptr, LAY_PLAIN, LAY_D))) return st;
    if ((st = run_col(p, 1, 0, COL_FILTER, b, a, s, dc_col, LAY_D, LAY_B))) return st;
    if ((st = run_col(p, 1, 1, COL_INV, a, b, s, nullptr, LAY_B, LAY_A))) return st;
    return run_zinv(p, 1, 0, epi, b, phi, nullptr, shiftp, s, LAY_A);
  }
  if ((st = run_col(p, 1, 1, COL_FWD, a, a, s))) return st;
  if ((st = run_col(p, 1, 0, COL_FILTER, a, a, s, dc_col))) return st;
  if ((st = run_col(p, 1, 1, COL_INV, a, a, s))) return st;
  return run_zinv(p, 1, 0, epi, a, phi, nullptr, shiftp, s);
}

static ff_status solve_fp32_refine(const ff_plan* p, const double* f, double* phi, hipStream_t s) {
  void* spec = p->spec;
  ff_status st;
  // phi0: FP32 passes, FP64 in and out
  if ((st = fp32_round_trip(p, 0, f, phi, EPI_STORE, p->dcf, nullptr, s))) return st;
  for (int it = 0; it < p->opt.refine_steps; ++it) {
    if (p->sym) {
      // finite-difference symbol: r = (f - mean f) - L_h phi with the stencil
      // in real space, one pass over f and phi; sum(phi) by a two-stage
      // reduction through the (idle) spectrum area
      if ((st = run_stencil(p, phi, f, p->res, false, p->dcf, static_cast<double*>(spec), p->dc, s))) return st;
    } else {
      // r = f - L_h phi in FP64, with the same spectral symbol the filter inverts;
      // the x pass also captures sum(phi) for the FP64 mean removal
      if ((st = run_zfwd(p, 0, 0, phi, spec, s))) return st;
      if ((st = run_col(p, 0, 1, COL_FWD, spec, spec, s))) return st;
      if ((st = run_col(p, 0, 0, COL_SYMBOL, spec, spec, s, p->dc))) return st;
      if ((st = run_col(p, 0, 1, COL_INV, spec, spec, s))) return st;
      // r = (f - mean f) - L_h phi: L_h phi has no mean, and removing f's mean
      // keeps the FP32 rounding of r relative to the residual itself
      if ((st = run_zinv(p, 0, 1, EPI_RESIDUAL, spec, p->res, f, p->dcf, s))) return st;
    }
    // correction in FP32: phi += L_h^-1 r - mean(phi)
    if ((st = fp32_round_trip(p, 1, p->res, phi, EPI_ACCUM, nullptr, p->dc, s))) return st;
  }
  return FF_OK;
}

ff_status solve_enqueue(const ff_plan* p, double* density, double* potential, hipStream_t s, hipEvent_t* ev) {
  if (p->opt.precision == FF_FP32_REFINE) {
    if (!p->spec) return set_error(FF_ERR_INVALID, "FP32 plan has no refinement work area bound");
    return solve_fp32_refine(p, density, potential, s);
  }
  ff_status st;
  // ping-pong density <-> potential (or potential <-> the plan's work array
  // with FF_PRESERVE_INPUT); the result lands in potential as in
  // ParisPencil.cpp:103-104,419
  void* w = p->work ? p->work : (void*)density;
  int li[5], lo[5];
  for (int k = 0; k < 5; ++k) pass_layouts(p, k, li[k], lo[k]);
  if (p->slab) {
    // Slab pipeline: each pair of passes runs slab by slab, so the second
    // pass reads the slab the first just wrote while it is cache-resident.
    // The spectrum of a plane (ny*nh complex) is larger than its real data
    // (ny*nz doubles), so when w aliases density the forward pair walks the
    // slabs from the top (a slab's y output never reaches real planes not yet
    // read) and the inverse pair from the bottom (a slab's real output never
    // reaches spectrum planes not yet read).
    const int nx = p->n[0], S = p->slab;
    const long long ny = p->n[1];
    const int nsl = (nx + S - 1) / S;
    for (int k = nsl - 1; k >= 0; --k) {
      const int i0 = k * S, ni = std::min(S, nx - i0);
      if ((st = run_zfwd(p, 0, 0, density, potential, s, i0 * ny, ni * ny))) return st;
      if ((st = run_col(p, 0, 1, COL_FWD, potential, w, s, nullptr, LAY_PLAIN, LAY_PLAIN, i0, ni))) return st;
    }
    if (ev && (st = hip_check(hipEventRecord(ev[0], s), "event"))) return st;
    if (ev && (st = hip_check(hipEventRecord(ev[1], s), "event"))) return st;
    if ((st = run_col(p, 0, 0, COL_FILTER, w, potential, s))) return st;
    if (ev && (st = hip_check(hipEventRecord(ev[2], s), "event"))) return st;
    for (int k = 0; k < nsl; ++k) {
      const int i0 = k * S, ni = std::min(S, nx - i0);
      if ((st = run_col(p, 0, 1, COL_INV, potential, w, s, nullptr, LAY_PLAIN, LAY_PLAIN, i0, ni))) return st;
      if ((st = run_zinv(p, 0, 0, EPI_STORE, w, potential, nullptr, nullptr, s, LAY_PLAIN, i0 * ny, ni * ny)))
        return st;
    }
    if (ev && (st = hip_check(hipEventRecord(ev[3], s), "event"))) return st;
    if (ev && (st = hip_check(hipEventRecord(ev[4], s), "event"))) return st;
    return FF_OK;
  }
  if ((st = run_zfwd(p, 0, 0, density, potential, s))) return st;
  if (ev && (st = hip_check(hipEventRecord(ev[0], s), "event"))) return st;
  if ((st = run_col(p, 0, 1, COL_FWD, potential, w, s, nullptr, li[1], lo[1]))) return st;
  if (ev && (st = hip_check(hipEventRecord(ev[1], s), "event"))) return st;
  if ((st = run_col(p, 0, 0, COL_FILTER, w, potential, s, nullptr, li[2], lo[2]))) return st;
  if (ev && (st = hip_check(hipEventRecord(ev[2], s), "event"))) return st;
  if ((st = run_col(p, 0, 1, COL_INV, potential, w, s, nullptr, li[3], lo[3]))) return st;
  if (ev && (st = hip_check(hipEventRecord(ev[3], s), "event"))) return st;
  if ((st = run_zinv(p, 0, 0, EPI_STORE, w, potential, nullptr, nullptr, s, li[4]))) return st;
  if (ev && (st = hip_check(hipEventRecord(ev[4], s), "event"))) return st;
  return FF_OK;
}

}  // namespace ff
