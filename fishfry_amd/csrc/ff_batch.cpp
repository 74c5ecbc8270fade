// The batched solve: many independent grids of uneven shapes, each with its
// own plan, launched stream-parallel, grouped by shape, or step by step
// through the multi-shape kernels (ff_multi.hpp), optionally captured in a
// HIP graph, and the autotuner.
#include <algorithm>
#include <climits>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>
#include <new>
#include <sstream>
#include <string>
#include <tuple>
#include <vector>

#include "ff_internal.hpp"
#include "ff_multi.hpp"
#include "ff_solver.hpp"

using namespace ff;

// The C-ABI batch.  Each instance runs the single-grid pass kernels with its
// own plan; the instances are spread over a few streams (fork on the caller's
// stream, join back onto it) so their kernels overlap on the device, and with
// FF_BATCH_GRAPH the whole fork/join sequence is captured once as a HIP graph
// and replayed while the buffers stay the same.
//
// With FF_BATCH_GROUPED, instances of one shape, precision and refinement
// count form a group that runs each of its passes as ONE launch, grid row
// blockIdx.y being the instance: the pass arguments of the group's instances
// are equal except for their buffers and domain constants, which the kernels
// read from a per-instance table (PassInst) in device memory.  The launch
// sequence comes from recording one ordinary solve per instance (the
// launchers' recorder), so a group runs exactly the kernels and arithmetic of
// its instances' single solves.
//
// FF_BATCH_PERSISTENT, a single persistent launch whose workgroups waited on
// each other's completion counters, is rejected: its kernel did not complete
// on MI355X, the cause was never established, and it was removed.  The
// step-batched mode below reaches one launch per pass step with no
// workgroup waiting on another.
//
// With FF_BATCH_STEPS, pass step s of every instance of a family (instances
// whose solves launch the same kernel roles step by step: same precision and
// refinement count) is ONE launch of a multi-shape kernel (ff_multi.hpp): a
// descriptor per instance holds the pass arguments its single solve would
// launch with, and each workgroup finds its instance's descriptor and tile.
// Instances whose line length has no compiled-in multi-shape variant run that
// step as grouped launches (one per shape) beside it, on the same stream.
struct BatchGroup {
  std::vector<int> inst;         // instances, in batch order
  std::vector<PassLaunch> ops;   // the group's launches (per-instance fields from the table)
};

struct StepLaunch {
  int kind = -1;  // MULTI_COL / MULTI_R2C / MULTI_C2R: a multi-shape launch; -1: op
  int f32 = 0, sel = 0, epi = 0;
  size_t lds = 0;
  MultiArgs m{};  // descriptor table and first workgroups (device)
  PassLaunch op{};
};

struct StepFamily {
  std::vector<int> inst;
  int steps = 0;
  int multi_launches = 0;
  std::vector<StepLaunch> ops;
};

struct ff_batch {
  int count = 0;
  int flags = 0;
  std::vector<ff_batch_item> items;
  std::vector<ff_plan*> plans;  // per instance
  std::vector<hipStream_t> streams;
  std::vector<hipEvent_t> joins;
  hipEvent_t fork = nullptr;
  std::vector<double*> bd, bp;
  bool bound = false;
  // FF_BATCH_GROUPED: groups and their per-instance tables, valid for the
  // plans and the buffers gd / gp they were recorded with
  std::vector<BatchGroup> groups;
  void* inst_dev = nullptr;
  size_t inst_cap = 0;
  bool groups_ok = false;
  std::vector<double*> gd, gp;
  // FF_BATCH_STEPS: families and their descriptor tables, valid for the
  // plans and the buffers sd / sp they were built with
  std::vector<StepFamily> fams;
  void* step_dev = nullptr;
  size_t step_cap = 0;
  bool steps_ok = false;
  std::vector<double*> sd, sp;
  hipGraphExec_t exec = nullptr;
  hipGraph_t graph = nullptr;
  hipStream_t graph_stream = nullptr;
};

namespace {

constexpr int kBatchStreams = 4;

// Instances the grouped and step modes solve on their own: grids whose z
// rows run in four steps (their row-local steps are not recorded launches).
bool solo(const ff_batch* b, int i) { return b->plans[i]->fz.n != 0; }

void drop_graph(ff_batch* b) {
  if (b->exec) (void)hipGraphExecDestroy(b->exec);
  if (b->graph) (void)hipGraphDestroy(b->graph);
  b->exec = nullptr;
  b->graph = nullptr;
}

void destroy_plans(ff_batch* b) {
  for (ff_plan* p : b->plans) plan_destroy(p);
  b->plans.clear();
}

ff_status make_plans(ff_batch* b) {
  destroy_plans(b);
  b->groups_ok = false;
  b->steps_ok = false;
  for (const ff_batch_item& it : b->items) {
    ff_options o{it.precision, it.refine_steps, 0};
    ff_plan* p = nullptr;
    ff_status st = plan_create(it.n, it.lo, it.hi, &o, &p);
    if (st != FF_OK) return st;
    b->plans.push_back(p);
  }
  return FF_OK;
}

// Records one solve per instance, forms the groups and uploads their
// per-instance tables (blocking: waits for the device before rewriting them).
ff_status prepare_groups(ff_batch* b, double* const* d, double* const* p) {
  b->groups_ok = false;  // the old groups' table may be freed below
  b->groups.clear();
  std::vector<std::vector<PassLaunch>> rec(b->count);
  for (int i = 0; i < b->count; ++i) {
    if (solo(b, i)) continue;
    set_launch_recorder(&rec[i]);
    const ff_status st = solve_enqueue(b->plans[i], d[i], p[i], nullptr, nullptr);
    set_launch_recorder(nullptr);
    if (st != FF_OK) return st;
  }
  std::vector<BatchGroup> groups;
  for (int i = 0; i < b->count; ++i) {
    if (solo(b, i)) continue;
    const ff_batch_item& it = b->items[i];
    BatchGroup* home = nullptr;
    for (BatchGroup& g : groups) {
      const int l = g.inst[0];
      const ff_batch_item& lt = b->items[l];
      if (lt.n[0] != it.n[0] || lt.n[1] != it.n[1] || lt.n[2] != it.n[2] || lt.precision != it.precision ||
          lt.refine_steps != it.refine_steps || rec[l].size() != rec[i].size())
        continue;
      bool same = true;
      for (size_t k = 0; k < rec[i].size() && same; ++k) same = launch_same(rec[l][k], rec[i][k]);
      if (same) {
        home = &g;
        break;
      }
    }
    if (!home) {
      groups.emplace_back();
      home = &groups.back();
    }
    home->inst.push_back(i);
  }
  std::vector<PassInst> tab;
  std::vector<size_t> off;
  for (const BatchGroup& g : groups) {
    off.push_back(tab.size());
    for (size_t k = 0; k < rec[g.inst[0]].size(); ++k)
      for (int i : g.inst) tab.push_back(launch_inst_fields(rec[i][k]));
  }
  // kernels of earlier solves may still read the old tables
  ff_status st = hip_check(hipDeviceSynchronize(), "batch table sync");
  const size_t bytes = std::max<size_t>(1, tab.size()) * sizeof(PassInst);
  if (st == FF_OK && bytes > b->inst_cap) {
    if (b->inst_dev) (void)hipFree(b->inst_dev);
    b->inst_dev = nullptr;
    b->inst_cap = 0;
    st = hip_check(hipMalloc(&b->inst_dev, bytes), "batch table hipMalloc");
    if (st == FF_OK) b->inst_cap = bytes;
  }
  if (st == FF_OK && !tab.empty())
    st = hip_check(hipMemcpy(b->inst_dev, tab.data(), tab.size() * sizeof(PassInst), hipMemcpyHostToDevice),
                   "batch table upload");
  if (st != FF_OK) return st;
  const PassInst* base = static_cast<const PassInst*>(b->inst_dev);
  for (size_t gi = 0; gi < groups.size(); ++gi) {
    BatchGroup& g = groups[gi];
    const std::vector<PassLaunch>& lead = rec[g.inst[0]];
    for (size_t k = 0; k < lead.size(); ++k) {
      PassLaunch r = lead[k];
      r.l.grid.y = (unsigned)g.inst.size();
      const PassInst* t = base + off[gi] + k * g.inst.size();
      if (r.kind == 0) r.col.inst = t;
      else r.row.inst = t;
      g.ops.push_back(r);
    }
  }
  b->groups.swap(groups);
  b->gd.assign(d, d + b->count);
  b->gp.assign(p, p + b->count);
  b->groups_ok = true;
  return FF_OK;
}

bool groups_current(const ff_batch* b, double* const* d, double* const* p) {
  return b->groups_ok && std::equal(b->gd.begin(), b->gd.end(), d) && std::equal(b->gp.begin(), b->gp.end(), p);
}

// ---- FF_BATCH_STEPS -----------------------------------------------------------

// The multi-shape descriptor of a recorded column pass, or false when no
// multi-shape variant runs it (Bluestein lines, tuning variants, blocked
// layouts, lengths without a compiled-in variant).  The descriptor keeps the
// recorded arguments and takes the variant's tile width; lds: the variant's
// LDS bytes.
bool col_desc(const PassLaunch& r, MultiColDesc& out, size_t& lds) {
  const ColArgs& c = r.col;
  if (r.blue || r.groups != 1 || r.split || c.outer_fast) return false;
  if (c.in.nlo || c.xtw || c.imap_n2 || c.jmap_n1) return false;  // four-step lines
  if (c.in.nb != 1 || c.out.nb != 1 || c.in.lgKT < 30 || c.out.lgKT < 30 || c.nline != c.lp.n) return false;
  if (!multi_kernel_exists(MULTI_COL, r.f32, r.sel, 0)) return false;
  const int v = multi_col_variant(c.lp.n, r.E, c.lp);
  if (v < 0) return false;
  const MultiVar mv = kMultiCols[v];
  const int T = multi_tile(mv);
  ColArgs a = c;
  a.inst = nullptr;
  a.T = T;
  a.lgT = 0;
  while ((1 << a.lgT) < T) a.lgT++;
  a.ps = kColPadShift;
  // rows of a plane 64-byte aligned but planes not: shift the tiles onto the
  // 64-byte grid, as the single-solve kernels with compiled-in plans do
  const long long mask = r.f32 ? 7 : 3;
  const long long W = a.out.Wlast;
  a.cshift_mask = a.cshift_mul = 0;
  if (((a.out.fr * W) & mask) == 0 && ((a.out.fo * W) & mask) != 0) {
    a.cshift_mask = (int)mask;
    a.cshift_mul = (int)((a.out.fo * W) & mask);
  }
  a.tiles = (a.ncols + a.cshift_mask + T - 1) / T;
  const long long nb = (long long)a.nouter * a.tiles;
  if (nb > INT_MAX / 2) return false;
  a.nblocks = (int)nb;
  memset(&out, 0, sizeof out);
  out.a = a;
  out.variant = v;
  lds = ((size_t)mv.n * T + ((mv.n - 1) >> kColPadShift) + 1) * (r.f32 ? 8 : 16);
  return true;
}

bool row_desc(const PassLaunch& r, MultiRowDesc& out, size_t& lds) {
  const RowArgs& c = r.row;
  if (r.blue || r.odd || c.nline != c.lp.n) return false;
  if (r.kind == 2 && c.spec.lgKT < 30) return false;  // z inverse of a blocked spectrum
  const int kind = r.kind == 1 ? MULTI_R2C : MULTI_C2R;
  if (!multi_kernel_exists(kind, r.f32, r.sel, r.kind == 2 ? r.epi : 0)) return false;
  const int v = multi_row_variant(c.lp.n, r.E, c.lp);
  if (v < 0) return false;
  const MultiVar mv = kMultiRows[v];
  const int T = multi_tile(mv);
  RowArgs a = c;
  a.inst = nullptr;
  a.T = T;
  a.ps = 4;
  a.pitch = (mv.n + 1) + (mv.n >> 4) + 1;
  const long long nb = (a.nrows + T - 1) / T;
  if (nb > INT_MAX / 2) return false;
  a.nblocks = (int)nb;
  memset(&out, 0, sizeof out);
  out.a = a;
  out.variant = v;
  lds = (size_t)T * a.pitch * (r.f32 ? 8 : 16);
  return true;
}

// Launch role of a recorded launch: kernel kind, precision, selector, epilogue.
std::tuple<int, int, int, int> launch_role(const PassLaunch& r) {
  return std::make_tuple(r.kind, r.f32, r.sel, r.kind == 2 ? r.epi : 0);
}

void push_bytes(std::vector<unsigned char>& img, const void* p, size_t n, size_t& off) {
  off = (img.size() + 255) / 256 * 256;
  img.resize(off + n);
  memcpy(img.data() + off, p, n);
}

// Records one solve per instance, forms the families, builds every step's
// descriptor table (and the grouped tables of the steps' remaining
// instances) and uploads them (blocking: waits for the device first).
ff_status prepare_steps(ff_batch* b, double* const* d, double* const* p) {
  b->steps_ok = false;
  b->fams.clear();
  std::vector<std::vector<PassLaunch>> rec(b->count);
  for (int i = 0; i < b->count; ++i) {
    if (solo(b, i)) continue;
    set_launch_recorder(&rec[i]);
    const ff_status st = solve_enqueue(b->plans[i], d[i], p[i], nullptr, nullptr);
    set_launch_recorder(nullptr);
    if (st != FF_OK) return st;
  }
  std::vector<StepFamily> fams;
  for (int i = 0; i < b->count; ++i) {
    if (solo(b, i)) continue;
    StepFamily* home = nullptr;
    for (StepFamily& f : fams) {
      const int l = f.inst[0];
      if (rec[l].size() != rec[i].size()) continue;
      bool same = true;
      for (size_t k = 0; k < rec[i].size() && same; ++k) same = launch_role(rec[l][k]) == launch_role(rec[i][k]);
      if (same) {
        home = &f;
        break;
      }
    }
    if (!home) {
      fams.emplace_back();
      home = &fams.back();
      home->steps = (int)rec[i].size();
    }
    home->inst.push_back(i);
  }
  // Device image: per step, the descriptors and first workgroups of the
  // multi-shape launch, then the per-instance tables of its grouped launches.
  // Pointers into the image are stored as offsets and rebased after upload.
  std::vector<unsigned char> img;
  struct Fix {
    size_t fam, op;
    size_t desc, block0, inst;
  };
  std::vector<Fix> fixes;
  for (size_t fi = 0; fi < fams.size(); ++fi) {
    StepFamily& f = fams[fi];
    for (int s = 0; s < f.steps; ++s) {
      const PassLaunch& lead = rec[f.inst[0]][s];
      struct Member {
        int variant, inst;
        size_t lds;
        double weight;  // bytes one tile moves
        MultiColDesc c;
        MultiRowDesc r;
      };
      std::vector<Member> multi;
      std::vector<int> rest;
      for (int i : f.inst) {
        Member m;
        m.inst = i;
        bool ok;
        const double eb = rec[i][s].f32 ? 8.0 : 16.0;
        if (lead.kind == 0) {
          ok = col_desc(rec[i][s], m.c, m.lds);
          m.variant = m.c.variant;
          if (ok) m.weight = 2.0 * eb * kMultiCols[m.variant].n * multi_tile(kMultiCols[m.variant]);
        } else {
          ok = row_desc(rec[i][s], m.r, m.lds);
          m.variant = m.r.variant;
          if (ok) m.weight = 2.0 * eb * kMultiRows[m.variant].n * multi_tile(kMultiRows[m.variant]);
        }
        if (ok) multi.push_back(m);
        else rest.push_back(i);
      }
      if (!multi.empty()) {
        // one variant's tiles after another: neighbouring workgroups run the same code
        std::stable_sort(multi.begin(), multi.end(),
                         [](const Member& x, const Member& y) { return x.variant < y.variant; });
        StepLaunch sl;
        sl.kind = lead.kind == 0 ? MULTI_COL : (lead.kind == 1 ? MULTI_R2C : MULTI_C2R);
        sl.f32 = lead.f32;
        sl.sel = lead.sel;
        sl.epi = lead.kind == 2 ? lead.epi : 0;
        std::vector<int> block0;
        long long nb = 0;
        std::vector<unsigned char> descs;
        for (const Member& m : multi) {
          block0.push_back((int)nb);
          sl.lds = std::max(sl.lds, m.lds);
          const unsigned char* src = lead.kind == 0 ? (const unsigned char*)&m.c : (const unsigned char*)&m.r;
          const size_t sz = lead.kind == 0 ? sizeof(MultiColDesc) : sizeof(MultiRowDesc);
          descs.insert(descs.end(), src, src + sz);
          nb += lead.kind == 0 ? m.c.a.nblocks : m.r.a.nblocks;
        }
        if (nb > INT_MAX / 2) return set_error(FF_ERR_UNSUPPORTED, "batch step too large for one launch");
        sl.m.ndesc = (int)multi.size();
        // cut the tiles into kXcds contiguous chunks of about equal work
        double total = 0.0;
        for (const Member& m : multi) total += m.weight * (lead.kind == 0 ? m.c.a.nblocks : m.r.a.nblocks);
        int x = 0;
        long long t = 0;
        double acc = 0.0;
        sl.m.chunk0[0] = 0;
        for (const Member& m : multi) {
          const long long cnt = lead.kind == 0 ? m.c.a.nblocks : m.r.a.nblocks;
          for (long long k = 0; k < cnt; ++k, ++t) {
            if (x + 1 < kXcds && acc >= total * (x + 1) / kXcds) {
              sl.m.chunkn[x] = (int)(t - sl.m.chunk0[x]);
              sl.m.chunk0[++x] = (int)t;
            }
            acc += m.weight;
          }
        }
        sl.m.chunkn[x] = (int)(t - sl.m.chunk0[x]);
        for (int y = x + 1; y < kXcds; ++y) {
          sl.m.chunk0[y] = (int)t;
          sl.m.chunkn[y] = 0;
        }
        int most = 0;
        for (int y = 0; y < kXcds; ++y) most = std::max(most, sl.m.chunkn[y]);
        sl.m.nblocks = kXcds * most;
        Fix fx{fi, f.ops.size(), 0, 0, SIZE_MAX};
        push_bytes(img, descs.data(), descs.size(), fx.desc);
        push_bytes(img, block0.data(), block0.size() * sizeof(int), fx.block0);
        fixes.push_back(fx);
        f.ops.push_back(sl);
        f.multi_launches++;
      }
      // the rest: one grouped launch per set of identical launches
      std::vector<std::vector<int>> clusters;
      for (int i : rest) {
        std::vector<int>* home = nullptr;
        for (auto& c : clusters)
          if (launch_same(rec[c[0]][s], rec[i][s])) {
            home = &c;
            break;
          }
        if (!home) {
          clusters.emplace_back();
          home = &clusters.back();
        }
        home->push_back(i);
      }
      for (const auto& c : clusters) {
        std::vector<PassInst> tab;
        for (int i : c) tab.push_back(launch_inst_fields(rec[i][s]));
        StepLaunch sl;
        sl.kind = -1;
        sl.op = rec[c[0]][s];
        sl.op.l.grid.y = (unsigned)c.size();
        Fix fx{fi, f.ops.size(), SIZE_MAX, SIZE_MAX, 0};
        push_bytes(img, tab.data(), tab.size() * sizeof(PassInst), fx.inst);
        fixes.push_back(fx);
        f.ops.push_back(sl);
      }
    }
  }
  // kernels of earlier solves may still read the old tables
  ff_status st = hip_check(hipDeviceSynchronize(), "batch table sync");
  const size_t bytes = std::max<size_t>(256, img.size());
  if (st == FF_OK && bytes > b->step_cap) {
    if (b->step_dev) (void)hipFree(b->step_dev);
    b->step_dev = nullptr;
    b->step_cap = 0;
    st = hip_check(hipMalloc(&b->step_dev, bytes), "batch step table hipMalloc");
    if (st == FF_OK) b->step_cap = bytes;
  }
  if (st == FF_OK && !img.empty())
    st = hip_check(hipMemcpy(b->step_dev, img.data(), img.size(), hipMemcpyHostToDevice), "batch step table upload");
  if (st != FF_OK) return st;
  unsigned char* base = static_cast<unsigned char*>(b->step_dev);
  for (const Fix& fx : fixes) {
    StepLaunch& sl = fams[fx.fam].ops[fx.op];
    if (sl.kind >= 0) {
      sl.m.desc = base + fx.desc;
      sl.m.block0 = reinterpret_cast<const int*>(base + fx.block0);
    } else {
      const PassInst* t = reinterpret_cast<const PassInst*>(base + fx.inst);
      if (sl.op.kind == 0) sl.op.col.inst = t;
      else sl.op.row.inst = t;
    }
  }
  b->fams.swap(fams);
  b->sd.assign(d, d + b->count);
  b->sp.assign(p, p + b->count);
  b->steps_ok = true;
  return FF_OK;
}

bool steps_current(const ff_batch* b, double* const* d, double* const* p) {
  return b->steps_ok && std::equal(b->sd.begin(), b->sd.end(), d) && std::equal(b->sp.begin(), b->sp.end(), p);
}

// Fork onto the batch's streams, enqueue every instance (or every group),
// join back onto s.
ff_status enqueue_all(ff_batch* b, double* const* d, double* const* p, hipStream_t s) {
  const bool grouped = (b->flags & FF_BATCH_GROUPED) != 0;
  const bool steps = (b->flags & FF_BATCH_STEPS) != 0;
  if (grouped && !groups_current(b, d, p)) {
    ff_status st = prepare_groups(b, d, p);
    if (st != FF_OK) return st;
  }
  if (steps && !steps_current(b, d, p)) {
    ff_status st = prepare_steps(b, d, p);
    if (st != FF_OK) return st;
  }
  ff_status st = hip_check(hipEventRecord(b->fork, s), "batch fork");
  for (size_t k = 0; k < b->streams.size() && st == FF_OK; ++k)
    st = hip_check(hipStreamWaitEvent(b->streams[k], b->fork, 0), "batch fork wait");
  if (steps) {
    // one stream per family: its steps in order, families side by side
    for (size_t fi = 0; fi < b->fams.size() && st == FF_OK; ++fi) {
      hipStream_t hs = b->streams[fi % b->streams.size()];
      for (const StepLaunch& op : b->fams[fi].ops) {
        if (op.kind >= 0) {
          st = hip_check(launch_multi(op.kind, op.f32, op.sel, op.epi, op.m, op.lds, hs), "multi-shape pass launch");
        } else {
          PassLaunch r = op.op;
          r.l.stream = hs;
          st = hip_check(replay_launch(r), "grouped pass launch");
        }
        if (st != FF_OK) break;
      }
    }
  } else if (grouped) {
    for (size_t gi = 0; gi < b->groups.size() && st == FF_OK; ++gi) {
      hipStream_t hs = b->streams[gi % b->streams.size()];
      for (const PassLaunch& op : b->groups[gi].ops) {
        PassLaunch r = op;
        r.l.stream = hs;
        st = hip_check(replay_launch(r), "grouped pass launch");
        if (st != FF_OK) break;
      }
    }
  }
  for (int i = 0; i < b->count && st == FF_OK; ++i)
    if ((!grouped && !steps) || solo(b, i))
      st = solve_enqueue(b->plans[i], d[i], p[i], b->streams[i % b->streams.size()], nullptr);
  for (size_t k = 0; k < b->streams.size() && st == FF_OK; ++k) {
    st = hip_check(hipEventRecord(b->joins[k], b->streams[k]), "batch join");
    if (st == FF_OK) st = hip_check(hipStreamWaitEvent(s, b->joins[k], 0), "batch join wait");
  }
  return st;
}

}  // namespace

extern "C" {

ff_status ff_batch_create(int count, const ff_batch_item* items, int flags, ff_batch** out) {
  clear_error();
  if (!out) return set_error(FF_ERR_INVALID, "ff_batch_create: null output");
  *out = nullptr;
  if (count < 1 || !items) return set_error(FF_ERR_INVALID, "ff_batch_create: need at least one item");
  if (flags & FF_BATCH_PERSISTENT)
    return set_error(FF_ERR_UNSUPPORTED,
                     "ff_batch_create: the single persistent launch is not available (its kernel did not "
                     "complete on MI355X and was removed); FF_BATCH_STEPS runs one launch per pass step");
  if (flags & ~(FF_BATCH_GRAPH | FF_BATCH_GROUPED | FF_BATCH_STEPS))
    return set_error(FF_ERR_INVALID, "ff_batch_create: bad flags");
  if ((flags & FF_BATCH_GROUPED) && (flags & FF_BATCH_STEPS))
    return set_error(FF_ERR_INVALID, "ff_batch_create: FF_BATCH_GROUPED and FF_BATCH_STEPS are exclusive");
  for (int i = 0; i < count; ++i)
    for (int a = 0; a < 3; ++a)
      if (items[i].n[a] < 2) return set_error(FF_ERR_INVALID, "ff_batch_create: extents must be >= 2");
  ff_batch* b = new (std::nothrow) ff_batch();
  if (!b) return set_error(FF_ERR_NOMEM, "host allocation failed");
  b->count = count;
  b->flags = flags;
  b->items.assign(items, items + count);
  ff_status st = make_plans(b);
  const int ns = std::min(kBatchStreams, count);
  for (int k = 0; k < ns && st == FF_OK; ++k) {
    hipStream_t s = nullptr;
    hipEvent_t e = nullptr;
    st = hip_check(hipStreamCreateWithFlags(&s, hipStreamNonBlocking), "batch stream");
    if (st == FF_OK) b->streams.push_back(s);
    if (st == FF_OK) st = hip_check(hipEventCreateWithFlags(&e, hipEventDisableTiming), "batch event");
    if (st == FF_OK) b->joins.push_back(e);
  }
  if (st == FF_OK) st = hip_check(hipEventCreateWithFlags(&b->fork, hipEventDisableTiming), "batch event");
  if (st != FF_OK) {
    ff_batch_destroy(b);
    return st;
  }
  *out = b;
  return FF_OK;
}

size_t ff_batch_bytes(const ff_batch* b, int index) {
  if (!b || index < 0 || index >= b->count) return 0;
  return b->plans[index]->bytes;
}

ff_status ff_batch_solve(ff_batch* b, double* const* density, double* const* potential, void* stream) {
  clear_error();
  if (!b || !density || !potential) return set_error(FF_ERR_INVALID, "ff_batch_solve: null argument");
  for (int i = 0; i < b->count; ++i)
    if (!density[i] || !potential[i] || density[i] == potential[i])
      return set_error(FF_ERR_INVALID, "ff_batch_solve: every instance needs two distinct buffers");
  hipStream_t s = (hipStream_t)stream;
  if (!(b->flags & FF_BATCH_GRAPH)) return enqueue_all(b, density, potential, s);
  if (!s) return set_error(FF_ERR_INVALID, "ff_batch_solve: graph capture needs a non-default stream");
  const bool same = b->bound && std::equal(b->bd.begin(), b->bd.end(), density) &&
                    std::equal(b->bp.begin(), b->bp.end(), potential);
  if (!same || !b->exec || b->graph_stream != s) {
    drop_graph(b);
    // grouped and step tables are recorded and uploaded before capture (blocking calls)
    if ((b->flags & FF_BATCH_GROUPED) && !groups_current(b, density, potential)) {
      ff_status st = prepare_groups(b, density, potential);
      if (st != FF_OK) return st;
    }
    if ((b->flags & FF_BATCH_STEPS) && !steps_current(b, density, potential)) {
      ff_status st = prepare_steps(b, density, potential);
      if (st != FF_OK) return st;
    }
    ff_status st = hip_check(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal), "begin capture");
    if (st != FF_OK) return st;
    st = enqueue_all(b, density, potential, s);
    hipGraph_t g = nullptr;
    const hipError_t e = hipStreamEndCapture(s, &g);
    if (st != FF_OK) {
      if (g) (void)hipGraphDestroy(g);
      return st;
    }
    st = hip_check(e, "end capture");
    if (st != FF_OK) return st;
    b->graph = g;
    st = hip_check(hipGraphInstantiate(&b->exec, g, nullptr, nullptr, 0), "graph instantiate");
    if (st != FF_OK) {
      drop_graph(b);
      return st;
    }
    b->graph_stream = s;
    b->bd.assign(density, density + b->count);
    b->bp.assign(potential, potential + b->count);
    b->bound = true;
  }
  return hip_check(hipGraphLaunch(b->exec, s), "graph launch");
}

ff_status ff_batch_wait(ff_batch* b, void* stream) {
  clear_error();
  if (!b) return set_error(FF_ERR_INVALID, "ff_batch_wait: null batch");
  return hip_check(hipStreamSynchronize((hipStream_t)stream), "batch synchronize");
}

ff_status ff_batch_autotune(ff_batch* b, double* const* density, double* const* potential, void* stream, int reps,
                            int* tuned, double* ms_before, double* ms_after) {
  clear_error();
  if (!b || !density || !potential || reps < 1) return set_error(FF_ERR_INVALID, "ff_batch_autotune: bad arguments");
  hipStream_t s = (hipStream_t)stream;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  ff_status st = hip_check(hipEventCreate(&e0), "event");
  if (st == FF_OK) st = hip_check(hipEventCreate(&e1), "event");
  auto time_batch = [&](double* out) {
    ff_status r = enqueue_all(b, density, potential, s);  // warm up
    if (r == FF_OK) r = hip_check(hipEventRecord(e0, s), "event");
    for (int k = 0; k < reps && r == FF_OK; ++k) r = enqueue_all(b, density, potential, s);
    if (r == FF_OK) r = hip_check(hipEventRecord(e1, s), "event");
    if (r == FF_OK) r = hip_check(hipEventSynchronize(e1), "event sync");
    float ms = 0.f;
    if (r == FF_OK) r = hip_check(hipEventElapsedTime(&ms, e0, e1), "elapsed");
    *out = ms / reps;
    return r;
  };
  double before = 0.0, after = 0.0;
  if (st == FF_OK) st = time_batch(&before);
  // every distinct (pass kind, length, precision): time each valid (E, T)
  int count = 0;
  std::map<std::tuple<int, int, int>, int> seen;  // key -> instance
  for (int i = 0; i < b->count && st == FF_OK; ++i) {
    const ff_plan* p = b->plans[i];
    const int nprec = p->opt.precision == FF_FP64 ? 1 : 2;
    for (int prec = 0; prec < nprec; ++prec) {
      const int M = p->n[2] % 2 ? p->n[2] : p->n[2] / 2;
      const int keys[4][2] = {{0, M}, {1, p->n[1]}, {2, p->n[0]}, {3, M}};
      for (auto& kv : keys) {
        // four-step lines have no single configuration to tune
        if ((kv[0] == 1 && p->fy.n) || (kv[0] == 2 && p->fx.n) || ((kv[0] == 0 || kv[0] == 3) && p->fz.n)) continue;
        if (!seen.count({kv[0], kv[1], prec})) seen[{kv[0], kv[1], prec}] = i;
      }
    }
  }
  for (auto& kv : seen) {
    if (st != FF_OK) break;
    const int kind = std::get<0>(kv.first), n = std::get<1>(kv.first), prec = std::get<2>(kv.first);
    ff_plan* p = b->plans[kv.second];
    const int eb = prec ? 8 : 16;
    double best = 1e30;
    int bE = 0, bT = 0, bF = 0;
    // column passes: split exchange, two groups, and (flags 4) the generic
    // kernel where one with the line plan compiled in exists
    const int nflags = (kind == 1 || kind == 2) ? 5 : 1;
    for (int E : kEValues) {
      for (int cand = 0; cand < 7 * nflags && st == FF_OK; ++cand) {
        const int T = 1 << (cand % 7);
        const int fl = cand / 7;
        // candidate configuration on a scratch copy of the plan's pass config
        ff_plan trial = *p;
        bool ok;
        if (kind == 0 || kind == 3) {
          RowCfg c;
          ok = row_cfg_with(p->n[2], eb, c, E, T, 0) && c.line.n == n;
          if (ok) (kind == 0 ? trial.zr : trial.zi)[prec] = c;
        } else {
          ColCfg c, c0;
          ok = col_cfg_with(n, p->nh, eb, c, E, T, fl, kind);
          if (ok && fl == FF_TUNE_GENERIC) ok = col_cfg_with(n, p->nh, eb, c0, E, T, 0, kind) && c0.stat;
          if (ok) (kind == 1 ? trial.cy : trial.cx)[prec] = c;
        }
        if (!ok) continue;
        const int ip = kv.second;
        double* src = density[ip];
        double* dst = potential[ip];
        auto launch = [&]() -> ff_status {
          if (kind == 0) return run_zfwd(&trial, prec, 0, src, dst, s);
          if (kind == 3) return run_zinv(&trial, prec, 0, EPI_STORE, src, dst, nullptr, nullptr, s);
          return run_col(&trial, prec, kind == 1 ? 1 : 0, kind == 2 ? COL_FILTER : COL_FWD, src, dst, s);
        };
        st = launch();
        if (st == FF_OK) st = hip_check(hipEventRecord(e0, s), "event");
        for (int r = 0; r < reps && st == FF_OK; ++r) st = launch();
        if (st == FF_OK) st = hip_check(hipEventRecord(e1, s), "event");
        if (st == FF_OK) st = hip_check(hipEventSynchronize(e1), "event sync");
        float ms = 0.f;
        if (st == FF_OK) st = hip_check(hipEventElapsedTime(&ms, e0, e1), "elapsed");
        if (st == FF_OK && ms < best) {
          best = ms;
          bE = E;
          bT = T;
          bF = fl;
        }
      }
    }
    if (st == FF_OK && bE) {
      const ff_tune t{kind, n, prec, bE, bT, bF};
      st = tune_set(&t);
      count++;
    }
  }
  if (st == FF_OK) {
    drop_graph(b);
    b->bound = false;
    st = make_plans(b);
  }
  if (st == FF_OK) st = time_batch(&after);
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  if (st != FF_OK) return st;
  if (tuned) *tuned = count;
  if (ms_before) *ms_before = before;
  if (ms_after) *ms_after = after;
  return FF_OK;
}

ff_status ff_batch_describe(const ff_batch* b, char* buf, size_t len) {
  clear_error();
  if (!b || !buf || !len) return set_error(FF_ERR_INVALID, "ff_batch_describe: bad arguments");
  std::ostringstream os;
  os << "{\"count\":" << b->count << ",\"flags\":" << b->flags << ",\"streams\":" << b->streams.size()
     << ",\"graph\":" << (b->exec ? 1 : 0) << ",\"groups\":[";
  for (size_t g = 0; g < b->groups.size(); ++g) {
    os << (g ? "," : "") << "{\"instances\":" << b->groups[g].inst.size() << ",\"launches\":" << b->groups[g].ops.size()
       << "}";
  }
  os << "],\"steps\":[";
  for (size_t f = 0; f < b->fams.size(); ++f) {
    const StepFamily& fm = b->fams[f];
    os << (f ? "," : "") << "{\"instances\":" << fm.inst.size() << ",\"steps\":" << fm.steps
       << ",\"launches\":" << fm.ops.size() << ",\"multi_launches\":" << fm.multi_launches << "}";
  }
  os << "],\"plans\":[";
  char tmp[4096];
  for (int i = 0; i < b->count; ++i) {
    if (ff_plan_describe(b->plans[i], tmp, sizeof tmp) != FF_OK) return FF_ERR_INVALID;
    os << (i ? "," : "") << tmp;
  }
  os << "]}";
  const std::string s = os.str();
  if (s.size() + 1 > len) return set_error(FF_ERR_INVALID, "ff_batch_describe: buffer too small");
  memcpy(buf, s.c_str(), s.size() + 1);
  return FF_OK;
}

ff_status ff_batch_destroy(ff_batch* b) {
  if (!b) return FF_OK;
  drop_graph(b);
  for (hipStream_t s : b->streams) (void)hipStreamSynchronize(s);
  destroy_plans(b);
  if (b->inst_dev) (void)hipFree(b->inst_dev);
  if (b->step_dev) (void)hipFree(b->step_dev);
  for (hipStream_t s : b->streams) (void)hipStreamDestroy(s);
  for (hipEvent_t e : b->joins) (void)hipEventDestroy(e);
  if (b->fork) (void)hipEventDestroy(b->fork);
  delete b;
  return FF_OK;
}

}  // extern "C"
