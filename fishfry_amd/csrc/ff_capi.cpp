// extern "C" entry points of libfishfry.so (declared in include/fishfry.h).
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>

#include "ff_solver.hpp"

namespace ff {
const char* last_error_text();
}

using namespace ff;

static double now_seconds() {
  using clk = std::chrono::steady_clock;
  return std::chrono::duration<double>(clk::now().time_since_epoch()).count();
}

extern "C" {

const char* ff_last_error(void) { return last_error_text(); }

size_t ff_reference_bytes(const int n[3], const int m[3]) {
  if (!n || !m) return 0;
  for (int a = 0; a < 3; ++a)
    if (n[a] < 1 || m[a] < 1) return 0;
  return reference_bytes(n, m);
}

ff_status ff_plan_create(const int n[3], const double lo[3], const double hi[3], const int m[3], const int id[3],
                         const ff_options* opts, ff_plan** out) {
  clear_error();
  if (!out) return set_error(FF_ERR_INVALID, "ff_plan_create: null output handle");
  *out = nullptr;
  if (!n || !lo || !hi) return set_error(FF_ERR_INVALID, "ff_plan_create: null n/lo/hi");
  if (m && (m[0] != 1 || m[1] != 1 || m[2] != 1))
    return set_error(FF_ERR_UNSUPPORTED, "ff_plan_create: only 1x1x1 tasks (one GPU) are supported");
  if (id && (id[0] != 0 || id[1] != 0 || id[2] != 0))
    return set_error(FF_ERR_INVALID, "ff_plan_create: id must be {0,0,0} at 1x1x1 tasks");
  return plan_create(n, lo, hi, opts, out);
}

size_t ff_plan_bytes(const ff_plan* plan) { return plan ? plan->bytes : 0; }

ff_status ff_plan_destroy(ff_plan* plan) {
  plan_destroy(plan);
  return FF_OK;
}

static ff_status check_solve(const ff_plan* plan, size_t bytes, const double* density, const double* potential) {
  if (!plan) return set_error(FF_ERR_INVALID, "ff_solve: null plan");
  if (!density || !potential) return set_error(FF_ERR_INVALID, "ff_solve: null density or potential");
  if (density == potential) return set_error(FF_ERR_INVALID, "ff_solve: density and potential must differ");
  if (bytes < plan->bytes) {
    char buf[160];
    snprintf(buf, sizeof buf, "ff_solve: bytes %zu < required %zu (ParisPencil::bytes)", bytes, plan->bytes);
    return set_error(FF_ERR_BYTES, buf);
  }
  return FF_OK;
}

ff_status ff_solve_async(const ff_plan* plan, size_t bytes, double* density, double* potential, void* stream) {
  clear_error();
  ff_status st = check_solve(plan, bytes, density, potential);
  if (st != FF_OK) return st;
  return solve_enqueue(plan, density, potential, (hipStream_t)stream, nullptr);
}

ff_status ff_solve_noalloc(const ff_plan* plan, size_t bytes, double* density, double* potential, void* stream) {
  clear_error();
  ff_status st = check_solve(plan, bytes, density, potential);
  if (st != FF_OK) return st;
  hipStream_t s = (hipStream_t)stream;
  size_t need = 0;
  if (plan->opt.precision == FF_FP32_REFINE) need = plan_scratch_bytes(plan);
  else need = plan_work_bytes(plan);
  if (need > kStaticWorkBytes) {
    char buf[200];
    snprintf(buf, sizeof buf,
             "ff_solve_noalloc: this plan needs a %zu-byte work area; the static work area holds %zu bytes", need,
             kStaticWorkBytes);
    return set_error(FF_ERR_INVALID, buf);
  }
  if (need == 0) return solve_enqueue(plan, density, potential, s, nullptr);  // no work area used
  int cur = -1;
  if ((st = hip_check(hipGetDevice(&cur), "hipGetDevice"))) return st;
  if (cur != plan->device)
    return set_error(FF_ERR_INVALID, "ff_solve_noalloc: the plan's device must be the current device");
  void* w = nullptr;
  if ((st = static_work_area(plan->device, &w))) return st;
  // a copy of the plan (plain data, on the stack) pointed at the static area
  ff_plan local = *plan;
  if (local.opt.precision == FF_FP32_REFINE) plan_bind_scratch(&local, w);
  else local.work = w;
  return solve_enqueue(&local, density, potential, s, nullptr);
}

ff_status ff_solve(const ff_plan* plan, size_t bytes, double* density, double* potential, void* stream,
                   ff_stamp* stamps, int* n_stamps) {
  clear_error();
  ff_status st = check_solve(plan, bytes, density, potential);
  if (st != FF_OK) return st;
  if (stamps && (!n_stamps || *n_stamps < 0)) return set_error(FF_ERR_INVALID, "ff_solve: n_stamps required");
  hipStream_t s = (hipStream_t)stream;
  const double t0 = now_seconds();
  const bool timed = stamps && plan->opt.precision == FF_FP64;
  hipEvent_t ev[6] = {};
  if (timed) {
    for (int e = 0; e < 6; ++e)
      if ((st = hip_check(hipEventCreate(&ev[e]), "hipEventCreate"))) return st;
    st = hip_check(hipEventRecord(ev[5], s), "event");
  }
  if (st == FF_OK) st = solve_enqueue(plan, density, potential, s, timed ? ev : nullptr);
  if (st == FF_OK) st = hip_check(hipStreamSynchronize(s), "solve synchronize");
  const double t1 = now_seconds();
  if (st == FF_OK && stamps) {
    // the reference's 14 stamps (ParisPencil.cpp:98,142,144,188,190,235,237,
    // 311,313,357,359,402,404,424); its reorders and self all-to-alls are
    // fused away here, so they take zero time at the end of their pass
    static const char* names[FF_SOLVE_STAMPS] = {"Total",
                                                 "Reorder 3D block into sub-pencils",
                                                 "Redistribute into Z pencils",
                                                 "Real-to-complex FFT in Z",
                                                 "Redistribute for Y pencils",
                                                 "Forward FFT in Y",
                                                 "Redistribute for X pencils",
                                                 "X FFTs and filter",
                                                 "Redistribute for Y pencils",
                                                 "Backward FFT in Y",
                                                 "Redistribute in Z pencils",
                                                 "Complex-to-real FFT in Z",
                                                 "Redistribute for 3D blocks",
                                                 "Rearrange into 3D blocks and normalize"};
    // pass boundary (0..5) reached by each stamp
    static const int at[FF_SOLVE_STAMPS] = {0, 0, 0, 1, 1, 2, 2, 3, 3, 4, 4, 5, 5, 5};
    double pass_t[6] = {t0, t0, t0, t0, t0, t1};
    if (timed) {
      for (int e = 0; e < 5; ++e) {
        float ms = 0.f;
        if (hipEventElapsedTime(&ms, ev[5], ev[e]) == hipSuccess) pass_t[e + 1] = t0 + 1e-3 * ms;
      }
      pass_t[5] = t1;
    }
    const int cnt = *n_stamps < FF_SOLVE_STAMPS ? *n_stamps : FF_SOLVE_STAMPS;
    for (int i = 0; i < cnt; ++i) {
      snprintf(stamps[i].name, sizeof stamps[i].name, "%s", names[i]);
      stamps[i].seconds = i == FF_SOLVE_STAMPS - 1 ? t1 : pass_t[at[i]];
    }
    *n_stamps = cnt;
  } else if (n_stamps) {
    *n_stamps = 0;
  }
  if (timed)
    for (int e = 0; e < 6; ++e)
      if (ev[e]) (void)hipEventDestroy(ev[e]);
  return st;
}

ff_status ff_plan_describe(const ff_plan* p, char* buf, size_t len) {
  clear_error();
  if (!p || !buf || !len) return set_error(FF_ERR_INVALID, "ff_plan_describe: bad arguments");
  std::string s = "{";
  char tmp[512];
  snprintf(tmp, sizeof tmp, "\"n\":[%d,%d,%d],\"nh\":%d,\"bytes\":%zu,\"precision\":%d,\"refine_steps\":%d,", p->n[0],
           p->n[1], p->n[2], p->nh, p->bytes, p->opt.precision, p->opt.refine_steps);
  s += tmp;
  snprintf(tmp, sizeof tmp,
           "\"dd\":[%.17g,%.17g,%.17g],\"kt\":%d,\"slab\":%d,\"symbol\":%d,\"pencil\":%d,\"four_step\":"
           "{\"y\":[%d,%d],\"x\":[%d,%d],\"z\":[%d,%d]},",
           p->ddi, p->ddj, p->ddk, p->lgKT ? 1 << p->lgKT : 0, p->slab, p->sym, p->pencil, p->fy.n1, p->fy.n2,
           p->fx.n1, p->fx.n2, p->fz.n1, p->fz.n2);
  s += tmp;
  auto line = [&](const char* key, const LineCfg& c, int T, size_t lds, int threads) {
    std::string r = "\"";
    r += key;
    r += "\":{";
    snprintf(tmp, sizeof tmp, "\"n\":%d,\"E\":%d,\"L\":%d,\"T\":%d,\"threads\":%d,\"lds\":%zu,\"radices\":[", c.n, c.E,
             c.lp.L, T, threads, lds);
    r += tmp;
    for (int i = 0; i < c.lp.np; ++i) {
      snprintf(tmp, sizeof tmp, "%s%d", i ? "," : "", c.lp.radix[i]);
      r += tmp;
    }
    r += "]}";
    return r;
  };
  // four-step axes: the n1-point and n2-point sub-line configurations
  auto sub = [&](const char* key, const ff_plan::FourStep& fs) {
    std::string k1 = std::string(key) + "1", k2 = std::string(key) + "2";
    return line(k1.c_str(), fs.c1[0].line, fs.c1[0].T, fs.c1[0].lds, fs.c1[0].threads) + "," +
           line(k2.c_str(), fs.c2[0].line, fs.c2[0].T, fs.c2[0].lds, fs.c2[0].threads);
  };
  if (p->fz.n) {
    s += sub("z", p->fz) + ",";
  } else {
    s += line("z", p->zr[0].line, p->zr[0].T, p->zr[0].lds, p->zr[0].threads) + ",";
    s += line("zi", p->zi[0].line, p->zi[0].T, p->zi[0].lds, p->zi[0].threads) + ",";
  }
  auto axis = [&](const char* key, const ff_plan::FourStep& fs, const ColCfg& c) {
    if (!fs.n) return line(key, c.line, c.T, c.lds, c.threads);
    return sub(key, fs);
  };
  s += axis("y", p->fy, p->cy[0]) + ",";
  s += axis("x", p->fx, p->cx[0]);
  s += "}";
  if (s.size() + 1 > len) return set_error(FF_ERR_INVALID, "ff_plan_describe: buffer too small");
  memcpy(buf, s.c_str(), s.size() + 1);
  return FF_OK;
}

ff_status ff_solve_pass(const ff_plan* p, int pass, const double* src, double* dst, void* stream) {
  clear_error();
  if (!p || !src || !dst) return set_error(FF_ERR_INVALID, "ff_solve_pass: null argument");
  if (p->opt.precision != FF_FP64) return set_error(FF_ERR_INVALID, "ff_solve_pass: FP64 plans only");
  hipStream_t s = (hipStream_t)stream;
  int li = LAY_PLAIN, lo = LAY_PLAIN;
  if (pass >= 0 && pass <= 4) pass_layouts(p, pass, li, lo);
  switch (pass) {
    case 0: return run_zfwd(p, 0, 0, src, dst, s);
    case 1: return run_col(p, 0, 1, COL_FWD, src, dst, s, nullptr, li, lo);
    case 2: return run_col(p, 0, 0, COL_FILTER, src, dst, s, nullptr, li, lo);
    case 3: return run_col(p, 0, 1, COL_INV, src, dst, s, nullptr, li, lo);
    case 4: return run_zinv(p, 0, 0, EPI_STORE, src, dst, nullptr, nullptr, s, li);
    default: return set_error(FF_ERR_INVALID, "ff_solve_pass: pass must be 0..4");
  }
}

ff_status ff_tune_set(const ff_tune* t) {
  clear_error();
  return tune_set(t);
}

void ff_tune_reset(void) { tune_reset(); }

ff_status ff_tune_query(int kind, int n, int precision, int ncols, ff_tune* out) {
  clear_error();
  if (!out || kind < 0 || kind > 3 || n < 1 || (precision != 0 && precision != 1) || ncols < 1)
    return set_error(FF_ERR_INVALID, "ff_tune_query: bad arguments");
  const int eb = precision ? 8 : 16;
  int E = 0, T = 0, flags = 0;
  if (kind == 0 || kind == 3) {
    RowCfg c;
    if (!make_row_cfg(2 * n, eb, c, kind, precision))
      return set_error(FF_ERR_UNSUPPORTED, "ff_tune_query: no plan for that length");
    E = c.line.E;
    T = c.T;
    flags = (c.stat ? 0 : FF_TUNE_GENERIC) | (c.stat == 2 ? FF_TUNE_LDS_TWIDDLES : 0);
  } else {
    ColCfg c;
    if (!make_col_cfg(n, ncols, eb, c, kind, precision))
      return set_error(FF_ERR_UNSUPPORTED, "ff_tune_query: no plan for that length");
    E = c.line.E;
    T = c.T;
    flags = (c.split ? FF_TUNE_SPLIT_LDS : 0) | (c.groups == 2 ? FF_TUNE_TWO_GROUPS : 0) |
            (c.stat ? 0 : FF_TUNE_GENERIC) | (c.stat == 2 ? FF_TUNE_LDS_TWIDDLES : 0);
  }
  *out = ff_tune{kind, n, precision, E, T, flags};
  return FF_OK;
}

void ff_field_default(const int n[3], ff_field* out) {
  if (!out) return;
  memset(out, 0, sizeof *out);
  for (int a = 0; a < 3; ++a) {
    out->n[a] = n ? n[a] : 0;
    out->d[a] = n && n[a] ? 2.0 * M_PI / n[a] : 0.0;  // FishFry.hpp:109-111
    out->lo[a] = -M_PI;                               // FishFry.hpp:133,137-139 at id = 0
    const double hi = M_PI - out->d[a];               // FishFry.hpp:134
    out->dd[a] = n && n[a] ? 2.0 * M_PI * double(n[a] - 1) / (double(n[a]) * (hi - out->lo[a])) : 0.0;
  }
  out->scale = 1.0;
}

ff_status ff_field_init(const ff_field* field, double* f, void* stream) {
  clear_error();
  return field_init(field, f, (hipStream_t)stream);
}

ff_status ff_field_diff(const ff_field* field, const double* phi, double out[3], void* stream) {
  clear_error();
  std::lock_guard<std::mutex> g(scratch_mutex());
  return field_diff(field, phi, out, (hipStream_t)stream);
}

}  // extern "C"
