// Host planner: radix plans, tile shapes, twiddles, kernel dispatch and the
// library's error / scratch bookkeeping.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <complex>
#include <map>
#include <mutex>
#include <tuple>

#include "ff_internal.hpp"

namespace ff {

// ---- errors ---------------------------------------------------------------

static thread_local std::string g_last_error;

ff_status set_error(ff_status st, const std::string& msg) {
  g_last_error = msg;
  return st;
}

void clear_error() { g_last_error.clear(); }

ff_status hip_check(hipError_t e, const char* what) {
  if (e == hipSuccess) return FF_OK;
  char buf[256];
  snprintf(buf, sizeof buf, "%s: %s (%s)", what, hipGetErrorName(e), hipGetErrorString(e));
  return set_error(e == hipErrorOutOfMemory ? FF_ERR_NOMEM : FF_ERR_HIP, buf);
}

const char* last_error_text() { return g_last_error.c_str(); }

// ---- scratch ----------------------------------------------------------------

std::mutex& scratch_mutex() {
  static std::mutex m;
  return m;
}

ff_status device_scratch(size_t ndoubles, double** out) {
  // per device, grown on demand; callers serialize through scratch_mutex()
  static std::map<int, std::pair<double*, size_t>> blocks;
  int dev = 0;
  ff_status st = hip_check(hipGetDevice(&dev), "hipGetDevice");
  if (st != FF_OK) return st;
  auto& b = blocks[dev];
  if (b.second < ndoubles) {
    if (b.first) {
      st = hip_check(hipDeviceSynchronize(), "scratch sync");
      if (st != FF_OK) return st;
      (void)hipFree(b.first);
      b.first = nullptr;
      b.second = 0;
    }
    st = hip_check(hipMalloc(&b.first, ndoubles * sizeof(double)), "scratch hipMalloc");
    if (st != FF_OK) return st;
    b.second = ndoubles;
  }
  *out = b.first;
  return FF_OK;
}

// ---- twiddles ---------------------------------------------------------------

void twiddle_ld(long long t, long long n, long double& c, long double& s) {
  // exp(-2 pi i t/n).  4t = q n + r (0 <= r < n): the angle is q quarter turns
  // plus (pi/2)(r/n); for r/n > 1/2 the complementary angle is used so the
  // long-double argument stays below pi/4 and mirrored points agree exactly.
  t %= n;
  if (t < 0) t += n;
  const long long v = 4 * t;
  const long long q = v / n;
  const long long r = v - q * n;
  const long double hp = 1.5707963267948966192313216916397514L;  // pi/2
  long double cb, sb;
  if (2 * r <= n) {
    const long double a = hp * ((long double)r / (long double)n);
    cb = cosl(a);
    sb = sinl(a);
  } else {
    const long double a = hp * ((long double)(n - r) / (long double)n);
    cb = sinl(a);
    sb = cosl(a);
  }
  long double cc, ss;
  switch (q & 3) {
    case 0: cc = cb; ss = sb; break;
    case 1: cc = -sb; ss = cb; break;
    case 2: cc = -cb; ss = -sb; break;
    default: cc = sb; ss = -cb; break;
  }
  c = cc;
  s = -ss;
}

// ---- Bluestein tables ---------------------------------------------------------

// In-place long-double DFT of a 2^a 3^b 5^c length by recursive
// decimation in time (exact twiddles from twiddle_ld).
static void fft_ld(std::vector<std::complex<long double>>& a, int sign) {
  const long long n = (long long)a.size();
  if (n == 1) return;
  const int p = n % 2 == 0 ? 2 : (n % 3 == 0 ? 3 : 5);
  const long long q = n / p;
  std::vector<std::vector<std::complex<long double>>> sub(p, std::vector<std::complex<long double>>(q));
  for (long long j = 0; j < q; ++j)
    for (int r = 0; r < p; ++r) sub[r][j] = a[j * p + r];
  for (int r = 0; r < p; ++r) fft_ld(sub[r], sign);
  for (long long k = 0; k < n; ++k) {
    std::complex<long double> acc = 0;
    for (int r = 0; r < p; ++r) {
      long double c, s;
      twiddle_ld((long long)r * k, n, c, s);  // exp(-2 pi i r k/n)
      acc += sub[r][k % q] * std::complex<long double>(c, sign < 0 ? s : -s);
    }
    a[k] = acc;
  }
}

static size_t push_cx(std::vector<unsigned char>& buf, const std::vector<std::complex<long double>>& v, int prec) {
  const size_t elem = prec ? 8 : 16;
  const size_t off = (buf.size() + 255) / 256 * 256;
  buf.resize(off + v.size() * elem);
  for (size_t t = 0; t < v.size(); ++t) {
    if (prec) {
      const float w[2] = {(float)v[t].real(), (float)v[t].imag()};
      memcpy(&buf[off + t * elem], w, sizeof w);
    } else {
      const double w[2] = {(double)v[t].real(), (double)v[t].imag()};
      memcpy(&buf[off + t * elem], w, sizeof w);
    }
  }
  return off;
}

void push_bluestein(std::vector<unsigned char>& buf, int n, int m, int prec, size_t& chirp, size_t& bsf,
                    size_t& bsi) {
  // w[t] = exp(-pi i t^2/n) = W_{2n}^(t^2 mod 2n), with the exact reduction
  std::vector<std::complex<long double>> w(n);
  for (long long t = 0; t < n; ++t) {
    long double c, s;
    twiddle_ld((t * t) % (2LL * n), 2LL * n, c, s);
    w[t] = {c, s};
  }
  chirp = push_cx(buf, w, prec);
  for (int dir = 0; dir < 2; ++dir) {
    std::vector<std::complex<long double>> b(m, 0);
    for (int t = 0; t < n; ++t) {
      const std::complex<long double> v = dir == 0 ? std::conj(w[t]) : w[t];
      b[t] = v;
      if (t) b[m - t] = v;
    }
    fft_ld(b, -1);
    for (auto& x : b) x /= (long double)m;
    (dir == 0 ? bsf : bsi) = push_cx(buf, b, prec);
  }
}

// ---- resident table cache ---------------------------------------------------

namespace {
struct TableKey {
  int dev, kind, n, m, prec;
  bool operator<(const TableKey& o) const {
    return std::tie(dev, kind, n, m, prec) < std::tie(o.dev, o.kind, o.n, o.m, o.prec);
  }
};
struct TableEntry {
  void* mem = nullptr;
  size_t off[3] = {0, 0, 0};
};
}  // namespace

// Builds a table on the host, uploads it once per device and keeps it until
// the process ends; kind 0: W_n[t], t < n; 1: exp(-2 pi i k/n), k < n/2 (at
// least one entry); 2: the Bluestein tables of length n over m; 3: the
// finite-difference factors of order m, k < n (FP64).
static const TableEntry* table_entry(int kind, int n, int m, int prec) {
  static std::mutex mu;
  static std::map<TableKey, TableEntry> cache;
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) return nullptr;
  std::lock_guard<std::mutex> g(mu);
  const TableKey key{dev, kind, n, m, prec};
  auto it = cache.find(key);
  if (it != cache.end()) return &it->second;
  std::vector<unsigned char> buf;
  TableEntry e;
  if (kind == 2) {
    push_bluestein(buf, n, m, prec, e.off[0], e.off[1], e.off[2]);
  } else if (kind == 3) {  // finite-difference factors of order m
    buf.resize((size_t)n * sizeof(double));
    for (int t = 0; t < n; ++t) {
      const double v = (double)fd_factor_ld(t, n, m);
      memcpy(&buf[(size_t)t * sizeof(double)], &v, sizeof v);
    }
  } else {
    const long long count = kind == 0 ? n : std::max(1, n / 2);
    const size_t elem = prec ? 8 : 16;
    buf.resize((size_t)count * elem);
    for (long long t = 0; t < count; ++t) {
      long double c, s;
      twiddle_ld(t, n, c, s);
      if (prec) {
        const float v[2] = {(float)c, (float)s};
        memcpy(&buf[t * elem], v, sizeof v);
      } else {
        const double v[2] = {(double)c, (double)s};
        memcpy(&buf[t * elem], v, sizeof v);
      }
    }
  }
  if (hipMalloc(&e.mem, buf.size()) != hipSuccess) return nullptr;
  if (hipMemcpy(e.mem, buf.data(), buf.size(), hipMemcpyHostToDevice) != hipSuccess) {
    (void)hipFree(e.mem);
    return nullptr;
  }
  return &(cache[key] = e);
}

// 4 s^2 (order 3) or 4 s^2 (1 + s^2/3) (order 5), s = sin(pi k/n): the
// eigenvalue magnitude (times h^2) of the 3- / 5-point second difference on
// exp(2 pi i k j/n).  sinl of pi min(k, n - k)/n keeps small angles accurate.
long double fd_factor_ld(long long k, long long n, int order) {
  long long kk = k % n;
  if (kk < 0) kk += n;
  if (n - kk < kk) kk = n - kk;
  const long double pi = 3.14159265358979323846264338327950288L;
  const long double sv = sinl(pi * (long double)kk / (long double)n);
  const long double s2 = sv * sv;
  return order == 5 ? 4.0L * s2 * (1.0L + s2 / 3.0L) : 4.0L * s2;
}

const double* table_symbol(int n, int order) {
  const TableEntry* e = table_entry(3, n, order, 0);
  return e ? static_cast<const double*>(e->mem) : nullptr;
}

const void* table_twiddle(int n, int prec) {
  const TableEntry* e = table_entry(0, n, 0, prec);
  return e ? e->mem : nullptr;
}

const void* table_post(int nz, int prec) {
  const TableEntry* e = table_entry(1, nz, 0, prec);
  return e ? e->mem : nullptr;
}

bool table_bluestein(int n, int m, int prec, const void* out[3]) {
  const TableEntry* e = table_entry(2, n, m, prec);
  if (!e) return false;
  for (int k = 0; k < 3; ++k) out[k] = static_cast<const unsigned char*>(e->mem) + e->off[k];
  return true;
}

// ---- radix plans ------------------------------------------------------------

static const int kRadices[] = {16, 12, 10, 8, 6, 5, 4, 3, 2};

// Minimum number of passes to factor m with radices dividing E; fills seq.
static int min_passes(int m, int E, std::map<int, std::pair<int, int>>& memo) {
  if (m == 1) return 0;
  auto it = memo.find(m);
  if (it != memo.end()) return it->second.first;
  int best = 1 << 20, bestR = 0;
  for (int R : kRadices) {
    if (R > E || E % R || m % R) continue;
    const int p = min_passes(m / R, E, memo);
    if (p + 1 < best) {
      best = p + 1;
      bestR = R;
    }
  }
  memo[m] = {best, bestR};
  return best;
}

static bool radix_seq(int n, int E, std::vector<int>& seq) {
  std::map<int, std::pair<int, int>> memo;
  const int p = min_passes(n, E, memo);
  if (p >= (1 << 20) || p > FF_MAXP) return false;
  seq.clear();
  int m = n;
  while (m > 1) {
    const int R = memo[m].second;
    seq.push_back(R);
    m /= R;
  }
  // largest radix first: the first pass needs no twiddles
  std::sort(seq.begin(), seq.end(), [](int a, int b) { return a > b; });
  return true;
}

static void fill_plan(int n, int E, const std::vector<int>& seq, LineCfg& out) {
  out.n = n;
  out.E = E;
  LinePlan& lp = out.lp;
  memset(&lp, 0, sizeof lp);
  lp.n = n;
  lp.L = n / E;
  lp.np = (int)seq.size();
  int ns = 1;
  for (int p = 0; p < lp.np; ++p) {
    lp.radix[p] = seq[p];
    lp.ns[p] = ns;
    lp.nsinv[p] = ns > 1 ? (uint32_t)(((1ull << 32) + (uint64_t)ns - 1) / (uint64_t)ns) : 0u;
    lp.twstep[p] = n / (ns * seq[p]);
    ns *= seq[p];
  }
}

static bool plan_smooth(int n, LineCfg& out, int forceE);

static bool smooth235(long long m) {
  for (int p : {2, 3, 5})
    while (m % p == 0) m /= p;
  return m == 1;
}

bool plan_line(int n, LineCfg& out, int forceE) {
  if (n < 1) return false;
  if (plan_smooth(n, out, forceE)) return true;
  if (smooth235(n)) return false;  // a 2^a 3^b 5^c length too long for one workgroup
  // Bluestein: of the 2^a 3^b 5^c engine lengths m >= 2n - 1 that have a plan,
  // the one with the least work, m * (passes + 0.04 E).  A slightly longer m
  // with fewer passes beats the shortest one (n = 458: m = 960 takes six
  // passes, m = 1000 three).
  double best = 1e300;
  bool found = false;
  LineCfg bestc;
  for (long long m = 2LL * n - 1; m <= 4LL * n; ++m) {
    if (!smooth235(m)) continue;
    LineCfg c;
    if (!plan_smooth((int)m, c, forceE)) continue;
    const double cost = double(m) * (double(c.lp.np) + 0.04 * c.E);
    if (cost < best) {
      best = cost;
      bestc = c;
      found = true;
    }
  }
  if (!found) return false;
  out = bestc;
  out.n = n;
  return true;
}

static bool plan_smooth(int n, LineCfg& out, int forceE) {
  if (n == 1) {
    fill_plan(1, 1, {}, out);
    return true;
  }
  double bestCost = 1e30;
  int bestE = 0;
  std::vector<int> bestSeq;
  for (int E : kEValues) {
    if (E == 1) continue;
    if (forceE && E != forceE) continue;
    if (n % E || n / E > max_threads_for(E)) continue;
    std::vector<int> seq;
    if (!radix_seq(n, E, seq)) continue;
    const double cost = (double)seq.size() + 0.04 * E;
    if (cost < bestCost) {
      bestCost = cost;
      bestE = E;
      bestSeq = seq;
    }
  }
  if (!bestE) return false;
  fill_plan(n, bestE, bestSeq, out);
  return true;
}

static int floor_pow2(int v) {
  int p = 1;
  while (p * 2 <= v) p *= 2;
  return p;
}

// ---- tuning overrides -------------------------------------------------------

static std::mutex& tune_mutex() {
  static std::mutex m;
  return m;
}
// (kind, n, precision) -> (E, T, flags)
static std::map<std::tuple<int, int, int>, std::tuple<int, int, int>>& tune_table() {
  static std::map<std::tuple<int, int, int>, std::tuple<int, int, int>> t;
  return t;
}

bool tune_lookup(int kind, int n, int prec, int& E, int& T, int& flags) {
  std::lock_guard<std::mutex> g(tune_mutex());
  auto it = tune_table().find({kind, n, prec});
  if (it == tune_table().end()) return false;
  std::tie(E, T, flags) = it->second;
  return true;
}

// Whether the compile-time plan of a kStaticCols entry for (n, E) is the plan
// the planner built at run time (the compiled-in kernels must do the same
// arithmetic as the generic ones).
template <int K>
static bool static_plan_matches(int n, int E, const LinePlan& lp) {
  if constexpr (K < kNumStaticCols) {
    constexpr StaticCol sc = kStaticCols[K];
    if (sc.n == n && sc.E == E) {
      constexpr LinePlan s = static_plan<sc.n, sc.E>();
      if (s.n != lp.n || s.L != lp.L || s.np != lp.np) return false;
      for (int p = 0; p < s.np; ++p)
        if (s.radix[p] != lp.radix[p] || s.ns[p] != lp.ns[p] || s.nsinv[p] != lp.nsinv[p] ||
            s.twstep[p] != lp.twstep[p])
          return false;
      return true;
    }
    return static_plan_matches<K + 1>(n, E, lp);
  } else {
    return false;
  }
}

bool col_cfg_with(int n, int ncols, int elem_bytes, ColCfg& out, int E, int T, int flags, int kind) {
  if (!plan_line(n, out.line, E)) return false;
  const int L = out.line.lp.L;
  if (T <= 0) {
    // default: about 256 threads per workgroup, widened until a tile row is
    // at least 64 contiguous bytes (FP32 columns) while the workgroup fits;
    // no wider than the columns
    T = floor_pow2(std::max(1, 256 / L));
    while (T * elem_bytes < 64 && 2 * T * L <= max_threads_for(out.line.E)) T *= 2;
    int cap = 1;
    while (cap < ncols) cap *= 2;
    if (T > cap) T = cap;
  }
  if (T < 1 || (T & (T - 1)) || T * L > max_threads_for(out.line.E)) return false;
  out.T = T;
  out.lgT = 0;
  while ((1 << out.lgT) < T) out.lgT++;
  out.ps = kColPadShift;
  out.threads = T * L;
  out.split = (flags & FF_TUNE_SPLIT_LDS) ? 1 : 0;
  out.groups = (flags & FF_TUNE_TWO_GROUPS) ? 2 : 1;
  if (out.groups == 2 && out.line.E != 8 && out.line.E != 16) return false;
  if (out.line.bluestein() && (out.split || out.groups != 1)) return false;  // Bluestein: default shape only
  out.stat = 0;
  const bool twl = (flags & FF_TUNE_LDS_TWIDDLES) != 0;
  if (!(flags & FF_TUNE_GENERIC) && out.groups == 1 && !out.line.bluestein() && out.ps == kColPadShift)
    for (const StaticCol& sc : kStaticCols)
      if (sc.n == n && sc.E == out.line.E && sc.T == T && sc.f32 == (elem_bytes == 8 ? 1 : 0) &&
          sc.split == out.split && (sc.kinds & (1 << kind)) && (!twl || sc.twl_waves > 0))
        out.stat = static_plan_matches<0>(n, out.line.E, out.line.lp) ? (twl ? 2 : 1) : 0;
  if (twl && out.stat != 2) return false;  // only kernels with a compiled-in plan
  const size_t ne = (size_t)out.line.lp.n;  // engine length (Bluestein: m >= 2n - 1)
  const size_t elems = ne * T + ((ne - 1) >> out.ps) + 1;
  out.lds = elems * (size_t)(out.split ? elem_bytes / 2 : elem_bytes);
  // LDS twiddles: the first half of the table after the exchange (16-byte
  // aligned)
  if (out.stat == 2) out.lds = (out.lds + 15) / 16 * 16 + ne / 2 * (size_t)elem_bytes;
  return out.lds <= 160 * 1024;
}

template <int K>
static bool static_row_plan_matches(int M, int E, const LinePlan& lp) {
  if constexpr (K < kNumStaticRows) {
    constexpr StaticRow sr = kStaticRows[K];
    if (sr.M == M && sr.E == E) {
      constexpr LinePlan s = static_plan<sr.M, sr.E>();
      if (s.n != lp.n || s.L != lp.L || s.np != lp.np) return false;
      for (int p = 0; p < s.np; ++p)
        if (s.radix[p] != lp.radix[p] || s.ns[p] != lp.ns[p] || s.nsinv[p] != lp.nsinv[p] ||
            s.twstep[p] != lp.twstep[p])
          return false;
      return true;
    }
    return static_row_plan_matches<K + 1>(M, E, lp);
  } else {
    return false;
  }
}

bool row_cfg_with(int nz, int elem_bytes, RowCfg& out, int E, int T, int flags) {
  const int M = (nz % 2 == 0) ? nz / 2 : nz;
  if (!plan_line(M, out.line, E)) return false;
  const int L = out.line.lp.L;
  // default: about 128 threads per workgroup (measured fastest for both z
  // passes at 1024^3 on MI355X: more workgroups per CU beat wider ones)
  if (T <= 0) T = std::max(1, 128 / L);
  if (T < 1 || T * L > max_threads_for(out.line.E)) return false;
  out.T = T;
  out.ps = 4;
  // a line holds the staged half spectrum (M + 1 entries) and the engine
  // transform (length M, or m >= 2M - 1 for Bluestein)
  if (out.line.bluestein()) {
    const int span = out.line.lp.n;
    out.pitch = span + (span >> out.ps) + 1;
  } else {
    out.pitch = (M + 1) + (M >> out.ps) + 1;
  }
  out.threads = T * L;
  out.Linv = L > 1 ? (uint32_t)(((1ull << 32) + (uint64_t)L - 1) / (uint64_t)L) : 0u;
  out.lds = (size_t)T * out.pitch * elem_bytes;
  out.stat = 0;
  if (!(flags & FF_TUNE_GENERIC) && nz % 2 == 0 && !out.line.bluestein() && out.ps == 4)
    out.stat = static_row_plan_matches<0>(M, out.line.E, out.line.lp) ? 1 : 0;
  if (flags & FF_TUNE_LDS_TWIDDLES) {
    // the first half of the W_M table follows the row tiles in LDS
    if (out.stat != 1 || (flags & FF_TUNE_GENERIC)) return false;
    out.stat = 2;
    out.lds += (size_t)(M / 2) * elem_bytes;
  }
  return out.lds <= 160 * 1024;
}

// Configurations measured fastest on MI355X where the default heuristic
// picks a slower one: elements per thread at 768^3 and 1000^3 (FP64), and
// the kernels with LDS twiddles at 1024 (column passes, FP64: 19.7 -> 18.4 ms
// per 1024^3 solve, FP32: 12.3 -> 11.0 ms); with the pencil layout the y passes at 1024
// also take the split exchange (three workgroups per CU: 18.4 -> 17.3 ms;
// in the plain layout the same kernel is slower).  lay: -1 any layout, 0
// plain, 1 pencil.  User overrides (ff_tune_set) take precedence.
struct BuiltinCfg {
  int kind, n, prec, lay, E, T, flags;
};
static const BuiltinCfg kBuiltin[] = {
    // z rows, nz = 1024: LDS twiddles (FP64 z forward 3.10 -> 2.86 ms, z
    // inverse 3.12 -> 2.97 ms plain, 3.53 -> 3.44 ms pencil; FP32 solve
    // 11.0 -> 10.6 ms)
    {0, 512, 0, -1, 8, 2, FF_TUNE_LDS_TWIDDLES},
    {3, 512, 0, -1, 8, 2, FF_TUNE_LDS_TWIDDLES},
    {0, 512, 1, -1, 8, 4, FF_TUNE_LDS_TWIDDLES},
    {3, 512, 1, -1, 8, 4, FF_TUNE_LDS_TWIDDLES},
    {0, 384, 0, -1, 12, 4, 0},   // z forward, nz = 768
    {3, 384, 0, -1, 12, 4, 0},   // z inverse, nz = 768
    {1, 768, 0, -1, 24, 4, 0},   // y, ny = 768
    {2, 768, 0, -1, 12, 4, 0},   // x, nx = 768
    {1, 1000, 0, -1, 20, 4, 0},  // y, ny = 1000
    {2, 1000, 0, -1, 20, 4, 0},  // x, nx = 1000 (compiled-in plan)
    {1, 1024, 0, 0, 16, 4, FF_TUNE_LDS_TWIDDLES},                      // y, ny = 1024
    {1, 1024, 0, 1, 16, 4, FF_TUNE_SPLIT_LDS | FF_TUNE_LDS_TWIDDLES},  // y, ny = 1024, pencil
    // x, nx = 1024: split exchange, three workgroups per CU
    {2, 1024, 0, -1, 16, 4, FF_TUNE_SPLIT_LDS | FF_TUNE_LDS_TWIDDLES},
    {1, 1024, 1, -1, 16, 8, FF_TUNE_LDS_TWIDDLES},  // FP32 y, ny = 1024
    {2, 1024, 1, -1, 16, 8, FF_TUNE_LDS_TWIDDLES},  // FP32 x, nx = 1024
};

static bool builtin_lookup(int kind, int n, int prec, int& E, int& T, int& flags, bool pencil = false) {
  for (const BuiltinCfg& b : kBuiltin)
    if (b.kind == kind && b.n == n && b.prec == prec && (b.lay < 0 || b.lay == (pencil ? 1 : 0))) {
      E = b.E;
      T = b.T;
      flags = b.flags;
      return true;
    }
  return false;
}

bool make_col_cfg(int n, int ncols, int elem_bytes, ColCfg& out, int kind, int prec, bool pencil) {
  int E = 0, T = 0, flags = 0;
  if (tune_lookup(kind, n, prec, E, T, flags) && col_cfg_with(n, ncols, elem_bytes, out, E, T, flags, kind))
    return true;
  if (builtin_lookup(kind, n, prec, E, T, flags, pencil) && col_cfg_with(n, ncols, elem_bytes, out, E, T, flags, kind))
    return true;
  return col_cfg_with(n, ncols, elem_bytes, out, 0, 0, 0, kind);
}

bool make_row_cfg(int nz, int elem_bytes, RowCfg& out, int kind, int prec) {
  const int M = (nz % 2 == 0) ? nz / 2 : nz;
  int E = 0, T = 0, flags = 0;
  if (tune_lookup(kind, M, prec, E, T, flags) && row_cfg_with(nz, elem_bytes, out, E, T, flags)) return true;
  if (builtin_lookup(kind, M, prec, E, T, flags) && row_cfg_with(nz, elem_bytes, out, E, T, flags)) return true;
  return row_cfg_with(nz, elem_bytes, out, 0, 0, 0);
}

ff_status tune_set(const ff_tune* t) {
  if (!t) return set_error(FF_ERR_INVALID, "ff_tune_set: null");
  if (t->kind < 0 || t->kind > 3 || t->n < 1 || (t->precision != 0 && t->precision != 1))
    return set_error(FF_ERR_INVALID, "ff_tune_set: kind 0..3, n >= 1, precision 0/1");
  if (t->E == 0) {
    std::lock_guard<std::mutex> g(tune_mutex());
    tune_table().erase({t->kind, t->n, t->precision});
    return FF_OK;
  }
  const int eb = t->precision ? 8 : 16;
  if (t->flags & ~(FF_TUNE_SPLIT_LDS | FF_TUNE_TWO_GROUPS | FF_TUNE_GENERIC | FF_TUNE_LDS_TWIDDLES))
    return set_error(FF_ERR_INVALID, "ff_tune_set: unknown flags");
  bool ok;
  if (t->kind == 0 || t->kind == 3) {
    RowCfg c;
    ok = (t->flags & ~(FF_TUNE_GENERIC | FF_TUNE_LDS_TWIDDLES)) == 0 &&
         row_cfg_with(2 * t->n, eb, c, t->E, t->T, t->flags);
  } else {
    ColCfg c;
    ok = col_cfg_with(t->n, 1 << 20, eb, c, t->E, t->T, t->flags, t->kind);
  }
  if (!ok) return set_error(FF_ERR_INVALID, "ff_tune_set: no valid configuration for that E/T/flags");
  std::lock_guard<std::mutex> g(tune_mutex());
  tune_table()[{t->kind, t->n, t->precision}] = std::make_tuple(t->E, t->T, t->flags);
  return FF_OK;
}

void tune_reset() {
  std::lock_guard<std::mutex> g(tune_mutex());
  tune_table().clear();
}

// ---- dispatch ---------------------------------------------------------------

#define FF_DECL(E)                                                                                         \
  template <>                                                                                              \
  hipError_t PassLaunchers<E>::col(int, int, int, int, int, int, const Launch&, const ColArgs&);                     \
  template <>                                                                                              \
  hipError_t PassLaunchers<E>::r2c(int, int, int, int, int, const Launch&, const RowArgs&);                \
  template <>                                                                                              \
  hipError_t PassLaunchers<E>::c2r(int, int, int, int, int, int, const Launch&, const RowArgs&);
FF_DECL(1)
FF_DECL(2)
FF_DECL(3)
FF_DECL(4)
FF_DECL(5)
FF_DECL(6)
FF_DECL(8)
FF_DECL(10)
FF_DECL(12)
FF_DECL(15)
FF_DECL(16)
FF_DECL(20)
FF_DECL(24)
FF_DECL(30)
#undef FF_DECL

#define FF_SWITCH_E(E, CALL)        \
  switch (E) {                      \
    case 1: return PassLaunchers<1>::CALL;   \
    case 2: return PassLaunchers<2>::CALL;   \
    case 3: return PassLaunchers<3>::CALL;   \
    case 4: return PassLaunchers<4>::CALL;   \
    case 5: return PassLaunchers<5>::CALL;   \
    case 6: return PassLaunchers<6>::CALL;   \
    case 8: return PassLaunchers<8>::CALL;   \
    case 10: return PassLaunchers<10>::CALL; \
    case 12: return PassLaunchers<12>::CALL; \
    case 15: return PassLaunchers<15>::CALL; \
    case 16: return PassLaunchers<16>::CALL; \
    case 20: return PassLaunchers<20>::CALL; \
    case 24: return PassLaunchers<24>::CALL; \
    case 30: return PassLaunchers<30>::CALL; \
    default: return hipErrorInvalidValue;    \
  }

static thread_local std::vector<PassLaunch>* t_recorder = nullptr;

void set_launch_recorder(std::vector<PassLaunch>* rec) { t_recorder = rec; }
bool launch_recording() { return t_recorder != nullptr; }

static hipError_t do_col(int E, int f32, int mode, int split, int groups, int blue, int stat, const Launch& l,
                         const ColArgs& a) {
  FF_SWITCH_E(E, col(f32, mode, split, groups, blue, stat, l, a))
}
static hipError_t do_r2c(int E, int f32, int in_f32, int odd, int blue, int stat, const Launch& l,
                         const RowArgs& a) {
  FF_SWITCH_E(E, r2c(f32, in_f32, odd, blue, stat, l, a))
}
static hipError_t do_c2r(int E, int f32, int out_kind, int odd, int epi, int blue, int stat, const Launch& l,
                         const RowArgs& a) {
  FF_SWITCH_E(E, c2r(f32, out_kind, odd, epi, blue, stat, l, a))
}

hipError_t launch_col(int E, int f32, int mode, int split, int groups, int blue, int stat, const Launch& l,
                      const ColArgs& a) {
  if (t_recorder) {
    PassLaunch r{};
    r.kind = 0, r.E = E, r.f32 = f32, r.sel = mode, r.split = split, r.groups = groups, r.blue = blue;
    r.stat = stat, r.l = l, r.col = a;
    t_recorder->push_back(r);
    return hipSuccess;
  }
  return do_col(E, f32, mode, split, groups, blue, stat, l, a);
}
hipError_t launch_r2c(int E, int f32, int in_f32, int odd, int blue, int stat, const Launch& l,
                      const RowArgs& a) {
  if (t_recorder) {
    PassLaunch r{};
    r.kind = 1, r.E = E, r.f32 = f32, r.sel = in_f32, r.odd = odd, r.blue = blue, r.stat = stat, r.l = l;
    r.row = a;
    t_recorder->push_back(r);
    return hipSuccess;
  }
  return do_r2c(E, f32, in_f32, odd, blue, stat, l, a);
}
hipError_t launch_c2r(int E, int f32, int out_kind, int odd, int epi, int blue, int stat, const Launch& l,
                      const RowArgs& a) {
  if (t_recorder) {
    PassLaunch r{};
    r.kind = 2, r.E = E, r.f32 = f32, r.sel = out_kind, r.odd = odd, r.epi = epi, r.blue = blue, r.stat = stat;
    r.l = l, r.row = a;
    t_recorder->push_back(r);
    return hipSuccess;
  }
  return do_c2r(E, f32, out_kind, odd, epi, blue, stat, l, a);
}

// The per-instance fields of a recorded launch.
PassInst launch_inst_fields(const PassLaunch& r) {
  if (r.kind == 0)
    return PassInst{r.col.src, r.col.dst, nullptr, r.col.dc_out, nullptr, r.col.ddi, r.col.ddj, r.col.ddk};
  return PassInst{r.row.src, r.row.dst, r.row.aux, nullptr, r.row.shiftp, 0.0, 0.0, 0.0};
}

// Whether two recorded launches differ only in their per-instance fields and
// stream (the recorded argument blocks are zero-filled before use, so their
// bytes compare).
bool launch_same(const PassLaunch& x, const PassLaunch& y) {
  PassLaunch a, b;
  memcpy(&a, &x, sizeof a);
  memcpy(&b, &y, sizeof b);
  for (PassLaunch* r : {&a, &b}) {
    r->l.stream = nullptr;
    r->col.src = nullptr;
    r->col.dst = nullptr;
    r->col.dc_out = nullptr;
    r->col.ddi = r->col.ddj = r->col.ddk = 0.0;
    r->row.src = nullptr;
    r->row.dst = nullptr;
    r->row.aux = nullptr;
    r->row.shiftp = nullptr;
  }
  return memcmp(&a, &b, sizeof a) == 0;
}

hipError_t replay_launch(const PassLaunch& r) {
  switch (r.kind) {
    case 0: return do_col(r.E, r.f32, r.sel, r.split, r.groups, r.blue, r.stat, r.l, r.col);
    case 1: return do_r2c(r.E, r.f32, r.sel, r.odd, r.blue, r.stat, r.l, r.row);
    case 2: return do_c2r(r.E, r.f32, r.sel, r.odd, r.epi, r.blue, r.stat, r.l, r.row);
    default: return hipErrorInvalidValue;
  }
}
#undef FF_SWITCH_E

}  // namespace ff
