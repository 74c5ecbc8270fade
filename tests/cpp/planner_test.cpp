// CPU test of the solver's host-side planning (fishfry_amd/csrc/ff_plan.cpp),
// built with AddressSanitizer and UndefinedBehaviorSanitizer: line plans for
// every length 1..4200 (direct radix-2/3/4/5/... plans and Bluestein), the
// column and row tile configurations, the tuning table, the long-double
// twiddles, Bluestein tables and finite-difference factors.  No GPU is
// touched: the pass launchers are link stand-ins that are never called.
#include <cmath>
#include <complex>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "ff_internal.hpp"

namespace ff {
#define FF_TEST_NOLAUNCH(E)                                                                                  \
  template <>                                                                                               \
  hipError_t PassLaunchers<E>::col(int, int, int, int, int, int, const Launch&, const ColArgs&) {          \
    return hipErrorInvalidValue;                                                                             \
  }                                                                                                          \
  template <>                                                                                               \
  hipError_t PassLaunchers<E>::r2c(int, int, int, int, int, const Launch&, const RowArgs&) {                 \
    return hipErrorInvalidValue;                                                                             \
  }                                                                                                          \
  template <>                                                                                               \
  hipError_t PassLaunchers<E>::c2r(int, int, int, int, int, int, const Launch&, const RowArgs&) {            \
    return hipErrorInvalidValue;                                                                             \
  }
FF_TEST_NOLAUNCH(1)
FF_TEST_NOLAUNCH(2)
FF_TEST_NOLAUNCH(3)
FF_TEST_NOLAUNCH(4)
FF_TEST_NOLAUNCH(5)
FF_TEST_NOLAUNCH(6)
FF_TEST_NOLAUNCH(8)
FF_TEST_NOLAUNCH(10)
FF_TEST_NOLAUNCH(12)
FF_TEST_NOLAUNCH(15)
FF_TEST_NOLAUNCH(16)
FF_TEST_NOLAUNCH(20)
FF_TEST_NOLAUNCH(24)
FF_TEST_NOLAUNCH(30)
#undef FF_TEST_NOLAUNCH
}  // namespace ff

#define REQUIRE(c)                                                     \
  do {                                                                 \
    if (!(c)) {                                                        \
      std::fprintf(stderr, "FAILED %s (line %d)\n", #c, __LINE__);     \
      std::exit(1);                                                    \
    }                                                                  \
  } while (0)

static bool smooth(long long m) {
  for (int p : {2, 3, 5})
    while (m % p == 0) m /= p;
  return m == 1;
}

// A line plan is a valid Stockham plan of its engine length.
static void check_plan(const ff::LineCfg& c, int n) {
  const ff::LinePlan& lp = c.lp;
  REQUIRE(c.n == n);
  REQUIRE(lp.n >= n && lp.L * c.E == lp.n && lp.np >= (lp.n > 1 ? 1 : 0) && lp.np <= FF_MAXP);
  long long prod = 1;
  for (int p = 0; p < lp.np; ++p) {
    REQUIRE(lp.radix[p] >= 2 && c.E % lp.radix[p] == 0);
    REQUIRE(lp.ns[p] == prod);
    REQUIRE((long long)lp.twstep[p] * lp.ns[p] * lp.radix[p] == lp.n);
    prod *= lp.radix[p];
  }
  REQUIRE(prod == lp.n);
  if (c.bluestein()) REQUIRE(lp.n >= 2 * n - 1 && smooth(lp.n) && !smooth(n));
  else REQUIRE(lp.n == n);
}

int main() {
  int direct = 0, blue = 0;
  for (int n = 1; n <= 4200; ++n) {
    ff::LineCfg c;
    const bool ok = ff::plan_line(n, c);
    if (n <= 4096) REQUIRE(ok);
    if (!ok) continue;
    check_plan(c, n);
    if (n <= 4096) (c.bluestein() ? blue : direct)++;
    for (int E : ff::kEValues) {  // forced elements per thread: valid or refused, never broken
      ff::LineCfg f;
      if (ff::plan_line(n, f, E)) check_plan(f, n);
    }
  }
  // 1..4096: the 137 lengths of the form 2^a 3^b 5^c are direct, the rest Bluestein
  REQUIRE(direct == 137 && blue == 3959);
  // tile configurations
  for (int n : {2, 3, 7, 32, 48, 96, 100, 160, 256, 500, 512, 768, 1000, 1024, 2048, 4096})
    for (int eb : {8, 16})
      for (int kind : {1, 2}) {
        ff::ColCfg c;
        if (ff::make_col_cfg(n, n / 2 + 1, eb, c, kind, eb == 8)) {
          REQUIRE(c.T >= 1 && (c.T & (c.T - 1)) == 0 && c.threads == c.T * c.line.lp.L && c.lds <= 160 * 1024);
          REQUIRE(c.threads <= 1024);
        }
        ff::RowCfg r;
        if (ff::make_row_cfg(n, eb, r, kind == 1 ? 0 : 3, eb == 8)) {
          // the row tiles, then half the twiddle table where it is read from LDS
          const size_t twl = r.stat == 2 ? (size_t)(r.line.lp.n / 2) * eb : 0;
          REQUIRE(r.T >= 1 && r.threads == r.T * r.line.lp.L && r.lds == (size_t)r.T * r.pitch * eb + twl);
          REQUIRE(r.stat != 2 || n == 1024);
          REQUIRE(r.lds <= 160 * 1024 && r.threads <= 1024);
        }
      }
  // tuning table round trip
  ff::tune_reset();
  const ff_tune t{2, 1024, 0, 16, 4, FF_TUNE_GENERIC};
  REQUIRE(ff::tune_set(&t) == FF_OK);
  int E = 0, T = 0, fl = 0;
  REQUIRE(ff::tune_lookup(2, 1024, 0, E, T, fl) && E == 16 && T == 4 && fl == FF_TUNE_GENERIC);
  const ff_tune bad{2, 1024, 0, 7, 4, 0};
  REQUIRE(ff::tune_set(&bad) != FF_OK);
  ff::tune_reset();
  REQUIRE(!ff::tune_lookup(2, 1024, 0, E, T, fl));
  // twiddles against std::polar in long double
  for (long long n : {1LL, 3LL, 8LL, 1000LL, 1024LL, 4095LL})
    for (long long t = 0; t < n; t += (n > 64 ? 7 : 1)) {
      long double c, s;
      ff::twiddle_ld(t, n, c, s);
      const long double a = -2.0L * 3.14159265358979323846264338327950288L * (long double)t / (long double)n;
      REQUIRE(std::fabs((double)(c - std::cos(a))) < 1e-17 && std::fabs((double)(s - std::sin(a))) < 1e-17);
    }
  // Bluestein tables: the chirp, and both kernel spectra (length m), finite
  for (int n : {7, 37, 101, 509}) {
    ff::LineCfg c;
    REQUIRE(ff::plan_line(n, c) && c.bluestein());
    for (int prec : {0, 1}) {
      std::vector<unsigned char> buf;
      size_t o[3];
      ff::push_bluestein(buf, n, c.lp.n, prec, o[0], o[1], o[2]);
      const size_t el = prec ? 8 : 16;
      REQUIRE(o[0] < o[1] && o[1] < o[2] && o[2] + (size_t)c.lp.n * el <= buf.size());
      REQUIRE(o[1] - o[0] >= (size_t)n * el);
      for (size_t b = 0; b + 8 <= buf.size(); b += 8) {
        double v;
        if (prec) {
          float w[2];
          std::memcpy(w, &buf[b], 8);
          v = (double)w[0] + (double)w[1];
        } else {
          std::memcpy(&v, &buf[b], 8);
        }
        REQUIRE(std::isfinite(v));
      }
    }
  }
  // finite-difference factors: symmetric, 0 at k = 0, at most 4 (16/3)
  for (int n : {2, 3, 64, 1000}) {
    for (int k = 0; k < n; ++k) {
      const long double g3 = ff::fd_factor_ld(k, n, 3), g5 = ff::fd_factor_ld(k, n, 5);
      REQUIRE(g3 >= 0 && g3 <= 4.0L + 1e-18L && g5 >= g3 && g5 <= 16.0L / 3 + 1e-18L);
      REQUIRE(ff::fd_factor_ld(n - k, n, 3) == g3 && ff::fd_factor_ld(n - k, n, 5) == g5);
    }
    REQUIRE(ff::fd_factor_ld(0, n, 3) == 0 && ff::fd_factor_ld(0, n, 5) == 0);
  }
  std::printf("planner ok: %d direct and %d Bluestein lengths\n", direct, blue);
  return 0;
}
