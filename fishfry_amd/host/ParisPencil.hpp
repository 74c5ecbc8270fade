// C++ facade with the reference's solver-plugin interface (ParisPencil.hpp:11-59)
// over the C-ABI of libfishfry.so.  FishFry<ParisPencil> instantiates it exactly
// as the reference's driver does (FishFry.hpp:135,141,180).
#pragma once

#include <cstdio>
#include <cstdlib>
#include <vector>

#include "../../include/fishfry.h"
#include "TimeStamp.hpp"

class ParisPencil {
 public:
  // n: global cells; lo/hi: physical bounds (hi one cell below the periodic
  // end); m: tasks per axis (1x1x1 only: one GPU); id: task coordinates.
  ParisPencil(const int n[3], const double lo[3], const double hi[3], const int m[3], const int id[3],
              const ff_options* opts = nullptr) {
    check(ff_plan_create(n, lo, hi, m, id, opts, &plan_), "ParisPencil");
  }

  ~ParisPencil() { ff_plan_destroy(plan_); }

  ParisPencil(const ParisPencil&) = delete;
  ParisPencil& operator=(const ParisPencil&) = delete;

  // Bytes needed for each of density and potential (ParisPencil.hpp:30).
  size_t bytes() const { return ff_plan_bytes(plan_); }

  // density: input, clobbered as work space (FP64); potential: output phi.
  // Returns after the device work completes, like the reference.
  void solve(size_t bytes, double* density, double* potential, std::vector<TimeStamp>& stamps) const {
    ff_stamp st[FF_SOLVE_STAMPS];
    int n = FF_SOLVE_STAMPS;
    check(ff_solve(plan_, bytes, density, potential, nullptr, st, &n), "ParisPencil::solve");
    stamps.clear();
    for (int i = 0; i < n; ++i) stamps.push_back({st[i].name, st[i].seconds});
  }

  const ff_plan* plan() const { return plan_; }

 private:
  // The reference reports failures and exits (gpu.hpp:46-60,87); the facade
  // keeps that behaviour on top of the library's status codes.
  static void check(ff_status s, const char* what) {
    if (s != FF_OK) {
      fprintf(stderr, "FISHFRY ERROR in %s: status %d: %s\n", what, (int)s, ff_last_error());
      fflush(stderr);
      exit((int)s);
    }
  }
  ff_plan* plan_ = nullptr;
};
