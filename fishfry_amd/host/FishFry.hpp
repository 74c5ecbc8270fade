// The benchmark driver FishFry<P> (FishFry.hpp:103-246) for one GPU.
//
// Same setup, timed loop, per-run output line and phase table as the
// reference; the MPI reductions collapse to the single rank, the init and
// diff kernels are the library's (ff_field_init / ff_field_diff, the latter
// with a deterministic reduction), and run() also reports the p50/p99 solve
// latency of the timed runs.
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <memory>
#include <vector>

#include "../../include/fishfry.h"
#include "TimeStamp.hpp"

template <typename P>
class FishFry {
 public:
  FishFry(const int nTasks[3], const int nPoints[3], const ff_options* opts = nullptr)
      : totalPoints_(double(nPoints[0]) * double(nPoints[1]) * double(nPoints[2])) {
    if (nTasks[0] != 1 || nTasks[1] != 1 || nTasks[2] != 1) {
      fprintf(stderr, "FishFry: only 1x1x1 tasks (one GPU) are supported\n");
      exit(2);
    }
    for (int a = 0; a < 3; ++a) n_[a] = nPoints[a];
    printf("FishFry<%s>: %dx%dx%d points over %dx%dx%d MPI processes\n", "ParisPencil", nPoints[0], nPoints[1],
           nPoints[2], nTasks[0], nTasks[1], nTasks[2]);
    fflush(stdout);
    ff_field_default(nPoints, &field_);  // dx = 2pi/n, lo = -pi (FishFry.hpp:109-111,133)
    const int id[3] = {0, 0, 0};
    double hi[3];
    for (int a = 0; a < 3; ++a) hi[a] = M_PI - field_.d[a];  // FishFry.hpp:134
    poisson_.reset(new P(nPoints, field_.lo, hi, nTasks, id, opts));
    bytes_ = std::max(poisson_->bytes(), size_t(totalPoints_) * sizeof(double));  // FishFry.hpp:141
    check(hipMalloc(&f_, bytes_), "hipMalloc f");
    check(hipMalloc(&phi_, bytes_), "hipMalloc phi");
    double s[3];
    check(ff_field_diff(&field_, nullptr, s, nullptr), "diff");  // FishFry.hpp:146-159
    scales_[0] = 1.0 / s[0];
    scales_[1] = 1.0 / s[1];
    scales_[2] = 1.0 / s[2];
  }

  ~FishFry() {
    poisson_.reset();
    (void)hipFree(phi_);
    (void)hipFree(f_);
  }

  void run(const int nIters) {
    std::vector<TimeStamp> totalStamps;
    std::vector<double> times;
    for (int i = 0; i <= nIters; i++) {
      check(ff_field_init(&field_, f_, nullptr), "init");  // FishFry.hpp:174
      check(hipDeviceSynchronize(), "sync");
      std::vector<TimeStamp> stamps;
      poisson_->solve(bytes_, f_, phi_, stamps);
      const double time = stamps.back().second - stamps.front().second;
      if (i == 0) {
        totalStamps = stamps;
        for (auto& s : totalStamps) s.second = 0;
      } else {
        for (unsigned j = 0; j < stamps.size(); j++) totalStamps.at(j).second += stamps.at(j).second;
        times.push_back(time);
      }
      double d[3];
      check(ff_field_diff(&field_, phi_, d, nullptr), "diff");
      printf("%d: Time %10.4e Rate %10.4e L_1 %10.4e L_2 %10.4e L_\\inf %10.4e", i, time, totalPoints_ / time,
             d[0] * scales_[0], sqrt(d[1] * scales_[1]), d[2] * scales_[2]);
      if (i == 0) printf(" (warmup, ignored)");
      printf("\n");
      fflush(stdout);
    }
    printf("\nTimes Averaged across Iterations\n");
    printf("Min Rank   | Avg Rank   | Max Rank   (%%Total) | Phase\n");
    printf("--------------------------------------------------------\n");
    const double perIter = 1.0 / double(nIters > 0 ? nIters : 1);
    double percent = 0;
    for (unsigned i = 0; i < totalStamps.size(); i++) {
      const double dt = (i == 0) ? totalStamps.back().second - totalStamps.front().second
                                 : totalStamps[i].second - totalStamps[i - 1].second;
      const double t = perIter * dt;
      if (i == 0) percent = t > 0 ? 100.0 / t : 0.0;
      printf("%10.4e | %10.4e | %10.4e (%5.1f%%) | %s\n", t, t, t, t * percent, totalStamps[i].first.c_str());
    }
    if (!times.empty()) {
      std::sort(times.begin(), times.end());
      auto pct = [&](double q) {
        const double pos = q * double(times.size() - 1);
        const size_t lo = size_t(pos);
        const size_t hi = std::min(lo + 1, times.size() - 1);
        return times[lo] + (pos - double(lo)) * (times[hi] - times[lo]);
      };
      printf("\nSolve latency p50 %10.4e s  p99 %10.4e s  Rate(p50) %10.4e points/s\n", pct(0.5), pct(0.99),
             totalPoints_ / pct(0.5));
    }
    fflush(stdout);
  }

 private:
  static void check(hipError_t e, const char* what) {
    if (e != hipSuccess) {
      fprintf(stderr, "HIP ERROR in %s: %s\n", what, hipGetErrorString(e));
      exit(int(e));
    }
  }
  static void check(ff_status s, const char* what) {
    if (s != FF_OK) {
      fprintf(stderr, "FISHFRY ERROR in %s: %s\n", what, ff_last_error());
      exit(int(s));
    }
  }

  int n_[3];
  double totalPoints_;
  std::unique_ptr<P> poisson_;
  ff_field field_;
  size_t bytes_ = 0;
  double* f_ = nullptr;
  double* phi_ = nullptr;
  double scales_[3] = {0, 0, 0};
};
