// fishfry mx my mz nx ny nz [runs] [--fp32-refine STEPS]
// The reference's CLI (main.cpp:12-23) for one GPU: the task grid must be 1 1 1.
#include <cstdio>
#include <cstdlib>
#include <cstring>

#include "FishFry.hpp"
#include "ParisPencil.hpp"

int main(int argc, char** argv) {
  if (argc <= 6) {
    fprintf(stderr, "usage: %s mx my mz nx ny nz [runs] [--fp32-refine STEPS]\n", argv[0]);
    return 1;
  }
  int nTasks[3] = {0, 0, 0};
  int nPoints[3] = {0, 0, 0};
  int nIters = 1;
  for (int i = 0; i < 3; i++) nTasks[i] = atoi(argv[i + 1]);
  for (int i = 0; i < 3; i++) nPoints[i] = atoi(argv[i + 4]);
  int arg = 7;
  if (argc > 7 && argv[7][0] != '-') {
    nIters = atoi(argv[7]);
    arg = 8;
  }
  ff_options opts{FF_FP64, 0, 0};
  for (; arg < argc; ++arg) {
    if (!strcmp(argv[arg], "--fp32-refine") && arg + 1 < argc) {
      opts.precision = FF_FP32_REFINE;
      opts.refine_steps = atoi(argv[++arg]);
    } else {
      fprintf(stderr, "unknown argument %s\n", argv[arg]);
      return 1;
    }
  }
  // main.cpp:16-21 preconditions
  if (nTasks[0] <= 0 || nTasks[1] <= 0 || nTasks[2] <= 0 || nPoints[0] < nTasks[0] || nPoints[1] < nTasks[1] ||
      nPoints[2] < nTasks[2] || nIters < 0) {
    fprintf(stderr, "invalid task grid, point grid or run count\n");
    return 1;
  }
  if (nTasks[0] * nTasks[1] * nTasks[2] != 1) {
    fprintf(stderr, "only 1x1x1 tasks are supported (one process, one GPU)\n");
    return 2;
  }
  FishFry<ParisPencil>(nTasks, nPoints, &opts).run(nIters);
  return 0;
}
