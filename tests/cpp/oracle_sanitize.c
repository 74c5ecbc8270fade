/* CPU run of the oracle (oracle/oracle.c) under AddressSanitizer and
 * UndefinedBehaviorSanitizer: every entry point on small odd, even and prime
 * shapes, with round-trip and residual checks so the run also computes
 * something checkable.  Test infrastructure only. */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>

#include "oracle.h"

#define REQUIRE(c)                                                   \
  do {                                                               \
    if (!(c)) {                                                      \
      fprintf(stderr, "FAILED %s (line %d)\n", #c, __LINE__);        \
      exit(1);                                                       \
    }                                                                \
  } while (0)

static double frand(unsigned *s) {
  *s = *s * 1103515245u + 12345u;
  return ((*s >> 8) & 0xffff) / 65536.0 - 0.5;
}

int main(void) {
  unsigned seed = 20261202u;
  const int shapes[][3] = {{5, 6, 4}, {7, 5, 11}, {8, 8, 8}, {12, 10, 9}, {2, 3, 13}, {3, 2, 2}};
  for (size_t s = 0; s < sizeof shapes / sizeof shapes[0]; s++) {
    const int *n = shapes[s];
    const long N = (long)n[0] * n[1] * n[2], Nh = (long)n[0] * n[1] * (n[2] / 2 + 1);
    double lo[3], hi[3], dx[3];
    for (int a = 0; a < 3; a++) {
      dx[a] = 2 * M_PI / n[a];
      lo[a] = -M_PI;
      hi[a] = M_PI - dx[a];
    }
    double *f = malloc(sizeof(double) * N), *phi = malloc(sizeof(double) * N), *lap = malloc(sizeof(double) * N);
    double *work = malloc(sizeof(double) * 2 * Nh), *back = malloc(sizeof(double) * N);
    double mean = 0;
    for (long t = 0; t < N; t++) mean += (f[t] = frand(&seed));
    mean /= N;
    /* rfft -> irfft round trip along z */
    REQUIRE(ora_rfft(n[2], f, work, (long)n[0] * n[1]) == 0);
    REQUIRE(ora_irfft(n[2], work, back, (long)n[0] * n[1]) == 0);
    for (long t = 0; t < N; t++) REQUIRE(fabs(back[t] / n[2] - f[t]) < 1e-13);
    /* complex fft forward and inverse on strided lines */
    for (long t = 0; t < 2 * Nh; t++) work[t] = frand(&seed);
    for (long t = 0; t < 2 * Nh; t++) back[t % N] = work[t];
    REQUIRE(ora_fft(n[0], work, Nh / n[0], Nh / n[0], 1, -1) == 0);
    REQUIRE(ora_fft(n[0], work, Nh / n[0], Nh / n[0], 1, +1) == 0);
    /* every symbol: L_h phi = f - mean f for the finite-difference ones */
    for (int order = 0; order <= 5; order += (order ? 2 : 3)) {
      REQUIRE(ora_solve_sym(n, lo, hi, order, f, phi, work, 1) == 0);
      if (order) {
        ora_stencil(n, lo, hi, order, phi, lap);
        for (long t = 0; t < N; t++) REQUIRE(fabs(lap[t] - (f[t] - mean)) < 1e-10);
      }
    }
    ora_filter(n, lo, hi, work);
    ora_filter_sym(n, lo, hi, 3, work);
    /* benchmark field and norms */
    ora_init_field(n[0], n[1], n[2], lo[0], lo[1], lo[2], dx[0], dx[1], dx[2], f);
    double e[3];
    ora_diff(n[0], n[1], n[2], lo[0], lo[1], lo[2], dx[0], dx[1], dx[2], phi, e);
    REQUIRE(e[0] >= 0 && e[1] >= 0 && e[2] >= 0);
    const int m1[3] = {1, 1, 1}, m2[3] = {2, 3, 4};
    REQUIRE(ora_bytes(n, m1) == (size_t)16 * n[0] * n[1] * (n[2] / 2 + 1));
    REQUIRE(ora_bytes(n, m2) > 0);
    REQUIRE(ora_dd(n[0], lo[0], hi[0]) > 0);
    free(f);
    free(phi);
    free(lap);
    free(work);
    free(back);
  }
  double c, sn;
  for (long t = 0; t < 97; t++) ora_twiddle(t, 97, &c, &sn);
  REQUIRE(ora_fd_factor(0, 8, 3) == 0.0 && ora_fd_factor(4, 8, 3) == 4.0);
  printf("oracle ok\n");
  return 0;
}
