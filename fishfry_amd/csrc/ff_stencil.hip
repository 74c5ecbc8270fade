// Finite-difference operator of FF_SYMBOL_3PT / FF_SYMBOL_5PT plans in real
// space: the 7-point (3-point per axis) or 13-point (5-point per axis)
// periodic stencil
//   3-point  (u[j+1] - 2u[j] + u[j-1]) / h^2
//   5-point  (-u[j+2] + 16u[j+1] - 30u[j] + 16u[j-1] - u[j-2]) / (12 h^2)
// summed over the three axes (h = (hi - lo)/(n - 1) per axis).  Its symbol is
// the one the x pass inverts, so refinement can take its FP64 residual
// r = (f - mean f) - L_h phi here, in one pass over f and phi (20 bytes per
// point with an FP32 r), instead of an FP64 transform round trip.
//
// A workgroup owns `rows` consecutive rows (i, j) of the [i][j][k] array;
// thread t handles positions k = t mod TPR, k + TPR, ... of row t / TPR of
// each group of 256 / TPR rows, so short rows still keep every lane busy.
// Neighbour rows along x and y are read directly (they are the rows this and
// nearby workgroups read as centres, so they come from the caches); the sum
// of phi for the mean removal is reduced in a fixed order: per thread, per
// workgroup (tree), then over workgroups (k_stencil_sum).
#include <algorithm>
#include <climits>
#include <cstring>

#include "ff_solver.hpp"

namespace ff {

constexpr int kStencilThreads = 256;

struct StencilArgs {
  const double* phi;
  const double* f;
  void* out;               // float r (residual) or double L_h phi
  int nx, ny, nz;
  int rows, nblocks;       // rows per workgroup, workgroups
  int lgtpr;               // log2 of the threads per row
  double cx, cy, cz;       // 1/h^2 per axis (5-point: divided by 12)
  double scale;            // 1/(nx ny nz)
  const double* sum_f;     // residual: sum of f
  double* partial;         // residual: sum of phi per workgroup
};

__device__ __forceinline__ int wrap_idx(int v, int n) {
  v %= n;
  return v < 0 ? v + n : v;
}

template <int ORDER, bool RESID>
__global__ __launch_bounds__(kStencilThreads) void k_stencil(const StencilArgs a) {
  __shared__ double red[kStencilThreads];
  const int tid = threadIdx.x;
  const int tpr = 1 << a.lgtpr;
  const int sub = tid >> a.lgtpr, k0 = tid & (tpr - 1);
  const int rpi = kStencilThreads >> a.lgtpr;  // rows per iteration
  const int b = xcd_remap(blockIdx.x, a.nblocks);
  const long long nrows = (long long)a.nx * a.ny;
  const long long nz = a.nz;
  const FF_GLB double* phi = glb<const double>(a.phi);
  double meanf = 0.0;
  if constexpr (RESID) meanf = *glb<const double>(a.sum_f) * a.scale;
  double acc = 0.0;
  for (int r0 = 0; r0 < a.rows; r0 += rpi) {
    const long long row = (long long)b * a.rows + r0 + sub;
    if (r0 + sub >= a.rows || row >= nrows) continue;
    const int i = (int)(row / a.ny), j = (int)(row - (long long)i * a.ny);
    const FF_GLB double* c0 = phi + row * nz;
    const FF_GLB double* xm1 = phi + ((long long)wrap_idx(i - 1, a.nx) * a.ny + j) * nz;
    const FF_GLB double* xp1 = phi + ((long long)wrap_idx(i + 1, a.nx) * a.ny + j) * nz;
    const FF_GLB double* ym1 = phi + ((long long)i * a.ny + wrap_idx(j - 1, a.ny)) * nz;
    const FF_GLB double* yp1 = phi + ((long long)i * a.ny + wrap_idx(j + 1, a.ny)) * nz;
    const FF_GLB double *xm2 = xm1, *xp2 = xp1, *ym2 = ym1, *yp2 = yp1;
    if constexpr (ORDER == 5) {
      xm2 = phi + ((long long)wrap_idx(i - 2, a.nx) * a.ny + j) * nz;
      xp2 = phi + ((long long)wrap_idx(i + 2, a.nx) * a.ny + j) * nz;
      ym2 = phi + ((long long)i * a.ny + wrap_idx(j - 2, a.ny)) * nz;
      yp2 = phi + ((long long)i * a.ny + wrap_idx(j + 2, a.ny)) * nz;
    }
    for (int k = k0; k < a.nz; k += tpr) {
      const double c = c0[k];
      const int km1 = wrap_idx(k - 1, a.nz), kp1 = wrap_idx(k + 1, a.nz);
      double dx, dy, dz;
      if constexpr (ORDER == 3) {
        dx = fmadd(-2.0, c, xp1[k] + xm1[k]);
        dy = fmadd(-2.0, c, yp1[k] + ym1[k]);
        dz = fmadd(-2.0, c, c0[kp1] + c0[km1]);
      } else {
        const int km2 = wrap_idx(k - 2, a.nz), kp2 = wrap_idx(k + 2, a.nz);
        dx = fmadd(-30.0, c, fmadd(16.0, xp1[k] + xm1[k], -(xp2[k] + xm2[k])));
        dy = fmadd(-30.0, c, fmadd(16.0, yp1[k] + ym1[k], -(yp2[k] + ym2[k])));
        dz = fmadd(-30.0, c, fmadd(16.0, c0[kp1] + c0[km1], -(c0[kp2] + c0[km2])));
      }
      const double lap = fmadd(a.cz, dz, fmadd(a.cy, dy, a.cx * dx));  // (x + y) + z
      if constexpr (RESID) {
        const double fv = glb<const double>(a.f)[row * nz + k];
        glb<float>(a.out)[row * nz + k] = float((fv - meanf) - lap);
        acc += c;
      } else {
        glb<double>(a.out)[row * nz + k] = lap;
      }
    }
  }
  if constexpr (RESID) {
    red[tid] = acc;
    __syncthreads();
    for (int w = kStencilThreads / 2; w > 0; w >>= 1) {
      if (tid < w) red[tid] += red[tid + w];
      __syncthreads();
    }
    if (tid == 0) glb<double>(a.partial)[b] = red[0];
  }
}

// out[0] = sum of partial[0 .. n), in a fixed order.
__global__ __launch_bounds__(kStencilThreads) void k_stencil_sum(const double* partial, int n, double* out) {
  __shared__ double red[kStencilThreads];
  const int tid = threadIdx.x;
  double acc = 0.0;
  for (int t = tid; t < n; t += kStencilThreads) acc += glb<const double>(partial)[t];
  red[tid] = acc;
  __syncthreads();
  for (int w = kStencilThreads / 2; w > 0; w >>= 1) {
    if (tid < w) red[tid] += red[tid + w];
    __syncthreads();
  }
  if (tid == 0) *glb<double>(out) = red[0];
}

ff_status run_stencil(const ff_plan* p, const double* phi, const double* f, void* out, bool r_f64,
                      const double* sum_f, double* partial, double* sum_phi, hipStream_t s) {
  if (p->sym != 3 && p->sym != 5) return set_error(FF_ERR_INVALID, "stencil: the plan has no finite-difference symbol");
  StencilArgs a;
  memset(&a, 0, sizeof a);
  a.phi = phi;
  a.f = f;
  a.out = out;
  a.nx = p->n[0];
  a.ny = p->n[1];
  a.nz = p->n[2];
  int tpr = 1;
  while (tpr < a.nz && tpr < kStencilThreads) tpr *= 2;
  a.lgtpr = 0;
  while ((1 << a.lgtpr) < tpr) a.lgtpr++;
  // about 8 points per thread
  const long long per = std::max(1LL, 8LL * kStencilThreads / a.nz);
  a.rows = (int)std::min<long long>(per, 1 << 16);
  const long long nrows = (long long)a.nx * a.ny;
  const long long nb = (nrows + a.rows - 1) / a.rows;
  if (nb > INT_MAX) return set_error(FF_ERR_UNSUPPORTED, "stencil: grid too large");
  a.nblocks = (int)nb;
  const double c12 = p->sym == 5 ? 12.0 : 1.0;
  a.cx = p->fd[0] / c12;
  a.cy = p->fd[1] / c12;
  a.cz = p->fd[2] / c12;
  a.scale = 1.0 / (double(p->n[0]) * double(p->n[1]) * double(p->n[2]));
  a.sum_f = sum_f;
  a.partial = partial;
  void* argv[] = {&a};
  const void* fn;
  if (r_f64) fn = p->sym == 3 ? (const void*)&k_stencil<3, false> : (const void*)&k_stencil<5, false>;
  else fn = p->sym == 3 ? (const void*)&k_stencil<3, true> : (const void*)&k_stencil<5, true>;
  ff_status st = hip_check(hipLaunchKernel(fn, dim3(a.nblocks), dim3(kStencilThreads), argv, 0, s), "stencil launch");
  if (st != FF_OK || r_f64) return st;
  const void* sargv[] = {&partial, &a.nblocks, &sum_phi};
  return hip_check(hipLaunchKernel((const void*)&k_stencil_sum, dim3(1), dim3(kStencilThreads),
                                   const_cast<void**>(sargv), 0, s),
                   "stencil sum launch");
}

}  // namespace ff

using namespace ff;

extern "C" ff_status ff_stencil_apply(const ff_plan* plan, const double* phi, double* out, void* stream) {
  clear_error();
  if (!plan || !phi || !out) return set_error(FF_ERR_INVALID, "ff_stencil_apply: null argument");
  if (phi == out) return set_error(FF_ERR_INVALID, "ff_stencil_apply: phi and out must differ");
  return run_stencil(plan, phi, nullptr, out, true, nullptr, nullptr, nullptr, (hipStream_t)stream);
}
