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
// A workgroup owns a tile of TJ rows (y) by TK positions (z) and marches
// along x through a chunk of planes.  Each thread keeps the x-line of its
// (j, k) in a register queue phi[i - H .. i + H] (H = 1 or 2), so every
// value of phi is read from memory once as it enters a queue; the tile of
// plane i, with an H-wide halo in y and z, goes to LDS (two buffers,
// alternating by plane, so one barrier per plane suffices) for the y and z
// neighbours.  Tiles and halos wrap periodically; lanes past the grid's
// edge load wrapped values so the LDS tile is consistent, and store nothing.
// The sum of phi for the mean removal is reduced in a fixed order: per
// thread (by plane), per workgroup (tree), then over workgroups
// (k_stencil_sum).
#include <algorithm>
#include <climits>
#include <cstring>

#include "ff_solver.hpp"

namespace ff {

constexpr int kStencilThreads = 256;

struct StencilArgs {
  const double* phi;
  const double* f;
  void* out;           // float r (residual) or double L_h phi
  int nx, ny, nz;
  int lgtk;            // log2 of TK (z positions per tile); TJ = 256 / TK rows
  int tiles_j, tiles_k, chunks, chunk;  // workgroup grid: tiles_j * tiles_k * chunks; planes per chunk
  double cx, cy, cz;   // 1/h^2 per axis (5-point: divided by 12)
  double scale;        // 1/(nx ny nz)
  const double* sum_f; // residual: sum of f
  double* partial;     // residual: sum of phi per workgroup
};

__device__ __forceinline__ int wrap_idx(int v, int n) {
  v %= n;
  return v < 0 ? v + n : v;
}

template <int ORDER, bool RESID>
__global__ __launch_bounds__(kStencilThreads) void k_stencil(const StencilArgs a) {
  constexpr int H = ORDER == 3 ? 1 : 2;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  double* lds = reinterpret_cast<double*>(smem_raw);
  __shared__ double red[kStencilThreads];
  const int tid = threadIdx.x;
  const int TK = 1 << a.lgtk, TJ = kStencilThreads >> a.lgtk;
  const int W = TK + 2 * H;                 // LDS tile pitch
  const int tile_cells = (TJ + 2 * H) * W;  // one buffer
  const int b = xcd_remap(blockIdx.x, a.tiles_j * a.tiles_k * a.chunks);
  const int c = b % a.chunks;
  const int t = b / a.chunks;
  const int tj = t / a.tiles_k, tk = t - (t / a.tiles_k) * a.tiles_k;
  const int jl = tid >> a.lgtk, kl = tid & (TK - 1);
  const int j0 = tj * TJ, k0 = tk * TK;
  const int j = j0 + jl, k = k0 + kl;
  const bool valid = j < a.ny && k < a.nz;
  const int jw = wrap_idx(j, a.ny), kw = wrap_idx(k, a.nz);
  const long long plane = (long long)a.ny * a.nz;
  const long long own = (long long)jw * a.nz + kw;  // offset of (j, k) in a plane
  // halo cells of this thread (the same every plane): plane offset and LDS index
  const int halo = tile_cells - TJ * TK;
  long long hoff[2];
  int hidx[2];
  int nh = 0;
  for (int e = tid; e < halo && nh < 2; e += kStencilThreads) {
    // enumerate the ring: rows -H..-1 and TJ..TJ+H-1 whole, then the side columns
    int jj, kk;
    const int top = 2 * H * W;
    if (e < top) {
      const int r = e / W;
      jj = r < H ? r - H : TJ + (r - H);
      kk = e - r * W - H;
    } else {
      const int e2 = e - top, r = e2 / (2 * H), q = e2 - r * 2 * H;
      jj = r;
      kk = q < H ? q - H : TK + (q - H);
    }
    hoff[nh] = (long long)wrap_idx(j0 + jj, a.ny) * a.nz + wrap_idx(k0 + kk, a.nz);
    hidx[nh] = (jj + H) * W + (kk + H);
    nh++;
  }
  const FF_GLB double* phi = glb<const double>(a.phi);
  double meanf = 0.0;
  if constexpr (RESID) meanf = *glb<const double>(a.sum_f) * a.scale;
  const int i0 = c * a.chunk, i1 = min(a.nx, i0 + a.chunk);
  double q[2 * H + 1];
#pragma unroll
  for (int m = 0; m < 2 * H; ++m) q[m] = phi[(long long)wrap_idx(i0 - H + m, a.nx) * plane + own];
  double acc = 0.0;
  const int me = (jl + H) * W + (kl + H);
  for (int i = i0; i < i1; ++i) {
    q[2 * H] = phi[(long long)wrap_idx(i + H, a.nx) * plane + own];
    double* buf = lds + (i & 1) * tile_cells;
    const long long pb = (long long)i * plane;
    buf[me] = q[H];
    for (int h = 0; h < nh; ++h) buf[hidx[h]] = phi[pb + hoff[h]];
    __syncthreads();
    const double cv = q[H];
    double dx, dy, dz;
    if constexpr (ORDER == 3) {
      dx = fmadd(-2.0, cv, q[2] + q[0]);
      dy = fmadd(-2.0, cv, buf[me + W] + buf[me - W]);
      dz = fmadd(-2.0, cv, buf[me + 1] + buf[me - 1]);
    } else {
      dx = fmadd(-30.0, cv, fmadd(16.0, q[3] + q[1], -(q[4] + q[0])));
      dy = fmadd(-30.0, cv, fmadd(16.0, buf[me + W] + buf[me - W], -(buf[me + 2 * W] + buf[me - 2 * W])));
      dz = fmadd(-30.0, cv, fmadd(16.0, buf[me + 1] + buf[me - 1], -(buf[me + 2] + buf[me - 2])));
    }
    const double lap = fmadd(a.cz, dz, fmadd(a.cy, dy, a.cx * dx));  // (x + y) + z
    if (valid) {
      const long long o = pb + own;
      if constexpr (RESID) {
        const double fv = glb<const double>(a.f)[o];
        glb<float>(a.out)[o] = float((fv - meanf) - lap);
        acc += cv;
      } else {
        glb<double>(a.out)[o] = lap;
      }
    }
#pragma unroll
    for (int m = 0; m < 2 * H; ++m) q[m] = q[m + 1];
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
  // z positions per tile: the power of two covering nz, 8 .. 64
  int tk = 8;
  while (tk < a.nz && tk < 64) tk *= 2;
  a.lgtk = 0;
  while ((1 << a.lgtk) < tk) a.lgtk++;
  const int tj = kStencilThreads / tk;
  a.tiles_j = (a.ny + tj - 1) / tj;
  a.tiles_k = (a.nz + tk - 1) / tk;
  const long long tiles = (long long)a.tiles_j * a.tiles_k;
  // split the x march into chunks until about 2048 workgroups run
  long long chunks = std::max(1LL, std::min<long long>(a.nx, (2048 + tiles - 1) / tiles));
  a.chunk = (int)((a.nx + chunks - 1) / chunks);
  a.chunks = (a.nx + a.chunk - 1) / a.chunk;
  const long long nb = tiles * a.chunks;
  if (nb > INT_MAX) return set_error(FF_ERR_UNSUPPORTED, "stencil: grid too large");
  const int H = p->sym == 3 ? 1 : 2;
  const size_t lds = 2 * (size_t)(tj + 2 * H) * (tk + 2 * H) * sizeof(double);
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
  ff_status st =
      hip_check(hipLaunchKernel(fn, dim3((unsigned)nb), dim3(kStencilThreads), argv, lds, s), "stencil launch");
  if (st != FF_OK || r_f64) return st;
  int n = (int)nb;
  const void* sargv[] = {&partial, &n, &sum_phi};
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
