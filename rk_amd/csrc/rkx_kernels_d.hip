// Unfused stage kernels: the GPU form of the reference's implementation D
// (src/impl/seq/D.c:58-78 with include/block-inline.h:253-320).
//
// One thread per cell (two interleaved components).  Stage i reads w_i with
// its stencil and scatters h*A[l][i]*F into w_l (l > i), dy and err, exactly
// the reference's order of accumulation; zero coefficients are not skipped
// (the reference's D does not skip them either).  The last stage writes
// y_new = y + dy into the other y buffer, so a rejected step needs no copy,
// and folds the relative error max-norm into the step state; the last
// workgroup runs the step control.  Full storage: y (2 buffers), w_1..w_{s-1},
// dy, err.
#include "rkx_device.hpp"

#include <algorithm>

namespace rkx {

namespace {

constexpr int kBlock = 256;
// the last stage runs grid-stride in 256 workgroups of 16 waves (4 per SIMD,
// enough to stream at full bandwidth): each workgroup folds its error maximum
// into the step state with atomics on one word, which queue at the end of a
// step, so few large workgroups finish sooner than many small ones
constexpr int kLastThreads = 1024;
constexpr int kLastBlocks = 256;

__device__ __forceinline__ double *ycur(state *st, double *y0, double *y1)
{
  return st->cur ? y1 : y0;
}

__global__ __launch_bounds__(kBlock) void k_d_first(dev_ode p, const dev_coef *__restrict__ m,
                                                    int s, state *st, double *y0, double *y1,
                                                    double *__restrict__ w,
                                                    double *__restrict__ dy,
                                                    double *__restrict__ err)
{
  if (st->done)
    return;
  const int cell = blockIdx.x * kBlock + threadIdx.x;
  if (cell >= p.cells)
    return;
  const double *y = ycur(st, y0, y1);
  double F0, F1;
  rhs_cell(p, y, cell, F0, F1);
  const size_t n = static_cast<size_t>(p.n);
  const int i = 2 * cell;
  const double2 Y = *reinterpret_cast<const double2 *>(y + i);
  for (int l = 1; l < s; ++l) {
    const double a = m->hA[l * RKX_MAX_STAGES + 0];
    double2 o;
    o.x = Y.x + a * F0;
    o.y = Y.y + a * F1;
    *reinterpret_cast<double2 *>(w + (l - 1) * n + i) = o;
  }
  const double hb = m->hb[0], he = m->he[0];
  *reinterpret_cast<double2 *>(dy + i) = make_double2(hb * F0, hb * F1);
  *reinterpret_cast<double2 *>(err + i) = make_double2(he * F0, he * F1);
}

__global__ __launch_bounds__(kBlock) void k_d_interm(dev_ode p, const dev_coef *__restrict__ m,
                                                     int s, int stage, state *st,
                                                     double *__restrict__ w,
                                                     double *__restrict__ dy,
                                                     double *__restrict__ err)
{
  if (st->done)
    return;
  const int cell = blockIdx.x * kBlock + threadIdx.x;
  if (cell >= p.cells)
    return;
  const size_t n = static_cast<size_t>(p.n);
  const double *wi = w + (stage - 1) * n;
  double F0, F1;
  rhs_cell(p, wi, cell, F0, F1);
  const int i = 2 * cell;
  for (int l = stage + 1; l < s; ++l) {
    const double a = m->hA[l * RKX_MAX_STAGES + stage];
    double2 *pw = reinterpret_cast<double2 *>(w + (l - 1) * n + i);
    double2 o = *pw;
    o.x += a * F0;
    o.y += a * F1;
    *pw = o;
  }
  const double hb = m->hb[stage], he = m->he[stage];
  double2 d = *reinterpret_cast<double2 *>(dy + i);
  d.x += hb * F0;
  d.y += hb * F1;
  *reinterpret_cast<double2 *>(dy + i) = d;
  double2 e = *reinterpret_cast<double2 *>(err + i);
  e.x += he * F0;
  e.y += he * F1;
  *reinterpret_cast<double2 *>(err + i) = e;
}

__global__ __launch_bounds__(kLastThreads) void k_d_last(dev_ode p,
                                                         const dev_coef *__restrict__ m, int s,
                                                         state *st, double *y0, double *y1,
                                                         double *__restrict__ w,
                                                         double *__restrict__ dy,
                                                         double *__restrict__ err)
{
  __shared__ double red[kLastThreads / 64];
  __shared__ unsigned int zred[kLastThreads / 64];
  if (st->done)
    return;                       // uniform: every block of the grid returns
  const int cur = st->cur;
  const double *y = cur ? y1 : y0;
  double *ynew = cur ? y0 : y1;
  double qmax = 0.0;
  unsigned int zero = 0u;
  // grid-stride: a bounded number of workgroups folds its maximum into the
  // step state (one atomic each), instead of one per 256 cells
  for (int cell = blockIdx.x * kLastThreads + threadIdx.x; cell < p.cells;
       cell += gridDim.x * kLastThreads) {
    const size_t n = static_cast<size_t>(p.n);
    const int stage = s - 1;
    double F0, F1;
    rhs_cell(p, w + (stage - 1) * n, cell, F0, F1);
    const int i = 2 * cell;
    const double hb = m->hb[stage], he = m->he[stage];
    double2 d = *reinterpret_cast<const double2 *>(dy + i);
    d.x += hb * F0;
    d.y += hb * F1;
    double2 e = *reinterpret_cast<const double2 *>(err + i);
    e.x += he * F0;
    e.y += he * F1;
    const double2 yo = *reinterpret_cast<const double2 *>(y + i);
    double2 yn;
    yn.x = yo.x + d.x;
    yn.y = yo.y + d.y;
    *reinterpret_cast<double2 *>(ynew + i) = yn;
    err_term(e.x, yn.x, yo.x, qmax, zero);
    err_term(e.y, yn.y, yo.y, qmax, zero);
  }
  block_finish(st, qmax, zero, red, zred);
}

__global__ __launch_bounds__(kBlock) void k_eval_rng(dev_ode p, int i0, int i1,
                                                     const double *__restrict__ y,
                                                     double *__restrict__ f)
{
  const int cell = blockIdx.x * kBlock + threadIdx.x + (i0 >> 1);
  if (cell >= p.cells || 2 * cell > i1)
    return;
  double F0, F1;
  rhs_cell(p, y, cell, F0, F1);
  if (2 * cell >= i0)
    f[2 * cell] = F0;
  if (2 * cell + 1 <= i1)
    f[2 * cell + 1] = F1;
}

}  // namespace

hipError_t launch_eval_rng(const dev_ode &p, int i, int j, const double *y, double *f,
                           hipStream_t s)
{
  const int c0 = i >> 1, c1 = j >> 1;
  const int cells = c1 - c0 + 1;
  const int grid = (cells + kBlock - 1) / kBlock;
  hipLaunchKernelGGL(k_eval_rng, dim3(grid), dim3(kBlock), 0, s, p, i, j, y, f);
  return hipGetLastError();
}

hipError_t launch_d_step(const dev_ode &p, const dev_coef *m, int s, state *st, double *y0,
                         double *y1, double *w, double *dy, double *err, hipStream_t stream)
{
  const int grid = (p.cells + kBlock - 1) / kBlock;
  hipLaunchKernelGGL(k_d_first, dim3(grid), dim3(kBlock), 0, stream, p, m, s, st, y0, y1, w,
                     dy, err);
  for (int i = 1; i < s - 1; ++i)
    hipLaunchKernelGGL(k_d_interm, dim3(grid), dim3(kBlock), 0, stream, p, m, s, i, st, w, dy,
                       err);
  const int last_grid = std::min((p.cells + kLastThreads - 1) / kLastThreads, kLastBlocks);
  hipLaunchKernelGGL(k_d_last, dim3(last_grid), dim3(kLastThreads), 0, stream, p, m, s, st, y0,
                     y1, w, dy, err);
  return hipGetLastError();
}

}  // namespace rkx
