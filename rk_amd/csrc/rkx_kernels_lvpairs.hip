// Stage-fused step for independent Lotka-Volterra pairs (src/problems/lv.c:
// ode_eval_comp of component i reads only its own pair, access distance 1).
//
// One thread per pair (grid-stride over a bounded number of workgroups)
// keeps every stage accumulator of the scatter form of D in registers:
// w_1..w_{s-1}, dy and err of both components receive the same products in
// the same order as the D kernels (block-inline.h:253-320; no zero skipping,
// as in seq D), so the step equals D's bitwise.  It reads y once and writes
// y_new once into the other y buffer: 16 bytes in and 16 out per pair per
// step, the low-storage form (two vectors of n) of a problem whose stencil
// needs no halo.  The error max-norm is folded per workgroup and the last
// workgroup runs the step control (block_finish).
#include "rkx_device.hpp"

#include <algorithm>

namespace rkx {
namespace {

// 256 workgroups of 16 waves: each workgroup folds its error maximum into the
// step state with atomics on one word, which queue at the end of a step, so
// fewer and larger workgroups finish faster (4e6 pairs, DOPRI54: 41 us per
// step against 90 us with 2048 workgroups of 4 waves)
constexpr int kPBlock = 1024;
constexpr int kPMaxBlocks = 256;

template <int S>
__global__ __launch_bounds__(kPBlock) void k_lv_step(dev_ode p, const dev_coef *__restrict__ m,
                                                    state *st, double *y0, double *y1)
{
  __shared__ double red[kPBlock / 64];
  __shared__ unsigned int zred[kPBlock / 64];
  const step_view sv = read_step(st);
  if (sv.done)
    return;                       // uniform: every block of the grid returns
  const int cur = sv.cur;
  const double *y = cur ? y1 : y0;
  double *ynew = cur ? y0 : y1;
  double qmax = 0.0;
  unsigned int zero = 0u;
  for (int pair = blockIdx.x * kPBlock + threadIdx.x; pair < p.cells;
       pair += gridDim.x * kPBlock) {
    const double2 Y = *reinterpret_cast<const double2 *>(y + 2 * pair);
    double wx[S], wy[S];          // [l]: w_l for l = 1..S-1
    double fx, fy;
    lv_pair(p, Y.x, Y.y, fx, fy);
#pragma unroll
    for (int l = 1; l < S; ++l) {
      const double a = m->hA[l * RKX_MAX_STAGES + 0];
      wx[l] = Y.x + a * fx;
      wy[l] = Y.y + a * fy;
    }
    double dx = m->hb[0] * fx, dyv = m->hb[0] * fy;
    double ex = m->he[0] * fx, ey = m->he[0] * fy;
#pragma unroll
    for (int i = 1; i < S; ++i) {
      lv_pair(p, wx[i], wy[i], fx, fy);
#pragma unroll
      for (int l = i + 1; l < S; ++l) {
        const double a = m->hA[l * RKX_MAX_STAGES + i];
        wx[l] += a * fx;
        wy[l] += a * fy;
      }
      dx += m->hb[i] * fx;
      dyv += m->hb[i] * fy;
      ex += m->he[i] * fx;
      ey += m->he[i] * fy;
    }
    const double2 yn = make_double2(Y.x + dx, Y.y + dyv);
    *reinterpret_cast<double2 *>(ynew + 2 * pair) = yn;
    err_term(ex, yn.x, Y.x, qmax, zero);
    err_term(ey, yn.y, Y.y, qmax, zero);
  }
  block_finish(st, qmax, zero, red, zred);
}

template <int S>
hipError_t launch_s(const dev_ode &p, const dev_coef *m, state *st, double *y0, double *y1,
                    hipStream_t stream)
{
  const int grid = std::max(1, std::min((p.cells + kPBlock - 1) / kPBlock, kPMaxBlocks));
  hipLaunchKernelGGL((k_lv_step<S>), dim3(grid), dim3(kPBlock), 0, stream, p, m, st, y0, y1);
  return hipGetLastError();
}

}  // namespace

bool pairs_supported(const dev_ode &p, const rkx_method &m)
{
  if (p.kind != RKX_LV)
    return false;
  switch (m.s) {
  case 2: case 3: case 4: case 7: case 8: case 13: return true;
  default: return false;
  }
}

hipError_t launch_lv_step(const dev_ode &p, const dev_coef *m, int s, state *st, double *y0,
                          double *y1, hipStream_t stream)
{
  switch (s) {
  case 2: return launch_s<2>(p, m, st, y0, y1, stream);
  case 3: return launch_s<3>(p, m, st, y0, y1, stream);
  case 4: return launch_s<4>(p, m, st, y0, y1, stream);
  case 7: return launch_s<7>(p, m, st, y0, y1, stream);
  case 8: return launch_s<8>(p, m, st, y0, y1, stream);
  case 13: return launch_s<13>(p, m, st, y0, y1, stream);
  default: return hipErrorInvalidValue;
  }
}

}  // namespace rkx
