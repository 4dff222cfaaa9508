// Stage kernels of the reference's other loop orders, full storage:
//
//   A  (src/impl/seq/A.c:58-81): stage l gathers w = y + sum_{i<l} hA[l][i] v_i
//      over the non-zero coefficients (block_gather_vec_interm_stage,
//      include/block-inline.h:76-92), then v_l = f(w) (block_rhs); the output
//      gathers dy, err over the non-zero weights (block_gather_vec_output,
//      :96-134).
//   E  (src/impl/seq/E.c:59-80): the same without skipping zero coefficients
//      (block_gather_interm_stage / block_gather_output, :38-72).
//   F  (src/impl/seq/F.c:61-85): stage l evaluates v_l = f(w_l) and in the same
//      pass gathers w_{l+1} = y + sum_{i<=l} hA[l+1][i] v_i
//      (block_rhs_gather_interm_stage, :165-182), with two w buffers.
//
// Every w_l, dy and err receives the same terms in the same ascending order as
// the scatter form of D (block-inline.h:253-320), so all three give D's
// results bitwise.  They differ in traffic: per DOPRI54 step A and E move
// about 54 vectors of n, F about 49, D 69 -- candidates for problems without
// a fused kernel (LV pairs, tables without a compiled pattern).  The output
// kernel also forms y_new = y + dy into the other y buffer and the error
// max-norm (A.c:85-92), and its last workgroup runs the step control.
#include "rkx_device.hpp"

#include <algorithm>

namespace rkx {
namespace {

constexpr int kGBlock = 256;
// the output kernel: 256 workgroups of 16 waves, grid-stride (few workgroups
// fold their error maxima into the step state: those atomics queue)
constexpr int kGOutThreads = 1024;
constexpr int kGOutBlocks = 256;

__device__ __forceinline__ const double *gcur(const state *st, const double *y0, const double *y1)
{
  return st->cur ? y1 : y0;
}

// v_l = f(w) on every cell (block_rhs); stage 0 reads the current y
__global__ __launch_bounds__(kGBlock) void k_g_rhs(dev_ode p, int l, state *st,
                                                   const double *y0, const double *y1,
                                                   const double *__restrict__ w,
                                                   double *__restrict__ v)
{
  if (st->done)
    return;
  const int cell = blockIdx.x * kGBlock + threadIdx.x;
  if (cell >= p.cells)
    return;
  const double *src = (l == 0) ? gcur(st, y0, y1) : w;
  double F0, F1;
  rhs_cell(p, src, cell, F0, F1);
  *reinterpret_cast<double2 *>(v + static_cast<size_t>(l) * p.n + 2 * cell) =
      make_double2(F0, F1);
}

// w = y + sum_{i<l} hA[l][i] v_i; SKIP: only the non-zero entries of row l of
// A (bit i of nzrow), the first non-zero entry forming y + a v as the first
// term does (block_gather_vec_interm_stage)
template <bool SKIP>
__global__ __launch_bounds__(kGBlock) void k_g_gather(dev_ode p, const dev_coef *__restrict__ m,
                                                      int l, unsigned nzrow, state *st,
                                                      const double *y0, const double *y1,
                                                      const double *__restrict__ v,
                                                      double *__restrict__ w)
{
  if (st->done)
    return;
  const int cell = blockIdx.x * kGBlock + threadIdx.x;
  if (cell >= p.cells)
    return;
  const size_t n = static_cast<size_t>(p.n);
  const int i = 2 * cell;
  const double2 Y = *reinterpret_cast<const double2 *>(gcur(st, y0, y1) + i);
  double a0 = Y.x, a1 = Y.y;
  for (int j = 0; j < l; ++j) {
    if (SKIP && !((nzrow >> j) & 1u))
      continue;
    const double a = m->hA[l * RKX_MAX_STAGES + j];
    const double2 V = *reinterpret_cast<const double2 *>(v + j * n + i);
    a0 += a * V.x;
    a1 += a * V.y;
  }
  *reinterpret_cast<double2 *>(w + i) = make_double2(a0, a1);
}

// F: v_l = f(w_l) and w_{l+1} = y + sum_{i<=l} hA[l+1][i] v_i in one pass
__global__ __launch_bounds__(kGBlock) void k_g_rhs_gather(dev_ode p,
                                                          const dev_coef *__restrict__ m, int l,
                                                          state *st, const double *y0,
                                                          const double *y1,
                                                          const double *__restrict__ wl,
                                                          double *__restrict__ v,
                                                          double *__restrict__ wn)
{
  if (st->done)
    return;
  const int cell = blockIdx.x * kGBlock + threadIdx.x;
  if (cell >= p.cells)
    return;
  const size_t n = static_cast<size_t>(p.n);
  const int i = 2 * cell;
  const double *y = gcur(st, y0, y1);
  double F0, F1;
  rhs_cell(p, (l == 0) ? y : wl, cell, F0, F1);
  *reinterpret_cast<double2 *>(v + l * n + i) = make_double2(F0, F1);
  const double2 Y = *reinterpret_cast<const double2 *>(y + i);
  double a0 = Y.x, a1 = Y.y;
  for (int j = 0; j <= l; ++j) {
    const double a = m->hA[(l + 1) * RKX_MAX_STAGES + j];
    double2 V;
    if (j == l)
      V = make_double2(F0, F1);
    else
      V = *reinterpret_cast<const double2 *>(v + j * n + i);
    a0 += a * V.x;
    a1 += a * V.y;
  }
  *reinterpret_cast<double2 *>(wn + i) = make_double2(a0, a1);
}

// dy = sum hb_i v_i, err = sum he_i v_i (SKIP: terms i >= 1 only where the
// weight's table entry is non-zero, bits of nzb / nze), y_new = y + dy into the
// other buffer, error max-norm, step control in the last workgroup
template <bool SKIP>
__global__ __launch_bounds__(kGOutThreads) void k_g_output(dev_ode p,
                                                           const dev_coef *__restrict__ m, int s,
                                                           unsigned nzb, unsigned nze, state *st,
                                                           double *y0, double *y1,
                                                           const double *__restrict__ v)
{
  __shared__ double red[kGOutThreads / 64];
  __shared__ unsigned int zred[kGOutThreads / 64];
  if (st->done)
    return;
  const int cur = st->cur;
  const double *y = cur ? y1 : y0;
  double *ynew = cur ? y0 : y1;
  const size_t n = static_cast<size_t>(p.n);
  double qmax = 0.0;
  unsigned int zero = 0u;
  for (int cell = blockIdx.x * kGOutThreads + threadIdx.x; cell < p.cells;
       cell += gridDim.x * kGOutThreads) {
    const int i = 2 * cell;
    const double2 V0 = *reinterpret_cast<const double2 *>(v + i);
    double e0 = m->he[0] * V0.x, e1 = m->he[0] * V0.y;
    double d0 = m->hb[0] * V0.x, d1 = m->hb[0] * V0.y;
    for (int j = 1; j < s; ++j) {
      const bool te = !SKIP || ((nze >> j) & 1u), tb = !SKIP || ((nzb >> j) & 1u);
      if (!te && !tb)
        continue;
      const double2 V = *reinterpret_cast<const double2 *>(v + j * n + i);
      if (te) {
        e0 += m->he[j] * V.x;
        e1 += m->he[j] * V.y;
      }
      if (tb) {
        d0 += m->hb[j] * V.x;
        d1 += m->hb[j] * V.y;
      }
    }
    const double2 yo = *reinterpret_cast<const double2 *>(y + i);
    const double2 yn = make_double2(yo.x + d0, yo.y + d1);
    *reinterpret_cast<double2 *>(ynew + i) = yn;
    err_term(e0, yn.x, yo.x, qmax, zero);
    err_term(e1, yn.y, yo.y, qmax, zero);
  }
  block_finish(st, qmax, zero, red, zred);
}

}  // namespace

hipError_t launch_gather_step(int form, const dev_ode &p, const dev_coef *m, int s,
                              const gather_masks &nz, state *st, double *y0, double *y1,
                              double *v, double *w0, double *w1, hipStream_t stream)
{
  const int grid = (p.cells + kGBlock - 1) / kGBlock;
  const dim3 g(grid), b(kGBlock);
  const dim3 go(std::min((p.cells + kGOutThreads - 1) / kGOutThreads, kGOutBlocks)),
      bo(kGOutThreads);
  if (form == RKX_IMPL_F) {
    double *wc = w0, *wn = w1;
    for (int l = 0; l < s - 1; ++l) {
      hipLaunchKernelGGL(k_g_rhs_gather, g, b, 0, stream, p, m, l, st, y0, y1, wc, v, wn);
      std::swap(wc, wn);
    }
    hipLaunchKernelGGL(k_g_rhs, g, b, 0, stream, p, s - 1, st, y0, y1, wc, v);
    hipLaunchKernelGGL((k_g_output<false>), go, bo, 0, stream, p,
                       m, s, nz.b, nz.e, st, y0, y1, v);
  } else {
    const bool skip = (form == RKX_IMPL_A);
    hipLaunchKernelGGL(k_g_rhs, g, b, 0, stream, p, 0, st, y0, y1, w0, v);
    for (int l = 1; l < s; ++l) {
      if (skip)
        hipLaunchKernelGGL((k_g_gather<true>), g, b, 0, stream, p, m, l, nz.rowA[l], st, y0, y1,
                           v, w0);
      else
        hipLaunchKernelGGL((k_g_gather<false>), g, b, 0, stream, p, m, l, nz.rowA[l], st, y0, y1,
                           v, w0);
      hipLaunchKernelGGL(k_g_rhs, g, b, 0, stream, p, l, st, y0, y1, w0, v);
    }
    if (skip)
      hipLaunchKernelGGL((k_g_output<true>), go, bo, 0, stream, p,
                         m, s, nz.b, nz.e, st, y0, y1, v);
    else
      hipLaunchKernelGGL((k_g_output<false>), go, bo, 0, stream,
                         p, m, s, nz.b, nz.e, st, y0, y1, v);
  }
  return hipGetLastError();
}

gather_masks make_gather_masks(const rkx_method &m)
{
  gather_masks z{};
  for (int l = 0; l < m.s; ++l)
    for (int i = 0; i < l; ++i)
      if (m.A[l][i] != 0.0)
        z.rowA[l] |= 1u << i;
  for (int i = 0; i < m.s; ++i) {
    if (m.b[i] != 0.0)
      z.b |= 1u << i;
    if (m.b[i] - m.b_hat[i] != 0.0)
      z.e |= 1u << i;
  }
  return z;
}

}  // namespace rkx
