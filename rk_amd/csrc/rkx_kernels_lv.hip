// Ensemble of independent Lotka-Volterra pairs: one thread integrates one
// 2-component system x' = c0 x - d0 x y, y' = d1 x y - c1 y (src/problems/lv.c
// with per-system constants) over [t0, te] with its own step size and
// accept/reject.  The arithmetic of every attempted step is that of the
// reference's seq D (block-inline.h:253-320, tools.h:191-288) on a system of
// size 2, so each system's result equals a plain single solve of that system
// (identical accepted/rejected sequences and bitwise the same y).  The step
// control runs in the thread, with libm_pow: the C library's pow restated, so
// every new h equals the reference's (thousands of systems step at once; no
// host round trip).
#include "rkx_device.hpp"

namespace rkx {
namespace {

constexpr int kLvBlock = 256;
constexpr int kLvMaxAttempts = 1 << 24;   // every thread stops after this many

// method table in constant memory: A row-major, b, e = b - b_hat
struct lv_method {
  double A[RKX_MAX_STAGES * RKX_MAX_STAGES];
  double b[RKX_MAX_STAGES];
  double e[RKX_MAX_STAGES];
  int s, ord;
};

template <int S>
__global__ __launch_bounds__(kLvBlock) void k_lv_ensemble(int count, lv_method m,
                                                          const double *__restrict__ prm,
                                                          const double *__restrict__ y0,
                                                          double *__restrict__ yout, double t0,
                                                          double te, double tol, double h0,
                                                          int step_limit, int *__restrict__ acc_out,
                                                          int *__restrict__ rej_out)
{
  const int kt = blockIdx.x * kLvBlock + threadIdx.x;
  if (kt >= count)
    return;
  const size_t k = static_cast<size_t>(kt);
  dev_ode p{};
  p.c0 = prm[4 * k + 0];
  p.d0 = prm[4 * k + 1];
  p.d1 = prm[4 * k + 2];
  p.c1 = prm[4 * k + 3];
  double y[2] = {y0[2 * k], y0[2 * k + 1]};
  state st{};
  st.t = t0;
  st.te = te;
  st.tol = tol;
  st.h = h0;
  st.ord = m.ord;
  st.step_limit = step_limit;
  st.done = !(t0 < te);
  int attempts = 0;
  while (!st.done && attempts < kLvMaxAttempts) {
    const double h = st.h;
    double w[S][2], dy[2], er[2];
    // first stage
    double f0, f1;
    lv_pair(p, y[0], y[1], f0, f1);
#pragma unroll
    for (int l = 1; l < S; ++l) {
      const double a = h * m.A[l * RKX_MAX_STAGES];
      w[l][0] = y[0] + a * f0;
      w[l][1] = y[1] + a * f1;
    }
    dy[0] = (h * m.b[0]) * f0;
    dy[1] = (h * m.b[0]) * f1;
    er[0] = (h * m.e[0]) * f0;
    er[1] = (h * m.e[0]) * f1;
    // intermediate and last stages
#pragma unroll
    for (int i = 1; i < S; ++i) {
      lv_pair(p, w[i][0], w[i][1], f0, f1);
#pragma unroll
      for (int l = i + 1; l < S; ++l) {
        const double a = h * m.A[l * RKX_MAX_STAGES + i];
        w[l][0] += a * f0;
        w[l][1] += a * f1;
      }
      const double hb = h * m.b[i], he = h * m.e[i];
      dy[0] += hb * f0;
      dy[1] += hb * f1;
      er[0] += he * f0;
      er[1] += he * f1;
    }
    double yn[2] = {y[0] + dy[0], y[1] + dy[1]};
    double q = 0.0;
    unsigned int zero = 0u;
    err_term(er[0], yn[0], y[0], q, zero);
    err_term(er[1], yn[1], y[1], q, zero);
    st.err = zero ? fmax(1.7E308, q) : q;
    const int cur_before = st.cur;
    const bool need = ctrl_decide(st);
    if (st.cur != cur_before) {          // accepted: y <- y_new
      y[0] = yn[0];
      y[1] = yn[1];
    }
    ctrl_finish(st, need, need ? libm_pow(st.pow_x, st.pow_y, dev_pow_tables) : 0.0);
    ++attempts;
  }
  yout[2 * k] = y[0];
  yout[2 * k + 1] = y[1];
  acc_out[k] = st.acc;
  rej_out[k] = st.rej;
}

}  // namespace
}  // namespace rkx

using namespace rkx;

extern "C" int rkx_lv_ensemble(int count, const double *params_dev, const double *y0_dev,
                               double *y_dev, double t0, double te, const rkx_method *m,
                               double tol, double h0, int step_limit, int *acc_dev, int *rej_dev,
                               void *stream)
{
  if (count < 1 || !params_dev || !y0_dev || !y_dev || !acc_dev || !rej_dev || h0 <= 0.0)
    return fail(RKX_EINVAL, "rkx_lv_ensemble: bad argument (h0 must be > 0)");
  if (count > (1 << 30))
    return fail(RKX_EINVAL, "rkx_lv_ensemble: at most 2^30 systems per call");
  if (int rc = check_method(m))
    return rc;
  lv_method lm{};
  for (int i = 0; i < m->s; ++i) {
    for (int j = 0; j < i; ++j)
      lm.A[i * RKX_MAX_STAGES + j] = m->A[i][j];
    lm.b[i] = m->b[i];
    lm.e[i] = m->b[i] - m->b_hat[i];
  }
  lm.s = m->s;
  lm.ord = m->ord;
  hipStream_t s = static_cast<hipStream_t>(stream);
  if (!s)
    if (int rc = default_stream(&s))
      return rc;
  const int grid = (count + kLvBlock - 1) / kLvBlock;
#define RKX_LV_CASE(S_)                                                                      \
  case S_:                                                                                   \
    hipLaunchKernelGGL(k_lv_ensemble<S_>, dim3(grid), dim3(kLvBlock), 0, s, count, lm,      \
                       params_dev, y0_dev, y_dev, t0, te, tol, h0, step_limit, acc_dev,      \
                       rej_dev);                                                             \
    break;
  switch (m->s) {
    RKX_LV_CASE(2)
    RKX_LV_CASE(3)
    RKX_LV_CASE(4)
    RKX_LV_CASE(7)
    RKX_LV_CASE(8)
    RKX_LV_CASE(13)
  default:
    return fail(RKX_EINVAL, "rkx_lv_ensemble: no kernel for this stage count");
  }
#undef RKX_LV_CASE
  RKX_HIP(hipGetLastError());
  RKX_HIP(hipStreamSynchronize(s));
  return RKX_OK;
}
