// Device-side building blocks: right-hand sides, step control, reductions.
//
// Every floating-point expression keeps the reference's operand order and is
// compiled with -ffp-contract=off, so results equal the reference bitwise.
// A fused multiply-add is written explicitly only where the product is exact
// (a power-of-two scale: 2*y, 4*y), where fma(a, b, c) == c + a*b bitwise.
#pragma once

#include "rkx_internal.hpp"
#include "rkx_pow.hpp"

namespace rkx {

// Tables of libm_pow in constant memory (one copy per kernel translation unit).
static __constant__ pow_tables dev_pow_tables =
#include "rkx_pow_tables.inc"
    ;

// Vectors reached through an instance table are generic pointers to the
// compiler, which then emits FLAT memory instructions; those count against
// the LDS counter too, so every wait for LDS (before each barrier) would also
// wait for the global loads in flight.  Kernels that prefetch through such
// pointers view them in the global address space instead.
using gconst = const __attribute__((address_space(1))) double;
using gbyte = const __attribute__((address_space(1))) char;
using gmut = __attribute__((address_space(1))) double;
__device__ __forceinline__ gconst *as_global(const double *p) { return (gconst *) p; }
__device__ __forceinline__ gmut *as_global(double *p) { return (gmut *) p; }

// What a step kernel reads of its state at its start.  Read through the
// global address space, before any store of the kernel, the loads are
// scalar (through the generic pointer they were vector loads), cur and done
// in one.  No workgroup of a state reads it after that state's control has
// run: the control runs in the last of the state's workgroups to finish,
// after all of them have started.
struct step_view {
  int done, cur;
  double h;
  const dev_method *meth;
};
__device__ __forceinline__ step_view read_step(const state *st)
{
  static_assert(offsetof(state, done) == offsetof(state, cur) + 4, "cur, done adjacent");
  using gstate = const __attribute__((address_space(1))) state;
  using gu64 = const __attribute__((address_space(1))) unsigned long long;
  gstate *g = (gstate *) st;
  const unsigned long long cd = *(gu64 *) &g->cur;     // cur, done in one load
  return step_view{static_cast<int>(cd >> 32), static_cast<int>(cd & 0xffffffffu), g->h, g->meth};
}
using gmethod = const __attribute__((address_space(1))) dev_method;
__device__ __forceinline__ gmethod *as_global(const dev_method *m) { return (gmethod *) m; }

// Stores v to row[idx] in the lanes where keep holds, without a branch: the
// store goes through a buffer resource of the row (len doubles); the other
// lanes address past its end, and the hardware drops out-of-range buffer
// stores.  row must be wave-uniform (the resource lives in scalar registers).
// A branch around the store would split the surrounding straight-line code
// into scheduling regions the compiler cannot interleave.
using u32x2 = unsigned int __attribute__((ext_vector_type(2)));
__device__ __forceinline__ void store_row_masked(gmut *row, int len, int idx, bool keep, double v)
{
  constexpr int kGfx9RawBuffer = 0x00020000;   // dword 3 of a raw gfx9 buffer resource
  // the row base in scalar registers (it is wave-uniform; said explicitly,
  // the resource needs no per-lane loop)
  const unsigned long long a = reinterpret_cast<unsigned long long>(row);
  const unsigned int lo = __builtin_amdgcn_readfirstlane(static_cast<unsigned int>(a));
  const unsigned int hi = __builtin_amdgcn_readfirstlane(static_cast<unsigned int>(a >> 32));
  void *base = reinterpret_cast<void *>((static_cast<unsigned long long>(hi) << 32) | lo);
  const __amdgpu_buffer_rsrc_t r = __builtin_amdgcn_make_buffer_rsrc(
      base, 0, len * static_cast<int>(sizeof(double)), kGfx9RawBuffer);
  const int off = keep ? idx * static_cast<int>(sizeof(double)) : 0x7ffffff0;
  __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(u32x2, v), r, off, 0, 0);
}

// A value that is the same in every lane (a step coefficient), moved to
// scalar registers: loaded through a generic pointer it would otherwise
// occupy a vector register pair in every lane for the whole kernel.
__device__ __forceinline__ double wave_uniform(double x)
{
  const long long b = __double_as_longlong(x);
  const int lo = __builtin_amdgcn_readfirstlane(static_cast<int>(b));
  const int hi = __builtin_amdgcn_readfirstlane(static_cast<int>(b >> 32));
  return __longlong_as_double((static_cast<long long>(hi) << 32) |
                              static_cast<unsigned int>(lo));
}

// ---- BRUSS2D-MIX (src/problems/bruss2d-mix.c:66-206) ------------------------
// Laplacian numerator of one component at grid point (k, j) given its
// neighbours.  vN/vS already carry the Neumann mirror on the top/bottom rows
// (the reference's 2.0*y[i+-N2] equals y[i+N2] + y[i+N2] exactly), so only
// the column boundaries need their own operand order:
//   interior column:  ((vN + vS) + vW) + vE - 4 v
//   j == 0:           (vN + vS) + 2 vE - 4 v
//   j == N-1:         (vN + vS) + 2 vW - 4 v
__device__ __forceinline__ double bruss_lap_col(double vN, double vS, double vW, double vE,
                                                double v, int j, int N)
{
  const double t1 = vN + vS;
  double t2;
  if (j == 0)
    t2 = __builtin_fma(2.0, vE, t1);
  else if (j == N - 1)
    t2 = __builtin_fma(2.0, vW, t1);
  else
    t2 = (t1 + vW) + vE;
  return __builtin_fma(-4.0, v, t2);
}

// f_u = ((1 + u*u*v) - 4.4 u) + coef*lap_u ;  f_v = (3.4 u - u*u*v) + coef*lap_v
__device__ __forceinline__ void bruss_react(double u, double v, double lu, double lv,
                                            double coef, double &fu, double &fv)
{
  const double uuv = (u * u) * v;
  fu = ((1.0 + uuv) - 4.4 * u) + coef * lu;
  fv = (3.4 * u - uuv) + coef * lv;
}

// RHS of grid point cell = k*N + j reading a full state vector y (D form).
__device__ __forceinline__ void bruss_cell(const dev_ode &p, const double *__restrict__ y,
                                           int cell, double &fu, double &fv)
{
  const int N = p.N;
  const int N2 = N + N;
  const int k = cell / N;
  const int j = cell - k * N;
  const int i = k * N2 + 2 * j;
  const int iN = (k == 0) ? i + N2 : i - N2;
  const int iS = (k == N - 1) ? i - N2 : i + N2;
  const double2 c = *reinterpret_cast<const double2 *>(y + i);
  const double2 nN = *reinterpret_cast<const double2 *>(y + iN);
  const double2 nS = *reinterpret_cast<const double2 *>(y + iS);
  const double2 nW = (j > 0) ? *reinterpret_cast<const double2 *>(y + i - 2) : c;
  const double2 nE = (j < N - 1) ? *reinterpret_cast<const double2 *>(y + i + 2) : c;
  const double lu = bruss_lap_col(nN.x, nS.x, nW.x, nE.x, c.x, j, N);
  const double lv = bruss_lap_col(nN.y, nS.y, nW.y, nE.y, c.y, j, N);
  bruss_react(c.x, c.y, lu, lv, p.coef, fu, fv);
}

// ---- LV pair (src/problems/lv.c:69-77) ---------------------------------------
__device__ __forceinline__ void lv_pair(const dev_ode &p, double x, double y, double &fx,
                                        double &fy)
{
  fx = p.c0 * x - p.d0 * x * y;
  fy = p.d1 * x * y - p.c1 * y;
}

// ---- LV chain (periodic, d = 2; no reference counterpart) -------------------
//   x' = (c0 x - d0 x y) + D ((x_l + x_r) - 2x), likewise y.
__device__ __forceinline__ void chain_cell(const dev_ode &p, const double *__restrict__ y,
                                           int cell, double &fx, double &fy)
{
  const int l = (cell == 0) ? p.cells - 1 : cell - 1;
  const int r = (cell == p.cells - 1) ? 0 : cell + 1;
  const double2 c = *reinterpret_cast<const double2 *>(y + 2 * cell);
  const double2 a = *reinterpret_cast<const double2 *>(y + 2 * l);
  const double2 b = *reinterpret_cast<const double2 *>(y + 2 * r);
  double gx, gy;
  lv_pair(p, c.x, c.y, gx, gy);
  fx = gx + p.diff * __builtin_fma(-2.0, c.x, a.x + b.x);
  fy = gy + p.diff * __builtin_fma(-2.0, c.y, a.y + b.y);
}

__device__ __forceinline__ void rhs_cell(const dev_ode &p, const double *__restrict__ y,
                                         int cell, double &f0, double &f1)
{
  if (p.kind == RKX_BRUSS2D) {
    bruss_cell(p, y, cell, f0, f1);
  } else if (p.kind == RKX_LV) {
    const double2 c = *reinterpret_cast<const double2 *>(y + 2 * cell);
    lv_pair(p, c.x, c.y, f0, f1);
  } else {
    chain_cell(p, y, cell, f0, f1);
  }
}

// ---- error norm (include/tools.h:280-288) ------------------------------------
// Per component q = |e| / max(|y_new|, |y_old|); a zero divisor raises the
// 1.7e308 sentinel, a NaN quotient is skipped (fmax semantics).
__device__ __forceinline__ void err_term(double e, double yn, double yo, double &qmax,
                                         unsigned int &zero)
{
  const double d = fmax(fabs(yn), fabs(yo));
  if (d == 0.0)
    zero = 1u;
  else
    qmax = fmax(qmax, fabs(e) / d);
}

// ---- step control (include/tools.h:217-278, 191-198) ------------------------

// Runs in exactly one thread after every workgroup has folded its maximum into
// st->err_bits / st->zero_div.  Resets the accumulators for the next step.
// The new h uses libm_pow, the C library's pow restated bitwise, so the
// device reproduces the reference's step sizes without a host round trip.
// Every workgroup thread t of nt computes its share of the premultiplied
// coefficients of the next step from the raw table: the same products h*A,
// h*b, h*(b - b_hat) as premult, each written to both places it is read from.
__device__ inline void premult_block(const dev_method &m, int s, double h, dev_coef &c, int t,
                                     int nt)
{
  const int total = s * s + 2 * s;
  for (int idx = t; idx < total; idx += nt) {
    if (idx < s * s) {
      const int i = idx / s, j = idx - (idx / s) * s;
      if (j < i) {
        const double v = h * m.A[i * RKX_MAX_STAGES + j];
        c.hA[i * RKX_MAX_STAGES + j] = v;
        c.col[j][i - j - 1] = v;
      }
    } else if (idx < s * s + s) {
      const int i = idx - s * s;
      const double v = h * m.b[i];
      c.hb[i] = v;
      c.col[i][s - i - 1] = v;
    } else {
      const int i = idx - s * s - s;
      const double v = h * m.e[i];
      c.he[i] = v;
      c.col[i][s - i] = v;
    }
  }
}

// Runs in exactly one thread after every workgroup has folded its maximum into
// st->err_bits / st->zero_div.  Resets the accumulators for the next step.
// The new h uses libm_pow, the C library's pow restated bitwise, so the
// device reproduces the reference's step sizes without a host round trip.
// Returns whether the next step's coefficients need premultiplying (device
// control), which the caller's workgroup does with premult_block.
__device__ inline bool dev_control(state *st)
{
  const unsigned long long bits = atomicOr(&st->err_bits, 0ull);
  const unsigned int zero = atomicOr(&st->zero_div, 0u);
  const double q = __longlong_as_double(static_cast<long long>(bits));
  st->err_bits = 0ull;
  st->zero_div = 0u;
  st->blocks_done = 0u;
  st->err = zero ? fmax(1.7E308, q) : q;
  if (st->ctrl == RKX_CTRL_HOST)
    return false;
  const int attempt = st->acc + st->rej;
  if (st->trace && attempt < st->trace_cap) {
    st->trace[3 * attempt + 0] = st->t;
    st->trace[3 * attempt + 1] = st->h;
    st->trace[3 * attempt + 2] = st->err;
  }
  const bool need = ctrl_decide(*st);
  const double p = need ? libm_pow(st->pow_x, st->pow_y, dev_pow_tables) : 0.0;
  ctrl_finish(*st, need, p);
  return true;
}

// ---- reductions -----------------------------------------------------------------
__device__ __forceinline__ double wave_max(double v)
{
  for (int off = 32; off > 0; off >>= 1)
    v = fmax(v, __shfl_xor(v, off, 64));
  return v;
}

// Folds each thread's (qmax, zero) into the step accumulators and runs the
// control in the last of the `nblocks` workgroups that work on this state to
// finish (nblocks = 0: the whole grid).  Must be called by all threads of the
// block.  red must hold blockDim.x/64 doubles of LDS.  coef = false: the
// kernels of this state form h * dev_method::col themselves, so the control
// writes no premultiplied coefficients.
__device__ inline void block_finish(state *st, double qmax, unsigned int zero, double *red,
                                    unsigned int *zred, unsigned int nblocks = 0u,
                                    bool coef = true)
{
  if (nblocks == 0u)
    nblocks = gridDim.x * gridDim.y * gridDim.z;
  const int lane = threadIdx.x & 63;
  const int wid = threadIdx.x >> 6;
  const int nw = (blockDim.x + 63) >> 6;
  qmax = wave_max(qmax);
  const unsigned long long zb = __ballot(zero != 0u);
  if (lane == 0) {
    red[wid] = qmax;
    zred[wid] = zb ? 1u : 0u;
  }
  __syncthreads();
  __shared__ int is_last;
  if (threadIdx.x == 0) {
    double m = red[0];
    unsigned int z = zred[0];
    for (int w = 1; w < nw; ++w) {
      m = fmax(m, red[w]);
      z |= zred[w];
    }
    atomicMax(&st->err_bits, static_cast<unsigned long long>(__double_as_longlong(m)));
    if (z)
      atomicOr(&st->zero_div, 1u);
    __threadfence();
    const unsigned int prev = atomicAdd(&st->blocks_done, 1u);
    is_last = (prev == nblocks - 1u);
  }
  __syncthreads();
  if (is_last) {                       // uniform over the workgroup
    __shared__ int pm;
    __shared__ double hn;
    if (threadIdx.x == 0) {
      __threadfence();
      pm = dev_control(st) ? 1 : 0;
      hn = st->h;
    }
    __syncthreads();
    if (pm && coef)
      premult_block(*st->meth, st->s, hn, *st->coef, static_cast<int>(threadIdx.x),
                    static_cast<int>(blockDim.x));
  }
}

}  // namespace rkx
