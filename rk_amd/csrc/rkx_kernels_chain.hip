// Stage-fused step of the diffusive periodic LV chain (cells c = 0..M-1, the
// pair (x_c, y_c) at components 2c, 2c+1, access distance 2):
//   x' = (c0 x - d0 x y) + D ((x_{c-1} + x_{c+1}) - 2x),
//   y' = (d1 x y - c1 y) + D ((y_{c-1} + y_{c+1}) - 2y).
// No reference counterpart; the scheme is the one-dimensional analogue of the
// BRUSS2D sweep.  A workgroup owns a segment of T = 256*C cells (C cells per
// lane, both components), evaluates all s stages of the scatter form
// (block-inline.h:253-320) with every accumulator in registers, exchanging
// the segment-internal edge cells of each lane through LDS once per stage.
// The segment overlaps its neighbours by s cells on each side (each stage's
// valid region shrinks by one cell), so a segment writes T - 2s cells of
// y_new.  Periodic ghost cells are loaded by index wrap.  Per step the kernel
// reads y once (plus the halo) and writes y_new once: two vectors resident.
// An instance gets at most kChainWG workgroups, each looping over segments,
// so the per-workgroup fold of the error maximum into the instance's state
// (atomics on one cache line) stays short for long chains.
#include "rkx_patterns.hpp"

namespace rkx {
namespace {

constexpr int kCB = 256;     // lanes per workgroup
constexpr int kCPT = 4;      // cells per lane
constexpr int kT = kCB * kCPT;
// workgroups per instance and step: each folds its error maximum into the
// instance's state with atomics on one word, which queue at the end of the
// step (5e5 cells: 20.5 us per step with 256, 28.6 us with 512)
constexpr int kChainWG = 256;

__device__ __forceinline__ int wrap(long long c, int M)
{
  long long m = c % M;
  return static_cast<int>(m < 0 ? m + M : m);
}

template <class P>
__global__ __launch_bounds__(kCB) void k_chain(const chain_inst *__restrict__ insts,
                                               const int2 *__restrict__ tiles, int ntiles)
{
  constexpr int s = P::s;
  constexpr int C = kCPT;
  __shared__ double2 edgeL[2][kCB], edgeR[2][kCB];
  __shared__ double red[kCB / 64];
  __shared__ unsigned int zred[kCB / 64];

  const int T = ntiles;
  const int b = blockIdx.x;
  const int x8 = b & 7, q = T >> 3, rm = T & 7;
  const int tile = x8 * q + min(x8, rm) + (b >> 3);
  const int2 td = tiles ? tiles[tile] : make_int2(0, tile);
  const chain_inst &I = insts[td.x];
  double *const y0p = I.y0, *const y1p = I.y1;
  const int M = I.cells;
  state *st = I.st;
  const step_view sv = read_step(st);
  if (sv.done)
    return;
  const int cur = sv.cur;
  const double2 *__restrict__ y = reinterpret_cast<const double2 *>(cur ? y1p : y0p);
  double2 *__restrict__ yn = reinterpret_cast<double2 *>(cur ? y0p : y1p);
  dev_ode p{};
  p.c0 = I.c0;
  p.d0 = I.d0;
  p.d1 = I.d1;
  p.c1 = I.c1;
  const double D = I.diff;
  gmethod *__restrict__ mt = as_global(sv.meth);
  const double h = sv.h;                         // this step's h (uniform)

  const int t = threadIdx.x;

  // h * col[i]: the premultiplied coefficients of stage i (premult's products)
  double kv[s][s + 1];
  static_for<s>([&](auto it) {
    constexpr int i = decltype(it)::value;
    static_for<s>([&](auto lt) {
      constexpr int l = decltype(lt)::value;
      if constexpr (l > i)
        kv[i][l - i - 1] = nzA<P>(l, i) ? wave_uniform(h * mt->col[i][l - i - 1]) : 0.0;
    });
    if constexpr (P::fsal && i < s - 1)
      kv[i][s - i - 1] = kv[i][s - i - 2];
    else
      kv[i][s - i - 1] = ((P::b >> i) & 1u) ? wave_uniform(h * mt->col[i][s - i - 1]) : 0.0;
    kv[i][s - i] = ((P::e >> i) & 1u) ? wave_uniform(h * mt->col[i][s - i]) : 0.0;
  });

  double qmax = 0.0;
  unsigned int zero = 0u;
  for (int seg = td.y; seg < I.ntiles; seg += I.nwg) {
  const long long out0 = static_cast<long long>(seg) * I.tile_out;    // first output cell
  const long long c0 = out0 - s + static_cast<long long>(t) * C;      // this lane's first cell
  double2 yv[C];
  double2 acc[C][s + 1];                 // [0] dy, [1..s-1] w_l, [s] err
  static_for<C>([&](auto ct) {
    constexpr int c = decltype(ct)::value;
    yv[c] = y[wrap(c0 + c, M)];
  });

  static_for<s>([&](auto itg) {
    constexpr int i = decltype(itg)::value;
    double2 cv[C];
    static_for<C>([&](auto ct) {
      constexpr int c = decltype(ct)::value;
      cv[c] = (i == 0) ? yv[c] : acc[c][i];
    });
    const int buf = i & 1;
    edgeL[buf][t] = cv[0];
    edgeR[buf][t] = cv[C - 1];
    __syncthreads();
    // the lane's outer neighbours (the segment's own edges read themselves:
    // those cells are halo and never reach the output)
    const double2 lft = edgeR[buf][t > 0 ? t - 1 : 0];
    const double2 rgt = edgeL[buf][t < kCB - 1 ? t + 1 : kCB - 1];
    static_for<C>([&](auto ct) {
      constexpr int c = decltype(ct)::value;
      const double2 l = (c == 0) ? lft : cv[c - 1];
      const double2 r = (c == C - 1) ? rgt : cv[c + 1];
      const double2 v = cv[c];
      double gx, gy;
      lv_pair(p, v.x, v.y, gx, gy);
      const double fx = gx + D * __builtin_fma(-2.0, v.x, l.x + r.x);
      const double fy = gy + D * __builtin_fma(-2.0, v.y, l.y + r.y);
      const double *kc = kv[i];
      if constexpr (i == 0) {
        static_for<s>([&](auto ltg) {
          constexpr int lq = decltype(ltg)::value;
          if constexpr (lq >= 1)
            acc[c][lq] = nzA<P>(lq, 0) ? make_double2(v.x + kc[lq - 1] * fx, v.y + kc[lq - 1] * fy)
                                       : v;
        });
        acc[c][0] = make_double2(kc[s - 1] * fx, kc[s - 1] * fy);
        acc[c][s] = make_double2(kc[s] * fx, kc[s] * fy);
      } else {
        static_for<s>([&](auto ltg) {
          constexpr int lq = decltype(ltg)::value;
          if constexpr (lq > i && nzA<P>(lq, i)) {
            acc[c][lq].x += kc[lq - i - 1] * fx;
            acc[c][lq].y += kc[lq - i - 1] * fy;
          }
        });
        if constexpr ((P::b >> i) & 1u) {
          acc[c][0].x += kc[s - i - 1] * fx;
          acc[c][0].y += kc[s - i - 1] * fy;
        }
        if constexpr ((P::e >> i) & 1u) {
          acc[c][s].x += kc[s - i] * fx;
          acc[c][s].y += kc[s - i] * fy;
        }
      }
      if constexpr (i == s - 1) {
        const long long cell = c0 + c;
        const int local = t * C + c;
        const bool out = (local >= s) && (local < kT - s) && (cell < M);
        double2 ynw;
        ynw.x = yv[c].x + acc[c][0].x;
        ynw.y = yv[c].y + acc[c][0].y;
        if (out) {
          yn[cell] = ynw;
          err_term(acc[c][s].x, ynw.x, yv[c].x, qmax, zero);
          err_term(acc[c][s].y, ynw.y, yv[c].y, qmax, zero);
        }
      }
    });
  });
  __syncthreads();                       // the next segment reuses the edge buffers
  }
  block_finish(st, qmax, zero, red, zred, static_cast<unsigned int>(I.nwg), false);
}

}  // namespace

bool chain_supported(const dev_ode &p, const rkx_method &m)
{
  const int pat = pattern_of(m);
  return p.kind == RKX_LV_CHAIN && pat >= 1 && pat <= 6 && kT > 2 * m.s;
}

chain_inst make_chain_inst(const dev_ode &p, const rkx_method &m, state *st, double *y0,
                           double *y1)
{
  chain_inst I{};
  I.cells = p.cells;
  I.tile_out = kT - 2 * m.s;
  I.ntiles = (p.cells + I.tile_out - 1) / I.tile_out;
  I.nwg = I.ntiles < kChainWG ? I.ntiles : kChainWG;
  I.c0 = p.c0;
  I.d0 = p.d0;
  I.d1 = p.d1;
  I.c1 = p.c1;
  I.diff = p.diff;
  I.y0 = y0;
  I.y1 = y1;
  I.st = st;
  return I;
}

hipError_t launch_chain(int pattern, const chain_inst *insts, const int2 *tiles, int ntiles,
                        hipStream_t stream)
{
  if (ntiles <= 0)
    return hipSuccess;
  return dispatch_pattern(pattern, [&](auto pat) {
    using P = decltype(pat);
    hipLaunchKernelGGL(k_chain<P>, dim3(ntiles), dim3(kCB), 0, stream, insts, tiles, ntiles);
    return hipGetLastError();
  });
}

}  // namespace rkx
