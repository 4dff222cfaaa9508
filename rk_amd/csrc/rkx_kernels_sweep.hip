// Stage-fused, temporally blocked BRUSS2D step: the GPU form of the
// reference's pipelined schedules PipeD / PipeDls (src/impl/seq/PipeD.c:76-124,
// src/impl/seq/PipeDls.c:81-129) with block size B = one grid row, which is
// the access distance d = 2N (bruss2d-mix.c:718-721).
//
// Work decomposition.  A workgroup of 64*W lanes owns a strip of grid columns
// (one lane per grid point, u and v interleaved) and a band of R rows.  It
// sweeps the band's rows in the block-diagonal order of PipeD: at row
// iteration r, stage i is applied to row r - i, so when stage i needs w_i on
// rows k-1, k, k+1, row k+1 has just been completed by stage i-1 in the same
// iteration.  All per-row accumulators (w_{i+1..s-1}, dy, err of the scatter
// form, block-inline.h:253-320) stay in registers in a ring of s+1 row slots;
// the ring is fully unrolled so slots are static registers.
//
// Neighbours.  Vertical neighbours are the lane's own registers; horizontal
// ones are exchanged through LDS once per row iteration (double buffered, one
// barrier).  Lanes at the strip edges and rows at the band edges are halo:
// each stage's valid region shrinks by one grid point per stage, so a strip
// produces 64*W - 2s output columns and a band recomputes s-1 rows on each
// side (the stage-fused temporal blocking).
//
// Boundaries.  Rows above 0 / below N-1 are ghost rows loaded by reflection;
// by symmetry they hold the mirrored values bitwise, which reproduces the
// reference's Neumann rows (2.0*y[i+-N2] == y + y).  The column boundary has
// its own operand order in the reference ((vN+vS) + 2 vE instead of
// ((vN+vS) + vW) + vE): lanes on column 0 read W from a zero slot and E from
// a slot holding 2*w of column 1 (and symmetrically on column N-1), so the
// common expression gives the reference's value bitwise.
//
// Traffic.  Per step each workgroup reads its strip of y over R + 2s rows and
// writes R rows of y_new; the step's y_new goes to the other y buffer, so a
// rejected step costs no copy.  Resident state is two vectors of n doubles.
#include "rkx_device.hpp"

namespace rkx {

namespace {

// Zero patterns of the compiled methods (bit i of row l: A[l][i] != 0).
struct pat_dopri54 {
  static constexpr int s = 7;
  static constexpr unsigned rowA[7] = {0x0, 0x1, 0x3, 0x7, 0xF, 0x1F, 0x3D};
  static constexpr unsigned b = 0x3D, e = 0x7D;
};
struct pat_verner65 {
  static constexpr int s = 8;
  static constexpr unsigned rowA[8] = {0x0, 0x1, 0x3, 0x7, 0xF, 0x1F, 0x1F, 0x5F};
  static constexpr unsigned b = 0x3D, e = 0xFD;
};
struct pat_rkf23 {
  static constexpr int s = 3;
  static constexpr unsigned rowA[3] = {0x0, 0x1, 0x3};
  static constexpr unsigned b = 0x3, e = 0x7;
};
struct pat_heun2 {
  static constexpr int s = 2;
  static constexpr unsigned rowA[2] = {0x0, 0x1};
  static constexpr unsigned b = 0x3, e = 0x3;
};
struct pat_rk4 {
  static constexpr int s = 4;
  static constexpr unsigned rowA[4] = {0x0, 0x1, 0x2, 0x4};
  static constexpr unsigned b = 0xF, e = 0xF;
};
struct pat_ssprk3 {
  static constexpr int s = 3;
  static constexpr unsigned rowA[3] = {0x0, 0x1, 0x3};
  static constexpr unsigned b = 0x7, e = 0x7;
};

template <class P>
__device__ __forceinline__ constexpr bool nzA(int l, int i)
{
  return (P::rowA[l] >> i) & 1u;
}

__device__ __forceinline__ int reflect(int k, int N)
{
  // even periodic extension about rows 0 and N-1
  if (N == 1)
    return 0;
  const int period = 2 * (N - 1);
  int m = k % period;
  if (m < 0)
    m += period;
  return (m < N) ? m : period - m;
}

__device__ __forceinline__ double2 ld2(const double *p)
{
  return *reinterpret_cast<const double2 *>(p);
}

template <class P, int WAVES>
struct sweep {
  static constexpr int s = P::s;
  static constexpr int L = 64 * WAVES;     // lanes = columns per strip
  static constexpr int NS = s + 1;         // row slots in the register ring
  static constexpr int LP = L + 3;         // + AUX_L, AUX_R, ZERO slots
  static constexpr int AUXL = L, AUXR = L + 1, ZERO = L + 2;
};

template <class P, int WAVES>
__global__ __launch_bounds__(64 * WAVES) void k_sweep(dev_ode p,
                                                      const dev_coef *__restrict__ cf,
                                                      state *st, double *y0, double *y1,
                                                      int R, int OC, int nstrips, int nbands)
{
  using SW = sweep<P, WAVES>;
  constexpr int s = SW::s, L = SW::L, NS = SW::NS, LP = SW::LP;
  __shared__ double2 xch[2][s][LP];
  __shared__ double red[WAVES];
  __shared__ unsigned int zred[WAVES];

  if (st->done | st->pow_wait)
    return;                                   // uniform over the whole grid
  const int cur = st->cur;
  const double *__restrict__ y = cur ? y1 : y0;
  double *__restrict__ yn = cur ? y0 : y1;
  const int N = p.N;
  const int N2 = N + N;
  const double coef = p.coef;

  // XCD-aware tile order: consecutive workgroups are dispatched round-robin
  // over the 8 XCDs; give each XCD a contiguous run of tiles (adjacent bands
  // of a strip share halo rows in that XCD's L2).
  const int T = gridDim.x;
  const int b = blockIdx.x;
  const int x = b & 7, q = T >> 3, rm = T & 7;
  const int tile = x * q + min(x, rm) + (b >> 3);
  const int strip = tile / nbands;
  const int band = tile - strip * nbands;

  const int lane = threadIdx.x;
  const int col = strip * OC - s + lane;
  const int ccol = reflect(col, N);          // in-bounds column for loads
  const bool out_col = (lane >= s) && (lane < L - s) && (col >= 0) && (col < N) &&
                       (col < (strip + 1) * OC);
  // LDS read slots for the W and E neighbours (see the file comment)
  const int slotW = (col == 0) ? SW::ZERO : ((col == N - 1) ? SW::AUXR : max(lane - 1, 0));
  const int slotE = (col == 0) ? SW::AUXL : ((col == N - 1) ? SW::ZERO : min(lane + 1, L - 1));
  const bool wr_auxl = (col == 1);           // 2*w of column 1 feeds column 0
  const bool wr_auxr = (col == N - 2);       // 2*w of column N-2 feeds column N-1
  if (lane == 0) {
    for (int bb = 0; bb < 2; ++bb)
      for (int i = 0; i < s; ++i)
        xch[bb][i][SW::ZERO] = make_double2(0.0, 0.0);
  }

  const int r0 = band * R;
  const int r1 = min(r0 + R, N);
  const int it0 = r0 - (s - 1);               // first row iteration
  const int it1 = r1 + s - 1;                 // one past the last

  // uniform premultiplied coefficients
  double hA[s][s];
  double hb[s], he[s];
#pragma unroll
  for (int l = 0; l < s; ++l) {
#pragma unroll
    for (int i = 0; i < s; ++i)
      hA[l][i] = (i < l && nzA<P>(l, i)) ? cf->hA[l * RKX_MAX_STAGES + i] : 0.0;
    hb[l] = ((P::b >> l) & 1u) ? cf->hb[l] : 0.0;
    he[l] = ((P::e >> l) & 1u) ? cf->he[l] : 0.0;
  }

  // register ring: slot -> [0] dy, [1..s-1] w_l, [s] err
  double2 acc[NS][s + 1];
  double2 yr[NS];                            // y rows, ring by row
  constexpr int PF = (s - 2 < 2) ? s - 2 : 2; // y rows loaded ahead of stage 0 (< NS-2)

  auto yrow = [&](int k) -> const double * {
    return y + static_cast<size_t>(reflect(k, N)) * N2 + 2 * ccol;
  };
  // rows it0-1 .. it0+PF are loaded before the loop
#pragma unroll
  for (int d = -1; d <= PF; ++d)
    yr[(d + NS) % NS] = ld2(yrow(it0 + d));

  double qmax = 0.0;
  unsigned int zero = 0u;

  for (int rb = it0; rb < it1; rb += NS) {
#pragma unroll
    for (int U = 0; U < NS; ++U) {
      const int rr = rb + U;
      if (rr >= it1)
        break;
      // prefetch y for stage 0 of a later row
      yr[(U + 1 + PF) % NS] = ld2(yrow(rr + 1 + PF));
      // y of the row finishing this iteration (stage s-1), for y_new and the norm
      const int kl = rr - (s - 1);
      const bool last_on = (rr >= it0 + 2 * (s - 1));
      double2 yold = make_double2(0.0, 0.0);
      if (last_on)
        yold = ld2(y + static_cast<size_t>(kl) * N2 + 2 * ccol);

      // publish the centre value of every active stage
      const int buf = (rr - it0) & 1;
#pragma unroll
      for (int i = 0; i < s; ++i) {
        if (rr >= it0 + 2 * i) {
          const double2 c = (i == 0) ? yr[U % NS] : acc[(U - i + NS) % NS][i];
          xch[buf][i][lane] = c;
          if (wr_auxl)
            xch[buf][i][SW::AUXL] = make_double2(c.x + c.x, c.y + c.y);
          if (wr_auxr)
            xch[buf][i][SW::AUXR] = make_double2(c.x + c.x, c.y + c.y);
        }
      }
      __syncthreads();

#pragma unroll
      for (int i = 0; i < s; ++i) {
        if (!(rr >= it0 + 2 * i))
          continue;
        const int sl = (U - i + NS) % NS;         // slot of row rr - i
        const int sn = (U - i - 1 + 2 * NS) % NS; // row rr - i - 1
        const int ss = (U - i + 1 + NS) % NS;     // row rr - i + 1
        double2 c, vn, vs;
        if (i == 0) {
          c = yr[U % NS];
          vn = yr[(U - 1 + NS) % NS];
          vs = yr[(U + 1) % NS];
        } else {
          c = acc[sl][i];
          vn = acc[sn][i];
          vs = acc[ss][i];
        }
        const double2 vw = xch[buf][i][slotW];
        const double2 ve = xch[buf][i][slotE];
        // Laplacian numerators: ((vN + vS) + vW) + vE - 4c
        const double lu = __builtin_fma(-4.0, c.x, ((vn.x + vs.x) + vw.x) + ve.x);
        const double lv = __builtin_fma(-4.0, c.y, ((vn.y + vs.y) + vw.y) + ve.y);
        double fu, fv;
        bruss_react(c.x, c.y, lu, lv, coef, fu, fv);

        if (i == 0) {
          // first stage: w_l = y + hA[l][0] F, dy = hb0 F, err = he0 F
#pragma unroll
          for (int l = 1; l < s; ++l) {
            if (nzA<P>(l, 0))
              acc[sl][l] = make_double2(c.x + hA[l][0] * fu, c.y + hA[l][0] * fv);
            else
              acc[sl][l] = c;
          }
          acc[sl][0] = make_double2(hb[0] * fu, hb[0] * fv);
          acc[sl][s] = make_double2(he[0] * fu, he[0] * fv);
        } else {
#pragma unroll
          for (int l = i + 1; l < s; ++l)
            if (nzA<P>(l, i)) {
              acc[sl][l].x += hA[l][i] * fu;
              acc[sl][l].y += hA[l][i] * fv;
            }
          if ((P::b >> i) & 1u) {
            acc[sl][0].x += hb[i] * fu;
            acc[sl][0].y += hb[i] * fv;
          }
          if ((P::e >> i) & 1u) {
            acc[sl][s].x += he[i] * fu;
            acc[sl][s].y += he[i] * fv;
          }
        }
        if (i == s - 1) {
          // y_new = y + dy; dy now holds y_old (block-inline.h:298-320)
          double2 ynw;
          ynw.x = yold.x + acc[sl][0].x;
          ynw.y = yold.y + acc[sl][0].y;
          if (out_col) {
            *reinterpret_cast<double2 *>(yn + static_cast<size_t>(kl) * N2 + 2 * col) = ynw;
            err_term(acc[sl][s].x, ynw.x, yold.x, qmax, zero);
            err_term(acc[sl][s].y, ynw.y, yold.y, qmax, zero);
          }
        }
      }
    }
  }
  block_finish(st, qmax, zero, red, zred);
}

template <class P, int WAVES>
hipError_t launch_one(const dev_ode &p, const dev_coef *cf, state *st, double *y0, double *y1,
                      int R, hipStream_t stream)
{
  constexpr int s = P::s;
  constexpr int L = 64 * WAVES;
  const int N = p.N;
  const int oc_max = L - 2 * s;
  const int nstrips = (N + oc_max - 1) / oc_max;
  const int OC = (N + nstrips - 1) / nstrips;
  const int nbands = (N + R - 1) / R;
  const int grid = nstrips * nbands;
  hipLaunchKernelGGL((k_sweep<P, WAVES>), dim3(grid), dim3(L), 0, stream, p, cf, st, y0, y1,
                     R, OC, nstrips, nbands);
  return hipGetLastError();
}

int pattern_of(const rkx_method &m)
{
  // 1 DOPRI54, 2 VERNER65, 3 RKF23, 4 HEUN2, 5 RK4, 6 SSPRK3; 0 none
  auto rowmask = [&](int l) {
    unsigned r = 0;
    for (int i = 0; i < l; ++i)
      if (m.A[l][i] != 0.0)
        r |= 1u << i;
    return r;
  };
  auto vmask = [&](const double *v, const double *w) {
    unsigned r = 0;
    for (int i = 0; i < m.s; ++i)
      if ((w ? v[i] - w[i] : v[i]) != 0.0)
        r |= 1u << i;
    return r;
  };
  auto match = [&](int s, const unsigned *rows, unsigned b, unsigned e) {
    if (m.s != s)
      return false;
    for (int l = 0; l < s; ++l)
      if (rowmask(l) != rows[l])
        return false;
    return vmask(m.b, nullptr) == b && vmask(m.b, m.b_hat) == e;
  };
  if (match(7, pat_dopri54::rowA, pat_dopri54::b, pat_dopri54::e)) return 1;
  if (match(8, pat_verner65::rowA, pat_verner65::b, pat_verner65::e)) return 2;
  if (match(3, pat_rkf23::rowA, pat_rkf23::b, pat_rkf23::e)) return 3;
  if (match(2, pat_heun2::rowA, pat_heun2::b, pat_heun2::e)) return 4;
  if (match(4, pat_rk4::rowA, pat_rk4::b, pat_rk4::e)) return 5;
  if (match(3, pat_ssprk3::rowA, pat_ssprk3::b, pat_ssprk3::e)) return 6;
  return 0;
}

template <class P>
hipError_t launch_pat(const dev_ode &p, const dev_coef *cf, state *st, double *y0, double *y1,
                      const sweep_cfg &c, hipStream_t stream)
{
  switch (c.waves) {
  case 1: return launch_one<P, 1>(p, cf, st, y0, y1, c.rows, stream);
  case 2: return launch_one<P, 2>(p, cf, st, y0, y1, c.rows, stream);
  default: return launch_one<P, 4>(p, cf, st, y0, y1, c.rows, stream);
  }
}

}  // namespace

bool sweep_supported(const dev_ode &p, const rkx_method &m)
{
  return p.kind == RKX_BRUSS2D && p.N >= 2 && pattern_of(m) != 0;
}

sweep_cfg sweep_default_cfg(const dev_ode &p, const rkx_method &m)
{
  (void) m;
  sweep_cfg c;
  c.waves = 4;
  c.rows = 32;
  if (p.N <= 256) {
    c.waves = 2;
    c.rows = 16;
  }
  return c;
}

hipError_t launch_sweep_step(const dev_ode &p, const rkx_method &m, const dev_coef *cf,
                             state *st, double *y0, double *y1, const sweep_cfg &c,
                             hipStream_t stream)
{
  switch (pattern_of(m)) {
  case 1: return launch_pat<pat_dopri54>(p, cf, st, y0, y1, c, stream);
  case 2: return launch_pat<pat_verner65>(p, cf, st, y0, y1, c, stream);
  case 3: return launch_pat<pat_rkf23>(p, cf, st, y0, y1, c, stream);
  case 4: return launch_pat<pat_heun2>(p, cf, st, y0, y1, c, stream);
  case 5: return launch_pat<pat_rk4>(p, cf, st, y0, y1, c, stream);
  case 6: return launch_pat<pat_ssprk3>(p, cf, st, y0, y1, c, stream);
  default: return hipErrorInvalidValue;
  }
}

}  // namespace rkx
