// Stage-fused, temporally blocked BRUSS2D step (the kernel template, shared
// by the translation units that instantiate it per method table): the GPU form of the
// reference's pipelined schedules PipeD / PipeDls (src/impl/seq/PipeD.c:76-124,
// src/impl/seq/PipeDls.c:81-129) with block size B = one grid row, which is
// the access distance d = 2N (bruss2d-mix.c:718-721).
//
// Work decomposition.  A workgroup owns a strip of L = 64*K grid columns and a
// band of R rows (or two bands of a narrow strip of L/2 lanes, one on each
// half of its waves).  It has 2K wavefronts: waves 0..K-1 carry the u
// component of the strip's columns (one lane per column), waves K..2K-1 the v
// component of the same columns, so each wave evaluates one of the two
// reaction formulas without divergence and holds half of a grid point's
// accumulators.  The
// workgroup sweeps the band's rows in the block-diagonal order of PipeD: at
// row iteration r, stage i is applied to row r - i, so when stage i needs w_i
// on rows k-1, k, k+1, row k+1 has just been completed by stage i-1 in the
// same iteration.  The per-row accumulators of the scatter form
// (w_{i+1..s-1}, dy, err; block-inline.h:253-320) stay in registers in a ring
// of s+1 row slots; the ring is fully unrolled so slots are static registers.
//
// Neighbours.  Vertical neighbours are the lane's own registers.  Horizontal
// neighbours and the other component of the same grid point are exchanged
// through LDS once per row iteration (double buffered, one barrier); each
// value is written to the next iteration's buffer as soon as it is final, so
// the barrier does not wait for a burst of writes.  Lanes
// at the strip edges and rows at the band edges are halo: each stage's valid
// region shrinks by one grid point per stage, so a strip produces L - 2s
// output columns and a band recomputes s-1 rows on each side (the
// stage-fused temporal blocking).
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
#pragma once

#include "rkx_patterns.hpp"

#include <algorithm>

namespace rkx {

// Internal linkage on purpose: each translation unit that includes this
// header instantiates the kernels of its own tables (rkx_kernels_sweep.hip:
// DOPRI54 ... SSPRK3; rkx_kernels_sweep_dopri87.hip: DOPRI87) and exports
// only the launch and occupancy functions declared in rkx_internal.hpp.
namespace {

// even periodic extension of a row/column index about 0 and N-1 (ghost
// rows and columns lie at most a few tiles' halo beyond the edge; the loop
// covers grids narrower than the halo as well)
__device__ __forceinline__ int reflect(int k, int N)
{
  while (k < 0 || k > N - 1)
    k = (k < 0) ? -k : 2 * (N - 1) - k;
  return k;
}

// GEN: ghost rows by the general reflection loop (grids too short for
// sweep_single_reflect); otherwise by a single reflection.
template <class P, int K, bool GEN>
__global__ __launch_bounds__(128 * K) __attribute__((amdgpu_waves_per_eu(P::occ))) void k_sweep(const sweep_inst *__restrict__ insts,
                                                   const int2 *__restrict__ tiles, int ntiles)
{
  constexpr int s = P::s;
  constexpr int L = 64 * K;          // columns per strip
  constexpr int NS = s + 1;          // row slots in the register ring
  constexpr int LP = L + 1;          // + a zero slot
  constexpr int ZERO = L;
  constexpr int PF = prefetch_rows(s);         // y rows loaded ahead of stage 0
  __shared__ double xch[2][s][2][LP];
  __shared__ double red[2 * K];
  __shared__ unsigned int zred[2 * K];

  // XCD-aware tile order: consecutive workgroups are dispatched round-robin
  // over the 8 XCDs; give each XCD a contiguous run of tiles (adjacent bands
  // of a strip share halo rows in that XCD's L2).
  const int T = ntiles;
  const int b = blockIdx.x;
  const int x = b & 7, q = T >> 3, rm = T & 7;
  const int tile = x * q + min(x, rm) + (b >> 3);
  // tile -> (instance, tile within the instance); no table: one instance
  const int2 td = tiles ? tiles[tile] : make_int2(0, tile);
  const sweep_inst &I = insts[td.x];
  double *const y0p = I.y0, *const y1p = I.y1;
  const int N = I.N;
  const int R = I.R, OC = I.OC;
  const int nwide = I.nwide, nbands = I.nbands;
  state *st = I.st;
  const step_view sv = read_step(st);
  if (sv.done)
    return;                                      // uniform over the instance's tiles
  // uniform, said explicitly: y, y_new and every row address below stay in
  // scalar registers
  const int cur = __builtin_amdgcn_readfirstlane(sv.cur);
  gconst *__restrict__ y = as_global(cur ? y1p : y0p);
  gmut *__restrict__ yn = as_global(cur ? y0p : y1p);
  const int N2 = N + N;
  const double coef = I.coef;
  gmethod *__restrict__ mt = as_global(sv.meth);
  const double h = sv.h;                         // this step's h (uniform)
  const int var = (threadIdx.x >= L) ? 1 : 0;   // wave-uniform: 0 = u, 1 = v
  const int cl = threadIdx.x - var * L;         // column lane within the strip

  // The tile's columns and band.  A wide tile is one band of a strip of L
  // lanes.  A paired tile carries two bands of the narrow last strip: lanes
  // [0, L/2) of each component the band 2p, lanes [L/2, L) the band 2p+1
  // (whole waves, as L/2 is a multiple of 64); without a band 2p+1 those
  // waves repeat band 2p and store nothing.  Both halves run the full band
  // height, so every wave passes the same barriers.
  const int nwt = nwide * nbands;
  const bool paired = td.y >= nwt;
  int band, lo, lsub, c0, c1;
  bool live = true;
  if (!paired) {
    const int strip = td.y / nbands;
    band = td.y - strip * nbands;
    lo = 0;
    lsub = L;
    c0 = strip * OC;
    c1 = c0 + OC;
  } else {
    const int half = __builtin_amdgcn_readfirstlane(cl >= L / 2 ? 1 : 0);
    band = 2 * (td.y - nwt) + half;
    if (band >= nbands) {
      band = nbands - 1;
      live = false;
    }
    lo = half * (L / 2);
    lsub = L / 2;
    c0 = nwide * OC;
    c1 = N;
  }
  const int cs = cl - lo;                       // column lane within the tile
  const int col = c0 - s + cs;
  const int ccol = reflect(col, N);             // in-bounds column for loads
  const bool out_col = live && (cs >= s) && (cs < lsub - s) && (col >= 0) && (col < N) &&
                       (col < c1);
  // Horizontal part of the Laplacian: ((vN + vS) + X) + m*Y.  Interior
  // columns read X = W, Y = E, m = 1.  Column 0 reads X from the zero slot and
  // Y = E with m = 2; column N-1 reads X from the zero slot and Y = W with
  // m = 2.  Scaling by 1 or 2 is exact, so every lane computes the
  // reference's expression for its column (bruss2d-mix.c:66-206).
  const int slotX = (col == 0 || col == N - 1) ? ZERO : max(cl - 1, 0);
  const int slotY = (col == N - 1) ? max(cl - 1, 0) : min(cl + 1, L - 1);
  const double mY = (col == 0 || col == N - 1) ? 2.0 : 1.0;
  if (threadIdx.x < 2 * s) {
    const int bb = threadIdx.x / s, i = threadIdx.x - bb * s;
    xch[bb][i][0][ZERO] = 0.0;
    xch[bb][i][1][ZERO] = 0.0;
  }

  const int r0 = band * R;
  const int r1 = min(r0 + R, N);
  const int it0 = r0 - (s - 1);                  // first row iteration
  const int it1 = paired ? it0 + R + 2 * (s - 1) : r1 + s - 1;   // one past the last

  // register ring: slot -> [0] dy, [1..s-1] w_l, [s] err (this lane's component)
  double acc[NS][s + 1];
  double yr[NS];                                 // y rows of stage 0, ring by row
  double yo[NS];                                 // y_old of finishing rows, ring by iteration
  constexpr int DY = 2;                          // y_old loaded DY iterations ahead
  // the first stored row (kl = r0) finishes at iteration it0 + 2(s-1); its
  // y_old must have been loaded by then, i.e. in an iteration >= it0
  static_assert(DY <= 2 * (s - 1) && DY + 1 <= NS, "y_old prefetch outruns the band");
  static_assert(PF + 3 <= NS, "y ring too small for the prefetch");
  // byte offset of this lane's component within a row (a row is addressed by
  // a scalar base, the lane by a 32-bit vector offset)
  const unsigned int lane_b = (2u * static_cast<unsigned int>(ccol) + var) * 8u;
  auto at_row = [&](int kr) -> double {
    const gbyte *row = reinterpret_cast<const gbyte *>(
        y + static_cast<unsigned>(kr) * static_cast<unsigned>(N2));
    return *reinterpret_cast<gconst *>(row + lane_b);
  };
  // Ghost rows: a single reflection about row 0 / N-1, min(|k|, 2(N-1) - |k|),
  // covers every row a band loads once the grid is taller than the halo plus
  // the prefetch (plus one trip of the ring with full trips, which may run
  // past the band's last iteration; sweep_single_reflect); shorter grids take
  // the general loop (GEN).  Offsets fit 32 bits (n = 2N^2 < 2^31).
  auto yrow = [&](int k) -> double {
    const int a = k < 0 ? -k : k;
    const int kr = GEN ? reflect(k, N) : min(a, 2 * (N - 1) - a);
    return at_row(kr);
  };
  static_for<PF + 2>([&](auto dt) {
    constexpr int d = decltype(dt)::value - 1;   // rows it0-1 .. it0+PF
    yr[(d + NS) % NS] = yrow(it0 + d);
  });
  static_for<NS>([&](auto dt) {
    constexpr int d = decltype(dt)::value;
    yo[d] = 0.0;
    static_for<s + 1>([&](auto et) { acc[d][decltype(et)::value] = 0.0; });
  });

  // The step's premultiplied coefficients, formed once (uniform, SGPRs):
  // kv[i] = h * col[i] = {hA[i+1..s-1][i], hb[i], he[i]} for the scatter of
  // stage i, the products premult forms (src/methods.c:43-58).  Only the
  // non-zero entries are formed; for a first-same-as-last table (A[s-1][j]
  // == b[j], e.g. DOPRI54) h*A[s-1][j] and h*b[j] are the same product, so
  // hb[j] reuses the register of hA[s-1][j].
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

  // The row loop, instantiated once for the u waves and once for the v
  // waves (VAR), so that each is straight-line code per row iteration.
  auto rows = [&](auto vt) {
    constexpr int VAR = decltype(vt)::value;
    // the first iteration's exchange: the band's first y row for stage 0,
    // zero accumulators of the rows above the band for the other stages
    static_for<s>([&](auto itg) {
      constexpr int i = decltype(itg)::value;
      xch[0][i][VAR][cl] = (i == 0) ? yr[0] : 0.0;
    });
    // One row iteration rr; U = (rr - it0) % NS names the ring slots.  Stage
    // i runs when 2i <= J: in the band's first 2(s-1) iterations (J = the
    // iteration's index) stage i would work on rows above the ones later
    // stages need (row rr - i < r0 - (s-1-i)), whose values reach no stored
    // row, so it is left out; J = 2(s-1) runs every stage.
    auto iter = [&](const int rr, auto ut, auto jt) {
        constexpr int U = decltype(ut)::value;
        constexpr int J = decltype(jt)::value;
        // Loads complete in issue order, so every load used in this
        // iteration was issued in an earlier one: y for stage 0 PF rows
        // ahead, and y_old of the row the last stage finishes DY iterations
        // ahead (clamped into the band: rows outside it are never stored).
        const int kf = min(max(rr + DY - (s - 1), r0), r1 - 1);
        yo[(U + DY) % NS] = at_row(kf);
        yr[(U + 1 + PF) % NS] = yrow(rr + 1 + PF);
        const int kl = rr - (s - 1);                 // row finishing this iteration
        const double yold = yo[U % NS];

        // With full trips the first 2(s-1) iterations run every stage (their
        // extra work on rows above the range is never read by an in-range
        // row and never stored); with full trips, and in a paired band
        // shorter than R, the last iterations work on rows below the band,
        // which are never stored either.
        const int buf = (rr - it0) & 1;
        __syncthreads();
        double nX[s], nY[s], nP[s];
        static_for<s>([&](auto itg) {
          constexpr int i = decltype(itg)::value;
          if constexpr (2 * i <= J) {
            nX[i] = xch[buf][i][VAR][slotX];
            nY[i] = xch[buf][i][VAR][slotY];
            nP[i] = xch[buf][i][VAR ^ 1][cl];           // the other component
          }
        });
        // The next iteration's exchange goes to the other buffer as each
        // value becomes final: its stage-0 row now, w_{i+1} of row rr - i
        // right after stage i's scatter.  Every wave read that buffer in the
        // previous iteration, before the barrier above.
        xch[buf ^ 1][0][VAR][cl] = yr[(U + 1) % NS];

        static_for<s>([&](auto itg) {
          constexpr int i = decltype(itg)::value;
          if constexpr (2 * i > J)
            return;
          constexpr int sl = (U - i + NS) % NS;              // slot of row rr - i
          constexpr int sn = (U - i - 1 + 2 * NS) % NS;      // row rr - i - 1
          constexpr int ss = (U - i + 1 + NS) % NS;          // row rr - i + 1
          double c, vn, vs;
          if constexpr (i == 0) {
            c = yr[U % NS];
            vn = yr[(U - 1 + NS) % NS];
            vs = yr[(U + 1) % NS];
          } else {
            c = acc[sl][i];
            vn = acc[sn][i];
            vs = acc[ss][i];
          }
          const double pt = nP[i];
          // Laplacian numerator ((vN + vS) + X) + m*Y - 4c, then the reaction.
          // m*Y and 4c are exact (m is 1 or 2), so each multiply-add rounds
          // once, as the separate add does.
          const double lap = __builtin_fma(-4.0, c, __builtin_fma(mY, nY[i], (vn + vs) + nX[i]));
          double f;
          if constexpr (VAR == 0) {
            const double uuv = (c * c) * pt;
            f = ((1.0 + uuv) - 4.4 * c) + coef * lap;
          } else {
            const double uuv = (pt * pt) * c;
            f = (3.4 * pt - uuv) + coef * lap;
          }
          const double *kc = kv[i];
          if constexpr (i == 0) {
            // first stage: w_l = y + hA[l][0] F, dy = hb0 F, err = he0 F
            static_for<s>([&](auto ltg) {
              constexpr int l = decltype(ltg)::value;
              if constexpr (l >= 1)
                acc[sl][l] = nzA<P>(l, 0) ? c + kc[l - 1] * f : c;
            });
            acc[sl][0] = kc[s - 1] * f;
            acc[sl][s] = kc[s] * f;
          } else {
            static_for<s>([&](auto ltg) {
              constexpr int l = decltype(ltg)::value;
              if constexpr (l > i && nzA<P>(l, i))
                acc[sl][l] += kc[l - i - 1] * f;
            });
            if constexpr ((P::b >> i) & 1u)
              acc[sl][0] += kc[s - i - 1] * f;
            if constexpr ((P::e >> i) & 1u)
              acc[sl][s] += kc[s - i] * f;
          }
          if constexpr (i + 1 < s)
            xch[buf ^ 1][i + 1][VAR][cl] = acc[sl][i + 1];
          if constexpr (i == s - 1) {
            // y_new = y + dy; the error term against y_old (block-inline.h:298-320)
            const double ynw = yold + acc[sl][0];
            const bool out = out_col && (kl >= r0) && (kl < r1);
            // Straight-line code, so the iteration stays one scheduling
            // region: the store is masked by buffer range checking, and the
            // division runs in every lane on selected operands.  q = |e| /
            // max(|y_new|, |y_old|); a zero divisor raises the sentinel; lanes
            // not stored contribute 0/1 (qmax >= 0 already); NaN quotients
            // are skipped by fmax.
            store_row_masked(yn + static_cast<unsigned>(kl) * static_cast<unsigned>(N2), N2,
                             2 * col + VAR, out, ynw);
            const double dv = fmax(fabs(ynw), fabs(yold));
            const bool dz = (dv == 0.0);
            // lanes not stored and zero divisors give 0/dv: 0, or NaN for
            // dv = 0 or NaN, which fmax skips
            const bool take = out && !dz;
            double num = take ? fabs(acc[sl][s]) : 0.0, den = dv;
            // opaque operands: otherwise the compiler folds 0/1 and puts the
            // division back under a branch
            __asm__("" : "+v"(num), "+v"(den));
            const double qv = num / den;
            zero |= (out && dz) ? 1u : 0u;
            qmax = fmax(qmax, qv);
          }
        });
    };
    constexpr int JALL = 2 * (s - 1);
    // the peeled iterations (every band runs more: R >= 1); with whole
    // trips the band height makes the rest fill whole trips
    constexpr int PRO = peeled_iters(P::full_trips, s);
    static_for<PRO>([&](auto jt) {
      constexpr int j = decltype(jt)::value;
      iter(it0 + j, std::integral_constant<int, j % NS>{}, jt);
    });
    for (int rb = it0 + PRO; rb < it1; rb += NS) {
      static_for<NS>([&](auto ut) {
        constexpr int U = decltype(ut)::value;
        const int rr = rb + U;
        if constexpr (!P::full_trips) {
          if (rr >= it1)
            return;
        }
        iter(rr, std::integral_constant<int, (U + PRO) % NS>{},
             std::integral_constant<int, JALL>{});
      });
    }
  };
  if (__builtin_amdgcn_readfirstlane(var) == 0)
    rows(std::integral_constant<int, 0>{});
  else
    rows(std::integral_constant<int, 1>{});
  block_finish(st, qmax, zero, red, zred, static_cast<unsigned int>(I.ntiles), false);
}

template <class P, int K>
hipError_t launch_one(bool gen, const sweep_inst *insts, const int2 *tiles, int ntiles,
                      hipStream_t stream)
{
  if (gen)
    hipLaunchKernelGGL((k_sweep<P, K, true>), dim3(ntiles), dim3(128 * K), 0, stream, insts,
                       tiles, ntiles);
  else
    hipLaunchKernelGGL((k_sweep<P, K, false>), dim3(ntiles), dim3(128 * K), 0, stream, insts,
                       tiles, ntiles);
  return hipGetLastError();
}

// Widest strip a table's kernel is built for: a workgroup's 2K waves must fit
// one CU (4 SIMDs) at the occupancy the kernel is register-allocated for.
template <class P>
constexpr int max_waves()
{
  return (2 * P::occ < 4) ? 2 * P::occ : 4;
}

template <class P>
hipError_t launch_pat(int K, bool gen, const sweep_inst *insts, const int2 *tiles, int ntiles,
                      hipStream_t stream)
{
  if (K < 1 || K > max_waves<P>())
    return hipErrorInvalidConfiguration;
  switch (K) {
  case 1: return launch_one<P, 1>(gen, insts, tiles, ntiles, stream);
  case 2: return launch_one<P, 2>(gen, insts, tiles, ntiles, stream);
  case 3:
    if constexpr (max_waves<P>() >= 3)
      return launch_one<P, 3>(gen, insts, tiles, ntiles, stream);
    return hipErrorInvalidConfiguration;
  default:
    if constexpr (max_waves<P>() >= 4)
      return launch_one<P, 4>(gen, insts, tiles, ntiles, stream);
    return hipErrorInvalidConfiguration;
  }
}

// Workgroups of k_sweep<P, K> resident per CU (0: the strip is too wide for
// the table's kernel).
template <class P, int K>
int blocks_per_cu()
{
  int nb = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_sweep<P, K, false>, 128 * K, 0) !=
      hipSuccess) {
    (void) hipGetLastError();
    nb = 1;
  }
  return nb > 0 ? nb : 0;
}

template <class P>
int blocks_per_cu_k(int K)
{
  if (K < 1 || K > max_waves<P>())
    return 0;
  switch (K) {
  case 1: return blocks_per_cu<P, 1>();
  case 2: return blocks_per_cu<P, 2>();
  case 3:
    if constexpr (max_waves<P>() >= 3)
      return blocks_per_cu<P, 3>();
    return 0;
  default:
    if constexpr (max_waves<P>() >= 4)
      return blocks_per_cu<P, 4>();
    return 0;
  }
}

}  // namespace
}  // namespace rkx
