// Stage-fused, temporally blocked BRUSS2D step: the GPU form of the
// reference's pipelined schedules PipeD / PipeDls (src/impl/seq/PipeD.c:76-124,
// src/impl/seq/PipeDls.c:81-129) with block size B = one grid row, which is
// the access distance d = 2N (bruss2d-mix.c:718-721).
//
// Work decomposition.  A workgroup owns a strip of L = 64*K grid columns and a
// band of R rows.  It has 2K wavefronts: waves 0..K-1 carry the u component of
// the strip's columns (one lane per column), waves K..2K-1 the v component of
// the same columns, so each wave evaluates one of the two reaction formulas
// without divergence and holds half of a grid point's accumulators.  The
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
#include "rkx_patterns.hpp"

#include <algorithm>

namespace rkx {

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
  state *st = I.st;
  if (st->done | st->pow_wait)
    return;                                      // uniform over the instance's tiles
  const int cur = st->cur;
  gconst *__restrict__ y = as_global(cur ? I.y1 : I.y0);
  gmut *__restrict__ yn = as_global(cur ? I.y0 : I.y1);
  const int N = I.N;
  const int N2 = N + N;
  const double coef = I.coef;
  const int R = I.R, OC = I.OC;
  const dev_coef *__restrict__ cf = I.cf;
  const int strip = td.y / I.nbands;
  const int band = td.y - strip * I.nbands;

  const int var = (threadIdx.x >= L) ? 1 : 0;   // wave-uniform: 0 = u, 1 = v
  const int cl = threadIdx.x - var * L;         // column lane within the strip
  const int col = strip * OC - s + cl;
  const int ccol = reflect(col, N);             // in-bounds column for loads
  const bool out_col = (cl >= s) && (cl < L - s) && (col >= 0) && (col < N) &&
                       (col < (strip + 1) * OC);
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
  const int it1 = r1 + s - 1;                    // one past the last

  // register ring: slot -> [0] dy, [1..s-1] w_l, [s] err (this lane's component)
  double acc[NS][s + 1];
  double yr[NS];                                 // y rows of stage 0, ring by row
  double yo[NS];                                 // y_old of finishing rows, ring by iteration
  constexpr int DY = 2;                          // y_old loaded DY iterations ahead
  // the first stored row (kl = r0) finishes at iteration it0 + 2(s-1); its
  // y_old must have been loaded by then, i.e. in an iteration >= it0
  static_assert(DY <= 2 * (s - 1) && DY + 1 <= NS, "y_old prefetch outruns the band");
  static_assert(PF + 3 <= NS, "y ring too small for the prefetch");
  gconst *yc = y + 2 * ccol + var;
  // Ghost rows: a single reflection about row 0 / N-1, min(|k|, 2(N-1) - |k|),
  // covers every row a band loads once the grid is taller than the halo plus
  // the prefetch (plus one trip of the ring with full trips, which may run
  // past the band's last iteration; sweep_single_reflect); shorter grids take
  // the general loop (GEN).  Offsets fit 32 bits (n = 2N^2 < 2^31).
  auto yrow = [&](int k) -> double {
    const int a = k < 0 ? -k : k;
    const int kr = GEN ? reflect(k, N) : min(a, 2 * (N - 1) - a);
    return yc[static_cast<unsigned>(kr) * static_cast<unsigned>(N2)];
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

  // The step's premultiplied coefficients, loaded once (uniform, SGPRs):
  // kv[i] = col[i] = {hA[i+1..s-1][i], hb[i], he[i]} for the scatter of stage
  // i.  Only the non-zero entries are loaded; for a first-same-as-last table
  // (A[s-1][j] == b[j], e.g. DOPRI54) h*A[s-1][j] and h*b[j] are the same
  // product, so hb[j] reuses the register of hA[s-1][j].
  double kv[s][s + 1];
  static_for<s>([&](auto it) {
    constexpr int i = decltype(it)::value;
    static_for<s>([&](auto lt) {
      constexpr int l = decltype(lt)::value;
      if constexpr (l > i)
        kv[i][l - i - 1] = nzA<P>(l, i) ? wave_uniform(cf->col[i][l - i - 1]) : 0.0;
    });
    if constexpr (P::fsal && i < s - 1)
      kv[i][s - i - 1] = kv[i][s - i - 2];
    else
      kv[i][s - i - 1] = ((P::b >> i) & 1u) ? wave_uniform(cf->col[i][s - i - 1]) : 0.0;
    kv[i][s - i] = ((P::e >> i) & 1u) ? wave_uniform(cf->col[i][s - i]) : 0.0;
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
    for (int rb = it0; rb < it1; rb += NS) {
      static_for<NS>([&](auto ut) {
        constexpr int U = decltype(ut)::value;
        const int rr = rb + U;
        if constexpr (!P::full_trips) {
          if (rr >= it1)
            return;
        }
        // Loads complete in issue order, so every load used in this
        // iteration was issued in an earlier one: y for stage 0 PF rows
        // ahead, and y_old of the row the last stage finishes DY iterations
        // ahead (clamped into the band: rows outside it are never stored).
        const int kf = min(max(rr + DY - (s - 1), r0), r1 - 1);
        yo[(U + DY) % NS] = yc[static_cast<unsigned>(kf) * static_cast<unsigned>(N2)];
        yr[(U + 1 + PF) % NS] = yrow(rr + 1 + PF);
        const int kl = rr - (s - 1);                 // row finishing this iteration
        const double yold = yo[U % NS];

        // Every stage runs every iteration: during the first 2(s-1)
        // iterations of a band, stage i works on rows above its range, whose
        // results are never read by an in-range row and never stored; with
        // full trips, iterations past it1 work on rows below the band, which
        // are never stored either.
        const int buf = (rr - it0) & 1;
        __syncthreads();
        double nX[s], nY[s], nP[s];
        static_for<s>([&](auto itg) {
          constexpr int i = decltype(itg)::value;
          nX[i] = xch[buf][i][VAR][slotX];
          nY[i] = xch[buf][i][VAR][slotY];
          nP[i] = xch[buf][i][VAR ^ 1][cl];           // the other component
        });
        // The next iteration's exchange goes to the other buffer as each
        // value becomes final: its stage-0 row now, w_{i+1} of row rr - i
        // right after stage i's scatter.  Every wave read that buffer in the
        // previous iteration, before the barrier above.
        xch[buf ^ 1][0][VAR][cl] = yr[(U + 1) % NS];

        static_for<s>([&](auto itg) {
          constexpr int i = decltype(itg)::value;
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
          // Laplacian numerator ((vN + vS) + X) + m*Y - 4c, then the reaction
          const double lap = __builtin_fma(-4.0, c, ((vn + vs) + nX[i]) + mY * nY[i]);
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
            const bool out = out_col && (kl >= r0) && (!P::full_trips || kl < r1);
            if (out)
              yn[static_cast<unsigned>(kl) * static_cast<unsigned>(N2) + 2 * col + VAR] = ynw;
            // without branches: q = |e| / max(|y_new|, |y_old|); a zero
            // divisor raises the sentinel, NaN quotients are skipped by fmax
            const double dv = fmax(fabs(ynw), fabs(yold));
            const double qv = fabs(acc[sl][s]) / dv;
            const bool dz = (dv == 0.0);
            zero |= (out && dz) ? 1u : 0u;
            qmax = (out && !dz) ? fmax(qmax, qv) : qmax;
          }
        });
      });
    }
  };
  if (__builtin_amdgcn_readfirstlane(var) == 0)
    rows(std::integral_constant<int, 0>{});
  else
    rows(std::integral_constant<int, 1>{});
  block_finish(st, qmax, zero, red, zred, static_cast<unsigned int>(I.ntiles));
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

template <class P>
hipError_t launch_pat(int K, bool gen, const sweep_inst *insts, const int2 *tiles, int ntiles,
                      hipStream_t stream)
{
  switch (K) {
  case 1: return launch_one<P, 1>(gen, insts, tiles, ntiles, stream);
  case 2: return launch_one<P, 2>(gen, insts, tiles, ntiles, stream);
  case 3: return launch_one<P, 3>(gen, insts, tiles, ntiles, stream);
  default: return launch_one<P, 4>(gen, insts, tiles, ntiles, stream);
  }
}

template <class P, int K>
int blocks_per_cu()
{
  int nb = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_sweep<P, K, false>, 128 * K, 0) !=
      hipSuccess) {
    (void) hipGetLastError();
    nb = 1;
  }
  return nb > 0 ? nb : 1;
}

template <class P>
int blocks_per_cu_k(int K)
{
  switch (K) {
  case 1: return blocks_per_cu<P, 1>();
  case 2: return blocks_per_cu<P, 2>();
  case 3: return blocks_per_cu<P, 3>();
  default: return blocks_per_cu<P, 4>();
  }
}

int resident_blocks(int pattern, int K)
{
  switch (pattern) {
  case 1: return blocks_per_cu_k<pat_dopri54>(K);
  case 2: return blocks_per_cu_k<pat_verner65>(K);
  case 3: return blocks_per_cu_k<pat_rkf23>(K);
  case 4: return blocks_per_cu_k<pat_heun2>(K);
  case 5: return blocks_per_cu_k<pat_rk4>(K);
  default: return blocks_per_cu_k<pat_ssprk3>(K);
  }
}

}  // namespace

bool sweep_supported(const dev_ode &p, const rkx_method &m)
{
  return p.kind == RKX_BRUSS2D && p.N >= 2 && pattern_of(m) != 0;
}

// Relative cost of one lane x row iteration of k_sweep for strips of 1..4
// waves per component, measured on MI355X (DOPRI54, N = 2000, one dispatch
// round): 1 wave pays a barrier and an LDS exchange for few lanes, 3 waves
// (6 per workgroup) do not pack two workgroups per CU.
constexpr double kTileCost[4] = {1.36, 1.0, 1.67, 0.97};

double sweep_tile_cost(int N, int s, int K, int R, const double *f, int trip)
{
  const int L = 64 * K;
  const int oc = L - 2 * s;
  if (oc < 1 || R < 1 || N < 1)
    return 1e300;
  const double fk = (f && f[K - 1] > 0.0) ? f[K - 1] : kTileCost[K - 1];
  const long long nstrips = (N + oc - 1) / oc;
  const long long nbands = (N + R - 1) / R;
  const long long reff = (N + nbands - 1) / nbands;
  long long iters = reff + 2 * (s - 1);
  if (trip > 0)
    iters = (iters + trip - 1) / trip * trip;
  // + 4: the prologue of a tile (coefficients, first rows) in row iterations
  return fk * static_cast<double>(nstrips * L) * static_cast<double>(nbands) *
         static_cast<double>(iters + 4);
}

namespace {

bool full_trips_of(int pattern)
{
  switch (pattern) {
  case 1: return pat_dopri54::full_trips;
  case 2: return pat_verner65::full_trips;
  case 3: return pat_rkf23::full_trips;
  case 4: return pat_heun2::full_trips;
  case 5: return pat_rk4::full_trips;
  default: return pat_ssprk3::full_trips;
  }
}

}  // namespace

bool sweep_single_reflect(int pattern, int s, int N)
{
  return N > s + prefetch_rows(s) + 3 + (full_trips_of(pattern) ? s + 1 : 0);
}

int sweep_band_rows(int pattern, int s, int rows)
{
  if (!full_trips_of(pattern))
    return rows;
  const int ns = s + 1;
  const int it = rows + 2 * (s - 1);
  return (it + ns - 1) / ns * ns - 2 * (s - 1);
}

// Default tile: strips of 64*K columns (2K waves per workgroup) and the band
// height R for which all strips x bands tiles are resident at once (one
// dispatch round: a second, partial round would leave most CUs idle while
// it runs), with bands as tall as that allows (the s-1 halo rows a band
// recomputes are a smaller share of taller bands).  Where more than one
// workgroup fits a CU, every count from the one that gives two waves per
// SIMD up to the resident maximum is a candidate: on MI355X the FP64 row
// loop gains little from more than two waves per SIMD (RKF23 at N = 2000:
// 58 us per step either way with the same bands), so the taller bands of
// fewer workgroups per CU can win (RKF23: 54.7 us with one per CU against
// 58.6 us filling both).  The candidate with the lowest modelled cost wins.
sweep_cfg sweep_default_cfg(const dev_ode &p, const rkx_method &m)
{
  const int s = m.s;
  const int N = p.N;
  const int pat = pattern_of(m);
  int cus = 0, dev = 0;
  if (hipGetDevice(&dev) != hipSuccess ||
      hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
      cus < 1) {
    (void) hipGetLastError();
    cus = 256;
  }
  sweep_cfg best{2, 4};
  double best_cost = 1e300;
  for (int K = 1; K <= 4; ++K) {
    const int oc = 64 * K - 2 * s;
    if (oc < 1)
      continue;
    const int nstrips = (N + oc - 1) / oc;
    const int res = resident_blocks(pat, K);
    const int per_cu_min = std::min(res, (4 + K - 1) / K);   // 2K waves each: 2 per SIMD
    for (int per_cu = per_cu_min; per_cu <= res; ++per_cu) {
      int bands = cus * per_cu / nstrips;
      if (bands < 1)
        bands = 1;
      int rows = (N + bands - 1) / bands;
      if (rows < 4)
        rows = 4;
      rows = sweep_band_rows(pat, s, rows);
      const double cost = sweep_tile_cost(N, s, K, rows, nullptr, full_trips_of(pat) ? s + 1 : 0);
      if (cost < best_cost) {
        best_cost = cost;
        best = sweep_cfg{K, rows};
      }
    }
  }
  return best;
}

// Geometry of one instance's tiles for strips of 64*K columns and bands of R rows.
sweep_inst make_sweep_inst(const dev_ode &p, const rkx_method &m, const sweep_cfg &c, state *st,
                           const dev_coef *cf, double *y0, double *y1)
{
  sweep_inst I{};
  const int L = 64 * c.waves;
  const int oc_max = L - 2 * m.s;
  I.N = p.N;
  I.R = sweep_band_rows(pattern_of(m), m.s, c.rows);
  I.nstrips = (p.N + oc_max - 1) / oc_max;
  I.OC = (p.N + I.nstrips - 1) / I.nstrips;
  I.nbands = (p.N + I.R - 1) / I.R;
  I.ntiles = I.nstrips * I.nbands;
  I.coef = p.coef;
  I.y0 = y0;
  I.y1 = y1;
  I.st = st;
  I.cf = cf;
  return I;
}

int sweep_pattern(const rkx_method &m) { return pattern_of(m); }

hipError_t launch_sweep(int pattern, int K, bool gen, const sweep_inst *insts, const int2 *tiles,
                        int ntiles, hipStream_t stream)
{
  if (ntiles <= 0)
    return hipSuccess;
  switch (pattern) {
  case 1: return launch_pat<pat_dopri54>(K, gen, insts, tiles, ntiles, stream);
  case 2: return launch_pat<pat_verner65>(K, gen, insts, tiles, ntiles, stream);
  case 3: return launch_pat<pat_rkf23>(K, gen, insts, tiles, ntiles, stream);
  case 4: return launch_pat<pat_heun2>(K, gen, insts, tiles, ntiles, stream);
  case 5: return launch_pat<pat_rk4>(K, gen, insts, tiles, ntiles, stream);
  case 6: return launch_pat<pat_ssprk3>(K, gen, insts, tiles, ntiles, stream);
  default: return hipErrorInvalidValue;
  }
}

}  // namespace rkx
