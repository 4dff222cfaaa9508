// Host side of the stage-fused BRUSS2D step (rkx_sweep_kernel.hpp): tile
// geometry and defaults, and the launches of the kernels of the tables with
// a compiled zero pattern (DOPRI87's are built in rkx_kernels_sweep_dopri87.hip).
#include "rkx_sweep_kernel.hpp"

#include <algorithm>

namespace rkx {

namespace {

int resident_blocks(int pattern, int K)
{
  switch (pattern) {
  case 1: return blocks_per_cu_k<pat_dopri54>(K);
  case 2: return blocks_per_cu_k<pat_verner65>(K);
  case 3: return blocks_per_cu_k<pat_rkf23>(K);
  case 4: return blocks_per_cu_k<pat_heun2>(K);
  case 5: return blocks_per_cu_k<pat_rk4>(K);
  case 6: return blocks_per_cu_k<pat_ssprk3>(K);
  case 7: return resident_blocks_dopri87(K);
  default: return 0;
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

namespace {

// Strips of an N-column grid for strips of 64K lanes: as many full-width
// strips (64K - 2s output columns) as fit, and the columns left over either
// in a narrow strip of 32K lanes whose bands are paired two to a workgroup
// (K even, and they fit its 32K - 2s output columns), or else in one more
// strip, the strips then evened out.
struct strip_geom {
  int nwide;      // full-width strips
  int oc;         // output columns per full-width strip
  int split;      // 1: a narrow last strip of paired bands
  // workgroups for nb bands per strip
  long long wgs(long long nb) const { return nwide * nb + split * ((nb + 1) / 2); }
};

strip_geom strips_of(int N, int s, int K)
{
  const int oc_max = 64 * K - 2 * s;
  const int oc_half = (K % 2 == 0) ? 32 * K - 2 * s : 0;
  const int nfull = N / oc_max;
  const int rem = N - nfull * oc_max;
  if (rem > 0 && rem <= oc_half)
    return strip_geom{nfull, oc_max, 1};
  const int nstrips = (N + oc_max - 1) / oc_max;
  return strip_geom{nstrips, (N + nstrips - 1) / nstrips, 0};
}

// Relative cost of one step of an N x N instance with strips of 64K columns
// and bands of R rows: lanes x row iterations (the balanced band height plus
// the 2(s-1) halo iterations and a prologue of about 4) x the measured
// per-lane factor of the strip width.  trip > 0 rounds a band's row
// iterations after the pro peeled ones up to whole trips of that many rows.
double tile_work(int N, int s, int K, int R, int trip, int pro)
{
  const int L = 64 * K;
  const int oc = L - 2 * s;
  if (oc < 1 || R < 1 || N < 1)
    return 1e300;
  const strip_geom g = strips_of(N, s, K);
  const long long nbands = (N + R - 1) / R;
  const long long reff = g.split ? R : (N + nbands - 1) / nbands;
  long long iters = reff + 2 * (s - 1);
  if (trip > 0)
    iters = pro + (iters - pro + trip - 1) / trip * trip;
  return kTileCost[K - 1] * static_cast<double>(g.wgs(nbands) * L) *
         static_cast<double>(iters + 4);
}

// Most bands per strip whose workgroups fit `slots`.
long long bands_for_slots(const strip_geom &g, long long slots)
{
  long long nb = static_cast<long long>(static_cast<double>(slots) / (g.nwide + 0.5 * g.split));
  while (nb > 1 && g.wgs(nb) > slots)
    --nb;
  return nb < 1 ? 1 : nb;
}

}  // namespace

namespace {

bool full_trips_of(int pattern)
{
  switch (pattern) {
  case 1: return pat_dopri54::full_trips;
  case 2: return pat_verner65::full_trips;
  case 3: return pat_rkf23::full_trips;
  case 4: return pat_heun2::full_trips;
  case 5: return pat_rk4::full_trips;
  case 6: return pat_ssprk3::full_trips;
  case 7: return pat_dopri87::full_trips;
  default: return false;
  }
}

}  // namespace

bool sweep_single_reflect(int pattern, int s, int N)
{
  return N > s + prefetch_rows(s) + 3 + (full_trips_of(pattern) ? s + 1 : 0);
}

bool sweep_inst_single_reflect(int pattern, int s, const sweep_inst &I)
{
  if (!sweep_single_reflect(pattern, s, I.N))
    return false;
  // the last paired band loads rows up to nbands*R + s - 1 + prefetch
  return !I.split ||
         static_cast<long long>(I.nbands) * I.R + s - 1 + prefetch_rows(s) <= 2LL * (I.N - 1);
}

int sweep_tile_iters(int pattern, int s, const sweep_inst &I, int t)
{
  if (t >= I.nwide * I.nbands)
    return I.R + 2 * (s - 1);                   // paired: the full band height
  const int band = t % I.nbands;
  const int rows = std::min(I.R, I.N - band * I.R);
  return sweep_band_rows(pattern, s, rows) + 2 * (s - 1);
}

int sweep_band_rows(int pattern, int s, int rows)
{
  if (!full_trips_of(pattern))
    return rows;
  const int ns = s + 1;
  const int pro = peeled_iters(true, s);
  const int it = rows + 2 * (s - 1) - pro;      // iterations in the loop
  return (it + ns - 1) / ns * ns - 2 * (s - 1) + pro;
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
// A grid so wide that one band per strip gives fewer tiles than CUs (N =
// 32767, K = 4: 136 strips) or more than fit at once (DOPRI87, N = 32767:
// 322 strips of 2 waves) instead takes shorter bands over several rounds
// when that shortens the step: a CU then works through its tiles in turn,
// ceil(tiles / CUs) x the row iterations of one.
sweep_cfg sweep_default_cfg(const dev_ode &p, const rkx_method &m)
{
  const int s = m.s;
  const int N = p.N;
  const int pat = pattern_of(m);
  const int trip = full_trips_of(pat) ? s + 1 : 0;
  const int pro = peeled_iters(true, s);      // used with trip only
  int cus = 0, dev = 0;
  if (hipGetDevice(&dev) != hipSuccess ||
      hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
      cus < 1) {
    (void) hipGetLastError();
    cus = 256;
  }
  auto band_rows_for = [&](long long bands) {
    if (bands < 1)
      bands = 1;
    int rows = static_cast<int>((N + bands - 1) / bands);
    return sweep_band_rows(pat, s, rows < 4 ? 4 : rows);
  };
  sweep_cfg best{2, 4};
  double best_cost = 1e300;
  for (int K = 1; K <= 4; ++K) {
    const int oc = 64 * K - 2 * s;
    if (oc < 1)
      continue;
    const strip_geom g = strips_of(N, s, K);
    const int res = resident_blocks(pat, K);
    if (res < 1)
      continue;                                     // too wide for this table's kernel
    const int per_cu_min = std::min(res, (4 + K - 1) / K);   // 2K waves each: 2 per SIMD
    for (int per_cu = per_cu_min; per_cu <= res; ++per_cu) {
      const int rows = band_rows_for(bands_for_slots(g, static_cast<long long>(cus) * per_cu));
      const double cost = tile_work(N, s, K, rows, trip, pro);
      if (cost < best_cost) {
        best_cost = cost;
        best = sweep_cfg{K, rows};
      }
    }
  }
  // several rounds for grids too wide for one band per strip to fill the CUs
  const strip_geom g = strips_of(N, s, best.waves);
  auto round_time = [&](int rows) {
    const long long tiles = g.wgs((N + rows - 1) / rows);
    long long iters = rows + 2 * (s - 1);
    if (trip > 0)
      iters = pro + (iters - pro + trip - 1) / trip * trip;
    return static_cast<double>((tiles + cus - 1) / cus) * static_cast<double>(iters + 4);
  };
  const long long tiles = g.wgs((N + best.rows - 1) / best.rows);
  if (tiles < cus || tiles > static_cast<long long>(cus) * resident_blocks(pat, best.waves)) {
    double best_time = round_time(best.rows);
    for (int k = 2; k <= 16; ++k) {
      const int rows = band_rows_for(bands_for_slots(g, static_cast<long long>(k) * cus));
      const double t = round_time(rows);
      if (t < best_time) {
        best_time = t;
        best.rows = rows;
      }
    }
  }
  return best;
}

// Geometry of one instance's tiles for strips of 64*K columns and bands of R rows.
sweep_inst make_sweep_inst(const dev_ode &p, const rkx_method &m, const sweep_cfg &c, state *st,
                           double *y0, double *y1)
{
  sweep_inst I{};
  const strip_geom g = strips_of(p.N, m.s, c.waves);
  I.N = p.N;
  I.R = sweep_band_rows(pattern_of(m), m.s, c.rows);
  I.OC = g.oc;
  I.nwide = g.nwide;
  I.split = g.split;
  I.nstrips = g.nwide + g.split;
  I.nbands = (p.N + I.R - 1) / I.R;
  I.ntiles = static_cast<int>(g.wgs(I.nbands));
  I.coef = p.coef;
  I.y0 = y0;
  I.y1 = y1;
  I.st = st;
  return I;
}

int sweep_pattern(const rkx_method &m) { return pattern_of(m); }

int sweep_resident_blocks(int pattern, int K) { return resident_blocks(pattern, K); }

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
  case 7: return launch_sweep_dopri87(K, gen, insts, tiles, ntiles, stream);
  default: return hipErrorInvalidValue;
  }
}

}  // namespace rkx
