// Batched solves: many instances of uneven size advanced together, each with
// its own t, h, accept/reject and step limit (include/rkx.h, rkx_batch_*).
//
// Instances with a fused kernel (BRUSS2D row sweep, LV chain segments; low
// storage) are grouped by (method pattern, kernel and strip width); one
// launch per group covers the tiles of all
// of the group's active instances through an (instance, tile) table, and the
// last tile of each instance runs that instance's step control.  Instances
// without a fused kernel (full storage) run the stage kernels of the D form.
// The host polls the instances' states every few batched steps, drops
// finished instances from the tile tables and lets callers add new ones in
// between (continuous batching).  An instance's arithmetic depends only on
// its own data, so its result is bitwise the same alone or in any mix.
#include "rkx_internal.hpp"

#include <algorithm>
#include <cstring>
#include <map>
#include <utility>
#include <vector>

namespace rkx {
namespace {

constexpr int kBatchCap = 1024;       // instances per batch
constexpr int kPollSteps = 16;        // batched steps between state polls

struct batch_inst {
  bool used = false;
  bool done = false;
  int storage = RKX_STORAGE_LOW;
  int pattern = 0;
  int K = 0;                            // strip waves of the row sweep; 0: chain kernel;
                                        // -1: LV-pairs step kernel (its own launch)
  rkx_ode ode{};
  rkx_method m{};
  dev_ode dp{};
  dev_method dm{};
  sweep_cfg base{1, 1};                 // tile shape chosen at admission
  sweep_cfg cfg{1, 1};                  // current shape (bands shortened as the batch drains)
  double *y[2] = {nullptr, nullptr};
  double *w = nullptr, *dy = nullptr, *err = nullptr;
  size_t bytes = 0;
  state snap{};
};

struct tile_group {
  int pattern = 0, K = 0;
  bool gen = false;      // some instance needs the general ghost-row form
  std::vector<int2> tiles;
  int2 *dev = nullptr;
  size_t dev_cap = 0;
};

}  // namespace
}  // namespace rkx

using namespace rkx;

struct rkx_batch {
  hipStream_t stream = nullptr;
  rkx_batch_opts opts{};
  state *st_dev = nullptr;
  dev_coef *cf_dev = nullptr;
  dev_method *dm_dev = nullptr;
  sweep_inst *si_dev = nullptr;
  chain_inst *ci_dev = nullptr;
  state *st_pinned = nullptr;          // 2 x kBatchCap: snapshots of two chunks in flight
  hipEvent_t ev[2] = {nullptr, nullptr};
  hipGraph_t graph = nullptr;           // kPollSteps batched steps, captured once per table set
  hipGraphExec_t graph_exec = nullptr;
  bool graph_ok = false;
  int cus = 256;                        // compute units of the device
  std::vector<batch_inst> inst;
  std::map<std::pair<int, int>, tile_group> groups;
  size_t bytes = 0;
  bool dirty = true;
  int hi_slot = 0;                      // one past the highest slot ever used
};

namespace {

void free_inst(batch_inst &b)
{
  for (double *p : {b.y[0], b.y[1], b.w, b.dy, b.err})
    if (p)
      (void) hipFree(p);
  b = batch_inst{};
}

// Strip width and band height of a batched BRUSS2D instance: one wavefront
// per component for grids that fit a 64-column strip, else two (measured
// fastest per lane when many instances share the CUs), and bands of 64 rows
// (long enough for the s-1 halo rows to be a small share, short enough that
// large instances split into many tiles).  The options override either; a
// strip width wider than the table's kernel is built for (DOPRI87: 2 waves)
// is narrowed to the widest it has.
sweep_cfg batch_cfg(const dev_ode &p, const rkx_method &m, const rkx_batch_opts &o)
{
  sweep_cfg c;
  c.waves = (p.N + 2 * m.s <= 64) ? 1 : 2;
  c.rows = 64;
  if (o.tile_waves > 0)
    c.waves = std::min(o.tile_waves, 4);
  if (o.tile_rows > 0)
    c.rows = o.tile_rows;
  const int pattern = sweep_pattern(m);
  while (c.waves > 1 && sweep_resident_blocks(pattern, c.waves) < 1)
    --c.waves;
  while (c.waves < 4 && 64 * c.waves <= 2 * m.s)
    ++c.waves;
  return c;
}

// Band height as the batch drains.  The admitted shape keeps halo rows a
// small share while many instances fill the CUs; as instances finish, the
// remaining fused instances get bands of a half, a quarter, ... of it (down
// to 8 rows), so that a few large instances still spread over the chip
// instead of a few workgroups sweeping all their rows.  The divisor is the
// one with the shortest modelled batched step: per (table, strip width)
// launch, the row iterations of all its tiles shared over the resident
// workgroup slots, but never less than its longest tile (bands recompute
// 2(s-1) halo rows, so shorter bands cost more work in total and take less
// time each).  The tiling never changes an instance's arithmetic, only how
// its rows are shared out.
int adapt_bands(rkx_batch *b)
{
  auto rows_at = [](const batch_inst &I, int div) {
    return std::max(8, I.base.rows / div);   // bands of at least 8 rows
  };
  auto step_time = [&](int div) {
    std::map<std::pair<int, int>, std::pair<double, double>> g;   // work, longest tile
    for (int i = 0; i < b->hi_slot; ++i) {
      const batch_inst &I = b->inst[static_cast<size_t>(i)];
      if (!I.used || I.done || I.storage != RKX_STORAGE_LOW || I.K <= 0)
        continue;
      const sweep_cfg c{I.base.waves, rows_at(I, div)};
      const sweep_inst si = make_sweep_inst(I.dp, I.m, c, nullptr, nullptr, nullptr);
      const int band = sweep_band_rows(I.pattern, I.m.s, std::min(si.R, I.dp.N));
      const double iters = band + 2.0 * (I.m.s - 1);
      auto &e = g[{I.pattern, I.K}];
      e.first += static_cast<double>(si.ntiles) * iters;
      e.second = std::max(e.second, iters);
    }
    double t = 0.0;
    for (const auto &kv : g) {
      const int slots = b->cus * std::max(1, sweep_resident_blocks(kv.first.first,
                                                                   kv.first.second));
      t += std::max(kv.second.first / slots, kv.second.second);
    }
    return t;
  };
  int div = 1;
  double best = step_time(1);
  for (int d = 2; d <= 64; d *= 2) {
    const double t = step_time(d);
    if (t < best) {
      best = t;
      div = d;
    }
  }
  for (int i = 0; i < b->hi_slot; ++i) {
    batch_inst &I = b->inst[static_cast<size_t>(i)];
    if (!I.used || I.done || I.storage != RKX_STORAGE_LOW || I.K <= 0)
      continue;
    const sweep_cfg c{I.base.waves, rows_at(I, div)};
    if (c.rows == I.cfg.rows)
      continue;
    I.cfg = c;
    const sweep_inst si = make_sweep_inst(I.dp, I.m, c, b->st_dev + i, I.y[0], I.y[1]);
    RKX_HIP(hipMemcpyAsync(b->si_dev + i, &si, sizeof(si), hipMemcpyHostToDevice, b->stream));
  }
  return RKX_OK;
}

// Row iterations of one tile of a fused instance (its band, or the shorter
// last band, plus the 2(s-1) halo rows).
int tile_iters(const batch_inst &I, int t)
{
  const sweep_inst si = make_sweep_inst(I.dp, I.m, I.cfg, nullptr, nullptr, nullptr);
  return sweep_tile_iters(I.pattern, I.m.s, si, t);
}

// Longest tiles first, so that a batched step ends on short tiles instead of
// leaving CUs idle behind a long one.  The sweep kernel gives each of the 8
// XCDs a contiguous run of the table (workgroup b -> run b % 8, in order), so
// the sorted tiles are dealt round-robin over the runs: every XCD works
// through a mix, longest first.
void order_tiles(const rkx_batch *b, std::vector<int2> &tiles)
{
  std::vector<std::pair<int, int2>> keyed;
  keyed.reserve(tiles.size());
  for (const int2 &t : tiles)
    keyed.push_back({tile_iters(b->inst[static_cast<size_t>(t.x)], t.y), t});
  std::stable_sort(keyed.begin(), keyed.end(),
                   [](const auto &a, const auto &c) { return a.first > c.first; });
  const size_t T = tiles.size(), q = T / 8, rm = T % 8;
  for (size_t j = 0; j < T; ++j) {
    const size_t x = j % 8, pos = j / 8;
    tiles[x * q + std::min(x, rm) + pos] = keyed[j].second;
  }
}

int rebuild_tables(rkx_batch *b)
{
  if (int rc = adapt_bands(b))
    return rc;
  for (auto &kv : b->groups) {
    kv.second.tiles.clear();
    kv.second.gen = false;
  }
  for (int i = 0; i < b->hi_slot; ++i) {
    const batch_inst &I = b->inst[static_cast<size_t>(i)];
    if (!I.used || I.done || I.storage != RKX_STORAGE_LOW || I.K < 0)
      continue;
    tile_group &g = b->groups[{I.pattern, I.K}];
    g.pattern = I.pattern;
    g.K = I.K;
    int nt = 0;
    if (I.K == 0) {
      nt = make_chain_inst(I.dp, I.m, nullptr, nullptr, nullptr).nwg;
    } else {
      const sweep_inst si = make_sweep_inst(I.dp, I.m, I.cfg, nullptr, nullptr, nullptr);
      nt = si.ntiles;
      if (!sweep_inst_single_reflect(I.pattern, I.m.s, si))
        g.gen = true;
    }
    for (int t = 0; t < nt; ++t)
      g.tiles.push_back(make_int2(i, t));
  }
  for (auto &kv : b->groups) {
    tile_group &g = kv.second;
    if (g.K != 0)
      order_tiles(b, g.tiles);
    if (g.tiles.size() > g.dev_cap) {
      if (g.dev)
        (void) hipFree(g.dev);
      g.dev = nullptr;
      g.dev_cap = std::max<size_t>(g.tiles.size(), 2 * g.dev_cap);
      if (hipMalloc(reinterpret_cast<void **>(&g.dev), g.dev_cap * sizeof(int2)) != hipSuccess) {
        (void) hipGetLastError();
        g.dev_cap = 0;
        return fail(RKX_ENOMEM, "rkx_batch: tile table allocation failed");
      }
    }
    if (!g.tiles.empty())
      RKX_HIP(hipMemcpyAsync(g.dev, g.tiles.data(), g.tiles.size() * sizeof(int2),
                             hipMemcpyHostToDevice, b->stream));
  }
  RKX_HIP(hipStreamSynchronize(b->stream));
  b->dirty = false;
  return RKX_OK;
}

int enqueue_batched_step(rkx_batch *b)
{
  for (auto &kv : b->groups) {
    const tile_group &g = kv.second;
    if (g.tiles.empty())
      continue;
    if (g.K == 0)
      RKX_HIP(launch_chain(g.pattern, b->ci_dev, g.dev, static_cast<int>(g.tiles.size()),
                           b->stream));
    else
      RKX_HIP(launch_sweep(g.pattern, g.K, g.gen, b->si_dev, g.dev,
                           static_cast<int>(g.tiles.size()), b->stream));
  }
  for (int i = 0; i < b->hi_slot; ++i) {
    batch_inst &I = b->inst[static_cast<size_t>(i)];
    if (!I.used || I.done)
      continue;
    if (I.storage == RKX_STORAGE_LOW && I.K < 0)
      RKX_HIP(launch_lv_step(I.dp, b->cf_dev + i, I.m.s, b->st_dev + i, I.y[0], I.y[1],
                             b->stream));
    else if (I.storage == RKX_STORAGE_FULL)
      RKX_HIP(launch_d_step(I.dp, b->cf_dev + i, I.m.s, b->st_dev + i, I.y[0], I.y[1], I.w,
                            I.dy, I.err, b->stream));
  }
  return RKX_OK;
}

void drop_graph(rkx_batch *b)
{
  if (b->graph_exec)
    (void) hipGraphExecDestroy(b->graph_exec);
  if (b->graph)
    (void) hipGraphDestroy(b->graph);
  b->graph_exec = nullptr;
  b->graph = nullptr;
  b->graph_ok = false;
}

// Capture kPollSteps batched steps of the current tables as one graph.  The
// capture is always ended, also when a launch fails inside it.
int capture_chunk(rkx_batch *b)
{
  drop_graph(b);
  RKX_HIP(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
  int rc = RKX_OK;
  for (int j = 0; j < kPollSteps && rc == RKX_OK; ++j)
    rc = enqueue_batched_step(b);
  hipGraph_t g = nullptr;
  const hipError_t e = hipStreamEndCapture(b->stream, &g);
  if (rc != RKX_OK) {
    if (g)
      (void) hipGraphDestroy(g);
    return rc;
  }
  RKX_HIP(e);
  b->graph = g;
  RKX_HIP(hipGraphInstantiate(&b->graph_exec, g, nullptr, nullptr, 0));
  b->graph_ok = true;
  return RKX_OK;
}

// Enqueue k batched steps and a snapshot of every instance's state into
// snapshot buffer `slot`; full chunks replay the captured graph.
int launch_chunk(rkx_batch *b, int slot, int k)
{
  if (b->dirty) {
    if (int rc = rebuild_tables(b))
      return rc;
    b->graph_ok = false;
  }
  if (k == kPollSteps) {
    if (!b->graph_ok)
      if (int rc = capture_chunk(b))
        return rc;
    RKX_HIP(hipGraphLaunch(b->graph_exec, b->stream));
  } else {
    for (int j = 0; j < k; ++j)
      if (int rc = enqueue_batched_step(b))
        return rc;
  }
  RKX_HIP(hipMemcpyAsync(b->st_pinned + static_cast<size_t>(slot) * kBatchCap, b->st_dev,
                         sizeof(state) * b->hi_slot, hipMemcpyDeviceToHost, b->stream));
  RKX_HIP(hipEventRecord(b->ev[slot], b->stream));
  return RKX_OK;
}

// Read the snapshot of `slot` once its chunk has finished: update the host
// copies and count newly finished instances.
int read_snapshot(rkx_batch *b, int slot, int *newly_done)
{
  RKX_HIP(hipEventSynchronize(b->ev[slot]));
  const state *snap = b->st_pinned + static_cast<size_t>(slot) * kBatchCap;
  for (int i = 0; i < b->hi_slot; ++i) {
    batch_inst &I = b->inst[static_cast<size_t>(i)];
    if (!I.used)
      continue;
    const state &s = snap[i];
    I.snap = s;
    if (s.done && !I.done) {
      I.done = true;
      ++*newly_done;
    }
  }
  return RKX_OK;
}

}  // namespace

extern "C" {

void rkx_batch_opts_default(rkx_batch_opts *o)
{
  if (!o)
    return;
  std::memset(o, 0, sizeof(*o));
  o->ctrl = RKX_CTRL_DEVICE;
}

int rkx_batch_create(rkx_batch **out, const rkx_batch_opts *o)
{
  if (!out)
    return fail(RKX_EINVAL, "rkx_batch_create: null output");
  rkx_batch_opts opts;
  if (o)
    opts = *o;
  else
    rkx_batch_opts_default(&opts);
  if (opts.ctrl != RKX_CTRL_DEVICE)
    return fail(RKX_EINVAL, "rkx_batch_create: ctrl must be RKX_CTRL_DEVICE");
  rkx_batch *b = new rkx_batch();
  b->opts = opts;
  b->stream = static_cast<hipStream_t>(opts.stream);
  int rc = RKX_OK;
  if (!b->stream)
    rc = default_stream(&b->stream);
  auto dalloc = [&](void **p, size_t bytes) {
    if (rc)
      return;
    if (hipMalloc(p, bytes) != hipSuccess) {
      (void) hipGetLastError();
      rc = fail(RKX_ENOMEM, "rkx_batch_create: device allocation failed");
    }
  };
  dalloc(reinterpret_cast<void **>(&b->st_dev), sizeof(state) * kBatchCap);
  dalloc(reinterpret_cast<void **>(&b->cf_dev), sizeof(dev_coef) * kBatchCap);
  dalloc(reinterpret_cast<void **>(&b->dm_dev), sizeof(dev_method) * kBatchCap);
  dalloc(reinterpret_cast<void **>(&b->si_dev), sizeof(sweep_inst) * kBatchCap);
  dalloc(reinterpret_cast<void **>(&b->ci_dev), sizeof(chain_inst) * kBatchCap);
  for (int k = 0; k < 2 && !rc; ++k)
    if (hipEventCreateWithFlags(&b->ev[k], hipEventDisableTiming) != hipSuccess) {
      (void) hipGetLastError();
      rc = fail(RKX_EDEVICE, "rkx_batch_create: event creation failed");
    }
  if (!rc && hipHostMalloc(reinterpret_cast<void **>(&b->st_pinned), 2 * sizeof(state) * kBatchCap,
                           hipHostMallocDefault) != hipSuccess) {
    (void) hipGetLastError();
    rc = fail(RKX_ENOMEM, "rkx_batch_create: pinned allocation failed");
  }
  if (rc) {
    rkx_batch_destroy(b);
    return rc;
  }
  b->inst.resize(kBatchCap);
  int dev = 0, cus = 0;
  if (hipGetDevice(&dev) == hipSuccess &&
      hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) == hipSuccess &&
      cus > 0)
    b->cus = cus;
  else
    (void) hipGetLastError();
  *out = b;
  return RKX_OK;
}

int rkx_batch_destroy(rkx_batch *b)
{
  if (!b)
    return RKX_OK;
  if (b->stream)
    (void) hipStreamSynchronize(b->stream);
  for (auto &I : b->inst)
    if (I.used)
      free_inst(I);
  for (auto &kv : b->groups)
    if (kv.second.dev)
      (void) hipFree(kv.second.dev);
  for (void *p : {static_cast<void *>(b->st_dev), static_cast<void *>(b->cf_dev),
                  static_cast<void *>(b->dm_dev), static_cast<void *>(b->si_dev),
                  static_cast<void *>(b->ci_dev)})
    if (p)
      (void) hipFree(p);
  if (b->st_pinned)
    (void) hipHostFree(b->st_pinned);
  drop_graph(b);
  for (hipEvent_t e : b->ev)
    if (e)
      (void) hipEventDestroy(e);
  delete b;
  return RKX_OK;
}

size_t rkx_batch_footprint(const rkx_ode *p, const rkx_method *m, int storage)
{
  if (check_ode(p) || check_method(m))
    return 0;
  const size_t vec = static_cast<size_t>(p->n) * sizeof(double);
  const dev_ode dp = make_dev_ode(*p);
  if (storage == RKX_STORAGE_AUTO)
    storage = (sweep_supported(dp, *m) || chain_supported(dp, *m) || pairs_supported(dp, *m))
                  ? RKX_STORAGE_LOW
                  : RKX_STORAGE_FULL;
  if (storage == RKX_STORAGE_LOW)
    return 2 * vec;
  return (2 + (m->s - 1) + 2) * vec;
}

int rkx_batch_add(rkx_batch *b, const rkx_ode *p, const rkx_method *m, double t0, double te,
                  const double *y0, double tol, double h0, int step_limit, int storage)
{
  if (!b || !y0)
    return fail(RKX_EINVAL, "rkx_batch_add: null argument");
  if (int rc = check_ode(p))
    return rc;
  if (int rc = check_method(m))
    return rc;
  const dev_ode dp = make_dev_ode(*p);
  const bool chain = chain_supported(dp, *m);
  const bool pairs = !chain && pairs_supported(dp, *m);
  const bool fused = chain || pairs || sweep_supported(dp, *m);
  if (storage == RKX_STORAGE_AUTO)
    storage = fused ? RKX_STORAGE_LOW : RKX_STORAGE_FULL;
  if (storage == RKX_STORAGE_LOW && !fused)
    return fail(RKX_EINVAL, "rkx_batch_add: no low-storage (fused) kernel for this problem/method");
  if (storage != RKX_STORAGE_LOW && storage != RKX_STORAGE_FULL)
    return fail(RKX_EINVAL, "rkx_batch_add: unknown storage mode");
  const size_t need = rkx_batch_footprint(p, m, storage);
  if (b->opts.hbm_budget > 0 && b->bytes + need > b->opts.hbm_budget)
    return fail(RKX_ELIMIT, "rkx_batch_add: HBM budget exhausted");
  int slot = -1;
  for (int i = 0; i < kBatchCap; ++i)
    if (!b->inst[static_cast<size_t>(i)].used) {
      slot = i;
      break;
    }
  if (slot < 0)
    return fail(RKX_ELIMIT, "rkx_batch_add: instance table full");
  batch_inst &I = b->inst[static_cast<size_t>(slot)];
  I = batch_inst{};
  I.storage = storage;
  I.ode = *p;
  I.m = *m;
  I.dp = dp;
  fill_dev_method(*m, I.dm);
  const size_t vec = static_cast<size_t>(p->n) * sizeof(double);
  auto dalloc = [&](double **q, size_t bytes) -> int {
    if (hipMalloc(reinterpret_cast<void **>(q), bytes) != hipSuccess) {
      (void) hipGetLastError();
      return fail(RKX_ENOMEM, "rkx_batch_add: device allocation failed");
    }
    return RKX_OK;
  };
  int rc = dalloc(&I.y[0], vec);
  if (!rc)
    rc = dalloc(&I.y[1], vec);
  if (!rc && storage == RKX_STORAGE_FULL) {
    rc = dalloc(&I.w, vec * (m->s - 1));
    if (!rc)
      rc = dalloc(&I.dy, vec);
    if (!rc)
      rc = dalloc(&I.err, vec);
  }
  if (rc) {
    free_inst(I);
    return rc;
  }
  I.used = true;
  I.bytes = need;

  state hs{};
  hs.t = t0;
  hs.te = te;
  hs.tol = tol;
  hs.h = initial_h(te - t0, tol, step_limit, h0);
  hs.step_limit = step_limit;
  hs.ord = m->ord;
  hs.ctrl = b->opts.ctrl;
  hs.s = m->s;
  hs.meth = b->dm_dev + slot;
  hs.coef = b->cf_dev + slot;
  hs.done = !(t0 < te);
  dev_coef hc{};
  premult(I.dm, m->s, hs.h, hc);
  I.snap = hs;
  I.done = hs.done;
  if (storage == RKX_STORAGE_LOW && pairs) {
    I.pattern = 0;
    I.K = -1;
  } else if (storage == RKX_STORAGE_LOW && chain) {
    I.pattern = sweep_pattern(*m);
    I.K = 0;
    const chain_inst ci =
        make_chain_inst(dp, *m, b->st_dev + slot, I.y[0], I.y[1]);
    RKX_HIP(hipMemcpyAsync(b->ci_dev + slot, &ci, sizeof(ci), hipMemcpyHostToDevice, b->stream));
  } else if (storage == RKX_STORAGE_LOW) {
    I.pattern = sweep_pattern(*m);
    I.cfg = batch_cfg(dp, *m, b->opts);
    I.base = I.cfg;
    I.K = I.cfg.waves;
    const sweep_inst si =
        make_sweep_inst(dp, *m, I.cfg, b->st_dev + slot, I.y[0], I.y[1]);
    RKX_HIP(hipMemcpyAsync(b->si_dev + slot, &si, sizeof(si), hipMemcpyHostToDevice, b->stream));
  }
  RKX_HIP(hipMemcpyAsync(b->dm_dev + slot, &I.dm, sizeof(I.dm), hipMemcpyHostToDevice, b->stream));
  RKX_HIP(hipMemcpyAsync(b->cf_dev + slot, &hc, sizeof(hc), hipMemcpyHostToDevice, b->stream));
  RKX_HIP(hipMemcpyAsync(b->st_dev + slot, &hs, sizeof(hs), hipMemcpyHostToDevice, b->stream));
  RKX_HIP(hipMemcpyAsync(I.y[0], y0, vec, hipMemcpyHostToDevice, b->stream));
  RKX_HIP(hipStreamSynchronize(b->stream));
  b->bytes += need;
  b->hi_slot = std::max(b->hi_slot, slot + 1);
  b->dirty = true;
  return slot;
}

int rkx_batch_run(rkx_batch *b, int max_steps, int *active)
{
  if (!b || max_steps < 0)
    return fail(RKX_EINVAL, "rkx_batch_run: bad argument");
  auto count_active = [&]() {
    int a = 0;
    for (int i = 0; i < b->hi_slot; ++i)
      a += (b->inst[static_cast<size_t>(i)].used && !b->inst[static_cast<size_t>(i)].done);
    return a;
  };
  // Chunks of kPollSteps batched steps (a captured graph), two in flight:
  // chunk c+1 and its state snapshot are enqueued before the host reads the
  // snapshot of chunk c.  Within a run the active set only shrinks, and the
  // tiles of a finished instance return at once, so a chunk enqueued before
  // the host saw an instance finish is still correct; the tables (and the
  // graph) are rebuilt before the next chunk after one does.
  int enqueued = 0;
  bool pending[2] = {false, false};
  int cur = 0;
  if (max_steps > 0 && count_active() > 0) {
    const int k = std::min(kPollSteps, max_steps);
    if (int rc = launch_chunk(b, cur, k))
      return rc;
    enqueued += k;
    pending[cur] = true;
  }
  while (pending[cur]) {
    const int nxt = cur ^ 1;
    if (enqueued < max_steps && count_active() > 0) {
      const int k = std::min(kPollSteps, max_steps - enqueued);
      if (int rc = launch_chunk(b, nxt, k))
        return rc;
      enqueued += k;
      pending[nxt] = true;
    }
    int newly = 0;
    if (int rc = read_snapshot(b, cur, &newly))
      return rc;
    pending[cur] = false;
    if (newly)
      b->dirty = true;
    cur = nxt;
  }
  RKX_HIP(hipStreamSynchronize(b->stream));
  if (active)
    *active = count_active();
  return RKX_OK;
}

int rkx_batch_status(rkx_batch *b, int id, rkx_stats *st, int *done)
{
  if (!b || id < 0 || id >= kBatchCap || !b->inst[static_cast<size_t>(id)].used)
    return fail(RKX_EINVAL, "rkx_batch_status: unknown instance");
  const batch_inst &I = b->inst[static_cast<size_t>(id)];
  if (st) {
    std::memset(st, 0, sizeof(*st));
    st->steps_acc = I.snap.acc;
    st->steps_rej = I.snap.rej;
    st->t = I.snap.t;
    st->h = I.snap.h;
    st->resident_bytes = I.bytes;
  }
  if (done)
    *done = I.done ? 1 : 0;
  return RKX_OK;
}

int rkx_batch_result(rkx_batch *b, int id, double *y)
{
  if (!b || !y || id < 0 || id >= kBatchCap || !b->inst[static_cast<size_t>(id)].used)
    return fail(RKX_EINVAL, "rkx_batch_result: unknown instance");
  const batch_inst &I = b->inst[static_cast<size_t>(id)];
  RKX_HIP(hipStreamSynchronize(b->stream));
  state s;
  RKX_HIP(hipMemcpy(&s, b->st_dev + id, sizeof(state), hipMemcpyDeviceToHost));
  RKX_HIP(hipMemcpy(y, I.y[s.cur], static_cast<size_t>(I.ode.n) * sizeof(double),
                    hipMemcpyDeviceToHost));
  return RKX_OK;
}

int rkx_batch_remove(rkx_batch *b, int id)
{
  if (!b || id < 0 || id >= kBatchCap || !b->inst[static_cast<size_t>(id)].used)
    return fail(RKX_EINVAL, "rkx_batch_remove: unknown instance");
  RKX_HIP(hipStreamSynchronize(b->stream));
  batch_inst &I = b->inst[static_cast<size_t>(id)];
  b->bytes -= I.bytes;
  free_inst(I);
  b->dirty = true;
  return RKX_OK;
}

size_t rkx_batch_bytes(const rkx_batch *b) { return b ? b->bytes : 0; }

int rkx_batch_tile(rkx_batch *b, int id, int *waves, int *rows)
{
  if (!b || !waves || !rows || id < 0 || id >= kBatchCap || !b->inst[static_cast<size_t>(id)].used)
    return fail(RKX_EINVAL, "rkx_batch_tile: unknown instance");
  const batch_inst &I = b->inst[static_cast<size_t>(id)];
  *waves = 0;
  *rows = 0;
  if (I.storage == RKX_STORAGE_LOW && I.K > 0) {
    *waves = I.cfg.waves;
    *rows = make_sweep_inst(I.dp, I.m, I.cfg, nullptr, nullptr, nullptr).R;
  }
  return RKX_OK;
}

}  // extern "C"
