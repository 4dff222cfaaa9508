// rkx_solve_noalloc: rkx_solve for callers that cannot allocate (code that
// runs under stream capture, tight serving loops, harnesses that forbid
// allocation).  It has no counterpart in the reference.
//
// The fused forms need two y buffers plus a few hundred bytes of step state,
// method, coefficients and instance descriptor.  Here the caller's y_dev is
// one y buffer; everything else lives in `scratch`, a file-scope device array
// of fixed size that every GPU the library runs on holds until the process
// ends.  Its device address comes from hipGetSymbolAddress, once per device
// (the array's name is only a host-side handle: host code never passes it to
// a kernel or a copy).  Steps are launched directly in chunks and the state
// is read back into host stack memory between chunks: no graph, event,
// pinned buffer or device allocation is created.
//
// Every call on a device reuses that device's array: concurrent calls on one
// device race, and each call writes all the state it reads before the first
// kernel runs (nothing is assumed to be zero from an earlier call).
#include "rkx_internal.hpp"

#include <chrono>
#include <cstring>
#include <mutex>

namespace rkx {
namespace {

constexpr size_t kScratchBytes = size_t(32) << 20;   // 32 MiB per device
constexpr size_t kHeaderBytes = 64 << 10;             // state, method, coefficients, instance
constexpr size_t kMaxDevices = 64;

static __device__ __align__(256) unsigned char scratch[kScratchBytes];

std::mutex g_mu;
unsigned char *g_ptr[kMaxDevices] = {};

// Device address of this device's scratch array (looked up once per device,
// with that device current).
int scratch_ptr(unsigned char **out)
{
  int dev = 0;
  RKX_HIP(hipGetDevice(&dev));
  if (dev < 0 || static_cast<size_t>(dev) >= kMaxDevices)
    return fail(RKX_ELIMIT, "rkx_solve_noalloc: device index beyond the supported range");
  std::lock_guard<std::mutex> lk(g_mu);
  if (!g_ptr[dev]) {
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) {
      (void) hipGetLastError();
      return fail(RKX_EDEVICE, std::string("rkx_solve_noalloc: hipGetSymbolAddress failed: ") +
                                   hipGetErrorString(e));
    }
    g_ptr[dev] = static_cast<unsigned char *>(p);
  }
  *out = g_ptr[dev];
  return RKX_OK;
}

size_t align_up(size_t x, size_t a) { return (x + a - 1) / a * a; }

}  // namespace
}  // namespace rkx

using namespace rkx;

static_assert((kScratchBytes - kHeaderBytes) / sizeof(double) == RKX_NOALLOC_MAX_COMPONENTS,
              "RKX_NOALLOC_MAX_COMPONENTS in rkx.h must match the scratch array");

extern "C" {

int rkx_solve_noalloc(const rkx_ode *p, const rkx_method *m, double t0, double te,
                      const double *y0_dev, double *y_dev, double tol, const rkx_opts *o_in,
                      rkx_stats *st_out)
{
  if (int rc = check_ode(p))
    return rc;
  if (int rc = check_method(m))
    return rc;
  if (!y0_dev || !y_dev)
    return fail(RKX_EINVAL, "rkx_solve_noalloc: null vector");
  rkx_opts o;
  if (o_in)
    o = *o_in;
  else
    rkx_opts_default(&o);
  if (o.impl != RKX_IMPL_PIPED && o.impl != RKX_IMPL_PIPEDLS)
    return fail(RKX_EINVAL, "rkx_solve_noalloc: only the fused low-storage forms "
                            "(RKX_IMPL_PIPEDLS / RKX_IMPL_PIPED) run without allocation");
  if (o.ctrl != RKX_CTRL_DEVICE && o.ctrl != RKX_CTRL_HOST)
    return fail(RKX_EINVAL, "rkx_solve_noalloc: unknown step-control mode");
  if (o.trace_cap != 0 || o.dump_path)
    return fail(RKX_EINVAL, "rkx_solve_noalloc: trace and dump output need rkx_solve");
  const dev_ode dp = make_dev_ode(*p);
  const bool chain = chain_supported(dp, *m);
  const bool pairs = !chain && pairs_supported(dp, *m);
  if (!chain && !pairs && !sweep_supported(dp, *m))
    return fail(RKX_EINVAL, "rkx_solve_noalloc: no fused kernel for this problem and method "
                            "(the D form needs full storage: use rkx_solve)");
  const size_t n = static_cast<size_t>(p->n);
  if (n > RKX_NOALLOC_MAX_COMPONENTS)
    return fail(RKX_ELIMIT, "rkx_solve_noalloc: " + std::to_string(n) +
                                " components exceed the scratch array's " +
                                std::to_string(RKX_NOALLOC_MAX_COMPONENTS) +
                                " (32 MiB per device; use rkx_solve)");
  hipStream_t stream = static_cast<hipStream_t>(o.stream);
  if (!stream)
    if (int rc = default_stream(&stream))
      return rc;
  unsigned char *base = nullptr;
  if (int rc = scratch_ptr(&base))
    return rc;

  // scratch layout: state | method | coefficients | instance | second y buffer
  size_t off = 0;
  state *st_dev = reinterpret_cast<state *>(base + off);
  off = align_up(off + sizeof(state), 256);
  dev_method *dm_dev = reinterpret_cast<dev_method *>(base + off);
  off = align_up(off + sizeof(dev_method), 256);
  dev_coef *cf_dev = reinterpret_cast<dev_coef *>(base + off);
  off = align_up(off + sizeof(dev_coef), 256);
  void *inst_dev = base + off;
  off = align_up(off + std::max(sizeof(sweep_inst), sizeof(chain_inst)), 256);
  static_assert(sizeof(state) + sizeof(dev_method) + sizeof(dev_coef) + sizeof(sweep_inst) +
                        sizeof(chain_inst) + 5 * 256 <= kHeaderBytes,
                "scratch header too small");
  double *ybuf[2] = {y_dev, reinterpret_cast<double *>(base + kHeaderBytes)};
  (void) off;

  const int s = m->s;
  dev_method dm;
  fill_dev_method(*m, dm);
  state hs{};
  hs.t = t0;
  hs.te = te;
  hs.tol = tol;
  hs.h = initial_h(te - t0, tol, o.step_limit, o.h0);
  hs.step_limit = o.step_limit;
  hs.ord = m->ord;
  hs.ctrl = o.ctrl;
  hs.s = s;
  hs.meth = dm_dev;
  hs.coef = cf_dev;
  hs.cur = 0;
  hs.done = !(t0 < te);
  dev_coef hc{};
  premult(dm, s, hs.h, hc);

  sweep_cfg cfg{1, 1};
  const int pattern = sweep_pattern(*m);
  int ntiles = 0;
  bool gen_rows = false;   // the general ghost-row form (sweep_inst_single_reflect)
  sweep_inst si{};
  chain_inst ci{};
  if (chain) {
    ci = make_chain_inst(dp, *m, st_dev, ybuf[0], ybuf[1]);
    ntiles = ci.nwg;
  } else if (!pairs) {
    cfg = sweep_default_cfg(dp, *m);
    if (o.tile_rows > 0)
      cfg.rows = o.tile_rows;
    if (o.tile_waves > 0)
      cfg.waves = o.tile_waves;
    if (cfg.waves < 1 || cfg.waves > 4 || cfg.rows < 1 || 64 * cfg.waves <= 2 * s ||
        sweep_resident_blocks(pattern, cfg.waves) < 1)
      return fail(RKX_EINVAL, "rkx_solve_noalloc: tile_waves must fit this table's fused kernel "
                              "(1..4, DOPRI87 1..2, wider than 2s), tile_rows >= 1");
    si = make_sweep_inst(dp, *m, cfg, st_dev, ybuf[0], ybuf[1]);
    ntiles = si.ntiles;
    gen_rows = !sweep_inst_single_reflect(pattern, s, si);
  }
  // everything the kernels read is written before the first launch
  if (y0_dev != y_dev)
    RKX_HIP(hipMemcpyAsync(ybuf[0], y0_dev, n * sizeof(double), hipMemcpyDeviceToDevice, stream));
  RKX_HIP(hipMemcpyAsync(dm_dev, &dm, sizeof(dm), hipMemcpyHostToDevice, stream));
  RKX_HIP(hipMemcpyAsync(cf_dev, &hc, sizeof(hc), hipMemcpyHostToDevice, stream));
  RKX_HIP(hipMemcpyAsync(st_dev, &hs, sizeof(hs), hipMemcpyHostToDevice, stream));
  if (chain)
    RKX_HIP(hipMemcpyAsync(inst_dev, &ci, sizeof(ci), hipMemcpyHostToDevice, stream));
  else if (!pairs)
    RKX_HIP(hipMemcpyAsync(inst_dev, &si, sizeof(si), hipMemcpyHostToDevice, stream));
  RKX_HIP(hipStreamSynchronize(stream));

  auto step = [&]() -> hipError_t {
    if (pairs)
      return launch_lv_step(dp, cf_dev, s, st_dev, ybuf[0], ybuf[1], stream);
    if (chain)
      return launch_chain(pattern, static_cast<const chain_inst *>(inst_dev), nullptr, ntiles,
                          stream);
    return launch_sweep(pattern, cfg.waves, gen_rows, static_cast<const sweep_inst *>(inst_dev),
                        nullptr, ntiles, stream);
  };

  const auto t_start = std::chrono::steady_clock::now();
  state cur = hs;
  if (!cur.done && o.ctrl == RKX_CTRL_HOST) {
    while (!cur.done) {
      RKX_HIP(step());
      RKX_HIP(hipMemcpyAsync(&cur, st_dev, sizeof(state), hipMemcpyDeviceToHost, stream));
      RKX_HIP(hipStreamSynchronize(stream));
      host_finish(cur, dm, hc);
      RKX_HIP(hipMemcpyAsync(cf_dev, &hc, sizeof(hc), hipMemcpyHostToDevice, stream));
      RKX_HIP(hipMemcpyAsync(st_dev, &cur, sizeof(state), hipMemcpyHostToDevice, stream));
      RKX_HIP(hipStreamSynchronize(stream));
    }
  } else if (!cur.done) {
    // chunks of directly launched steps; kernels past the end return at once
    constexpr int kChunk = 32;
    int last = -1, stalled = 0;
    for (;;) {
      for (int k = 0; k < kChunk; ++k)
        RKX_HIP(step());
      RKX_HIP(hipMemcpyAsync(&cur, st_dev, sizeof(state), hipMemcpyDeviceToHost, stream));
      RKX_HIP(hipStreamSynchronize(stream));
      if (cur.done)
        break;
      const int att = cur.acc + cur.rej;
      stalled = (att == last) ? stalled + 1 : 0;
      last = att;
      if (stalled >= 3)
        return fail(RKX_EDEVICE, "rkx_solve_noalloc: device step control made no progress");
    }
  }
  const auto t_end = std::chrono::steady_clock::now();
  if (cur.cur == 1)
    RKX_HIP(hipMemcpyAsync(y_dev, ybuf[1], n * sizeof(double), hipMemcpyDeviceToDevice, stream));
  RKX_HIP(hipStreamSynchronize(stream));
  if (st_out) {
    st_out->steps_acc = cur.acc;
    st_out->steps_rej = cur.rej;
    st_out->t = cur.t;
    st_out->h = cur.h;
    st_out->kernel_time = std::chrono::duration<double>(t_end - t_start).count();
    st_out->resident_bytes = kHeaderBytes + n * sizeof(double);
  }
  return RKX_OK;
}

}  // extern "C"
