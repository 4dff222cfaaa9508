// rkx host runtime: the C ABI of include/rkx.h over HIP.
//
// The step loop of the reference's solver() (src/impl/seq/D.c:58-78) runs on
// the device: every step's kernels read t, h and the current-buffer index from
// a device-resident state and the last workgroup of the step's final kernel
// decides accept/reject and the next h.  The host only enqueues chunks of
// steps, captured once as a hipGraph, and polls the state between chunks;
// kernels of a chunk that run past the end of the integration return at once.
#include "rkx_internal.hpp"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <atomic>
#include <mutex>
#include <thread>
#include <string>
#include <vector>

namespace rkx {

namespace {
thread_local std::string g_last_error;
}

int fail(int code, const std::string &msg)
{
  g_last_error = msg;
  return code;
}

dev_ode make_dev_ode(const rkx_ode &p)
{
  dev_ode d{};
  d.kind = p.kind;
  d.n = p.n;
  d.cells = p.n / 2;
  d.N = (p.kind == RKX_BRUSS2D) ? p.grid : 0;
  if (p.kind == RKX_BRUSS2D) {
    const double N1 = static_cast<double>(p.grid) - 1.0;
    d.coef = p.alpha * N1 * N1;   // ((alpha*N1)*N1), bruss2d-mix.c:77
  }
  d.c0 = p.c0;
  d.d0 = p.d0;
  d.d1 = p.d1;
  d.c1 = p.c1;
  d.diff = p.diff;
  return d;
}

namespace {
hipStream_t g_stream = nullptr;
std::mutex g_init_mu;
}  // namespace

int default_stream(hipStream_t *out)
{
  std::lock_guard<std::mutex> lk(g_init_mu);
  if (!g_stream) {
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1)
      return fail(RKX_EDEVICE, "no HIP device available");
    RKX_HIP(hipStreamCreateWithFlags(&g_stream, hipStreamNonBlocking));
  }
  *out = g_stream;
  return RKX_OK;
}

int check_ode(const rkx_ode *p)
{
  if (!p)
    return fail(RKX_EINVAL, "null problem descriptor");
  switch (p->kind) {
  case RKX_BRUSS2D:
    if (p->grid < 2 || p->n != 2 * p->grid * p->grid)
      return fail(RKX_EINVAL, "BRUSS2D needs N >= 2 and n = 2*N*N");
    if (p->grid > 32767)
      return fail(RKX_ELIMIT, "BRUSS2D grid above 32767 (n = 2N^2 must fit in int)");
    break;
  case RKX_LV:
    if (p->grid < 1 || p->n != 2 * p->grid)
      return fail(RKX_EINVAL, "LV needs pairs >= 1 and n = 2*pairs");
    break;
  case RKX_LV_CHAIN:
    if (p->grid < 3 || p->n != 2 * p->grid)
      return fail(RKX_EINVAL, "LV chain needs cells >= 3 and n = 2*cells");
    break;
  default:
    return fail(RKX_EINVAL, "unknown problem kind");
  }
  return RKX_OK;
}

int check_method(const rkx_method *m)
{
  if (!m || m->s < 2 || m->s > RKX_MAX_STAGES)
    return fail(RKX_EINVAL, "method needs 2 <= s <= 16 stages");
  for (int i = 0; i < m->s; ++i)
    for (int j = i; j < m->s; ++j)
      if (m->A[i][j] != 0.0)
        return fail(RKX_EINVAL, "method is not explicit (A[i][j] != 0 for j >= i)");
  return RKX_OK;
}

void fill_dev_method(const rkx_method &m, dev_method &d)
{
  std::memset(&d, 0, sizeof(d));
  for (int i = 0; i < m.s; ++i) {
    for (int j = 0; j < i; ++j)
      d.A[i * RKX_MAX_STAGES + j] = m.A[i][j];
    d.b[i] = m.b[i];
    d.e[i] = m.b[i] - m.b_hat[i];      // seq/D.c:40-41
    d.c[i] = m.c[i];
  }
  for (int i = 0; i < m.s; ++i) {
    for (int l = i + 1; l < m.s; ++l)
      d.col[i][l - i - 1] = d.A[l * RKX_MAX_STAGES + i];
    d.col[i][m.s - i - 1] = d.b[i];
    d.col[i][m.s - i] = d.e[i];
  }
}

// initial_stepsize (src/tools.c:25-36)
double initial_h(double H, double tol, int limit, double h0)
{
  if (h0 > 0.0)
    return h0;
  return tol > 0.0 ? std::fmin(10.0 * tol, H / static_cast<double>(limit))
                   : H / static_cast<double>(limit);
}

namespace {

// All device memory of one solve, released on scope exit.
struct solve_mem {
  std::vector<void *> ptrs;
  ~solve_mem()
  {
    for (void *p : ptrs)
      (void) hipFree(p);
  }
  int alloc(void **out, size_t bytes)
  {
    void *p = nullptr;
    if (hipMalloc(&p, bytes) != hipSuccess) {
      (void) hipGetLastError();
      return fail(RKX_ENOMEM, "hipMalloc of " + std::to_string(bytes) + " bytes failed");
    }
    ptrs.push_back(p);
    *out = p;
    return RKX_OK;
  }
};

struct graph_holder {
  hipGraph_t g = nullptr;
  hipGraphExec_t ge = nullptr;
  ~graph_holder()
  {
    if (ge)
      (void) hipGraphExecDestroy(ge);
    if (g)
      (void) hipGraphDestroy(g);
  }
};

}  // namespace

// The step control on the host: ctrl_decide/ctrl_finish with the C library's
// pow, then the premultiplied coefficients of the next step.
void host_finish(state &st, const dev_method &dm, dev_coef &hc)
{
  const bool need = ctrl_decide(st);
  const double p = need ? std::pow(st.pow_x, st.pow_y) : 0.0;
  ctrl_finish(st, need, p);
  premult(dm, st.s, st.h, hc);
}

// DUMP output (include/tools.h:239-256): one line per attempted step, t and
// h then every component of the step's new y, "%.20e" separated by tabs.  The
// reference creates the file at the first attempt and appends afterwards.
struct dump_file {
  FILE *f = nullptr;
  ~dump_file()
  {
    if (f)
      std::fclose(f);
  }
  int append(const char *path, double t, double h, const std::vector<double> &y)
  {
    if (!f && !(f = std::fopen(path, "w")))
      return fail(RKX_EINVAL, std::string("rkx_solve: cannot open dump file ") + path);
    std::fprintf(f, "%.20e\t%.20e", t, h);
    for (double v : y)
      std::fprintf(f, "\t%.20e", v);
    std::fputc('\n', f);
    if (std::ferror(f))
      return fail(RKX_EINVAL, std::string("rkx_solve: write to dump file failed: ") + path);
    return RKX_OK;
  }
};

}  // namespace rkx

using namespace rkx;

extern "C" {

const char *rkx_last_error(void) { return g_last_error.c_str(); }

int rkx_device_sync(void)
{
  RKX_HIP(hipDeviceSynchronize());
  return RKX_OK;
}

void rkx_opts_default(rkx_opts *o)
{
  if (!o)
    return;
  std::memset(o, 0, sizeof(*o));
  o->impl = RKX_IMPL_PIPEDLS;
  o->ctrl = RKX_CTRL_DEVICE;
}

int rkx_ode_bruss2d(rkx_ode *p, int N, double alpha)
{
  if (!p || N < 2 || N > 32767)
    return fail(RKX_EINVAL, "rkx_ode_bruss2d: need 2 <= N <= 32767 (n = 2N^2 fits in int)");
  std::memset(p, 0, sizeof(*p));
  p->kind = RKX_BRUSS2D;
  p->grid = N;
  p->n = 2 * N * N;
  p->alpha = alpha;
  return RKX_OK;
}

int rkx_ode_lv(rkx_ode *p, int pairs, double c0, double d0, double d1, double c1)
{
  if (!p || pairs < 1 || pairs > (1 << 29))
    return fail(RKX_EINVAL, "rkx_ode_lv: bad pair count");
  std::memset(p, 0, sizeof(*p));
  p->kind = RKX_LV;
  p->grid = pairs;
  p->n = 2 * pairs;
  p->c0 = c0;
  p->d0 = d0;
  p->d1 = d1;
  p->c1 = c1;
  return RKX_OK;
}

int rkx_ode_lv_chain(rkx_ode *p, int cells, double c0, double d0, double d1, double c1,
                     double diff)
{
  if (!p || cells < 3 || cells > (1 << 29))
    return fail(RKX_EINVAL, "rkx_ode_lv_chain: need 3 <= cells");
  std::memset(p, 0, sizeof(*p));
  p->kind = RKX_LV_CHAIN;
  p->grid = cells;
  p->n = 2 * cells;
  p->c0 = c0;
  p->d0 = d0;
  p->d1 = d1;
  p->c1 = c1;
  p->diff = diff;
  return RKX_OK;
}

int rkx_ode_acc_dist(const rkx_ode *p)
{
  if (check_ode(p))
    return RKX_EINVAL;
  switch (p->kind) {
  case RKX_BRUSS2D: return 2 * p->grid;
  case RKX_LV: return 1;
  default: return 2;
  }
}

int rkx_ode_start(const rkx_ode *p, double t, double *y0)
{
  (void) t;
  if (int rc = check_ode(p))
    return rc;
  if (!y0)
    return fail(RKX_EINVAL, "rkx_ode_start: null y0");
  if (p->kind == RKX_BRUSS2D) {
    const int N = p->grid;
    for (int i = 0; i < N; ++i)
      for (int j = 0; j < N; ++j) {
        y0[2 * N * i + 2 * j] = 0.5 + static_cast<double>(j) / static_cast<double>(N - 1);
        y0[2 * N * i + 2 * j + 1] =
            1.0 + (5.0 * static_cast<double>(i)) / static_cast<double>(N - 1);
      }
  } else {
    for (int i = 0; i < p->n; i += 2) {   // lv.c:58-62
      y0[i] = 0.2;
      y0[i + 1] = 1.1;
    }
  }
  return RKX_OK;
}

// splitmix64 finaliser over (seed, instance, index): a counter-based generator,
// so every value is reproducible independently of generation order.
static inline unsigned long long rkx_mix(unsigned long long z)
{
  z += 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

static inline double rkx_unit(unsigned long long seed, int inst, long long idx)
{
  const unsigned long long k =
      rkx_mix(seed ^ rkx_mix(static_cast<unsigned long long>(inst) * 0x100000001B3ull) ^
              rkx_mix(static_cast<unsigned long long>(idx) + 0x632BE59BD9B4E019ull));
  return static_cast<double>(k >> 11) * 0x1.0p-53;   // [0, 1)
}

int rkx_ode_start_seeded(const rkx_ode *p, unsigned long long seed, int instance, double amp,
                         double *y0)
{
  if (int rc = rkx_ode_start(p, 0.0, y0))
    return rc;
  if (p->kind == RKX_BRUSS2D) {
    const int N = p->grid;
    for (long long c = 0; c < static_cast<long long>(N) * N; ++c)
      y0[2 * c] = y0[2 * c] + amp * (2.0 * rkx_unit(seed, instance, c) - 1.0);
  } else {
    for (int c = 0; c < p->n / 2; ++c) {
      y0[2 * c] = 0.2 + 0.1 * rkx_unit(seed, instance, 2 * c);
      y0[2 * c + 1] = 1.1 + 0.1 * rkx_unit(seed, instance, 2 * c + 1);
    }
  }
  return RKX_OK;
}

int rkx_eval_rng_device(const rkx_ode *p, int i, int j, double t, const double *y_dev,
                        double *f_dev, void *stream)
{
  (void) t;
  if (int rc = check_ode(p))
    return rc;
  if (i < 0 || j >= p->n || j < i || !y_dev || !f_dev)
    return fail(RKX_EINVAL, "rkx_eval_rng_device: bad range or pointer");
  hipStream_t s = static_cast<hipStream_t>(stream);
  if (!s)
    if (int rc = default_stream(&s))
      return rc;
  const dev_ode d = make_dev_ode(*p);
  RKX_HIP(launch_eval_rng(d, i, j, y_dev, f_dev, s));
  RKX_HIP(hipStreamSynchronize(s));
  return RKX_OK;
}

int rkx_eval_rng_host(const rkx_ode *p, int i, int j, double t, const double *y, double *f)
{
  (void) t;
  if (int rc = check_ode(p))
    return rc;
  if (i < 0 || j >= p->n || j < i || !y || !f)
    return fail(RKX_EINVAL, "rkx_eval_rng_host: bad range or pointer");
  // Only the components the stencils of f[i..j] read travel to the device:
  // [i - d - 1, j + d + 1] clipped to the vector (the periodic chain reads
  // across its ends, so it sends everything).  The kernel indexes absolutely,
  // so it gets the window's base shifted back by lo, as the reference's tiled
  // stages pass v - first to ode_eval_rng (include/block-inline.h:338).
  const int d = rkx_ode_acc_dist(p);
  int lo = 0, hi = p->n - 1;
  if (p->kind != RKX_LV_CHAIN) {
    lo = std::max(0, (i & ~1) - d);
    hi = std::min(p->n - 1, (j | 1) + d);
  }
  const size_t win = static_cast<size_t>(hi - lo + 1);
  // per-thread device staging, grown on demand and kept for the thread's
  // later calls (a problem object's callbacks are called many times)
  thread_local double *yd = nullptr, *fd = nullptr;
  thread_local size_t cap = 0;
  if (win > cap) {
    if (yd)
      (void) hipFree(yd);
    if (fd)
      (void) hipFree(fd);
    yd = fd = nullptr;
    cap = 0;
    if (hipMalloc(reinterpret_cast<void **>(&yd), win * sizeof(double)) != hipSuccess ||
        hipMalloc(reinterpret_cast<void **>(&fd), win * sizeof(double)) != hipSuccess) {
      (void) hipGetLastError();
      return fail(RKX_ENOMEM, "rkx_eval_rng_host: device staging allocation failed");
    }
    cap = win;
  }
  hipStream_t s = nullptr;
  if (int rc = default_stream(&s))
    return rc;
  RKX_HIP(hipMemcpyAsync(yd, y + lo, win * sizeof(double), hipMemcpyHostToDevice, s));
  const dev_ode dp = make_dev_ode(*p);
  RKX_HIP(launch_eval_rng(dp, i, j, yd - lo, fd - lo, s));
  RKX_HIP(hipMemcpyAsync(f + i, fd + (i - lo), static_cast<size_t>(j - i + 1) * sizeof(double),
                         hipMemcpyDeviceToHost, s));
  RKX_HIP(hipStreamSynchronize(s));
  return RKX_OK;
}

int rkx_eval_comp_host(const rkx_ode *p, int i, double t, const double *y, double *out)
{
  if (!out)
    return fail(RKX_EINVAL, "rkx_eval_comp_host: null output");
  double v[2] = {0.0, 0.0};
  // a one-component range written at absolute index i into v - (i & ~1)
  if (int rc = rkx_eval_rng_host(p, i, i, t, y, v - (i & ~1)))
    return rc;
  *out = v[i & 1];
  return RKX_OK;
}

// The full-storage loop order that moves the fewest vectors of n per step
// for a table: D (block-inline.h:253-320) scatters into s+1 accumulators, A
// and E gather each stage from the stage derivatives (A skipping the zero
// entries of the table), F fuses each RHS with the next stage's gather.  All
// give the same results bitwise.  Counts: reads + writes of n doubles.
static int least_traffic_form(const rkx_method &m)
{
  const int s = m.s;
  int d = (s + 2) + 5;                                   // first and last stage
  for (int i = 1; i < s - 1; ++i)
    d += 5 + 2 * (s - 1 - i);
  int a = 2 + 1, e = 2 + 1;                              // stage 0; output reads y
  for (int l = 1; l < s; ++l) {
    int nz = 0;
    for (int i = 0; i < l; ++i)
      nz += (m.A[l][i] != 0.0);
    a += 4 + nz;
    e += 4 + l;
  }
  for (int i = 0; i < s; ++i) {
    a += (i == 0 || m.b[i] != 0.0 || m.b[i] - m.b_hat[i] != 0.0) ? 1 : 0;
    e += 1;
  }
  a += 1;                                                // y_new
  e += 1;
  int f = 3 + 2 + (s + 2);                               // stage 0, last stage, output
  for (int l = 1; l < s - 1; ++l)
    f += l + 4;
  int best = RKX_IMPL_D, cost = d;
  if (f < cost) { best = RKX_IMPL_F; cost = f; }
  if (a < cost) { best = RKX_IMPL_A; cost = a; }
  if (e < cost) { best = RKX_IMPL_E; cost = e; }
  return best;
}

size_t rkx_solve_footprint(const rkx_ode *p, const rkx_method *m, int impl)
{
  if (check_ode(p) || check_method(m))
    return 0;
  const size_t vec = static_cast<size_t>(p->n) * sizeof(double);
  const size_t fixed = sizeof(state) + sizeof(dev_method) + sizeof(dev_coef);
  const dev_ode dp = make_dev_ode(*p);
  const bool fused =
      sweep_supported(dp, *m) || chain_supported(dp, *m) || pairs_supported(dp, *m);
  const size_t s = static_cast<size_t>(m->s);
  if ((impl == RKX_IMPL_PIPED || impl == RKX_IMPL_PIPEDLS) && !fused)
    impl = least_traffic_form(*m);
  if (impl == RKX_IMPL_A || impl == RKX_IMPL_E)
    return (2 + s + 1) * vec + fixed;            // y x2, v_0..v_{s-1}, w
  if (impl == RKX_IMPL_F)
    return (2 + s + 2) * vec + fixed;            // y x2, v_0..v_{s-1}, w x2
  if (impl == RKX_IMPL_D)
    return (2 + (s - 1) + 2) * vec + fixed;      // y x2, w_1..w_{s-1}, dy, err
  return 2 * vec + fixed;                        // y x2
}

int rkx_solve(const rkx_ode *p, const rkx_method *m, double t0, double te,
              const double *y0_dev, double *y_dev, double tol, const rkx_opts *o_in,
              rkx_stats *st_out)
{
  if (int rc = check_ode(p))
    return rc;
  if (int rc = check_method(m))
    return rc;
  if (!y0_dev || !y_dev)
    return fail(RKX_EINVAL, "rkx_solve: null vector");
  rkx_opts o;
  if (o_in)
    o = *o_in;
  else
    rkx_opts_default(&o);
  if (o.impl < RKX_IMPL_D || o.impl > RKX_IMPL_F)
    return fail(RKX_EINVAL, "rkx_solve: unknown impl");
  if (o.ctrl != RKX_CTRL_DEVICE && o.ctrl != RKX_CTRL_HOST)
    return fail(RKX_EINVAL, "rkx_solve: unknown step-control mode");
  if (o.trace_cap < 0 || (o.trace_cap > 0 && !o.trace))
    return fail(RKX_EINVAL, "rkx_solve: trace_cap without a trace buffer");
  if (o.dump_path)
    o.ctrl = RKX_CTRL_HOST;          // the dump needs every attempted step on the host
  hipStream_t stream = static_cast<hipStream_t>(o.stream);
  if (!stream)
    if (int rc = default_stream(&stream))
      return rc;

  const dev_ode dp = make_dev_ode(*p);
  const int s = m->s;
  int impl = o.impl;
  // fused forms: the row sweep for BRUSS2D, the segment kernel for the LV
  // chain, the per-pair kernel for LV pairs; anything else (tables without a
  // compiled pattern or stage count) runs the full-storage loop order with
  // the least traffic.  D, A, E and F run as asked on every problem.
  bool fused = impl == RKX_IMPL_PIPED || impl == RKX_IMPL_PIPEDLS;
  const bool chain = fused && chain_supported(dp, *m);
  const bool pairs = fused && !chain && pairs_supported(dp, *m);
  if (fused && !chain && !pairs && !sweep_supported(dp, *m)) {
    impl = least_traffic_form(*m);
    fused = false;
  }
  const bool sweep = fused && !chain && !pairs;
  const bool gather = impl == RKX_IMPL_A || impl == RKX_IMPL_E || impl == RKX_IMPL_F;
  const size_t n = static_cast<size_t>(p->n);
  const size_t vec = n * sizeof(double);

  solve_mem mem;
  double *ybuf[2] = {nullptr, nullptr};
  double *w = nullptr, *dy = nullptr, *err = nullptr, *trace_dev = nullptr;
  double *v = nullptr, *w1 = nullptr;   // gather forms: v_0..v_{s-1}; F's second w
  state *st_dev = nullptr;
  dev_method *dm_dev = nullptr;
  dev_coef *cf_dev = nullptr;
  if (int rc = mem.alloc(reinterpret_cast<void **>(&st_dev), sizeof(state)))
    return rc;
  if (int rc = mem.alloc(reinterpret_cast<void **>(&dm_dev), sizeof(dev_method)))
    return rc;
  if (int rc = mem.alloc(reinterpret_cast<void **>(&cf_dev), sizeof(dev_coef)))
    return rc;
  for (int b = 0; b < 2; ++b)
    if (int rc = mem.alloc(reinterpret_cast<void **>(&ybuf[b]), vec))
      return rc;
  if (impl == RKX_IMPL_D) {
    if (int rc = mem.alloc(reinterpret_cast<void **>(&w), vec * (s - 1)))
      return rc;
    if (int rc = mem.alloc(reinterpret_cast<void **>(&dy), vec))
      return rc;
    if (int rc = mem.alloc(reinterpret_cast<void **>(&err), vec))
      return rc;
  } else if (gather) {
    if (int rc = mem.alloc(reinterpret_cast<void **>(&v), vec * s))
      return rc;
    if (int rc = mem.alloc(reinterpret_cast<void **>(&w), vec))
      return rc;
    if (impl == RKX_IMPL_F)
      if (int rc = mem.alloc(reinterpret_cast<void **>(&w1), vec))
        return rc;
  }
  const gather_masks nz = make_gather_masks(*m);
  if (o.trace_cap > 0)
    if (int rc = mem.alloc(reinterpret_cast<void **>(&trace_dev),
                           sizeof(double) * 3 * static_cast<size_t>(o.trace_cap)))
      return rc;

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
  hs.trace = trace_dev;
  hs.trace_cap = o.trace_cap;
  hs.s = s;
  hs.meth = dm_dev;
  hs.coef = cf_dev;
  dev_coef hc{};
  premult(dm, s, hs.h, hc);
  hs.cur = 0;
  hs.done = !(t0 < te);
  RKX_HIP(hipMemcpyAsync(dm_dev, &dm, sizeof(dm), hipMemcpyHostToDevice, stream));
  RKX_HIP(hipMemcpyAsync(cf_dev, &hc, sizeof(hc), hipMemcpyHostToDevice, stream));
  RKX_HIP(hipMemcpyAsync(st_dev, &hs, sizeof(hs), hipMemcpyHostToDevice, stream));
  RKX_HIP(hipMemcpyAsync(ybuf[0], y0_dev, vec, hipMemcpyDeviceToDevice, stream));

  sweep_cfg cfg{1, 1};
  if (sweep)
    cfg = sweep_default_cfg(dp, *m);
  if (o.tile_rows > 0)
    cfg.rows = o.tile_rows;
  if (o.tile_waves > 0)
    cfg.waves = o.tile_waves;
  if (cfg.waves < 1 || cfg.waves > 4 || cfg.rows < 1 || 64 * cfg.waves <= 2 * s)
    return fail(RKX_EINVAL, "rkx_solve: tile_waves must be 1..4 (and wider than 2s), "
                            "tile_rows >= 1");
  sweep_inst *inst_dev = nullptr;
  chain_inst *chain_dev = nullptr;
  int ntiles = 0;
  bool gen_rows = false;   // the general ghost-row form (sweep_inst_single_reflect)
  const int pattern = sweep_pattern(*m);
  if (sweep && sweep_resident_blocks(pattern, cfg.waves) < 1)
    return fail(RKX_EINVAL, "rkx_solve: tile_waves wider than this table's fused kernel "
                            "(DOPRI87: at most 2)");
  if (chain) {
    if (int rc = mem.alloc(reinterpret_cast<void **>(&chain_dev), sizeof(chain_inst)))
      return rc;
    const chain_inst hi = make_chain_inst(dp, *m, st_dev, ybuf[0], ybuf[1]);
    ntiles = hi.nwg;
    RKX_HIP(hipMemcpyAsync(chain_dev, &hi, sizeof(hi), hipMemcpyHostToDevice, stream));
    RKX_HIP(hipStreamSynchronize(stream));
  } else if (sweep) {
    if (int rc = mem.alloc(reinterpret_cast<void **>(&inst_dev), sizeof(sweep_inst)))
      return rc;
    const sweep_inst hi = make_sweep_inst(dp, *m, cfg, st_dev, ybuf[0], ybuf[1]);
    ntiles = hi.ntiles;
    gen_rows = !sweep_inst_single_reflect(pattern, s, hi);
    RKX_HIP(hipMemcpyAsync(inst_dev, &hi, sizeof(hi), hipMemcpyHostToDevice, stream));
    RKX_HIP(hipStreamSynchronize(stream));
  }

  auto enqueue_step = [&]() -> hipError_t {
    if (impl == RKX_IMPL_D)
      return launch_d_step(dp, cf_dev, s, st_dev, ybuf[0], ybuf[1], w, dy, err, stream);
    if (gather)
      return launch_gather_step(impl, dp, cf_dev, s, nz, st_dev, ybuf[0], ybuf[1], v, w, w1,
                                stream);
    if (pairs)
      return launch_lv_step(dp, cf_dev, s, st_dev, ybuf[0], ybuf[1], stream);
    if (chain)
      return launch_chain(pattern, chain_dev, nullptr, ntiles, stream);
    return launch_sweep(pattern, cfg.waves, gen_rows, inst_dev, nullptr, ntiles, stream);
  };

  state *hs_pinned = nullptr;   // [0], [1]: snapshots of two chunks in flight
  RKX_HIP(hipHostMalloc(reinterpret_cast<void **>(&hs_pinned), 2 * sizeof(state),
                        hipHostMallocDefault));
  struct pinned_free {
    state *p;
    ~pinned_free() { (void) hipHostFree(p); }
  } pf{hs_pinned};

  RKX_HIP(hipStreamSynchronize(stream));
  const auto t_start = std::chrono::steady_clock::now();

  if (hs.done) {
    hs_pinned[0] = hs;
  } else if (o.ctrl == RKX_CTRL_HOST) {
    // One host round trip per step: the device computes the error norm, the
    // host runs the reference's step control with the C library's pow.
    state cur = hs;
    dump_file dump;
    std::vector<double> ydump;
    while (!cur.done) {
      RKX_HIP(enqueue_step());
      RKX_HIP(hipMemcpyAsync(hs_pinned, st_dev, sizeof(state), hipMemcpyDeviceToHost, stream));
      RKX_HIP(hipStreamSynchronize(stream));
      cur = hs_pinned[0];
      const int attempt = cur.acc + cur.rej;
      if (o.dump_path) {
        // the step's new y is in the other buffer until the decision flips cur
        ydump.resize(n);
        RKX_HIP(hipMemcpy(ydump.data(), ybuf[cur.cur ^ 1], vec, hipMemcpyDeviceToHost));
        if (int rc = dump.append(o.dump_path, cur.t, cur.h, ydump))
          return rc;
      }
      if (o.trace && attempt < o.trace_cap) {
        o.trace[3 * attempt + 0] = cur.t;
        o.trace[3 * attempt + 1] = cur.h;
        o.trace[3 * attempt + 2] = cur.err;
      }
      host_finish(cur, dm, hc);
      RKX_HIP(hipMemcpyAsync(cf_dev, &hc, sizeof(hc), hipMemcpyHostToDevice, stream));
      hs_pinned[0] = cur;
      RKX_HIP(hipMemcpyAsync(st_dev, hs_pinned, sizeof(state), hipMemcpyHostToDevice, stream));
    }
    RKX_HIP(hipStreamSynchronize(stream));
    hs_pinned[0] = cur;
  } else {
    // Chunks of steps captured once as a graph.  The host keeps two chunks in
    // flight: it enqueues chunk k+1 (and a snapshot copy of the state) before
    // waiting on the snapshot taken after chunk k.  The snapshot copy leaves
    // the GPU idle for about 18 us, so a solve that is still running after
    // its first kChunkGrowAfter chunks launches the graph kChunkReps times
    // per snapshot (a short solve keeps short chunks: after the last step
    // every launch in flight returns at once, but still costs its launch).
    constexpr int kChunkGrowAfter = 8, kChunkReps = 4;
    const int chunk = fused ? 32 : 16;
    graph_holder gh;
    RKX_HIP(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal));
    hipError_t cap_err = hipSuccess;
    for (int k = 0; k < chunk && cap_err == hipSuccess; ++k)
      cap_err = enqueue_step();
    hipError_t end_err = hipStreamEndCapture(stream, &gh.g);
    if (cap_err != hipSuccess)
      return fail(RKX_EDEVICE, std::string("kernel launch during capture failed: ") +
                                   hipGetErrorString(cap_err));
    RKX_HIP(end_err);
    RKX_HIP(hipGraphInstantiate(&gh.ge, gh.g, nullptr, nullptr, 0));
    hipEvent_t ev[2];
    RKX_HIP(hipEventCreateWithFlags(&ev[0], hipEventDisableTiming));
    RKX_HIP(hipEventCreateWithFlags(&ev[1], hipEventDisableTiming));
    struct ev_free {
      hipEvent_t *e;
      ~ev_free()
      {
        (void) hipEventDestroy(e[0]);
        (void) hipEventDestroy(e[1]);
      }
    } evf{ev};
    int launched = 0;
    auto launch_chunk = [&](int slot) -> int {
      const int reps = (launched++ < kChunkGrowAfter) ? 1 : kChunkReps;
      for (int r = 0; r < reps; ++r)
        RKX_HIP(hipGraphLaunch(gh.ge, stream));
      RKX_HIP(hipMemcpyAsync(&hs_pinned[slot], st_dev, sizeof(state), hipMemcpyDeviceToHost,
                             stream));
      RKX_HIP(hipEventRecord(ev[slot], stream));
      return RKX_OK;
    };
    if (int rc = launch_chunk(0))
      return rc;
    int last_attempts = -1, stalled = 0;
    for (int k = 0;; ++k) {
      const int a = k & 1, b = a ^ 1;
      if (int rc = launch_chunk(b))
        return rc;
      RKX_HIP(hipEventSynchronize(ev[a]));
      if (hs_pinned[a].done)
        break;
      // every chunk must advance the attempt counter; a chunk that does not
      // means the device-side step control is not running -- stop, do not spin
      const int att = hs_pinned[a].acc + hs_pinned[a].rej;
      stalled = (att == last_attempts) ? stalled + 1 : 0;
      last_attempts = att;
      if (stalled >= 3) {
        (void) hipStreamSynchronize(stream);
        return fail(RKX_EDEVICE, "device step control made no progress");
      }
    }
    RKX_HIP(hipStreamSynchronize(stream));
    RKX_HIP(hipMemcpy(&hs_pinned[0], st_dev, sizeof(state), hipMemcpyDeviceToHost));
  }
  const auto t_end = std::chrono::steady_clock::now();
  const state fin = hs_pinned[0];
  RKX_HIP(hipMemcpyAsync(y_dev, ybuf[fin.cur], vec, hipMemcpyDeviceToDevice, stream));
  if (o.trace_cap > 0 && o.ctrl != RKX_CTRL_HOST) {
    const int ntr = std::min(o.trace_cap, fin.acc + fin.rej);
    if (ntr > 0)
      RKX_HIP(hipMemcpyAsync(o.trace, trace_dev, sizeof(double) * 3 * ntr, hipMemcpyDeviceToHost,
                             stream));
  }
  RKX_HIP(hipStreamSynchronize(stream));
  if (st_out) {
    st_out->steps_acc = fin.acc;
    st_out->steps_rej = fin.rej;
    st_out->t = fin.t;
    st_out->h = fin.h;
    st_out->kernel_time = std::chrono::duration<double>(t_end - t_start).count();
    st_out->resident_bytes = rkx_solve_footprint(p, m, impl);
  }
  return RKX_OK;
}

int rkx_solve_host(const rkx_ode *p, const rkx_method *m, double t0, double te, const double *y0,
                   double *y, double tol, const rkx_opts *o, rkx_stats *st)
{
  if (int rc = check_ode(p))
    return rc;
  if (!y0 || !y)
    return fail(RKX_EINVAL, "rkx_solve_host: null vector");
  const size_t vec = static_cast<size_t>(p->n) * sizeof(double);
  solve_mem mem;
  double *d0 = nullptr, *d1 = nullptr;
  if (int rc = mem.alloc(reinterpret_cast<void **>(&d0), vec))
    return rc;
  if (int rc = mem.alloc(reinterpret_cast<void **>(&d1), vec))
    return rc;
  RKX_HIP(hipMemcpy(d0, y0, vec, hipMemcpyHostToDevice));
  if (int rc = rkx_solve(p, m, t0, te, d0, d1, tol, o, st))
    return rc;
  RKX_HIP(hipMemcpy(y, d1, vec, hipMemcpyDeviceToHost));
  return RKX_OK;
}

}  // extern "C"
