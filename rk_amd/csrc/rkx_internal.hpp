// Internal declarations shared by the rkx host runtime and the HIP kernels.
#pragma once

#include <hip/hip_runtime.h>

#include <atomic>
#include <cstddef>
#include <cstdint>
#include <string>
#include <thread>
#include <vector>

#include "../../include/rkx.h"

namespace rkx {

int fail(int code, const std::string &msg);

#define RKX_HIP(call)                                                              \
  do {                                                                             \
    hipError_t e_ = (call);                                                        \
    if (e_ != hipSuccess)                                                          \
      return ::rkx::fail(RKX_EDEVICE, std::string(#call " failed: ") +             \
                                          hipGetErrorString(e_));                 \
  } while (0)

// Problem descriptor as the kernels see it (by value in kernel arguments).
struct dev_ode {
  int kind;
  int n;        // components
  int cells;    // n / 2: grid points (BRUSS2D), pairs (LV), cells (chain)
  int N;        // BRUSS2D grid size
  double coef;  // BRUSS2D (alpha*(N-1))*(N-1), the reference's alpha*N1*N1
  double c0, d0, d1, c1;
  double diff;
};

dev_ode make_dev_ode(const rkx_ode &p);

// Method coefficients in device memory: A row-major [RKX_MAX_STAGES^2], then
// b, e = b - b_hat (src/impl/seq/D.c:40-41), c.
struct dev_method {
  double A[RKX_MAX_STAGES * RKX_MAX_STAGES];
  double b[RKX_MAX_STAGES];
  double e[RKX_MAX_STAGES];
  double c[RKX_MAX_STAGES];
  // grouped by the stage that scatters with them, as dev_coef::col but not
  // multiplied by h: col[i] = {A[i+1][i], ..., A[s-1][i], b[i], e[i]}.  The
  // fused kernels form h * col themselves at their start (the same products
  // as premult), so their steps need no premultiplication in the control.
  double col[RKX_MAX_STAGES][RKX_MAX_STAGES + 2];
};

// Premultiplied coefficients of the current step (premult, src/methods.c:43-58):
// hA = h*A (j < i), hb = h*b, he = h*e.  Written by the step control together
// with the new h; read as uniform values by every kernel of the next step.
struct dev_coef {
  double hA[RKX_MAX_STAGES * RKX_MAX_STAGES];
  double hb[RKX_MAX_STAGES];
  double he[RKX_MAX_STAGES];
  // the same values grouped by the stage that consumes them: stage i scatters
  // with col[i] = {hA[i+1][i], ..., hA[s-1][i], hb[i], he[i]} (contiguous, so
  // a fused kernel loads one stage's coefficients with one scalar load)
  double col[RKX_MAX_STAGES][RKX_MAX_STAGES + 2];
};

// Integration state of one solve, resident in device memory.  Written only by
// the step control (the last workgroup of a step, or the host).
struct alignas(64) state {
  double t, h, te, tol;
  double err;                        // error norm of the last attempt
  unsigned long long err_bits;       // atomicMax accumulator (bits of a double >= 0)
  unsigned int zero_div;             // some divisor max(|y|,|y_old|) was 0
  unsigned int blocks_done;          // last-workgroup detection
  int acc, rej;
  int cur;                           // which of the two y buffers holds y
  int done;                          // the step loop has ended
  int step_limit;
  int ord;
  int ctrl;                          // RKX_CTRL_*
  int trace_cap;
  int s;                             // stages
  double pow_x, pow_y;               // arguments of the step's pow (ctrl_decide)
  double *trace;                     // device buffer, 3 doubles per attempt
  const dev_method *meth;            // raw coefficients for the premultiplication
  dev_coef *coef;                    // premultiplied coefficients of the next step
};

// premult of the raw table by h (identical on host and device)
__host__ __device__ inline void premult(const dev_method &m, int s, double h, dev_coef &c)
{
  for (int i = 0; i < s; ++i) {
    for (int j = 0; j < i; ++j)
      c.hA[i * RKX_MAX_STAGES + j] = h * m.A[i * RKX_MAX_STAGES + j];
    c.hb[i] = h * m.b[i];
    c.he[i] = h * m.e[i];
  }
  for (int i = 0; i < s; ++i) {
    for (int l = i + 1; l < s; ++l)
      c.col[i][l - i - 1] = c.hA[l * RKX_MAX_STAGES + i];
    c.col[i][s - i - 1] = c.hb[i];
    c.col[i][s - i] = c.he[i];
  }
}

// Step control of include/tools.h:235-278, split at the one libm call.
// ctrl_decide: accept/reject with st.err, advance t and the counters, and set
// pow_x = tol/err, pow_y = 1/(ord+1).  Returns true when the new h needs
// p = pow(pow_x, pow_y) (variable step size).
__host__ __device__ inline bool ctrl_decide(state &st)
{
  if (st.tol <= 0.0) {                 // fixed step size (tools.h:258-263)
    st.t += st.h;
    st.acc += 1;
    st.cur ^= 1;
    return false;
  }
  if (st.err <= st.tol) {              // accept (tools.h:264-270)
    st.t += st.h;
    st.acc += 1;
    st.cur ^= 1;
  } else {                             // reject: y stays in the other buffer
    st.rej += 1;
  }
  st.pow_x = st.tol / st.err;
  st.pow_y = 1.0 / (st.ord + 1.0);
  return true;
}

// ctrl_finish: h_new_acc/h_new_rej with p (tools.h:221-233), then the
// for-increment h = fmin(h, te - t) and the loop test of FOR_ALL_GRIDPOINTS
// (tools.h:191-198).  The caller premultiplies the coefficients with the new h.
__host__ __device__ inline void ctrl_finish(state &st, bool used_pow, double p)
{
  if (used_pow)
    st.h = st.h * fmin(2.0, fmax(0.5, 0.9 * p));
  st.h = fmin(st.h, st.te - st.t);
  const int attempts = st.acc + st.rej;
  st.done = !((st.t < st.te) && (st.step_limit <= 0 || attempts < st.step_limit));
}

// ---- host runtime helpers (rkx_host.cpp) ------------------------------------
int check_ode(const rkx_ode *p);
int check_method(const rkx_method *m);
void fill_dev_method(const rkx_method &m, dev_method &d);
double initial_h(double H, double tol, int limit, double h0);   // src/tools.c:25-36
int default_stream(hipStream_t *out);
// the step control on the host (ctrl_decide/ctrl_finish with the C library's
// pow) and the next step's premultiplied coefficients
void host_finish(state &st, const dev_method &dm, dev_coef &hc);

// Kernel launchers (rkx_kernels_*.hip).  All return hipError_t of the launch.
hipError_t launch_eval_rng(const dev_ode &p, int i, int j, const double *y, double *f,
                           hipStream_t s);
hipError_t launch_d_step(const dev_ode &p, const dev_coef *cf, int s, state *st,
                         double *y0, double *y1, double *w, double *dy, double *err,
                         hipStream_t stream);
// Stage-fused step of independent LV pairs (rkx_kernels_lvpairs.hip): every
// table with 2, 3, 4, 7, 8 or 13 stages
bool pairs_supported(const dev_ode &p, const rkx_method &m);
hipError_t launch_lv_step(const dev_ode &p, const dev_coef *cf, int s, state *st, double *y0,
                          double *y1, hipStream_t stream);

// Gather forms A, E, F (src/impl/seq/A.c, E.c, F.c): non-zero patterns of the
// raw table (zero_pattern, src/methods.c:24-39) as bit masks, bit i of rowA[l]
// for A[l][i], of b for b[i], of e for b[i] - b_hat[i]
struct gather_masks {
  unsigned rowA[RKX_MAX_STAGES];
  unsigned b, e;
};
gather_masks make_gather_masks(const rkx_method &m);
// one attempted step of form RKX_IMPL_A/E/F: v holds s vectors of n, w0 (and
// w1 for F) one each
hipError_t launch_gather_step(int form, const dev_ode &p, const dev_coef *cf, int s,
                              const gather_masks &nz, state *st, double *y0, double *y1,
                              double *v, double *w0, double *w1, hipStream_t stream);

struct sweep_cfg {
  int waves;     // wavefronts per workgroup (strip width 64*waves lanes)
  int rows;      // band height in grid rows
};
// One BRUSS2D instance as the fused kernel sees it (device memory).  A launch
// covers a table of (instance, tile) pairs, so one kernel steps many
// instances, each with its own state, coefficients and step control.
//
// Columns: nwide strips of OC output columns, each band of a strip one
// workgroup.  When the columns left over after the widest strips fit a strip
// of half the lanes (split), they form a narrow last strip whose bands are
// paired: one workgroup carries two bands of it, each on half of its waves,
// so the narrow strip costs half a strip of workgroups.
struct sweep_inst {
  int N;                 // grid size
  int R;                 // band height (rows)
  int OC;                // output columns per wide strip
  int nstrips, nbands;   // strips (wide + narrow), bands per strip
  int nwide;             // wide strips: tiles 0 .. nwide*nbands-1
  int split;             // 1: a narrow last strip, (nbands+1)/2 paired tiles
  int ntiles;            // workgroups that share this state per step
  double coef;           // (alpha*(N-1))*(N-1)
  double *y0, *y1;       // the two y buffers
  state *st;             // step state; the kernel forms h*A etc. from st->meth
};

bool sweep_supported(const dev_ode &p, const rkx_method &m);
int sweep_pattern(const rkx_method &m);   // compiled zero pattern (0: none)
sweep_cfg sweep_default_cfg(const dev_ode &p, const rkx_method &m);
// Band height a tile gets for `rows` requested rows: unchanged, or for tables
// whose fused kernel runs whole trips of its s+1-row ring, rounded up so that
// the band's rows + 2(s-1) row iterations, less the peeled ones
// (peeled_iters), fill whole trips.
int sweep_band_rows(int pattern, int s, int rows);
// Whether every ghost row a band of an N-row grid loads lies within a single
// reflection about row 0 / N-1 (the kernel's fast form); false for grids
// shorter than the halo, prefetch and ring: those launch the general form.
bool sweep_single_reflect(int pattern, int s, int N);
// The same for one instance's tiles: the paired bands of a narrow strip both
// run the full band height, so the last of them may load rows further past
// N-1 than a band of its own height.
bool sweep_inst_single_reflect(int pattern, int s, const sweep_inst &I);
// Row iterations of tile t of an instance (its band, or the shorter last
// band, plus the 2(s-1) halo rows; paired bands run the full band height).
int sweep_tile_iters(int pattern, int s, const sweep_inst &I, int t);
sweep_inst make_sweep_inst(const dev_ode &p, const rkx_method &m, const sweep_cfg &c, state *st,
                           double *y0, double *y1);
// tiles == nullptr: a single instance (insts[0]) with ntiles tiles; gen: the
// general reflection form (some instance fails sweep_single_reflect)
hipError_t launch_sweep(int pattern, int K, bool gen, const sweep_inst *insts, const int2 *tiles,
                        int ntiles, hipStream_t stream);
// the DOPRI87 kernels (their own translation unit: 13 stages take long to build)
hipError_t launch_sweep_dopri87(int K, bool gen, const sweep_inst *insts, const int2 *tiles,
                                int ntiles, hipStream_t stream);
int resident_blocks_dopri87(int K);
// workgroups of the fused kernel resident per CU for a pattern and strip
// width; 0: the strip is too wide for that table's kernel
int sweep_resident_blocks(int pattern, int K);

// One LV chain instance as the fused chain kernel sees it (device memory).
struct chain_inst {
  int cells;             // M
  int tile_out;          // output cells per segment
  int ntiles;            // segments of 1024 cells per step
  int nwg;               // workgroups per step (each loops over segments)
  double c0, d0, d1, c1, diff;
  double *y0, *y1;
  state *st;             // step state; the kernel forms h*A etc. from st->meth
};

bool chain_supported(const dev_ode &p, const rkx_method &m);
chain_inst make_chain_inst(const dev_ode &p, const rkx_method &m, state *st, double *y0,
                           double *y1);
hipError_t launch_chain(int pattern, const chain_inst *insts, const int2 *tiles, int ntiles,
                        hipStream_t stream);

}  // namespace rkx
