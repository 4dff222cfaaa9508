/*
 * rkx -- MI355X-native embedded Runge-Kutta solver: the C ABI.
 *
 * Plain C types only (pointers, sizes, doubles); no HIP or torch types cross
 * this boundary (a stream is passed as void*).  Device pointers are marked
 * "_dev"; everything else lives in host memory.  Every entry point returns 0
 * on success or a negative RKX_E* code; rkx_last_error() gives the message
 * of the calling thread's last failure.
 *
 * The reference (UBT-AI2/rk) binds its three plugin interfaces at compile and
 * link time: the problem (include/ode.h:29-36), the method table
 * (include/methods.h:92-93) and the solver (include/solver.h:47).  A GPU
 * kernel cannot call host ode_eval_* callbacks, so the problem crosses this
 * ABI as a descriptor (rkx_ode) naming one of the device right-hand sides,
 * and the method crosses it as data with the layout of alloc_emb_rk_method
 * (include/methods.h:31-42).  The link-compatible solver() of solver.h:47 is
 * provided on top of this ABI by rk_amd/csrc/dropin/solver_hip.c.
 */
#ifndef RKX_H_
#define RKX_H_

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define RKX_MAX_STAGES 16

#define RKX_OK 0
#define RKX_EINVAL (-1)        /* bad argument / unsupported combination */
#define RKX_ENOMEM (-2)        /* device or host allocation failed */
#define RKX_EDEVICE (-3)       /* HIP runtime error */
#define RKX_ELIMIT (-4)        /* request exceeds a documented limit */

/* ---- problems (replaces include/ode.h:29-36) ----------------------------- */
enum {
  RKX_BRUSS2D = 1,   /* src/problems/bruss2d-mix.c: 2N*N comps, (u,v) interleaved, d = 2N */
  RKX_LV = 2,        /* src/problems/lv.c: pairs (x,y), d = 1 */
  RKX_LV_CHAIN = 3   /* diffusive periodic LV chain (no reference counterpart), d = 2 */
};

typedef struct rkx_ode {
  int kind;
  int n;                 /* ode_size (ode.h:29) */
  int grid;              /* BRUSS2D: N; LV: pairs; LV_CHAIN: cells */
  double alpha;          /* BRUSS2D diffusion (config.h:34-36 BRUSS_ALPHA) */
  double c0, d0, d1, c1; /* LV: x' = c0 x - d0 x y, y' = d1 x y - c1 y (lv.c:31-34) */
  double diff;           /* LV_CHAIN diffusion coefficient */
} rkx_ode;

/* Fill a descriptor; replaces the compile-time BRUSS_GRID_SIZE/BRUSS_ALPHA
 * (include/config.h:30-36) and LV_COUNT (config.h:42-44). */
int rkx_ode_bruss2d(rkx_ode *p, int N, double alpha);
int rkx_ode_lv(rkx_ode *p, int pairs, double c0, double d0, double d1, double c1);
int rkx_ode_lv_chain(rkx_ode *p, int cells, double c0, double d0, double d1, double c1,
                     double diff);

/* ode_acc_dist() (ode.h:36; bruss2d-mix.c:718-721, lv.c:127-130) */
int rkx_ode_acc_dist(const rkx_ode *p);

/* ode_start(t, y0) (ode.h:33; bruss2d-mix.c:35-60, lv.c:43-63), host memory. */
int rkx_ode_start(const rkx_ode *p, double t, double *y0);

/* Start vector of a seeded batch instance: the BRUSS2D start vector plus
 * amp*U[-1,1) noise on u from a counter-based hash of (seed, instance, index);
 * LV chain: x = 0.2 + 0.1*U[0,1), y = 1.1 + 0.1*U[0,1).  No reference
 * counterpart (the reference has no noise, bruss2d-mix.c:51-59). */
int rkx_ode_start_seeded(const rkx_ode *p, unsigned long long seed, int instance,
                         double amp, double *y0);

/* ode_eval_rng(i, j, t, y, f) (ode.h:35) on the device: writes f_dev[i..j]
 * inclusive with absolute indexing, reading y_dev.  The values equal the
 * reference's ode_eval_comp bitwise. */
int rkx_eval_rng_device(const rkx_ode *p, int i, int j, double t, const double *y_dev,
                        double *f_dev, void *stream);

/* The same on host buffers (y: all n components; writes f[i..j] only): the
 * host-side ode_eval_rng callback (ode.h:35) of the drop-in problem plugins.
 * Only the components the stencils read (i - d - 1 .. j + d + 1, d the
 * access distance) are staged through device memory.  Synchronous. */
int rkx_eval_rng_host(const rkx_ode *p, int i, int j, double t, const double *y, double *f);

/* ode_eval_comp(i, t, y) (ode.h:34) on host buffers: f_i into *out. */
int rkx_eval_comp_host(const rkx_ode *p, int i, double t, const double *y, double *out);

/* ---- methods (replaces rk_method_t, include/methods.h:92-118) ------------ */
typedef struct rkx_method {
  char name[32];
  int s, ord, has_emb;   /* has_emb = HAS_EMB_SOL(M) (methods.h:95-118) */
  double A[RKX_MAX_STAGES][RKX_MAX_STAGES];
  double b[RKX_MAX_STAGES], b_hat[RKX_MAX_STAGES], c[RKX_MAX_STAGES];
} rkx_method;

/* HEUN2, RK4, SSPRK3, RKF23, DOPRI54, VERNER65, DOPRI87 (src/methods.c:62-724) */
int rkx_method_get(const char *name, rkx_method *m);

/* ---- single solve (replaces solver(), include/solver.h:47) --------------- */
enum {
  RKX_IMPL_D = 0,        /* unfused stage kernels, full storage (src/impl/seq/D.c) */
  RKX_IMPL_PIPED = 1,    /* the same stage-fused sweep as RKX_IMPL_PIPEDLS: the fused form
                            keeps every stage vector of src/impl/seq/PipeD.c in registers,
                            so it has no full-storage variant (two vectors of n) */
  RKX_IMPL_PIPEDLS = 2,  /* stage-fused sweep, low storage: two vectors of n
                            (src/impl/seq/PipeDls.c); without a fused kernel for the
                            problem and table, the full-storage loop order below with
                            the fewest vector passes per step */
  RKX_IMPL_A = 3,        /* gather form skipping zero coefficients: w_l = y + sum hA v_i,
                            then v_l = f(w_l) (src/impl/seq/A.c); s + 3 vectors of n */
  RKX_IMPL_E = 4,        /* the gather form without the zero skip (src/impl/seq/E.c) */
  RKX_IMPL_F = 5         /* RHS and the next stage's gather in one pass, two w vectors
                            (src/impl/seq/F.c); s + 4 vectors of n.  D, A, E and F run on
                            every problem and give the same results bitwise */
};

enum {
  RKX_CTRL_DEVICE = 0,     /* step control on the device, in the last workgroup of each
                              step; the one pow() of the new h (include/tools.h:226) is
                              the C library's pow restated bitwise (rkx_pow_device), so
                              every h equals the reference's (default) */
  RKX_CTRL_HOST = 1        /* step control on the host after each step with the C
                              library's pow (one host round trip per step; bitwise the
                              same steps) */
};

typedef struct rkx_opts {
  double h0;             /* > 0: H_INIT; else initial_stepsize() (src/tools.c:25-36) */
  int step_limit;        /* STEP_LIMIT (config.h:70-72); <= 0: unlimited */
  int impl;              /* RKX_IMPL_* */
  int ctrl;              /* RKX_CTRL_* */
  int tile_rows;         /* sweep band height in grid rows (0: tuned default; VERNER65
                            rounds it up so that a band's row iterations after its
                            first 2(s-1) fill whole trips of its ring) */
  int tile_waves;        /* sweep strip width in wavefronts per component, 1..4 (DOPRI87:
                            1..2, its kernel runs one wave per SIMD); 0: tuned default.
                            With an even width, columns left over after the full
                            strips that fit half a strip run as a narrow strip whose
                            bands share workgroups two at a time */
  double *trace;         /* optional host buffer, 3 doubles (t, h, err) per attempt */
  int trace_cap;         /* attempts the trace buffer holds */
  void *stream;          /* hipStream_t to run on; NULL: the library's own stream */
  const char *dump_path; /* optional: every attempted step appended to this file in the
                            reference's DUMP format (include/tools.h:239-256): t, h and
                            the step's new y before the accept/reject decision, "%.20e"
                            tab-separated, one line per attempt; the file is created at
                            the first attempt.  Implies RKX_CTRL_HOST (bitwise the same
                            steps) and one device-to-host copy of y per attempt. */
} rkx_opts;

typedef struct rkx_stats {
  int steps_acc, steps_rej;   /* print_statistics (src/tools.c:62-79) */
  double t, h;                /* final t and h */
  double kernel_time;         /* seconds over the step loop (tools.c:77-78) */
  size_t resident_bytes;      /* device bytes held by the solve */
} rkx_stats;

void rkx_opts_default(rkx_opts *o);

/* Integrate y' = f(t, y) from t0 to te starting at y0_dev; the result goes to
 * y_dev (may equal y0_dev).  Both are device pointers of p->n doubles.  tol <= 0
 * means fixed step size (tools.h:223-224, 258-263). */
int rkx_solve(const rkx_ode *p, const rkx_method *m, double t0, double te,
              const double *y0_dev, double *y_dev, double tol, const rkx_opts *o,
              rkx_stats *st);

/* The same with host buffers (copies in and out; the PCIe copies are not
 * inside kernel_time). */
int rkx_solve_host(const rkx_ode *p, const rkx_method *m, double t0, double te,
                   const double *y0, double *y, double tol, const rkx_opts *o,
                   rkx_stats *st);

/* ---- batched solves (no reference counterpart: the reference integrates one
 * system per process).  Many instances of uneven size share each launch: one
 * batched step advances every active instance by one attempted step with its
 * own t, h, accept/reject and step limit, so each instance's result is the
 * same bitwise whether it runs alone or in a mix.  Instances can be added and
 * removed between steps (continuous batching). ------------------------------ */
typedef struct rkx_batch rkx_batch;

enum {
  RKX_STORAGE_AUTO = 0,   /* low storage when a fused kernel exists, else full */
  RKX_STORAGE_LOW = 1,    /* fused kernel, two vectors of n (PipeDls form) */
  RKX_STORAGE_FULL = 2    /* stage kernels, (s+3) vectors of n (D form) */
};

typedef struct rkx_batch_opts {
  size_t hbm_budget;      /* device bytes the batch may hold for instance data (0: no limit) */
  int ctrl;               /* RKX_CTRL_DEVICE */
  int tile_waves;         /* fused-kernel strip width for all instances (0: chosen per size;
                             rk_amd/autotune.py picks both by timing the batch's own work) */
  int tile_rows;          /* fused-kernel band height (0: chosen per size; rounded up as
                             rkx_opts.tile_rows) */
  void *stream;           /* hipStream_t, NULL: the library's own stream */
} rkx_batch_opts;

void rkx_batch_opts_default(rkx_batch_opts *o);
int rkx_batch_create(rkx_batch **b, const rkx_batch_opts *o);
int rkx_batch_destroy(rkx_batch *b);

/* Device bytes an instance would hold in the given storage mode. */
size_t rkx_batch_footprint(const rkx_ode *p, const rkx_method *m, int storage);

/* Add an instance starting from host y0 at t0.  Returns its id (>= 0), or
 * RKX_ELIMIT when it does not fit the HBM budget now (retry after removing
 * finished instances), or another negative code. */
int rkx_batch_add(rkx_batch *b, const rkx_ode *p, const rkx_method *m, double t0, double te,
                  const double *y0, double tol, double h0, int step_limit, int storage);

/* Run up to max_steps batched steps (stops early when every instance is done).
 * *active receives the number of instances still integrating. */
int rkx_batch_run(rkx_batch *b, int max_steps, int *active);

/* Per-instance statistics; *done = 1 once the instance reached te or its
 * step limit. */
int rkx_batch_status(rkx_batch *b, int id, rkx_stats *st, int *done);

/* Copy an instance's current y (n doubles) to host memory. */
int rkx_batch_result(rkx_batch *b, int id, double *y);

/* Release an instance's device memory (its id may be reused). */
int rkx_batch_remove(rkx_batch *b, int id);

/* Device bytes held by the batch's instances. */
size_t rkx_batch_bytes(const rkx_batch *b);

/* The row-sweep tile an instance runs with now: strip width in wavefronts
 * per component and band height in rows (both 0 for instances without the
 * row sweep: LV chains, full storage).  Band heights shrink as the batch
 * drains, when fewer than two tiles per compute unit would remain. */
int rkx_batch_tile(rkx_batch *b, int id, int *waves, int *rows);

/* ---- ensemble of independent LV pairs (no reference counterpart; the
 * reference's LV has one shared h for identical pairs, lv.c:31-37) ---------
 * `count` systems x' = c0 x - d0 x y, y' = d1 x y - c1 y, constants
 * params_dev[4k..4k+3] = (c0, d0, d1, c1), start y0_dev[2k..2k+1], each
 * integrated from t0 to te with its own step size, accept/reject and the
 * stage arithmetic of seq D and the step control of RKX_CTRL_DEVICE, so each
 * system's accepted/rejected sequence and result equal a plain single solve
 * of that system bitwise.  h0 > 0 is the initial step of every system;
 * 1 <= count <= 2^30.  Writes
 * y_dev[2k..2k+1] and the per-system accepted/rejected counts (device). */
int rkx_lv_ensemble(int count, const double *params_dev, const double *y0_dev, double *y_dev,
                    double t0, double te, const rkx_method *m, double tol, double h0,
                    int step_limit, int *acc_dev, int *rej_dev, void *stream);

/* ---- the step control's pow ------------------------------------------------
 * pow(x, y) of the C library (glibc 2.35: ARM optimized-routines' algorithm,
 * its x86-64 FMA build), restated so the device step control reproduces the
 * reference's h_new_acc/h_new_rej (include/tools.h:221-233) bitwise.  Domain:
 * x >= 0 (subnormals, +0, +inf included), 0 < y <= 1/2.  Elementwise over n
 * values: on the device (x_dev, y_dev, p_dev device pointers; synchronous)
 * and on the host CPU (the same code compiled for the host). */
int rkx_pow_device(const double *x_dev, const double *y_dev, double *p_dev, size_t n,
                   void *stream);
int rkx_pow_host(const double *x, const double *y, double *p, size_t n);

/* ---- solve without allocation ------------------------------------------------
 * No counterpart in the reference.  For callers that cannot allocate device
 * memory (tight serving loops, harnesses that forbid allocation): rkx_solve
 * for the fused low-storage forms (impl RKX_IMPL_PIPEDLS or RKX_IMPL_PIPED;
 * BRUSS2D, LV pairs and LV chains with a table that has a fused kernel), the same
 * arguments and bitwise the same results.  y_dev serves as one of the two y
 * buffers; the other and the step state live in a 32 MiB device array the
 * library holds on each device it runs on until the process ends.  Makes no
 * device or pinned allocation and creates no graph or event; with o->stream
 * set it also uses no stream of its own.  More than
 * RKX_NOALLOC_MAX_COMPONENTS components: RKX_ELIMIT; trace or dump output,
 * the D form, or a problem without a fused kernel: RKX_EINVAL.  Calls that
 * overlap on one device share that array and race. */
#define RKX_NOALLOC_MAX_COMPONENTS 4186112   /* (32 MiB - 64 KiB) / 8 */
int rkx_solve_noalloc(const rkx_ode *p, const rkx_method *m, double t0, double te,
                      const double *y0_dev, double *y_dev, double tol, const rkx_opts *o,
                      rkx_stats *st);

/* ---- runtime -------------------------------------------------------------- */
const char *rkx_last_error(void);
int rkx_device_sync(void);
/* Device bytes rkx_solve would hold for this problem/method/impl. */
size_t rkx_solve_footprint(const rkx_ode *p, const rkx_method *m, int impl);

#ifdef __cplusplus
}
#endif

#endif /* RKX_H_ */
