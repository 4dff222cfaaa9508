/*
 * CPU oracle for the embedded Runge-Kutta hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load this library, and only as the checker
 * (or the timed CPU baseline) -- never as a product code path.
 *
 * This is a from-scratch C restatement of the reference's sequential
 * implementation D (src/impl/seq/D.c:23-90) with its stage kernels
 * (include/block-inline.h:253-320), step control (include/tools.h:191-288),
 * premultiplication (src/methods.c:43-58), the method tables
 * (src/methods.c:62-724) and the two problems (src/problems/bruss2d-mix.c,
 * src/problems/lv.c).  The pipelined schedules PipeD / PipeDls
 * (src/impl/seq/PipeD.c, src/impl/seq/PipeDls.c), the gather forms A, E, F
 * (src/impl/seq/A.c, E.c, F.c) and the pthreads D (src/impl/pthreads/D.c) are
 * restated as well; they must agree bitwise with D.
 *
 * Every floating-point expression keeps the reference's operand order; the
 * library is compiled with -O3 -ffp-contract=off and without -march, so no
 * multiply-add is ever fused (the reference relies on baseline x86-64).
 *
 * Pinning (how each part is checked, see tests/test_oracle_*.py):
 *   - orc_solve_D / PipeD / PipeDls / A / E / F: known-answer footprints and
 *     step counts produced by the compiled reference, and DUMP trajectories
 *     (every attempted step, full vectors) at tiny grids, bitwise; the
 *     reference's own A, E and F binaries reproduce D's trajectories.
 *   - BRUSS2D RHS: brute-force dense Neumann Laplacian on tiny grids; the
 *     homogeneous steady state u=1, v=3.4.
 *   - LV RHS: closed form; the conserved quantity of the pair.
 *   - method tables: order conditions with exact rationals.
 *   - step control: synthetic error values (0, inf, 0/0 divisor, NaN).
 *   - LV chain, generalized LV pairs and noisy BRUSS2D start values have no
 *     reference counterpart; they are pinned by the invariants above
 *     (chain with zero diffusion == independent pairs; noise 0 == reference
 *     start vector) and by the convergence check against a tight-tolerance
 *     long-double run (orc_solve_ld, oracle_ld.c).
 */
#ifndef RK_ORACLE_H_
#define RK_ORACLE_H_

#include <stdint.h>
#include <stdio.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ORC_MAX_S 16

enum {
  ORC_BRUSS2D = 1,   /* src/problems/bruss2d-mix.c */
  ORC_LV = 2,        /* src/problems/lv.c (pairs; per-problem constants) */
  ORC_LV_CHAIN = 3   /* diffusive periodic predator-prey chain (no reference counterpart) */
};

typedef struct orc_ode {
  int kind;
  int n;         /* ode_size */
  int grid;      /* BRUSS2D: N; LV: number of pairs; LV_CHAIN: cells M */
  double alpha;  /* BRUSS2D diffusion alpha */
  double c0, d0, d1, c1; /* LV: x' = c0 x - d0 x y, y' = d1 x y - c1 y (lv.c:31-34) */
  double diff;   /* LV_CHAIN diffusion coefficient */
} orc_ode;

typedef struct orc_method {
  char name[32];
  int s, ord, has_emb;
  double A[ORC_MAX_S][ORC_MAX_S];
  double b[ORC_MAX_S], b_hat[ORC_MAX_S], c[ORC_MAX_S];
} orc_method;

typedef struct orc_opts {
  double h0;         /* > 0: fixed initial step (H_INIT); else initial_stepsize() rule */
  int step_limit;    /* STEP_LIMIT (<= 0: unlimited) */
  int block;         /* PipeD/PipeDls block size B (0: access distance) */
  /* optional per-attempt trace (t, h, err) of at most trace_cap attempts */
  double *trace;
  int trace_cap;
  /* optional DUMP-format file (include/tools.h:239-256) */
  FILE *dump;
} orc_opts;

typedef struct orc_stats {
  int steps_acc, steps_rej;
  double t, h;
  double kernel_time;   /* seconds, the step loop only (src/tools.c:77-78) */
} orc_stats;

/* problems ---------------------------------------------------------------- */
int orc_ode_init_bruss2d(orc_ode *p, int N, double alpha);
int orc_ode_init_lv(orc_ode *p, int pairs, double c0, double d0, double d1, double c1);
int orc_ode_init_lv_chain(orc_ode *p, int cells, double c0, double d0, double d1,
                          double c1, double diff);
double orc_eval_comp(const orc_ode *p, int i, double t, const double *y);
void orc_eval_rng(const orc_ode *p, int i, int j, double t, const double *y, double *f);
int orc_acc_dist(const orc_ode *p);
void orc_ode_start(const orc_ode *p, double t, double *y0);

/* methods ------------------------------------------------------------------ */
int orc_method_by_name(const char *name, orc_method *m);

/* step control (exported for the synthetic-value tests) -------------------- */
double orc_h_new(double h, double err, double tol, int ord);
double orc_update_error_max(double emax, double ej, double yj, double yj_old);
void orc_pow(const double *x, const double *y, double *p, size_t n);
double orc_initial_stepsize(double H, double tol, int step_limit, double h0);

/* solvers ------------------------------------------------------------------- */
int orc_solve_D(const orc_ode *p, const orc_method *m, double t0, double te,
                const double *y0, double *y, double tol, const orc_opts *o,
                orc_stats *st);
int orc_solve_PipeD(const orc_ode *p, const orc_method *m, double t0, double te,
                    const double *y0, double *y, double tol, const orc_opts *o,
                    orc_stats *st);
int orc_solve_PipeDls(const orc_ode *p, const orc_method *m, double t0, double te,
                      const double *y0, double *y, double tol, const orc_opts *o,
                      orc_stats *st);
/* gather forms (src/impl/seq/A.c, E.c, F.c) */
int orc_solve_A(const orc_ode *p, const orc_method *m, double t0, double te, const double *y0,
                double *y, double tol, const orc_opts *o, orc_stats *st);
int orc_solve_E(const orc_ode *p, const orc_method *m, double t0, double te, const double *y0,
                double *y, double tol, const orc_opts *o, orc_stats *st);
int orc_solve_F(const orc_ode *p, const orc_method *m, double t0, double te, const double *y0,
                double *y, double tol, const orc_opts *o, orc_stats *st);
int orc_solve_pthreads_D(const orc_ode *p, const orc_method *m, double t0, double te,
                         const double *y0, double *y, double tol, const orc_opts *o,
                         int threads, orc_stats *st);
/* independent LV pairs with per-system constants, each through orc_solve_D */
int orc_lv_ensemble(int count, const double *params, const double *y0, double *y, double t0,
                    double te, const orc_method *m, double tol, double h0, int step_limit,
                    int *acc, int *rej, int threads);
/* accuracy reference: the D loop in long double at any tolerance (not a
 * parity reference); y0 and y are double, the state is long double */
int orc_solve_ld(const orc_ode *p, const char *method, double t0, double te, const double *y0,
                 double *y, double tol, double h0, int step_limit, orc_stats *st);

#ifdef __cplusplus
}
#endif

#endif
