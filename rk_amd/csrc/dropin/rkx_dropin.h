/*
 * The reference's three plugin interfaces, provided on top of the rkx C ABI
 * so that a reference-style driver links against the GPU solver unchanged.
 *
 *   problem (include/ode.h:29-36):   ode_size, ode_start, ode_eval_comp,
 *                                    ode_eval_rng, ode_acc_dist
 *   methods (include/methods.h:92-118): rk_method_t HEUN2 ... DOPRI87,
 *                                    allocated in the layout of
 *                                    alloc_emb_rk_method (methods.h:31-42)
 *   solver  (include/solver.h:47):   solver(t0, te, y0, y, tol)
 *
 * The problem and the method are chosen at build time exactly as in the
 * reference (problem object at link time; METHOD, BRUSS_GRID_SIZE,
 * BRUSS_ALPHA, LV_COUNT, T_START, T_END, TOL, H_INIT, STEP_LIMIT as -D
 * defines with the reference's defaults, include/config.h:30-90).  A GPU
 * kernel cannot call the host ode_eval_* callbacks, so each problem object
 * also names its device right-hand side: rkx_dropin_problem().
 */
#ifndef RKX_DROPIN_H_
#define RKX_DROPIN_H_

#include "rkx.h"

/* build-time parameters, defaults of include/config.h:30-90 */
#ifndef BRUSS_GRID_SIZE
#define BRUSS_GRID_SIZE 250
#endif
#ifndef BRUSS_ALPHA
#define BRUSS_ALPHA 2E-3
#endif
#ifndef LV_COUNT
#define LV_COUNT 10000
#endif
#ifndef T_START
#define T_START 0.0
#endif
#ifndef T_END
#define T_END 0.1
#endif
#ifndef TOL
#define TOL 1E-9
#endif
#ifndef STEP_LIMIT
#define STEP_LIMIT 200
#endif
#ifndef METHOD
#define METHOD DOPRI54
#endif
/* PipeDls (fused, two vectors resident) or D (stage kernels, s+2 vectors) */
#ifndef RKX_DROPIN_IMPL
#define RKX_DROPIN_IMPL RKX_IMPL_PIPEDLS
#endif

#ifdef __cplusplus
extern "C" {
#endif

/* ---- problem plugin (ode.h:29-36) ---------------------------------------- */
extern int ode_size;
void ode_start(double t, double *y0);
double ode_eval_comp(int i, double t, const double *y);
void ode_eval_rng(int i, int j, double t, const double *y, double *f);
int ode_acc_dist(void);
/* the device descriptor of the linked problem */
void rkx_dropin_problem(rkx_ode *p);

/* ---- method plugin (methods.h:92-118) ------------------------------------- */
typedef void(rk_method_t)(double ***A, double **b, double **b_hat, double **c, int *s, int *ord);
rk_method_t HEUN2, RK4, SSPRK3, RKF23, DOPRI54, VERNER65, DOPRI87;
#define HAS_EMB_SOL(MNAME) RKX_HAS_EMB_SOL_(MNAME)
#define RKX_HAS_EMB_SOL_(MNAME) HAS_EMB_SOL_##MNAME()
#define HAS_EMB_SOL_HEUN2() 0
#define HAS_EMB_SOL_RK4() 0
#define HAS_EMB_SOL_SSPRK3() 0
#define HAS_EMB_SOL_RKF23() 1
#define HAS_EMB_SOL_DOPRI54() 1
#define HAS_EMB_SOL_VERNER65() 1
#define HAS_EMB_SOL_DOPRI87() 1
/* release a table returned by a method function (the layout of
 * alloc_emb_rk_method: row pointers in one block, coefficients in another) */
void rkx_dropin_free_method(double ***A, double **b, double **b_hat, double **c);

/* ---- solver plugin (solver.h:47) ------------------------------------------ */
void solver(double t0, double te, double *y0, double *y, double tol);

#ifdef __cplusplus
}
#endif

#endif
