/*
 * solver() of include/solver.h:47 on the GPU: the link-time replacement of a
 * reference implementation object (src/impl/seq/D.c:23-90).  Same contract:
 * y0 is read, the result is returned in y, the caller owns both; the method
 * is METHOD and the step-size rule is initial_stepsize (src/tools.c:25-36)
 * with H_INIT / STEP_LIMIT.  The same lines as the reference's solvers are
 * printed (solver banner, "Method:", "Initial step size:", and
 * print_statistics, src/tools.c:62-79), so log2tab reads the output.
 * Kernel time is the device step loop only, as the reference's timer
 * covers the step loop only.
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>

#include "rkx_dropin.h"

static const char *variant_name(int impl)
{
  switch (impl) {
  case RKX_IMPL_D:
    return "hipD (stage kernels of D on the device)";
  case RKX_IMPL_PIPED:
    return "hipPipeD (stage-fused temporally blocked sweep, two vectors resident; "
           "the PipeDls kernel)";
  case RKX_IMPL_A:
    return "hipA (gather stage kernels skipping zero coefficients on the device)";
  case RKX_IMPL_E:
    return "hipE (gather stage kernels on the device)";
  case RKX_IMPL_F:
    return "hipF (fused RHS and gather stage kernels on the device)";
  default:
    return "hipPipeDls (stage-fused temporally blocked sweep, two vectors resident)";
  }
}

void solver(double t0, double te, double *y0, double *y, double tol)
{
  double **A, *b, *b_hat, *c;
  int s, ord;
  printf("Solver type: device embedded Runge-Kutta method\n");
  printf("Implementation variant: %s\n", variant_name(RKX_DROPIN_IMPL));
  METHOD(&A, &b, &b_hat, &c, &s, &ord);

  rkx_method m;
  if (rkx_method_get("DOPRI54", &m) != RKX_OK) /* any table: overwritten below */
    abort();
  for (int i = 0; i < RKX_MAX_STAGES; i++) {
    for (int j = 0; j < RKX_MAX_STAGES; j++)
      m.A[i][j] = (i < s && j < s) ? A[i][j] : 0.0;
    m.b[i] = i < s ? b[i] : 0.0;
    m.b_hat[i] = i < s ? b_hat[i] : 0.0;
    m.c[i] = i < s ? c[i] : 0.0;
  }
  m.s = s;
  m.ord = ord;
  m.has_emb = HAS_EMB_SOL(METHOD);
  rkx_dropin_free_method(&A, &b, &b_hat, &c);

  /* initial_stepsize (src/tools.c:25-36) */
#ifdef H_INIT
  const double h = H_INIT;
#else
  const double H = te - t0;
  const double h = tol > 0.0 ? fmin(10.0 * tol, H / (STEP_LIMIT)) : H / (STEP_LIMIT);
#endif
  printf("Initial step size: %.e\n", h);

  rkx_ode p;
  rkx_dropin_problem(&p);
  rkx_opts o;
  rkx_opts_default(&o);
  o.h0 = h;
  o.step_limit = STEP_LIMIT;
  o.impl = RKX_DROPIN_IMPL;
#ifdef DUMP
  o.dump_path = "DUMP"; /* every attempted step, as step_control does (tools.h:239-256) */
#endif
  rkx_stats st;
  if (rkx_solve_host(&p, &m, t0, te, y0, y, tol, &o, &st) != RKX_OK) {
    fprintf(stderr, "rkx: %s\n", rkx_last_error());
    exit(1);
  }

  /* print_statistics (src/tools.c:62-79) */
  const int steps = st.steps_acc + st.steps_rej;
  printf("Number of steps: %d (%d accepted, %d rejected)\n", steps, st.steps_acc,
         st.steps_rej);
  if (STEP_LIMIT > 0 && steps >= STEP_LIMIT)
    printf("Step limit reached.\n");
  printf("Kernel time: %.2e s (%.2e s per step)\n", st.kernel_time,
         st.kernel_time / (double) steps);
}
