/*
 * rk_oracle: command-line driver over the oracle library, printing the
 * reference's output contract (src/main.c:75-109, src/tools.c:69-78).
 * TEST INFRASTRUCTURE ONLY.
 *
 *   rk_oracle --problem bruss2d|lv|lvchain --grid N [--alpha A] --method M
 *             --tol TOL --t0 T0 --te TE [--h0 H] [--step-limit K]
 *             [--impl D|PipeD|PipeDls|pthreads_D] [--threads P]
 */
#include "oracle.h"

#include <stdio.h>
#include <stdlib.h>
#include <string.h>

int main(int argc, char **argv)
{
  const char *problem = "bruss2d", *method = "DOPRI54", *impl = "D";
  int grid = 128, step_limit = 0, threads = 1;
  double alpha = 2e-3, tol = 1e-6, t0 = 0.0, te = 4.0, h0 = 1e-3, diff = 1e-3;
  for (int i = 1; i + 1 < argc; i += 2) {
    const char *k = argv[i], *v = argv[i + 1];
    if (!strcmp(k, "--problem")) problem = v;
    else if (!strcmp(k, "--grid")) grid = atoi(v);
    else if (!strcmp(k, "--alpha")) alpha = atof(v);
    else if (!strcmp(k, "--diff")) diff = atof(v);
    else if (!strcmp(k, "--method")) method = v;
    else if (!strcmp(k, "--tol")) tol = atof(v);
    else if (!strcmp(k, "--t0")) t0 = atof(v);
    else if (!strcmp(k, "--te")) te = atof(v);
    else if (!strcmp(k, "--h0")) h0 = atof(v);
    else if (!strcmp(k, "--step-limit")) step_limit = atoi(v);
    else if (!strcmp(k, "--impl")) impl = v;
    else if (!strcmp(k, "--threads")) threads = atoi(v);
    else { fprintf(stderr, "unknown option %s\n", k); return 2; }
  }
  orc_ode p;
  int rc;
  if (!strcmp(problem, "bruss2d")) rc = orc_ode_init_bruss2d(&p, grid, alpha);
  else if (!strcmp(problem, "lv")) rc = orc_ode_init_lv(&p, grid, 0.1, 0.2, 0.4, 0.2);
  else rc = orc_ode_init_lv_chain(&p, grid, 0.1, 0.2, 0.4, 0.2, diff);
  if (rc) { fprintf(stderr, "bad problem\n"); return 2; }
  orc_method m;
  if (orc_method_by_name(method, &m)) { fprintf(stderr, "bad method\n"); return 2; }
  if (!m.has_emb) tol = 0.0;     /* src/main.c:47 */
  double *y0 = malloc((size_t) p.n * sizeof(double));
  double *y = malloc((size_t) p.n * sizeof(double));
  orc_ode_start(&p, t0, y0);
  orc_opts o = {0};
  o.h0 = h0;
  o.step_limit = step_limit;
  orc_stats st;
  if (!strcmp(impl, "D")) rc = orc_solve_D(&p, &m, t0, te, y0, y, tol, &o, &st);
  else if (!strcmp(impl, "PipeD")) rc = orc_solve_PipeD(&p, &m, t0, te, y0, y, tol, &o, &st);
  else if (!strcmp(impl, "PipeDls")) rc = orc_solve_PipeDls(&p, &m, t0, te, y0, y, tol, &o, &st);
  else rc = orc_solve_pthreads_D(&p, &m, t0, te, y0, y, tol, &o, threads, &st);
  if (rc) { fprintf(stderr, "solve failed (%d)\n", rc); return 1; }
  printf("Number of steps: %d (%d accepted, %d rejected)\n",
         st.steps_acc + st.steps_rej, st.steps_acc, st.steps_rej);
  if (step_limit > 0 && st.steps_acc + st.steps_rej >= step_limit)
    printf("Step limit reached.\n");
  printf("Kernel time: %.2e s (%.2e s per step)\n", st.kernel_time,
         st.kernel_time / (double) (st.steps_acc + st.steps_rej));
  double mn = y[0], mx = y[0];
  for (int i = 1; i < p.n; i++)
    if (y[i] < mn) mn = y[i];
    else if (y[i] > mx) mx = y[i];
  printf("Result: y[0] = %.20e, ..., y[%d] = %.20e; min = %.20e, max = %.20e\n",
         y[0], p.n - 1, y[p.n - 1], mn, mx);
  free(y0);
  free(y);
  return 0;
}
