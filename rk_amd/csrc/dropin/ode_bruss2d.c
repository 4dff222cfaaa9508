/*
 * BRUSS2D-MIX problem object (src/problems/bruss2d-mix.c) for the drop-in
 * build: ode.h's host interface over the device right-hand side.
 * BRUSS_GRID_SIZE and BRUSS_ALPHA are build-time defines as in the
 * reference; the device descriptor comes from desc_bruss2d.c.  ode_start
 * prints the reference's banner (bruss2d-mix.c:43-46) and fills the start
 * vector of bruss2d-mix.c:51-59; the ode_eval_* callbacks evaluate on the
 * device (rkx_eval_rng_host, rkx_eval_comp_host: only the stencil window of
 * the requested components travels) and equal the reference's values
 * bitwise.
 */
#include <stdio.h>
#include <stdlib.h>

#include "rkx_dropin.h"

int ode_size = 2 * BRUSS_GRID_SIZE * BRUSS_GRID_SIZE;

void ode_start(double t, double *y0)
{
  rkx_ode p;
  rkx_dropin_problem(&p);
  printf("ODE: BRUSS2D-MIX (2D Brusselator with mixed row-oriented ordering)\n");
  printf("Grid size: %d\n", BRUSS_GRID_SIZE);
  printf("Alpha: %.2e\n", BRUSS_ALPHA);
  if (rkx_ode_start(&p, t, y0) != RKX_OK) {
    fprintf(stderr, "rkx: %s\n", rkx_last_error());
    exit(1);
  }
}

void ode_eval_rng(int i, int j, double t, const double *y, double *f)
{
  rkx_ode p;
  rkx_dropin_problem(&p);
  if (rkx_eval_rng_host(&p, i, j, t, y, f) != RKX_OK) {
    fprintf(stderr, "rkx: %s\n", rkx_last_error());
    exit(1);
  }
}

double ode_eval_comp(int i, double t, const double *y)
{
  rkx_ode p;
  rkx_dropin_problem(&p);
  double v;
  if (rkx_eval_comp_host(&p, i, t, y, &v) != RKX_OK) {
    fprintf(stderr, "rkx: %s\n", rkx_last_error());
    exit(1);
  }
  return v;
}

int ode_acc_dist(void)
{
  return 2 * BRUSS_GRID_SIZE;
}
