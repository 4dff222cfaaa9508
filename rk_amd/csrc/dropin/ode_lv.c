/*
 * Lotka-Volterra problem object (src/problems/lv.c) for the drop-in build:
 * LV_COUNT identical pairs x' = 0.1x - 0.2xy, y' = 0.4xy - 0.2y starting at
 * (0.2, 1.1) (lv.c:31-37).  Banner of lv.c:51-53.  The device descriptor
 * comes from desc_lv.c; the ode_eval_* callbacks evaluate on the device.
 */
#include <stdio.h>
#include <stdlib.h>

#include "rkx_dropin.h"

int ode_size = 2 * LV_COUNT;

void ode_start(double t, double *y0)
{
  rkx_ode p;
  rkx_dropin_problem(&p);
  printf("ODE: LV (Lotka-Volterra, scaled)\n");
  printf("System size: %d\n", ode_size);
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
  return 1;
}
