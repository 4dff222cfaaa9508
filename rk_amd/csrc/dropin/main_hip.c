/*
 * Driver with the reference's command-line contract and output
 * (src/main.c:40-122): prints the problem banner, the integration interval,
 * the start vector's footprint and the tolerance, calls solver(), then the
 * total time and the result footprint at %.20e, which is what make check
 * style comparisons and log2tab read.  The lock type and pad size lines are
 * printed for log2tab's sake (the device solver has neither).
 */
#include <stdio.h>
#include <stdlib.h>
#include <time.h>

#include "rkx_dropin.h"

static double wall(void)
{
  struct timespec ts;
  clock_gettime(CLOCK_MONOTONIC, &ts);
  return (double) ts.tv_sec + 1e-9 * (double) ts.tv_nsec;
}

int main(void)
{
  const double tol = HAS_EMB_SOL(METHOD) ? TOL : 0.0;
  double *y0 = malloc(sizeof(double) * (size_t) ode_size);
  double *y = malloc(sizeof(double) * (size_t) ode_size);
  if (!y0 || !y)
    return 1;
  ode_start(T_START, y0);
  double mn = y0[0], mx = y0[0];
  for (int i = 1; i < ode_size; i++) {
    mn = y0[i] < mn ? y0[i] : mn;
    mx = y0[i] > mx ? y0[i] : mx;
  }
  printf("Integration interval: [%.2e,%.2e]\n", T_START, T_END);
  printf("Initial value: y0[0] = %e, ..., y0[%d] = %e; min = %e, max = %e\n", y0[0],
         ode_size - 1, y0[ode_size - 1], mn, mx);
  printf("Tolerance: %.2e\n", tol);
  printf("Lock type: none\n");
  printf("Pad size: 0\n");

  const double t_start = wall();
  solver(T_START, T_END, y0, y, tol);
  printf("Total time:  %.2e s\n", wall() - t_start);

  mn = mx = y[0];
  for (int i = 1; i < ode_size; i++) {
    mn = y[i] < mn ? y[i] : mn;
    mx = y[i] > mx ? y[i] : mx;
  }
  printf("Result: y[0] = %.20e, ..., y[%d] = %.20e; min = %.20e, max = %.20e\n", y[0],
         ode_size - 1, y[ode_size - 1], mn, mx);
  free(y0);
  free(y);
  return 0;
}
