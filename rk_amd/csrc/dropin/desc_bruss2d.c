/*
 * Device descriptor of the BRUSS2D-MIX problem (src/problems/bruss2d-mix.c)
 * for solver_hip.c, from the same build-time defines as the reference's
 * problem object (BRUSS_GRID_SIZE, BRUSS_ALPHA; include/config.h:30-36).
 * It defines rkx_dropin_problem() only, so it links beside either problem
 * object: the reference's own bruss2d-mix.o or ode_bruss2d.c of this
 * directory.
 */
#include <stdio.h>
#include <stdlib.h>

#include "rkx_dropin.h"

void rkx_dropin_problem(rkx_ode *p)
{
  if (rkx_ode_bruss2d(p, BRUSS_GRID_SIZE, BRUSS_ALPHA) != RKX_OK) {
    fprintf(stderr, "rkx: %s\n", rkx_last_error());
    exit(1);
  }
}
