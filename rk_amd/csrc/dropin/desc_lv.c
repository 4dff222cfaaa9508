/*
 * Device descriptor of the LV problem (src/problems/lv.c) for solver_hip.c:
 * LV_COUNT pairs x' = c0 x - d0 x y, y' = d1 x y - c1 y with the reference's
 * constants c0 = 0.1, d0 = 0.2, d1 = 0.4, c1 = 0.2 (lv.c:31-34).  It defines
 * rkx_dropin_problem() only, so it links beside either problem object: the
 * reference's own lv.o or ode_lv.c of this directory.
 */
#include <stdio.h>
#include <stdlib.h>

#include "rkx_dropin.h"

void rkx_dropin_problem(rkx_ode *p)
{
  if (rkx_ode_lv(p, LV_COUNT, 0.1, 0.2, 0.4, 0.2) != RKX_OK) {
    fprintf(stderr, "rkx: %s\n", rkx_last_error());
    exit(1);
  }
}
