/*
 * Oracle problems: BRUSS2D-MIX and LV restated from the reference, plus the
 * diffusive LV chain (no reference counterpart).  TEST INFRASTRUCTURE ONLY.
 */
#include "oracle.h"

#include <stdlib.h>

int orc_ode_init_bruss2d(orc_ode *p, int N, double alpha)
{
  if (N < 2)
    return -1;
  p->kind = ORC_BRUSS2D;
  p->grid = N;
  p->n = 2 * N * N;               /* bruss2d-mix.c:29 */
  p->alpha = alpha;
  p->c0 = p->d0 = p->d1 = p->c1 = 0.0;
  p->diff = 0.0;
  return 0;
}

int orc_ode_init_lv(orc_ode *p, int pairs, double c0, double d0, double d1, double c1)
{
  if (pairs < 1)
    return -1;
  p->kind = ORC_LV;
  p->grid = pairs;
  p->n = 2 * pairs;               /* lv.c:29 */
  p->alpha = 0.0;
  p->c0 = c0;
  p->d0 = d0;
  p->d1 = d1;
  p->c1 = c1;
  p->diff = 0.0;
  return 0;
}

int orc_ode_init_lv_chain(orc_ode *p, int cells, double c0, double d0, double d1,
                          double c1, double diff)
{
  if (cells < 3)
    return -1;
  p->kind = ORC_LV_CHAIN;
  p->grid = cells;
  p->n = 2 * cells;
  p->alpha = 0.0;
  p->c0 = c0;
  p->d0 = d0;
  p->d1 = d1;
  p->c1 = c1;
  p->diff = diff;
  return 0;
}

/* ---- BRUSS2D-MIX ------------------------------------------------------------
 * Component i = 2N*row + 2*col + var (var 0 = u, 1 = v).  The Laplacian keeps
 * the reference's per-case operand order (bruss2d-mix.c:66-206): the mirrored
 * Neumann neighbour appears as 2.0*y[...] and is summed first on the top and
 * bottom rows, and last on the left and right columns.
 */
static double bruss_lap(int N, int k, int j, int i, const double *y)
{
  const int N2 = N + N;
  if (k == 0) {
    if (j == 0)
      return 2.0 * y[i + N2] + 2.0 * y[i + 2] - 4.0 * y[i];
    if (j == N - 1)
      return 2.0 * y[i + N2] + 2.0 * y[i - 2] - 4.0 * y[i];
    return 2.0 * y[i + N2] + y[i - 2] + y[i + 2] - 4.0 * y[i];
  } else if (k == N - 1) {
    if (j == 0)
      return 2.0 * y[i - N2] + 2.0 * y[i + 2] - 4.0 * y[i];
    if (j == N - 1)
      return 2.0 * y[i - N2] + 2.0 * y[i - 2] - 4.0 * y[i];
    return 2.0 * y[i - N2] + y[i - 2] + y[i + 2] - 4.0 * y[i];
  } else {
    if (j == 0)
      return y[i - N2] + y[i + N2] + 2.0 * y[i + 2] - 4.0 * y[i];
    if (j == N - 1)
      return y[i - N2] + y[i + N2] + 2.0 * y[i - 2] - 4.0 * y[i];
    return y[i - N2] + y[i + N2] + y[i - 2] + y[i + 2] - 4.0 * y[i];
  }
}

static double bruss_comp(const orc_ode *p, int i, const double *y)
{
  const int N = p->grid;
  const double N1 = (double) N - 1.0;
  const int N2 = N + N;
  const int k = i / N2;
  const int j = (i - k * N2) / 2;
  /* alpha * N1 * N1 * (...) parses as ((alpha*N1)*N1)*(...) */
  const double coef = p->alpha * N1 * N1;
  if (!(i & 1)) {
    const double u = y[i], v = y[i + 1];
    return 1.0 + u * u * v - 4.4 * u + coef * bruss_lap(N, k, j, i, y);
  } else {
    const double u = y[i - 1], v = y[i];
    return 3.4 * u - u * u * v + coef * bruss_lap(N, k, j, i, y);
  }
}

/* ---- LV pairs (lv.c:69-87) ---------------------------------------------- */
static inline double lv_f0(const orc_ode *p, double x, double y)
{
  return p->c0 * x - p->d0 * x * y;
}

static inline double lv_f1(const orc_ode *p, double x, double y)
{
  return p->d1 * x * y - p->c1 * y;
}

/* ---- LV chain ----------------------------------------------------------------
 * Cell c holds (x_c, y_c) at components 2c, 2c+1; periodic in c.
 *   x_c' = (c0 x - d0 x y) + D * ((x_{c-1} + x_{c+1}) - 2 x)
 *   y_c' = (d1 x y - c1 y) + D * ((y_{c-1} + y_{c+1}) - 2 y)
 * Access distance 2.
 */
static double chain_comp(const orc_ode *p, int i, const double *y)
{
  const int n = p->n;
  const int il = (i >= 2) ? i - 2 : i - 2 + n;
  const int ir = (i + 2 < n) ? i + 2 : i + 2 - n;
  const double lap = y[il] + y[ir] - 2.0 * y[i];
  if (!(i & 1))
    return lv_f0(p, y[i], y[i + 1]) + p->diff * lap;
  else
    return lv_f1(p, y[i - 1], y[i]) + p->diff * lap;
}

double orc_eval_comp(const orc_ode *p, int i, double t, const double *y)
{
  (void) t;
  switch (p->kind) {
  case ORC_BRUSS2D:
    return bruss_comp(p, i, y);
  case ORC_LV:
    return (i & 1) ? lv_f1(p, y[i - 1], y[i]) : lv_f0(p, y[i], y[i + 1]);
  case ORC_LV_CHAIN:
    return chain_comp(p, i, y);
  default:
    abort();
  }
}

/* Range form: writes f[i..j] inclusive with absolute indexing, the contract of
 * ode_eval_rng (ode.h:35).  The BRUSS2D interior is a straight loop over
 * rows so the pipelined schedules run at a comparable speed to the
 * reference's range evaluator (bruss2d-mix.c:670-712); every value equals
 * orc_eval_comp bitwise. */
void orc_eval_rng(const orc_ode *p, int i, int j, double t, const double *y, double *f)
{
  if (p->kind == ORC_BRUSS2D) {
    const int N = p->grid;
    const int N2 = N + N;
    const double N1 = (double) N - 1.0;
    const double coef = p->alpha * N1 * N1;
    int q = i;
    while (q <= j) {
      const int k = q / N2;
      const int rowbase = k * N2;
      const int col = (q - rowbase) / 2;
      if (k > 0 && k < N - 1 && col > 0 && col < N - 1 && !(q & 1) && q + 1 <= j) {
        /* a run of interior (u,v) pairs on an interior row: columns
           col .. N-2, both components inside [i, j] */
        int last = rowbase + 2 * (N - 2);   /* u of column N-2 */
        if (last > j - 1)
          last = j - 1;
        for (; q <= last; q += 2) {
          const double u = y[q], v = y[q + 1];
          const double uuv = u * u * v;
          f[q] = 1.0 + uuv - 4.4 * u +
            coef * (y[q - N2] + y[q + N2] + y[q - 2] + y[q + 2] - 4.0 * u);
          f[q + 1] = 3.4 * u - uuv +
            coef * (y[q + 1 - N2] + y[q + 1 + N2] + y[q - 1] + y[q + 3] - 4.0 * v);
        }
        continue;
      }
      f[q] = bruss_comp(p, q, y);
      q++;
    }
    return;
  }
  for (int q = i; q <= j; q++)
    f[q] = orc_eval_comp(p, q, t, y);
}

int orc_acc_dist(const orc_ode *p)
{
  switch (p->kind) {
  case ORC_BRUSS2D:
    return 2 * p->grid;          /* bruss2d-mix.c:718-721 */
  case ORC_LV:
    return 1;                    /* lv.c:127-130 */
  case ORC_LV_CHAIN:
    return 2;
  default:
    abort();
  }
}

void orc_ode_start(const orc_ode *p, double t, double *y0)
{
  (void) t;
  if (p->kind == ORC_BRUSS2D) {
    const int N = p->grid;
    /* bruss2d-mix.c:51-59 */
    for (int i = 0; i < N; i++)
      for (int j = 0; j < N; j++)
        y0[2 * N * i + 2 * j] = 0.5 + (double) j / (double) (N - 1);
    for (int i = 0; i < N; i++)
      for (int j = 0; j < N; j++)
        y0[2 * N * i + 2 * j + 1] = 1.0 + (5.0 * (double) i) / (double) (N - 1);
  } else if (p->kind == ORC_LV) {
    /* lv.c:36-37, 58-62 */
    for (int i = 0; i < p->n; i += 2) {
      y0[i] = 0.2;
      y0[i + 1] = 1.1;
    }
  } else {
    for (int i = 0; i < p->n; i += 2) {
      y0[i] = 0.2;
      y0[i + 1] = 1.1;
    }
  }
}
