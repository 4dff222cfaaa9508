/*
 * Accuracy reference: the seq D step loop (src/impl/seq/D.c:23-90, step
 * control include/tools.h:217-288) carried out entirely in long double (x87,
 * 64-bit significand) with the method's rationals evaluated in long double.
 * TEST INFRASTRUCTURE ONLY.
 *
 * This is not a parity reference (its rounding differs from the FP64 path
 * by design); it is run at a tight tolerance to give a reference solution
 * whose own error is far below the FP64 solver's tolerance, so the global
 * error of an FP64 solve at tol 1e-6 / 1e-8 can be stated against it.  The
 * right-hand sides are the same formulas as oracle_problems.c (BRUSS2D with
 * the Neumann mirror, LV pairs, LV chain); operand order is irrelevant here.
 */
#include "oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

int orc_method_ld(const char *name, int *s, int *ord, long double A[ORC_MAX_S][ORC_MAX_S],
                  long double b[ORC_MAX_S], long double b_hat[ORC_MAX_S],
                  long double c[ORC_MAX_S]);

static long double ld_comp(const orc_ode *p, int i, const long double *y)
{
  if (p->kind == ORC_BRUSS2D) {
    const int N = p->grid, N2 = 2 * N;
    const int k = i / N2, j = (i - k * N2) / 2;
    const long double N1 = (long double) N - 1.0L;
    const long double yn = (k == 0) ? y[i + N2] : y[i - N2];
    const long double ys = (k == N - 1) ? y[i - N2] : y[i + N2];
    const long double yw = (j == 0) ? y[i + 2] : y[i - 2];
    const long double ye = (j == N - 1) ? y[i - 2] : y[i + 2];
    const long double lap = yn + ys + yw + ye - 4.0L * y[i];
    const long double coef = (long double) p->alpha * N1 * N1;
    const long double u = (i & 1) ? y[i - 1] : y[i];
    const long double v = (i & 1) ? y[i] : y[i + 1];
    if (!(i & 1))
      return 1.0L + u * u * v - 4.4L * u + coef * lap;
    return 3.4L * u - u * u * v + coef * lap;
  }
  const long double x = (i & 1) ? y[i - 1] : y[i];
  const long double z = (i & 1) ? y[i] : y[i + 1];
  long double f = (i & 1) ? (long double) p->d1 * x * z - (long double) p->c1 * z
                          : (long double) p->c0 * x - (long double) p->d0 * x * z;
  if (p->kind == ORC_LV_CHAIN) {
    const int n = p->n;
    const int il = (i >= 2) ? i - 2 : i - 2 + n;
    const int ir = (i + 2 < n) ? i + 2 : i + 2 - n;
    f += (long double) p->diff * (y[il] + y[ir] - 2.0L * y[i]);
  }
  return f;
}

int orc_solve_ld(const orc_ode *p, const char *method, double t0, double te, const double *y0,
                 double *y_out, double tol, double h0, int step_limit, orc_stats *st)
{
  int s, ord;
  long double A[ORC_MAX_S][ORC_MAX_S], b[ORC_MAX_S], bh[ORC_MAX_S], c[ORC_MAX_S];
  if (!p || orc_method_ld(method, &s, &ord, A, b, bh, c))
    return -1;
  const int n = p->n;
  long double *y = malloc(sizeof(long double) * (size_t) n);
  long double *yn = malloc(sizeof(long double) * (size_t) n);
  long double *w = malloc(sizeof(long double) * (size_t) n);
  long double *F = malloc(sizeof(long double) * (size_t) n * (size_t) s);
  if (!y || !yn || !w || !F) {
    free(y); free(yn); free(w); free(F);
    return -1;
  }
  for (int j = 0; j < n; j++)
    y[j] = y0[j];
  long double t = t0, h = h0 > 0.0 ? h0 : fmin(10.0 * tol, te - t0);
  const long double T = te;
  int acc = 0, rej = 0;
  while (t < T && (step_limit <= 0 || acc + rej < step_limit)) {
    for (int i = 0; i < s; i++) {
      for (int j = 0; j < n; j++) {
        long double v = y[j];
        for (int l = 0; l < i; l++)
          v += h * A[i][l] * F[(size_t) l * n + j];
        w[j] = v;
      }
      for (int j = 0; j < n; j++)
        F[(size_t) i * n + j] = ld_comp(p, j, w);
    }
    long double emax = 0.0L;
    for (int j = 0; j < n; j++) {
      long double dy = 0.0L, e = 0.0L;
      for (int i = 0; i < s; i++) {
        dy += h * b[i] * F[(size_t) i * n + j];
        e += h * (b[i] - bh[i]) * F[(size_t) i * n + j];
      }
      yn[j] = y[j] + dy;
      const long double d = fmaxl(fabsl(yn[j]), fabsl(y[j]));
      emax = (d == 0.0L) ? 1.7E308L : fmaxl(emax, fabsl(e) / d);
    }
    const long double tl = tol;
    if (tol <= 0.0 || emax <= tl) {
      t += h;
      memcpy(y, yn, sizeof(long double) * (size_t) n);
      acc++;
    } else {
      rej++;
    }
    if (tol > 0.0)
      h *= fminl(2.0L, fmaxl(0.5L, 0.9L * powl(tl / emax, 1.0L / (ord + 1.0L))));
    h = fminl(h, T - t);
  }
  for (int j = 0; j < n; j++)
    y_out[j] = (double) y[j];
  if (st) {
    st->steps_acc = acc;
    st->steps_rej = rej;
    st->t = (double) t;
    st->h = (double) h;
    st->kernel_time = 0.0;
  }
  free(y); free(yn); free(w); free(F);
  return 0;
}
