/*
 * Oracle solvers: seq D (the parity reference), the pipelined PipeD and
 * low-storage PipeDls schedules, the gather forms A, E and F, and pthreads D.
 * TEST INFRASTRUCTURE ONLY.
 */

#include "oracle.h"

#include <math.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>
#include <sys/time.h>

static double now_s(void)
{
  struct timeval tv;
  gettimeofday(&tv, NULL);
  return (double) tv.tv_sec + (double) tv.tv_usec * 1E-6;
}

/* ---- step control (include/tools.h:217-288, src/tools.c:25-36) ---------- */

double orc_h_new(double h, double err, double tol, int ord)
{
  if (tol <= 0.0)
    return h;
  return h * fmin(2.0, fmax(0.5, 0.9 * pow(tol / err, 1.0 / (ord + 1.0))));
}

/* The C library's pow elementwise: the reference's one libm call
 * (include/tools.h:226), as the checker of the device's restatement. */
void orc_pow(const double *x, const double *y, double *p, size_t n)
{
  for (size_t i = 0; i < n; ++i)
    p[i] = pow(x[i], y[i]);
}

double orc_update_error_max(double emax, double ej, double yj, double yj_old)
{
  const double divisor = fmax(fabs(yj), fabs(yj_old));
  if (divisor == 0.0)
    return 1.7E308;
  return fmax(emax, fabs(ej) / divisor);
}

double orc_initial_stepsize(double H, double tol, int step_limit, double h0)
{
  if (h0 > 0.0)
    return h0;                                    /* H_INIT */
  return tol > 0.0 ? fmin(10.0 * tol, H / (double) step_limit)
                   : H / (double) step_limit;
}

/* returns 1 on accept; restores y from y_old on reject (tools.h:235-278) */
static int step_control(double *t, double *h, double err, int ord, double tol,
                        double *y, const double *y_old, int n, int *acc, int *rej,
                        const orc_opts *o)
{
  const int attempt = *acc + *rej;
  if (o && o->trace && attempt < o->trace_cap) {
    o->trace[3 * attempt + 0] = *t;
    o->trace[3 * attempt + 1] = *h;
    o->trace[3 * attempt + 2] = err;
  }
  if (o && o->dump) {
    fprintf(o->dump, "%.20e\t%.20e", *t, *h);
    for (int i = 0; i < n; i++)
      fprintf(o->dump, "\t%.20e", y[i]);
    fprintf(o->dump, "\n");
  }
  if (tol <= 0.0) {
    *t += *h;
    (*acc)++;
    return 1;
  } else if (err <= tol) {
    *t += *h;
    *h = orc_h_new(*h, err, tol, ord);
    (*acc)++;
    return 1;
  } else {
    memcpy(y, y_old, (size_t) n * sizeof(double));
    *h = orc_h_new(*h, err, tol, ord);
    (*rej)++;
    return 0;
  }
}

static inline int keep_going(double t, double te, int acc, int rej, int limit)
{
  return (t < te) && (limit <= 0 || acc + rej < limit);
}

/* premult (src/methods.c:43-58) with e = b - b_hat folded in (seq/D.c:40-41) */
typedef struct {
  int s, ord;
  double hA[ORC_MAX_S][ORC_MAX_S], hb[ORC_MAX_S], he[ORC_MAX_S], hc[ORC_MAX_S];
  int nzA[ORC_MAX_S][ORC_MAX_S], nzb[ORC_MAX_S], nze[ORC_MAX_S];
  double e[ORC_MAX_S];
} coef_t;

static void coef_init(coef_t *k, const orc_method *m)
{
  memset(k, 0, sizeof(*k));
  k->s = m->s;
  k->ord = m->ord;
  for (int i = 0; i < m->s; i++)
    k->e[i] = m->b[i] - m->b_hat[i];
  /* zero pattern (src/methods.c:24-39): 1 means non-zero */
  for (int i = 0; i < m->s; i++) {
    for (int j = 0; j < m->s; j++)
      k->nzA[i][j] = (m->A[i][j] != 0.0);
    k->nzb[i] = (m->b[i] != 0.0);
    k->nze[i] = (k->e[i] != 0.0);
  }
}

static void premult(coef_t *k, const orc_method *m, double h)
{
  for (int i = 0; i < m->s; i++) {
    for (int j = 0; j < i; j++)
      k->hA[i][j] = h * m->A[i][j];
    k->hb[i] = h * m->b[i];
    k->he[i] = h * k->e[i];
    k->hc[i] = h * m->c[i];
  }
}

static int check_args(const orc_ode *p, const orc_method *m)
{
  if (!p || !m || m->s < 2 || m->s > ORC_MAX_S || p->n < 1)
    return -1;
  return 0;
}

/* ---- seq D (src/impl/seq/D.c:23-90, block-inline.h:253-320) -------------- */

int orc_solve_D(const orc_ode *p, const orc_method *m, double t0, double te,
                const double *y0, double *y, double tol, const orc_opts *o,
                orc_stats *st)
{
  if (check_args(p, m))
    return -1;
  const int n = p->n, s = m->s;
  const int limit = o ? o->step_limit : 0;
  coef_t k;
  coef_init(&k, m);
  double *w[ORC_MAX_S];
  for (int l = 0; l < s; l++)
    w[l] = (double *) malloc((size_t) n * sizeof(double));
  double *err = (double *) malloc((size_t) n * sizeof(double));
  double *dy = (double *) malloc((size_t) n * sizeof(double));
  double *y_old = dy;
  int acc = 0, rej = 0;
  double h = orc_initial_stepsize(te - t0, tol, limit, o ? o->h0 : 0.0);
  double t;
  memcpy(y, y0, (size_t) n * sizeof(double));
  const double t_start = now_s();
  for (t = t0; keep_going(t, te, acc, rej, limit); h = fmin(h, te - t)) {
    premult(&k, m, h);
    double err_max = 0.0;
    /* first stage */
    for (int j = 0; j < n; j++) {
      const double F = orc_eval_comp(p, j, t + k.hc[0], y);
      const double Y = y[j];
      for (int l = 1; l < s; l++)
        w[l][j] = Y + k.hA[l][0] * F;
      dy[j] = k.hb[0] * F;
      err[j] = k.he[0] * F;
    }
    /* intermediate stages */
    for (int i = 1; i < s - 1; i++) {
      for (int j = 0; j < n; j++) {
        const double F = orc_eval_comp(p, j, t + k.hc[i], w[i]);
        for (int l = i + 1; l < s; l++)
          w[l][j] += k.hA[l][i] * F;
        dy[j] += k.hb[i] * F;
        err[j] += k.he[i] * F;
      }
    }
    /* last stage, y update, error norm */
    {
      const int i = s - 1;
      for (int j = 0; j < n; j++) {
        const double F = orc_eval_comp(p, j, t + k.hc[i], w[i]);
        dy[j] += k.hb[i] * F;
        err[j] += k.he[i] * F;
        const double yj_old = y[j];
        y[j] += dy[j];
        dy[j] = yj_old;
        err_max = orc_update_error_max(err_max, err[j], y[j], yj_old);
      }
    }
    step_control(&t, &h, err_max, m->ord, tol, y, y_old, n, &acc, &rej, o);
  }
  if (st) {
    st->steps_acc = acc;
    st->steps_rej = rej;
    st->t = t;
    st->h = h;
    st->kernel_time = now_s() - t_start;
  }
  for (int l = 0; l < s; l++)
    free(w[l]);
  free(err);
  free(dy);
  return 0;
}

/* ---- tiled stage kernels (block-inline.h:324-428) ------------------------ */

typedef struct {
  const orc_ode *p;
  const coef_t *k;
  int s, B;
  double *y, *err, *dy, **w, *v;
  double t, err_max;
} pipe_ctx;

static void tiled_first(pipe_ctx *c, int first)
{
  const coef_t *k = c->k;
  const int B = c->B;
  orc_eval_rng(c->p, first, first + B - 1, c->t + k->hc[0], c->y, c->v - first);
  for (int l = 1; l < c->s; l++) {
    double *wl = c->w[l];
    if (k->nzA[l][0])
      for (int jj = 0; jj < B; jj++)
        wl[first + jj] = c->y[first + jj] + k->hA[l][0] * c->v[jj];
    else
      for (int jj = 0; jj < B; jj++)
        wl[first + jj] = c->y[first + jj];
  }
  for (int jj = 0; jj < B; jj++)
    c->dy[first + jj] = k->hb[0] * c->v[jj];
  for (int jj = 0; jj < B; jj++)
    c->err[first + jj] = k->he[0] * c->v[jj];
}

static void tiled_interm(pipe_ctx *c, int i, int first)
{
  const coef_t *k = c->k;
  const int B = c->B;
  orc_eval_rng(c->p, first, first + B - 1, c->t + k->hc[i], c->w[i], c->v - first);
  for (int l = i + 1; l < c->s; l++)
    if (k->nzA[l][i])
      for (int jj = 0; jj < B; jj++)
        c->w[l][first + jj] += k->hA[l][i] * c->v[jj];
  if (k->nzb[i])
    for (int jj = 0; jj < B; jj++)
      c->dy[first + jj] += k->hb[i] * c->v[jj];
  if (k->nze[i])
    for (int jj = 0; jj < B; jj++)
      c->err[first + jj] += k->he[i] * c->v[jj];
}

static void tiled_last(pipe_ctx *c, int first)
{
  const coef_t *k = c->k;
  const int B = c->B, i = c->s - 1;
  orc_eval_rng(c->p, first, first + B - 1, c->t + k->hc[i], c->w[i], c->v - first);
  if (k->nzb[i])
    for (int jj = 0; jj < B; jj++)
      c->dy[first + jj] += k->hb[i] * c->v[jj];
  if (k->nze[i])
    for (int jj = 0; jj < B; jj++)
      c->err[first + jj] += k->he[i] * c->v[jj];
  for (int jj = 0; jj < B; jj++) {
    const int j = first + jj;
    const double yj_old = c->y[j];
    c->y[j] += c->dy[j];
    c->dy[j] = yj_old;
    c->err_max = orc_update_error_max(c->err_max, c->err[j], c->y[j], yj_old);
  }
}

/* The block-diagonal wavefront of seq/PipeD.c:76-124 (identical loops in
 * seq/PipeDls.c:81-129): init triangle, sweep, finalize. */
static void pipe_step(pipe_ctx *c, int n)
{
  const int B = c->B, s = c->s;
  const int k1 = (s - 1) * B, k2 = n - B, k4 = k1 + B;
  int i, j, l;
  for (j = 0; j < k1; j += B) {
    tiled_first(c, j);
    for (l = j, i = 1; l > 0; i++) {
      l -= B;
      tiled_interm(c, i, l);
    }
  }
  for (i = k1; i < n; i += k4) {
    tiled_first(c, i);
    i -= B;
    for (j = 1; j < s - 1; j++) {
      tiled_interm(c, j, i);
      i -= B;
    }
    tiled_last(c, i);
  }
  for (i = 1; i < s; i++) {
    for (j = i, l = k2; j < s - 1; j++, l -= B)
      tiled_interm(c, j, l);
    tiled_last(c, l);
  }
}

static int solve_pipe(const orc_ode *p, const orc_method *m, double t0, double te,
                      const double *y0, double *y, double tol, const orc_opts *o,
                      orc_stats *st, int low_storage)
{
  if (check_args(p, m))
    return -1;
  const int n = p->n, s = m->s;
  const int B = (o && o->block > 0) ? o->block : orc_acc_dist(p);
  if (B < orc_acc_dist(p) || n % B != 0 || n / B < s)
    return -2;
  const int limit = o ? o->step_limit : 0;
  coef_t k;
  coef_init(&k, m);
  double *w[ORC_MAX_S] = {0};
  double *buf = NULL, *err, *dy;
  if (low_storage) {
    /* one overlapped buffer of n + (s^2+5s-4)B/2 doubles (seq/PipeDls.c:49-59) */
    buf = (double *) calloc((size_t) n + (size_t) (s * s + 5 * s - 4) * B / 2, sizeof(double));
    dy = buf;
    err = dy + s * B;
    w[1] = err + 3 * B;
    for (int i = 2; i < s; i++)
      w[i] = w[i - 1] + (i + 2) * B;
  } else {
    for (int l = 0; l < s; l++)
      w[l] = (double *) malloc((size_t) n * sizeof(double));
    err = (double *) malloc((size_t) n * sizeof(double));
    dy = (double *) malloc((size_t) n * sizeof(double));
  }
  double *v = (double *) malloc((size_t) B * sizeof(double));
  pipe_ctx c = {p, &k, s, B, y, err, dy, w, v, 0.0, 0.0};
  int acc = 0, rej = 0;
  double h = orc_initial_stepsize(te - t0, tol, limit, o ? o->h0 : 0.0);
  double t;
  memcpy(y, y0, (size_t) n * sizeof(double));
  const double t_start = now_s();
  for (t = t0; keep_going(t, te, acc, rej, limit); h = fmin(h, te - t)) {
    premult(&k, m, h);
    c.t = t;
    c.err_max = 0.0;
    pipe_step(&c, n);
    step_control(&t, &h, c.err_max, m->ord, tol, y, dy, n, &acc, &rej, o);
  }
  if (st) {
    st->steps_acc = acc;
    st->steps_rej = rej;
    st->t = t;
    st->h = h;
    st->kernel_time = now_s() - t_start;
  }
  if (low_storage) {
    free(buf);
  } else {
    for (int l = 0; l < s; l++)
      free(w[l]);
    free(err);
    free(dy);
  }
  free(v);
  return 0;
}

int orc_solve_PipeD(const orc_ode *p, const orc_method *m, double t0, double te,
                    const double *y0, double *y, double tol, const orc_opts *o,
                    orc_stats *st)
{
  return solve_pipe(p, m, t0, te, y0, y, tol, o, st, 0);
}

int orc_solve_PipeDls(const orc_ode *p, const orc_method *m, double t0, double te,
                      const double *y0, double *y, double tol, const orc_opts *o,
                      orc_stats *st)
{
  return solve_pipe(p, m, t0, te, y0, y, tol, o, st, 1);
}

/* ---- gather forms A, E, F (src/impl/seq/A.c:23-103, E.c:23-101, F.c:23-106)
 * Stage vectors v_0..v_{s-1} are kept whole.  A and E evaluate v_0 = f(y),
 * then per stage gather w = y + sum_{i<l} hA[l][i] v_i (A: non-zero entries
 * only, block-inline.h:76-92; E: every entry, :38-50) and v_l = f(w).  F
 * evaluates v_l = f(w_l) and gathers w_{l+1} in the same loop over j
 * (block_rhs_gather_interm_stage, :165-182, w_0 = y), the last stage on its
 * own.  The output gathers err and dy (A: block_gather_vec_output, :96-134;
 * E and F: block_gather_output, :54-72) and updates y (A.c:85-92). */

enum { GATHER_A, GATHER_E, GATHER_F };

static int solve_gather(const orc_ode *p, const orc_method *m, double t0, double te,
                        const double *y0, double *y, double tol, const orc_opts *o,
                        orc_stats *st, int form)
{
  if (check_args(p, m))
    return -1;
  const int n = p->n, s = m->s;
  const int limit = o ? o->step_limit : 0;
  const int skip = (form == GATHER_A);
  coef_t k;
  coef_init(&k, m);
  double *v[ORC_MAX_S];
  for (int l = 0; l < s; l++)
    v[l] = (double *) malloc((size_t) n * sizeof(double));
  double *wa = (double *) malloc((size_t) n * sizeof(double));
  double *wb = (double *) malloc((size_t) n * sizeof(double));
  double *err = (double *) malloc((size_t) n * sizeof(double));
  double *dy = (double *) malloc((size_t) n * sizeof(double));
  int acc = 0, rej = 0;
  double h = orc_initial_stepsize(te - t0, tol, limit, o ? o->h0 : 0.0);
  double t;
  memcpy(y, y0, (size_t) n * sizeof(double));
  const double t_start = now_s();
  for (t = t0; keep_going(t, te, acc, rej, limit); h = fmin(h, te - t)) {
    premult(&k, m, h);
    if (form == GATHER_F) {
      double *wbuf[2] = {wa, wb};  /* w_cur / w_next, swapped per stage */
      const double *wl = y;        /* w_0 = y */
      for (int l = 0; l < s - 1; l++) {
        double *wn = wbuf[l & 1];
        for (int j = 0; j < n; j++) {
          v[l][j] = orc_eval_comp(p, j, t + k.hc[l], wl);
          double a = y[j];
          for (int i = 0; i <= l; i++)
            a += k.hA[l + 1][i] * v[i][j];
          wn[j] = a;
        }
        wl = wn;
      }
      orc_eval_rng(p, 0, n - 1, t + k.hc[s - 1], wl, v[s - 1]);
    } else {
      orc_eval_rng(p, 0, n - 1, t + k.hc[0], y, v[0]);
      for (int l = 1; l < s; l++) {
        if (skip) {
          if (k.nzA[l][0])
            for (int j = 0; j < n; j++)
              wa[j] = y[j] + k.hA[l][0] * v[0][j];
          else
            memcpy(wa, y, (size_t) n * sizeof(double));
          for (int i = 1; i < l; i++)
            if (k.nzA[l][i])
              for (int j = 0; j < n; j++)
                wa[j] += k.hA[l][i] * v[i][j];
        } else {
          for (int j = 0; j < n; j++) {
            double a = y[j];
            for (int i = 0; i < l; i++)
              a += k.hA[l][i] * v[i][j];
            wa[j] = a;
          }
        }
        orc_eval_rng(p, 0, n - 1, t + k.hc[l], wa, v[l]);
      }
    }
    for (int j = 0; j < n; j++) {
      err[j] = k.he[0] * v[0][j];
      dy[j] = k.hb[0] * v[0][j];
    }
    for (int i = 1; i < s; i++) {
      const int te_ = !skip || k.nze[i], tb = !skip || k.nzb[i];
      for (int j = 0; j < n; j++) {
        if (te_)
          err[j] += k.he[i] * v[i][j];
        if (tb)
          dy[j] += k.hb[i] * v[i][j];
      }
    }
    double err_max = 0.0;
    for (int j = 0; j < n; j++) {
      const double yj_old = y[j];
      y[j] += dy[j];
      dy[j] = yj_old;              /* dy now holds y_old (A.c:85-90) */
      err_max = orc_update_error_max(err_max, err[j], y[j], yj_old);
    }
    step_control(&t, &h, err_max, m->ord, tol, y, dy, n, &acc, &rej, o);
  }
  if (st) {
    st->steps_acc = acc;
    st->steps_rej = rej;
    st->t = t;
    st->h = h;
    st->kernel_time = now_s() - t_start;
  }
  for (int l = 0; l < s; l++)
    free(v[l]);
  free(wa);
  free(wb);
  free(err);
  free(dy);
  return 0;
}

int orc_solve_A(const orc_ode *p, const orc_method *m, double t0, double te, const double *y0,
                double *y, double tol, const orc_opts *o, orc_stats *st)
{
  return solve_gather(p, m, t0, te, y0, y, tol, o, st, GATHER_A);
}

int orc_solve_E(const orc_ode *p, const orc_method *m, double t0, double te, const double *y0,
                double *y, double tol, const orc_opts *o, orc_stats *st)
{
  return solve_gather(p, m, t0, te, y0, y, tol, o, st, GATHER_E);
}

int orc_solve_F(const orc_ode *p, const orc_method *m, double t0, double te, const double *y0,
                double *y, double tol, const orc_opts *o, orc_stats *st)
{
  return solve_gather(p, m, t0, te, y0, y, tol, o, st, GATHER_F);
}

/* ---- pthreads D (src/impl/pthreads/D.c:50-218) --------------------------- */

typedef struct {
  const orc_ode *p;
  const orc_method *m;
  double t0, te, tol;
  const double *y0;
  double *y, **w, *err, *dy;
  int threads, limit;
  double h0;
  pthread_barrier_t bar;
  double *partial;          /* per-thread err_max for the max reduction */
  int *first, *size;
  int acc, rej;
  double t_final, h_final, kernel_time;
} pth_shared;

typedef struct {
  pth_shared *sh;
  int me;
} pth_arg;

static void *pth_worker(void *arg)
{
  pth_shared *sh = ((pth_arg *) arg)->sh;
  const int me = ((pth_arg *) arg)->me;
  const orc_ode *p = sh->p;
  const orc_method *m = sh->m;
  const int s = m->s;
  const int lo = sh->first[me], hi = sh->first[me] + sh->size[me];
  coef_t k;
  coef_init(&k, m);
  double **w = sh->w, *err = sh->err, *dy = sh->dy, *y = sh->y;
  const double tol = sh->tol, te = sh->te;
  int acc = 0, rej = 0;
  double h = orc_initial_stepsize(sh->te - sh->t0, tol, sh->limit, sh->h0);
  double t;
  memcpy(y + lo, sh->y0 + lo, (size_t) (hi - lo) * sizeof(double));
  pthread_barrier_wait(&sh->bar);
  const double t_start = now_s();
  for (t = sh->t0; keep_going(t, te, acc, rej, sh->limit); h = fmin(h, te - t)) {
    premult(&k, m, h);
    double err_max = 0.0;
    for (int j = lo; j < hi; j++) {
      const double F = orc_eval_comp(p, j, t + k.hc[0], y);
      const double Y = y[j];
      for (int l = 1; l < s; l++)
        w[l][j] = Y + k.hA[l][0] * F;
      dy[j] = k.hb[0] * F;
      err[j] = k.he[0] * F;
    }
    for (int i = 1; i < s - 1; i++) {
      pthread_barrier_wait(&sh->bar);
      for (int j = lo; j < hi; j++) {
        const double F = orc_eval_comp(p, j, t + k.hc[i], w[i]);
        for (int l = i + 1; l < s; l++)
          w[l][j] += k.hA[l][i] * F;
        dy[j] += k.hb[i] * F;
        err[j] += k.he[i] * F;
      }
    }
    pthread_barrier_wait(&sh->bar);
    {
      const int i = s - 1;
      for (int j = lo; j < hi; j++) {
        const double F = orc_eval_comp(p, j, t + k.hc[i], w[i]);
        dy[j] += k.hb[i] * F;
        err[j] += k.he[i] * F;
        const double yj_old = y[j];
        y[j] += dy[j];
        dy[j] = yj_old;
        err_max = orc_update_error_max(err_max, err[j], y[j], yj_old);
      }
    }
    /* reduction_max (src/threads.c:53-87) */
    sh->partial[me] = err_max;
    pthread_barrier_wait(&sh->bar);
    double red = sh->partial[0];
    for (int q = 1; q < sh->threads; q++)
      red = fmax(red, sh->partial[q]);
    step_control(&t, &h, red, m->ord, tol, y + lo, dy + lo, hi - lo, &acc, &rej, NULL);
    pthread_barrier_wait(&sh->bar);
  }
  if (me == 0) {
    sh->acc = acc;
    sh->rej = rej;
    sh->t_final = t;
    sh->h_final = h;
    sh->kernel_time = now_s() - t_start;
  }
  return NULL;
}

int orc_solve_pthreads_D(const orc_ode *p, const orc_method *m, double t0, double te,
                         const double *y0, double *y, double tol, const orc_opts *o,
                         int threads, orc_stats *st)
{
  if (check_args(p, m) || threads < 1)
    return -1;
  const int n = p->n, s = m->s;
  if (threads > n)
    threads = n;
  pth_shared sh;
  memset(&sh, 0, sizeof(sh));
  sh.p = p;
  sh.m = m;
  sh.t0 = t0;
  sh.te = te;
  sh.tol = tol;
  sh.y0 = y0;
  sh.y = y;
  sh.threads = threads;
  sh.limit = o ? o->step_limit : 0;
  sh.h0 = o ? o->h0 : 0.0;
  double *w[ORC_MAX_S];
  for (int l = 0; l < s; l++)
    w[l] = (double *) malloc((size_t) n * sizeof(double));
  sh.w = w;
  sh.err = (double *) malloc((size_t) n * sizeof(double));
  sh.dy = (double *) malloc((size_t) n * sizeof(double));
  sh.partial = (double *) calloc((size_t) threads, sizeof(double));
  sh.first = (int *) malloc((size_t) threads * sizeof(int));
  sh.size = (int *) malloc((size_t) threads * sizeof(int));
  /* blockwise_distribution (src/tools.c:42-56) */
  {
    const int q = n / threads, r = n % threads;
    for (int i = 0; i < threads; i++) {
      sh.first[i] = i < r ? i * (q + 1) : r * (q + 1) + (i - r) * q;
      sh.size[i] = i < r ? q + 1 : q;
    }
  }
  pthread_barrier_init(&sh.bar, NULL, (unsigned) threads);
  pthread_t *tid = (pthread_t *) malloc((size_t) threads * sizeof(pthread_t));
  pth_arg *args = (pth_arg *) malloc((size_t) threads * sizeof(pth_arg));
  for (int i = 0; i < threads; i++) {
    args[i].sh = &sh;
    args[i].me = i;
  }
  for (int i = 1; i < threads; i++)
    pthread_create(&tid[i], NULL, pth_worker, &args[i]);
  pth_worker(&args[0]);
  for (int i = 1; i < threads; i++)
    pthread_join(tid[i], NULL);
  pthread_barrier_destroy(&sh.bar);
  if (st) {
    st->steps_acc = sh.acc;
    st->steps_rej = sh.rej;
    st->t = sh.t_final;
    st->h = sh.h_final;
    st->kernel_time = sh.kernel_time;
  }
  for (int l = 0; l < s; l++)
    free(w[l]);
  free(sh.err);
  free(sh.dy);
  free(sh.partial);
  free(sh.first);
  free(sh.size);
  free(tid);
  free(args);
  return 0;
}

/* ---- ensemble of independent LV pairs (checker of rkx_lv_ensemble) --------
 * Each system k is the reference's LV problem (src/problems/lv.c:69-77) with
 * one pair and its own constants params[4k..4k+3] = (c0, d0, d1, c1), solved
 * by orc_solve_D from y0[2k..2k+1]; systems are shared out over `threads`
 * POSIX threads in contiguous ranges. */
typedef struct {
  int first, count;
  const double *params, *y0;
  double *y;
  double t0, te, tol, h0;
  int step_limit;
  const orc_method *m;
  int *acc, *rej;
  int rc;
} ens_arg;

static void *ens_worker(void *vp)
{
  ens_arg *a = (ens_arg *) vp;
  orc_opts o;
  memset(&o, 0, sizeof(o));
  o.h0 = a->h0;
  o.step_limit = a->step_limit;
  for (int k = a->first; k < a->first + a->count && !a->rc; k++) {
    orc_ode p;
    const double *q = a->params + 4 * (size_t) k;
    orc_ode_init_lv(&p, 1, q[0], q[1], q[2], q[3]);
    orc_stats st;
    a->rc = orc_solve_D(&p, a->m, a->t0, a->te, a->y0 + 2 * (size_t) k, a->y + 2 * (size_t) k,
                        a->tol, &o, &st);
    a->acc[k] = st.steps_acc;
    a->rej[k] = st.steps_rej;
  }
  return NULL;
}

int orc_lv_ensemble(int count, const double *params, const double *y0, double *y, double t0,
                    double te, const orc_method *m, double tol, double h0, int step_limit,
                    int *acc, int *rej, int threads)
{
  if (count < 1 || threads < 1)
    return -1;
  if (threads > count)
    threads = count;
  ens_arg *args = (ens_arg *) calloc((size_t) threads, sizeof(ens_arg));
  pthread_t *tid = (pthread_t *) malloc((size_t) threads * sizeof(pthread_t));
  const int q = count / threads, r = count % threads;
  for (int i = 0; i < threads; i++) {
    ens_arg *a = &args[i];
    a->first = i < r ? i * (q + 1) : r * (q + 1) + (i - r) * q;
    a->count = i < r ? q + 1 : q;
    a->params = params;
    a->y0 = y0;
    a->y = y;
    a->t0 = t0;
    a->te = te;
    a->tol = tol;
    a->h0 = h0;
    a->step_limit = step_limit;
    a->m = m;
    a->acc = acc;
    a->rej = rej;
  }
  for (int i = 1; i < threads; i++)
    pthread_create(&tid[i], NULL, ens_worker, &args[i]);
  ens_worker(&args[0]);
  int rc = args[0].rc;
  for (int i = 1; i < threads; i++) {
    pthread_join(tid[i], NULL);
    if (args[i].rc)
      rc = args[i].rc;
  }
  free(args);
  free(tid);
  return rc;
}
