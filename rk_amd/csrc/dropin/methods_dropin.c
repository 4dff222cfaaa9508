/*
 * The reference's method functions (rk_method_t, include/methods.h:92-118)
 * over the library's tables (rkx_method_get): each allocates the table in
 * the layout of alloc_emb_rk_method (methods.h:31-42: a block of s row
 * pointers, and one block of (s+3)*s doubles holding A's rows, then b,
 * b_hat, c), so a solver written against methods.h frees it as usual, and
 * prints the reference's "Method:" line (src/methods.c:79-494), which
 * log2tab parses.
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "rkx_dropin.h"

static void load(const char *name, const char *label, double ***A, double **b, double **b_hat,
                 double **c, int *s, int *ord)
{
  rkx_method m;
  if (rkx_method_get(name, &m) != RKX_OK) {
    fprintf(stderr, "rkx: %s\n", rkx_last_error());
    exit(1);
  }
  const int n = m.s;
  double **rows = malloc(sizeof(double *) * (size_t) n);
  double *blk = malloc(sizeof(double) * (size_t) ((n + 3) * n));
  if (!rows || !blk)
    abort();
  for (int i = 0; i < n; i++) {
    rows[i] = blk + (size_t) i * n;
    for (int j = 0; j < n; j++)
      rows[i][j] = m.A[i][j];
  }
  *A = rows;
  *b = blk + (size_t) n * n;
  *b_hat = *b + n;
  *c = *b_hat + n;
  memcpy(*b, m.b, sizeof(double) * (size_t) n);
  memcpy(*b_hat, m.b_hat, sizeof(double) * (size_t) n);
  memcpy(*c, m.c, sizeof(double) * (size_t) n);
  *s = n;
  *ord = m.ord;
  printf("Method: %s (order %d, %d stages)\n", label, m.ord, m.s);
}

void rkx_dropin_free_method(double ***A, double **b, double **b_hat, double **c)
{
  free((*A)[0]);
  free(*A);
  *A = NULL;
  *b = *b_hat = *c = NULL;
}

#define RKX_METHOD_FN(NAME, LABEL)                                                  \
  void NAME(double ***A, double **b, double **b_hat, double **c, int *s, int *ord) \
  {                                                                                 \
    load(#NAME, LABEL, A, b, b_hat, c, s, ord);                                     \
  }

RKX_METHOD_FN(HEUN2, "HEUN2")
RKX_METHOD_FN(RK4, "RK4")
RKX_METHOD_FN(SSPRK3, "SSPRK3")
RKX_METHOD_FN(RKF23, "RKF 2(3)")
RKX_METHOD_FN(DOPRI54, "DOPRI 5(4)")
RKX_METHOD_FN(VERNER65, "VERNER65")
RKX_METHOD_FN(DOPRI87, "DOPRI 8(7)")
