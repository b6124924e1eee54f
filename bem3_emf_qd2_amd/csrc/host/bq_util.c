/*
 * bq_util.c -- small vector helpers and the checked allocator that the reference's
 * callers use (same names and behaviour as com_src/my_utils.c:15-78).
 */
#include <stdio.h>
#include <stdlib.h>
#include <math.h>
#include "bem3_emf_qd2.h"
#include "../bq_solver.h"

void vadd_d(double *r, double *v0, double *v1) { for (int i = 0; i < 3; i++) r[i] = v0[i] + v1[i]; }
void vsub_d(double *r, double *v0, double *v1) { for (int i = 0; i < 3; i++) r[i] = v0[i] - v1[i]; }
double vabs_d(double *v) { return sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]); }
double vdot_d(double *v0, double *v1) { return v0[0] * v1[0] + v0[1] * v1[1] + v0[2] * v1[2]; }

double vabs_2dm(double *v0, double *v1)
{
  double d[3] = {v0[0] - v1[0], v0[1] - v1[1], v0[2] - v1[2]};
  return vabs_d(d);
}

int vuni_d(double *r)
{
  double a = vabs_d(r);
  if (a == 0.0) return -1;
  r[0] /= a;
  r[1] /= a;
  r[2] /= a;
  return 0;
}

void vcrs_d(double *r, double *v0, double *v1)
{
  double t0 = v0[1] * v1[2] - v0[2] * v1[1];
  double t1 = v0[2] * v1[0] - v0[0] * v1[2];
  double t2 = v0[0] * v1[1] - v0[1] * v1[0];
  r[0] = t0;
  r[1] = t1;
  r[2] = t2;
}

/* zero-initialised allocation that exits with the caller's message on failure */
void *m_alloc2(size_t num, size_t size, char *txt)
{
  void *p = calloc(num, size);
  if (p == NULL) {
    printf("memory allocation error!\n%s. Exit..\n", txt ? txt : "");
    exit(1);
  }
  return p;
}

int bq_sizeof_dqd2(void) { return (int)sizeof(DQD2); }

void bq_pack_elem(double *const *cr, double *const *cw, int tri, double *dst)
{
  const double js = tri ? 2.0 / (3.0 * sqrt(3.0)) : 1.0;
  for (int d = 0; d < 3; d++) {
    for (int c = 0; c < 4; c++) dst[4 * d + c] = cr[d][c];
    for (int c = 0; c < 3; c++) dst[12 + 3 * d + c] = cw[d][c] * js;
  }
  dst[21] = tri ? 1.0 : 0.0;
  dst[22] = dst[23] = 0.0;
}
