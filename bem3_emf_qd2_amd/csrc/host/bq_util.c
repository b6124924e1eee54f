/*
 * bq_util.c -- small vector helpers and the checked allocator that the reference's
 * callers use (same names and behaviour as com_src/my_utils.c:15-78).
 */
#include <stdio.h>
#include <stdlib.h>
#include <math.h>
#include "bem3_emf_qd2.h"
#include "../bq_solver.h"
#include "../bq_assemble.h"

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

/* Gauss-Legendre rule of order n on [-1, 1]: roots of P_n by Newton iteration from the
 * Chebyshev-like start cos(pi (i + 3/4) / (n + 1/2)), weights 2 / ((1 - x^2) P_n'(x)^2). */
void bq_gauss_legendre(int n, double *x, double *w)
{
  for (int i = 0; i < (n + 1) / 2; i++) {
    double z = cos(M_PI * (i + 0.75) / (n + 0.5)), dp = 1.0;
    for (int it = 0; it < 100; it++) {
      double p0 = 1.0, p1 = z;
      for (int j = 2; j <= n; j++) {
        double p2 = ((2.0 * j - 1.0) * z * p1 - (j - 1.0) * p0) / j;
        p0 = p1;
        p1 = p2;
      }
      dp = n * (z * p1 - p0) / (z * z - 1.0);
      double dz = p1 / dp;
      z -= dz;
      if (fabs(dz) < 1e-17) break;
    }
    x[i] = -z;
    x[n - 1 - i] = z;
    w[i] = w[n - 1 - i] = 2.0 / ((1.0 - z * z) * dp * dp);
  }
}
