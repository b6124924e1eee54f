/*
 * or_field.c -- domain identification and field evaluation at arbitrary points.
 *
 * TEST INFRASTRUCTURE ONLY (see oracle.h).
 *
 * Follows d3qd2_src/d3qd2_setup.c (domain_id :157-202, q0_domain_id_l :238-245,
 * fid_calc_solid_angle :1351-1422) and d3qd2_src/d3qd2_field.c (EH_s :203-269,
 * EH_t :271-285).
 */
#include <math.h>
#include <complex.h>
#include <stdlib.h>
#include "oracle.h"

static void cross(double *c, const double *a, const double *b)
{
  c[0] = a[1] * b[2] - a[2] * b[1];
  c[1] = a[2] * b[0] - a[0] * b[2];
  c[2] = a[0] * b[1] - a[1] * b[0];
}
static double dot(const double *a, const double *b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }
static double nrm(const double *a) { return sqrt(dot(a, a)); }
static int unit(double *a)
{
  double n = nrm(a);
  if (n == 0.0) return -1;
  for (int i = 0; i < 3; i++) a[i] /= n;
  return 0;
}

/* signed solid angle of the spherical triangle(s) spanned by unit vectors to the corners */
static double tri_angle(const double *r0, const double *r1, const double *r2, int *flg, double *n20out)
{
  double n01[3], n12[3], n20[3], nt[3];
  cross(n01, r0, r1);
  cross(n12, r1, r2);
  cross(n20, r2, r0);
  cross(nt, n01, n20);
  double a0 = atan2(nrm(nt), -dot(n01, n20));
  cross(nt, n12, n01);
  double a1 = atan2(nrm(nt), -dot(n12, n01));
  cross(nt, n20, n12);
  double a2 = atan2(nrm(nt), -dot(n20, n12));
  double D = dot(r0, n12);
  if (n20out) for (int i = 0; i < 3; i++) n20out[i] = n20[i];
  if (D > 0.0) return a0 + a1 + a2 - M_PI;
  if (D < 0.0) return -(a0 + a1 + a2 - M_PI);
  *flg = -1;
  return 0.0;
}

static double solid_angle(int quad, double r[4][3], int *flg)
{
  *flg = 0;
  int st = 0;
  for (int i = 0; i < (quad ? 4 : 3); i++) st += unit(r[i]);
  if (st < 0) { *flg = -1; return 0.0; }
  double n20[3];
  double om = tri_angle(r[0], r[1], r[2], flg, n20);
  if (*flg < 0) return 0.0;
  if (quad) {
    /* second triangle (0, 2, 3) reusing n01 = -n20 of the first (setup.c:1390-1420) */
    double n01[3] = {-n20[0], -n20[1], -n20[2]}, n12[3], nb[3], nt[3];
    cross(n12, r[2], r[3]);
    cross(nb, r[3], r[0]);
    cross(nt, n01, nb);
    double a0 = atan2(nrm(nt), -dot(n01, nb));
    cross(nt, n12, n01);
    double a1 = atan2(nrm(nt), -dot(n12, n01));
    cross(nt, nb, n12);
    double a2 = atan2(nrm(nt), -dot(nb, n12));
    double D = dot(r[0], n12);
    if (D > 0.0) om += a0 + a1 + a2 - M_PI;
    else if (D < 0.0) om -= a0 + a1 + a2 - M_PI;
    else { *flg = -1; return 0.0; }
  }
  return om;
}

int or_domain_id(const or_model *m, const double *rt)
{
  double *Og = calloc(m->MN + 1, sizeof(double));
  int res = -1;
  for (int d = 0; d <= m->MN && res < 0; d++) {
    for (int i = 1; i <= m->dNe[d]; i++) {
      int s = m->sid[d][i];
      if (s <= 0) continue;
      double rv[4][3];
      for (int j = 0; j < 4; j++)
        for (int k = 0; k < 3; k++) rv[j][k] = m->rn[3 * m->ed[4 * s + j] + k] - rt[k];
      int flg;
      double om = solid_angle(m->ed[4 * s + 3] != 0, rv, &flg);
      if (flg < 0) { res = d; break; }
      Og[d] += om;
      Og[m->sd[s]] -= om;
    }
    if (res >= 0) break;
    if (d == 0 && fabs(Og[d]) < 2.0 * M_PI) res = d;
    else if (Og[d] > 2.0 * M_PI) res = d;
  }
  free(Og);
  return res;
}

int or_EH_t(const or_model *m, double complex **E, double complex **dE, double complex **H,
            double complex **dH, const double *r, int type, double complex *Eo, double complex *Ho)
{
  int l1 = 0, l2 = 0;
  double r0[3] = {r[0], r[1], r[2]};
  if (m->ps == 1) {
    const or_qpd2 *q = &m->qd;
    double idd = 1.0 / (q->vd1[0] * q->vd2[1] - q->vd2[0] * q->vd1[1]);
    l1 = (int)floor((r[0] * q->vd2[1] - r[1] * q->vd2[0]) * idd + 0.5);
    l2 = (int)floor(-(r[0] * q->vd1[1] - r[1] * q->vd1[0]) * idd + 0.5);
    r0[0] = r[0] - l1 * q->vd1[0] - l2 * q->vd2[0];
    r0[1] = r[1] - l1 * q->vd1[1] - l2 * q->vd2[1];
  }
  int did = or_domain_id(m, r0);
  for (int i = 0; i < 3; i++) Eo[i] = Ho[i] = 0.0;
  if (did < 0) return did;
  double F = 0;
  for (int s = 1; s <= m->dNe[did]; s++) {
    double complex CC[9];
    or_coef_rt(m, did, r0, m->sid[did][s], type, CC);
    for (int n = 0; n < 4; n++)
      for (int i = 0; i < 3; i++) {
        Eo[i] += CC[n] * dE[did][12 * s + 3 * n + i] - CC[n + 4] * E[did][12 * s + 3 * n + i];
        Ho[i] += CC[n] * dH[did][12 * s + 3 * n + i] - CC[n + 4] * H[did][12 * s + 3 * n + i];
      }
    F += creal(CC[8]);
  }
  double C = (did == 0) ? 1.0 + F : F;
  for (int i = 0; i < 3; i++) { Eo[i] /= C; Ho[i] /= C; }
  if (l1 != 0 || l2 != 0) {
    const or_qpd2 *q = &m->qd;
    double arg = q->vk[0] * (l1 * q->vd1[0] + l2 * q->vd2[0]) + q->vk[1] * (l1 * q->vd1[1] + l2 * q->vd2[1]);
    double complex ce = cos(arg) + I * sin(arg);
    for (int i = 0; i < 3; i++) { Eo[i] *= ce; Ho[i] *= ce; }
  }
  if (did == 0) {
    double complex e[3], h[3];
    or_ipw_EH(m, r, e, h);
    for (int i = 0; i < 3; i++) { Eo[i] += e[i]; Ho[i] += h[i]; }
  }
  return did;
}
