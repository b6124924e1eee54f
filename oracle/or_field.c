/*
 * or_field.c -- domain identification and field evaluation at arbitrary points.
 *
 * TEST INFRASTRUCTURE ONLY (see oracle.h).
 *
 * Follows d3qd2_src/d3qd2_setup.c (domain_id :157-202, q0_domain_id_l :238-245,
 * fid_calc_solid_angle :1351-1422), d3qd2_src/d3qd2_field.c (mEMP_s/t :10-73, EH_mEMP_s/t :90-190,
 * EH_s :203-269,
 * EH_t :271-285, EH_s_bd/EH_t_bd :299-391) and d3qd2_src/d3qd2_force.c (force_FN).
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

/* modified electromagnetic potential U = (A, phi) at r (mEMP_s / mEMP_t,
 * d3qd2_src/d3qd2_field.c:10-73): the representation sum of EH_s over the stored U, dU
 * ([Ne+1][4][4] per domain); total adds the incident E to A in domain 0 */
int or_mEMP(const or_model *m, double complex **U, double complex **dU, const double *r, int type, int total,
            double complex *Uo)
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
  for (int i = 0; i < 4; i++) Uo[i] = 0.0;
  if (did < 0) return did;
  double F = 0;
  for (int s = 1; s <= m->dNe[did]; s++) {
    double complex CC[9];
    or_coef_rt(m, did, r0, m->sid[did][s], type, CC);
    for (int n = 0; n < 4; n++)
      for (int i = 0; i < 4; i++) Uo[i] += CC[n] * dU[did][16 * s + 4 * n + i] - CC[n + 4] * U[did][16 * s + 4 * n + i];
    F += creal(CC[8]);
  }
  double C = (did == 0) ? 1.0 + F : F;
  for (int i = 0; i < 4; i++) Uo[i] /= C;
  if (l1 != 0 || l2 != 0) {
    const or_qpd2 *q = &m->qd;
    double arg = q->vk[0] * (l1 * q->vd1[0] + l2 * q->vd2[0]) + q->vk[1] * (l1 * q->vd1[1] + l2 * q->vd2[1]);
    double complex ce = cos(arg) + I * sin(arg);
    for (int i = 0; i < 4; i++) Uo[i] *= ce;
  }
  if (total && did == 0) {
    double complex e[3], h[3];
    or_ipw_EH(m, r, e, h);
    for (int i = 0; i < 3; i++) Uo[i] += e[i];
  }
  return did;
}

/* E, H from the modified potential and its gradient (EH_mEMP_s / EH_mEMP_t,
 * d3qd2_src/d3qd2_field.c:90-190): E = A - grad phi, H = lambda0 / (2 pi i) curl A.  A point
 * where a solid-angle derivative exceeds CBD_CDF = 10 (too close to the boundary) gives zero
 * fields and returns -did (-99 for the open region). */
int or_EH_mEMP(const or_model *m, double complex **U, double complex **dU, const double *r, int type, int total,
               double complex *Eo, double complex *Ho)
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
  double complex Us[4] = {0, 0, 0, 0}, dUs[3][4] = {{0}};
  double F = 0, dF[3] = {0, 0, 0};
  for (int s = 1; s <= m->dNe[did]; s++) {
    double complex CC[9], dC[3][9];
    or_coef_rt_grad(m, did, r0, m->sid[did][s], type, CC, dC);
    for (int n = 0; n < 4; n++)
      for (int c = 0; c < 4; c++) {
        double complex u = U[did][16 * s + 4 * n + c], du = dU[did][16 * s + 4 * n + c];
        Us[c] += CC[n] * du - CC[n + 4] * u;
        for (int d = 0; d < 3; d++) dUs[d][c] += dC[d][n] * du - dC[d][n + 4] * u;
      }
    F += creal(CC[8]);
    for (int d = 0; d < 3; d++) dF[d] += creal(dC[d][8]);
  }
  if (fabs(dF[0]) > 10.0 || fabs(dF[1]) > 10.0 || fabs(dF[2]) > 10.0) return did == 0 ? -99 : -did;
  double iC = (did == 0) ? 1.0 / (1.0 + F) : 1.0 / F;
  for (int c = 0; c < 4; c++) Us[c] *= iC;
  for (int d = 0; d < 3; d++)
    for (int c = 0; c < 4; c++) dUs[d][c] = (dUs[d][c] - Us[c] * dF[d]) * iC;
  double complex ch = m->lambda0 / (2.0 * M_PI * I);
  for (int d = 0; d < 3; d++) Eo[d] = -dUs[d][3] + Us[d];
  Ho[0] = ch * (dUs[1][2] - dUs[2][1]);
  Ho[1] = ch * (dUs[2][0] - dUs[0][2]);
  Ho[2] = ch * (dUs[0][1] - dUs[1][0]);
  if (l1 != 0 || l2 != 0) {
    const or_qpd2 *q = &m->qd;
    double arg = q->vk[0] * (l1 * q->vd1[0] + l2 * q->vd2[0]) + q->vk[1] * (l1 * q->vd1[1] + l2 * q->vd2[1]);
    double complex ce = cos(arg) + I * sin(arg);
    for (int i = 0; i < 3; i++) { Eo[i] *= ce; Ho[i] *= ce; }
  }
  if (total && did == 0) {
    double complex e[3], h[3];
    or_ipw_EH(m, r, e, h);
    for (int i = 0; i < 3; i++) { Eo[i] += e[i]; Ho[i] += h[i]; }
  }
  return did;
}

void or_EH_bd(const or_model *m, double complex **E, double complex **dE, double complex **H,
              double complex **dH, int did, int t, double zt, double et, int type, int total,
              double complex *Eo, double complex *Ho)
{
  int td = m->sid[did][t];
  double F = 0.0;
  for (int i = 0; i < 3; i++) Eo[i] = Ho[i] = 0.0;
  for (int s = 1; s <= m->dNe[did]; s++) {
    double complex CC[9];
    or_coef_bd(m, did, td, zt, et, m->sid[did][s], type, CC);
    for (int n = 0; n < 4; n++)
      for (int i = 0; i < 3; i++) {
        Eo[i] += CC[n] * dE[did][12 * s + 3 * n + i] - CC[n + 4] * E[did][12 * s + 3 * n + i];
        Ho[i] += CC[n] * dH[did][12 * s + 3 * n + i] - CC[n + 4] * H[did][12 * s + 3 * n + i];
      }
    F += creal(CC[8]);
  }
  double C = (did == 0) ? 1.0 + F : F;
  for (int i = 0; i < 3; i++) { Eo[i] /= C; Ho[i] /= C; }
  if (total && did == 0) {
    double r[3];
    double complex e[3], h[3];
    or_r_bd(m, td, zt, et, r);
    or_ipw_EH(m, r, e, h);
    for (int i = 0; i < 3; i++) { Eo[i] += e[i]; Ho[i] += h[i]; }
  }
}

/* Maxwell stress T = 1/2 Re(eps E E* + H H*) - 1/4 (eps |E|^2 + |H|^2) I contracted with the
 * surface element w and its moment about rc, eps = Re(n0)^2 (d3qd2_force.c:60-98) */
static void stress_add(const or_model *m, const double complex *e, const double complex *h, const double *r,
                       const double *w, const double *rc, double wt, double *F, double *N)
{
  double epd = creal(m->n[0]) * creal(m->n[0]);
  double aE2 = 0, aH2 = 0, T[9];
  for (int j = 0; j < 3; j++) { aE2 += creal(e[j] * conj(e[j])); aH2 += creal(h[j] * conj(h[j])); }
  double Tc = 0.25 * (epd * aE2 + aH2);
  for (int j = 0; j < 3; j++)
    for (int k = 0; k < 3; k++) T[3 * j + k] = 0.5 * (epd * creal(e[j] * conj(e[k])) + creal(h[j] * conj(h[k])));
  T[0] -= Tc; T[4] -= Tc; T[8] -= Tc;
  double d[3] = {r[0] - rc[0], r[1] - rc[1], r[2] - rc[2]};
  for (int j = 0; j < 3; j++) {
    F[j] += (T[3 * j] * w[0] + T[3 * j + 1] * w[1] + T[3 * j + 2] * w[2]) * wt;
    N[j] += ((d[1] * T[3 * j + 2] - d[2] * T[3 * j + 1]) * w[0] + (d[2] * T[3 * j] - d[0] * T[3 * j + 2]) * w[1] +
             (d[0] * T[3 * j + 1] - d[1] * T[3 * j]) * w[2]) * wt;
  }
}

int or_force_FN(const or_model *m, double complex **E, double complex **dE, double complex **H,
                double complex **dH, const double *rc, int type, double *F, double *N)
{
  for (int j = 0; j < 3; j++) F[j] = N[j] = 0.0;
  if (m->dNe[0] == 0) return -1;
  int ne = m->dNe[0];
  double *part = calloc((size_t)6 * (ne + 1), sizeof(double));
#pragma omp parallel for schedule(dynamic)
  for (int t = 1; t <= ne; t++) {
    int at = abs(m->sid[0][t]), tri = (m->ed[4 * at + 3] == 0);
    double f[3] = {0, 0, 0}, nn[3] = {0, 0, 0};
    int np = type == 0 ? 4 : (tri ? 7 : 9);
    for (int i = 0; i < np; i++) {
      double r[3], w[3], wt, zt, et;
      double complex e[3], h[3];
      const double *cw = m->cw + 9 * at;
      if (type == 0 && !(tri && i == 3)) {
        /* stored node fields plus the incident wave */
        for (int j = 0; j < 3; j++) {
          r[j] = m->ren[12 * at + 3 * i + j];
          w[j] = m->wen[12 * at + 3 * i + j];
          e[j] = E[0][12 * t + 3 * i + j];
          h[j] = H[0][12 * t + 3 * i + j];
        }
        double complex te[3], th[3];
        or_ipw_EH(m, r, te, th);
        for (int j = 0; j < 3; j++) { e[j] += te[j]; h[j] += th[j]; }
        wt = tri ? m->wt34[i] : m->wt44[i];
      } else {
        zt = type == 0 ? m->zt34[i] : (tri ? m->zt37[i] : m->zt49[i]);
        et = type == 0 ? m->et34[i] : (tri ? m->et37[i] : m->et49[i]);
        wt = type == 0 ? m->wt34[i] : (tri ? m->wt37[i] : m->wt49[i]);
        or_r_bd(m, at, zt, et, r);
        for (int j = 0; j < 3; j++) w[j] = cw[3 * j] + cw[3 * j + 1] * zt + cw[3 * j + 2] * et;
        /* PREC_DEF_FN = 0: node rule for the other elements (bem3_emf_qd2.h:14) */
        or_EH_bd(m, E, dE, H, dH, 0, t, zt, et, 0, 1, e, h);
      }
      stress_add(m, e, h, r, w, rc, wt, f, nn);
    }
    double sc = tri ? -0.5 : -1.0;
    for (int j = 0; j < 3; j++) { part[6 * t + j] = sc * f[j]; part[6 * t + 3 + j] = sc * nn[j]; }
  }
  for (int t = 1; t <= ne; t++)
    for (int j = 0; j < 3; j++) { F[j] += part[6 * t + j]; N[j] += part[6 * t + 3 + j]; }
  free(part);
  return 0;
}
