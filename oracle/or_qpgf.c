/*
 * or_qpgf.c -- quasi-periodic Green's function of the Helmholtz equation with 2-D
 * periodicity (value, gradient, Hessian), restated for the oracle.
 *
 * TEST INFRASTRUCTURE ONLY (see oracle.h).
 *
 * Follows qgf2_src/d3_qpgf_d2.c:
 *   method choice (Ewald vs Fourier-domain sum)         :124-176
 *   fold into the unit cell + Bloch phase                :290-319, 409-438
 *   Fourier-domain sum and its terms                      :745-885, 887-1044
 *   Ewald split: Ewald parameter rule                     :1671-1712
 *   Ewald spatial sum and its terms                       :1714-1838, 2052-2200
 *   Ewald spectral sum and its terms                      :1840-2009, 2201-2399
 *   adaptive stopping tests (EFS, comp/diff)              :471-482
 * The reference variants (_v1, _v2, _qG, _dqG) share one implementation here, each with
 * its own stopping tests (sp_* / sn_* below).  Results differ from the reference only in
 * terms below eps*EFS of the sum -- including the reference's early stop of the _v1/_dqG
 * spectral sum when r_z == 0 exactly, whose stop test then compares zero with zero.
 */
#include <dlfcn.h>
#include <math.h>
#include <pthread.h>
#include <stdio.h>
#include <stdlib.h>
#include <complex.h>
#include "oracle.h"

#define S_LIMIT 15
#define FD_LIMIT 100
#define EW_LIMIT 100
#define EW_EXPM 3.0
#define EFS 0.1

static int same_z(double complex a, double complex b, double eps)
{
  return 2.0 * fabs(creal(a) - creal(b)) <= eps * (fabs(creal(a)) + fabs(creal(b))) &&
         2.0 * fabs(cimag(a) - cimag(b)) <= eps * (fabs(cimag(a)) + fabs(cimag(b)));
}

static int small_z(double complex t, double complex s, double eps)
{
  return fabs(creal(t)) <= fabs(creal(s)) * EFS * eps && fabs(cimag(t)) <= fabs(cimag(s)) * EFS * eps;
}

typedef struct {
  double complex q, d[3], h[6];
} acc_t;

static void acc_add(acc_t *a, const acc_t *t)
{
  a->q += t->q;
  for (int i = 0; i < 3; i++) a->d[i] += t->d[i];
  for (int i = 0; i < 6; i++) a->h[i] += t->h[i];
}

/* ---- Fourier-domain sum ------------------------------------------------------------ */
static int fd_term(const or_qpd2 *qd, int n1, int n2, const double *g1, const double *g2,
                   const double *r, acc_t *t)
{
  double an0 = -qd->vk[0] + n1 * g1[0] + n2 * g2[0];
  double an1 = -qd->vk[1] + n1 * g1[1] + n2 * g2[1];
  double bn2 = qd->k * qd->k - an0 * an0 - an1 * an1;
  double complex dz, F;
  int kind;
  if (bn2 > 0.0) {
    double bn = sqrt(bn2), arg = an0 * r[0] + an1 * r[1] + bn * fabs(r[2]);
    dz = cos(arg) + I * sin(arg);
    F = dz / bn;
    kind = 0;
  } else if (bn2 < 0.0) {
    double bn = sqrt(-bn2), arg = an0 * r[0] + an1 * r[1];
    dz = (cos(arg) + I * sin(arg)) * exp(-bn * fabs(r[2]));
    F = -I * dz / bn;
    kind = 1;
  } else return -2;
  t->q = F;
  t->d[0] = F * an0;
  t->d[1] = F * an1;
  t->d[2] = dz;
  t->h[0] = an0 * t->d[0];
  t->h[1] = an1 * t->d[1];
  t->h[2] = bn2 * F;
  t->h[3] = an0 * t->d[1];
  t->h[4] = an1 * t->d[2];
  t->h[5] = an0 * t->d[2];
  return kind;
}

static int fd_sum(int v2, const double *r, double eps, const or_qpd2 *qd, acc_t *res)
{
  if (r[2] == 0.0) return -4;
  double det = qd->vd1[0] * qd->vd2[1] - qd->vd2[0] * qd->vd1[1];
  if (det == 0.0) return -3;
  double idd = 1.0 / det;
  double g1[2] = {2.0 * M_PI * idd * qd->vd2[1], -2.0 * M_PI * idd * qd->vd2[0]};
  double g2[2] = {-2.0 * M_PI * idd * qd->vd1[1], 2.0 * M_PI * idd * qd->vd1[0]};
  int nc = (int)round((qd->vk[0] * qd->vd1[0] + qd->vk[1] * qd->vd1[1]) / (2.0 * M_PI));
  acc_t S = {0}, T;
  int cnt = 0;
  for (int dir = 0; dir < 3; dir++) {
    int start = (dir == 0) ? nc : (dir == 1 ? nc + 1 : nc - 1);
    int step = (dir == 2) ? -1 : 1;
    int n1 = start;
    for (int it = 0;; it++, n1 += step) {
      if (dir > 0 && it >= FD_LIMIT - 1) return -1;
      double complex old = v2 ? S.h[2] : S.d[2];
      int n2c = (int)round((qd->vk[0] * g2[0] + qd->vk[1] * g2[1] - n1 * (g1[0] * g2[0] + g1[1] * g2[1])) /
                           (g2[0] * g2[0] + g2[1] * g2[1]));
      int k = fd_term(qd, n1, n2c, g1, g2, r, &T);
      if (k < 0) return k;
      acc_add(&S, &T);
      cnt++;
      for (int sd = 0; sd < 2; sd++) {
        int n2 = n2c + (sd == 0 ? 1 : -1), m;
        for (m = 1; m < FD_LIMIT; m++, n2 += (sd == 0 ? 1 : -1)) {
          k = fd_term(qd, n1, n2, g1, g2, r, &T);
          if (k < 0) return k;
          acc_add(&S, &T);
          cnt++;
          if (k > 0 && (v2 ? small_z(T.h[2], S.h[2], eps) : small_z(T.d[2], S.d[2], eps))) break;
        }
        if (m == FD_LIMIT) return -1;
      }
      if (dir == 0) break;
      if (same_z(v2 ? S.h[2] : S.d[2], old, eps)) break;
    }
  }
  double ai = fabs(idd), sg = r[2] / fabs(r[2]);
  res->q = S.q * 0.5 * I * ai;
  res->d[0] = S.d[0] * (-0.5 * ai);
  res->d[1] = S.d[1] * (-0.5 * ai);
  res->d[2] = S.d[2] * (-0.5 * ai * sg);
  for (int i = 0; i < 4; i++) res->h[i] = S.h[i] * (-0.5 * I * ai);
  res->h[4] = S.h[4] * (-0.5 * I * ai * sg);
  res->h[5] = S.h[5] * (-0.5 * I * ai * sg);
  return cnt;
}

/* ---- Ewald method ------------------------------------------------------------------ */
static void ew_spatial_term(const or_qpd2 *qd, int l1, int l2, const double *r, double eps,
                            double veps, acc_t *t)
{
  double R0 = l1 * qd->vd1[0] + l2 * qd->vd2[0], R1 = l1 * qd->vd1[1] + l2 * qd->vd2[1];
  double ra[3] = {r[0] + R0, r[1] + R1, r[2]};
  double rl2 = ra[0] * ra[0] + ra[1] * ra[1] + ra[2] * ra[2];
  double rl = sqrt(rl2), irl = 1.0 / rl;
  double ke = qd->k * veps, ive = 1.0 / veps, y = rl * 0.5 * ive;
  double arg = qd->vk[0] * R0 + qd->vk[1] * R1;
  double complex cef = (cos(arg) + I * sin(arg)) * exp(ke * ke - y * y);
  double complex w = or_w(ke + I * y, eps);
  double complex cf1 = cef * irl * irl * (irl * creal(w) - qd->k * cimag(w) + ive / sqrt(M_PI));
  double complex cf2 = cef * irl * irl * (qd->k * qd->k * irl * 2.0 * creal(w) - ive * ive * ive / sqrt(M_PI));
  double complex cf3 = -6.0 * irl * irl * cf1 + cf2;
  t->q = cef * irl * creal(w);
  t->d[0] = ra[0] * cf1;
  t->d[1] = ra[1] * cf1;
  t->d[2] = cf1;
  t->h[0] = 2.0 * cf1 + ra[0] * ra[0] * cf3;
  t->h[1] = 2.0 * cf1 + ra[1] * ra[1] * cf3;
  t->h[2] = 2.0 * cf1 + ra[2] * ra[2] * cf3;
  t->h[3] = ra[0] * ra[1] * cf3;
  t->h[4] = ra[1] * cf3;
  t->h[5] = ra[0] * cf3;
}

/* stop rules per reference variant (mode): 0 _v1 and 2 _qG stop the spatial sum on the
 * value, 3 _dqG on the x/y gradient, 1 _v2 on the xx/yy Hessian (d3_qpgf_d2.c:1082-1168,
 * :1377-1480, :1714-1812, :2052-2157) */
static int sp_small(int mode, const acc_t *T, const acc_t *S, double eps)
{
  if (mode == 1) return small_z(T->h[0], S->h[0], eps) && small_z(T->h[1], S->h[1], eps);
  if (mode == 3) return small_z(T->d[0], S->d[0], eps) && small_z(T->d[1], S->d[1], eps);
  return small_z(T->q, S->q, eps);
}
static int sp_same(int mode, const acc_t *S, const acc_t *O, double eps)
{
  if (mode == 1) return same_z(S->h[0], O->h[0], eps) && same_z(S->h[1], O->h[1], eps);
  if (mode == 3) return same_z(S->d[0], O->d[0], eps) && same_z(S->d[1], O->d[1], eps);
  return same_z(S->q, O->q, eps);
}
/* spectral sums: _v1 and _dqG stop on dqG_z (which is exactly zero at z = 0, so those
 * variants truncate there), _qG on the value, _v2 on the xx/yy Hessian
 * (d3_qpgf_d2.c:1190-1335, :1505-1670, :1840-2009, :2201-2399) */
static int sn_small(int mode, const acc_t *T, const acc_t *S, double eps)
{
  if (mode == 1) return small_z(T->h[0], S->h[0], eps) && small_z(T->h[1], S->h[1], eps);
  if (mode == 2) return small_z(T->q, S->q, eps);
  return small_z(T->d[2], S->d[2], eps);
}
static int sn_same(int mode, const acc_t *S, const acc_t *O, double eps)
{
  if (mode == 1) return same_z(S->h[0], O->h[0], eps) && same_z(S->h[1], O->h[1], eps);
  if (mode == 2) return same_z(S->q, O->q, eps);
  return same_z(S->d[2], O->d[2], eps);
}

static int ew_spatial(int mode, const double *r, double eps, double veps, const or_qpd2 *qd, acc_t *res)
{
  acc_t S = {0}, T;
  int cnt = 0;
  double d22 = qd->vd2[0] * qd->vd2[0] + qd->vd2[1] * qd->vd2[1];
  for (int dir = 0; dir < 3; dir++) {
    int l1 = (dir == 0) ? 0 : (dir == 1 ? 1 : -1), step = (dir == 2) ? -1 : 1;
    for (int it = 0;; it++, l1 += step) {
      if (dir > 0 && it >= EW_LIMIT - 1) return -1;
      acc_t old = S;
      int lc = -(int)round(((r[0] + l1 * qd->vd1[0]) * qd->vd2[0] + (r[1] + l1 * qd->vd1[1]) * qd->vd2[1]) / d22);
      ew_spatial_term(qd, l1, lc, r, eps, veps, &T);
      acc_add(&S, &T);
      cnt++;
      for (int sd = 0; sd < 2; sd++) {
        int l2 = lc + (sd == 0 ? 1 : -1), m;
        for (m = 1; m < EW_LIMIT; m++, l2 += (sd == 0 ? 1 : -1)) {
          ew_spatial_term(qd, l1, l2, r, eps, veps, &T);
          acc_add(&S, &T);
          cnt++;
          if (sp_small(mode, &T, &S, eps)) break;
        }
        if (m == EW_LIMIT) return -1;
      }
      if (dir == 0) break;
      if (sp_same(mode, &S, &old, eps)) break;
    }
  }
  res->q = S.q / (4.0 * M_PI);
  res->d[0] = S.d[0] / (-4.0 * M_PI);
  res->d[1] = S.d[1] / (-4.0 * M_PI);
  res->d[2] = S.d[2] * (-r[2] / (4.0 * M_PI));
  for (int i = 0; i < 4; i++) res->h[i] = S.h[i] / (-8.0 * M_PI);
  res->h[4] = S.h[4] * (-r[2] / (8.0 * M_PI));
  res->h[5] = S.h[5] * (-r[2] / (8.0 * M_PI));
  return cnt;
}

static int ew_spectral_term(const or_qpd2 *qd, int n1, int n2, const double *g1, const double *g2,
                            const double *r, double eps, double veps, acc_t *t)
{
  double an0 = -qd->vk[0] + n1 * g1[0] + n2 * g2[0];
  double an1 = -qd->vk[1] + n1 * g1[1] + n2 * g2[1];
  double bn2 = qd->k * qd->k - an0 * an0 - an1 * an1;
  double arg = an0 * r[0] + an1 * r[1];
  double complex ce = cos(arg) + I * sin(arg);
  double az = r[2] / (2.0 * veps);
  double ef = exp(bn2 * veps * veps - az * az);
  double complex F, G;
  int kind;
  double bn;
  if (bn2 > 0.0) {
    bn = sqrt(bn2);
    double complex ff = cos(bn * r[2]) + I * sin(bn * r[2]);
    double complex w = or_w(bn * veps + I * az, eps);
    F = 2.0 * (I * ef * cimag(w) + ff);
    G = 2.0 * (ef * creal(w) - ff);
    t->q = ce / bn * F;
    t->h[2] = ce * (I * bn * F + 2.0 / (sqrt(M_PI) * veps) * ef);
    kind = 0;
  } else if (bn2 < 0.0) {
    bn = sqrt(-bn2);
    double ebz = exp(bn * r[2]);
    double f1 = ebz * erfc(bn * veps + az), f2 = erfc(bn * veps - az) / ebz;
    F = f1 + f2;
    G = f1 - f2;
    t->q = -I * ce / bn * F;
    t->h[2] = ce * (-bn * F + 2.0 / (sqrt(M_PI) * veps) * ef);
    kind = 1;
  } else return -3;
  t->d[0] = t->q * an0;
  t->d[1] = t->q * an1;
  t->d[2] = ce * G;
  t->h[0] = t->d[0] * an0;
  t->h[1] = t->d[1] * an1;
  t->h[3] = t->d[0] * an1;
  t->h[4] = an1 * t->d[2];
  t->h[5] = an0 * t->d[2];
  return kind;
}

static int ew_spectral(int mode, const double *r, double eps, double veps, const or_qpd2 *qd, acc_t *res)
{
  double det = qd->vd1[0] * qd->vd2[1] - qd->vd2[0] * qd->vd1[1];
  if (det == 0.0) return -4;
  double idd = 1.0 / det;
  double g1[2] = {2.0 * M_PI * idd * qd->vd2[1], -2.0 * M_PI * idd * qd->vd2[0]};
  double g2[2] = {-2.0 * M_PI * idd * qd->vd1[1], 2.0 * M_PI * idd * qd->vd1[0]};
  int nc = (int)round((qd->vk[0] * qd->vd1[0] + qd->vk[1] * qd->vd1[1]) / (2.0 * M_PI));
  acc_t S = {0}, T;
  int cnt = 0;
  double g22 = g2[0] * g2[0] + g2[1] * g2[1];
  for (int dir = 0; dir < 3; dir++) {
    int n1 = (dir == 0) ? nc : (dir == 1 ? nc + 1 : nc - 1), step = (dir == 2) ? -1 : 1;
    for (int it = 0;; it++, n1 += step) {
      if (dir > 0 && it >= EW_LIMIT - 1) return -2;
      acc_t old = S;
      int n2c = (int)round(-((-qd->vk[0] + n1 * g1[0]) * g2[0] + (-qd->vk[1] + n1 * g1[1]) * g2[1]) / g22);
      int k = ew_spectral_term(qd, n1, n2c, g1, g2, r, eps, veps, &T);
      if (k < 0) return k;
      acc_add(&S, &T);
      cnt++;
      int nprop = 0;
      for (int sd = 0; sd < 2; sd++) {
        int n2 = n2c + (sd == 0 ? 1 : -1), m;
        for (m = 1; m < EW_LIMIT; m++, n2 += (sd == 0 ? 1 : -1)) {
          k = ew_spectral_term(qd, n1, n2, g1, g2, r, eps, veps, &T);
          if (k < 0) return k;
          acc_add(&S, &T);
          cnt++;
          if (k > 0) {
            if (sn_small(mode, &T, &S, eps)) break;
          } else nprop++;
        }
        if (m == EW_LIMIT) return -2;
      }
      if (dir == 0) break;
      if (nprop == 0 && sn_same(mode, &S, &old, eps)) break;
    }
  }
  double cd = 0.25 * fabs(idd);
  double complex ci = cd * I;
  res->q = S.q * ci;
  res->d[0] = S.d[0] * (-cd);
  res->d[1] = S.d[1] * (-cd);
  res->d[2] = S.d[2] * cd;
  res->h[0] = S.h[0] * (-ci);
  res->h[1] = S.h[1] * (-ci);
  res->h[2] = S.h[2] * (-cd);
  res->h[3] = S.h[3] * (-ci);
  res->h[4] = S.h[4] * ci;
  res->h[5] = S.h[5] * ci;
  return cnt;
}

static int ew_sum(int mode, const double *r, double eps, const or_qpd2 *qd, acc_t *res)
{
  double det = qd->vd1[0] * qd->vd2[1] - qd->vd2[0] * qd->vd1[1];
  if (det == 0.0) return -4;
  double adp = hypot(qd->vd1[0] + qd->vd2[0], qd->vd1[1] + qd->vd2[1]);
  double adm = hypot(qd->vd1[0] - qd->vd2[0], qd->vd1[1] - qd->vd2[1]);
  double veps = sqrt(fabs(det) / (4.0 * M_PI) * adp / adm);
  double r2 = r[0] * r[0] + r[1] * r[1] + r[2] * r[2];
  if (qd->k * qd->k * veps * veps - r2 / (4.0 * veps * veps) > EW_EXPM)
    veps = sqrt(EW_EXPM + sqrt(qd->k * qd->k * r2 + EW_EXPM * EW_EXPM)) / (sqrt(2.0) * qd->k);
  acc_t a, b;
  int c1 = ew_spatial(mode, r, eps, veps, qd, &a);
  if (c1 < 0) return c1;
  int c2 = ew_spectral(mode, r, eps, veps, qd, &b);
  if (c2 < 0) return c2;
  *res = a;
  acc_add(res, &b);
  return c1 + c2;
}

/* Ewald/Fourier choice of d3_qpgf_d2.c:124-176: Fourier only when no reciprocal order
 * within S_LIMIT of the nearest one is propagating and its decay over |z| reaches eps. */
static int use_fourier(const double *r, double eps, const or_qpd2 *qd)
{
  double k2 = qd->k * qd->k;
  double idd = 1.0 / (qd->vd1[0] * qd->vd2[1] - qd->vd2[0] * qd->vd1[1]);
  double cpd = 2.0 * M_PI * idd;
  double ad12 = qd->vd1[0] * qd->vd1[0] + qd->vd1[1] * qd->vd1[1];
  double ad22 = qd->vd2[0] * qd->vd2[0] + qd->vd2[1] * qd->vd2[1];
  int n1 = (int)round((qd->vk[0] * qd->vd1[0] + qd->vk[1] * qd->vd1[1]) / (2.0 * M_PI));
  int n2 = (int)round((qd->vk[0] * qd->vd2[0] + qd->vk[1] * qd->vd2[1]) / (2.0 * M_PI));
#define AX(a, b) (-qd->vk[0] + (double)(a) * cpd * qd->vd2[1] - (double)(b) * cpd * qd->vd1[1])
#define AY(a, b) (-qd->vk[1] - (double)(a) * cpd * qd->vd2[0] + (double)(b) * cpd * qd->vd1[0])
  double bN = sqrt(k2 - AX(n1, n2) * AX(n1, n2) - AY(n1, n2) * AY(n1, n2));
  int p1, p2, m1, m2;
  if (ad22 < ad12) { p1 = n1 + S_LIMIT; p2 = n2; m1 = n1 - S_LIMIT; m2 = n2; }
  else { p1 = n1; p2 = n2 + S_LIMIT; m1 = n1; m2 = n2 - S_LIMIT; }
  double bNp = k2 - AX(p1, p2) * AX(p1, p2) - AY(p1, p2) * AY(p1, p2);
  double bNm = k2 - AX(m1, m2) * AX(m1, m2) - AY(m1, m2) * AY(m1, m2);
#undef AX
#undef AY
  if (bNp > 0.0 || bNm > 0.0) return 0;
  bNp = sqrt(-bNp);
  bNm = sqrt(-bNm);
  return (bN / bNp * exp(-bNp * fabs(r[2])) < eps && bN / bNm * exp(-bNm * fabs(r[2])) < eps);
}

/* OR_QPGF_REF = path of oracle/_ref/libqpgf_ref.so (the reference's own qgf2_src compiled
 * here): value + gradient evaluations (mode 0) then call its d3hm_qpgf_d2_v1 instead of the
 * restatement, so that a complete oracle solve shows where the reference's own code lands on
 * this machine's arithmetic.  Test instrument only. */
typedef int (*ref_v1_fn)(double complex *qG, double complex *dqG, double *r, double eps, void *qd);
static ref_v1_fn g_ref_v1;
static pthread_once_t g_ref_once = PTHREAD_ONCE_INIT;
static void ref_load(void)
{
  const char *p = getenv("OR_QPGF_REF");
  if (!p || !*p) return;
  void *h = dlopen(p, RTLD_NOW);
  if (!h) {
    fprintf(stderr, "oracle: OR_QPGF_REF %s: %s\n", p, dlerror());
    exit(1);
  }
  g_ref_v1 = (ref_v1_fn)dlsym(h, "d3hm_qpgf_d2_v1");
  if (!g_ref_v1) {
    fprintf(stderr, "oracle: d3hm_qpgf_d2_v1 not in %s\n", p);
    exit(1);
  }
}

int or_qpgf(int mode, const double *r, double eps, const or_qpd2 *qd, double complex *out)
{
  pthread_once(&g_ref_once, ref_load);
  if (mode == 0 && g_ref_v1) {
    double rr[3] = {r[0], r[1], r[2]};
    for (int i = 0; i < 10; i++) out[i] = 0.0;
    return g_ref_v1(&out[0], &out[1], rr, eps, (void *)qd);
  }
  int v2 = (mode == 1);
  double idd = 1.0 / (qd->vd1[0] * qd->vd2[1] - qd->vd2[0] * qd->vd1[1]);
  int l1 = (int)floor((r[0] * qd->vd2[1] - r[1] * qd->vd2[0]) * idd + 0.5);
  int l2 = (int)floor(-(r[0] * qd->vd1[1] - r[1] * qd->vd1[0]) * idd + 0.5);
  double rt[3] = {r[0] - l1 * qd->vd1[0] - l2 * qd->vd2[0], r[1] - l1 * qd->vd1[1] - l2 * qd->vd2[1], r[2]};
  acc_t res;
  int fd = use_fourier(r, eps, qd), err;
  if (fd) {
    err = fd_sum(v2, rt, eps, qd, &res);
    if (err < 0) return err - 10;
  } else {
    err = ew_sum(mode, rt, eps, qd, &res);
    if (err < 0) return err - 20;
  }
  if (l1 != 0 || l2 != 0) {
    double arg = -(qd->vk[0] * (l1 * qd->vd1[0] + l2 * qd->vd2[0]) + qd->vk[1] * (l1 * qd->vd1[1] + l2 * qd->vd2[1]));
    double complex cf = cos(arg) + I * sin(arg);
    res.q *= cf;
    for (int i = 0; i < 3; i++) res.d[i] *= cf;
    for (int i = 0; i < 6; i++) res.h[i] *= cf;
  }
  out[0] = res.q;
  for (int i = 0; i < 3; i++) out[1 + i] = res.d[i];
  for (int i = 0; i < 6; i++) out[4 + i] = res.h[i];
  return err;
}

int or_qpgf_batch(int mode, int n, const double *r, double eps, const or_qpd2 *qd, double *out)
{
  int worst = 1;
#pragma omp parallel for schedule(dynamic, 16) reduction(min : worst)
  for (int i = 0; i < n; i++) {
    double complex o[10];
    int e = or_qpgf(mode, r + 3 * i, eps, qd, o);
    if (e < worst) worst = e;
    for (int j = 0; j < 10; j++) {
      out[20 * i + 2 * j] = creal(o[j]);
      out[20 * i + 2 * j + 1] = cimag(o[j]);
    }
  }
  return worst;
}
