/*
 * bq_field.c -- field API of the reference (d3qd2_src/d3qd2_field.c) on the GPU.
 *
 * The solved model is uploaded once per DQD2 (the first field call after solve_bieq or
 * dat_read) and kept in HBM until finalize_dqd2.  The scalar entry points (EH_s, EH_t ...)
 * are safe to call from several host threads at once, as the reference's examples do from
 * OpenMP loops: calls are serialised on one lock and each evaluates its point on the GPU.
 * bq_EH_t_batch evaluates many points in one launch.
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <math.h>
#include <complex.h>
#include <pthread.h>
#include "bem3_emf_qd2.h"
#include "bq_amd.h"
#include "../bq_field.h"
#include "../bq_solver.h"

static pthread_mutex_t g_lock = PTHREAD_MUTEX_INITIALIZER;
static struct {
  DQD2 *qd;
  const void *key;  /* identity of the solution arrays the context was built from */
  void *ctx;
} g_cache;

static void field_fail(const char *what)
{
  printf("%s failed on the GPU: %s. Exit...\n", what, bq_last_error());
  exit(1);
}

static void flat_elem(BOUD *bd, double *E, double *ren, double *wen)
{
  for (int i = 1; i <= bd->Ne; i++) {
    bq_pack_elem(bd->cr[i], bd->cw[i], bd->ed[i][3] == 0, E + 24 * (size_t)i);
    for (int j = 0; j < 4; j++)
      for (int d = 0; d < 3; d++) {
        ren[12 * i + 3 * j + d] = bd->ren[i][j][d];
        wen[12 * i + 3 * j + d] = bd->wen[i][j][d];
      }
  }
}

static double *flat_z3(double complex ***a, int n1, int m)
{
  double *o = calloc((size_t)n1 * 4 * m * 2, sizeof(double));
  for (int i = 0; i < n1; i++)
    for (int j = 0; j < 4; j++)
      for (int c = 0; c < m; c++) {
        o[2 * ((size_t)i * 4 * m + j * m + c)] = creal(a[i][j][c]);
        o[2 * ((size_t)i * 4 * m + j * m + c) + 1] = cimag(a[i][j][c]);
      }
  return o;
}

/* caller holds g_lock */
static void *context_for(DQD2 *qd)
{
  const void *key = qd->bd.sb ? (const void *)qd->bd.sb[0].E : NULL;
  if (g_cache.ctx && g_cache.qd == qd && g_cache.key == key) return g_cache.ctx;
  if (g_cache.ctx) {
    bq_field_ctx_destroy(g_cache.ctx);
    g_cache.ctx = NULL;
  }
  BOUD *bd = &qd->bd;
  int n1 = bd->Ne + 1, MN = qd->MN;
  double *E = calloc((size_t)n1 * 24, sizeof(double)), *ren = calloc((size_t)n1 * 12, sizeof(double));
  double *wen = calloc((size_t)n1 * 12, sizeof(double)), *kn = calloc((size_t)(MN + 1) * 2, sizeof(double));
  flat_elem(bd, E, ren, wen);
  int *dNe = calloc(MN + 1, sizeof(int));
  const int **sid = calloc(MN + 1, sizeof(int *));
  double **Eb = calloc(MN + 1, sizeof(double *)), **dEb = calloc(MN + 1, sizeof(double *));
  double **Hb = calloc(MN + 1, sizeof(double *)), **dHb = calloc(MN + 1, sizeof(double *));
  for (int d = 0; d <= MN; d++) {
    kn[2 * d] = creal(qd->kn[d]);
    kn[2 * d + 1] = cimag(qd->kn[d]);
    SUBD *s = &bd->sb[d];
    dNe[d] = s->Ne;
    sid[d] = s->sid;
    Eb[d] = flat_z3(s->E, s->Ne + 1, 3);
    dEb[d] = flat_z3(s->dE, s->Ne + 1, 3);
    Hb[d] = flat_z3(s->H, s->Ne + 1, 3);
    dHb[d] = flat_z3(s->dH, s->Ne + 1, 3);
  }
  bq_fdata f;
  memset(&f, 0, sizeof f);
  f.Ne = bd->Ne;
  f.MN = MN;
  f.ps = bd->ps;
  f.qd[0] = bd->qd.k;
  for (int i = 0; i < 3; i++) f.qd[1 + i] = bd->qd.vk[i];
  f.qd[4] = bd->qd.vd1[0];
  f.qd[5] = bd->qd.vd1[1];
  f.qd[6] = bd->qd.vd2[0];
  f.qd[7] = bd->qd.vd2[1];
  f.kn = kn;
  f.E = E;
  f.ren = ren;
  f.wen = wen;
  f.dNe = dNe;
  f.sid = (const int *const *)sid;
  f.Eb = (const double *const *)Eb;
  f.dEb = (const double *const *)dEb;
  f.Hb = (const double *const *)Hb;
  f.dHb = (const double *const *)dHb;
  f.xli = bd->xli; f.wli = bd->wli; f.xhi = bd->xhi; f.whi = bd->whi;
  f.z49 = bd->zt_49; f.e49 = bd->et_49; f.w49 = bd->wt_49;
  f.w34 = bd->wt_34; f.z37 = bd->zt_37; f.e37 = bd->et_37; f.w37 = bd->wt_37;
  void *ctx = NULL;
  int rc = bq_field_ctx_create(&f, &ctx);
  for (int d = 0; d <= MN; d++) {
    free(Eb[d]); free(dEb[d]); free(Hb[d]); free(dHb[d]);
  }
  free(E); free(ren); free(wen); free(kn); free(dNe); free(sid); free(Eb); free(dEb); free(Hb); free(dHb);
  if (rc) field_fail("field set-up");
  g_cache.qd = qd;
  g_cache.key = key;
  g_cache.ctx = ctx;
  return ctx;
}

void bq_release_device_state(DQD2 *qd)
{
  pthread_mutex_lock(&g_lock);
  if (g_cache.ctx && g_cache.qd == qd) {
    bq_field_ctx_destroy(g_cache.ctx);
    g_cache.ctx = NULL;
    g_cache.qd = NULL;
    g_cache.key = NULL;
  }
  pthread_mutex_unlock(&g_lock);
}

static void fold(DQD2 *qd, const double *rt, double *r0, int *l1, int *l2)
{
  if (qd->bd.ps == 1) {
    q0_domain_id_l(l1, l2, (double *)rt, qd);
    r0[0] = rt[0] - *l1 * qd->bd.qd.vd1[0] - *l2 * qd->bd.qd.vd2[0];
    r0[1] = rt[1] - *l1 * qd->bd.qd.vd1[1] - *l2 * qd->bd.qd.vd2[1];
  } else {
    *l1 = *l2 = 0;
    r0[0] = rt[0];
    r0[1] = rt[1];
  }
  r0[2] = rt[2];
}

/* Scattered (or interior) E, H at n points; did[] receives the domain ids. */
static void eh_scattered(DQD2 *qd, int n, const double *r, int type, double complex *E, double complex *H, int *did)
{
  double *r0 = malloc(sizeof(double) * 3 * (n > 0 ? n : 1));
  int *l1 = malloc(sizeof(int) * (n > 0 ? n : 1)), *l2 = malloc(sizeof(int) * (n > 0 ? n : 1));
  int *dd = malloc(sizeof(int) * (n > 0 ? n : 1));
  int m = 0;
  for (int p = 0; p < n; p++) {
    fold(qd, r + 3 * p, r0 + 3 * m, &l1[p], &l2[p]);
    did[p] = domain_id(r0 + 3 * m, qd);
    if (did[p] >= 0) dd[m++] = did[p];
  }
  double *out = calloc((size_t)13 * (m > 0 ? m : 1), sizeof(double));
  if (m > 0) {
    pthread_mutex_lock(&g_lock);
    void *ctx = context_for(qd);
    int rc = bq_field_eval(ctx, m, r0, dd, type, out);
    pthread_mutex_unlock(&g_lock);
    if (rc) field_fail("field evaluation");
  }
  for (int p = 0, q = 0; p < n; p++) {
    for (int i = 0; i < 3; i++) E[3 * p + i] = H[3 * p + i] = 0.0;
    if (did[p] < 0) continue;
    const double *o = out + 13 * q++;
    double C = (did[p] == 0) ? 1.0 + o[12] : o[12];
    double complex ce = 1.0;
    if (l1[p] != 0 || l2[p] != 0) {
      double arg = qd->bd.qd.vk[0] * (l1[p] * qd->bd.qd.vd1[0] + l2[p] * qd->bd.qd.vd2[0]) +
                   qd->bd.qd.vk[1] * (l1[p] * qd->bd.qd.vd1[1] + l2[p] * qd->bd.qd.vd2[1]);
      ce = cos(arg) + I * sin(arg);
    }
    for (int i = 0; i < 3; i++) {
      E[3 * p + i] = (o[2 * i] + I * o[2 * i + 1]) / C * ce;
      H[3 * p + i] = (o[6 + 2 * i] + I * o[6 + 2 * i + 1]) / C * ce;
    }
  }
  free(r0); free(l1); free(l2); free(dd); free(out);
}

int EH_s(double complex *E, double complex *H, double *rt, int type, DQD2 *qd)
{
  int did;
  eh_scattered(qd, 1, rt, type, E, H, &did);
  return did;
}

int EH_t(double complex *E, double complex *H, double *rt, int type, DQD2 *qd)
{
  int did = EH_s(E, H, rt, type, qd);
  if (did == 0) {
    double complex e[3], h[3];
    calc_mipw_EH(e, h, rt, &qd->pw);
    for (int i = 0; i < 3; i++) {
      E[i] += e[i];
      H[i] += h[i];
    }
  }
  return did;
}

int EH_i(double complex *E, double complex *H, double *rt, int type, DQD2 *qd)
{
  (void)type;
  int did = (qd->bd.ps == 1) ? q0_domain_id(rt, qd) : domain_id(rt, qd);
  calc_mipw_EH(E, H, rt, &qd->pw);
  return did;
}

/* Total fields at n points in one GPU launch; E, H: [n][3] complex, did: [n]. */
int bq_EH_t_batch(void *qdv, int n, const double *r, int type, double *E, double *H, int *did)
{
  DQD2 *qd = (DQD2 *)qdv;
  double complex *Ez = (double complex *)E, *Hz = (double complex *)H;
  eh_scattered(qd, n, r, type, Ez, Hz, did);
  for (int p = 0; p < n; p++) {
    if (did[p] != 0) continue;
    double complex e[3], h[3];
    calc_mipw_EH(e, h, (double *)(r + 3 * p), &qd->pw);
    for (int i = 0; i < 3; i++) {
      Ez[3 * p + i] += e[i];
      Hz[3 * p + i] += h[i];
    }
  }
  return 0;
}

int mEMP_i(double complex *U, double *rt, int type, DQD2 *qd)
{
  (void)type;
  double complex e[3], h[3];
  int did = (qd->bd.ps == 1) ? q0_domain_id(rt, qd) : domain_id(rt, qd);
  calc_mipw_EH(e, h, rt, &qd->pw);
  for (int i = 0; i < 3; i++) U[i] = e[i];
  U[3] = 0.0;
  return did;
}

int EH_mEMP_i(double complex *E, double complex *H, double *rt, int type, DQD2 *qd)
{
  return EH_i(E, H, rt, type, qd);
}

void EH_i_bd(double complex *E, double complex *H, int did, int t, double zeta_t, double eta_t, int type, DQD2 *qd)
{
  (void)type;
  int td = qd->bd.sb[did].sid[t], s = abs(td);
  double r[3];
  for (int i = 0; i < 3; i++)
    r[i] = qd->bd.cr[s][i][0] + qd->bd.cr[s][i][1] * zeta_t + qd->bd.cr[s][i][2] * eta_t +
           qd->bd.cr[s][i][3] * zeta_t * eta_t;
  if (did == 0) calc_mipw_EH(E, H, r, &qd->pw);
  else for (int i = 0; i < 3; i++) E[i] = H[i] = 0.0;
}

static void not_yet(const char *name)
{
  printf("%s(): not available in this build of the GPU library. Exit...\n", name);
  exit(1);
}

int mEMP_s(double complex *U, double *rt, int type, DQD2 *qd) { (void)U; (void)rt; (void)type; (void)qd; not_yet("mEMP_s"); return -1; }
int mEMP_t(double complex *U, double *rt, int type, DQD2 *qd) { (void)U; (void)rt; (void)type; (void)qd; not_yet("mEMP_t"); return -1; }
int EH_mEMP_s(double complex *E, double complex *H, double *rt, int type, DQD2 *qd)
{ (void)E; (void)H; (void)rt; (void)type; (void)qd; not_yet("EH_mEMP_s"); return -1; }
int EH_mEMP_t(double complex *E, double complex *H, double *rt, int type, DQD2 *qd)
{ (void)E; (void)H; (void)rt; (void)type; (void)qd; not_yet("EH_mEMP_t"); return -1; }
void EH_s_bd(double complex *E, double complex *H, int did, int t, double zeta_t, double eta_t, int type, DQD2 *qd)
{ (void)E; (void)H; (void)did; (void)t; (void)zeta_t; (void)eta_t; (void)type; (void)qd; not_yet("EH_s_bd"); }
void EH_t_bd(double complex *E, double complex *H, int did, int t, double zeta_t, double eta_t, int type, DQD2 *qd)
{ (void)E; (void)H; (void)did; (void)t; (void)zeta_t; (void)eta_t; (void)type; (void)qd; not_yet("EH_t_bd"); }
