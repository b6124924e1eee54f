/*
 * bq_field.c -- field and force API of the reference (d3qd2_src/d3qd2_field.c,
 * d3qd2_src/d3qd2_force.c) on the GPU.
 *
 * The solved model is uploaded once per DQD2 (the first field call after solve_bieq or
 * dat_read) and kept in HBM until finalize_dqd2.  The scalar entry points (EH_s, EH_t ...)
 * are safe to call from several host threads at once, as the reference's examples do from
 * OpenMP loops: concurrent calls are gathered into one launch (point_call below).
 * bq_EH_t_batch, bq_EH_t_bd_batch and bq_EH_mEMP_t_batch evaluate many points in one launch;
 * force_FN is one launch of boundary queries followed by the stress sums on the host.
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

/* [n1][4][m] complex -> interleaved doubles; zeros when the array is absent */
static double *flat_z3(double complex ***a, int n1, int m)
{
  double *o = calloc((size_t)n1 * 4 * m * 2, sizeof(double));
  if (!a) return o;
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
  double **Ub = calloc(MN + 1, sizeof(double *)), **dUb = calloc(MN + 1, sizeof(double *));
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
    Ub[d] = flat_z3(s->U, s->Ne + 1, 4);
    dUb[d] = flat_z3(s->dU, s->Ne + 1, 4);
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
  f.Ub = (const double *const *)Ub;
  f.dUb = (const double *const *)dUb;
  f.xli = bd->xli; f.wli = bd->wli; f.xhi = bd->xhi; f.whi = bd->whi;
  f.z49 = bd->zt_49; f.e49 = bd->et_49; f.w49 = bd->wt_49;
  f.w34 = bd->wt_34; f.z37 = bd->zt_37; f.e37 = bd->et_37; f.w37 = bd->wt_37;
  /* node coordinates and element connectivity for the device domain test */
  double *rnf = calloc((size_t)(bd->Nn + 1) * 3, sizeof(double));
  int *edf = calloc((size_t)n1 * 4, sizeof(int));
  for (int i = 0; i <= bd->Nn; i++)
    for (int k = 0; k < 3; k++) rnf[3 * i + k] = bd->rn[i][k];
  for (int i = 0; i < n1; i++)
    for (int j = 0; j < 4; j++) edf[4 * i + j] = bd->ed[i][j];
  f.Nn = bd->Nn;
  f.rn = rnf;
  f.ed = edf;
  f.md = bd->md;
  f.sd = bd->sd;
  void *ctx = NULL;
  int rc = bq_field_ctx_create(&f, &ctx);
  free(rnf);
  free(edf);
  for (int d = 0; d <= MN; d++) {
    free(Eb[d]); free(dEb[d]); free(Hb[d]); free(dHb[d]); free(Ub[d]); free(dUb[d]);
  }
  free(E); free(ren); free(wen); free(kn); free(dNe); free(sid);
  free(Eb); free(dEb); free(Hb); free(dHb); free(Ub); free(dUb);
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

/* one launch over m queries (free points r0 or boundary points bt/bze); out [m][13 | 9 | 36]
 * for kinds 0, 1, 2 */
static double *launch(DQD2 *qd, int kind, int m, const double *r0, const int *dd, const int *bt, const double *bze,
                      int type)
{
  const size_t nv = kind == 0 ? 13 : (kind == 1 ? 9 : 36);
  double *out = calloc(nv * (m > 0 ? m : 1), sizeof(double));
  if (m > 0) {
    pthread_mutex_lock(&g_lock);
    void *ctx = context_for(qd);
    int rc = bq_field_eval(ctx, kind, m, r0, dd, bt, bze, type, out);
    pthread_mutex_unlock(&g_lock);
    if (rc) field_fail("field evaluation");
  }
  return out;
}

static void fold(DQD2 *qd, const double *rt, double *r0, int *l1, int *l2);

/* domain ids of n folded points in one device launch (the reference's domain_id per point) */
static void domain_ids(DQD2 *qd, int n, const double *r0, int *did)
{
  if (n <= 0) return;
  pthread_mutex_lock(&g_lock);
  void *ctx = context_for(qd);
  int rc = bq_field_domain_ids(ctx, n, r0, did);
  pthread_mutex_unlock(&g_lock);
  if (rc) field_fail("domain identification");
}

int bq_domain_id_batch(void *qdv, int n, const double *r, int *did)
{
  DQD2 *qd = (DQD2 *)qdv;
  if (n < 0 || (n > 0 && (!r || !did))) return -2;
  if (n == 0) return 0;
  double *r0 = malloc(sizeof(double) * 3 * (size_t)n);
  for (int p = 0; p < n; p++) {
    int l1, l2;
    fold(qd, r + 3 * p, r0 + 3 * p, &l1, &l2);
  }
  domain_ids(qd, n, r0, did);
  free(r0);
  return 0;
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

static double complex bloch(DQD2 *qd, int l1, int l2)
{
  if (l1 == 0 && l2 == 0) return 1.0;
  double arg = qd->bd.qd.vk[0] * (l1 * qd->bd.qd.vd1[0] + l2 * qd->bd.qd.vd2[0]) +
               qd->bd.qd.vk[1] * (l1 * qd->bd.qd.vd1[1] + l2 * qd->bd.qd.vd2[1]);
  return cos(arg) + I * sin(arg);
}

/* Scattered (or interior) fields at n points (EH_s, mEMP_s, d3qd2_src/d3qd2_field.c:10-59,
 * :203-270).  kind 0: X = E [n][3], Y = H [n][3]; kind 1: X = U [n][4], Y unused.  did[]
 * receives the domain ids; points outside every domain get zeros. */
static void scattered(DQD2 *qd, int kind, int n, const double *r, int type, double complex *X, double complex *Y,
                      int *did)
{
  const int nx = kind == 0 ? 3 : 4, nv = kind == 0 ? 13 : 9;
  size_t nn = n > 0 ? n : 1;
  double *r0 = malloc(sizeof(double) * 3 * nn);
  int *l1 = malloc(sizeof(int) * nn), *l2 = malloc(sizeof(int) * nn), *dd = malloc(sizeof(int) * nn);
  int m = 0;
  for (int p = 0; p < n; p++) fold(qd, r + 3 * p, r0 + 3 * p, &l1[p], &l2[p]);
  domain_ids(qd, n, r0, did);
  for (int p = 0; p < n; p++)
    if (did[p] >= 0) {
      for (int k = 0; k < 3; k++) r0[3 * m + k] = r0[3 * p + k];
      dd[m++] = did[p];
    }
  double *out = launch(qd, kind, m, r0, dd, NULL, NULL, type);
  for (int p = 0, q = 0; p < n; p++) {
    for (int i = 0; i < nx; i++) {
      X[nx * p + i] = 0.0;
      if (kind == 0) Y[nx * p + i] = 0.0;
    }
    if (did[p] < 0) continue;
    const double *o = out + (size_t)nv * q++;
    double C = (did[p] == 0) ? 1.0 + o[nv - 1] : o[nv - 1];
    double complex ce = bloch(qd, l1[p], l2[p]);
    for (int i = 0; i < nx; i++) X[nx * p + i] = (o[2 * i] + I * o[2 * i + 1]) / C * ce;
    if (kind == 0)
      for (int i = 0; i < 3; i++) Y[3 * p + i] = (o[6 + 2 * i] + I * o[6 + 2 * i + 1]) / C * ce;
  }
  free(r0); free(l1); free(l2); free(dd); free(out);
}

/* ---- concurrent single-point calls ------------------------------------------------
 * The reference's examples call EH_t from OpenMP loops, one point per thread
 * (example2.c:30-36, example3.c:120).  Each call joins a queue; a caller that finds no
 * evaluation in flight becomes the leader, takes every queued request like the one at the
 * head (same model, entry point and rule), evaluates them in one launch and wakes the
 * others.  Every point is computed by its own block in a fixed order, so a result does not
 * depend on which requests shared its launch. */
typedef struct preq {
  DQD2 *qd;
  int op, type; /* op 0: EH_s, 1: mEMP_s, 2: EH_mEMP_s */
  double r[3];
  double complex x[4], y[3];
  int ret, done;
  struct preq *next;
} preq;

static pthread_mutex_t q_mu = PTHREAD_MUTEX_INITIALIZER;
static pthread_cond_t q_cv = PTHREAD_COND_INITIALIZER;
static preq *q_head, *q_tail;
static int q_busy;

static void eh_memp(DQD2 *qd, int n, const double *r, int type, int total, double complex *E, double complex *H,
                    int *ret);

static void run_batch(preq **b, int n)
{
  DQD2 *qd = b[0]->qd;
  const int op = b[0]->op, type = b[0]->type, nx = op == 1 ? 4 : 3;
  double *r = malloc(sizeof(double) * 3 * n);
  int *ret = malloc(sizeof(int) * n);
  double complex *X = malloc(sizeof(double complex) * 4 * n), *Y = malloc(sizeof(double complex) * 3 * n);
  for (int i = 0; i < n; i++) memcpy(r + 3 * i, b[i]->r, sizeof(double) * 3);
  if (op == 2) eh_memp(qd, n, r, type, 0, X, Y, ret);
  else scattered(qd, op, n, r, type, X, Y, ret);
  for (int i = 0; i < n; i++) {
    for (int c = 0; c < nx; c++) b[i]->x[c] = X[nx * i + c];
    if (op != 1)
      for (int c = 0; c < 3; c++) b[i]->y[c] = Y[3 * i + c];
    b[i]->ret = ret[i];
  }
  free(r); free(ret); free(X); free(Y);
}

static int point_call(DQD2 *qd, int op, int type, const double *rt, double complex *X, double complex *Y)
{
  preq me;
  memset(&me, 0, sizeof me);
  me.qd = qd;
  me.op = op;
  me.type = type;
  memcpy(me.r, rt, sizeof(double) * 3);
  pthread_mutex_lock(&q_mu);
  if (q_tail) q_tail->next = &me;
  else q_head = &me;
  q_tail = &me;
  while (!me.done) {
    if (q_busy) {
      pthread_cond_wait(&q_cv, &q_mu);
      continue;
    }
    q_busy = 1;
    const preq *h = q_head;
    int n = 0;
    for (preq *p = q_head; p; p = p->next)
      if (p->qd == h->qd && p->op == h->op && p->type == h->type) n++;
    preq **b = malloc(sizeof(preq *) * n);
    DQD2 *hq = h->qd;
    int hop = h->op, hty = h->type, m = 0;
    preq *keep = NULL, *kt = NULL;
    for (preq *p = q_head, *nx; p; p = nx) {
      nx = p->next;
      p->next = NULL;
      if (p->qd == hq && p->op == hop && p->type == hty) {
        b[m++] = p;
      } else {
        if (kt) kt->next = p;
        else keep = p;
        kt = p;
      }
    }
    q_head = keep;
    q_tail = kt;
    pthread_mutex_unlock(&q_mu);
    run_batch(b, m);
    pthread_mutex_lock(&q_mu);
    for (int i = 0; i < m; i++) b[i]->done = 1;
    free(b);
    q_busy = 0;
    pthread_cond_broadcast(&q_cv);
  }
  pthread_mutex_unlock(&q_mu);
  for (int c = 0; c < (op == 1 ? 4 : 3); c++) X[c] = me.x[c];
  if (op != 1 && Y)
    for (int c = 0; c < 3; c++) Y[c] = me.y[c];
  return me.ret;
}

int EH_s(double complex *E, double complex *H, double *rt, int type, DQD2 *qd)
{
  return point_call(qd, 0, type, rt, E, H);
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
  scattered(qd, 0, n, r, type, Ez, Hz, did);
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

int mEMP_s(double complex *U, double *rt, int type, DQD2 *qd)
{
  return point_call(qd, 1, type, rt, U, NULL);
}

int mEMP_t(double complex *U, double *rt, int type, DQD2 *qd)
{
  int did = mEMP_s(U, rt, type, qd);
  if (did == 0) {
    double complex e[3], h[3];
    calc_mipw_EH(e, h, rt, &qd->pw);
    for (int i = 0; i < 3; i++) U[i] += e[i];
  }
  return did;
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

/* position on element seid at (zeta, eta) (r_bd, d3qd2_src/d3qd2_elem.c:18-29) */
static void r_on(const BOUD *bd, int seid, double zt, double et, double *r)
{
  int s = abs(seid);
  for (int i = 0; i < 3; i++) {
    r[i] = bd->cr[s][i][0] + bd->cr[s][i][1] * zt + bd->cr[s][i][2] * et;
    if (bd->ed[s][3] != 0) r[i] += bd->cr[s][i][3] * zt * et;
  }
}

static int bd_query_ok(DQD2 *qd, int did, int t)
{
  return did >= 0 && did <= qd->MN && t >= 1 && t <= qd->bd.sb[did].Ne;
}

/* Fields at n boundary points (EH_s_bd / EH_t_bd, d3qd2_src/d3qd2_field.c:299-391): point p
 * is (zt[p], et[p]) on the t[p]-th element (1-based) of domain did[p]; total adds the
 * incident wave in domain 0.  Returns -1 (nothing evaluated) on an invalid query. */
static int bd_fields(DQD2 *qd, int n, const int *did, const int *t, const double *zt, const double *et, int type,
                     int total, double complex *E, double complex *H)
{
  for (int p = 0; p < n; p++)
    if (!bd_query_ok(qd, did[p], t[p])) return -1;
  size_t nn = n > 0 ? n : 1;
  double *ze = malloc(sizeof(double) * 2 * nn);
  for (int p = 0; p < n; p++) {
    ze[2 * p] = zt[p];
    ze[2 * p + 1] = et[p];
  }
  double *out = launch(qd, 0, n, NULL, did, t, ze, type);
  for (int p = 0; p < n; p++) {
    const double *o = out + 13 * (size_t)p;
    double C = (did[p] == 0) ? 1.0 + o[12] : o[12];
    for (int i = 0; i < 3; i++) {
      E[3 * p + i] = (o[2 * i] + I * o[2 * i + 1]) / C;
      H[3 * p + i] = (o[6 + 2 * i] + I * o[6 + 2 * i + 1]) / C;
    }
    if (total && did[p] == 0) {
      double r[3];
      double complex e[3], h[3];
      r_on(&qd->bd, qd->bd.sb[0].sid[t[p]], zt[p], et[p], r);
      calc_mipw_EH(e, h, r, &qd->pw);
      for (int i = 0; i < 3; i++) {
        E[3 * p + i] += e[i];
        H[3 * p + i] += h[i];
      }
    }
  }
  free(ze);
  free(out);
  return 0;
}

static void bd_one(DQD2 *qd, const char *name, double complex *E, double complex *H, int did, int t, double zt,
                   double et, int type, int total)
{
  if (bd_fields(qd, 1, &did, &t, &zt, &et, type, total, E, H)) {
    printf("%s(): element %d of domain %d does not exist. Exit...\n", name, t, did);
    exit(1);
  }
}

void EH_s_bd(double complex *E, double complex *H, int did, int t, double zeta_t, double eta_t, int type, DQD2 *qd)
{
  bd_one(qd, "EH_s_bd", E, H, did, t, zeta_t, eta_t, type, 0);
}

void EH_t_bd(double complex *E, double complex *H, int did, int t, double zeta_t, double eta_t, int type, DQD2 *qd)
{
  bd_one(qd, "EH_t_bd", E, H, did, t, zeta_t, eta_t, type, 1);
}

int bq_EH_t_bd_batch(void *qdv, int n, const int *did, const int *t, const double *zt, const double *et, int type,
                     double *E, double *H)
{
  if (n < 0 || type < 0 || type > 3) return -1;
  return bd_fields((DQD2 *)qdv, n, did, t, zt, et, type, 1, (double complex *)E, (double complex *)H);
}

void EH_i_bd(double complex *E, double complex *H, int did, int t, double zeta_t, double eta_t, int type, DQD2 *qd)
{
  (void)type;
  double r[3];
  r_on(&qd->bd, qd->bd.sb[did].sid[t], zeta_t, eta_t, r);
  if (did == 0) calc_mipw_EH(E, H, r, &qd->pw);
  else for (int i = 0; i < 3; i++) E[i] = H[i] = 0.0;
}

/* Maxwell stress tensor T = 1/2 Re(eps E E* + H H*) - 1/4 (eps |E|^2 + |H|^2) I with
 * eps = Re(n0)^2, applied to the surface element w: adds wt T w to F and wt (r - rc) x (T w)
 * to N (the stress sums of bil_force_4p, d3qd2_src/d3qd2_force.c:56-110). */
static void stress_add(double epd, const double complex *e, const double complex *h, const double *r,
                       const double *w, const double *rc, double wt, double *F, double *N)
{
  double aE2 = 0.0, aH2 = 0.0, T[9];
  for (int j = 0; j < 3; j++) {
    aE2 += creal(e[j] * conj(e[j]));
    aH2 += creal(h[j] * conj(h[j]));
  }
  double Tc = 0.25 * (epd * aE2 + aH2);
  for (int j = 0; j < 3; j++)
    for (int k = 0; k < 3; k++) T[3 * j + k] = 0.5 * (epd * creal(e[j] * conj(e[k])) + creal(h[j] * conj(h[k])));
  T[0] -= Tc;
  T[4] -= Tc;
  T[8] -= Tc;
  double d[3] = {r[0] - rc[0], r[1] - rc[1], r[2] - rc[2]};
  for (int j = 0; j < 3; j++) {
    F[j] += (T[3 * j] * w[0] + T[3 * j + 1] * w[1] + T[3 * j + 2] * w[2]) * wt;
    N[j] += ((d[1] * T[3 * j + 2] - d[2] * T[3 * j + 1]) * w[0] + (d[2] * T[3 * j] - d[0] * T[3 * j + 2]) * w[1] +
             (d[0] * T[3 * j + 1] - d[1] * T[3 * j]) * w[2]) * wt;
  }
}

/* Force F and torque N about rc on the objects (force_FN, d3qd2_src/d3qd2_force.c:10-53):
 * Maxwell stress integrated over domain 0's boundary.  type 0: the stored node fields
 * plus the incident wave at the element nodes (a triangle's centroid point needs a boundary
 * query); type 1: boundary queries at the 9 (quad) or 7 (triangle) point rule.  Every
 * boundary query is evaluated in one launch with the node rule (PREC_DEF_FN) for the other
 * elements.  Element sums are added in element order; returns -1 if domain 0 is empty. */
int force_FN(double *F, double *N, double *rc, int type, DQD2 *qd)
{
  BOUD *bd = &qd->bd;
  SUBD *s0 = &bd->sb[0];
  int ne = s0->Ne;
  for (int j = 0; j < 3; j++) F[j] = N[j] = 0.0;
  if (ne <= 0) return -1;
  /* the queries */
  int nq = 0;
  for (int t = 1; t <= ne; t++) {
    int tri = bd->ed[abs(s0->sid[t])][3] == 0;
    nq += type == 0 ? (tri ? 1 : 0) : (tri ? 7 : 9);
  }
  size_t qn = nq > 0 ? nq : 1;
  int *qd_ = malloc(sizeof(int) * qn), *qt = malloc(sizeof(int) * qn);
  double *qz = malloc(sizeof(double) * qn), *qe = malloc(sizeof(double) * qn);
  double complex *qE = malloc(sizeof(double complex) * 3 * qn), *qH = malloc(sizeof(double complex) * 3 * qn);
  for (int t = 1, q = 0; t <= ne; t++) {
    int tri = bd->ed[abs(s0->sid[t])][3] == 0;
    if (type == 0) {
      if (!tri) continue;
      qd_[q] = 0; qt[q] = t; qz[q] = bd->zt_34[3]; qe[q] = bd->et_34[3]; q++;
    } else {
      int np = tri ? 7 : 9;
      for (int i = 0; i < np; i++, q++) {
        qd_[q] = 0;
        qt[q] = t;
        qz[q] = tri ? bd->zt_37[i] : bd->zt_49[i];
        qe[q] = tri ? bd->et_37[i] : bd->et_49[i];
      }
    }
  }
  if (nq > 0 && bd_fields(qd, nq, qd_, qt, qz, qe, PREC_DEF_FN, 1, qE, qH)) {
    printf("force_FN(): invalid boundary query. Exit...\n");
    exit(1);
  }
  double epd = creal(qd->n[0]) * creal(qd->n[0]);
  for (int t = 1, q = 0; t <= ne; t++) {
    int as = abs(s0->sid[t]), tri = bd->ed[as][3] == 0;
    int np = type == 0 ? 4 : (tri ? 7 : 9);
    double f[3] = {0.0, 0.0, 0.0}, nn[3] = {0.0, 0.0, 0.0};
    for (int i = 0; i < np; i++) {
      double r[3], w[3], wt;
      double complex e[3], h[3];
      if (type == 0 && !(tri && i == 3)) {
        double complex te[3], th[3];
        for (int j = 0; j < 3; j++) {
          r[j] = bd->ren[as][i][j];
          w[j] = bd->wen[as][i][j];
        }
        calc_mipw_EH(te, th, r, &qd->pw);
        for (int j = 0; j < 3; j++) {
          e[j] = s0->E[t][i][j] + te[j];
          h[j] = s0->H[t][i][j] + th[j];
        }
        wt = tri ? bd->wt_34[i] : bd->wt_44[i];
      } else {
        double zt = qz[q], et = qe[q];
        wt = type == 0 ? bd->wt_34[3] : (tri ? bd->wt_37[i] : bd->wt_49[i]);
        r_on(bd, as, zt, et, r);
        for (int j = 0; j < 3; j++) {
          w[j] = bd->cw[as][j][0] + bd->cw[as][j][1] * zt + bd->cw[as][j][2] * et;
          e[j] = qE[3 * q + j];
          h[j] = qH[3 * q + j];
        }
        q++;
      }
      stress_add(epd, e, h, r, w, rc, wt, f, nn);
    }
    double sc = tri ? -0.5 : -1.0;
    for (int j = 0; j < 3; j++) {
      F[j] += sc * f[j];
      N[j] += sc * nn[j];
    }
  }
  free(qd_); free(qt); free(qz); free(qe); free(qE); free(qH);
  return 0;
}

/* E, H from the modified potential and its gradient at n points (EH_mEMP_s,
 * d3qd2_src/d3qd2_field.c:90-173): E = A - grad phi, H = lambda0 / (2 pi i) curl A.  A point
 * where a solid-angle derivative exceeds CBD_CDF = 10 (too close to the boundary) gets zero
 * fields and ret = -did (-99 for the open region); otherwise ret = did (-1 outside every
 * domain).  total adds the incident wave in the open region (EH_mEMP_t). */
static void eh_memp(DQD2 *qd, int n, const double *r, int type, int total, double complex *E, double complex *H,
                    int *ret)
{
  size_t nn = n > 0 ? n : 1;
  double *r0 = malloc(sizeof(double) * 3 * nn);
  int *l1 = malloc(sizeof(int) * nn), *l2 = malloc(sizeof(int) * nn), *dd = malloc(sizeof(int) * nn);
  int m = 0;
  for (int p = 0; p < n; p++) fold(qd, r + 3 * p, r0 + 3 * p, &l1[p], &l2[p]);
  domain_ids(qd, n, r0, ret);
  for (int p = 0; p < n; p++)
    if (ret[p] >= 0) {
      for (int k = 0; k < 3; k++) r0[3 * m + k] = r0[3 * p + k];
      dd[m++] = ret[p];
    }
  double *out = launch(qd, 2, m, r0, dd, NULL, NULL, type);
  const double complex ch = qd->pw.lambda_0 / (2.0 * M_PI * I);
  for (int p = 0, q = 0; p < n; p++) {
    for (int i = 0; i < 3; i++) E[3 * p + i] = H[3 * p + i] = 0.0;
    if (ret[p] < 0) continue;
    const double *o = out + 36 * (size_t)q++;
    const int did = ret[p];
    if (fabs(o[33]) > 10.0 || fabs(o[34]) > 10.0 || fabs(o[35]) > 10.0) {
      ret[p] = did == 0 ? -99 : -did;
      continue;
    }
    double iC = (did == 0) ? 1.0 / (1.0 + o[32]) : 1.0 / o[32];
    double complex U[4], dU[3][4];
    for (int c = 0; c < 4; c++) U[c] = (o[2 * c] + I * o[2 * c + 1]) * iC;
    for (int d = 0; d < 3; d++)
      for (int c = 0; c < 4; c++)
        dU[d][c] = ((o[8 + 8 * d + 2 * c] + I * o[9 + 8 * d + 2 * c]) - U[c] * o[33 + d]) * iC;
    double complex ce = bloch(qd, l1[p], l2[p]);
    for (int d = 0; d < 3; d++) E[3 * p + d] = (-dU[d][3] + U[d]) * ce;
    H[3 * p + 0] = ch * (dU[1][2] - dU[2][1]) * ce;
    H[3 * p + 1] = ch * (dU[2][0] - dU[0][2]) * ce;
    H[3 * p + 2] = ch * (dU[0][1] - dU[1][0]) * ce;
    if (total && did == 0) {
      double complex e[3], h[3];
      calc_mipw_EH(e, h, (double *)(r + 3 * p), &qd->pw);
      for (int i = 0; i < 3; i++) {
        E[3 * p + i] += e[i];
        H[3 * p + i] += h[i];
      }
    }
  }
  free(r0); free(l1); free(l2); free(dd); free(out);
}

int EH_mEMP_s(double complex *E, double complex *H, double *rt, int type, DQD2 *qd)
{
  return point_call(qd, 2, type, rt, E, H);
}

int EH_mEMP_t(double complex *E, double complex *H, double *rt, int type, DQD2 *qd)
{
  int ret = EH_mEMP_s(E, H, rt, type, qd);
  if (ret == 0) {
    double complex e[3], h[3];
    calc_mipw_EH(e, h, rt, &qd->pw);
    for (int i = 0; i < 3; i++) {
      E[i] += e[i];
      H[i] += h[i];
    }
  }
  return ret;
}

int bq_EH_mEMP_t_batch(void *qdv, int n, const double *r, int type, double *E, double *H, int *ret)
{
  if (n < 0 || type < 0 || type > 3) return -1;
  eh_memp((DQD2 *)qdv, n, r, type, 1, (double complex *)E, (double complex *)H, ret);
  return 0;
}
