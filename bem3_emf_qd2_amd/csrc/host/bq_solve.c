/*
 * bq_solve.c -- solve_bieq: the reference's pipeline (d3qd2_src/d3qd2_solve_bieq.c:36-84)
 * with every O(N^2) and O(N^3) step on the GPU through the solver session, and the
 * per-node bookkeeping on the host:
 *   coefficient matrices   -> device assembly of every domain
 *   global system + solve  -> device block elimination (two dense LUs of order 2 NN)
 *   tmatrix_bd_store       -> here (:456-544)
 *   q1_solve_eh_bv         -> device products dG U' - dH U, per-node 3x3 algebra here
 *                             (:931-1030, dudt_fcm :617-660, inv_matrix_33 :911-929)
 *   solve_deh_bv           -> device G^{-1} (H [E H]) per domain (:662-791)
 * Progress lines match the reference's stdout.
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <math.h>
#include <complex.h>
#include <time.h>
#include <pthread.h>
#include "bem3_emf_qd2.h"
#include "bq_amd.h"
#include "../bq_solver.h"

static bq_stats g_last;

const void *bq_solve_last_stats(void) { return &g_last; }

static double now_s(void)
{
  struct timespec ts;
  clock_gettime(CLOCK_MONOTONIC, &ts);
  return ts.tv_sec + 1e-9 * ts.tv_nsec;
}

static void gpu_fail(const char *what)
{
  printf("\n%s failed on the GPU: %s. Exit...\n", what, bq_last_error());
  exit(1);
}

/* flat host copies of the model for the device session */
typedef struct {
  bq_model m;
  double *E, *ren, *nrm, *Ui, *dUi, *kn;
  int *eni, *md, *sd, *dNe;
  int **sid;
} flat_model;

static void flatten(DQD2 *qd, flat_model *f)
{
  BOUD *bd = &qd->bd;
  int Ne = bd->Ne, n1 = Ne + 1;
  f->E = calloc((size_t)n1 * 24, sizeof(double));
  f->ren = calloc((size_t)n1 * 12, sizeof(double));
  f->nrm = calloc((size_t)n1 * 12, sizeof(double));
  f->Ui = calloc((size_t)n1 * 32, sizeof(double));
  f->dUi = calloc((size_t)n1 * 32, sizeof(double));
  f->kn = calloc((size_t)(qd->MN + 1) * 2, sizeof(double));
  f->eni = calloc((size_t)n1 * 4, sizeof(int));
  f->md = calloc(n1, sizeof(int));
  f->sd = calloc(n1, sizeof(int));
  f->dNe = calloc(qd->MN + 1, sizeof(int));
  f->sid = calloc(qd->MN + 1, sizeof(int *));
  for (int i = 1; i <= Ne; i++) {
    bq_pack_elem(bd->cr[i], bd->cw[i], bd->ed[i][3] == 0, f->E + 24 * (size_t)i);
    for (int j = 0; j < 4; j++) {
      double w[3] = {bd->wen[i][j][0], bd->wen[i][j][1], bd->wen[i][j][2]};
      vuni_d(w);
      for (int d = 0; d < 3; d++) {
        f->ren[12 * i + 3 * j + d] = bd->ren[i][j][d];
        f->nrm[12 * i + 3 * j + d] = w[d];
      }
      for (int c = 0; c < 4; c++) {
        f->Ui[32 * i + 8 * j + 2 * c] = creal(bd->Ui[i][j][c]);
        f->Ui[32 * i + 8 * j + 2 * c + 1] = cimag(bd->Ui[i][j][c]);
        f->dUi[32 * i + 8 * j + 2 * c] = creal(bd->dUi[i][j][c]);
        f->dUi[32 * i + 8 * j + 2 * c + 1] = cimag(bd->dUi[i][j][c]);
      }
      f->eni[4 * i + j] = bd->eni[i][j];
    }
    f->md[i] = bd->md[i];
    f->sd[i] = bd->sd[i];
  }
  for (int d = 0; d <= qd->MN; d++) {
    f->kn[2 * d] = creal(qd->kn[d]);
    f->kn[2 * d + 1] = cimag(qd->kn[d]);
    f->dNe[d] = bd->sb[d].Ne;
    f->sid[d] = bd->sb[d].sid;
  }
  bq_model *m = &f->m;
  m->Ne = Ne;
  m->NN = bd->NN;
  m->MN = qd->MN;
  m->ps = bd->ps;
  m->qd[0] = bd->qd.k;
  for (int i = 0; i < 3; i++) m->qd[1 + i] = bd->qd.vk[i];
  m->qd[4] = bd->qd.vd1[0];
  m->qd[5] = bd->qd.vd1[1];
  m->qd[6] = bd->qd.vd2[0];
  m->qd[7] = bd->qd.vd2[1];
  m->kn = f->kn;
  m->E = f->E;
  m->ren = f->ren;
  m->nrm = f->nrm;
  m->eni = f->eni;
  m->md = f->md;
  m->sd = f->sd;
  m->Ui = f->Ui;
  m->dUi = f->dUi;
  m->dNe = f->dNe;
  m->sid = (const int *const *)f->sid;
  m->xli = bd->xli;
  m->wli = bd->wli;
  m->xhi = bd->xhi;
  m->whi = bd->whi;
}

static void free_flat(flat_model *f)
{
  free(f->E); free(f->ren); free(f->nrm); free(f->Ui); free(f->dUi); free(f->kn);
  free(f->eni); free(f->md); free(f->sd); free(f->dNe); free(f->sid);
}

static int is_tri(BOUD *bd, int s) { return bd->ed[abs(s)][3] == 0; }

/* scatter the global solution into the per-domain boundary values */
static void store_solution(DQD2 *md, const double complex *X)
{
  size_t nn = md->bd.NN;
  for (int d = 0; d <= md->MN; d++) {
    for (int s = 1; s <= md->bd.sb[d].Ne; s++) {
      int sd = md->bd.sb[d].sid[s], asd = abs(sd), mdid = md->bd.md[asd], sdid = md->bd.sd[asd];
      int nl = is_tri(&md->bd, sd) ? 3 : 4;
      for (int l = 0; l < nl; l++) {
        size_t e = (size_t)md->bd.eni[asd][l];
        double complex *U = md->bd.sb[d].U[s][l], *dU = md->bd.sb[d].dU[s][l];
        if (d == mdid) {
          for (int i = 0; i < 4; i++) {
            U[i] = X[nn * i + e];
            dU[i] = X[nn * (i + 4) + e];
          }
          continue;
        }
        double complex k1 = md->kn[mdid], k2 = md->kn[sdid];
        double complex delt = k1 * k1 - k2 * k2, beps = k1 * k1 / (k2 * k2), sigm = 1.0 - beps;
        double n[3] = {md->bd.wen[asd][l][0], md->bd.wen[asd][l][1], md->bd.wen[asd][l][2]};
        vuni_d(n);
        double complex tU[4], tUi[3] = {0, 0, 0};
        tU[3] = X[nn * 3 + e];
        for (int i = 0; i < 3; i++) {
          tU[i] = X[nn * i + e];
          U[i] = tU[i];
          dU[i] = -(X[nn * (i + 4) + e] + delt * tU[3] * n[i]);
          if (mdid == 0) {
            tUi[i] = md->bd.Ui[asd][l][i];
            U[i] += tUi[i];
            dU[i] += -md->bd.dUi[asd][l][i];
          }
        }
        U[3] = tU[3];
        dU[3] = -(beps * X[nn * 7 + e] + sigm * (n[0] * tU[0] + n[1] * tU[1] + n[2] * tU[2]));
        if (mdid == 0) dU[3] += -sigm * (n[0] * tUi[0] + n[1] * tUi[1] + n[2] * tUi[2]);
      }
    }
  }
}

static void inv33(double *a, const double *vn, const double *vtz, const double *vte)
{
  double idet = 1.0 / (vn[0] * vtz[1] * vte[2] + vn[1] * vtz[2] * vte[0] + vn[2] * vtz[0] * vte[1] -
                       (vn[2] * vtz[1] * vte[0] + vn[1] * vtz[0] * vte[2] + vn[0] * vtz[2] * vte[1]));
  a[0] = vtz[1] * vte[2] - vtz[2] * vte[1];
  a[1] = -vn[1] * vte[2] + vn[2] * vte[1];
  a[2] = vn[1] * vtz[2] - vn[2] * vtz[1];
  a[3] = -vtz[0] * vte[2] + vtz[2] * vte[0];
  a[4] = vn[0] * vte[2] - vn[2] * vte[0];
  a[5] = -vn[0] * vtz[2] + vn[2] * vtz[0];
  a[6] = vtz[0] * vte[1] - vtz[1] * vte[0];
  a[7] = -vn[0] * vte[1] + vn[1] * vte[0];
  a[8] = vn[0] * vtz[1] - vn[1] * vtz[0];
  for (int i = 0; i < 9; i++) a[i] *= idet;
}

static void node_tangents(BOUD *bd, int s, int n, double *vtz, double *vte)
{
  for (int i = 0; i < 3; i++) {
    vtz[i] = bd->cr[s][i][1] + bd->cr[s][i][3] * bd->et_44[n];
    vte[i] = bd->cr[s][i][2] + bd->cr[s][i][3] * bd->zt_44[n];
  }
  vuni_d(vtz);
  vuni_d(vte);
}

/* E = -grad(phi) + A and H = (lambda0 / 2 pi i) curl A from the normal derivative (dUn)
 * and the two tangential derivatives; a = inverse of the [n, tz, te] frame */
static void eh_from(const double *a, const double complex *U, const double complex *dUn, const double complex *dUz,
                    const double complex *dUe, double complex ch, double complex *E, double complex *H)
{
#define GR(c, i) (dUn[c] * a[3 * (i)] + dUz[c] * a[3 * (i) + 1] + dUe[c] * a[3 * (i) + 2])
  for (int i = 0; i < 3; i++) E[i] = -GR(3, i) + U[i];
  H[0] = ch * (GR(2, 1) - GR(1, 2));
  H[1] = ch * (GR(0, 2) - GR(2, 0));
  H[2] = ch * (GR(1, 0) - GR(0, 1));
#undef GR
}

static int boundary_EH(DQD2 *md, bq_sess *S)
{
  double complex ch = md->pw.lambda_0 / (2.0 * M_PI * I);
  BOUD *bd = &md->bd;
  /* position of each element in the open region's list (0: not there) */
  int *pos0 = calloc((size_t)bd->Ne + 1, sizeof(int));
  for (int s = 1; s <= bd->sb[0].Ne; s++) {
    int e = bd->sb[0].sid[s];
    if (e > 0 && e <= bd->Ne && pos0[e] == 0) pos0[e] = s;
  }
  for (int d = 1; d <= md->MN; d++) {
    SUBD *sb = &bd->sb[d];
    int Ne = sb->Ne;
    size_t N = 4 * (size_t)Ne;
    double complex *u = calloc(N * 4, sizeof(double complex)), *du = calloc(N * 4, sizeof(double complex));
    double complex *raw = calloc(2 * N * 4, sizeof(double complex));
    double *tdF = calloc(3 * N, sizeof(double));
    for (int s = 1; s <= Ne; s++)
      for (int sn = 0; sn < 4; sn++)
        for (int i = 0; i < 4; i++) {
          u[(size_t)i * N + 4 * (s - 1) + sn] = sb->U[s][sn][i];
          du[(size_t)i * N + 4 * (s - 1) + sn] = sb->dU[s][sn][i];
        }
    if (bq_sess_dudt(S, d, (double *)u, (double *)du, (double *)raw) || bq_sess_get_dGH(S, d, NULL, NULL, tdF)) {
      free(pos0);
      free(u);
      free(du);
      free(raw);
      free(tdF);
      return -1;
    }
    for (int t = 1; t <= Ne; t++) {
      int at = sb->sid[t], atd = abs(at);
      double sig = at < 0 ? -1.0 : 1.0;
      for (int tn = 0; tn < 4; tn++) {
        if (tn == 3 && is_tri(bd, atd)) continue;
        size_t row = 4 * (size_t)(t - 1) + tn;
        double vtz[3], vte[3], vn[3], a[9];
        node_tangents(bd, atd, tn, vtz, vte);
        for (int i = 0; i < 3; i++) vn[i] = sig * bd->wen[atd][tn][i];
        vuni_d(vn);
        inv33(a, vn, vtz, vte);
        const double *f = tdF + 3 * row;
        double complex dUz[4], dUe[4];
        for (int i = 0; i < 4; i++) {
          dUz[i] = (raw[(size_t)i * 2 * N + 2 * row] - sb->U[t][tn][i] * f[1]) / f[0];
          dUe[i] = (raw[(size_t)i * 2 * N + 2 * row + 1] - sb->U[t][tn][i] * f[2]) / f[0];
        }
        eh_from(a, sb->U[t][tn], sb->dU[t][tn], dUz, dUe, ch, sb->E[t][tn], sb->H[t][tn]);
        if (bd->md[atd] != 0) continue;
        /* open-region side of the same node: scattered part, incident derivative removed */
        double complex e[3], h[3], dUiz[3], dUie[3], dh[3];
        calc_mipw_EH_dEHdv(e, h, dUiz, dh, bd->ren[atd][tn], vtz, &md->pw);
        calc_mipw_EH_dEHdv(e, h, dUie, dh, bd->ren[atd][tn], vte, &md->pw);
        int s = pos0[atd];
        if (s <= 0) {
          printf("solve_bieq(), open region id error! id=%d. Exit...\n", s);
          exit(0);
        }
        double complex mdn[4], z2[4], e2[4];
        for (int i = 0; i < 4; i++) {
          mdn[i] = -bd->sb[0].dU[s][tn][i];
          z2[i] = i < 3 ? dUz[i] - dUiz[i] : dUz[i];
          e2[i] = i < 3 ? dUe[i] - dUie[i] : dUe[i];
        }
        eh_from(a, bd->sb[0].U[s][tn], mdn, z2, e2, ch, bd->sb[0].E[s][tn], bd->sb[0].H[s][tn]);
      }
    }
    free(u);
    free(du);
    free(raw);
    free(tdF);
  }
  free(pos0);
  return 0;
}

static int boundary_dEH(DQD2 *md, bq_sess *S)
{
  BOUD *bd = &md->bd;
  for (int d = 0; d <= md->MN; d++) {
    SUBD *sb = &bd->sb[d];
    int Ne = sb->Ne;
    if (Ne == 0) continue;
    size_t N = 4 * (size_t)Ne;
    double complex *eh = calloc(N * 6, sizeof(double complex)), *x = calloc(N * 6, sizeof(double complex));
    for (int s = 1; s <= Ne; s++)
      for (int sn = 0; sn < 4; sn++)
        for (int i = 0; i < 3; i++) {
          eh[(size_t)i * N + 4 * (s - 1) + sn] = sb->E[s][sn][i];
          eh[(size_t)(i + 3) * N + 4 * (s - 1) + sn] = sb->H[s][sn][i];
        }
    if (bq_sess_deh(S, d, (double *)eh, (double *)x)) {
      free(eh);
      free(x);
      return -1;
    }
    for (int s = 1; s <= Ne; s++)
      for (int sn = 0; sn < 4; sn++)
        for (int i = 0; i < 3; i++) {
          sb->dE[s][sn][i] = x[(size_t)i * N + 4 * (s - 1) + sn];
          sb->dH[s][sn][i] = x[(size_t)(i + 3) * N + 4 * (s - 1) + sn];
        }
    free(eh);
    free(x);
  }
  return 0;
}

void bq_release_device_state(DQD2 *qd); /* bq_field.c */

/* Device sessions kept per model between solves: the mesh, element tables, matrices, factors
 * and work buffers stay resident, so that a sweep over incidences (bq_set_incidence between
 * solves) allocates nothing after its first point and a polarisation-only change reuses the
 * matrices and factors.  Keyed by the DQD2 address and its per-domain arrays (replaced when the
 * model is set up again); released by finalize_dqd2 or bq_release_solver. */
typedef struct sess_entry {
  DQD2 *qd;
  const void *key;
  bq_sess *S;
  flat_model f;
  bq_stats last; /* of the model's last solve_bieq */
  struct sess_entry *next;
} sess_entry;
static sess_entry *g_sess;
static pthread_mutex_t g_smu = PTHREAD_MUTEX_INITIALIZER;

static const void *model_key(DQD2 *qd) { return qd->bd.sb ? (const void *)qd->bd.sb[0].E : NULL; }

static void entry_free(sess_entry *e)
{
  if (e->S) bq_sess_destroy(e->S);
  free_flat(&e->f);
  free(e);
}

/* the model's session, created or brought up to date with its current incident wave */
static bq_sess *session_for(DQD2 *qd)
{
  pthread_mutex_lock(&g_smu);
  sess_entry **pp = &g_sess, *e = NULL;
  for (; *pp; pp = &(*pp)->next)
    if ((*pp)->qd == qd) break;
  if (*pp && (*pp)->key != model_key(qd)) { /* set up again since: a new session */
    sess_entry *old = *pp;
    *pp = old->next;
    entry_free(old);
  }
  e = *pp;
  pthread_mutex_unlock(&g_smu);
  if (e) {
    free_flat(&e->f);
    memset(&e->f, 0, sizeof e->f);
    flatten(qd, &e->f);
    if (bq_sess_update(e->S, &e->f.m) >= 0) return e->S;
    bq_release_solver(qd);
  }
  e = calloc(1, sizeof *e);
  e->qd = qd;
  e->key = model_key(qd);
  flatten(qd, &e->f);
  e->S = bq_sess_create(&e->f.m);
  if (!e->S) {
    free_flat(&e->f);
    free(e);
    return NULL;
  }
  pthread_mutex_lock(&g_smu);
  e->next = g_sess;
  g_sess = e;
  pthread_mutex_unlock(&g_smu);
  return e->S;
}

void bq_release_solver(void *qdv)
{
  DQD2 *qd = (DQD2 *)qdv;
  pthread_mutex_lock(&g_smu);
  sess_entry **pp = &g_sess;
  while (*pp && (*pp)->qd != qd) pp = &(*pp)->next;
  sess_entry *e = *pp;
  if (e) *pp = e->next;
  pthread_mutex_unlock(&g_smu);
  if (e) entry_free(e);
}

long long bq_solver_bytes(void *qdv)
{
  long long b = 0;
  pthread_mutex_lock(&g_smu);
  for (sess_entry *e = g_sess; e; e = e->next)
    if (!qdv || e->qd == (DQD2 *)qdv) b += (long long)bq_sess_bytes(e->S);
  pthread_mutex_unlock(&g_smu);
  return b;
}

/* Optional gate around the two heavy phases of a solve (0: assembly, FP64 vector work; 1:
 * factorisation, solve and boundary fields, matrix work), set by the request loop so that
 * concurrent solves take the phases in turn and one's assembly runs beside another's
 * factorisation instead of beside another assembly. */
static void (*g_gate)(int phase, int enter);

void bq_set_phase_gate(void (*fn)(int phase, int enter)) { g_gate = fn; }

static void gate(int phase, int enter)
{
  if (g_gate) g_gate(phase, enter);
}

/* The reference's solve_bieq pipeline; progress lines when verbose.  Returns 0, or a negative
 * code (the reference's QPGF codes -12 / -23, -1 device failure, -3 singular system) with the
 * reason in bq_last_error(). */
static int solve_core(DQD2 *qd, int verbose, const char **stage)
{
  double t0 = now_s(), ts;
  /* the solution arrays are rewritten in place: the field evaluator's copy of the previous
     solution must not be reused */
  bq_release_device_state(qd);
  if (verbose) {
    printf("\nsolve modified electromagnetic potential boundary value \n");
    printf("  coefficient matrix          ");
    fflush(stdout);
  }
  ts = now_s();
  *stage = "device set-up";
  bq_sess *S = session_for(qd);
  if (!S) return -1;
  int reused = bq_sess_assembled(S), rc = 0;
  *stage = "coefficient matrix assembly";
  if (!reused) {
    gate(0, 1);
    rc = bq_sess_assemble(S);
    gate(0, 0);
    if (rc) return rc < 0 ? rc : -1;
  }
  if (verbose) {
    printf("finished. Elapsed time : %5g (sec)\n", now_s() - ts);
    printf("  solve mEMP boundary value   ");
    fflush(stdout);
  }
  ts = now_s();
  size_t na = 8 * (size_t)qd->bd.NN;
  double complex *X = calloc(na, sizeof(double complex));
  *stage = "global system solve";
  gate(1, 1);
  rc = bq_sess_solve_global(S, (double *)X);
  if (rc) {
    gate(1, 0);
    free(X);
    return rc < 0 ? rc : -1;
  }
  store_solution(qd, X);
  free(X);
  if (verbose) {
    printf("finished. Elapsed time : %5g (sec)\n", now_s() - ts);
    printf("solve electromagnteic field boundary value \n");
    printf("  solve EH boundary value     ");
    fflush(stdout);
  }
  ts = now_s();
  *stage = "boundary field";
  rc = (boundary_EH(qd, S) || boundary_dEH(qd, S)) ? -1 : 0;
  gate(1, 0);
  if (rc) return rc;
  double tf = now_s() - ts;
  if (verbose) printf("finished. Elapsed time : %5g (sec)\n", tf);
  bq_stats st = *bq_sess_stats(S);
  st.t_fields_ms = tf * 1e3;
  st.reused = reused;
  if (reused) st.t_assemble_ms = 0.0;
  pthread_mutex_lock(&g_smu);
  g_last = st;
  for (sess_entry *e = g_sess; e; e = e->next)
    if (e->qd == qd) e->last = st;
  pthread_mutex_unlock(&g_smu);
  if (verbose) printf("Total elapsed time : %g (sec)\n", now_s() - t0);
  return 0;
}

void solve_bieq(DQD2 *qd)
{
  const char *stage = "";
  if (solve_core(qd, 1, &stage)) gpu_fail(stage);
}

int bq_solve_bieq_rc(void *qd)
{
  const char *stage = "";
  return solve_core((DQD2 *)qd, 0, &stage);
}

int bq_solve_stats(void *qdv, void *out)
{
  int rc = -1;
  pthread_mutex_lock(&g_smu);
  for (sess_entry *e = g_sess; e; e = e->next)
    if (e->qd == (DQD2 *)qdv) {
      memcpy(out, &e->last, sizeof e->last);
      rc = 0;
    }
  pthread_mutex_unlock(&g_smu);
  return rc;
}

/* Assemble the matrices of domain d for a set-up model and copy them to host buffers
 * (any pointer may be NULL): G, H [N][N], dG, dH [2N][N] complex, tdF [N][3].  Used by the
 * parity tests to compare the device assembly with the CPU restatement entry by entry. */
int bq_assemble_domain_host(void *qdv, int d, double *G, double *H, double *dG, double *dH, double *tdF)
{
  DQD2 *qd = (DQD2 *)qdv;
  flat_model f;
  memset(&f, 0, sizeof f);
  flatten(qd, &f);
  bq_sess *S = bq_sess_create(&f.m);
  int rc = S ? 0 : -1;
  if (!rc) rc = bq_sess_assemble(S);
  if (!rc && (G || H)) rc = bq_sess_get_GH(S, d, G, H);
  if (!rc && (dG || dH || tdF)) rc = bq_sess_get_dGH(S, d, dG, dH, tdF);
  if (S) {
    g_last = *bq_sess_stats(S);
    bq_sess_destroy(S);
  }
  free_flat(&f);
  return rc;
}
