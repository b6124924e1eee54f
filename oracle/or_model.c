/*
 * or_model.c -- model set-up for the oracle: incident plane waves, element constants,
 * quadrature tables, sub-domain lists and incident field at element nodes.
 *
 * TEST INFRASTRUCTURE ONLY (see oracle.h).
 *
 * Follows d3qd2_src/d3qd2_setup.c (initialize_dqd2 :105-139, init_elem_const :1139-1222,
 * malloc/init_sub_domain :1224-1307, init_boundary_data :1309-1349), d3qd2_src/ipw.c
 * (setup_ipw :121-145, calc_ipw_dEHdv :229-256) and com_src/gauleg.c:3-36.
 */
#include <stdlib.h>
#include <string.h>
#include <math.h>
#include <complex.h>
#include "oracle.h"

#define P44_N 0.5773502691896257645091
#define P49_N 0.7745966692414833770359
#define P49_W0 0.3086419753086419753086
#define P49_W1 0.4938271604938271604938
#define P49_W2 0.7901234567901234567901
#define MEPS 1.0e-15
/* linear-triangle collocation nodes (0.4 of the circumradius towards each vertex) plus the
 * centroid, and the 7-point rule, on the reference triangle with vertices (-1/2, -sqrt3/2),
 * (1, 0), (-1/2, sqrt3/2) (reference d3qd2_src/d3qd2_const.h:63-73, d3qd2_setup.c:1208-1217) */
#define P34_N0 0.2
#define P34_N1 0.3464101615137754587055
#define P34_W0 0.5208333333333333333333
#define P34_W1 0.5625
#define P37_N0 0.3480702390148154917985
#define P37_N1 0.2052130961576726346557
#define P37_N2 0.6028753385763033130539
#define P37_N3 0.3554395089236065568357
#define P37_W0 0.1259391805448271525957
#define P37_W1 0.1323941527885061807376
#define P37_W2 0.225
#define SQ3 1.732050807568877293528

/* Gauss-Legendre nodes/weights on [x1,x2] by Newton iteration on P_n (ascending). */
void or_gauleg(double x1, double x2, double *x, double *w, int n)
{
  int m = (n + 1) / 2;
  double xm = 0.5 * (x2 + x1), xl = 0.5 * (x2 - x1);
  for (int i = 1; i <= m; i++) {
    double z = cos(M_PI * (i - 0.25) / (n + 0.5)), z1, pp;
    do {
      double p1 = 1.0, p2 = 0.0, p3;
      for (int j = 1; j <= n; j++) {
        p3 = p2;
        p2 = p1;
        p1 = ((2.0 * j - 1.0) * z * p2 - (j - 1.0) * p3) / j;
      }
      pp = n * (z * p1 - p2) / (z * z - 1.0);
      z1 = z;
      z = z1 - p1 / pp;
    } while (fabs(z - z1) > 3.0e-16);
    x[i - 1] = xm - xl * z;
    x[n - i] = xm + xl * z;
    w[i - 1] = 2.0 * xl / ((1.0 - z * z) * pp * pp);
    w[n - i] = w[i - 1];
  }
  if (x1 == -x2 && n % 2 == 1) x[n / 2] = 0.0;
}

static void setup_wave(or_ipw *p)
{
  double cs = 1.0 / sqrt(pow(cabs(p->e0x), 2) + pow(cabs(p->e0y), 2));
  double complex ex = p->e0x * cs, ey = p->e0y * cs, hx = -ey * p->ni, hy = ex * p->ni;
  p->E0 = sqrt(p->power * 2.0 / p->ni);
  p->ki = 2.0 * M_PI * p->ni / p->lambda0;
  p->st = sin(p->theta); p->ct = cos(p->theta);
  p->sp = sin(p->phi); p->cp = cos(p->phi);
  double st = p->st, ct = p->ct, sp = p->sp, cp = p->cp;
  p->ex = ex * (sp * sp + cp * cp * ct) + ey * (sp * cp * ct - sp * cp);
  p->ey = ex * (sp * cp * ct - sp * cp) + ey * (sp * sp * ct + cp * cp);
  p->ez = -(ex * cp + ey * sp) * st;
  p->hx = hx * (sp * sp + cp * cp * ct) + hy * (sp * cp * ct - sp * cp);
  p->hy = hx * (sp * cp * ct - sp * cp) + hy * (sp * sp * ct + cp * cp);
  p->hz = -(hx * cp + hy * sp) * st;
}

void or_ipw_EH_dv(const or_model *m, const double *x, const double *v, double complex *e,
                  double complex *h, double complex *de, double complex *dh)
{
  for (int i = 0; i < 3; i++) e[i] = h[i] = de[i] = dh[i] = 0.0;
  for (int w = 0; w < m->n_ipw; w++) {
    const or_ipw *p = &m->ipw[w];
    double kx = p->st * p->cp, ky = p->st * p->sp, kz = p->ct;
    double te = p->ki * ((x[0] - p->fx) * kx + (x[1] - p->fy) * ky + (x[2] - p->fz) * kz);
    double complex Ee = (cos(te) + I * sin(te)) * p->E0;
    double complex g = v ? I * p->ki * (kx * v[0] + ky * v[1] + kz * v[2]) : 0.0;
    double complex ev[3] = {p->ex * Ee, p->ey * Ee, p->ez * Ee};
    double complex hv[3] = {p->hx * Ee, p->hy * Ee, p->hz * Ee};
    for (int i = 0; i < 3; i++) {
      e[i] += ev[i];
      h[i] += hv[i];
      de[i] += ev[i] * g;
      dh[i] += hv[i] * g;
    }
  }
}

void or_ipw_EH(const or_model *m, const double *x, double complex *e, double complex *h)
{
  double complex de[3], dh[3];
  or_ipw_EH_dv(m, x, NULL, e, h, de, dh);
}

static void rule_alloc(or_rule *r, int n)
{
  r->n = n;
  r->z = malloc(sizeof(double) * n);
  r->e = malloc(sizeof(double) * n);
  r->w = malloc(sizeof(double) * n);
}

/* Tensor Gauss-Legendre rule on the square [-1,1]^2 (reference bil_*_GL/_GH loops). */
static void rule_quad_gl(or_rule *r, int n, const double *x, const double *w)
{
  rule_alloc(r, n * n);
  for (int i = 0; i < n; i++)
    for (int j = 0; j < n; j++) {
      r->z[n * i + j] = x[i];
      r->e[n * i + j] = x[j];
      r->w[n * i + j] = w[i] * w[j];
    }
}

/* Gauss-Legendre rule on the reference triangle: the triangle is split into three
 * quadrilaterals at the edge midpoints and the centroid, each mapped bilinearly from
 * (a, b) in [-1,1]^2 with Jacobian (sqrt3/16)(1 + a/4 + b/4)
 * (reference lit_dcoef_v2_GL d3qd2_src/lit_dcoef_v2.c:213-236). */
static void rule_tri_gl(or_rule *r, int n, const double *x, const double *w)
{
  rule_alloc(r, 3 * n * n);
  int q = 0;
  for (int i = 0; i < n; i++)
    for (int j = 0; j < n; j++) {
      double a = x[i], b = x[j], wt = 0.0625 * SQ3 * (1.0 + 0.25 * a + 0.25 * b) * w[i] * w[j];
      r->z[q] = 0.125 * (3.0 + 2.0 * a + 2.0 * b + a * b);
      r->e[q] = 0.125 * SQ3 * (-a + b);
      r->w[q++] = wt;
      r->z[q] = 0.0625 * (-3.0 + a - 5.0 * b - a * b);
      r->e[q] = 0.0625 * SQ3 * (3.0 + 3.0 * a + b + a * b);
      r->w[q++] = wt;
      r->z[q] = 0.0625 * (-3.0 - 5.0 * a + b - a * b);
      r->e[q] = 0.0625 * SQ3 * (-3.0 - a - 3.0 * b - a * b);
      r->w[q++] = wt;
    }
}

static void rule_list(or_rule *r, int n, const double *z, const double *e, const double *w, double scale)
{
  rule_alloc(r, n);
  for (int i = 0; i < n; i++) { r->z[i] = z[i]; r->e[i] = e[i]; r->w[i] = w[i] * scale; }
}

or_model *or_model_new(int MN, const double *n_re_im, int ps, const double *vd1, const double *vd2,
                       int n_ipw, const double *ip, int Nn, const double *rn, int Ne,
                       const int *ed, const int *md, const int *sd)
{
  or_model *m = calloc(1, sizeof(or_model));
  m->MN = MN; m->ps = ps; m->Nn = Nn; m->Ne = Ne; m->n_ipw = n_ipw;
  m->ipw = calloc(n_ipw, sizeof(or_ipw));
  for (int w = 0; w < n_ipw; w++) {
    const double *q = ip + 12 * w;
    or_ipw *p = &m->ipw[w];
    p->lambda0 = q[0]; p->ni = q[1]; p->power = q[2];
    p->e0x = q[3] + I * q[4]; p->e0y = q[5] + I * q[6];
    p->fx = q[7]; p->fy = q[8]; p->fz = q[9]; p->theta = q[10]; p->phi = q[11];
    setup_wave(p);
  }
  m->lambda0 = m->ipw[0].lambda0;
  m->n0 = m->ipw[0].ni;
  /* QPD2 from the first wave (d3qd2_setup.c:121-126) */
  m->qd.k = m->ipw[0].ki;
  m->qd.vk[0] = m->qd.k * m->ipw[0].st * m->ipw[0].cp;
  m->qd.vk[1] = m->qd.k * m->ipw[0].st * m->ipw[0].sp;
  m->qd.vk[2] = m->qd.k * m->ipw[0].ct;
  if (m->qd.vk[0] == 0.0) m->qd.vk[0] = m->qd.k * MEPS;
  if (m->qd.vk[1] == 0.0) m->qd.vk[1] = m->qd.k * MEPS;
  m->qd.vd1[0] = vd1[0]; m->qd.vd1[1] = vd1[1];
  m->qd.vd2[0] = vd2[0]; m->qd.vd2[1] = vd2[1];
  m->n = calloc(MN + 1, sizeof(double complex));
  m->kn = calloc(MN + 1, sizeof(double complex));
  m->n[0] = m->n0;
  for (int i = 1; i <= MN; i++) m->n[i] = n_re_im[2 * (i - 1)] + I * n_re_im[2 * (i - 1) + 1];
  for (int i = 0; i <= MN; i++) m->kn[i] = 2.0 * M_PI / m->lambda0 * m->n[i];

  m->rn = malloc(sizeof(double) * 3 * (Nn + 1));
  memcpy(m->rn, rn, sizeof(double) * 3 * (Nn + 1));
  m->ed = malloc(sizeof(int) * 4 * (Ne + 1));
  memcpy(m->ed, ed, sizeof(int) * 4 * (Ne + 1));
  m->md = malloc(sizeof(int) * (Ne + 1));
  m->sd = malloc(sizeof(int) * (Ne + 1));
  memcpy(m->md, md, sizeof(int) * (Ne + 1));
  memcpy(m->sd, sd, sizeof(int) * (Ne + 1));
  /* element node numbering (d3qd2_setup.c:829-840) */
  m->eni = calloc(4 * (Ne + 1), sizeof(int));
  int nc = 0;
  for (int i = 1; i <= Ne; i++) {
    int tri = (m->ed[4 * i + 3] == 0);
    for (int j = 0; j < 4; j++) m->eni[4 * i + j] = (tri && j == 3) ? -1 : nc++;
  }
  m->NN = nc;

  /* element constants (d3qd2_setup.c:1139-1185) */
  m->cr = calloc(12 * (Ne + 1), sizeof(double));
  m->cw = calloc(9 * (Ne + 1), sizeof(double));
  for (int i = 1; i <= Ne; i++) {
    double rc[3][4];
    for (int d = 0; d < 3; d++)
      for (int j = 0; j < 4; j++) rc[d][j] = m->ed[4 * i + j] ? m->rn[3 * m->ed[4 * i + j] + d] : 0.0;
    double *cr = m->cr + 12 * i, *cw = m->cw + 9 * i;
    if (m->ed[4 * i + 3] != 0) {
      for (int d = 0; d < 3; d++) {
        cr[4 * d + 0] = 0.25 * (rc[d][0] + rc[d][1] + rc[d][2] + rc[d][3]);
        cr[4 * d + 1] = 0.25 * (-rc[d][0] + rc[d][1] + rc[d][2] - rc[d][3]);
        cr[4 * d + 2] = 0.25 * (-rc[d][0] - rc[d][1] + rc[d][2] + rc[d][3]);
        cr[4 * d + 3] = 0.25 * (rc[d][0] - rc[d][1] + rc[d][2] - rc[d][3]);
      }
      int pa[3] = {1, 2, 0}, pb[3] = {2, 0, 1};
      for (int d = 0; d < 3; d++) {
        int a = pa[d], b = pb[d];
        cw[3 * d + 0] = 0.125 * ((rc[a][0] - rc[a][2]) * (rc[b][1] - rc[b][3]) - (rc[a][1] - rc[a][3]) * (rc[b][0] - rc[b][2]));
        cw[3 * d + 1] = 0.125 * ((rc[a][2] - rc[a][3]) * (rc[b][0] - rc[b][1]) - (rc[a][0] - rc[a][1]) * (rc[b][2] - rc[b][3]));
        cw[3 * d + 2] = 0.125 * ((rc[a][1] - rc[a][2]) * (rc[b][0] - rc[b][3]) - (rc[a][0] - rc[a][3]) * (rc[b][1] - rc[b][2]));
      }
    } else {
      for (int d = 0; d < 3; d++) {
        cr[4 * d + 0] = 1.0 / 3.0 * (rc[d][0] + rc[d][1] + rc[d][2]);
        cr[4 * d + 1] = 1.0 / 3.0 * (-rc[d][0] + 2.0 * rc[d][1] - rc[d][2]);
        cr[4 * d + 2] = 1.0 / sqrt(3.0) * (rc[d][2] - rc[d][0]);
        cr[4 * d + 3] = 0.0;
      }
      cw[0] = (rc[1][1] - rc[1][0]) * (rc[2][2] - rc[2][0]) - (rc[2][1] - rc[2][0]) * (rc[1][2] - rc[1][0]);
      cw[3] = (rc[2][1] - rc[2][0]) * (rc[0][2] - rc[0][0]) - (rc[0][1] - rc[0][0]) * (rc[2][2] - rc[2][0]);
      cw[6] = (rc[0][1] - rc[0][0]) * (rc[1][2] - rc[1][0]) - (rc[1][1] - rc[1][0]) * (rc[0][2] - rc[0][0]);
    }
  }
  const double z44[4] = {-P44_N, P44_N, P44_N, -P44_N}, e44[4] = {-P44_N, -P44_N, P44_N, P44_N};
  const double z49[9] = {-P49_N, P49_N, P49_N, -P49_N, 0.0, P49_N, 0.0, -P49_N, 0.0};
  const double e49[9] = {-P49_N, -P49_N, P49_N, P49_N, -P49_N, 0.0, P49_N, 0.0, 0.0};
  const double w49[9] = {P49_W0, P49_W0, P49_W0, P49_W0, P49_W1, P49_W1, P49_W1, P49_W1, P49_W2};
  for (int j = 0; j < 4; j++) { m->zt44[j] = z44[j]; m->et44[j] = e44[j]; m->wt44[j] = 1.0; }
  for (int j = 0; j < 9; j++) { m->zt49[j] = z49[j]; m->et49[j] = e49[j]; m->wt49[j] = w49[j]; }
  const double z34[4] = {-P34_N0, 2.0 * P34_N0, -P34_N0, 0.0}, e34[4] = {-P34_N1, 0.0, P34_N1, 0.0};
  const double w34[4] = {P34_W0, P34_W0, P34_W0, -P34_W1};
  const double z37[7] = {-P37_N0, 2.0 * P37_N0, -P37_N0, P37_N1, -2.0 * P37_N1, P37_N1, 0.0};
  const double e37[7] = {-P37_N2, 0.0, P37_N2, P37_N3, 0.0, -P37_N3, 0.0};
  const double w37[7] = {P37_W0, P37_W0, P37_W0, P37_W1, P37_W1, P37_W1, P37_W2};
  for (int j = 0; j < 4; j++) { m->zt34[j] = z34[j]; m->et34[j] = e34[j]; m->wt34[j] = w34[j]; }
  for (int j = 0; j < 7; j++) { m->zt37[j] = z37[j]; m->et37[j] = e37[j]; m->wt37[j] = w37[j]; }
  or_gauleg(-1.0, 1.0, m->xli, m->wli, 12);
  or_gauleg(-1.0, 1.0, m->xhi, m->whi, 64);
  /* the 4- and 7-point triangle weights are fractions of the area 3 sqrt3 / 4 */
  rule_list(&m->rule[0][0], 4, m->zt44, m->et44, m->wt44, 1.0);
  rule_list(&m->rule[0][1], 9, m->zt49, m->et49, m->wt49, 1.0);
  rule_quad_gl(&m->rule[0][2], 12, m->xli, m->wli);
  rule_quad_gl(&m->rule[0][3], 64, m->xhi, m->whi);
  rule_list(&m->rule[1][0], 4, m->zt34, m->et34, m->wt34, 0.75 * SQ3);
  rule_list(&m->rule[1][1], 7, m->zt37, m->et37, m->wt37, 0.75 * SQ3);
  rule_tri_gl(&m->rule[1][2], 12, m->xli, m->wli);
  rule_tri_gl(&m->rule[1][3], 64, m->xhi, m->whi);

  /* sub-domain lists: main side positive, sub side negative (d3qd2_setup.c:1290-1307) */
  m->dNe = calloc(MN + 1, sizeof(int));
  for (int i = 1; i <= Ne; i++) { m->dNe[m->md[i]]++; m->dNe[m->sd[i]]++; }
  m->sid = calloc(MN + 1, sizeof(int *));
  for (int d = 0; d <= MN; d++) {
    m->sid[d] = calloc(m->dNe[d] + 1, sizeof(int));
    int c = 1;
    for (int i = 1; i <= Ne; i++) {
      if (m->md[i] == d) m->sid[d][c++] = i;
      else if (m->sd[i] == d) m->sid[d][c++] = -i;
    }
  }

  /* element nodes and incident field (d3qd2_setup.c:1309-1349): quads at the 2x2 Gauss
   * points, triangles at the three collocation nodes and the centroid */
  m->ren = calloc(12 * (Ne + 1), sizeof(double));
  m->wen = calloc(12 * (Ne + 1), sizeof(double));
  m->Ui = calloc(16 * (Ne + 1), sizeof(double complex));
  m->dUi = calloc(16 * (Ne + 1), sizeof(double complex));
  for (int i = 1; i <= Ne; i++) {
    const double *cr = m->cr + 12 * i, *cw = m->cw + 9 * i;
    int tri = (m->ed[4 * i + 3] == 0);
    for (int j = 0; j < 4; j++) {
      double z = tri ? m->zt34[j] : m->zt44[j], e = tri ? m->et34[j] : m->et44[j], r[3], w[3];
      for (int d = 0; d < 3; d++) {
        r[d] = cr[4 * d] + cr[4 * d + 1] * z + cr[4 * d + 2] * e + cr[4 * d + 3] * z * e;
        w[d] = cw[3 * d] + cw[3 * d + 1] * z + cw[3 * d + 2] * e;
        m->ren[12 * i + 3 * j + d] = r[d];
        m->wen[12 * i + 3 * j + d] = w[d];
      }
      double aw = sqrt(w[0] * w[0] + w[1] * w[1] + w[2] * w[2]);
      for (int d = 0; d < 3; d++) w[d] /= aw;
      double complex e3[3], h3[3], de[3], dh[3];
      or_ipw_EH_dv(m, r, w, e3, h3, de, dh);
      for (int d = 0; d < 3; d++) {
        m->Ui[16 * i + 4 * j + d] = e3[d];
        m->dUi[16 * i + 4 * j + d] = de[d];
      }
      m->Ui[16 * i + 4 * j + 3] = 0.0;
      m->dUi[16 * i + 4 * j + 3] = 0.0;
    }
  }
  return m;
}

void or_model_free(or_model *m)
{
  if (!m) return;
  for (int d = 0; d <= m->MN; d++) free(m->sid[d]);
  free(m->sid); free(m->dNe); free(m->ipw); free(m->n); free(m->kn);
  free(m->rn); free(m->ed); free(m->eni); free(m->md); free(m->sd);
  free(m->cr); free(m->cw); free(m->ren); free(m->wen); free(m->Ui); free(m->dUi);
  for (int a = 0; a < 2; a++)
    for (int b = 0; b < 4; b++) { free(m->rule[a][b].z); free(m->rule[a][b].e); free(m->rule[a][b].w); }
  free(m);
}

void or_model_export(const or_model *m, double *ren, double *wen, double *Ui, double *dUi,
                     double *qd, double *kn)
{
  int n = m->Ne + 1;
  if (ren) memcpy(ren, m->ren, sizeof(double) * 12 * n);
  if (wen) memcpy(wen, m->wen, sizeof(double) * 12 * n);
  if (Ui) memcpy(Ui, m->Ui, sizeof(double complex) * 16 * n);
  if (dUi) memcpy(dUi, m->dUi, sizeof(double complex) * 16 * n);
  if (qd) memcpy(qd, &m->qd, sizeof(or_qpd2));
  if (kn) memcpy(kn, m->kn, sizeof(double complex) * (m->MN + 1));
}

/* Replace the element node data by stored arrays (same layout as or_model_export): a model
 * read back from a result file carries the node points, normals and incident field the run
 * that wrote it computed (dat_read, d3qd2_src/d3qd2_setup.c:420-512). */
void or_model_set_nodes(or_model *m, const double *ren, const double *wen, const double *Ui, const double *dUi)
{
  int n = m->Ne + 1;
  memcpy(m->ren, ren, sizeof(double) * 12 * n);
  memcpy(m->wen, wen, sizeof(double) * 12 * n);
  memcpy(m->Ui, Ui, sizeof(double complex) * 16 * n);
  memcpy(m->dUi, dUi, sizeof(double complex) * 16 * n);
}

int or_model_dNe(const or_model *m, int d) { return m->dNe[d]; }
int or_model_NN(const or_model *m) { return m->NN; }
int or_model_sid(const or_model *m, int d, int i) { return m->sid[d][i]; }
