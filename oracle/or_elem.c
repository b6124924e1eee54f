/*
 * or_elem.c -- element-pair boundary-integral coefficients for bilinear quads and linear
 * triangles.
 *
 * TEST INFRASTRUCTURE ONLY (see oracle.h).
 *
 * Restates, for PREC_DEF_BE = 2 (d3qd2_src/bem3_emf_qd2.h:13) and SW_* = 0
 * (d3qd2_src/d3qd2_const.h:80-90):
 *   periodic regular pair, GL/GH         q0_bil_dcoef_v2_cc_GL/_GH  d3qd2_src/q0_bil_dcoef_v2_cc.c:129-290
 *                                        q0_lit_dcoef_v2_cc_GL/_GH  d3qd2_src/q0_lit_dcoef_v2_cc.c:155-301
 *   periodic singular pair (polar + FD)  q0_bil_dcoef_bd_vt2_cc_DV  :292-301, q0_bil_coef_bd
 *                                        d3qd2_src/q0_bil_coef.c:246-338, q0_bil_coef_bd_F
 *                                        q0_bil_dcoef_v2_cc.c:336-454;
 *                                        q0_lit_dcoef_bd_vt2_cc_DV q0_lit_dcoef_v2_cc.c:331-342,
 *                                        q0_lit_coef_bd d3qd2_src/q0_lit_coef.c:297-371, :549-602
 *   free-space regular pair, GL/GH       bil_dcoef_v2_GL/_GH  d3qd2_src/bil_dcoef_v2.c:141-325,
 *                                        lit_dcoef_v2_GL/_GH  d3qd2_src/lit_dcoef_v2.c:183-357, :381-413
 *   free-space singular pair (PV)        bil_coef_bd  d3qd2_src/bil_coef.c:244-342,
 *                                        bil_dcoef_bd_tz0_PV/_te0_PV  d3qd2_src/bil_dcoef.c:369-739,
 *                                        integrands :916-1630;
 *                                        lit_coef_bd d3qd2_src/lit_coef.c:290-349, :543-578,
 *                                        lit_dcoef_bd_tz0_PV/_te0_PV d3qd2_src/lit_dcoef.c:461-637,
 *                                        integrands :845-952
 *   dispatch + GL->GH fallback           d3qd2_src/d3qd2_elem.c:515-572, 830-886
 *   boundary-point coefficients          q0_coef_bd / coef_bd d3qd2_src/d3qd2_elem.c:684-770, :187-230:
 *                                        q0_bil_coef_bd q0_bil_coef.c:246-338, q0_lit_coef_bd
 *                                        q0_lit_coef.c:297-371, bil_coef_bd, lit_coef_bd
 *   field-point coefficients             q0_bil_coef_{4p,9p,GL,GH} d3qd2_src/q0_bil_coef.c:10-198,
 *                                        bil_coef_{4p,9p,GL,GH} d3qd2_src/bil_coef.c:25-202,
 *                                        q0_lit_coef_{4p,7p,GL,GH} d3qd2_src/q0_lit_coef.c:10-242,
 *                                        lit_coef_{4p,7p,GL,GH} d3qd2_src/lit_coef.c:26-242
 *
 * One restatement serves both shapes.  Every rule is a list of (zeta, eta, weight) points
 * with weights in reference-area units, and every integrand carries the surface Jacobian
 * |W| js, where W is the element's normal vector field (cw) and js = 1 for the bilinear map
 * (whose cw already is dr/dzeta x dr/deta) and js = 2/(3 sqrt3) for the triangle (whose cw
 * is (r1-r0)x(r2-r0) while dr/dzeta x dr/deta = cw 2/(3 sqrt3)).  The reference writes the
 * same constants as per-rule prefactors (1/24, 0.5, I_EP, I_SPSQ3 ...); they are equal.
 */
#include <stdio.h>
#include <stdlib.h>
#include <math.h>
#include <complex.h>
#include "oracle.h"

#define I_FP 7.957747154594766788444e-2 /* 1/(4 pi) */
#define P44_N 0.5773502691896257645091
#define P34_N0 0.2
#define SQ3 1.732050807568877293528
#define MEPS 1.0e-15
#define GEPS 1.0e-15
#define DEPS 0.1
#define CDFH 1.0e-3
#define CBD_SDF 0.5
#define GLN 12
#define GHN 64

typedef struct {
  double cr[3][4], cw[3][3];
  int tri;   /* 1 for a linear triangle */
  double js; /* surface Jacobian per |W|: 1 (quad) or 2/(3 sqrt3) (triangle) */
} elc;

static int is_tri_id(const or_model *m, int s) { return m->ed[4 * abs(s) + 3] == 0; }

static void get_el(const or_model *m, int s, elc *e)
{
  for (int d = 0; d < 3; d++) {
    for (int c = 0; c < 4; c++) e->cr[d][c] = m->cr[12 * s + 4 * d + c];
    for (int c = 0; c < 3; c++) e->cw[d][c] = m->cw[9 * s + 3 * d + c];
  }
  e->tri = is_tri_id(m, s);
  e->js = e->tri ? 2.0 / (3.0 * SQ3) : 1.0;
}

static int n_shape(const elc *e) { return e->tri ? 3 : 4; }

/* interpolation functions: bilinear ones equal to 1 at their own 2x2 Gauss point
 * (d3qd2_elem.c:903-915); linear ones equal to 1 at their own triangle collocation node
 * (lit_Nn d3qd2_elem.c:1047-1058) */
static double Nn(const elc *e, int n, double z, double t)
{
  if (e->tri) {
    double ia = 0.5 / P34_N0;
    switch (n) {
    case 0: return (1.0 - ia * (z + SQ3 * t)) / 3.0;
    case 1: return (1.0 + 2.0 * ia * z) / 3.0;
    default: return (1.0 - ia * (z - SQ3 * t)) / 3.0;
    }
  }
  double ia = 1.0 / P44_N;
  switch (n) {
  case 0: return 0.25 * (1.0 - z * ia) * (1.0 - t * ia);
  case 1: return 0.25 * (1.0 + z * ia) * (1.0 - t * ia);
  case 2: return 0.25 * (1.0 + z * ia) * (1.0 + t * ia);
  default: return 0.25 * (1.0 - z * ia) * (1.0 + t * ia);
  }
}

/* r * dN/dr of the linear triangle functions along direction (c, s) (lit_dNn_pr_cs,
 * d3qd2_src/lit_dcoef.c:845-856) */
static double tri_dN(int n, double c, double s)
{
  double ia = 0.5 / P34_N0;
  switch (n) {
  case 0: return -2.0 / 3.0 * ia * (c + SQ3 * s);
  case 1: return 4.0 / 3.0 * ia * c;
  default: return -2.0 / 3.0 * ia * (c - SQ3 * s);
  }
}

static void rw_at(const elc *e, double z, double t, double *r, double *w)
{
  for (int d = 0; d < 3; d++) {
    r[d] = e->cr[d][0] + e->cr[d][1] * z + e->cr[d][2] * t + e->cr[d][3] * z * t;
    w[d] = e->cw[d][0] + e->cw[d][1] * z + e->cw[d][2] * t;
  }
}

static double dot3(const double *a, const double *b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }
static double nrm3(const double *a) { return sqrt(dot3(a, a)); }

/* flat element test (bil_check_plane d3qd2_elem.c:987-996); triangles are flat */
static int is_plane(const elc *e)
{
  if (e->tri) return 1;
  double b[3] = {e->cr[0][1], e->cr[1][1], e->cr[2][1]}, c[3] = {e->cw[0][2], e->cw[1][2], e->cw[2][2]};
  return fabs(dot3(b, c) / (nrm3(b) * nrm3(c))) < GEPS;
}

static double bdC_of(const elc *e) { return e->cr[0][1] * e->cw[0][2] + e->cr[1][1] * e->cw[1][2] + e->cr[2][1] * e->cw[2][2]; }

/* collocation node coordinates of node nid on element s */
static void node_zeta_eta(const or_model *m, int s, int nid, double *zt, double *et)
{
  if (is_tri_id(m, s)) { *zt = m->zt34[nid]; *et = m->et34[nid]; }
  else { *zt = m->zt44[nid]; *et = m->et44[nid]; }
}

/* unit tangents at element node nid (d3qd2_elem.c:122-143) */
void or_node_tangents(const or_model *m, int s, int nid, double *vtz, double *vte)
{
  elc e;
  get_el(m, abs(s), &e);
  double zt, et;
  node_zeta_eta(m, abs(s), nid, &zt, &et);
  for (int d = 0; d < 3; d++) {
    vtz[d] = e.cr[d][1] + e.cr[d][3] * et;
    vte[d] = e.cr[d][2] + e.cr[d][3] * zt;
  }
  double a = nrm3(vtz), b = nrm3(vte);
  for (int d = 0; d < 3; d++) { vtz[d] /= a; vte[d] /= b; }
}

/* H and F change sign for the sub-domain side of an element (bil_convert_CC / lit_convert_CC,
 * d3qd2_elem.c:927-947, :1075-1093) */
static void flip_sign(double complex *C) { for (int i = 4; i < 9; i++) C[i] = -C[i]; }

static void qpgf_or_die(const or_model *m, int mode, const double *R, double complex *o)
{
  int err = or_qpgf(mode, R, MEPS, &m->qd, o);
  if (err < 0) {
    fprintf(stderr, "oracle: qpgf error %d at R=(%g,%g,%g)\n", err, R[0], R[1], R[2]);
    exit(1);
  }
}

static const or_rule *rule_of(const or_model *m, const elc *e, int type) { return &m->rule[e->tri][type]; }

/* ---- periodic regular pair ---------------------------------------------------------- */
static void q0_regular(const or_model *m, const double *rt, const double *v0, const double *v1,
                       int s, int type, double complex *CC, double complex *dC0, double complex *dC1)
{
  elc e;
  get_el(m, s, &e);
  const or_rule *q = rule_of(m, &e, type);
  int ns = n_shape(&e);
  for (int l = 0; l < 9; l++) CC[l] = dC0[l] = dC1[l] = 0.0;
  double F = 0, dF0 = 0, dF1 = 0;
  for (int p = 0; p < q->n; p++) {
    double z = q->z[p], t = q->e[p], wq = q->w[p], R[3], W[3];
    rw_at(&e, z, t, R, W);
    for (int d = 0; d < 3; d++) { R[d] -= rt[d]; W[d] *= e.js; }
    double aR = nrm3(R), iR = 1.0 / aR, aW = nrm3(W);
    double rdW = dot3(R, W) * iR, rdv0 = dot3(R, v0) * iR, rdv1 = dot3(R, v1) * iR;
    double complex o[10];
    qpgf_or_die(m, 0, R, o);
    double complex hn = o[1] * W[0] + o[2] * W[1] + o[3] * W[2];
    for (int ep = 0; ep < ns; ep++) {
      double Np = Nn(&e, ep, z, t);
      CC[ep] += wq * o[0] * Np * aW;
      CC[ep + 4] += wq * hn * Np;
    }
    F += wq * rdW * iR * iR;
    dF0 += wq * iR * iR * iR * (3.0 * rdv0 * rdW - dot3(W, v0));
    dF1 += wq * iR * iR * iR * (3.0 * rdv1 * rdW - dot3(W, v1));
  }
  CC[8] = F * I_FP;
  dC0[8] = dF0 * I_FP;
  dC1[8] = dF1 * I_FP;
}

/* ---- polar rule around (zt, et): visits every (r, theta) node of the sub-triangles ---- */
typedef void (*polar_fn)(void *ctx, double r, double cth, double sth, double wgt);

/* corners of the reference element seen from (zt, et): distances r[] and angles th[],
 * closed with th[n] = th[0] + 2 pi (bil_calc_node_r_th d3qd2_elem.c:1014-1033,
 * lit_calc_node_r_th :1149-1162).  Returns the number of sub-triangles. */
static int node_r_th(int tri, double zt, double et, double *r, double *th)
{
  if (tri) {
    double ztp = 1.0 - zt, ztm = -0.5 - zt, etp = 0.5 * SQ3 - et, etm = -0.5 * SQ3 - et;
    r[0] = sqrt(ztm * ztm + etm * etm); r[1] = sqrt(ztp * ztp + et * et); r[2] = sqrt(ztm * ztm + etp * etp);
    r[3] = r[0];
    th[0] = atan2(etm, ztm); th[1] = atan2(-et, ztp); th[2] = atan2(etp, ztm);
    th[3] = th[0] + 2.0 * M_PI;
    return 3;
  }
  double ztp = 1.0 - zt, ztm = -1.0 - zt, etp = 1.0 - et, etm = -1.0 - et;
  r[0] = sqrt(ztm * ztm + etm * etm); r[1] = sqrt(ztp * ztp + etm * etm);
  r[2] = sqrt(ztp * ztp + etp * etp); r[3] = sqrt(ztm * ztm + etp * etp);
  r[4] = r[0];
  th[0] = atan2(etm, ztm); th[1] = atan2(etm, ztp); th[2] = atan2(etp, ztp); th[3] = atan2(etp, ztm);
  th[4] = th[0] + 2.0 * M_PI;
  return 4;
}

/* theta: two GHN halves per sub-triangle; r: rule (nr, xr, wr) on [r_lo, rm(theta)] */
static void polar_visit(const or_model *m, int tri, double zt, double et, double rlo, int nr,
                        const double *xr, const double *wr, polar_fn f, void *ctx)
{
  double r[5], th[5];
  int nsub = node_r_th(tri, zt, et, r, th);
  for (int j = 0; j < nsub; j++) {
    double r0 = r[j], r1 = r[j + 1], t0 = th[j], t1 = th[j + 1];
    for (int h = 0; h < 2; h++) {
      for (int mm = 0; mm < GHN; mm++) {
        double theta = 0.25 * (t1 - t0) * m->xhi[mm] + (h == 0 ? 0.25 * (t1 + 3.0 * t0) : 0.25 * (3.0 * t1 + t0));
        double wt = 0.25 * (t1 - t0) * m->whi[mm];
        double rm = r0 * r1 * sin(t1 - t0) / (r0 * sin(theta - t0) - r1 * sin(theta - t1));
        double c = cos(theta), s = sin(theta);
        for (int i = 0; i < nr; i++) {
          double rr = 0.5 * (rm - rlo) * xr[i] + 0.5 * (rm + rlo);
          f(ctx, rr, c, s, wt * 0.5 * (rm - rlo) * wr[i]);
        }
      }
    }
  }
}

/* K = (R(zt + r c, et + r s) - R(zt, et)) / r and the Jacobian-scaled W at that point */
static void polar_KW(const elc *e, double zt, double et, double r, double c, double s, double *K, double *W,
                     double *z, double *t)
{
  *z = r * c + zt;
  *t = r * s + et;
  for (int d = 0; d < 3; d++) {
    K[d] = e->cr[d][1] * c + e->cr[d][2] * s + e->cr[d][3] * (r * c * s + zt * s + et * c);
    W[d] = (e->cw[d][0] + e->cw[d][1] * (*z) + e->cw[d][2] * (*t)) * e->js;
  }
}

typedef struct {
  const or_model *m;
  elc e;
  double zt, et;
  double complex k;
  double complex G[4], H[4];
  double F;
} pctx;

static void q0_sing_fn(void *vp, double r, double c, double s, double w)
{
  pctx *p = vp;
  double K[3], W[3], z, t, R[3];
  polar_KW(&p->e, p->zt, p->et, r, c, s, K, W, &z, &t);
  for (int d = 0; d < 3; d++) R[d] = r * K[d];
  /* the reference evaluates qG and dqG separately here (q0_bil_coef.c:444, :489;
   * q0_lit_coef.c:569, :596); their stop rules differ only for R_z == 0 exactly */
  double complex o[10], og[10];
  qpgf_or_die(p->m, 3, R, o);
  qpgf_or_die(p->m, 2, R, og);
  double aW = nrm3(W);
  double complex hn = o[1] * W[0] + o[2] * W[1] + o[3] * W[2];
  for (int n = 0; n < n_shape(&p->e); n++) {
    double Np = Nn(&p->e, n, z, t);
    p->G[n] += w * og[0] * Np * aW * r;
    p->H[n] += w * hn * Np * r;
  }
}

static void F_fn(void *vp, double r, double c, double s, double w)
{
  pctx *p = vp;
  double K[3];
  for (int d = 0; d < 3; d++) K[d] = p->e.cr[d][1] * c + p->e.cr[d][2] * s + p->e.cr[d][3] * (r * c * s + p->zt * s + p->et * c);
  double aK = nrm3(K);
  p->F += w * c * s / (aK * aK * aK);
}

/* F of the element at its own point (zt, et) by the polar rule (free-space solid angle) */
static double polar_F(const or_model *m, int s, double zt, double et)
{
  pctx p = {0};
  p.m = m;
  get_el(m, s, &p.e);
  if (is_plane(&p.e)) return 0.0;
  p.zt = zt; p.et = et;
  polar_visit(m, p.e.tri, zt, et, 0.0, GLN, m->xli, m->wli, F_fn, &p);
  return p.F * I_FP * bdC_of(&p.e);
}

static void q0_singular(const or_model *m, int s, int nid, double complex *CC, double complex *dCz,
                        double complex *dCe)
{
  pctx p = {0};
  p.m = m;
  get_el(m, s, &p.e);
  node_zeta_eta(m, s, nid, &p.zt, &p.et);
  polar_visit(m, p.e.tri, p.zt, p.et, 0.0, GLN, m->xli, m->wli, q0_sing_fn, &p);
  for (int i = 0; i < 9; i++) CC[i] = dCz[i] = dCe[i] = 0.0;
  for (int n = 0; n < n_shape(&p.e); n++) { CC[n] = p.G[n]; CC[n + 4] = p.H[n]; }
  /* a flat element has F = 0 and dF/dt = 0 at its own nodes; the triangle path sets them
   * so without computing (q0_lit_dcoef_v2_cc.c:331-342) */
  if (p.e.tri) return;
  CC[8] = polar_F(m, s, p.zt, p.et);
  double Tz[3], Te[3];
  for (int d = 0; d < 3; d++) {
    Tz[d] = p.e.cr[d][1] + p.e.cr[d][3] * p.et;
    Te[d] = p.e.cr[d][2] + p.e.cr[d][3] * p.zt;
  }
  double Fp = polar_F(m, s, p.zt + CDFH, p.et), Fm = polar_F(m, s, p.zt - CDFH, p.et);
  dCz[8] = (Fp - Fm) / (2.0 * CDFH) / nrm3(Tz);
  Fp = polar_F(m, s, p.zt, p.et + CDFH);
  Fm = polar_F(m, s, p.zt, p.et - CDFH);
  dCe[8] = (Fp - Fm) / (2.0 * CDFH) / nrm3(Te);
}

/* ---- free-space regular pair ------------------------------------------------------- */
static void fs_regular(const or_model *m, const double *rt, const double *v0, const double *v1,
                       int ss, double complex k, int type, double complex *CC, double complex *dC0,
                       double complex *dC1)
{
  int s = abs(ss);
  elc e;
  get_el(m, s, &e);
  const or_rule *q = rule_of(m, &e, type);
  int ns = n_shape(&e);
  for (int l = 0; l < 9; l++) CC[l] = dC0[l] = dC1[l] = 0.0;
  double F = 0, dF0 = 0, dF1 = 0;
  for (int p = 0; p < q->n; p++) {
    double z = q->z[p], t = q->e[p], wij = q->w[p], R[3], W[3];
    rw_at(&e, z, t, R, W);
    for (int d = 0; d < 3; d++) { R[d] -= rt[d]; W[d] *= e.js; }
    double aR = nrm3(R), iR = 1.0 / aR, aW = nrm3(W);
    double complex ca = I * k * aR, cb = ca - 1.0, ce = cexp(ca);
    double rdW = dot3(R, W) * iR, tD = rdW * iR * iR;
    double rdv0 = dot3(R, v0) * iR, rdv1 = dot3(R, v1) * iR, Wdv0 = dot3(W, v0), Wdv1 = dot3(W, v1);
    for (int ep = 0; ep < ns; ep++) {
      double Np = Nn(&e, ep, z, t);
      CC[ep] += wij * ce * iR * aW * Np;
      CC[ep + 4] += wij * cb * ce * tD * Np;
      dC0[ep] += wij * cb * ce * iR * iR * rdv0 * aW * Np;
      dC1[ep] += wij * cb * ce * iR * iR * rdv1 * aW * Np;
      dC0[ep + 4] += wij * ce * iR * iR * iR * ((k * k * aR * aR + 3.0 * cb) * rdv0 * rdW - cb * Wdv0) * Np;
      dC1[ep + 4] += wij * ce * iR * iR * iR * ((k * k * aR * aR + 3.0 * cb) * rdv1 * rdW - cb * Wdv1) * Np;
    }
    F += wij * tD;
    dF0 += wij * iR * iR * iR * (3.0 * rdv0 * rdW - Wdv0);
    dF1 += wij * iR * iR * iR * (3.0 * rdv1 * rdW - Wdv1);
  }
  for (int ep = 0; ep < 8; ep++) {
    CC[ep] *= I_FP;
    dC0[ep] *= (ep < 4 ? -I_FP : I_FP);
    dC1[ep] *= (ep < 4 ? -I_FP : I_FP);
  }
  CC[8] = F * I_FP;
  dC0[8] = dF0 * I_FP;
  dC1[8] = dF1 * I_FP;
  if (ss < 0) { flip_sign(CC); flip_sign(dC0); flip_sign(dC1); }
}

/* ---- free-space singular pair: G,H,F by polar rule (bil_coef_bd / lit_coef_bd) ------- */
static void fs_sing_GHF_fn(void *vp, double r, double c, double s, double w)
{
  pctx *p = vp;
  double K[3], W[3], z, t;
  polar_KW(&p->e, p->zt, p->et, r, c, s, K, W, &z, &t);
  double aK = nrm3(K), aW = nrm3(W);
  double complex kR = p->k * r * aK, ce = cexp(I * kR);
  for (int n = 0; n < n_shape(&p->e); n++) {
    double Np = Nn(&p->e, n, z, t);
    p->G[n] += w * ce / aK * aW * Np;
    p->H[n] += w * (I * kR - 1.0) * ce / (aK * aK * aK) * c * s * Np;
  }
  p->F += w * c * s / (aK * aK * aK);
}

/* PV tangential derivative (direction 0 = zeta, 1 = eta), bil_dcoef.c:369-739 */
typedef struct {
  const or_model *m;
  elc e;
  double zt, et, T[3];
  int dir;
  double complex k;
  double complex dG[4], dH[4];
  double dF;
} pvctx;

static double KdT_of(const pvctx *p, const double *K) { return dot3(K, p->T); }

/* D-bar region: r in [DEPS, rm] with GHN points */
static void pv_Db_fn(void *vp, double r, double c, double s, double w)
{
  pvctx *p = vp;
  double K[3], W[3], z, t;
  polar_KW(&p->e, p->zt, p->et, r, c, s, K, W, &z, &t);
  double aK = nrm3(K), iK = 1.0 / aK, aW = nrm3(W), KT = KdT_of(p, K);
  double complex kR = p->k * r * aK, cb = I * kR - 1.0, ce = cexp(I * kR);
  double cs_dir = (p->dir == 0) ? c : s;   /* (K.T) multiplies cos for tz, sin for te */
  double outer = (p->dir == 0) ? s : c;    /* theta factor outside the r integral */
  for (int n = 0; n < n_shape(&p->e); n++) {
    double Np = Nn(&p->e, n, z, t);
    p->dG[n] += w * cb * ce * iK * iK / r * (KT * iK) * Np * aW;
    p->dH[n] += w * outer * ce / (r * pow(aK, 5)) * ((kR * kR + 3.0 * cb) * KT * cs_dir - cb * aK * aK) * Np;
  }
  p->dF += w * outer / (r * pow(aK, 5)) * (3.0 * KT * cs_dir - aK * aK);
}

/* D-eps disk: theta in 4 quarters of [-pi, pi], r in [0, DEPS] with GLN points */
static void pv_De(pvctx *p)
{
  const or_model *m = p->m;
  double complex sG[4] = {0}, sH[4] = {0};
  for (int q = 0; q < 4; q++) {
    double off = -0.75 * M_PI + 0.5 * M_PI * q;
    for (int mm = 0; mm < GHN; mm++) {
      double theta = 0.25 * M_PI * m->xhi[mm] + off, wt = 0.25 * M_PI * m->whi[mm];
      double c = cos(theta), s = sin(theta);
      double cs_dir = (p->dir == 0) ? c : s, outer = (p->dir == 0) ? s : c;
      double complex g[4] = {0}, h0[4] = {0}, h1[4] = {0};
      for (int i = 0; i < GLN; i++) {
        double r = 0.5 * DEPS * (m->xli[i] + 1.0), wr = 0.5 * DEPS * m->wli[i];
        double K[3], W[3], z, t;
        polar_KW(&p->e, p->zt, p->et, r, c, s, K, W, &z, &t);
        double aK = nrm3(K), iK = 1.0 / aK, aW = nrm3(W), KT = KdT_of(p, K);
        double complex kR = p->k * aK * r, cb = I * kR - 1.0;
        double complex cc = ccos(kR), sn = csin(kR), ce = cexp(I * kR);
        for (int n = 0; n < 4; n++) {
          double Np = Nn(&p->e, n, z, t);
          g[n] += wr * (cb * sn + kR * cc) * iK * iK / r * (KT * iK) * aW * Np;
          h0[n] += wr * ce / pow(aK, 4) * ((p->k * kR + 3.0 * I * p->k) * KT * cs_dir - I * p->k * aK * aK) * Np;
          h1[n] += wr * sn / (pow(aK, 5) * r) * (3.0 * KT * cs_dir - aK * aK) * Np;
        }
      }
      for (int n = 0; n < 4; n++) {
        sG[n] += wt * g[n];
        sH[n] += wt * outer * (h0[n] - I * h1[n]);
      }
    }
  }
  for (int n = 0; n < 4; n++) {
    p->dG[n] += I * sG[n];
    p->dH[n] += sH[n];
  }
}

/* D-eps principal-value pairing (+r, -r): theta in [0, pi] as two halves */
static void pv_DP(pvctx *p)
{
  const or_model *m = p->m;
  double complex sG[4] = {0}, sH[4] = {0};
  double sF = 0;
  for (int h = 0; h < 2; h++) {
    for (int mm = 0; mm < GHN; mm++) {
      double theta = 0.25 * M_PI * m->xhi[mm] + (h == 0 ? 0.25 * M_PI : 0.75 * M_PI);
      double wt = 0.25 * M_PI * m->whi[mm];
      double c = cos(theta), s = sin(theta);
      double cs_dir = (p->dir == 0) ? c : s, outer = (p->dir == 0) ? s : c;
      double complex g[4] = {0}, hh[4] = {0};
      double f = 0;
      for (int i = 0; i < GLN; i++) {
        double r = 0.5 * DEPS * (m->xli[i] + 1.0), wr = 0.5 * DEPS * m->wli[i];
        double Kp[3], Km[3], Wp[3], Wm[3];
        double zp = r * c + p->zt, ep = r * s + p->et, zm = -r * c + p->zt, em = -r * s + p->et;
        for (int d = 0; d < 3; d++) {
          Kp[d] = p->e.cr[d][1] * c + p->e.cr[d][2] * s + p->e.cr[d][3] * (r * c * s + p->zt * s + p->et * c);
          Km[d] = p->e.cr[d][1] * c + p->e.cr[d][2] * s + p->e.cr[d][3] * (-r * c * s + p->zt * s + p->et * c);
          Wp[d] = p->e.cw[d][0] + p->e.cw[d][1] * zp + p->e.cw[d][2] * ep;
          Wm[d] = p->e.cw[d][0] + p->e.cw[d][1] * zm + p->e.cw[d][2] * em;
        }
        double aKp = nrm3(Kp), aKm = nrm3(Km), aWp = nrm3(Wp), aWm = nrm3(Wm);
        double KpT = KdT_of(p, Kp), KmT = KdT_of(p, Km);
        double complex ccp = ccos(p->k * aKp * r), ccm = ccos(-p->k * aKm * r);
        for (int n = 0; n < 4; n++) {
          double Np = Nn(&p->e, n, zp, ep), Nm = Nn(&p->e, n, zm, em);
          g[n] += wr / r * (ccp / (aKp * aKp) * (KpT / aKp) * Np * aWp - ccm / (aKm * aKm) * (KmT / aKm) * Nm * aWm);
          hh[n] += wr / r * (ccp / pow(aKp, 5) * (3.0 * KpT * cs_dir - aKp * aKp) * Np -
                             ccm / pow(aKm, 5) * (3.0 * KmT * cs_dir - aKm * aKm) * Nm);
        }
        f += wr / r * ((3.0 * KpT * cs_dir - aKp * aKp) / pow(aKp, 5) - (3.0 * KmT * cs_dir - aKm * aKm) / pow(aKm, 5));
      }
      for (int n = 0; n < 4; n++) {
        sG[n] += wt * g[n];
        sH[n] += wt * outer * hh[n];
      }
      sF += wt * outer * f;
    }
  }
  for (int n = 0; n < 4; n++) {
    p->dG[n] -= sG[n];
    p->dH[n] -= sH[n];
  }
  p->dF += sF;
}

/* Linear triangle, D-eps disk and its PV pairing.  The interpolation function is
 * N(zt + r c, et + r s) = N(zt, et) + r tri_dN(c, s); the constant part cancels in the
 * principal value and the r integrals of the linear part have closed forms, leaving theta
 * rules only (lit_sdG_De_theta, lit_sdG_DP_theta d3qd2_src/lit_dcoef.c:906-952; the
 * reference adds i/k times the first and subtracts 1/k times the second, :503-524). */
static void tri_pv_disk(pvctx *p)
{
  const or_model *m = p->m;
  double aW = nrm3((double[3]){p->e.cw[0][0], p->e.cw[1][0], p->e.cw[2][0]}) * p->e.js;
  double complex k = p->k;
  for (int n = 0; n < 3; n++) {
    double complex tDe = 0.0, tDP = 0.0;
    for (int q = 0; q < 4; q++) {
      double off = -0.75 * M_PI + 0.5 * M_PI * q;
      for (int mm = 0; mm < GHN; mm++) {
        double theta = 0.25 * M_PI * m->xhi[mm] + off, c = cos(theta), s = sin(theta);
        double K[3];
        for (int d = 0; d < 3; d++) K[d] = p->e.cr[d][1] * c + p->e.cr[d][2] * s;
        double aK = nrm3(K), iK = 1.0 / aK, KT = KdT_of(p, K) * iK;
        double complex keK = k * DEPS * aK, cs = csin(keK), cc = ccos(keK);
        tDe += 0.25 * M_PI * m->whi[mm] * tri_dN(n, c, s) * iK * iK * iK * KT *
               ((keK + I) * cs + (2.0 - I * keK) * cc - 2.0);
      }
    }
    for (int h = 0; h < 2; h++) {
      for (int mm = 0; mm < GHN; mm++) {
        double theta = 0.25 * M_PI * m->xhi[mm] + (h == 0 ? 0.25 * M_PI : 0.75 * M_PI);
        double c = cos(theta), s = sin(theta), K[3];
        for (int d = 0; d < 3; d++) K[d] = p->e.cr[d][1] * c + p->e.cr[d][2] * s;
        double aK = nrm3(K), iK = 1.0 / aK, KT = KdT_of(p, K) * iK;
        tDP += 0.25 * M_PI * m->whi[mm] * tri_dN(n, c, s) * csin(k * aK * DEPS) * iK * iK * iK * KT;
      }
    }
    p->dG[n] += (I * tDe / k - tDP / k) * aW;
  }
}

static void fs_pv_deriv(const or_model *m, int s, double zt, double et, double complex k, int dir,
                        double complex *dCC)
{
  pvctx p = {0};
  p.m = m;
  get_el(m, s, &p.e);
  p.zt = zt; p.et = et; p.k = k; p.dir = dir;
  for (int d = 0; d < 3; d++)
    p.T[d] = (dir == 0) ? p.e.cr[d][1] + p.e.cr[d][3] * et : p.e.cr[d][2] + p.e.cr[d][3] * zt;
  double iaT = 1.0 / nrm3(p.T);
  polar_visit(m, p.e.tri, zt, et, DEPS, GHN, m->xhi, m->whi, pv_Db_fn, &p);
  for (int i = 0; i < 9; i++) dCC[i] = 0.0;
  if (p.e.tri) {
    /* flat: dH/dt = dF/dt = 0 (lit_dcoef_bd_tz0_PV computes dG only) */
    tri_pv_disk(&p);
    for (int n = 0; n < 3; n++) dCC[n] = p.dG[n] * (-I_FP * iaT);
    return;
  }
  pv_De(&p);
  pv_DP(&p);
  int plane = is_plane(&p.e);
  double bdC = bdC_of(&p.e);
  for (int n = 0; n < 4; n++) {
    dCC[n] = p.dG[n] * (-I_FP * iaT);
    dCC[n + 4] = plane ? 0.0 : p.dH[n] * bdC * I_FP * iaT;
  }
  dCC[8] = plane ? 0.0 : p.dF * bdC * I_FP * iaT;
}

static void fs_singular(const or_model *m, int ss, int nid, double complex k, double complex *CC,
                        double complex *dCz, double complex *dCe)
{
  int s = abs(ss);
  pctx p = {0};
  p.m = m;
  get_el(m, s, &p.e);
  node_zeta_eta(m, s, nid, &p.zt, &p.et);
  p.k = k;
  polar_visit(m, p.e.tri, p.zt, p.et, 0.0, GLN, m->xli, m->wli, fs_sing_GHF_fn, &p);
  int plane = is_plane(&p.e);
  double bdC = bdC_of(&p.e);
  for (int i = 0; i < 9; i++) CC[i] = 0.0;
  for (int n = 0; n < n_shape(&p.e); n++) {
    CC[n] = p.G[n] * I_FP;
    CC[n + 4] = plane ? 0.0 : p.H[n] * I_FP * bdC;
  }
  CC[8] = plane ? 0.0 : p.F * I_FP * bdC;
  fs_pv_deriv(m, s, p.zt, p.et, k, 0, dCz);
  fs_pv_deriv(m, s, p.zt, p.et, k, 1, dCe);
  if (ss < 0) { flip_sign(CC); flip_sign(dCz); flip_sign(dCe); }
}

int or_pair(const or_model *m, int did, int t, int tn, int s, int type, double complex *CC,
            double complex *dCz, double complex *dCe)
{
  double vtz[3], vte[3];
  or_node_tangents(m, abs(t), tn, vtz, vte);
  const double *rt = m->ren + 12 * abs(t) + 3 * tn;
  int rule = (type >= 3) ? 3 : 2;
  if (did == 0 && m->ps == 1) {
    if (t != s) {
      q0_regular(m, rt, vtz, vte, s, rule, CC, dCz, dCe);
      if (type < 3 && (fabs(creal(dCz[8])) > CBD_SDF || fabs(creal(dCe[8])) > CBD_SDF)) {
        q0_regular(m, rt, vtz, vte, s, 3, CC, dCz, dCe);
        return 1;
      }
      return 0;
    }
    q0_singular(m, s, tn, CC, dCz, dCe);
    return 2;
  }
  double complex k = m->kn[did];
  if (t != s) {
    fs_regular(m, rt, vtz, vte, s, k, rule, CC, dCz, dCe);
    if (type < 3 && (fabs(creal(dCz[8])) > CBD_SDF || fabs(creal(dCe[8])) > CBD_SDF)) {
      fs_regular(m, rt, vtz, vte, s, k, 3, CC, dCz, dCe);
      return 1;
    }
    return 0;
  }
  fs_singular(m, s, tn, k, CC, dCz, dCe);
  return 2;
}

/* ---- field-point coefficients (no derivatives) ------------------------------------- */
void or_coef_rt(const or_model *m, int did, const double *rt, int ss, int type, double complex *CC)
{
  int s = abs(ss);
  elc e;
  get_el(m, s, &e);
  int periodic = (did == 0 && m->ps == 1), ns = n_shape(&e);
  double complex k = m->kn[did];
  const or_rule *q = rule_of(m, &e, type);
  for (int l = 0; l < 9; l++) CC[l] = 0.0;
  double F = 0;
  for (int p = 0; p < q->n; p++) {
    double R[3], W[3];
    rw_at(&e, q->z[p], q->e[p], R, W);
    if (type == 0) {
      /* the node rule reads the stored node data (q0_bil_coef.c:28-30, q0_lit_coef.c:29-31) */
      for (int d = 0; d < 3; d++) { R[d] = m->ren[12 * s + 3 * p + d]; W[d] = m->wen[12 * s + 3 * p + d]; }
    }
    for (int d = 0; d < 3; d++) { R[d] -= rt[d]; W[d] *= e.js; }
    double aR = nrm3(R), iR = 1.0 / aR, aW = nrm3(W);
    double tD = dot3(R, W) * iR * iR * iR;
    F += q->w[p] * tD;
    double complex g, h;
    if (periodic) {
      double complex o[10];
      qpgf_or_die(m, 0, R, o);
      g = o[0] * aW;
      h = o[1] * W[0] + o[2] * W[1] + o[3] * W[2];
    } else {
      double complex ca = I * k * aR, ce = cexp(ca);
      g = I_FP * ce * iR * aW;
      h = I_FP * (ca - 1.0) * ce * tD;
    }
    for (int ep = 0; ep < ns; ep++) {
      /* node points carry their own function only; the triangle centroid carries 1/3 each */
      double Np = (type == 0 && p < ns) ? (ep == p ? 1.0 : 0.0) : Nn(&e, ep, q->z[p], q->e[p]);
      CC[ep] += q->w[p] * g * Np;
      CC[ep + 4] += q->w[p] * h * Np;
    }
  }
  CC[8] = F * I_FP;
  if (!periodic && ss < 0) flip_sign(CC);
}

/* ---- field-point coefficients and their gradient in the field point ----------------
 * dC[p][.] = d CC / d rt_p (q0_dcoef_rt_grad / dcoef_rt_grad, d3qd2_src/d3qd2_elem.c:316-349,
 * :650-683; kernels q0_bil_dcoef_grad.c:10-328, q0_lit_dcoef_grad.c:10-389,
 * bil_dcoef_grad.c:10-298, lit_dcoef_grad.c:10-352).  Same rules and weights as or_coef_rt;
 * the periodic kernel takes the Hessian from the _v2 lattice sum. */
void or_coef_rt_grad(const or_model *m, int did, const double *rt, int ss, int type, double complex *CC,
                     double complex (*dC)[9])
{
  int s = abs(ss);
  elc e;
  get_el(m, s, &e);
  int periodic = (did == 0 && m->ps == 1), ns = n_shape(&e);
  double complex k = m->kn[did];
  const or_rule *q = rule_of(m, &e, type);
  for (int l = 0; l < 9; l++) CC[l] = dC[0][l] = dC[1][l] = dC[2][l] = 0.0;
  double F = 0, dF[3] = {0, 0, 0};
  for (int p = 0; p < q->n; p++) {
    double R[3], W[3];
    rw_at(&e, q->z[p], q->e[p], R, W);
    if (type == 0)
      for (int d = 0; d < 3; d++) { R[d] = m->ren[12 * s + 3 * p + d]; W[d] = m->wen[12 * s + 3 * p + d]; }
    for (int d = 0; d < 3; d++) { R[d] -= rt[d]; W[d] *= e.js; }
    double aR = nrm3(R), iR = 1.0 / aR, aW = nrm3(W);
    double rdW = dot3(R, W) * iR, tD = rdW * iR * iR;
    F += q->w[p] * tD;
    for (int d = 0; d < 3; d++) dF[d] += q->w[p] * iR * iR * iR * (3.0 * R[d] * iR * rdW - W[d]);
    double complex g, h, dg[3], dh[3];
    if (periodic) {
      double complex o[10];
      qpgf_or_die(m, 1, R, o);
      /* Hessian order xx, yy, zz, xy, yz, zx */
      double complex Hm[3][3] = {{o[4], o[7], o[9]}, {o[7], o[5], o[8]}, {o[9], o[8], o[6]}};
      g = o[0] * aW;
      h = o[1] * W[0] + o[2] * W[1] + o[3] * W[2];
      for (int d = 0; d < 3; d++) {
        dg[d] = -o[1 + d] * aW;
        dh[d] = -(Hm[d][0] * W[0] + Hm[d][1] * W[1] + Hm[d][2] * W[2]);
      }
    } else {
      double complex ca = I * k * aR, cb = ca - 1.0, ce = cexp(ca);
      g = I_FP * ce * iR * aW;
      h = I_FP * cb * ce * tD;
      for (int d = 0; d < 3; d++) {
        dg[d] = -I_FP * cb * ce * iR * iR * (R[d] * iR) * aW;
        dh[d] = I_FP * ce * iR * iR * iR * ((k * k * aR * aR + 3.0 * cb) * (R[d] * iR) * rdW - cb * W[d]);
      }
    }
    for (int ep = 0; ep < ns; ep++) {
      double Np = (type == 0 && p < ns) ? (ep == p ? 1.0 : 0.0) : Nn(&e, ep, q->z[p], q->e[p]);
      CC[ep] += q->w[p] * g * Np;
      CC[ep + 4] += q->w[p] * h * Np;
      for (int d = 0; d < 3; d++) {
        dC[d][ep] += q->w[p] * dg[d] * Np;
        dC[d][ep + 4] += q->w[p] * dh[d] * Np;
      }
    }
  }
  CC[8] = F * I_FP;
  for (int d = 0; d < 3; d++) dC[d][8] = dF[d] * I_FP;
  if (!periodic && ss < 0) { flip_sign(CC); flip_sign(dC[0]); flip_sign(dC[1]); flip_sign(dC[2]); }
}

/* ---- boundary-point coefficients: (zt, et) on element t, source s ---------------------- */
void or_coef_bd(const or_model *m, int did, int t, double zt, double et, int ss, int type, double complex *CC)
{
  int s = abs(ss), at = abs(t);
  elc e;
  get_el(m, at, &e);
  if (s != at) {
    double rt[3], W[3];
    rw_at(&e, zt, et, rt, W);
    or_coef_rt(m, did, rt, ss, type, CC);
    return;
  }
  pctx p = {0};
  p.m = m;
  p.e = e;
  p.zt = zt;
  p.et = et;
  for (int i = 0; i < 9; i++) CC[i] = 0.0;
  if (did == 0 && m->ps == 1) {
    polar_visit(m, e.tri, zt, et, 0.0, GLN, m->xli, m->wli, q0_sing_fn, &p);
    for (int n = 0; n < n_shape(&e); n++) { CC[n] = p.G[n]; CC[n + 4] = p.H[n]; }
    if (!e.tri) CC[8] = polar_F(m, s, zt, et);
    return;
  }
  p.k = m->kn[did];
  polar_visit(m, e.tri, zt, et, 0.0, GLN, m->xli, m->wli, fs_sing_GHF_fn, &p);
  int plane = is_plane(&e);
  double bdC = bdC_of(&e);
  for (int n = 0; n < n_shape(&e); n++) {
    CC[n] = p.G[n] * I_FP;
    CC[n + 4] = plane ? 0.0 : p.H[n] * I_FP * bdC;
  }
  CC[8] = plane ? 0.0 : p.F * I_FP * bdC;
  if (ss < 0) flip_sign(CC);
}

/* position on element s at (zt, et) (r_bd, d3qd2_src/d3qd2_elem.c:18-29) */
void or_r_bd(const or_model *m, int s, double zt, double et, double *r)
{
  elc e;
  double W[3];
  get_el(m, abs(s), &e);
  rw_at(&e, zt, et, r, W);
}
