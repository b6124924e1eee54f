/*
 * oracle.h -- CPU restatement of the bem3_emf_qd2 boundary-integral path.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product library links, loads or calls this
 * code; only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg use it, as a
 * checker and as the timed CPU port.  It follows the reference's algorithm step by step
 * (file:line cited per function) in plain C99 so that the HIP path can be compared with it
 * on the same inputs.
 *
 * Conventions: element ids are 1-based as in the reference (index 0 unused); complex
 * numbers cross this interface as interleaved (re, im) doubles.
 */
#ifndef BQ_ORACLE_H
#define BQ_ORACLE_H

#include <complex.h>

/* Same field order and meaning as the reference's QPD2 (qgf2_src/d3_qpgf_d2.h:32-36). */
typedef struct { double k, vk[3], vd1[2], vd2[2]; } or_qpd2;

/* ---- or_faddeeva.c ---------------------------------------------------------------- */
double complex or_w(double complex z, double relerr);
double or_erfcx(double x);
void or_faddeeva_w(const double *z, double relerr, double *w); /* ctypes helper */

/* ---- or_qpgf.c -------------------------------------------------------------------- */
/* mode: 0 = value+gradient (d3hm_qpgf_d2_v1), 1 = +Hessian (_v2), 2 = value only (_qG),
 *       3 = gradient only (_dqG).  out = qG[1], dqG[3], d2qG[6] as complex.
 * Return: term count (>0) or the reference's negative codes (-11..-24). */
int or_qpgf(int mode, const double *r, double eps, const or_qpd2 *qd, double complex *out);
int or_qpgf_batch(int mode, int n, const double *r, double eps, const or_qpd2 *qd, double *out);

/* ---- or_model.c ------------------------------------------------------------------- */
typedef struct {
  double lambda0, ni, power;
  double complex e0x, e0y;
  double fx, fy, fz, theta, phi;
  /* derived (reference d3qd2_src/ipw.c:121-145) */
  double E0, ki, ct, st, cp, sp;
  double complex ex, ey, ez, hx, hy, hz;
} or_ipw;

typedef struct { int n; double *z, *e, *w; } or_rule;

typedef struct {
  int MN, ps, Nn, Ne, NN, n_ipw;
  double lambda0, n0;
  double complex *n, *kn;     /* [MN+1] */
  or_qpd2 qd;
  or_ipw *ipw;                /* [n_ipw] */
  double *rn;                 /* [(Nn+1)*3] */
  int *ed, *eni;              /* [(Ne+1)*4] */
  int *md, *sd;               /* [Ne+1] */
  double *cr, *cw;            /* [(Ne+1)*12], [(Ne+1)*9]  (dim-major like cr[s][d][c]) */
  double *ren, *wen;          /* [(Ne+1)*4*3] */
  double complex *Ui, *dUi;   /* [(Ne+1)*4*4] */
  int *dNe;                   /* [MN+1] elements per domain */
  int **sid;                  /* sid[d][1..dNe[d]] signed element ids */
  double xli[12], wli[12], xhi[64], whi[64];
  double zt44[4], et44[4], wt44[4], zt49[9], et49[9], wt49[9];
  double zt34[4], et34[4], wt34[4], zt37[7], et37[7], wt37[7];
  /* Quadrature point lists over the reference element in (zeta, eta), weights in
   * reference-area units (sum = 4 for the square, 3 sqrt(3)/4 for the triangle):
   * rule[shape][type], shape 0 = bilinear quad, 1 = linear triangle; type 0 = element
   * nodes (4p), 1 = 9p / 7p, 2 = GL (12), 3 = GH (64). */
  or_rule rule[2][4];
} or_model;

or_model *or_model_new(int MN, const double *n_re_im, int ps, const double *vd1, const double *vd2,
                       int n_ipw, const double *ipw_params, int Nn, const double *rn, int Ne,
                       const int *ed, const int *md, const int *sd);
void or_model_free(or_model *m);
void or_model_export(const or_model *m, double *ren, double *wen, double *Ui, double *dUi,
                     double *qd, double *kn);
void or_model_set_nodes(or_model *m, const double *ren, const double *wen, const double *Ui, const double *dUi);
int or_model_dNe(const or_model *m, int d);
int or_model_NN(const or_model *m);
int or_model_sid(const or_model *m, int d, int i);
void or_gauleg(double x1, double x2, double *x, double *w, int n);
void or_ipw_EH(const or_model *m, const double *x, double complex *e, double complex *h);
void or_ipw_EH_dv(const or_model *m, const double *x, const double *v, double complex *e,
                  double complex *h, double complex *de, double complex *dh);

/* ---- or_elem.c -------------------------------------------------------------------- */
/* Pair coefficients for collocation node (t, tn) against element s of domain did:
 * CC[0..3]=G, CC[4..7]=H, CC[8]=F, dCz/dCe = tangential derivatives (reference
 * q0_dcoef_bd_node_t2_cc d3qd2_src/d3qd2_elem.c:830 and dcoef_bd_node_t2 :515).
 * t and s are signed element ids as stored in sid lists. Returns 1 if the GH rule was
 * used, 2 if singular, 0 otherwise. */
int or_pair(const or_model *m, int did, int t, int tn, int s, int type, double complex *CC,
            double complex *dCz, double complex *dCe);
/* Unit tangents at node nid of element s (reference tz_te_bd_node d3qd2_src/d3qd2_elem.c:122-143). */
void or_node_tangents(const or_model *m, int s, int nid, double *vtz, double *vte);
/* Coefficients for an arbitrary field point (reference q0_coef_rt / coef_rt). */
void or_coef_rt(const or_model *m, int did, const double *rt, int s, int type, double complex *CC);
/* as or_coef_rt, plus dC[p][.] = d CC / d rt_p (q0_dcoef_rt_grad / dcoef_rt_grad,
 * d3qd2_src/d3qd2_elem.c:316-349, :650-683) */
void or_coef_rt_grad(const or_model *m, int did, const double *rt, int s, int type, double complex *CC,
                     double complex (*dC)[9]);

/* Coefficients for the point (zt, et) of element t (signed ids as in sid lists): the polar
 * singular rule when s is t, else the rule `type` at that point (reference q0_coef_bd /
 * coef_bd, d3qd2_src/d3qd2_elem.c:684-770, :187-230). */
void or_coef_bd(const or_model *m, int did, int t, double zt, double et, int s, int type, double complex *CC);
void or_r_bd(const or_model *m, int s, double zt, double et, double *r);

/* ---- or_system.c ------------------------------------------------------------------ */
/* Assemble rows [t0, t1] (1-based element indices of domain did, inclusive) of the domain
 * matrices; N = 4*dNe.  tG,tH: [N][N]; tdG,tdH: [N][2][N]; tdF: [N][3]
 * (reference create_cmatrix_domain d3qd2_src/d3qd2_solve_bieq.c:165-256). */
void or_assemble_rows(const or_model *m, int did, int t0, int t1, double complex *tG,
                      double complex *tH, double complex *tdG, double complex *tdH, double *tdF);
/* Global system na x na (row-major) and RHS from the domain matrices
 * (reference create_matrix_csr_dac :284-421). Gs/Hs: per-domain matrices. */
void or_global_system(const or_model *m, double complex **Gs, double complex **Hs,
                      double complex *A, double complex *B);
/* Scatter the solution into per-domain U/dU (reference tmatrix_bd_store :456-544).
 * U, dU: per domain [(dNe+1)*4*4]. */
void or_store_solution(const or_model *m, const double complex *X, double complex **U,
                       double complex **dU);
/* Boundary E/H for domains >= 1 and the open side (reference q1_solve_eh_bv :931-1030). */
void or_boundary_EH(const or_model *m, double complex **U, double complex **dU,
                    double complex **tdG, double complex **tdH, double **tdF,
                    double complex **E, double complex **H);
/* Right-hand sides H_d [E H] for the dE/dH systems (reference solve_deh_bv :662-791). */
void or_deh_rhs(const or_model *m, int d, const double complex *tH, double complex *E,
                double complex *H, double complex *B);

/* ---- or_field.c ------------------------------------------------------------------- */
int or_domain_id(const or_model *m, const double *rt);
int or_EH_t(const or_model *m, double complex **E, double complex **dE, double complex **H,
            double complex **dH, const double *r, int type, double complex *Eo, double complex *Ho);
/* E, H from (A, phi) and its gradient at r (EH_mEMP_s / EH_mEMP_t, d3qd2_src/d3qd2_field.c:90-190);
 * returns the domain id, or -did (-99 for the open region) with zero fields when the point is
 * too close to the boundary. */
int or_EH_mEMP(const or_model *m, double complex **U, double complex **dU, const double *r, int type, int total,
               double complex *Eo, double complex *Ho);
/* Modified electromagnetic potential (A, phi) at r from the stored U, dU ([Ne+1][4][4] complex
 * per domain); total = 1 adds the incident E to A in the open region (mEMP_s / mEMP_t,
 * d3qd2_src/d3qd2_field.c:10-73).  Returns the domain id. */
int or_mEMP(const or_model *m, double complex **U, double complex **dU, const double *r, int type, int total,
            double complex *Uo);
/* Field at the point (zt, et) of element t (1-based position in domain did's list); total = 1
 * adds the incident field in the open region (EH_s_bd / EH_t_bd, d3qd2_src/d3qd2_field.c:299-403). */
void or_EH_bd(const or_model *m, double complex **E, double complex **dE, double complex **H,
              double complex **dH, int did, int t, double zt, double et, int type, int total,
              double complex *Eo, double complex *Ho);
/* Radiation force and torque about rc on the objects in the open region from the Maxwell
 * stress tensor (force_FN, d3qd2_src/d3qd2_force.c:10-240). */
int or_force_FN(const or_model *m, double complex **E, double complex **dE, double complex **H,
                double complex **dH, const double *rc, int type, double *F, double *N);

#endif
