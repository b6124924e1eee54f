/*
 * bem3_emf_qd2.h -- drop-in C API of libbem3qd2 (MI355X implementation of the
 * bem3_emf_qd2 boundary-integral path).
 *
 * Every prototype below replaces the reference function of the same name; the data types
 * keep the reference's field order and sizes because callers touch fields directly
 * (example2.c uses qd.pw.lambda_0) and ex.dat stores MIPW, Ipw and QPD2 as raw bytes.
 *   types        d3qd2_src/d3qd2_elem.h:13-60, d3qd2_src/m_ipw.h:14-27, d3qd2_src/ipw.h:19-37,
 *                qgf2_src/d3_qpgf_d2.h:32-36
 *   API          d3qd2_src/bem3_emf_qd2.h:17-76
 *   helpers      com_src/my_utils.h:9-36 (m_alloc2 and the vector helpers examples use)
 * Error behaviour follows the reference: input and set-up errors print a message and exit;
 * the field functions return the domain id (negative near a boundary, -1 if no domain).
 * The compute-heavy calls (solve_bieq, field evaluation, QPGF) run on the GPU; they exit
 * with a message when no gfx950 device is usable -- there is no CPU fallback.
 */
#ifndef BEM3_EMF_QD2_H_
#define BEM3_EMF_QD2_H_

#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <math.h>
#include <complex.h>

#define PREC_DEF_BE 2 /* element-pair rule used by solve_bieq: 12x12 Gauss-Legendre */
#define PREC_DEF_FN 0 /* rule used by force_FN's field calls */

#define GLN 12
#define GHN 64
#define OPENDID 99

/* ---- quasi-periodic Green's function data (qgf2_src/d3_qpgf_d2.h:32-36) ---- */
typedef struct qp_data_v2 {
  double k;             /* wave number */
  double vk[3];         /* wave vector */
  double vd1[2], vd2[2]; /* lattice vectors */
} QPD2;

/* ---- incident plane wave (d3qd2_src/ipw.h:19-37) ---- */
typedef struct incident_planewave_data {
  double E0;
  double ki;
  double cos_t, sin_t;
  double cos_p, sin_p;
  double complex ex, ey, ez;
  double complex hx, hy, hz;
} IpwD;

typedef struct incident_planewave {
  double lambda0;
  double ni;
  double power;
  double complex e0x;
  double complex e0y;
  double fx, fy, fz;
  double theta;
  double phi;
  IpwD data;
} Ipw;

/* ---- multiple plane waves (d3qd2_src/m_ipw.h:14-27) ---- */
typedef struct beam_object {
  Ipw *ipw;
} BOBJ;

typedef struct ipw_data {
  char fn_ipw[64];
  int n_ipw;
  double n_0;
  double lambda_0;
  double omega;
  BOBJ bd;
} MIPW;

/* ---- model (d3qd2_src/d3qd2_elem.h:13-60) ---- */
typedef struct sub_domain_data {
  int Ne;
  int *sid;
  double complex ***U, ***dU;
  double complex ***E, ***dE;
  double complex ***H, ***dH;
} SUBD;

typedef struct boundary_data {
  int ps;
  QPD2 qd;
  int Nn;
  int Ne;
  double **rn;
  int **ed;
  int NN;
  int **eni;
  int *md;
  int *sd;
  int *gd;
  double ***cr, ***cw;
  double ***ren, ***wen;
  double complex ***Ui, ***dUi;
  double zt_44[4], et_44[4], wt_44[4], zt_49[9], et_49[9], wt_49[9];
  double zt_34[4], et_34[4], wt_34[4], zt_37[7], et_37[7], wt_37[7];
  double xli[GLN], wli[GLN];
  double xhi[GHN], whi[GHN];
  SUBD *sb;
} BOUD;

typedef struct domain_data {
  char ped_fn[128], med_fn[128], msh_fn[128];
  double rv[3], th;
  double tv[3];
  int MN;
  double complex *n;
  double complex *kn;
  MIPW pw;
  BOUD bd;
} DQD2;

/* -- d3qd2_setup.c (reference d3qd2_src/bem3_emf_qd2.h:17-27) -- */
void read_dqd2(int argc, char **argv, DQD2 *dq1);
void print_dqd2(DQD2 *qd);
void print_dqd2_mksa(DQD2 *qd);
void initialize_dqd2(DQD2 *qd);
void finalize_dqd2(DQD2 *qd);
int domain_id(double *rt, DQD2 *qd);
int q0_domain_id(double *rt, DQD2 *qd);
void q0_domain_id_l(int *l1, int *l2, double *rt, DQD2 *qd);
void dat_write(char *filename, DQD2 *qd);
void dat_read(char *filename, DQD2 *qd);
void output_node_particles(char *fname, DQD2 *qd);

/* -- d3qd2_solve_bieq.c (reference :31) -- */
void solve_bieq(DQD2 *qd);

/* -- d3qd2_field.c (reference :36-69); type 0: 4-point, 1: 9-point, 2: 12x12, 3: 64x64 -- */
int mEMP_s(double complex *U, double *rt, int type, DQD2 *qd);
int mEMP_t(double complex *U, double *rt, int type, DQD2 *qd);
int mEMP_i(double complex *U, double *rt, int type, DQD2 *qd);
int EH_mEMP_s(double complex *E, double complex *H, double *rt, int type, DQD2 *qd);
int EH_mEMP_t(double complex *E, double complex *H, double *rt, int type, DQD2 *qd);
int EH_mEMP_i(double complex *E, double complex *H, double *rt, int type, DQD2 *qd);
int EH_s(double complex *E, double complex *H, double *rt, int type, DQD2 *qd);
int EH_t(double complex *E, double complex *H, double *rt, int type, DQD2 *qd);
int EH_i(double complex *E, double complex *H, double *rt, int type, DQD2 *qd);
void EH_s_bd(double complex *E, double complex *H, int did, int t, double zeta_t, double eta_t, int type, DQD2 *qd);
void EH_t_bd(double complex *E, double complex *H, int did, int t, double zeta_t, double eta_t, int type, DQD2 *qd);
void EH_i_bd(double complex *E, double complex *H, int did, int t, double zeta_t, double eta_t, int type, DQD2 *qd);

/* -- d3qd2_force.c (reference bem3_emf_qd2.h:75-80): force F and torque N about rc on the
 * objects from the Maxwell stress on domain 0's boundary; type 0: 4-point, 1: 9-/7-point
 * rule; returns -1 if domain 0 has no element -- */
int force_FN(double *F, double *N, double *rc, int type, DQD2 *qd);

/* -- incident field helpers (d3qd2_src/m_ipw.h:30-38) -- */
void read_data_mipw(char *fname, MIPW *obj);
void print_data_mipw(MIPW *obj);
void print_data_mipw_mksa(MIPW *obj);
void check_param_mipw(MIPW *obj);
void setup_mipw(MIPW *obj);
void free_mipw(MIPW *obj);
void calc_mipw_EH(double complex *e, double complex *h, double *x, MIPW *obj);
void calc_mipw_EH_dEHdv(double complex *e, double complex *h, double complex *dedv, double complex *dhdv, double *x,
                        double *v, MIPW *obj);

/* -- utilities (com_src/my_utils.h:9-36) -- */
void vadd_d(double *r, double *v0, double *v1);
void vsub_d(double *r, double *v0, double *v1);
int vuni_d(double *r);
double vabs_d(double *v);
double vabs_2dm(double *v0, double *v1);
double vdot_d(double *v0, double *v1);
void vcrs_d(double *r, double *v0, double *v1);
void *m_alloc2(size_t num, size_t size, char *txt);

/* -- quasi-periodic Green's function (qgf2_src/d3_qpgf_d2.h:101-126), one point each -- */
int d3hm_qpgf_d2_v1(double complex *qG, double complex *dqG, double *r, double eps, QPD2 *qd);
/* d2qG[6] = xx, yy, zz, xy, yz, zx second derivatives (qgf2_src/d3_qpgf_d2.h:124) */
int d3hm_qpgf_d2_v2(double complex *qG, double complex *dqG, double complex *d2qG, double *r, double eps, QPD2 *qd);
int d3hm_qpgf_d2_qG(double complex *qG, double *r, double eps, QPD2 *qd);
int d3hm_qpgf_d2_dqG(double complex *dqG, double *r, double eps, QPD2 *qd);

/* -- Faddeeva w (qgf2_src/Faddeeva.h:39) -- */
double complex Faddeeva_w(double complex z, double relerr);

#endif /* BEM3_EMF_QD2_H_ */
