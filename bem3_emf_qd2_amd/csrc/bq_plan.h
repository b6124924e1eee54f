/*
 * bq_plan.h -- lattice/Bloch constants for the quasi-periodic Green's function, computed
 * once per (lattice, incidence) on the host and passed to kernels by value.
 *
 * Shared by the host C code and the HIP kernels, so it holds plain doubles and ints only.
 */
#ifndef BQ_PLAN_H
#define BQ_PLAN_H

#ifdef __cplusplus
extern "C" {
#endif

/* Default exponent cut for truncated lattice sums: terms below exp(-BQ_EXP_CUT) (4e-18) of
 * the leading term are dropped (the plan's cut; a tuned plan may take a smaller one). */
#define BQ_EXP_CUT 40.0
/* Default error bound of the distance-chosen Gauss rule relative to the integrand: log(1e-15) */
#define BQ_RULE_LTOL (-34.538776394910684)
#define BQ_EW_EXPM 3.0 /* Ewald amplification limit, reference qgf2_src/d3_qpgf_d2.h:26 */
#define BQ_S_LIMIT 15  /* reference qgf2_src/d3_qpgf_d2.h:19 */

typedef struct bq_qplan {
  double k, k2;            /* real wave number of the open region */
  double vk0, vk1;         /* in-plane Bloch wave vector */
  double d1x, d1y, d2x, d2y;
  double idd, aidd;        /* 1/det(d1,d2), |1/det| */
  double g1x, g1y, g2x, g2y; /* reciprocal vectors, g_i . d_j = 2 pi delta_ij */
  double b1x, b1y, b2x, b2y; /* dual vectors, b_i . d_j = delta_ij */
  double veps0;            /* lattice default Ewald parameter */
  double fd_bN, fd_bNp, fd_bNm; /* Fourier-domain selection data (bNp/bNm = 0 -> Ewald only) */
  int fd_allowed;
  double eps;              /* relative accuracy used in the Fourier/Ewald choice */
  double ph2re, ph2im;     /* exp(i vk . d2): Bloch phase step along d2 */
  double cell, rcell;      /* |det(d1,d2)| and |det(g1,g2)| (unit-cell areas) */
  double cut;              /* lattice-sum truncation: terms below exp(-cut) are dropped */
  double rule_ltol;        /* log of the error bound of the distance-chosen Gauss rule */
  /* Faddeeva-series weights of the wavefront-uniform spatial sums (device, NULL: none), for
   * the Ewald parameters v_j = veps0 exp(-j wtab_lrho), j < wtab_nv: rows of BQ_WTAB_ROW
   * doubles, [v_j, x_j = k v_j, exp(-x_j^2), term count, 2a/pi, 1 if a single Newton step
   * suffices for the terms' reciprocals else 0], then for n = 1 .. BQ_WTAB_N:
   * (a n)^2, exp(-(a n)^2 - x^2), exp(-(a n + x)^2), a n exp(-(a n + x)^2), exp(-(a n - x)^2),
   * a n exp(-(a n - x)^2), a = pi / sqrt(-ln(relerr / 2)) the series parameter of Algorithm 916 */
  const double *wtab;
  int wtab_nv;
  double wtab_lrho;
  int *err;                /* device word (may be NULL): a kernel stores the reference's
                              negative QPGF code there (atomic minimum) when a reciprocal
                              order has beta_n^2 = k^2 - |k_t + G|^2 == 0 exactly (a Wood
                              anomaly): -12 in the Fourier-domain sum, -23 in the Ewald
                              spectral sum (qgf2_src/d3_qpgf_d2.h:113-123) */
} bq_qplan;

#define BQ_WTAB_N 60
#define BQ_WTAB_ROW (6 * (BQ_WTAB_N + 1)) /* header, then BQ_WTAB_N terms */
#define BQ_WTAB_MAXV 128

/* Relative cost of one lattice term of each kind, for choosing the Fourier-domain or the
 * Ewald sum per point (both converge to the same value): Fourier-domain term 1, Ewald
 * spectral term 2.1, Ewald spatial term (one Faddeeva evaluation) 6.1, as timed on MI355X
 * for the QPGF kernel on 2^20-point batches. */
#ifndef BQ_COST_FD
#define BQ_COST_FD 1.0
#endif
#ifndef BQ_COST_SPEC
#define BQ_COST_SPEC 2.1
#endif
#ifndef BQ_COST_SPAT
#define BQ_COST_SPAT 6.1
#endif
/* Weight of the Fourier-domain term in the wavefront-uniform form, where a wavefront takes the
 * Fourier-domain sum only when every lane's own choice is it.  Timed on the 6144-quad sphere's
 * assembly (k_q0_gl per point): 0.4 3.36 s, 0.55 3.25 s, 0.7 3.22 s, 1.0 3.24 s, 1.4 3.38 s. */
#ifndef BQ_COST_FD_WAVE
#define BQ_COST_FD_WAVE 0.7
#endif

/* Fill a plan from the reference's QPD2 fields (k, vk[3], vd1[2], vd2[2]).
 * Returns 0, or -1 if the lattice is degenerate. */
int bq_qplan_init(bq_qplan *p, double k, const double *vk, const double *vd1, const double *vd2, double eps);

#ifdef __cplusplus
}
#endif
#endif
