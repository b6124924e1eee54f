/*
 * bq_assemble.h -- device-side domain assembly interface (internal, C ABI).
 */
#ifndef BQ_ASSEMBLE_H
#define BQ_ASSEMBLE_H
#include <stddef.h>
#include "bq_plan.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct bq_adom {
  int dNe;            /* elements of the domain */
  int periodic;       /* 1: quasi-periodic Green's function (open region, ps = 1) */
  int open_domain;    /* 1: diagonal term 1 + F (domain 0), else F */
  double kr, ki;      /* wave number of the domain (free-space kernel) */
  bq_qplan plan;      /* lattice data (periodic kernel) */
  const double *E;    /* device: element constants, 24 doubles per element id (1-based) */
  const double *ren;  /* device: element node coordinates [(Ne+1)][4][3] */
  const int *sid;     /* device: signed element ids of the domain [dNe+1] */
  double *G, *H;      /* device out: N x N complex */
  double *dG, *dH;    /* device out: 2N x N complex (free-space domains) */
  double *tdF;        /* device out: N x 3 (may be NULL) */
  int *n_gh_out;      /* host out: number of pairs that used the 64x64 rule (may be NULL) */
  long long *n_points_out; /* host out [3]: integrand points of the regular (6- to 12-point),
                              64-point and singular rules (may be NULL) */
  int full_rule;      /* 1: every regular periodic pair takes the 12-point rule */
} bq_adom;

size_t bq_assemble_work_bytes(int dNe);
int bq_assemble_domain_dev(const bq_adom *a, void *work, void *stream);
int bq_set_quadrature(const double *xl, const double *wl, const double *xh, const double *wh);

/* Gauss-Legendre nodes and weights of order n on [-1, 1] (Newton iteration on P_n). */
void bq_gauss_legendre(int n, double *x, double *w);

#ifdef __cplusplus
}
#endif
#endif
