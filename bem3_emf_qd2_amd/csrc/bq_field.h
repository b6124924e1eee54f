/*
 * bq_field.h -- device field evaluation (internal, C ABI).
 */
#ifndef BQ_FIELD_H
#define BQ_FIELD_H
#include "bq_plan.h"
#ifdef __cplusplus
extern "C" {
#endif

#define BQ_MAXD 16 /* domains supported by the field context */

/* device-side view passed to the field kernel by value */
typedef struct bq_fctx {
  int ps, MN;
  int has_tri;                           /* any linear triangles */
  bq_qplan plan;
  const double *E, *ren, *wen, *kn;      /* device */
  int dNe[BQ_MAXD];
  const int *sid[BQ_MAXD];               /* device, per domain */
  const double *Eb[BQ_MAXD], *dEb[BQ_MAXD], *Hb[BQ_MAXD], *dHb[BQ_MAXD]; /* device, [(dNe+1)][4][3] complex */
  const double *Ub[BQ_MAXD], *dUb[BQ_MAXD]; /* device, [(dNe+1)][4][4] complex */
} bq_fctx;

/* host-side description of a solved model */
typedef struct bq_fdata {
  int Ne, MN, ps;
  double qd[8];
  const double *kn;                      /* [MN+1] complex */
  const double *E, *ren, *wen;           /* [(Ne+1)][24] (bq_pack_elem), [(Ne+1)][12], [(Ne+1)][12] */
  const int *dNe;
  const int *const *sid;
  const double *const *Eb, *const *dEb, *const *Hb, *const *dHb, *const *Ub, *const *dUb;
  const double *xli, *wli, *xhi, *whi, *z49, *e49, *w49;
  const double *w34, *z37, *e37, *w37;   /* triangle node-rule weights and 7-point rule */
  int Nn;                                /* mesh nodes, for the domain test */
  const double *rn;                      /* [(Nn+1)][3] node coordinates */
  const int *ed, *md, *sd;               /* [(Ne+1)][4] element nodes (4th 0 for a triangle), main and sub-domain */
} bq_fdata;

int bq_field_ctx_create(const bq_fdata *f, void **handle);
/* kind 0: out[n][13] = sum E (3 complex), sum H (3 complex), F; kind 1: out[n][9] = sum U
 * (4 complex), F; kind 2: out[n][36] = sum U, sum dU/dr_j (3 x 4 complex), F, dF/dr_j; for
 * folded points pts[n][3] with known domain ids, or, when bd_t is not NULL (kinds 0 and 1),
 * for the boundary points (bd_ze[2p], bd_ze[2p+1]) of elements bd_t[p] (1-based positions
 * in the domain's list; pts unused). */
int bq_field_eval(void *handle, int kind, int n, const double *pts, const int *did, const int *bd_t,
                  const double *bd_ze, int type, double *out);
void bq_field_ctx_destroy(void *handle);
/* domain id of n points (the reference's domain_id on folded points, d3qd2_src/d3qd2_setup.c:
 * 157-245: signed solid angles of the elements of each domain; -1 outside every domain) */
int bq_field_domain_ids(void *handle, int n, const double *pts, int *did);

int bq_field_eval_dev(const bq_fctx *C, int kind, int npts, const double *d_pts, const int *d_did, const int *d_bd_t,
                      const double *d_bd_ze, int type, double *d_out, void *stream);
int bq_field_rules_init(const double *z49, const double *e49, const double *w49, const double *w34,
                        const double *z37, const double *e37, const double *w37);

#ifdef __cplusplus
}
#endif
#endif
