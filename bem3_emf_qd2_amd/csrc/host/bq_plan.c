/*
 * bq_plan.c -- host set-up of the lattice constants used by the QPGF kernels.
 *
 * The Ewald parameter default and the Fourier/Ewald selection data follow the reference
 * (qgf2_src/d3_qpgf_d2.c:124-176 for the selection, :1680-1693 for the Ewald parameter).
 */
#include <math.h>
#include "../bq_plan.h"

int bq_qplan_init(bq_qplan *p, double k, const double *vk, const double *vd1, const double *vd2, double eps)
{
  double det = vd1[0] * vd2[1] - vd2[0] * vd1[1];
  if (det == 0.0 || !(k > 0.0)) return -1;
  p->k = k;
  p->k2 = k * k;
  p->err = 0;
  p->wtab = 0;
  p->wtab_nv = 0;
  p->wtab_lrho = 0.0;
  p->cut = BQ_EXP_CUT;
  p->rule_ltol = BQ_RULE_LTOL;
  p->vk0 = vk[0];
  p->vk1 = vk[1];
  p->d1x = vd1[0]; p->d1y = vd1[1];
  p->d2x = vd2[0]; p->d2y = vd2[1];
  p->idd = 1.0 / det;
  p->aidd = fabs(p->idd);
  double cpd = 2.0 * M_PI * p->idd;
  p->g1x = cpd * vd2[1];  p->g1y = -cpd * vd2[0];
  p->g2x = -cpd * vd1[1]; p->g2y = cpd * vd1[0];
  p->b1x = p->g1x / (2.0 * M_PI); p->b1y = p->g1y / (2.0 * M_PI);
  p->b2x = p->g2x / (2.0 * M_PI); p->b2y = p->g2y / (2.0 * M_PI);
  /* same expression as d3_qpgf_d2.c:1680-1682, so the value is bit-identical */
  double sp0 = vd1[0] + vd2[0], sp1 = vd1[1] + vd2[1], sm0 = vd1[0] - vd2[0], sm1 = vd1[1] - vd2[1];
  double adp = sqrt(sp0 * sp0 + sp1 * sp1);
  double adm = sqrt(sm0 * sm0 + sm1 * sm1);
  p->veps0 = sqrt(fabs(det) / (4.0 * M_PI) * adp / adm);
  p->eps = eps;
  p->ph2re = cos(vk[0] * vd2[0] + vk[1] * vd2[1]);
  p->ph2im = sin(vk[0] * vd2[0] + vk[1] * vd2[1]);
  p->cell = fabs(det);
  p->rcell = fabs(p->g1x * p->g2y - p->g1y * p->g2x);

  /* Fourier-domain sum is allowed only if the orders S_LIMIT away from the nearest one
   * are evanescent on both sides; the per-point test then checks their decay over |z|. */
  int n1 = (int)round((vk[0] * vd1[0] + vk[1] * vd1[1]) / (2.0 * M_PI));
  int n2 = (int)round((vk[0] * vd2[0] + vk[1] * vd2[1]) / (2.0 * M_PI));
#define AX(a, b) (-vk[0] + (double)(a) * cpd * vd2[1] - (double)(b) * cpd * vd1[1])
#define AY(a, b) (-vk[1] - (double)(a) * cpd * vd2[0] + (double)(b) * cpd * vd1[0])
  double bN2 = p->k2 - AX(n1, n2) * AX(n1, n2) - AY(n1, n2) * AY(n1, n2);
  double ad12 = vd1[0] * vd1[0] + vd1[1] * vd1[1], ad22 = vd2[0] * vd2[0] + vd2[1] * vd2[1];
  int p1, p2, m1, m2;
  if (ad22 < ad12) { p1 = n1 + BQ_S_LIMIT; p2 = n2; m1 = n1 - BQ_S_LIMIT; m2 = n2; }
  else { p1 = n1; p2 = n2 + BQ_S_LIMIT; m1 = n1; m2 = n2 - BQ_S_LIMIT; }
  double bNp = p->k2 - AX(p1, p2) * AX(p1, p2) - AY(p1, p2) * AY(p1, p2);
  double bNm = p->k2 - AX(m1, m2) * AX(m1, m2) - AY(m1, m2) * AY(m1, m2);
#undef AX
#undef AY
  p->fd_allowed = 0;
  p->fd_bN = p->fd_bNp = p->fd_bNm = 0.0;
  if (!(bNp > 0.0 || bNm > 0.0) && bN2 >= 0.0) {
    p->fd_allowed = 1;
    p->fd_bN = sqrt(bN2);
    p->fd_bNp = sqrt(-bNp);
    p->fd_bNm = sqrt(-bNm);
  }
  return 0;
}
