/*
 * bq_qgf.c -- the reference's single-point QPGF and Faddeeva entry points
 * (qgf2_src/d3_qpgf_d2.h:39-126, qgf2_src/Faddeeva.h:39) as thin wrappers over the GPU
 * batch kernels.  They keep the reference's return convention (term count > 0, negative on
 * failure); the eps argument is accepted for compatibility, the kernels always work to
 * ~1e-15 relative accuracy.
 */
#include <stdio.h>
#include <complex.h>
#include "bem3_emf_qd2.h"
#include "bq_amd.h"

static int qpgf_one(int variant, double complex *qG, double complex *dqG, double *r, QPD2 *qd)
{
  double q8[8] = {qd->k, qd->vk[0], qd->vk[1], qd->vk[2], qd->vd1[0], qd->vd1[1], qd->vd2[0], qd->vd2[1]}, out[8];
  if (bq_qpgf_batch(q8, variant, 1, r, out)) return -30;
  if (qG) *qG = out[0] + I * out[1];
  if (dqG)
    for (int i = 0; i < 3; i++) dqG[i] = out[2 + 2 * i] + I * out[3 + 2 * i];
  return 1;
}

int d3hm_qpgf_d2_v1(double complex *qG, double complex *dqG, double *r, double eps, QPD2 *qd)
{
  (void)eps;
  return qpgf_one(0, qG, dqG, r, qd);
}

int d3hm_qpgf_d2_v2(double complex *qG, double complex *dqG, double complex *d2qG, double *r, double eps, QPD2 *qd)
{
  (void)eps;
  double q8[8] = {qd->k, qd->vk[0], qd->vk[1], qd->vk[2], qd->vd1[0], qd->vd1[1], qd->vd2[0], qd->vd2[1]}, out[20];
  if (bq_qpgf_batch(q8, 2, 1, r, out)) return -30;
  if (qG) *qG = out[0] + I * out[1];
  if (dqG)
    for (int i = 0; i < 3; i++) dqG[i] = out[2 + 2 * i] + I * out[3 + 2 * i];
  if (d2qG)
    for (int i = 0; i < 6; i++) d2qG[i] = out[8 + 2 * i] + I * out[9 + 2 * i];
  return 1;
}

int d3hm_qpgf_d2_qG(double complex *qG, double *r, double eps, QPD2 *qd)
{
  (void)eps;
  return qpgf_one(1, qG, NULL, r, qd);
}

int d3hm_qpgf_d2_dqG(double complex *dqG, double *r, double eps, QPD2 *qd)
{
  (void)eps;
  return qpgf_one(0, NULL, dqG, r, qd);
}

double complex Faddeeva_w(double complex z, double relerr)
{
  (void)relerr;
  double zz[2] = {creal(z), cimag(z)}, w[2];
  if (bq_faddeeva_w_batch(1, zz, w)) {
    printf("Faddeeva_w(): GPU evaluation failed: %s\n", bq_last_error());
    return NAN + I * NAN;
  }
  return w[0] + I * w[1];
}
