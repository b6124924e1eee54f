/*
 * bq_amd.h -- entry points of libbem3qd2 beyond the reference's own API.
 *
 * Plain C ABI: pointers and sizes, complex numbers as interleaved (re, im) doubles.
 * Every function returns 0 on success and a negative code on failure; bq_last_error()
 * then describes the failure.  None of these has a CPU fallback: without a usable gfx950
 * device they return -1.
 */
#ifndef BQ_AMD_H
#define BQ_AMD_H

#ifdef __cplusplus
extern "C" {
#endif

const char *bq_last_error(void);

/* 0 if a gfx950 device is usable by this process. */
int bq_device_ok(void);

/* w(z) for n points; z and w are [n][2] (re, im).  Batched form of the reference's
 * Faddeeva_w (qgf2_src/Faddeeva.h:39, Faddeeva.cc:679), relative accuracy ~1e-15. */
int bq_faddeeva_w_batch(int n, const double *z, double *w);

/* Quasi-periodic Green's function value and gradient for n relative positions r[n][3].
 * qd8 = the reference's QPD2 {k, vk[3], vd1[2], vd2[2]} as 8 doubles.
 * out[n][8] = qG, dqG/dx, dqG/dy, dqG/dz (complex).  Batched form of d3hm_qpgf_d2_v1
 * (qgf2_src/d3_qpgf_d2.h:101-112, d3_qpgf_d2.c:124). */
int bq_qpgf_v1_batch(const double *qd8, int n, const double *r, double *out);

#ifdef __cplusplus
}
#endif
#endif
