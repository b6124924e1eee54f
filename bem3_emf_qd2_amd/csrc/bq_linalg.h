/*
 * bq_linalg.h -- device dense complex linear algebra (internal, C ABI).
 * Matrices are column-major interleaved complex doubles in device memory unless the
 * function name says _host.  Streams are passed as void* (a hipStream_t).
 */
#ifndef BQ_LINALG_H
#define BQ_LINALG_H
#ifdef __cplusplus
extern "C" {
#endif

/* LU with partial pivoting; dinv (device, bq_lu_inv_doubles(n) doubles) receives the
 * inverted diagonal blocks that bq_zgetrs_dev uses. */
long bq_lu_inv_doubles(int n);
int bq_zgetrf_dev(void *stream, int n, double *A, long lda, int *ipiv, double *dinv);
int bq_zgetrs_dev(void *stream, int n, int nrhs, const double *A, long lda, const int *ipiv, const double *dinv,
                  double *B, long ldb);
int bq_zgemm_dev(void *stream, int m, int n, int k, const double *alpha, const double *A, long lda, const double *B,
                 long ldb, const double *beta, double *C, long ldc);
int bq_zgemm_op_dev(void *stream, int ta, int tb, int m, int n, int k, const double *alpha, const double *A,
                    long lda, const double *B, long ldb, const double *beta, double *C, long ldc);
int bq_zgesv_host(int n, int nrhs, double *A, double *B);

#ifdef __cplusplus
}
#endif
#endif
