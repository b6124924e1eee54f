/*
 * bq_linalg.h -- device dense complex linear algebra (internal, C ABI).
 * Matrices are column-major interleaved complex doubles (z) or floats (c) in device memory
 * unless the function name says _host.  Streams are passed as void* (a hipStream_t).
 */
#ifndef BQ_LINALG_H
#define BQ_LINALG_H
#ifdef __cplusplus
extern "C" {
#endif

/* LU with partial pivoting; dinv (device, bq_lu_inv_doubles(n) doubles) receives the
 * inverted diagonal blocks and the row permutation that bq_zgetrs_dev uses. */
long bq_lu_inv_doubles(int n);
int bq_zgetrf_dev(void *stream, int n, double *A, long lda, int *ipiv, double *dinv);
int bq_zgetrs_dev(void *stream, int n, int nrhs, const double *A, long lda, const int *ipiv, const double *dinv,
                  double *B, long ldb);
int bq_zgemm_dev(void *stream, int m, int n, int k, const double *alpha, const double *A, long lda, const double *B,
                 long ldb, const double *beta, double *C, long ldc);
int bq_zgemm_op_dev(void *stream, int ta, int tb, int m, int n, int k, const double *alpha, const double *A,
                    long lda, const double *B, long ldb, const double *beta, double *C, long ldc);
int bq_zgesv_host(int n, int nrhs, double *A, double *B);
/* B -= A X for a few columns (nrhs <= 8): A m x k read once; X and B must not overlap. */
int bq_zgemm_skinny_dev(void *stream, int m, int nrhs, long k, const double *A, long lda, const double *X, long ldx,
                        double *B, long ldb);

/* The same in complex FP32 (interleaved complex floats; dinv holds bq_lu_inv_doubles(n)
 * floats). */
int bq_cgetrf_dev(void *stream, int n, float *A, long lda, int *ipiv, float *dinv);
int bq_cgetrs_dev(void *stream, int n, int nrhs, const float *A, long lda, const int *ipiv, const float *dinv,
                  float *B, long ldb);
int bq_cgemm_op_dev(void *stream, int ta, int tb, int m, int n, int k, const float *alpha, const float *A,
                    long lda, const float *B, long ldb, const float *beta, float *C, long ldc);
int bq_cgesv_host(int n, int nrhs, float *A, float *B);
/* C -= A B in complex FP32 with bf16x3 products on the matrix cores (bq_mfma.hip); the
 * _store variant writes C = -A B and allows B == C (in place). */
int bq_cgemm_split_dev(void *stream, int m, int n, int k, const float *A, long lda, const float *B, long ldb,
                       float *C, long ldc);
int bq_cgemm_split_store_dev(void *stream, int m, int n, int k, const float *A, long lda, const float *B, long ldb,
                             float *C, long ldc);
/* 1: FP32 factorization updates on the bf16x3 kernel, 0: rocBLAS cgemm; < 0 queries.
 * Returns the previous mode. */
int bq_lu_update_mode(int mode);
/* 1: look-ahead of the factorizations (a second stream per stream), 0: off; < 0 queries.
 * Returns the previous mode. */
int bq_lu_lookahead(int mode);
/* Free the per-stream state of the solves (rocBLAS handle, split-GEMM workspace, look-ahead
 * stream) before the stream is destroyed. */
void bq_lu_release_stream(void *stream);
void bq_cgemm_split_release(void *stream);

/* Few-column solves (nrhs <= 8) through the inverted 512-row diagonal blocks of both factors:
 * W (device) holds bq_lu_blkinv_elems(n) complex elements, filled by bq_[zc]lu_blkinv_dev from
 * the factors and dinv; n must be a multiple of 512 (bq_lu_blkinv_elems returns 0 otherwise). */
long bq_lu_blkinv_elems(int n);
int bq_zlu_blkinv_dev(void *stream, int n, const double *A, long lda, const double *dinv, double *W);
int bq_clu_blkinv_dev(void *stream, int n, const float *A, long lda, const float *dinv, float *W);
int bq_zgetrs_blkinv_dev(void *stream, int n, int nrhs, const double *A, long lda, const double *dinv,
                         const double *W, double *B, long ldb);
int bq_cgetrs_blkinv_dev(void *stream, int n, int nrhs, const float *A, long lda, const float *dinv,
                         const float *W, float *B, long ldb);

/* The row permutation the factorization stored in dinv (perm[i] = original row placed at
 * row i; device int[n]); single != 0 for the FP32 buffer. */
const int *bq_lu_perm_dev(const void *dinv, int n, int single);
/* X (device, n x n, ldx) = U^-1 L^-1 from the factors: the inverse with its columns permuted,
 * A^-1[:, perm[i]] = X[:, i] (perm from bq_lu_perm_dev).  Its forward sweep skips the zeros of
 * L^-1, so it takes two thirds of the flops of a solve with n right-hand sides. */
int bq_zlu_inverse_dev(void *stream, int n, const double *Ad, long lda, const double *dinv, double *Xd, long ldx);
int bq_clu_inverse_dev(void *stream, int n, const float *Ad, long lda, const float *dinv, float *Xd, long ldx);
/* Solves whose right-hand sides come in column groups (host arrays): group g, columns
 * g0[g] .. g0[g]+gn[g]-1, is zero in permuted rows 0 .. gf[g]-1, and its forward solve starts
 * there.  The groups cover 0 .. nrhs-1. */
int bq_zgetrs_groups_dev(void *stream, int n, int nrhs, const double *A, long lda, const int *ipiv, const double *dinv,
                         double *B, long ldb, int ng, const int *g0, const int *gn, const int *gf);
int bq_cgetrs_groups_dev(void *stream, int n, int nrhs, const float *A, long lda, const int *ipiv, const float *dinv,
                         float *B, long ldb, int ng, const int *g0, const int *gn, const int *gf);

#ifdef __cplusplus
}
#endif
#endif
