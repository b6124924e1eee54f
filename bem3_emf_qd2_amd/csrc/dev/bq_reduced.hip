// bq_reduced.hip -- global solve of a single object (one bounded domain inside the open
// region) by reduction to three dense systems of order NN, with the G0 and G1 factors that
// the dE/dH step of the reference (solve_deh_bv, d3qd2_src/d3qd2_solve_bieq.c:662-791,
// LAPACKE_zgesv at :753) needs anyway.
//
// Unknowns X_0..X_3 = (A_x, A_y, A_z, phi) on the open-region side and X_4..X_7 their normal
// derivatives, per element node (global order, n = NN).  The reference's rows
// (create_matrix_csr_dac, d3qd2_src/d3qd2_solve_bieq.c:284-421), with G0, H0 the open-region
// matrices and G1, H1 the object's, k1 = k(open), k2 = k(object), D = k1^2 - k2^2,
// b = k1^2 / k2^2, s = 1 - b, N_c = diag(n_c) (unit normal components at the nodes):
//   open region   H0 X_c - G0 X_{c+4} = 0                                   (c = 0..3)
//   object, c<3   H1 X_c + G1 X_{c+4} + D G1 N_c X_3 = -H1 Ui_c - G1 dUi_c
//   object, c=3   s G1 sum_c N_c X_c + H1 X_3 + b G1 X_7 = -s G1 (n . Ui)
// With M = G0^-1 H0 (X_{c+4} = M X_c) and P = G1^-1 H1 the object rows become
//   (P + M) X_c + D N_c X_3 = g_c,    s sum_c N_c X_c + (P + b M) X_3 = g_3,
//   g_c = -P Ui_c - dUi_c,  g_3 = -s (n . Ui),
// and with B = P + M, X_c = B^-1 g_c - D B^-1 N_c X_3, the Schur system for X_3 is
//   S X_3 = g_3 - s sum_c n_c o (B^-1 g_c),   S = P + b M - s D (B^-1 o n n^T),
// since sum_c N_c B^-1 N_c = B^-1 o (n_i . n_j).  Work: LUs of G0, G1, B and S (4 x 8/3 n^3)
// and M, P, B^-1 (3 x 8 n^3) against 2 x 8/3 (2n)^3 + 8 (2n)^2 n for the block elimination of
// order 2 NN; the G0, G1 factors also give dE/dH.
//
// The same elimination solves the system for any right-hand side (r0 on the open-region rows,
// r1 on the object rows): u0 = G0^-1 r0, u1 = G1^-1 r1, g_c = u1_c + u0_c (c < 3),
// g_3 = u1_3 + b u0_3, then as above, and X_{c+4} = M X_c - u0_c.  In FP64 that is the
// direct solve, followed by up to two steps of iterative refinement as the reference's PARDISO
// call takes (d3qd2_src/d3qd2_mkl_solver.c:53-58).  With complex FP32 factors it is the
// preconditioner of flexible GMRES (right preconditioning, classical Gram-Schmidt with one
// re-orthogonalisation) on the unreduced system in FP64: residuals b - A x are formed from the
// FP64 matrices G0, H0, G1, H1 as assembled (Carson and Higham, SIAM J. Sci. Comput. 40(2),
// 2018, GMRES-IR).  The stopping test is the normwise backward error of the unreduced system,
//   eta = max_i |b - A x|_i / (||A||_inf ||x||_inf + ||b||_inf),
// with ||A||_inf from exact row sums of |A|.
#include <cmath>
#include <cfloat>
#include <cstring>
#include <algorithm>
#include <string>
#include <thread>
#include "bq_runtime.hpp"
#include "bq_complex.hpp"
#include "bq_sess.hpp"
#include "../bq_linalg.h"

extern "C" const char* bq_last_error(void);

namespace {

__device__ __forceinline__ cplx ldz(const double2* p, long i) { return mk(p[i].x, p[i].y); }
__device__ __forceinline__ void stz(double2* p, long i, cplx v) { p[i] = make_double2(v.re, v.im); }

template <class F2>
__device__ __forceinline__ F2 mkv(double x, double y);
template <>
__device__ __forceinline__ double2 mkv<double2>(double x, double y) { return make_double2(x, y); }
template <>
__device__ __forceinline__ float2 mkv<float2>(double x, double y) { return make_float2((float)x, (float)y); }
template <class F2>
__device__ __forceinline__ cplx ldv(const F2* p, long i) { return mk((double)p[i].x, (double)p[i].y); }
template <class F2>
__device__ __forceinline__ void stv(F2* p, long i, cplx v) { p[i] = mkv<F2>(v.re, v.im); }

inline unsigned nblk(long n, int b) { return (unsigned)((n + b - 1) / b); }

// out (n x n column-major, global node order) = A (N x N row-major, domain-local rows and
// columns) through ginv (global node -> local index); 32 x 32 tiles through LDS so that both
// the row-major reads and the column-major writes are contiguous.
template <class F2>
__global__ void __launch_bounds__(256) k_to_global(int n, long N, const int* __restrict__ ginv,
                                                   const double2* __restrict__ A, F2* __restrict__ out)
{
  __shared__ double2 tile[32][33];
  const int gr0 = blockIdx.x * 32, gc0 = blockIdx.y * 32;
  const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
  const int gc = gc0 + tx;
  const long c = gc < n ? ginv[gc] : 0;
  for (int i = ty; i < 32; i += 8) {
    const int gr = gr0 + i;
    double2 v = make_double2(0.0, 0.0);
    if (gr < n && gc < n) v = A[(long)ginv[gr] * N + c];
    tile[i][tx] = v;
  }
  __syncthreads();
  for (int j = ty; j < 32; j += 8) {
    const int gcc = gc0 + j, grr = gr0 + tx;
    if (gcc < n && grr < n) {
      const double2 v = tile[tx][j];
      out[(long)gcc * n + grr] = mkv<F2>(v.x, v.y);
    }
  }
}

template <class F2>
__global__ void k_add(long len, const F2* __restrict__ a, const F2* __restrict__ b, F2* __restrict__ c)
{
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < len) c[i] = mkv<F2>((double)a[i].x + (double)b[i].x, (double)a[i].y + (double)b[i].y);
}

// S = P + b M - sD (Binv o n n^T), written over S's buffer (which may hold the LU of B), from
// X = U^-1 L^-1 of B's factors: Binv[:, perm[c]] = X[:, c].  Binv in natural order replaces P
// (each element of P is read, then overwritten, by the same thread).
template <class F2>
__global__ void k_schur(int n, F2* __restrict__ P, const F2* __restrict__ M, const F2* __restrict__ X,
                        const int* __restrict__ perm, const double* __restrict__ nn3, cplx beps, cplx sd,
                        F2* __restrict__ S)
{
  const long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (long)n * n) return;
  const long c = idx / n, i = idx - c * n, j = perm[c];
  const long ij = j * n + i;
  const double g = nn3[3 * i] * nn3[3 * j] + nn3[3 * i + 1] * nn3[3 * j + 1] + nn3[3 * i + 2] * nn3[3 * j + 2];
  const cplx x = ldv(X, idx);
  const cplx v = ldv(P, ij) + beps * ldv(M, ij) - (sd * x) * g;
  stv(S, ij, v);
  stv(P, ij, x);
}

template <class F2>
__global__ void k_cvt(long len, const double2* __restrict__ a, F2* __restrict__ b)
{
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < len) b[i] = mkv<F2>(a[i].x, a[i].y);
}

template <class F2>
__global__ void k_cvt_back(long len, const F2* __restrict__ a, double2* __restrict__ b)
{
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < len) b[i] = make_double2((double)a[i].x, (double)a[i].y);
}

// g_c = u1_c + u0_c (c < 3), g_3 = u1_3 + b u0_3, over u1's buffer
template <class F2>
__global__ void k_gcomb(int n, const F2* __restrict__ u0, F2* __restrict__ u1, cplx beps)
{
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= 4L * n) return;
  const cplx a = ldv(u0, i), c = ldv(u1, i);
  stv(u1, i, i < 3L * n ? c + a : c + beps * a);
}

// Y[3n + e] = g3[e] - s sum_c n_c(e) Y[c n + e]
template <class F2>
__global__ void k_rhs3(int n, const F2* __restrict__ g3, const double* __restrict__ nn3, cplx sigm, F2* __restrict__ Y)
{
  const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n) return;
  cplx a = mk(0.0, 0.0);
  for (int c = 0; c < 3; ++c) a += ldv(Y, (long)c * n + e) * nn3[3 * e + c];
  stv(Y, 3L * n + e, ldv(g3, e) - sigm * a);
}

// T[c n + e] = n_c(e) X3[e]
template <class F2>
__global__ void k_tnorm(int n, const F2* __restrict__ X3, const double* __restrict__ nn3, F2* __restrict__ T)
{
  const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n) return;
  const cplx x = ldv(X3, e);
  for (int c = 0; c < 3; ++c) stv(T, (long)c * n + e, x * nn3[3 * e + c]);
}

// Domain-local operand vectors loc = [a_0..a_3, b_0..b_3] (length N each) of the row products
// y_q = H a_q + G b_q of domain d (0: open region, 1: object).  mode 0: the unreduced
// operator applied to X (8 vectors [q][n], global order); mode 1: the right-hand side from
// the incident field (a_q = -Ui_q, b_q = -dUi_q, a_3 = 0, b_3 = -s n . Ui).
__global__ void k_op_prep(int n, int d, long N, int mode, const int* __restrict__ gidx, const double2* __restrict__ X,
                          const double* __restrict__ nn3, const double2* __restrict__ Uig,
                          const double2* __restrict__ dUig, cplx delt, cplx sigm, cplx beps,
                          double2* __restrict__ loc)
{
  const long c = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= N) return;
  const int e = gidx[c];
  cplx a[4], b[4];
  if (e < 0) {
    for (int q = 0; q < 4; ++q) a[q] = b[q] = mk(0.0, 0.0);
  } else if (mode == 1) {
    cplx nu = mk(0.0, 0.0);
    for (int q = 0; q < 3; ++q) {
      const cplx ui = ldz(Uig, (long)q * n + e);
      a[q] = -ui;
      b[q] = -ldz(dUig, (long)q * n + e);
      nu += ui * nn3[3L * e + q];
    }
    a[3] = mk(0.0, 0.0);
    b[3] = -(sigm * nu);
  } else if (d == 0) {
    for (int q = 0; q < 4; ++q) {
      a[q] = ldz(X, (long)q * n + e);
      b[q] = -ldz(X, (long)(q + 4) * n + e);
    }
  } else {
    const cplx x3 = ldz(X, 3L * n + e);
    cplx nx = mk(0.0, 0.0);
    for (int q = 0; q < 3; ++q) {
      const cplx xq = ldz(X, (long)q * n + e);
      a[q] = xq;
      b[q] = ldz(X, (long)(q + 4) * n + e) + (delt * x3) * nn3[3L * e + q];
      nx += xq * nn3[3L * e + q];
    }
    a[3] = x3;
    b[3] = sigm * nx + beps * ldz(X, 7L * n + e);
  }
  for (int q = 0; q < 4; ++q) {
    stz(loc, (long)q * N + c, a[q]);
    stz(loc, (long)(4 + q) * N + c, b[q]);
  }
}

// yl[q N + r] = sum_c H[r, c] a_q[c] + G[r, c] b_q[c] (q < 4) for kMvRows rows per
// workgroup of the row-major N x N matrices: each matrix entry is read once.
constexpr int kMvRows = 4;
__global__ void __launch_bounds__(256) k_rm_mv(long N, const double2* __restrict__ H, const double2* __restrict__ G,
                                              const double2* __restrict__ loc, double2* __restrict__ yl)
{
  __shared__ double red[4][kMvRows * 8];
  const long r0 = (long)blockIdx.x * kMvRows;
  const int nr = (int)min((long)kMvRows, N - r0);
  double acc[kMvRows * 8];
#pragma unroll
  for (int k = 0; k < kMvRows * 8; ++k) acc[k] = 0.0;
  for (long c = threadIdx.x; c < N; c += 256) {
    double2 a[4], b[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      a[q] = loc[(long)q * N + c];
      b[q] = loc[(long)(4 + q) * N + c];
    }
#pragma unroll
    for (int i = 0; i < kMvRows; ++i) {
      if (i < nr) {
        const double2 h = H[(r0 + i) * N + c], g = G[(r0 + i) * N + c];
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          acc[i * 8 + 2 * q] += h.x * a[q].x - h.y * a[q].y + g.x * b[q].x - g.y * b[q].y;
          acc[i * 8 + 2 * q + 1] += h.x * a[q].y + h.y * a[q].x + g.x * b[q].y + g.y * b[q].x;
        }
      }
    }
  }
#pragma unroll
  for (int k = 0; k < kMvRows * 8; ++k) {
    double v = acc[k];
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
    acc[k] = v;
  }
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  if (lane == 0)
#pragma unroll
    for (int k = 0; k < kMvRows * 8; ++k) red[wv][k] = acc[k];
  __syncthreads();
  if (threadIdx.x < kMvRows * 8) {
    const int k = threadIdx.x, i = k / 8, q = (k % 8) / 2, part = k & 1;
    if (i < nr) {
      const double v = red[0][k] + red[1][k] + red[2][k] + red[3][k];
      double* o = (double*)(yl + (long)q * N + r0 + i);
      o[part] = v;
    }
  }
}

// out[(q0 + q) n + gidx[r]] = sub - yl (sub != null) or yl, for live rows r
__global__ void k_scatter_rows(int n, long N, const int* __restrict__ gidx, const double2* __restrict__ yl, int q0,
                               const double2* __restrict__ sub, double2* __restrict__ out)
{
  const long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= 4 * N) return;
  const int q = (int)(idx / N);
  const long r = idx - q * N;
  const int e = gidx[r];
  if (e < 0) return;
  const long o = (long)(q0 + q) * n + e;
  const cplx y = ldz(yl, (long)q * N + r);
  stz(out, o, sub ? ldz(sub, o) - y : y);
}

// Row sums of |A| of the unreduced operator for the rows of domain d, into rn[(4 d + q) n + e]:
// open region sum|H| + sum|G|; object rows q < 3: sum|H| + sum|G| (1 + |D| |n_q|), q = 3:
// sum|H| + sum|G| (|s| sum_c |n_c| + |b|)  (nn3 in domain-local column order: nloc)
__global__ void __launch_bounds__(256) k_rownorm(long N, int n, int d, const int* __restrict__ gidx,
                                                const double2* __restrict__ H, const double2* __restrict__ G,
                                                const double* __restrict__ nloc, double ad, double as_, double ab,
                                                double* __restrict__ rn)
{
  __shared__ double red[4][5];
  const long r = blockIdx.x;
  double v[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
  for (long c = threadIdx.x; c < N; c += 256) {
    const double2 h = H[r * N + c], g = G[r * N + c];
    const double ah = sqrt(h.x * h.x + h.y * h.y), ag = sqrt(g.x * g.x + g.y * g.y);
    v[0] += ah;
    v[1] += ag;
    v[2] += ag * fabs(nloc[3 * c]);
    v[3] += ag * fabs(nloc[3 * c + 1]);
    v[4] += ag * fabs(nloc[3 * c + 2]);
  }
  for (int k = 0; k < 5; ++k) {
    double x = v[k];
    for (int o = 32; o > 0; o >>= 1) x += __shfl_xor(x, o);
    v[k] = x;
  }
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  if (lane == 0)
    for (int k = 0; k < 5; ++k) red[wv][k] = v[k];
  __syncthreads();
  if (threadIdx.x != 0) return;
  const int e = gidx[r];
  if (e < 0) return;
  double s[5];
  for (int k = 0; k < 5; ++k) s[k] = red[0][k] + red[1][k] + red[2][k] + red[3][k];
  for (int q = 0; q < 4; ++q) {
    double t;
    if (d == 0) t = s[0] + s[1];
    else if (q < 3) t = s[0] + s[1] + ad * s[2 + q];
    else t = s[0] + as_ * (s[2] + s[3] + s[4]) + ab * s[1];
    rn[(long)(4 * d + q) * n + e] = t;
  }
}

// normals in domain-local column order: nloc[3 c + k] = nn3[3 gidx[c] + k] (0 for no node)
__global__ void k_nloc(long N, const int* __restrict__ gidx, const double* __restrict__ nn3, double* __restrict__ nloc)
{
  const long c = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= N) return;
  const int e = gidx[c];
  for (int k = 0; k < 3; ++k) nloc[3 * c + k] = e < 0 ? 0.0 : nn3[3L * e + k];
}

// part[v nb + b] = sum over block b's share of conj(V_v) . w  (grid nb x nv)
__global__ void __launch_bounds__(256) k_mdot(long len, const double2* __restrict__ V, long ldv,
                                             const double2* __restrict__ w, double2* __restrict__ part)
{
  __shared__ double red[2][4];
  const int v = blockIdx.y;
  const double2* a = V + (long)v * ldv;
  double re = 0.0, im = 0.0;
  for (long i = (long)blockIdx.x * 256 + threadIdx.x; i < len; i += (long)gridDim.x * 256) {
    const double2 x = a[i], y = w[i];
    re += x.x * y.x + x.y * y.y;
    im += x.x * y.y - x.y * y.x;
  }
  for (int o = 32; o > 0; o >>= 1) {
    re += __shfl_xor(re, o);
    im += __shfl_xor(im, o);
  }
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  if (lane == 0) {
    red[0][wv] = re;
    red[1][wv] = im;
  }
  __syncthreads();
  if (threadIdx.x == 0)
    part[(long)v * gridDim.x + blockIdx.x] =
        make_double2(red[0][0] + red[0][1] + red[0][2] + red[0][3], red[1][0] + red[1][1] + red[1][2] + red[1][3]);
}

// per block: max_i |w_i| over its share, and of a second vector (or 0)
__global__ void __launch_bounds__(256) k_vmax(long len, const double2* __restrict__ a, const double2* __restrict__ b,
                                             double* __restrict__ part)
{
  __shared__ double red[2][4];
  double ma = 0.0, mb = 0.0;
  for (long i = (long)blockIdx.x * 256 + threadIdx.x; i < len; i += (long)gridDim.x * 256) {
    ma = fmax(ma, sqrt(a[i].x * a[i].x + a[i].y * a[i].y));
    if (b) mb = fmax(mb, sqrt(b[i].x * b[i].x + b[i].y * b[i].y));
  }
  for (int o = 32; o > 0; o >>= 1) {
    ma = fmax(ma, __shfl_xor(ma, o));
    mb = fmax(mb, __shfl_xor(mb, o));
  }
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  if (lane == 0) {
    red[0][wv] = ma;
    red[1][wv] = mb;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    part[2 * blockIdx.x] = fmax(fmax(red[0][0], red[0][1]), fmax(red[0][2], red[0][3]));
    part[2 * blockIdx.x + 1] = fmax(fmax(red[1][0], red[1][1]), fmax(red[1][2], red[1][3]));
  }
}

// w -= sum_v h_v V_v  (h on the device, nv entries)
__global__ void k_vupdate(long len, int nv, const double2* __restrict__ V, long ldv, const double2* __restrict__ h,
                          double2* __restrict__ w)
{
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= len) return;
  cplx a = ldz(w, i);
  for (int v = 0; v < nv; ++v) a -= ldz(h, v) * ldz(V, (long)v * ldv + i);
  stz(w, i, a);
}

__global__ void k_vscale(long len, const double2* __restrict__ a, double s, double2* __restrict__ b)
{
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < len) b[i] = make_double2(a[i].x * s, a[i].y * s);
}

__global__ void k_vaxpy(long len, const double2* __restrict__ a, double2* __restrict__ b)
{
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < len) b[i] = make_double2(b[i].x + a[i].x, b[i].y + a[i].y);
}

// global <-> domain-local columns of up to 6 vectors: to_glob 0: out[c n + gidx[r]] = in[c N + r];
// to_glob 1 (back): out[c N + r] = in[c n + gidx[r]] (0 where there is no node)
__global__ void k_perm6(int n, long N, int ncol, const int* __restrict__ gidx, const double2* __restrict__ in,
                        double2* __restrict__ out, int back)
{
  const long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= N * ncol) return;
  const long c = idx / N, r = idx - c * N;
  const int e = gidx[r];
  if (back) out[c * N + r] = e < 0 ? make_double2(0.0, 0.0) : in[c * n + e];
  else if (e >= 0) out[c * n + e] = in[c * N + r];
}

// ---- precision dispatch of the dense kernels --------------------------------------------
template <class F>
struct LA;
template <>
struct LA<double> {
  using V = double2;
  static int getrf(hipStream_t s, int n, void* A, int* ip, void* inv)
  {
    return bq_zgetrf_dev(s, n, (double*)A, n, ip, (double*)inv);
  }
  static int getrs(hipStream_t s, int n, int nrhs, const void* A, const int* ip, const void* inv, void* B)
  {
    return bq_zgetrs_dev(s, n, nrhs, (const double*)A, n, ip, (const double*)inv, (double*)B, n);
  }
  static int blkinv(hipStream_t s, int n, const void* A, const void* inv, void* W)
  {
    return W ? bq_zlu_blkinv_dev(s, n, (const double*)A, n, (const double*)inv, (double*)W) : 0;
  }
  static int inverse(hipStream_t s, int n, const void* A, const void* inv, void* X)
  {
    return bq_zlu_inverse_dev(s, n, (const double*)A, n, (const double*)inv, (double*)X, n);
  }
  static const int* perm(const void* inv, int n) { return bq_lu_perm_dev(inv, n, 0); }
  // few columns through the inverted diagonal blocks W when there are any
  static int getrs_w(hipStream_t s, int n, int nrhs, const void* A, const int* ip, const void* inv, const void* W,
                     void* B)
  {
    if (W && nrhs <= 8)
      return bq_zgetrs_blkinv_dev(s, n, nrhs, (const double*)A, n, (const double*)inv, (const double*)W, (double*)B, n);
    return getrs(s, n, nrhs, A, ip, inv, B);
  }
  static int gemm(hipStream_t s, int m, int nc, int k, cplx al, const void* A, const void* B, cplx be, void* C)
  {
    const double a[2] = {al.re, al.im}, b[2] = {be.re, be.im};
    return bq_zgemm_op_dev(s, 0, 0, m, nc, k, a, (const double*)A, m, (const double*)B, k, b, (double*)C, m);
  }
};
template <>
struct LA<float> {
  using V = float2;
  static int getrf(hipStream_t s, int n, void* A, int* ip, void* inv)
  {
    return bq_cgetrf_dev(s, n, (float*)A, n, ip, (float*)inv);
  }
  static int getrs(hipStream_t s, int n, int nrhs, const void* A, const int* ip, const void* inv, void* B)
  {
    return bq_cgetrs_dev(s, n, nrhs, (const float*)A, n, ip, (const float*)inv, (float*)B, n);
  }
  static int blkinv(hipStream_t s, int n, const void* A, const void* inv, void* W)
  {
    return W ? bq_clu_blkinv_dev(s, n, (const float*)A, n, (const float*)inv, (float*)W) : 0;
  }
  static int inverse(hipStream_t s, int n, const void* A, const void* inv, void* X)
  {
    return bq_clu_inverse_dev(s, n, (const float*)A, n, (const float*)inv, (float*)X, n);
  }
  static const int* perm(const void* inv, int n) { return bq_lu_perm_dev(inv, n, 1); }
  static int getrs_w(hipStream_t s, int n, int nrhs, const void* A, const int* ip, const void* inv, const void* W,
                     void* B)
  {
    if (W && nrhs <= 8)
      return bq_cgetrs_blkinv_dev(s, n, nrhs, (const float*)A, n, (const float*)inv, (const float*)W, (float*)B, n);
    return getrs(s, n, nrhs, A, ip, inv, B);
  }
  static int gemm(hipStream_t s, int m, int nc, int k, cplx al, const void* A, const void* B, cplx be, void* C)
  {
    const float a[2] = {(float)al.re, (float)al.im}, b[2] = {(float)be.re, (float)be.im};
    return bq_cgemm_op_dev(s, 0, 0, m, nc, k, a, (const float*)A, m, (const float*)B, k, b, (float*)C, m);
  }
};

// ---- host helpers -----------------------------------------------------------------------
struct Coef {
  cplx delt, beps, sigm;
};

Coef coefs(const bq_sess* S)
{
  const cplx k1 = {S->kn_h[0], S->kn_h[1]}, k2 = {S->kn_h[2], S->kn_h[3]};
  // host arithmetic on the device complex type (plain doubles)
  const cplx k12 = {k1.re * k1.re - k1.im * k1.im, 2.0 * k1.re * k1.im};
  const cplx k22 = {k2.re * k2.re - k2.im * k2.im, 2.0 * k2.re * k2.im};
  Coef c;
  c.delt = {k12.re - k22.re, k12.im - k22.im};
  const double den = k22.re * k22.re + k22.im * k22.im;
  c.beps = {(k12.re * k22.re + k12.im * k22.im) / den, (k12.im * k22.re - k12.re * k22.im) / den};
  c.sigm = {1.0 - c.beps.re, -c.beps.im};
  return c;
}

cplx hmul(cplx a, cplx b) { return {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re}; }
double habs(cplx a) { return std::hypot(a.re, a.im); }

int dalloc(bq_sess* S, void** p, size_t bytes)
{
  *p = nullptr;
  if (bytes == 0) return 0;
  if (hipMalloc(p, bytes) != hipSuccess) {
    (void)hipGetLastError();
    bq::set_error("device memory: %zu bytes for the single-object reduction", bytes);
    return -1;
  }
  S->red.bytes += bytes;
  return 0;
}

constexpr int kRestart = 20;   // FGMRES restart length
constexpr int kMaxCycles = 4;  // FGMRES cycles before the FP64 factors take over
constexpr int kDotBlocks = 256;

// allocate the reduction's buffers once per session (complex FP64 size if cap64)
int red_alloc(bq_sess* S, bool cap64)
{
  bq_reduced& R = S->red;
  if (R.G0 && (R.cap64 || !cap64)) return 0;
  if (R.G0) bq::reduced_free(S);
  const int n = S->NN;
  const size_t es = cap64 ? 16 : 8, sq = (size_t)n * n * es;
  const size_t invb = (size_t)bq_lu_inv_doubles(n) * (cap64 ? 8 : 4);
  const long N0 = 4L * S->dNe[0], N1 = 4L * S->dNe[1];
  const long Nm = std::max(N0, N1);
  const size_t v8 = (size_t)8 * n * 16;
  R.cap64 = cap64;
  if (dalloc(S, &R.G0, sq) || dalloc(S, &R.G1, sq) || dalloc(S, &R.Mx, sq) || dalloc(S, &R.P, sq) ||
      dalloc(S, &R.B, sq) || dalloc(S, &R.Binv, sq) || dalloc(S, (void**)&R.ip0, (size_t)n * 4) ||
      dalloc(S, (void**)&R.ip1, (size_t)n * 4) || dalloc(S, (void**)&R.ipB, (size_t)n * 4) ||
      dalloc(S, &R.inv0, invb) || dalloc(S, &R.inv1, invb) || dalloc(S, &R.invB, invb) ||
      dalloc(S, (void**)&R.ginv[0], (size_t)n * 4) || dalloc(S, (void**)&R.ginv[1], (size_t)n * 4) ||
      dalloc(S, (void**)&R.gidx[0], (size_t)N0 * 4) || dalloc(S, (void**)&R.gidx[1], (size_t)N1 * 4) ||
      dalloc(S, (void**)&R.nn3, (size_t)n * 24) || dalloc(S, (void**)&R.Uig, (size_t)n * 48) ||
      dalloc(S, (void**)&R.dUig, (size_t)n * 48) || dalloc(S, (void**)&R.rownorm, (size_t)n * 64) ||
      dalloc(S, (void**)&R.b, v8) || dalloc(S, (void**)&R.x, v8) || dalloc(S, (void**)&R.r, v8) ||
      dalloc(S, (void**)&R.t, v8) || dalloc(S, (void**)&R.loc, (size_t)Nm * 12 * 16) ||
      dalloc(S, (void**)&R.nloc, (size_t)Nm * 24) ||
      dalloc(S, (void**)&R.part, (size_t)kDotBlocks * (kRestart + 2) * 16) ||
      dalloc(S, &R.vf, (size_t)15 * n * 16) || dalloc(S, (void**)&R.Vb, v8 * (kRestart + 1)) ||
      dalloc(S, (void**)&R.Zb, v8 * kRestart) || dalloc(S, &R.W0, (size_t)bq_lu_blkinv_elems(n) * es) ||
      dalloc(S, &R.W1, (size_t)bq_lu_blkinv_elems(n) * es) || dalloc(S, &R.WB, (size_t)bq_lu_blkinv_elems(n) * es)) {
    bq::reduced_free(S);
    return -1;
  }
  // node maps and normals (host)
  std::vector<int> gi[2], gx[2];
  std::vector<double> nn((size_t)n * 3, 0.0);
  for (int d = 0; d < 2; ++d) {
    const long N = 4L * S->dNe[d];
    gi[d].assign(n, -1);
    gx[d].assign(N, -1);
    for (int t = 1; t <= S->dNe[d]; ++t) {
      const int at = std::abs(S->sid_h[d][t]);
      const bool tri = S->m.E[24 * (size_t)at + 21] != 0.0;
      for (int tn = 0; tn < 4; ++tn) {
        if (tn == 3 && tri) continue;
        const int e = S->m.eni[4 * at + tn];
        const long r = 4L * (t - 1) + tn;
        if (e < 0 || e >= n) continue;
        gi[d][e] = (int)r;
        gx[d][r] = e;
        for (int k = 0; k < 3; ++k) nn[3 * (size_t)e + k] = S->m.nrm[12 * (size_t)at + 3 * tn + k];
      }
    }
    for (int e = 0; e < n; ++e)
      if (gi[d][e] < 0) {
        bq::set_error("single-object reduction: node %d has no row in domain %d", e, d);
        return -1;
      }
  }
  hipStream_t s = S->s;
  for (int d = 0; d < 2; ++d) {
    BQ_HIP(hipMemcpyAsync(R.ginv[d], gi[d].data(), (size_t)n * 4, hipMemcpyHostToDevice, s));
    BQ_HIP(hipMemcpyAsync(R.gidx[d], gx[d].data(), gx[d].size() * 4, hipMemcpyHostToDevice, s));
  }
  BQ_HIP(hipMemcpyAsync(R.nn3, nn.data(), nn.size() * 8, hipMemcpyHostToDevice, s));
  BQ_HIP(hipStreamSynchronize(s));
  return 0;
}

// incident Ui, dUi of the object side per global node (from the element-node arrays)
int upload_incident(bq_sess* S)
{
  const int n = S->NN;
  std::vector<double> ui((size_t)n * 6, 0.0), dui((size_t)n * 6, 0.0);
  for (int t = 1; t <= S->dNe[1]; ++t) {
    const int at = std::abs(S->sid_h[1][t]);
    for (int l = 0; l < 4; ++l) {
      const int e = S->m.eni[4 * at + l];
      if (e < 0 || e >= n) continue;
      for (int c = 0; c < 3; ++c) {
        ui[2 * ((size_t)c * n + e)] = S->m.Ui[32 * (size_t)at + 8 * l + 2 * c];
        ui[2 * ((size_t)c * n + e) + 1] = S->m.Ui[32 * (size_t)at + 8 * l + 2 * c + 1];
        dui[2 * ((size_t)c * n + e)] = S->m.dUi[32 * (size_t)at + 8 * l + 2 * c];
        dui[2 * ((size_t)c * n + e) + 1] = S->m.dUi[32 * (size_t)at + 8 * l + 2 * c + 1];
      }
    }
  }
  BQ_HIP(hipMemcpyAsync(S->red.Uig, ui.data(), ui.size() * 8, hipMemcpyHostToDevice, S->s));
  BQ_HIP(hipMemcpyAsync(S->red.dUig, dui.data(), dui.size() * 8, hipMemcpyHostToDevice, S->s));
  BQ_HIP(hipStreamSynchronize(S->s));
  return 0;
}

// y (8n, equation rows) = A X (sub == null) or sub - A X, with the FP64 matrices as assembled;
// mode 1 ignores X and forms the right-hand side from the incident field into y (object rows;
// open-region rows 0)
int op_apply(bq_sess* S, int mode, const double2* X, const double2* sub, double2* y)
{
  bq_reduced& R = S->red;
  hipStream_t s = S->s;
  const int n = S->NN;
  const Coef k = coefs(S);
  double2* loc = (double2*)R.loc;
  if (mode == 1) BQ_HIP(hipMemsetAsync(y, 0, (size_t)4 * n * 16, s));
  const int kt = bq::ktime_begin(mode == 1 ? "op_rhs" : "op_residual", s);
  for (int d = (mode == 1 ? 1 : 0); d < 2; ++d) {
    const long N = 4L * S->dNe[d];
    double2* yl = loc + 8 * N;  // the 4 row products after the 8 operand vectors
    k_op_prep<<<nblk(N, 256), 256, 0, s>>>(n, d, N, mode, R.gidx[d], X, R.nn3, (const double2*)R.Uig,
                                           (const double2*)R.dUig, k.delt, k.sigm, k.beps, loc);
    BQ_HIP(hipGetLastError());
    k_rm_mv<<<nblk(N, kMvRows), 256, 0, s>>>(N, (const double2*)S->H[d], (const double2*)S->G[d], loc, yl);
    BQ_HIP(hipGetLastError());
    k_scatter_rows<<<nblk(4 * N, 256), 256, 0, s>>>(n, N, R.gidx[d], yl, 4 * d, sub, y);
    BQ_HIP(hipGetLastError());
  }
  bq::ktime_end(kt, s);
  return 0;
}

// max |a_i| and max |b_i| over 8 n entries
int vmax2(bq_sess* S, const double2* a, const double2* b, double& ma, double& mb)
{
  const long len = 8L * S->NN;
  k_vmax<<<kDotBlocks, 256, 0, S->s>>>(len, a, b, S->red.part);
  BQ_HIP(hipGetLastError());
  std::vector<double> h(2 * kDotBlocks);
  BQ_HIP(hipMemcpyAsync(h.data(), S->red.part, h.size() * 8, hipMemcpyDeviceToHost, S->s));
  BQ_HIP(hipStreamSynchronize(S->s));
  ma = mb = 0.0;
  for (int i = 0; i < kDotBlocks; ++i) {
    ma = std::max(ma, h[2 * i]);
    mb = std::max(mb, h[2 * i + 1]);
  }
  return 0;
}

// h[v] = V_v^H w for v < nv (V: nv vectors of 8 n), summed on the host in a fixed order
int mdot(bq_sess* S, int nv, const double2* V, long ldv, const double2* w, std::vector<cplx>& h)
{
  const long len = 8L * S->NN;
  k_mdot<<<dim3(kDotBlocks, nv), 256, 0, S->s>>>(len, V, ldv, w, (double2*)S->red.part);
  BQ_HIP(hipGetLastError());
  std::vector<double> p((size_t)2 * kDotBlocks * nv);
  BQ_HIP(hipMemcpyAsync(p.data(), S->red.part, p.size() * 8, hipMemcpyDeviceToHost, S->s));
  BQ_HIP(hipStreamSynchronize(S->s));
  h.assign(nv, cplx{0.0, 0.0});
  for (int v = 0; v < nv; ++v)
    for (int b = 0; b < kDotBlocks; ++b) {
      h[v].re += p[2 * ((size_t)v * kDotBlocks + b)];
      h[v].im += p[2 * ((size_t)v * kDotBlocks + b) + 1];
    }
  return 0;
}

// the reduced elimination for 8 n right-hand-side rows r (FP64, equation rows) -> X (FP64)
template <class F>
int red_apply(bq_sess* S, const double2* r, double2* X)
{
  using V = typename LA<F>::V;
  bq_reduced& R = S->red;
  hipStream_t s = S->s;
  const int n = S->NN;
  const long n4 = 4L * n;
  const Coef k = coefs(S);
  V* R0 = (V*)R.vf;   // open-region rows -> u0
  V* R1 = R0 + n4;    // object rows -> u1 -> g
  V* Y4 = R1 + n4;    // [y_0 y_1 y_2 X_3]
  V* T = Y4 + n4;     // n_c X_3
  const int kt = bq::ktime_begin(sizeof(F) == 8 ? "reduced_apply_z" : "reduced_apply_c", s);
  k_cvt<V><<<nblk(2 * n4, 256), 256, 0, s>>>(2 * n4, r, R0);
  BQ_HIP(hipGetLastError());
  if (LA<F>::getrs_w(s, n, 4, R.G0, R.ip0, R.inv0, R.W0, R0)) return -1;
  if (LA<F>::getrs_w(s, n, 4, R.G1, R.ip1, R.inv1, R.W1, R1)) return -1;
  k_gcomb<V><<<nblk(n4, 256), 256, 0, s>>>(n, R0, R1, k.beps);
  BQ_HIP(hipGetLastError());
  if (LA<F>::gemm(s, n, 3, n, cplx{1.0, 0.0}, R.Binv, R1, cplx{0.0, 0.0}, Y4)) return -1;
  k_rhs3<V><<<nblk(n, 256), 256, 0, s>>>(n, R1 + 3L * n, R.nn3, k.sigm, Y4);
  BQ_HIP(hipGetLastError());
  if (LA<F>::getrs_w(s, n, 1, R.B, R.ipB, R.invB, R.WB, Y4 + 3L * n)) return -1;
  k_tnorm<V><<<nblk(n, 256), 256, 0, s>>>(n, Y4 + 3L * n, R.nn3, T);
  BQ_HIP(hipGetLastError());
  const cplx md = {-k.delt.re, -k.delt.im};
  if (LA<F>::gemm(s, n, 3, n, md, R.Binv, T, cplx{1.0, 0.0}, Y4)) return -1;
  if (LA<F>::gemm(s, n, 4, n, cplx{1.0, 0.0}, R.Mx, Y4, cplx{-1.0, 0.0}, R0)) return -1;
  k_cvt_back<V><<<nblk(n4, 256), 256, 0, s>>>(n4, Y4, X);
  k_cvt_back<V><<<nblk(n4, 256), 256, 0, s>>>(n4, R0, X + n4);
  BQ_HIP(hipGetLastError());
  bq::ktime_end(kt, s);
  return 0;
}

int red_apply_p(bq_sess* S, const double2* r, double2* X)
{
  return S->red.prec == 1 ? red_apply<float>(S, r, X) : red_apply<double>(S, r, X);
}

template <class F>
int red_factor_t(bq_sess* S)
{
  using V = typename LA<F>::V;
  bq_reduced& R = S->red;
  hipStream_t s = S->s;
  const int n = S->NN;
  const long nn = (long)n * n;
  const Coef k = coefs(S);
  const dim3 tg(nblk(n, 32), nblk(n, 32));
  for (int d = 0; d < 2; ++d) {
    const long N = 4L * S->dNe[d];
    k_to_global<V><<<tg, 256, 0, s>>>(n, N, R.ginv[d], (const double2*)S->G[d], (V*)(d ? R.G1 : R.G0));
    k_to_global<V><<<tg, 256, 0, s>>>(n, N, R.ginv[d], (const double2*)S->H[d], (V*)(d ? R.P : R.Mx));
    BQ_HIP(hipGetLastError());
  }
  // The two domains' chains, LU(G0) -> M = G0^-1 H0 and LU(G1) -> P = G1^-1 H1, are
  // independent: the object's runs on the session's second stream from a helper thread
  // (getrf waits on the host for its pivot check), so that the latency-bound panel steps of
  // one overlap the matrix-core updates of the other.
  if (!S->s2 && hipStreamCreateWithFlags(&S->s2, hipStreamNonBlocking) != hipSuccess) {
    S->s2 = nullptr;
    bq::set_error("second session stream creation failed");
    return -1;
  }
  hipStream_t s2 = S->s2;
  hipEvent_t ev[2] = {nullptr, nullptr};
  for (auto& e : ev) BQ_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
  struct EvFree {
    hipEvent_t* e;
    ~EvFree()
    {
      for (int i = 0; i < 2; ++i)
        if (e[i]) (void)hipEventDestroy(e[i]);
    }
  } ev_free{ev};
  BQ_HIP(hipEventRecord(ev[0], s));  // G1, H1 converted
  BQ_HIP(hipStreamWaitEvent(s2, ev[0], 0));
  int dev = 0;
  BQ_HIP(hipGetDevice(&dev));
  int info1 = 0, rc1 = 0;
  std::string err1;
  std::thread helper([&] {
    if (hipSetDevice(dev) != hipSuccess) {
      rc1 = -1;
      err1 = "helper thread: hipSetDevice failed";
      return;
    }
    info1 = LA<F>::getrf(s2, n, R.G1, R.ip1, R.inv1);
    if (info1 == 0 && (LA<F>::getrs(s2, n, n, R.G1, R.ip1, R.inv1, R.P) ||  // P = G1^-1 H1
                       LA<F>::blkinv(s2, n, R.G1, R.inv1, R.W1)))
      rc1 = -1;
    if (info1 || rc1) err1 = bq_last_error();
    if (hipEventRecord(ev[1], s2) != hipSuccess && !rc1) {
      rc1 = -1;
      err1 = "helper thread: event record failed";
    }
  });
  int info = LA<F>::getrf(s, n, R.G0, R.ip0, R.inv0);
  int rc0 = (info == 0) ? (LA<F>::getrs(s, n, n, R.G0, R.ip0, R.inv0, R.Mx) ||  // M = G0^-1 H0
                           LA<F>::blkinv(s, n, R.G0, R.inv0, R.W0))
                        : 0;
  helper.join();
  if (info) {
    bq::set_error("single-object reduction: G of the open region is singular or not finite (info %d)", info);
    (void)hipStreamSynchronize(s2);
    return info > 0 ? -3 : info;
  }
  if (rc0) {
    (void)hipStreamSynchronize(s2);
    return -1;
  }
  if (info1) {
    bq::set_error("single-object reduction: G of the object is singular or not finite (info %d)", info1);
    (void)hipStreamSynchronize(s2);
    return info1 > 0 ? -3 : info1;
  }
  if (rc1) {
    bq::set_error("%s", err1.c_str());
    (void)hipStreamSynchronize(s2);
    return -1;
  }
  BQ_HIP(hipStreamWaitEvent(s, ev[1], 0));
  k_add<V><<<nblk(nn, 256), 256, 0, s>>>(nn, (const V*)R.P, (const V*)R.Mx, (V*)R.B);
  BQ_HIP(hipGetLastError());
  info = LA<F>::getrf(s, n, R.B, R.ipB, R.invB);
  if (info) {
    bq::set_error("single-object reduction: P + M is singular (info %d)", info);
    return info > 0 ? -3 : info;
  }
  // B^-1 with permuted columns, then the Schur complement, which also writes B^-1 in natural
  // order over P (no longer needed); the two buffers then trade places
  if (LA<F>::inverse(s, n, R.B, R.invB, R.Binv)) return -1;
  const cplx sd = hmul(k.sigm, k.delt);
  k_schur<V><<<nblk(nn, 256), 256, 0, s>>>(n, (V*)R.P, (const V*)R.Mx, (const V*)R.Binv, LA<F>::perm(R.invB, n),
                                           R.nn3, k.beps, sd, (V*)R.B);
  BQ_HIP(hipGetLastError());
  std::swap(R.P, R.Binv);
  info = LA<F>::getrf(s, n, R.B, R.ipB, R.invB);
  if (info) {
    bq::set_error("single-object reduction: Schur complement is singular (info %d)", info);
    return info > 0 ? -3 : info;
  }
  return LA<F>::blkinv(s, n, R.B, R.invB, R.WB);
}

// ||A||_inf of the unreduced operator from exact row sums
int op_norm(bq_sess* S)
{
  bq_reduced& R = S->red;
  hipStream_t s = S->s;
  const int n = S->NN;
  const Coef k = coefs(S);
  double* nloc = R.nloc;
  BQ_HIP(hipMemsetAsync(R.rownorm, 0, (size_t)n * 64, s));
  for (int d = 0; d < 2; ++d) {
    const long N = 4L * S->dNe[d];
    k_nloc<<<nblk(N, 256), 256, 0, s>>>(N, R.gidx[d], R.nn3, nloc);
    k_rownorm<<<(unsigned)N, 256, 0, s>>>(N, n, d, R.gidx[d], (const double2*)S->H[d], (const double2*)S->G[d], nloc,
                                           habs(k.delt), habs(k.sigm), habs(k.beps), R.rownorm);
    BQ_HIP(hipGetLastError());
  }
  std::vector<double> h((size_t)8 * n);
  BQ_HIP(hipMemcpyAsync(h.data(), R.rownorm, h.size() * 8, hipMemcpyDeviceToHost, s));
  BQ_HIP(hipStreamSynchronize(s));
  double a = 0.0;
  for (double v : h) a = std::max(a, v);
  R.anorm = a;
  return 0;
}

// r = b - A x and its backward error
int residual(bq_sess* S, const double2* b, const double2* x, double2* r, double bmax, double& eta, double& rmax)
{
  if (op_apply(S, 0, x, b, r)) return -1;
  double xmax = 0.0;
  if (vmax2(S, r, x, rmax, xmax)) return -1;
  const double den = S->red.anorm * xmax + bmax;
  eta = den > 0.0 ? rmax / den : 0.0;
  return 0;
}

}  // namespace

namespace bq {

bool single_object(const bq_sess* S)
{
  if (S->MN != 1 || S->dNe.size() != 2) return false;
  if (S->dNe[0] != S->Ne || S->dNe[1] != S->Ne) return false;
  for (int i = 1; i <= S->Ne; ++i)
    if (S->m.md[i] != 0 || S->m.sd[i] != 1) return false;
  return S->nlive[0] == S->NN && S->nlive[1] == S->NN;
}

void reduced_free(bq_sess* S)
{
  bq_reduced& R = S->red;
  auto f = [](void* p) { if (p) (void)hipFree(p); };
  for (void* p : {R.G0, R.G1, R.Mx, R.P, R.B, R.Binv, (void*)R.ip0, (void*)R.ip1, (void*)R.ipB, R.inv0, R.inv1,
                  R.invB, (void*)R.ginv[0], (void*)R.ginv[1], (void*)R.gidx[0], (void*)R.gidx[1], (void*)R.nn3,
                  (void*)R.Uig, (void*)R.dUig, (void*)R.rownorm, (void*)R.b, (void*)R.x, (void*)R.r, (void*)R.t,
                  (void*)R.loc, (void*)R.nloc, (void*)R.part, R.vf, (void*)R.Vb, (void*)R.Zb, R.W0, R.W1, R.WB})
    f(p);
  R = bq_reduced();
}

int reduced_factor(bq_sess* S, int prec)
{
  hipStream_t s = S->s;
  if (red_alloc(S, prec == 0)) return -1;
  hipEvent_t e0, e1;
  BQ_HIP(hipEventCreate(&e0));
  BQ_HIP(hipEventCreate(&e1));
  BQ_HIP(hipEventRecord(e0, s));
  S->red.prec = -1;
  const int kt = bq::ktime_begin(prec ? "reduced_factor_c" : "reduced_factor_z", s);
  const int rc = prec ? red_factor_t<float>(S) : red_factor_t<double>(S);
  bq::ktime_end(kt, s);
  if (rc) return rc;
  if (op_norm(S)) return -1;
  S->red.prec = prec;
  BQ_HIP(hipEventRecord(e1, s));
  BQ_HIP(hipEventSynchronize(e1));
  float ms = 0.f;
  BQ_HIP(hipEventElapsedTime(&ms, e0, e1));
  S->st.t_factor_ms = ms;
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  return 0;
}

int reduced_solve(bq_sess* S, double* Xh)
{
  bq_reduced& R = S->red;
  hipStream_t s = S->s;
  const int n = S->NN;
  const long len = 8L * n;
  if (R.prec < 0) {
    bq::set_error("single-object reduction: no factors");
    return -1;
  }
  if (upload_incident(S)) return -1;
  double2 *b = (double2*)R.b, *x = (double2*)R.x, *r = (double2*)R.r;
  if (op_apply(S, 1, nullptr, nullptr, b)) return -1;
  double bmax = 0.0, dummy = 0.0;
  if (vmax2(S, b, nullptr, bmax, dummy)) return -1;
  if (red_apply_p(S, b, x)) return -1;
  double eta = 0.0, rmax = 0.0;
  if (residual(S, b, x, r, bmax, eta, rmax)) return -1;
  S->st.ir_iterations = 0;
  S->st.gmres_iterations = 0;
  if (R.prec == 0) {
    // up to two refinement steps with the FP64 factors, while they shrink the residual
    for (int it = 0; it < 2 && eta > 1e-16; ++it) {
      double2* d = (double2*)R.t;
      if (red_apply_p(S, r, d)) return -1;
      k_vaxpy<<<nblk(len, 256), 256, 0, s>>>(len, d, x);
      BQ_HIP(hipGetLastError());
      double e2 = 0.0, r2 = 0.0;
      if (residual(S, b, x, r, bmax, e2, r2)) return -1;
      S->st.ir_iterations = it + 1;
      if (!(e2 < 0.5 * eta)) {
        eta = e2;
        break;
      }
      eta = e2;
    }
  } else {
    // flexible GMRES(kRestart) on A x = b, right-preconditioned by the FP32 elimination; the
    // Krylov basis V and the preconditioned vectors Z = P V are contiguous (ld = len)
    const double kEta = 1e-15;
    double2 *Vb = (double2*)R.Vb, *Zb = (double2*)R.Zb;
    double prev = eta;
    int total = 0;
    for (int cyc = 0; cyc < kMaxCycles && eta > kEta; ++cyc) {
      std::vector<cplx> h0;
      if (mdot(S, 1, r, len, r, h0)) return -1;
      const double beta = std::sqrt(h0[0].re);
      if (!(beta > 0.0)) break;
      double xmax = 0.0, dm = 0.0;
      if (vmax2(S, x, nullptr, xmax, dm)) return -1;
      // ||r||_inf <= ||r||_2: the cycle stops once the 2-norm estimate meets the target
      const double stop2 = kEta * (R.anorm * xmax + bmax);
      k_vscale<<<nblk(len, 256), 256, 0, s>>>(len, r, 1.0 / beta, Vb);
      BQ_HIP(hipGetLastError());
      std::vector<std::vector<cplx>> Hm(kRestart, std::vector<cplx>(kRestart + 1, cplx{0, 0}));
      std::vector<cplx> gv(kRestart + 1, cplx{0, 0}), sn(kRestart);
      std::vector<double> csr(kRestart);
      gv[0] = cplx{beta, 0.0};
      int kdim = 0;
      for (int j = 0; j < kRestart; ++j) {
        double2* w = Vb + (long)(j + 1) * len;
        if (red_apply_p(S, Vb + (long)j * len, Zb + (long)j * len)) return -1;
        if (op_apply(S, 0, Zb + (long)j * len, nullptr, w)) return -1;
        // classical Gram-Schmidt against V_0..V_j with one re-orthogonalisation
        std::vector<cplx> hc(j + 1, cplx{0, 0});
        for (int pass = 0; pass < 2; ++pass) {
          std::vector<cplx> hp;
          if (mdot(S, j + 1, Vb, len, w, hp)) return -1;
          std::vector<double> hv(2 * (j + 1));
          for (int v = 0; v <= j; ++v) {
            hv[2 * v] = hp[v].re;
            hv[2 * v + 1] = hp[v].im;
            hc[v].re += hp[v].re;
            hc[v].im += hp[v].im;
          }
          BQ_HIP(hipMemcpyAsync(R.part, hv.data(), hv.size() * 8, hipMemcpyHostToDevice, s));
          k_vupdate<<<nblk(len, 256), 256, 0, s>>>(len, j + 1, Vb, len, (const double2*)R.part, w);
          BQ_HIP(hipGetLastError());
        }
        std::vector<cplx> hn;
        if (mdot(S, 1, w, len, w, hn)) return -1;
        const double hnorm = std::sqrt(hn[0].re);
        for (int v = 0; v <= j; ++v) Hm[j][v] = hc[v];
        if (hnorm > 0.0) {
          k_vscale<<<nblk(len, 256), 256, 0, s>>>(len, w, 1.0 / hnorm, w);
          BQ_HIP(hipGetLastError());
        }
        // previous rotations [c s; -conj(s) c] (c real) on column j, then a new one that
        // zeroes its subdiagonal entry hnorm
        for (int v = 0; v < j; ++v) {
          const cplx a = Hm[j][v], c2 = Hm[j][v + 1];
          Hm[j][v] = cplx{csr[v] * a.re + (sn[v].re * c2.re - sn[v].im * c2.im),
                          csr[v] * a.im + (sn[v].re * c2.im + sn[v].im * c2.re)};
          Hm[j][v + 1] = cplx{-(sn[v].re * a.re + sn[v].im * a.im) + csr[v] * c2.re,
                              -(sn[v].re * a.im - sn[v].im * a.re) + csr[v] * c2.im};
        }
        {
          const cplx a = Hm[j][j];
          const double b2 = hnorm, aa = habs(a), den = std::hypot(aa, b2);
          double c = 1.0;
          cplx sv = {0.0, 0.0};
          if (den > 0.0 && aa == 0.0) {
            c = 0.0;
            sv = cplx{1.0, 0.0};
          } else if (den > 0.0) {
            c = aa / den;
            sv = cplx{a.re / aa * b2 / den, a.im / aa * b2 / den};  // (a / |a|) b / den
          }
          csr[j] = c;
          sn[j] = sv;
          const cplx g = gv[j];
          Hm[j][j] = cplx{c * a.re + sv.re * b2, c * a.im + sv.im * b2};
          gv[j] = cplx{c * g.re, c * g.im};
          gv[j + 1] = cplx{-(sv.re * g.re + sv.im * g.im), -(sv.re * g.im - sv.im * g.re)};
        }
        kdim = j + 1;
        ++total;
        if (habs(gv[j + 1]) <= stop2 || hnorm == 0.0) break;
      }
      // y = R^-1 g (upper triangular, R[i][j] = Hm[j][i]), x += Z y
      std::vector<cplx> y(kdim);
      for (int i = kdim - 1; i >= 0; --i) {
        cplx a = gv[i];
        for (int j = i + 1; j < kdim; ++j) {
          const cplx p = hmul(Hm[j][i], y[j]);
          a.re -= p.re;
          a.im -= p.im;
        }
        const cplx d = Hm[i][i];
        const double dd = d.re * d.re + d.im * d.im;
        y[i] = cplx{(a.re * d.re + a.im * d.im) / dd, (a.im * d.re - a.re * d.im) / dd};
      }
      std::vector<double> hv(2 * kdim);
      for (int j = 0; j < kdim; ++j) {
        hv[2 * j] = -y[j].re;
        hv[2 * j + 1] = -y[j].im;
      }
      BQ_HIP(hipMemcpyAsync(R.part, hv.data(), hv.size() * 8, hipMemcpyHostToDevice, s));
      k_vupdate<<<nblk(len, 256), 256, 0, s>>>(len, kdim, Zb, len, (const double2*)R.part, x);
      BQ_HIP(hipGetLastError());
      if (residual(S, b, x, r, bmax, eta, rmax)) return -1;
      if (cyc > 0 && !(eta < 0.5 * prev)) break;  // stagnation at the FP64 rounding level
      prev = eta;
    }
    S->st.gmres_iterations = total;
  }
  S->st.backward_error = eta;
  if (!(eta <= 1e-10)) {
    bq::set_error("single-object reduction: backward error %.3g above 1e-10", eta);
    return 1;
  }
  // to the reference's X layout X[c NN + e], c = 0..7
  std::vector<double> hx((size_t)2 * len);
  BQ_HIP(hipMemcpyAsync(hx.data(), x, hx.size() * 8, hipMemcpyDeviceToHost, s));
  BQ_HIP(hipStreamSynchronize(s));
  std::memcpy(Xh, hx.data(), hx.size() * 8);
  return 0;
}

int reduced_deh(bq_sess* S, int d, const double* EH, double* dEH)
{
  bq_reduced& R = S->red;
  hipStream_t s = S->s;
  const int n = S->NN;
  const long N = 4L * S->dNe[d];
  double2 *eh = (double2*)R.loc, *bl = eh + 6 * N;                        // 12 N
  double2 *bg = (double2*)R.r, *xg = (double2*)R.x, *rg = (double2*)R.t;  // 6 n <= 8 n each
  BQ_HIP(hipMemcpyAsync(eh, EH, (size_t)N * 6 * 16, hipMemcpyHostToDevice, s));
  const double one[2] = {1, 0}, zero[2] = {0, 0};
  // H [E H] over all local rows (rows and columns without a node contribute 0)
  if (bq_zgemm_op_dev(s, 1, 0, (int)N, 6, (int)N, one, S->H[d], N, (const double*)eh, N, zero, (double*)bl, N))
    return -1;
  k_perm6<<<nblk(N * 6, 256), 256, 0, s>>>(n, N, 6, R.gidx[d], bl, bg, 0);
  BQ_HIP(hipGetLastError());
  const void* Gf = d ? R.G1 : R.G0;
  const int* ip = d ? R.ip1 : R.ip0;
  const void* inv = d ? R.inv1 : R.inv0;
  const void* Wd = d ? R.W1 : R.W0;
  const long n6 = 6L * n;
  if (R.prec == 0) {
    BQ_HIP(hipMemcpyAsync(xg, bg, (size_t)n6 * 16, hipMemcpyDeviceToDevice, s));
    if (LA<double>::getrs_w(s, n, 6, Gf, ip, inv, Wd, xg)) return -1;
  } else {
    // FP32 factors + refinement with the FP64 G_d: x += G^-1 (b - G x)
    float2* vf = (float2*)R.vf;
    BQ_HIP(hipMemsetAsync(xg, 0, (size_t)n6 * 16, s));
    BQ_HIP(hipMemcpyAsync(rg, bg, (size_t)n6 * 16, hipMemcpyDeviceToDevice, s));
    double prev = 1e300;
    for (int it = 0; it < 10; ++it) {
      k_cvt<float2><<<nblk(n6, 256), 256, 0, s>>>(n6, rg, vf);
      BQ_HIP(hipGetLastError());
      if (LA<float>::getrs_w(s, n, 6, Gf, ip, inv, Wd, vf)) return -1;
      k_cvt_back<float2><<<nblk(n6, 256), 256, 0, s>>>(n6, vf, rg);
      k_vaxpy<<<nblk(n6, 256), 256, 0, s>>>(n6, rg, xg);
      BQ_HIP(hipGetLastError());
      // correction size relative to x
      std::vector<double> hd((size_t)2 * n6), hx((size_t)2 * n6);
      BQ_HIP(hipMemcpyAsync(hd.data(), rg, hd.size() * 8, hipMemcpyDeviceToHost, s));
      BQ_HIP(hipMemcpyAsync(hx.data(), xg, hx.size() * 8, hipMemcpyDeviceToHost, s));
      BQ_HIP(hipStreamSynchronize(s));
      double dm = 0.0, xm = 0.0;
      for (size_t i = 0; i < hd.size(); ++i) {
        dm = std::max(dm, std::fabs(hd[i]));
        xm = std::max(xm, std::fabs(hx[i]));
      }
      const double corr = xm > 0.0 ? dm / xm : 0.0;
      if (!(corr == corr)) {
        bq::set_error("reduced dE/dH: non-finite correction");
        return -1;
      }
      if (corr <= 1e-15 || (it > 1 && corr > 0.5 * prev)) break;
      prev = corr;
      // r = b - G x with the FP64 domain matrix (local order)
      k_perm6<<<nblk(N * 6, 256), 256, 0, s>>>(n, N, 6, R.gidx[d], xg, eh, 1);
      BQ_HIP(hipGetLastError());
      if (bq_zgemm_op_dev(s, 1, 0, (int)N, 6, (int)N, one, S->G[d], N, (const double*)eh, N, zero, (double*)bl, N))
        return -1;
      k_perm6<<<nblk(N * 6, 256), 256, 0, s>>>(n, N, 6, R.gidx[d], bl, rg, 0);
      BQ_HIP(hipGetLastError());
      // rg = bg - rg
      k_vscale<<<nblk(n6, 256), 256, 0, s>>>(n6, rg, -1.0, rg);
      k_vaxpy<<<nblk(n6, 256), 256, 0, s>>>(n6, bg, rg);
      BQ_HIP(hipGetLastError());
    }
  }
  k_perm6<<<nblk(N * 6, 256), 256, 0, s>>>(n, N, 6, R.gidx[d], xg, bl, 1);
  BQ_HIP(hipGetLastError());
  BQ_HIP(hipMemcpyAsync(dEH, bl, (size_t)N * 6 * 16, hipMemcpyDeviceToHost, s));
  BQ_HIP(hipStreamSynchronize(s));
  return 0;
}

}  // namespace bq
