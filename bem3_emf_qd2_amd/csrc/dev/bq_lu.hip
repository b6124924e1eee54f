// bq_lu.hip -- dense complex LU with partial pivoting and the matching solves, on the
// device, column-major (LAPACK zgetrf/zgetrs and cgetrf/cgetrs semantics, 0-based pivots),
// in FP64 and FP32 from one templated set of kernels (the FP32 factors precondition the
// mixed-precision solve).
//
// Right-looking LU blocked on two levels.  An outer panel of 512 columns is factored in
// steps of nb = 32 columns:
//   1. k_tslu       chooses the step's nb pivot rows by tournament (partial pivoting on
//                   256-row blocks in registers, then 8-way reductions of candidate sets);
//                   the last level also yields the factored diagonal block,
//   2. k_swap       applies the interchanges to the outer panel's columns,
//   3. k_diag_finish  stores L11, U11 and their inverses inv(L11), inv(U11),
//   4. k_l21        L21 = A21 inv(U11),   k_apply_blk  U12 = inv(L11) A12 inside the panel,
//   5. rocBLAS zgemm updates the rest of the outer panel.
// Then the panel's interchanges go to all other columns, U12 of the outer block row is
// solved with the panel's unit L11 (32-column steps), and one rank-512 zgemm updates the
// trailing matrix (MFMA on gfx950: FP64, or FP32 for the single-precision factors); with the
// look-ahead the next panel's columns are updated first and the rest on a second stream,
// concurrently with the next panel's steps.  A rank-32 update per 32 columns would stream the
// trailing matrix through HBM 8x as often; at order 36 648 that traffic alone is ~10 s.
// Solves: the permutation getrf stored applies the pivots in one gather, then, per 512-row
// block, 32-row diagonal steps X_j = inv(T_jj) X_j (k_apply_blk) with in-block zgemm
// updates, and one rank-512 zgemm for the rows outside the block; with at most 8 right-hand
// sides one kernel runs a whole 256-row diagonal block and one skinny kernel the rows
// outside it.
// Replaces the reference's MKL PARDISO and LAPACKE_zgesv calls
// (d3qd2_src/d3qd2_mkl_solver.c:18-148, d3qd2_src/d3qd2_solve_bieq.c:753).
#include <rocblas/rocblas.h>
#include <atomic>
#include <cstring>
#include <mutex>
#include <type_traits>
#include <vector>
#include "bq_runtime.hpp"
#include "bq_complex.hpp"
#include "../bq_linalg.h"

namespace {

// complex float with the operations the kernels use (as cplx for double)
struct cplxf {
  float re, im;
};
__device__ __forceinline__ cplxf operator+(cplxf a, cplxf b) { return cplxf{a.re + b.re, a.im + b.im}; }
__device__ __forceinline__ cplxf operator-(cplxf a, cplxf b) { return cplxf{a.re - b.re, a.im - b.im}; }
__device__ __forceinline__ cplxf operator*(cplxf a, cplxf b)
{
  return cplxf{fmaf(a.re, b.re, -a.im * b.im), fmaf(a.re, b.im, a.im * b.re)};
}
__device__ __forceinline__ cplxf cdivf(cplxf a, cplxf b)
{
  if (fabsf(b.re) >= fabsf(b.im)) {
    float r = b.im / b.re, d = b.re + b.im * r;
    return cplxf{(a.re + a.im * r) / d, (a.im - a.re * r) / d};
  }
  float r = b.re / b.im, d = b.im + b.re * r;
  return cplxf{(a.re * r + a.im) / d, (a.im * r - a.re) / d};
}

// precision traits: storage pair V, register complex C, conversions and the rocBLAS gemm
template <class R>
struct P;
template <>
struct P<double> {
  using V = double2;
  using C = cplx;
  static __device__ __forceinline__ C mkc(double r, double i) { return mk(r, i); }
  static __device__ __forceinline__ C c(V a) { return mk(a.x, a.y); }
  static __host__ __device__ __forceinline__ V v(double r, double i) { return make_double2(r, i); }
  static __device__ __forceinline__ V v(C a) { return make_double2(a.re, a.im); }
  static __device__ __forceinline__ C div(C a, C b) { return cdiv(a, b); }
  static rocblas_status gemm(rocblas_handle h, rocblas_operation ta, rocblas_operation tb, int m, int n, int k,
                             const V* al, const V* A, int lda, const V* B, int ldb, const V* be, V* Cm, int ldc)
  {
    return rocblas_zgemm(h, ta, tb, m, n, k, (const rocblas_double_complex*)al, (const rocblas_double_complex*)A, lda,
                         (const rocblas_double_complex*)B, ldb, (const rocblas_double_complex*)be,
                         (rocblas_double_complex*)Cm, ldc);
  }
  static rocblas_status gemm_sb(rocblas_handle h, int m, int n, int k, const V* al, const V* A, int lda, long sa,
                                const V* B, int ldb, long sb, const V* be, V* Cm, int ldc, long sc, int batch)
  {
    return rocblas_zgemm_strided_batched(h, rocblas_operation_none, rocblas_operation_none, m, n, k,
                                         (const rocblas_double_complex*)al, (const rocblas_double_complex*)A, lda, sa,
                                         (const rocblas_double_complex*)B, ldb, sb, (const rocblas_double_complex*)be,
                                         (rocblas_double_complex*)Cm, ldc, sc, batch);
  }
  static constexpr const char* kt_f = "zgetrf";
  static constexpr const char* kt_p = "zgetrf_panel";
  static constexpr const char* kt_u = "zgetrf_update";
  static constexpr const char* kt_s = "zgetrs";
};
template <>
struct P<float> {
  using V = float2;
  using C = cplxf;
  static __device__ __forceinline__ C mkc(float r, float i) { return cplxf{r, i}; }
  static __device__ __forceinline__ C c(V a) { return cplxf{a.x, a.y}; }
  static __host__ __device__ __forceinline__ V v(float r, float i) { return make_float2(r, i); }
  static __device__ __forceinline__ V v(C a) { return make_float2(a.re, a.im); }
  static __device__ __forceinline__ C div(C a, C b) { return cdivf(a, b); }
  static rocblas_status gemm(rocblas_handle h, rocblas_operation ta, rocblas_operation tb, int m, int n, int k,
                             const V* al, const V* A, int lda, const V* B, int ldb, const V* be, V* Cm, int ldc)
  {
    return rocblas_cgemm(h, ta, tb, m, n, k, (const rocblas_float_complex*)al, (const rocblas_float_complex*)A, lda,
                         (const rocblas_float_complex*)B, ldb, (const rocblas_float_complex*)be,
                         (rocblas_float_complex*)Cm, ldc);
  }
  static rocblas_status gemm_sb(rocblas_handle h, int m, int n, int k, const V* al, const V* A, int lda, long sa,
                                const V* B, int ldb, long sb, const V* be, V* Cm, int ldc, long sc, int batch)
  {
    return rocblas_cgemm_strided_batched(h, rocblas_operation_none, rocblas_operation_none, m, n, k,
                                         (const rocblas_float_complex*)al, (const rocblas_float_complex*)A, lda, sa,
                                         (const rocblas_float_complex*)B, ldb, sb, (const rocblas_float_complex*)be,
                                         (rocblas_float_complex*)Cm, ldc, sc, batch);
  }
  static constexpr const char* kt_f = "cgetrf";
  static constexpr const char* kt_p = "cgetrf_panel";
  static constexpr const char* kt_u = "cgetrf_update";
  static constexpr const char* kt_s = "cgetrs";
};

constexpr int kNB = 32;
constexpr int kNB2 = 256;  // row blocks of the solves (multiple of kNB)
constexpr int kNBF = 512;  // outer panel of the factorization (multiple of kNB)
constexpr int kBlkElems = 2 * kNB * kNB;  // inv(L11) then inv(U11), complex, per diagonal block


// ---- panel: tournament pivoting ---------------------------------------------------------
// The m x nb panel is split into blocks of 256 rows.  One 256-thread workgroup runs partial
// pivoting on a block held in registers (one row of nb columns per lane) and keeps only the
// nb chosen rows (leaf level).  Up to 8 candidate sets are then reduced the same way on their original rows
// (merge levels) until nb rows remain; they become the panel's pivots in the order the last
// reduction chose them, and that reduction's eliminated rows are the factors L11, U11 of the
// diagonal block.  This replaces a column-by-column pivot search over the whole panel (a
// global reduction per column) by log8(m/256) + 1 small factorizations; the pivots are those
// of communication-avoiding LU (Grigori, Demmel, Xiang, SIAM J. Matrix Anal. Appl. 32(4), 2011).
constexpr int kTsRows = 256;  // rows per tournament workgroup
constexpr int kTsFan = kTsRows / kNB;  // candidate sets merged per workgroup

// Wave-wide maximum (every lane receives it): four DPP steps give each 16-lane row its maximum,
// then the four rows' values are read as scalars.
template <int CTRL>
__device__ __forceinline__ float dpp_mov(float x)
{
  return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(x), CTRL, 0xF, 0xF, false));
}
template <int CTRL>
__device__ __forceinline__ double dpp_mov(double x)
{
  const long long b = __double_as_longlong(x);
  const int lo = __builtin_amdgcn_update_dpp(0, (int)b, CTRL, 0xF, 0xF, false);
  const int hi = __builtin_amdgcn_update_dpp(0, (int)(b >> 32), CTRL, 0xF, 0xF, false);
  return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}
__device__ __forceinline__ float rd_lane(float x, int l) { return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(x), l)); }
__device__ __forceinline__ double rd_lane(double x, int l)
{
  const long long b = __double_as_longlong(x);
  const int lo = __builtin_amdgcn_readlane((int)b, l), hi = __builtin_amdgcn_readlane((int)(b >> 32), l);
  return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}
template <class R>
__device__ __forceinline__ R wave_max_all(R x)
{
  x = fmax(x, dpp_mov<0xB1>(x));   // quad permutation [1,0,3,2]
  x = fmax(x, dpp_mov<0x4E>(x));   // quad permutation [2,3,0,1]
  x = fmax(x, dpp_mov<0x141>(x));  // half-row mirror: the two quads of each 8 lanes
  x = fmax(x, dpp_mov<0x140>(x));  // row mirror: the two halves of each 16 lanes
  return fmax(fmax(rd_lane(x, 0), rd_lane(x, 16)), fmax(rd_lane(x, 32), rd_lane(x, 48)));
}

// 1/p with p scaled by a power of two so that |p|^2 neither overflows nor underflows (zero for
// p = 0): the multipliers of a column are a * (1/p), as LAPACK's cgetf2/zgetf2 scale by the
// reciprocal of the pivot
__device__ __forceinline__ cplxf recip_c(cplxf p)
{
  const float m = fmaxf(fabsf(p.re), fabsf(p.im));
  int e;
  (void)frexpf(m, &e);
  const float r = ldexpf(p.re, -e), i = ldexpf(p.im, -e);
  const float d = m > 0.f ? __builtin_amdgcn_rcpf(fmaf(r, r, i * i)) : 0.f;
  return cplxf{ldexpf(r * d, -e), ldexpf(-i * d, -e)};
}
__device__ __forceinline__ cplx recip_c(cplx p)
{
  const double m = fmax(fabs(p.re), fabs(p.im));
  int e;
  (void)frexp(m, &e);
  const double r = ldexp(p.re, -e), i = ldexp(p.im, -e), q = m > 0.0 ? fma(r, r, i * i) : 1.0;
  double d = __builtin_amdgcn_rcp(q);  // estimate refined by two Newton steps
  d = fma(d, fma(-q, d, 1.0), d);
  d = fma(d, fma(-q, d, 1.0), d);
  if (!(m > 0.0)) d = 0.0;
  return mk(ldexp(r * d, -e), ldexp(-i * d, -e));
}

// mode 0 (leaf): rows j + 256 blk + slot.  mode 1 (merge): slot -> set 8 blk + slot / nb.
// cout[blk][nb] receives the chosen rows.  When gridDim.x == 1 the call is the last level:
// F (nb x nb, column-major) receives the eliminated pivot rows (unit L below the diagonal,
// U on and above) and ipiv[j..j+jb) the LAPACK-style interchanges for those rows.
// Four wavefronts, one row of nb columns per lane.  A column step: the wavefront's maximum of
// |re| + |im| by DPP, its first lane holding it writes its value, row number and the row's
// remaining columns to LDS, one barrier, and every lane takes the largest of the four (ties to
// the lowest wavefront), forms its own multiplier and updates its row.  The LDS slots alternate
// between two buffers, so the next column's writes need no second barrier.  Ties go to the
// lowest thread: the lowest row in a leaf block.
template <class R>
__global__ void __launch_bounds__(256) k_tslu(int mode, int n, const typename P<R>::V* __restrict__ A, long lda, int j, int jb,
                                              const int* __restrict__ cin, int nin, int* __restrict__ cout,
                                              typename P<R>::V* __restrict__ F, int* __restrict__ ipiv)
{
  using V = typename P<R>::V;
  using C = typename P<R>::C;
  using T = P<R>;
  __shared__ R vbuf[2][4];
  __shared__ int ibuf[2][4];
  __shared__ V prow[2][4][kNB];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  int row_id = -1;
  if (mode == 0) {
    const int r = j + blockIdx.x * kTsRows + tid;
    row_id = r < n ? r : -1;
  } else {
    const int set = kTsFan * blockIdx.x + tid / kNB, k = tid % kNB;
    row_id = (set < nin && k < jb) ? cin[set * kNB + k] : -1;
  }
  C a[kNB];
#pragma unroll
  for (int col = 0; col < kNB; ++col)
    a[col] = (row_id >= 0 && col < jb) ? T::c(A[(long)(j + col) * lda + row_id]) : T::mkc(0.0, 0.0);
  bool live = row_id >= 0;
  int mystep = -1, myord = -1;
#pragma unroll
  for (int c = 0; c < kNB; ++c) {
    if (c < jb) {
      const int b = c & 1;
      const R v = live ? fabs(a[c].re) + fabs(a[c].im) : R(-1.0);
      const R wm = wave_max_all(v);
      const unsigned long long hit = __ballot(v == wm);
      const int wl = hit ? __ffsll((long long)hit) - 1 : 0;
      if (lane == wl) {
        vbuf[b][wave] = v;
        ibuf[b][wave] = row_id;
#pragma unroll
        for (int k = 0; k < kNB; ++k)
          if (k >= c) prow[b][wave][k] = T::v(a[k].re, a[k].im);
      }
      __syncthreads();
      R vv[4];
      int iv[4];
#pragma unroll
      for (int h = 0; h < 4; ++h) {
        vv[h] = vbuf[b][h];
        iv[h] = ibuf[b][h];
      }
      R pv = vv[0];
      int pw = 0, pi = iv[0];
#pragma unroll
      for (int h = 1; h < 4; ++h)
        if (vv[h] > pv) {
          pv = vv[h];
          pw = h;
          pi = iv[h];
        }
      // the pivot row (the same LDS address for every lane)
      V pr[kNB];
#pragma unroll
      for (int k = 0; k < kNB; ++k)
        if (k >= c) pr[k] = prow[b][pw][k];
      if (wave == 0 && lane == c) myord = pv >= 0.0 ? pi : -1;
      if (pv >= 0.0 && live && row_id == pi) {
        live = false;
        mystep = c;
      }
      if (pv >= 0.0 && live) {
        const C l = a[c] * recip_c(T::c(pr[c]));
        a[c] = l;
#pragma unroll
        for (int k = 0; k < kNB; ++k)
          if (k > c) {  // a[k] -= l pr[k]
            a[k].re = fma(-l.re, pr[k].x, fma(l.im, pr[k].y, a[k].re));
            a[k].im = fma(-l.re, pr[k].y, fma(-l.im, pr[k].x, a[k].im));
          }
      }
    }
  }
  if (wave == 0 && lane < kNB) cout[blockIdx.x * kNB + lane] = myord;
  if (gridDim.x == 1) {
    if (mystep >= 0) {
#pragma unroll
      for (int k = 0; k < kNB; ++k) F[k * kNB + mystep] = T::v(a[k].re, a[k].im);
    }
    if (wave == 0) {
      // interchanges: step c swaps row j + c with the current place of its pivot row; lane c2
      // follows where the pivot row of step c2 sits
      int pos = myord, mine = 0;
      for (int c = 0; c < jb; ++c) {
        const int t = j + c;
        const int pc = __builtin_amdgcn_readlane(pos, c);
        const int p = pc >= 0 ? pc : t;
        if (lane == c) mine = p;
        if (p != t && lane > c && pos == t) pos = p;
      }
      if (lane < jb) ipiv[j + lane] = mine;
    }
  }
}

// Write the factored diagonal block F (from the last tournament level) into A[j, j] (rows
// already interchanged) and store inv(L11), inv(U11) into inv (kNB x kNB each,
// column-major, identity beyond jb).  The inverses use recursive doubling: with the inverses
// X11, X22 of the diagonal s x s blocks of a 2s x 2s lower block, X21 = -X22 L21 X11 (upper:
// X12 = -X11 U12 X22), one element per thread per level.
template <class R>
__global__ void __launch_bounds__(1024) k_diag_finish(typename P<R>::V* __restrict__ A, long lda, int j, int jb,
                                                      const typename P<R>::V* __restrict__ F, typename P<R>::V* __restrict__ inv)
{
  using V = typename P<R>::V;
  using C = typename P<R>::C;
  using T = P<R>;
  __shared__ V Lx[kNB * kNB];
  __shared__ V Ux[kNB * kNB];
  __shared__ V Tm[kNB * kNB];
  const int tid = threadIdx.x, r = tid % kNB, c = tid / kNB;
  {
    const bool in = r < jb && c < jb;
    const V f = in ? F[c * kNB + r] : T::v(r == c ? 1.0 : 0.0, 0.0);
    if (in) A[(long)(j + c) * lda + j + r] = f;
    Lx[c * kNB + r] = (c < r) ? f : T::v(c == r ? 1.0 : 0.0, 0.0);
    C u = (c >= r) ? T::c(f) : T::mkc(0.0, 0.0);
    if (r == c) u = (u.re != 0.0 || u.im != 0.0) ? T::div(T::mkc(1.0, 0.0), u) : T::mkc(0.0, 0.0);
    Ux[c * kNB + r] = T::v(u);
  }
  __syncthreads();
  for (int s = 1; s < kNB; s <<= 1) {
    const int b = (r / (2 * s)) * (2 * s), bc = (c / (2 * s)) * (2 * s);
    const bool lo = r >= b + s && c >= b && c < b + s;    // lower, 2s block at b
    const bool up = c >= bc + s && r >= bc && r < bc + s;  // upper, 2s block at bc
    C t = T::mkc(0, 0);
    if (lo)
      for (int k = b; k < b + s; ++k) t = t + T::c(Lx[k * kNB + r]) * T::c(Lx[c * kNB + k]);  // L21 X11
    if (up)
      for (int k = bc + s; k < bc + 2 * s; ++k) t = t + T::c(Ux[k * kNB + r]) * T::c(Ux[c * kNB + k]);  // U12 X22
    if (lo || up) Tm[c * kNB + r] = T::v(t);
    __syncthreads();
    t = T::mkc(0, 0);
    if (lo)
      for (int k = b + s; k < b + 2 * s; ++k) t = t + T::c(Lx[k * kNB + r]) * T::c(Tm[c * kNB + k]);  // X22 T
    if (up)
      for (int k = bc; k < bc + s; ++k) t = t + T::c(Ux[k * kNB + r]) * T::c(Tm[c * kNB + k]);  // X11 T
    __syncthreads();
    if (lo) Lx[c * kNB + r] = T::v(-t.re, -t.im);
    if (up) Ux[c * kNB + r] = T::v(-t.re, -t.im);
    __syncthreads();
  }
  inv[c * kNB + r] = Lx[c * kNB + r];
  inv[kNB * kNB + c * kNB + r] = Ux[c * kNB + r];
}

// L21 = A21 inv(U11) in place for rows j+jb..n-1 of the panel; 64-row tiles in LDS
template <class R>
__global__ void __launch_bounds__(256) k_l21(int n, typename P<R>::V* __restrict__ A, long lda, int j, int jb,
                                             const typename P<R>::V* __restrict__ invU)
{
  using V = typename P<R>::V;
  using C = typename P<R>::C;
  using T = P<R>;
  __shared__ V Us[kNB * kNB];
  __shared__ V As[64 * kNB];
  const int tid = threadIdx.x;
  const int r0 = j + jb + blockIdx.x * 64;
  for (int q = tid; q < kNB * kNB; q += 256) Us[q] = invU[q];
  for (int q = tid; q < 64 * kNB; q += 256) {
    const int rr = q % 64, c = q / 64;
    As[q] = (r0 + rr < n && c < jb) ? A[(long)(j + c) * lda + r0 + rr] : T::v(0.0, 0.0);
  }
  __syncthreads();
  for (int q = tid; q < 64 * kNB; q += 256) {
    const int rr = q % 64, c = q / 64;
    if (r0 + rr >= n || c >= jb) continue;
    C acc = T::mkc(0, 0);
#pragma unroll 8
    for (int k = 0; k < kNB; ++k) acc = acc + T::c(As[k * 64 + rr]) * T::c(Us[c * kNB + k]);
    A[(long)(j + c) * lda + r0 + rr] = T::v(acc);
  }
}

// apply the interchanges ipiv[j..j+jb) to columns col0 .. col0+ncol-1
template <class R>
__global__ void __launch_bounds__(256) k_swap(typename P<R>::V* __restrict__ A, long lda, int col0, int ncol, int j, int jb,
                                              const int* __restrict__ ipiv)
{
  using V = typename P<R>::V;
  using C = typename P<R>::C;
  using T = P<R>;
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= ncol) return;
  V* cp = A + (long)(col0 + c) * lda;
  for (int r = j; r < j + jb; ++r) {
    const int p = ipiv[r];
    if (p != r) {
      const V t = cp[r];
      cp[r] = cp[p];
      cp[p] = t;
    }
  }
}

// B[0:jb, 0:ncol] <- Tm B for a kNB x kNB matrix Tm (column-major, identity beyond jb);
// 64 columns per workgroup, both operands staged in LDS.
template <class R>
__global__ void __launch_bounds__(256) k_apply_blk(const typename P<R>::V* __restrict__ Tm, int jb, typename P<R>::V* __restrict__ B,
                                                   long ldb, int ncol)
{
  using V = typename P<R>::V;
  using C = typename P<R>::C;
  using T = P<R>;
  __shared__ V Ts[kNB * kNB];
  __shared__ V Bs[kNB * 64];
  const int tid = threadIdx.x, col0 = blockIdx.x * 64;
  for (int q = tid; q < kNB * kNB; q += 256) Ts[q] = Tm[q];
  for (int q = tid; q < kNB * 64; q += 256) {
    const int r = q % kNB, c = q / kNB;
    Bs[q] = (r < jb && col0 + c < ncol) ? B[(long)(col0 + c) * ldb + r] : T::v(0.0, 0.0);
  }
  __syncthreads();
  for (int q = tid; q < kNB * 64; q += 256) {
    const int r = q % kNB, c = q / kNB;
    if (r >= jb || col0 + c >= ncol) continue;
    C acc = T::mkc(0, 0);
#pragma unroll 8
    for (int k = 0; k < kNB; ++k) acc = acc + T::c(Ts[k * kNB + r]) * T::c(Bs[c * kNB + k]);
    B[(long)(col0 + c) * ldb + r] = T::v(acc);
  }
}

// k_apply_blk on the diagonal blocks of several column panels at once: block y applies
// Tm + y * tstride to B + y * bstride (same jb, ldb and ncol for all)
template <class R>
__global__ void __launch_bounds__(256) k_apply_blk_batched(const typename P<R>::V* __restrict__ Tm, long tstride,
                                                           int jb, typename P<R>::V* __restrict__ B, long bstride,
                                                           long ldb, int ncol)
{
  using V = typename P<R>::V;
  using C = typename P<R>::C;
  using T = P<R>;
  __shared__ V Ts[kNB * kNB];
  __shared__ V Bs[kNB * 64];
  const int tid = threadIdx.x, col0 = blockIdx.x * 64;
  const V* tm = Tm + blockIdx.y * tstride;
  V* b = B + blockIdx.y * bstride;
  for (int q = tid; q < kNB * kNB; q += 256) Ts[q] = tm[q];
  for (int q = tid; q < kNB * 64; q += 256) {
    const int r = q % kNB, c = q / kNB;
    Bs[q] = (r < jb && col0 + c < ncol) ? b[(long)(col0 + c) * ldb + r] : T::v(0.0, 0.0);
  }
  __syncthreads();
  for (int q = tid; q < kNB * 64; q += 256) {
    const int r = q % kNB, c = q / kNB;
    if (r >= jb || col0 + c >= ncol) continue;
    C acc = T::mkc(0, 0);
#pragma unroll 8
    for (int k = 0; k < kNB; ++k) acc = acc + T::c(Ts[k * kNB + r]) * T::c(Bs[c * kNB + k]);
    b[(long)(col0 + c) * ldb + r] = T::v(acc);
  }
}

// W[b] = -I for nb square blocks of order m (column-major, contiguous)
template <class R>
__global__ void k_neg_eye(long total, int m, typename P<R>::V* __restrict__ W)
{
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= total) return;
  const long e = i % ((long)m * m);
  W[i] = P<R>::v((e % m) == (e / m) ? -1.0 : 0.0, 0.0);
}

// the interchanges ipiv[0..n) applied in order to 0..n-1 give the row permutation
__global__ void k_ipiv_perm(int n, const int* __restrict__ ipiv, int* __restrict__ perm)
{
  if (blockIdx.x != 0 || threadIdx.x != 0) return;
  for (int r = 0; r < n; ++r) perm[r] = r;
  for (int r = 0; r < n; ++r) {
    const int p = ipiv[r];
    if (p != r) {
      const int t = perm[r];
      perm[r] = perm[p];
      perm[p] = t;
    }
  }
}

// Bt[r, c] = B[perm[r], c]
template <class R>
__global__ void __launch_bounds__(256) k_gather_rows(int n, int nrhs, const int* __restrict__ perm,
                                                     const typename P<R>::V* __restrict__ B, long ldb, typename P<R>::V* __restrict__ Bt)
{
  using V = typename P<R>::V;
  using C = typename P<R>::C;
  using T = P<R>;
  const long q = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= (long)n * nrhs) return;
  const int r = (int)(q % n), c = (int)(q / n);
  Bt[(long)c * n + r] = B[(long)c * ldb + perm[r]];
}

template <class R>
__global__ void k_identity(int n, typename P<R>::V* __restrict__ X, long ldx)
{
  const long q = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= (long)n * n) return;
  const long c = q / n, r = q - c * n;
  X[c * ldx + r] = P<R>::v(r == c ? 1.0 : 0.0, 0.0);
}

template <class R>
__global__ void k_check_diag(int n, const typename P<R>::V* __restrict__ A, long lda, int* __restrict__ info)
{
  using V = typename P<R>::V;
  using C = typename P<R>::C;
  using T = P<R>;
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  V d = A[(long)i * lda + i];
  // an exactly zero pivot, or one that is Inf / NaN (a non-finite matrix entry upstream)
  if ((d.x == 0.0 && d.y == 0.0) || !isfinite((double)d.x) || !isfinite((double)d.y)) atomicMin(info, i + 1);
}

// the row permutation getrf leaves after the inverted diagonal blocks
template <class V>
int* perm_of(V* inv, int n)
{
  return (int*)(inv + (long)((n + kNB - 1) / kNB) * kBlkElems);
}

// ---- solves with few right-hand sides (nrhs <= kSmallRhs) --------------------------------
// The blocked solve below issues two launches per 32-row step; with a handful of columns each
// launch does almost nothing and the solve is bound by launch latency.  Here one workgroup
// runs all 32-row steps of a 256-row diagonal block with its rows of B in LDS, and one kernel
// applies the block's columns of L (or U) to the rows outside it, reading them once.
constexpr int kSmallRhs = 8;
constexpr int kPermChunk = 512;  // columns per pass of the in-place row permutation
constexpr int kSolveGrp = 4;     // 512-row blocks per rank-2048 update of the many-column solves

// rows [j0, je) of B (nrhs columns): LOWER: X = inv(L_blk) X (unit lower, steps downwards);
// else X = inv(U_blk) X (steps upwards).  inv: the 32x32 inverses getrf stored.
template <class R, bool LOWER>
__global__ void __launch_bounds__(256) k_trsv_blk(int n, const typename P<R>::V* __restrict__ A, long lda,
                                                  const typename P<R>::V* __restrict__ inv,
                                                  typename P<R>::V* __restrict__ B, long ldb, int nrhs, int j0)
{
  using V = typename P<R>::V;
  using C = typename P<R>::C;
  using T = P<R>;
  __shared__ V xs[kSmallRhs * kNB2];
  __shared__ V ts[kNB * kNB];
  const int tid = threadIdx.x, je = min(n, j0 + kNB2), nb = je - j0;
  for (int q = tid; q < nb * nrhs; q += 256) {
    const int r = q % nb, c = q / nb;
    xs[c * kNB2 + r] = B[(long)c * ldb + j0 + r];
  }
  const int nst = (nb + kNB - 1) / kNB;
  for (int t = 0; t < nst; ++t) {
    const int st = LOWER ? t : nst - 1 - t;
    const int j = j0 + st * kNB, jb = min(kNB, je - j);
    const V* Tm = inv + (long)(j / kNB) * kBlkElems + (LOWER ? 0 : kNB * kNB);
    // this step's inverse to LDS and this thread's row of the step's columns to registers,
    // all loads issued before any is used (the row update has at most 224 rows: one per thread)
    const int lo = LOWER ? j + jb : j0, hi = LOWER ? je : j;
    const int rr = lo + tid;
    const bool rowact = rr < hi;
    C a[kNB];
#pragma unroll
    for (int k = 0; k < kNB; ++k) a[k] = (rowact && k < jb) ? T::c(A[(long)(j + k) * lda + rr]) : T::mkc(0, 0);
#pragma unroll
    for (int q = 0; q < kNB * kNB / 256; ++q) ts[q * 256 + tid] = Tm[q * 256 + tid];
    __syncthreads();
    const int r = tid % kNB, c = tid / kNB;
    const bool act = c < nrhs && r < jb;
    C acc = T::mkc(0, 0);
    if (act)
#pragma unroll 8
      for (int k = 0; k < jb; ++k) acc = acc + T::c(ts[k * kNB + r]) * T::c(xs[c * kNB2 + (j - j0) + k]);
    __syncthreads();
    if (act) xs[c * kNB2 + (j - j0) + r] = T::v(acc);
    __syncthreads();
    if (rowact) {
#pragma unroll
      for (int cc = 0; cc < kSmallRhs; ++cc) {
        if (cc >= nrhs) continue;
        C v = T::c(xs[cc * kNB2 + (rr - j0)]);
#pragma unroll
        for (int k = 0; k < kNB; ++k) v = v - a[k] * T::c(xs[cc * kNB2 + (j - j0) + min(k, jb - 1)]);
        xs[cc * kNB2 + (rr - j0)] = T::v(v);
      }
    }
    __syncthreads();
  }
  for (int q = tid; q < nb * nrhs; q += 256) {
    const int r = q % nb, c = q / nb;
    B[(long)c * ldb + j0 + r] = xs[c * kNB2 + r];
  }
}

// Bo[r, c] -= sum_k A[r, k] X[k, c] for r < m, k < kk (<= 256), c < nrhs.  64 rows per
// workgroup, the k range split over 4 wavefronts and reduced in LDS.  X and Bo are disjoint.
template <class R>
__global__ void __launch_bounds__(256) k_skinny_update(int m, int kk, const typename P<R>::V* __restrict__ A, long lda,
                                                       const typename P<R>::V* __restrict__ X, long ldx, int nrhs,
                                                       typename P<R>::V* __restrict__ Bo, long ldb)
{
  using V = typename P<R>::V;
  using C = typename P<R>::C;
  using T = P<R>;
  __shared__ V xs[kSmallRhs * kNB2];
  __shared__ V part[3 * 64 * kSmallRhs];
  const int tid = threadIdx.x, lane = tid & 63, q = tid >> 6;
  // columns past nrhs staged as zeros: no per-column branch in the inner loop, so the loads of
  // A are issued 16 at a time
  for (int e = tid; e < kk * kSmallRhs; e += 256) {
    const int k = e % kk, c = e / kk;
    xs[c * kNB2 + k] = c < nrhs ? X[(long)c * ldx + k] : T::v(T::mkc(0, 0));
  }
  __syncthreads();
  const int r = blockIdx.x * 64 + lane;
  C acc[kSmallRhs];
#pragma unroll
  for (int cc = 0; cc < kSmallRhs; ++cc) acc[cc] = T::mkc(0, 0);
  if (r < m) {
    const int k0 = q * 64, k1 = min(kk, k0 + 64);
    int k = k0;
    for (; k + 16 <= k1; k += 16) {
      C a[16];
#pragma unroll
      for (int u = 0; u < 16; ++u) a[u] = T::c(A[(long)(k + u) * lda + r]);
#pragma unroll
      for (int u = 0; u < 16; ++u)
#pragma unroll
        for (int cc = 0; cc < kSmallRhs; ++cc) acc[cc] = acc[cc] + a[u] * T::c(xs[cc * kNB2 + k + u]);
    }
    for (; k < k1; ++k) {
      const C a = T::c(A[(long)k * lda + r]);
#pragma unroll
      for (int cc = 0; cc < kSmallRhs; ++cc) acc[cc] = acc[cc] + a * T::c(xs[cc * kNB2 + k]);
    }
  }
  if (q > 0) {
#pragma unroll
    for (int cc = 0; cc < kSmallRhs; ++cc) part[((q - 1) * kSmallRhs + cc) * 64 + lane] = T::v(acc[cc]);
  }
  __syncthreads();
  if (q == 0 && r < m) {
#pragma unroll
    for (int cc = 0; cc < kSmallRhs; ++cc) {
      if (cc >= nrhs) continue;
      C t = acc[cc];
      for (int w = 0; w < 3; ++w) t = t + T::c(part[(w * kSmallRhs + cc) * 64 + lane]);
      V* o = Bo + (long)cc * ldb + r;
      *o = T::v(T::c(*o) - t);
    }
  }
}

// Bo[r, c] -= sum_k A[r, k] X[k, c] (STORE: Bo = -sum) for r < m, k < kk <= kNBF, c < nrhs: the
// products of the solves through the inverted 512-row diagonal blocks.  X (kk x nrhs) is staged
// in LDS once per workgroup; 64 rows per workgroup, the k range split over its 4 wavefronts
// (16 loads of A in flight per lane) and reduced in LDS.  X and Bo are disjoint.
template <class R, bool STORE>
__global__ void __launch_bounds__(256) k_skinny_blk(int m, int kk, const typename P<R>::V* __restrict__ A, long lda,
                                                    const typename P<R>::V* __restrict__ X, long ldx, int nrhs,
                                                    typename P<R>::V* __restrict__ Bo, long ldb)
{
  using V = typename P<R>::V;
  using C = typename P<R>::C;
  using T = P<R>;
  __shared__ V xs[kSmallRhs * kNBF];
  __shared__ V part[3 * 64 * kSmallRhs];
  const int tid = threadIdx.x, lane = tid & 63, q = tid >> 6;
  // columns past nrhs staged as zeros, so that the inner loop has no per-column branch and the
  // loads of A can be issued together
  for (int e = tid; e < kk * kSmallRhs; e += 256) {
    const int k = e % kk, c = e / kk;
    xs[c * kNBF + k] = c < nrhs ? X[(long)c * ldx + k] : T::v(T::mkc(0, 0));
  }
  __syncthreads();
  const int r = blockIdx.x * 64 + lane;
  C acc[kSmallRhs];
#pragma unroll
  for (int cc = 0; cc < kSmallRhs; ++cc) acc[cc] = T::mkc(0, 0);
  if (r < m) {
    const int kq = (kk + 3) / 4, k0 = q * kq, k1 = min(kk, k0 + kq);
    int k = k0;
    for (; k + 16 <= k1; k += 16) {
      C a[16];
#pragma unroll
      for (int u = 0; u < 16; ++u) a[u] = T::c(A[(long)(k + u) * lda + r]);
#pragma unroll
      for (int u = 0; u < 16; ++u)
#pragma unroll
        for (int cc = 0; cc < kSmallRhs; ++cc) acc[cc] = acc[cc] + a[u] * T::c(xs[cc * kNBF + k + u]);
    }
    for (; k < k1; ++k) {
      const C a = T::c(A[(long)k * lda + r]);
#pragma unroll
      for (int cc = 0; cc < kSmallRhs; ++cc) acc[cc] = acc[cc] + a * T::c(xs[cc * kNBF + k]);
    }
  }
  if (q > 0) {
#pragma unroll
    for (int cc = 0; cc < kSmallRhs; ++cc) part[((q - 1) * kSmallRhs + cc) * 64 + lane] = T::v(acc[cc]);
  }
  __syncthreads();
  if (q == 0 && r < m) {
#pragma unroll
    for (int cc = 0; cc < kSmallRhs; ++cc) {
      if (cc >= nrhs) continue;
      C t = acc[cc];
      for (int w = 0; w < 3; ++w) t = t + T::c(part[(w * kSmallRhs + cc) * 64 + lane]);
      V* o = Bo + (long)cc * ldb + r;
      *o = STORE ? T::v(T::mkc(0, 0) - t) : T::v(T::c(*o) - t);
    }
  }
}

// Bo[r, c] -= sum_k A[r, k] X[k, c] for r < m, k < kk (any length), c < nrhs <= kSmallRhs:
// a matrix-times-few-vectors product that reads A once.  64 rows per workgroup, the k range
// split over its 4 wavefronts (reduced in LDS); X is read at the same address by every lane.
template <class R>
__global__ void __launch_bounds__(256) k_skinny_gemm(int m, long kk, const typename P<R>::V* __restrict__ A, long lda,
                                                     const typename P<R>::V* __restrict__ X, long ldx, int nrhs,
                                                     typename P<R>::V* __restrict__ Bo, long ldb)
{
  using V = typename P<R>::V;
  using C = typename P<R>::C;
  using T = P<R>;
  __shared__ V part[3 * 64 * kSmallRhs];
  const int tid = threadIdx.x, lane = tid & 63, q = tid >> 6;
  const int r = blockIdx.x * 64 + lane;
  C acc[kSmallRhs];
#pragma unroll
  for (int cc = 0; cc < kSmallRhs; ++cc) acc[cc] = T::mkc(0, 0);
  const long kq = (kk + 3) / 4, k0 = q * kq, k1 = min(kk, k0 + kq);
  if (r < m) {
#pragma unroll 8
    for (long k = k0; k < k1; ++k) {
      const C a = T::c(A[k * lda + r]);
#pragma unroll
      for (int cc = 0; cc < kSmallRhs; ++cc)
        if (cc < nrhs) acc[cc] = acc[cc] + a * T::c(X[(long)cc * ldx + k]);
    }
  }
  if (q > 0) {
#pragma unroll
    for (int cc = 0; cc < kSmallRhs; ++cc) part[((q - 1) * kSmallRhs + cc) * 64 + lane] = T::v(acc[cc]);
  }
  __syncthreads();
  if (q == 0 && r < m) {
#pragma unroll
    for (int cc = 0; cc < kSmallRhs; ++cc) {
      if (cc >= nrhs) continue;
      C t = acc[cc];
      for (int w = 0; w < 3; ++w) t = t + T::c(part[(w * kSmallRhs + cc) * 64 + lane]);
      V* o = Bo + (long)cc * ldb + r;
      *o = T::v(T::c(*o) - t);
    }
  }
}

// One rocBLAS handle per (device, stream), created on first use and kept for the process:
// sessions on different streams (concurrent requests) never share a handle, so no caller
// changes the stream of a handle another thread is using.
struct HandleSlot {
  int dev;
  hipStream_t s;
  rocblas_handle h;
};
std::mutex g_hmu;
std::vector<HandleSlot> g_handles;

rocblas_handle handle_for(hipStream_t s)
{
  int dev = 0;
  (void)hipGetDevice(&dev);
  std::lock_guard<std::mutex> lk(g_hmu);
  for (const HandleSlot& q : g_handles)
    if (q.dev == dev && q.s == s) return q.h;
  rocblas_handle h = nullptr;
  if (rocblas_create_handle(&h) != rocblas_status_success) return nullptr;
  rocblas_set_stream(h, s);
  g_handles.push_back(HandleSlot{dev, s, h});
  return h;
}

// The look-ahead stream of getrf per (device, stream): a second stream at the default priority
// with two events (panel factored on the main stream; right-hand update done on this one),
// created on first use and kept for the process.
struct SideSlot {
  int dev;
  hipStream_t s;
  hipStream_t side;
  hipEvent_t ev[2];
};
std::vector<SideSlot> g_side;

bool side_for(hipStream_t s, SideSlot* out)
{
  int dev = 0;
  (void)hipGetDevice(&dev);
  std::lock_guard<std::mutex> lk(g_hmu);
  for (const SideSlot& q : g_side)
    if (q.dev == dev && q.s == s) {
      *out = q;
      return true;
    }
  SideSlot q{dev, s, nullptr, {nullptr, nullptr}};
  if (hipStreamCreateWithFlags(&q.side, hipStreamNonBlocking) != hipSuccess) return false;
  for (hipEvent_t& e : q.ev)
    if (hipEventCreateWithFlags(&e, hipEventDisableTiming) != hipSuccess) {
      for (hipEvent_t& f : q.ev)
        if (f) (void)hipEventDestroy(f);
      (void)hipStreamDestroy(q.side);
      return false;
    }
  g_side.push_back(q);
  *out = q;
  return true;
}

// Look-ahead of the factorization (1, default; BQ_LU_LOOKAHEAD=0 sets 0 at start-up)
std::atomic<int> g_lookahead{-1};

int lookahead_mode()
{
  int m = g_lookahead.load();
  if (m < 0) {
    const char* e = getenv("BQ_LU_LOOKAHEAD");
    m = (e && strcmp(e, "0") == 0) ? 0 : 1;
    g_lookahead.store(m);
  }
  return m;
}

template <class R>
int gemm_op(hipStream_t s, int ta, int tb, int m, int n, int k, typename P<R>::V alpha, const typename P<R>::V* A,
            long lda, const typename P<R>::V* B, long ldb, typename P<R>::V beta, typename P<R>::V* C, long ldc)
{
  if (m <= 0 || n <= 0 || k <= 0) return 0;
  rocblas_handle h = handle_for(s);
  if (!h) {
    bq::set_error("rocblas handle creation failed");
    return -1;
  }
  rocblas_status st = P<R>::gemm(h, ta ? rocblas_operation_transpose : rocblas_operation_none,
                                 tb ? rocblas_operation_transpose : rocblas_operation_none, m, n, k, &alpha, A,
                                 (int)lda, B, (int)ldb, &beta, C, (int)ldc);
  if (st != rocblas_status_success) {
    bq::set_error("rocblas complex gemm failed (%d)", (int)st);
    return -1;
  }
  return 0;
}

// Updates C -= A B of the FP32 factorization and solves of order >= kSplitMin in every
// dimension go to the bf16x3 matrix-core kernel (bq_mfma.hip) unless the update mode is 0.
constexpr int kSplitMin = 128;
std::atomic<int> g_update_mode{-1};

int update_mode()
{
  int m = g_update_mode.load();
  if (m < 0) {
    const char* e = getenv("BQ_LU_UPDATE");
    m = (e && strcmp(e, "fp32") == 0) ? 0 : 1;
    g_update_mode.store(m);
  }
  return m;
}

template <class R>
int gemm(hipStream_t s, int m, int n, int k, typename P<R>::V alpha, const typename P<R>::V* A, long lda,
         const typename P<R>::V* B, long ldb, typename P<R>::V beta, typename P<R>::V* C, long ldc)
{
  if constexpr (std::is_same<R, float>::value) {
    if (alpha.x == -1.f && alpha.y == 0.f && beta.x == 1.f && beta.y == 0.f && m >= kSplitMin && n >= kSplitMin &&
        k >= kSplitMin && update_mode() == 1)
      return bq_cgemm_split_dev(s, m, n, k, (const float*)A, lda, (const float*)B, ldb, (float*)C, ldc);
  }
  return gemm_op<R>(s, 0, 0, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc);
}

// -inv(T_bb) of nfull consecutive diagonal blocks of order kNBF of the factored A: the unit
// lower factor (upper = false) or the upper one, into W (nfull blocks of kNBF x kNBF,
// column-major, contiguous).  Built for all blocks at once by the in-block substitution of
// the solves (32-row steps with the inverted 32 x 32 diagonal blocks), starting from -I:
// 16 batched launches of each kind in all.
template <class R>
int block_inverses(hipStream_t s, int nfull, const typename P<R>::V* A, long lda, const typename P<R>::V* inv,
                   bool upper, typename P<R>::V* W)
{
  using V = typename P<R>::V;
  const V one = P<R>::v(1.0, 0.0), mone = P<R>::v(-1.0, 0.0);
  const long bs = (long)kNBF * kNBF, tot = bs * nfull;
  k_neg_eye<R><<<(unsigned)((tot + 255) / 256), 256, 0, s>>>(tot, kNBF, W);
  BQ_HIP(hipGetLastError());
  rocblas_handle h = handle_for(s);
  if (!h) {
    bq::set_error("rocblas handle creation failed");
    return -1;
  }
  const int steps = kNBF / kNB;
  const dim3 g(kNBF / 64, nfull);
  const long tstride = (long)steps * kBlkElems, astride = (long)kNBF * lda + kNBF;
  for (int q = 0; q < steps; ++q) {
    const int t = upper ? steps - 1 - q : q, j = t * kNB;
    k_apply_blk_batched<R><<<g, 256, 0, s>>>(inv + (long)t * kBlkElems + (upper ? kNB * kNB : 0), tstride, kNB, W + j,
                                             bs, kNBF, kNBF);
    BQ_HIP(hipGetLastError());
    rocblas_status st = rocblas_status_success;
    if (!upper && kNBF - j - kNB > 0)
      st = P<R>::gemm_sb(h, kNBF - j - kNB, kNBF, kNB, &mone, A + (long)j * lda + (j + kNB), (int)lda, astride, W + j,
                         kNBF, bs, &one, W + j + kNB, kNBF, bs, nfull);
    if (upper && j > 0)
      st = P<R>::gemm_sb(h, j, kNBF, kNB, &mone, A + (long)j * lda, (int)lda, astride, W + j, kNBF, bs, &one, W, kNBF,
                         bs, nfull);
    if (st != rocblas_status_success) {
      bq::set_error("rocblas strided batched gemm failed (%d)", (int)st);
      return -1;
    }
  }
  return 0;
}

// X (kNBF x nc, ldx) <- inv(T_bb) X = -Wneg X in place on the matrix cores (Wneg = -inv(T_bb))
inline int apply_block_inverse(hipStream_t s, int nc, const float2* Wneg, float2* X, long ldx)
{
  return bq_cgemm_split_store_dev(s, kNBF, nc, kNBF, (const float*)Wneg, kNBF, (const float*)X, ldx, (float*)X, ldx);
}

// Solve A X = B with at most kSmallRhs right-hand sides through the inverted 512-row diagonal
// blocks of both factors (W: nfull blocks -inv(L_bb), then nfull blocks -inv(U_bb), from
// block_inverses; n a multiple of kNBF).  Each block is one product with its inverse and one
// product of the factor's block column with the rows outside it, so a triangle takes 2 n / 512
// launches that each read their part of the factor once, where the 256-row substitution of the
// few-column getrs takes 2 n / 256 launches bound by their 32-row step chains.  The right-hand
// sides alternate between B and a work copy T: the forward sweep reads T and writes the
// block solutions into B, the backward sweep reads B and writes into T.
template <class R>
int getrs_blkinv(hipStream_t s, int n, int nrhs, const typename P<R>::V* A, long lda, const typename P<R>::V* inv,
                 const typename P<R>::V* W, typename P<R>::V* B, long ldb)
{
  using V = typename P<R>::V;
  if (n <= 0 || nrhs <= 0) return 0;
  if (nrhs > kSmallRhs || n % kNBF != 0) {
    bq::set_error("getrs_blkinv: %d right-hand sides, order %d (at most %d, a multiple of %d)", nrhs, n, kSmallRhs,
                  kNBF);
    return -1;
  }
  const int kt_all = bq::ktime_begin(P<R>::kt_s, s);
  const int nfull = n / kNBF;
  const long bs = (long)kNBF * kNBF;
  const V* WL = W;
  const V* WU = W + bs * nfull;
  const int* perm = perm_of(inv, n);
  V* T = nullptr;
  BQ_HIP(hipMallocAsync((void**)&T, sizeof(V) * (size_t)n * nrhs, s));
  const long tot = (long)n * nrhs;
  k_gather_rows<R><<<(unsigned)((tot + 255) / 256), 256, 0, s>>>(n, nrhs, perm, B, ldb, T);
  BQ_HIP(hipGetLastError());
  const unsigned gb = kNBF / 64;
  for (int b = 0; b < nfull; ++b) {
    const int j0 = b * kNBF, je = j0 + kNBF;
    k_skinny_blk<R, true><<<gb, 256, 0, s>>>(kNBF, kNBF, WL + bs * b, kNBF, T + j0, n, nrhs, B + j0, ldb);
    if (n - je > 0)
      k_skinny_blk<R, false><<<(unsigned)((n - je + 63) / 64), 256, 0, s>>>(n - je, kNBF, A + (long)j0 * lda + je,
                                                                            lda, B + j0, ldb, nrhs, T + je, n);
    BQ_HIP(hipGetLastError());
  }
  for (int b = nfull - 1; b >= 0; --b) {
    const int j0 = b * kNBF;
    k_skinny_blk<R, true><<<gb, 256, 0, s>>>(kNBF, kNBF, WU + bs * b, kNBF, B + j0, ldb, nrhs, T + j0, n);
    if (j0 > 0)
      k_skinny_blk<R, false><<<(unsigned)((j0 + 63) / 64), 256, 0, s>>>(j0, kNBF, A + (long)j0 * lda, lda, T + j0, n,
                                                                         nrhs, B, ldb);
    BQ_HIP(hipGetLastError());
  }
  BQ_HIP(hipMemcpy2DAsync(B, ldb * sizeof(V), T, (size_t)n * sizeof(V), (size_t)n * sizeof(V), nrhs,
                          hipMemcpyDeviceToDevice, s));
  BQ_HIP(hipFreeAsync(T, s));
  bq::ktime_end(kt_all, s);
  return 0;
}

// The factored outer panel j0 .. j0+jb2-1 applied to the columns c0 .. c0+w-1 right of it,
// on stream st: the panel's interchanges, U12 = inv(L11) A12 (one product with
// Wneg = -inv(L11) when given and w >= kSplitMin, else 32-row steps) and the trailing
// update A22 -= L21 U12.
template <class R>
int panel_right(hipStream_t st, int n, typename P<R>::V* A, long lda, const int* ipiv, const typename P<R>::V* inv,
                int j0, int jb2, int c0, int w, const typename P<R>::V* Wneg)
{
  using V = typename P<R>::V;
  const V one = P<R>::v(1.0, 0.0), mone = P<R>::v(-1.0, 0.0);
  const int je = j0 + jb2;
  V* Ar = A + (long)c0 * lda;
  k_swap<R><<<(w + 255) / 256, 256, 0, st>>>(A, lda, c0, w, j0, jb2, ipiv);
  BQ_HIP(hipGetLastError());
  bool done = false;
  if constexpr (std::is_same<R, float>::value) {
    if (Wneg && w >= kSplitMin) {
      if (apply_block_inverse(st, w, Wneg, Ar + j0, lda)) return -1;
      done = true;
    }
  }
  for (int j = j0; j < je && !done; j += kNB) {
    const int jb = min(kNB, je - j);
    k_apply_blk<R><<<(w + 63) / 64, 256, 0, st>>>(inv + (long)(j / kNB) * kBlkElems, jb, Ar + j, lda, w);
    BQ_HIP(hipGetLastError());
    if (je - j - jb > 0 &&
        gemm<R>(st, je - j - jb, w, jb, mone, A + (long)j * lda + (j + jb), lda, Ar + j, lda, one, Ar + j + jb, lda))
      return -1;
  }
  int rc = 0;
  if (n - je > 0)
    BQ_KTIME(P<R>::kt_u, st,
             rc = gemm<R>(st, n - je, w, jb2, mone, A + (long)j0 * lda + je, lda, Ar + j0, lda, one, Ar + je, lda));
  return rc ? -1 : 0;
}

// In-place LU of the n x n column-major matrix A (device); inv receives the inverted
// diagonal blocks.  Returns 0, a positive k if U(k,k) is exactly zero (1-based), or < 0.
// With the look-ahead (n > 1024), the columns right of each outer panel are split: the next
// panel's 512 columns are brought up to date on s, ahead of the rest, which the look-ahead
// stream updates while s already factors that next panel; the latency-bound 32-column steps
// of one panel thus overlap the matrix-core update of the previous one.  Before the next
// panel's interchanges reach the columns left of it (whose L21 the update reads) and its
// own right-hand work starts, s waits for that update.
template <class R>
int getrf(hipStream_t s, int n, typename P<R>::V* A, long lda, int* ipiv, typename P<R>::V* inv)
{
  using V = typename P<R>::V;
  const V one = P<R>::v(1.0, 0.0), mone = P<R>::v(-1.0, 0.0);
  const int nleaf = (n + kTsRows - 1) / kTsRows;
  SideSlot side{};
  const bool la = lookahead_mode() == 1 && n > 2 * kNBF;
  if (la && !side_for(s, &side)) {
    bq::set_error("getrf: look-ahead stream creation failed");
    return -1;
  }
  int* cand = nullptr;
  V* F = nullptr;
  BQ_HIP(hipMallocAsync((void**)&cand, sizeof(int) * (size_t)2 * nleaf * kNB, s));
  BQ_HIP(hipMallocAsync((void**)&F, sizeof(V) * kNB * kNB, s));
  // FP32 with the matrix-core updates: U12 of each full panel by one product with the
  // panel's inverted L11 (-inv(L11) in Wp; two alternating buffers with the look-ahead,
  // whose stream may still read the previous panel's)
  bool blk_inv = false;
  V* Wp = nullptr;
  const long wsz = (long)kNBF * kNBF;
  if constexpr (std::is_same<R, float>::value) {
    if (update_mode() == 1 && n >= 2 * kNBF) {
      blk_inv = true;
      BQ_HIP(hipMallocAsync((void**)&Wp, sizeof(V) * wsz * (la ? 2 : 1), s));
    }
  }
  const int kt_all = bq::ktime_begin(P<R>::kt_f, s);
  if (la) {  // the look-ahead stream starts from A as s has it
    BQ_HIP(hipEventRecord(side.ev[0], s));
    BQ_HIP(hipStreamWaitEvent(side.side, side.ev[0], 0));
  }
  bool pend = false;  // a right-hand update outstanding on the look-ahead stream
  for (int j0 = 0, pi = 0; j0 < n; j0 += kNBF, ++pi) {
    const int jb2 = min(kNBF, n - j0), je = j0 + jb2;
    const int kt_p = bq::ktime_begin(P<R>::kt_p, s);
    for (int j = j0; j < je; j += kNB) {
      const int jb = min(kNB, je - j);
      V* invj = inv + (long)(j / kNB) * kBlkElems;
      // pivots of the step by tournament; the last level also factors the diagonal block
      int nset = (n - j + kTsRows - 1) / kTsRows;
      int* cur = cand;
      int* nxt = cand + (size_t)nleaf * kNB;
      k_tslu<R><<<nset, kTsRows, 0, s>>>(0, n, A, lda, j, jb, nullptr, 0, cur, F, ipiv);
      while (nset > 1) {
        const int nout = (nset + kTsFan - 1) / kTsFan;
        k_tslu<R><<<nout, kTsRows, 0, s>>>(1, n, A, lda, j, jb, cur, nset, nxt, F, ipiv);
        int* t = cur;
        cur = nxt;
        nxt = t;
        nset = nout;
      }
      k_swap<R><<<(jb2 + 255) / 256, 256, 0, s>>>(A, lda, j0, jb2, j, jb, ipiv);
      k_diag_finish<R><<<1, 1024, 0, s>>>(A, lda, j, jb, F, invj);
      const int rest = n - j - jb, wide = je - j - jb;
      if (rest > 0) k_l21<R><<<(rest + 63) / 64, 256, 0, s>>>(n, A, lda, j, jb, invj + kNB * kNB);
      BQ_HIP(hipGetLastError());
      if (wide > 0) {  // the rest of the outer panel
        k_apply_blk<R><<<(wide + 63) / 64, 256, 0, s>>>(invj, jb, A + (long)(j + jb) * lda + j, lda, wide);
        BQ_HIP(hipGetLastError());
        if (gemm<R>(s, rest, wide, jb, mone, A + (long)j * lda + (j + jb), lda, A + (long)(j + jb) * lda + j, lda, one,
                    A + (long)(j + jb) * lda + (j + jb), lda))
          return -1;
      }
    }
    if (pend) {
      BQ_HIP(hipStreamWaitEvent(s, side.ev[1], 0));
      pend = false;
    }
    // the panel's interchanges on the columns left of it
    if (j0 > 0) k_swap<R><<<(j0 + 255) / 256, 256, 0, s>>>(A, lda, 0, j0, j0, jb2, ipiv);
    BQ_HIP(hipGetLastError());
    const int right = n - je;
    V* Wcur = nullptr;
    if (right > 0 && blk_inv && jb2 == kNBF && right >= kSplitMin) {
      Wcur = Wp + (la ? (pi & 1) * wsz : 0);
      if (block_inverses<R>(s, 1, A + (long)j0 * lda + j0, lda, inv + (long)(j0 / kNB) * kBlkElems, false, Wcur))
        return -1;
    }
    bq::ktime_end(kt_p, s);
    const int r1 = la ? min(right, kNBF) : right, r2 = right - r1;
    if (r2 > 0) {
      BQ_HIP(hipEventRecord(side.ev[0], s));
      BQ_HIP(hipStreamWaitEvent(side.side, side.ev[0], 0));
      if (panel_right<R>(side.side, n, A, lda, ipiv, inv, j0, jb2, je + r1, r2, Wcur)) return -1;
      BQ_HIP(hipEventRecord(side.ev[1], side.side));
      pend = true;
    }
    if (r1 > 0 && panel_right<R>(s, n, A, lda, ipiv, inv, j0, jb2, je, r1, Wcur)) return -1;
  }
  if (pend) BQ_HIP(hipStreamWaitEvent(s, side.ev[1], 0));
  k_ipiv_perm<<<1, 1, 0, s>>>(n, ipiv, perm_of(inv, n));
  BQ_HIP(hipGetLastError());
  bq::ktime_end(kt_all, s);
  BQ_HIP(hipFreeAsync(cand, s));
  BQ_HIP(hipFreeAsync(F, s));
  if (Wp) BQ_HIP(hipFreeAsync(Wp, s));
  int* dinfo = nullptr;
  BQ_HIP(hipMallocAsync((void**)&dinfo, sizeof(int), s));
  int big = 0x7fffffff;
  BQ_HIP(hipMemcpyAsync(dinfo, &big, sizeof(int), hipMemcpyHostToDevice, s));
  k_check_diag<R><<<(n + 255) / 256, 256, 0, s>>>(n, A, lda, dinfo);
  int info = 0;
  BQ_HIP(hipMemcpyAsync(&info, dinfo, sizeof(int), hipMemcpyDeviceToHost, s));
  BQ_HIP(hipFreeAsync(dinfo, s));
  BQ_HIP(hipStreamSynchronize(s));
  return info == big ? 0 : info;
}

// Solve A X = B with the factors from getrf; B is n x nrhs column-major (device).
// With ident, B holds the n x n identity, unpermuted: the solve yields U^-1 L^-1 (the inverse
// with its columns permuted, A^-1[:, perm[i]] = X[:, i]), and the forward sweep of block row
// j0 .. je-1 touches only columns 0 .. je-1 (column c of L^-1 is zero above row c): n^3 / 6
// complex multiply-adds instead of n^3 / 2.
template <class R>
int getrs(hipStream_t s, int n, int nrhs, const typename P<R>::V* A, long lda, const int* ipiv,
          const typename P<R>::V* inv, typename P<R>::V* B, long ldb, int ng = 0, const int* g0 = nullptr,
          const int* gn = nullptr, const int* gf = nullptr, bool ident = false)
{
  using V = typename P<R>::V;
  const V one = P<R>::v(1.0, 0.0), mone = P<R>::v(-1.0, 0.0);
  if (n <= 0 || nrhs <= 0) return 0;
  if (ident && (nrhs != n || ng > 0)) {
    bq::set_error("getrs: the identity form takes n right-hand sides and no column groups");
    return -1;
  }
  const int kt_all = bq::ktime_begin(P<R>::kt_s, s);
  const int* perm = perm_of(inv, n);  // stored by getrf (ipiv is kept for the interface)
  (void)ipiv;
  // rows permuted in place, kPermChunk columns at a time through one n x kPermChunk buffer
  // (not a copy of all of B: with n right-hand sides that would double the solve's memory)
  if (!ident) {
    const int chunk = min(nrhs, kPermChunk);
    V* Bt = nullptr;
    BQ_HIP(hipMallocAsync((void**)&Bt, sizeof(V) * (size_t)n * chunk, s));
    for (int c0 = 0; c0 < nrhs; c0 += chunk) {
      const int nc = min(chunk, nrhs - c0);
      const long tot = (long)n * nc;
      k_gather_rows<R><<<(unsigned)((tot + 255) / 256), 256, 0, s>>>(n, nc, perm, B + (long)c0 * ldb, ldb, Bt);
      BQ_HIP(hipGetLastError());
      BQ_HIP(hipMemcpy2DAsync(B + (long)c0 * ldb, ldb * sizeof(V), Bt, (size_t)n * sizeof(V), (size_t)n * sizeof(V),
                              nc, hipMemcpyDeviceToDevice, s));
    }
    BQ_HIP(hipFreeAsync(Bt, s));
  }
  if (nrhs <= kSmallRhs) {
    for (int j0 = 0; j0 < n; j0 += kNB2) {
      const int je = min(n, j0 + kNB2);
      k_trsv_blk<R, true><<<1, 256, 0, s>>>(n, A, lda, inv, B, ldb, nrhs, j0);
      if (n - je > 0)
        k_skinny_update<R><<<(n - je + 63) / 64, 256, 0, s>>>(n - je, je - j0, A + (long)j0 * lda + je, lda, B + j0,
                                                               ldb, nrhs, B + je, ldb);
      BQ_HIP(hipGetLastError());
    }
    for (int j0 = ((n - 1) / kNB2) * kNB2; j0 >= 0; j0 -= kNB2) {
      const int je = min(n, j0 + kNB2);
      k_trsv_blk<R, false><<<1, 256, 0, s>>>(n, A, lda, inv, B, ldb, nrhs, j0);
      if (j0 > 0)
        k_skinny_update<R><<<(j0 + 63) / 64, 256, 0, s>>>(j0, je - j0, A + (long)j0 * lda, lda, B + j0, ldb, nrhs, B,
                                                           ldb);
      BQ_HIP(hipGetLastError());
    }
    bq::ktime_end(kt_all, s);
    return 0;
  }
  const unsigned cblk = (unsigned)((nrhs + 63) / 64);
  // FP32 with the matrix-core updates and many right-hand sides: the full 512-row diagonal
  // blocks are inverted once (all at a time) and applied as products on the matrix cores
  const int nfull = n / kNBF;
  V *WL = nullptr, *WU = nullptr;
  bool blk_inv = false;
  if constexpr (std::is_same<R, float>::value) {
    if (update_mode() == 1 && nrhs >= kSplitMin && nfull > 0) {
      blk_inv = true;
      const size_t bs = (size_t)kNBF * kNBF;
      BQ_HIP(hipMallocAsync((void**)&WL, sizeof(V) * bs * nfull, s));
      BQ_HIP(hipMallocAsync((void**)&WU, sizeof(V) * bs * nfull, s));
      if (block_inverses<R>(s, nfull, A, lda, inv, false, WL) || block_inverses<R>(s, nfull, A, lda, inv, true, WU))
        return -1;
    }
  }
  // forward: unit lower L, 512-row blocks of 32-row steps.  With column groups, group g
  // (columns g0[g] .. g0[g] + gn[g]) starts at the block holding row gf[g]: the caller states
  // that its permuted right-hand sides are zero above that row, so are its solutions there.
  const int ngrp = ng > 0 ? ng : 1;
  for (int g = 0; g < ngrp; ++g) {
    const int c0 = ng > 0 ? g0[g] : 0, nc = ng > 0 ? gn[g] : nrhs;
    const int first = ng > 0 ? (max(0, gf[g]) / kNBF) * kNBF : 0;
    if (nc <= 0) continue;
    V* Bg = B + (long)c0 * ldb;
    // kSolveGrp blocks at a time: in-group updates of the group's later rows, then one update of
    // all rows below the group with rank kSolveGrp x 512 (each row below is read and written
    // once per group rather than once per block)
    for (int ga = first; ga < n; ga += kSolveGrp * kNBF) {
      const int gz = min(n, ga + kSolveGrp * kNBF);
      for (int j0 = ga; j0 < gz; j0 += kNBF) {
        const int je = min(n, j0 + kNBF);
        const int ncb = ident ? min(nc, je) : nc;  // columns with nonzeros in this block row
        const unsigned gblk = (unsigned)((ncb + 63) / 64);
        bool done = false;
        if constexpr (std::is_same<R, float>::value) {
          if (blk_inv && je - j0 == kNBF && ncb >= kSplitMin) {
            if (apply_block_inverse(s, ncb, WL + (size_t)(j0 / kNBF) * kNBF * kNBF, Bg + j0, ldb)) return -1;
            done = true;
          }
        }
        for (int j = j0; j < je && !done; j += kNB) {
          const int jb = min(kNB, je - j);
          k_apply_blk<R><<<gblk, 256, 0, s>>>(inv + (long)(j / kNB) * kBlkElems, jb, Bg + j, ldb, ncb);
          BQ_HIP(hipGetLastError());
          if (je - j - jb > 0 && gemm<R>(s, je - j - jb, ncb, jb, mone, A + (long)j * lda + (j + jb), lda, Bg + j,
                                         ldb, one, Bg + j + jb, ldb))
            return -1;
        }
        if (gz - je > 0 &&
            gemm<R>(s, gz - je, ncb, je - j0, mone, A + (long)j0 * lda + je, lda, Bg + j0, ldb, one, Bg + je, ldb))
          return -1;
      }
      const int ncg = ident ? min(nc, gz) : nc;
      if (n - gz > 0 &&
          gemm<R>(s, n - gz, ncg, gz - ga, mone, A + (long)ga * lda + gz, lda, Bg + ga, ldb, one, Bg + gz, ldb))
        return -1;
    }
  }
  // backward: U, 512-row blocks from the bottom, 32-row steps upwards inside each, grouped as
  // the forward sweep
  const int nblk = (n + kNBF - 1) / kNBF;
  for (int bh = nblk - 1; bh >= 0; bh -= kSolveGrp) {
    const int bl = max(0, bh - kSolveGrp + 1), ga = bl * kNBF, gz = min(n, (bh + 1) * kNBF);
    for (int b = bh; b >= bl; --b) {
      const int j0 = b * kNBF, je = min(n, j0 + kNBF);
      bool done = false;
      if constexpr (std::is_same<R, float>::value) {
        if (blk_inv && je - j0 == kNBF) {
          if (apply_block_inverse(s, nrhs, WU + (size_t)(j0 / kNBF) * kNBF * kNBF, B + j0, ldb)) return -1;
          done = true;
        }
      }
      for (int j = ((je - 1 - j0) / kNB) * kNB + j0; j >= j0 && !done; j -= kNB) {
        const int jb = min(kNB, je - j);
        k_apply_blk<R><<<cblk, 256, 0, s>>>(inv + (long)(j / kNB) * kBlkElems + kNB * kNB, jb, B + j, ldb, nrhs);
        BQ_HIP(hipGetLastError());
        if (j > j0 && gemm<R>(s, j - j0, nrhs, jb, mone, A + (long)j * lda + j0, lda, B + j, ldb, one, B + j0, ldb))
          return -1;
      }
      if (j0 > ga &&
          gemm<R>(s, j0 - ga, nrhs, je - j0, mone, A + (long)j0 * lda + ga, lda, B + j0, ldb, one, B + ga, ldb))
        return -1;
    }
    if (ga > 0 && gemm<R>(s, ga, nrhs, gz - ga, mone, A + (long)ga * lda, lda, B + ga, ldb, one, B, ldb)) return -1;
  }
  if (WL) BQ_HIP(hipFreeAsync(WL, s));
  if (WU) BQ_HIP(hipFreeAsync(WU, s));
  bq::ktime_end(kt_all, s);
  return 0;
}

// -inv of the 512-row diagonal blocks of both factors into W (lower ones, then upper ones)
template <class R>
int blkinv_both(hipStream_t s, int n, const typename P<R>::V* A, long lda, const typename P<R>::V* inv,
                typename P<R>::V* W)
{
  if (!(n > 0 && n % kNBF == 0)) {
    bq::set_error("block inverses: order %d is not a multiple of %d", n, kNBF);
    return -1;
  }
  const int nfull = n / kNBF;
  if (block_inverses<R>(s, nfull, A, lda, inv, false, W)) return -1;
  return block_inverses<R>(s, nfull, A, lda, inv, true, W + (long)kNBF * kNBF * nfull);
}

}  // namespace

extern "C" {

// Release what the library keeps per stream on the current device (the rocBLAS handle bound
// to it and the split-GEMM workspace); call before destroying a stream the solves used.
void bq_lu_release_stream(void* stream)
{
  int dev = 0;
  (void)hipGetDevice(&dev);
  {
    std::lock_guard<std::mutex> lk(g_hmu);
    for (size_t i = 0; i < g_handles.size(); ++i)
      if (g_handles[i].dev == dev && g_handles[i].s == (hipStream_t)stream) {
        (void)rocblas_destroy_handle(g_handles[i].h);
        g_handles.erase(g_handles.begin() + i);
        break;
      }
  }
  bq_cgemm_split_release(stream);
  // the look-ahead stream of this stream, with its handle and workspace
  SideSlot q{};
  bool found = false;
  {
    std::lock_guard<std::mutex> lk(g_hmu);
    for (size_t i = 0; i < g_side.size(); ++i)
      if (g_side[i].dev == dev && g_side[i].s == (hipStream_t)stream) {
        q = g_side[i];
        g_side.erase(g_side.begin() + i);
        found = true;
        break;
      }
  }
  if (found) {
    (void)hipStreamSynchronize(q.side);
    bq_lu_release_stream(q.side);
    for (hipEvent_t e : q.ev) (void)hipEventDestroy(e);
    (void)hipStreamDestroy(q.side);
  }
}

// Look-ahead of the factorization: 1 on (default), 0 off; mode < 0 only queries.  Returns the
// mode in force before.
int bq_lu_lookahead(int mode)
{
  const int old = lookahead_mode();
  if (mode == 0 || mode == 1) g_lookahead.store(mode);
  return old;
}

// Update mode of the FP32 factorization: 1 bf16x3 matrix-core kernel (default), 0 rocBLAS
// cgemm (FP32 matrix cores); mode < 0 only queries.  Returns the mode in force before.
int bq_lu_update_mode(int mode)
{
  const int old = update_mode();
  if (mode == 0 || mode == 1) g_update_mode.store(mode);
  return old;
}

// C = alpha op(A) op(B) + beta C, op = transpose when the flag is set (rocBLAS, column-major)
int bq_zgemm_op_dev(void* stream, int ta, int tb, int m, int n, int k, const double* alpha, const double* A, long lda,
                    const double* B, long ldb, const double* beta, double* C, long ldc)
{
  double2 a = make_double2(alpha[0], alpha[1]), b = make_double2(beta[0], beta[1]);
  return gemm_op<double>((hipStream_t)stream, ta, tb, m, n, k, a, (const double2*)A, lda, (const double2*)B, ldb, b,
                         (double2*)C, ldc);
}

int bq_zgemm_dev(void* stream, int m, int n, int k, const double* alpha, const double* A, long lda, const double* B,
                 long ldb, const double* beta, double* C, long ldc)
{
  double2 a = make_double2(alpha[0], alpha[1]), b = make_double2(beta[0], beta[1]);
  return gemm<double>((hipStream_t)stream, m, n, k, a, (const double2*)A, lda, (const double2*)B, ldb, b,
                      (double2*)C, ldc);
}

int bq_cgemm_op_dev(void* stream, int ta, int tb, int m, int n, int k, const float* alpha, const float* A, long lda,
                    const float* B, long ldb, const float* beta, float* C, long ldc)
{
  float2 a = make_float2(alpha[0], alpha[1]), b = make_float2(beta[0], beta[1]);
  return gemm_op<float>((hipStream_t)stream, ta, tb, m, n, k, a, (const float2*)A, lda, (const float2*)B, ldb, b,
                        (float2*)C, ldc);
}

// B -= A X with A m x k (column-major, lda), X k x nrhs (ldx), B m x nrhs (ldb), nrhs <= 8,
// complex double; X and B must not overlap.  Returns 0, or -1 for nrhs out of range.
int bq_zgemm_skinny_dev(void* stream, int m, int nrhs, long k, const double* A, long lda, const double* X, long ldx,
                        double* B, long ldb)
{
  if (nrhs < 1 || nrhs > kSmallRhs) {
    bq::set_error("bq_zgemm_skinny_dev: nrhs %d outside 1..%d", nrhs, kSmallRhs);
    return -1;
  }
  if (m <= 0 || k <= 0) return 0;
  k_skinny_gemm<double><<<(m + 63) / 64, 256, 0, (hipStream_t)stream>>>(m, k, (const double2*)A, lda,
                                                                       (const double2*)X, ldx, nrhs, (double2*)B, ldb);
  BQ_HIP(hipGetLastError());
  return 0;
}

// The row permutation getrf stored in its inverse-block buffer: perm[i] = the original row
// that the factorization placed at row i.  dinv as passed to getrf (doubles or floats).
const int* bq_lu_perm_dev(const void* dinv, int n, int single)
{
  return single ? perm_of((const float2*)dinv, n) : perm_of((const double2*)dinv, n);
}

// X (n x n, ldx) = U^-1 L^-1 from getrf's factors: the inverse of A with its columns permuted,
// A^-1[:, perm[i]] = X[:, i] for perm from bq_lu_perm_dev.
int bq_zlu_inverse_dev(void* stream, int n, const double* Ad, long lda, const double* dinv, double* Xd, long ldx)
{
  if (n <= 0) return 0;
  hipStream_t s = (hipStream_t)stream;
  const long tot = (long)n * n;
  k_identity<double><<<(unsigned)((tot + 255) / 256), 256, 0, s>>>(n, (double2*)Xd, ldx);
  BQ_HIP(hipGetLastError());
  return getrs<double>(s, n, n, (const double2*)Ad, lda, nullptr, (const double2*)dinv, (double2*)Xd, ldx, 0, nullptr,
                       nullptr, nullptr, true);
}
int bq_clu_inverse_dev(void* stream, int n, const float* Ad, long lda, const float* dinv, float* Xd, long ldx)
{
  if (n <= 0) return 0;
  hipStream_t s = (hipStream_t)stream;
  const long tot = (long)n * n;
  k_identity<float><<<(unsigned)((tot + 255) / 256), 256, 0, s>>>(n, (float2*)Xd, ldx);
  BQ_HIP(hipGetLastError());
  return getrs<float>(s, n, n, (const float2*)Ad, lda, nullptr, (const float2*)dinv, (float2*)Xd, ldx, 0, nullptr,
                      nullptr, nullptr, true);
}

// bq_zgetrs_dev / bq_cgetrs_dev with column groups: group g (columns g0[g] .. g0[g]+gn[g]-1,
// host arrays) has permuted right-hand sides that are zero in rows 0 .. gf[g]-1.
int bq_zgetrs_groups_dev(void* stream, int n, int nrhs, const double* Ad, long lda, const int* ipiv,
                         const double* dinv, double* Bd, long ldb, int ng, const int* g0, const int* gn, const int* gf)
{
  return getrs<double>((hipStream_t)stream, n, nrhs, (const double2*)Ad, lda, ipiv, (const double2*)dinv,
                       (double2*)Bd, ldb, ng, g0, gn, gf);
}

int bq_cgetrs_groups_dev(void* stream, int n, int nrhs, const float* Ad, long lda, const int* ipiv,
                         const float* dinv, float* Bd, long ldb, int ng, const int* g0, const int* gn, const int* gf)
{
  return getrs<float>((hipStream_t)stream, n, nrhs, (const float2*)Ad, lda, ipiv, (const float2*)dinv,
                      (float2*)Bd, ldb, ng, g0, gn, gf);
}

long bq_lu_inv_doubles(int n)
{
  // inverted diagonal blocks, then the row permutation (n int32, which fit in n reals)
  return (long)((n + kNB - 1) / kNB) * kBlkElems * 2 + n;
}

// In-place LU of the n x n column-major matrix A (device).  ipiv: device int[n]; dinv: device
// buffer of bq_lu_inv_doubles(n) doubles for the inverted diagonal blocks (used by getrs).
// Returns 0, a positive k if U(k,k) is exactly zero (1-based, like LAPACK), or < 0.
int bq_zgetrf_dev(void* stream, int n, double* Ad, long lda, int* ipiv, double* dinv)
{
  return getrf<double>((hipStream_t)stream, n, (double2*)Ad, lda, ipiv, (double2*)dinv);
}

// Solve A X = B with the factors from bq_zgetrf_dev; B is n x nrhs column-major (device).
int bq_zgetrs_dev(void* stream, int n, int nrhs, const double* Ad, long lda, const int* ipiv, const double* dinv,
                  double* Bd, long ldb)
{
  return getrs<double>((hipStream_t)stream, n, nrhs, (const double2*)Ad, lda, ipiv, (const double2*)dinv,
                       (double2*)Bd, ldb);
}

// The same in complex FP32: A, B interleaved complex floats; dinv holds bq_lu_inv_doubles(n)
// floats.
int bq_cgetrf_dev(void* stream, int n, float* Ad, long lda, int* ipiv, float* dinv)
{
  return getrf<float>((hipStream_t)stream, n, (float2*)Ad, lda, ipiv, (float2*)dinv);
}

int bq_cgetrs_dev(void* stream, int n, int nrhs, const float* Ad, long lda, const int* ipiv, const float* dinv,
                  float* Bd, long ldb)
{
  return getrs<float>((hipStream_t)stream, n, nrhs, (const float2*)Ad, lda, ipiv, (const float2*)dinv, (float2*)Bd,
                      ldb);
}

// The inverted 512-row diagonal blocks of both factors for bq_[zc]getrs_blkinv_dev: complex
// elements W must hold (0 when n is not a multiple of 512), and their construction.
long bq_lu_blkinv_elems(int n) { return (n > 0 && n % kNBF == 0) ? 2L * (n / kNBF) * kNBF * kNBF : 0; }

int bq_zlu_blkinv_dev(void* stream, int n, const double* Ad, long lda, const double* dinv, double* W)
{
  return blkinv_both<double>((hipStream_t)stream, n, (const double2*)Ad, lda, (const double2*)dinv, (double2*)W);
}
int bq_clu_blkinv_dev(void* stream, int n, const float* Ad, long lda, const float* dinv, float* W)
{
  return blkinv_both<float>((hipStream_t)stream, n, (const float2*)Ad, lda, (const float2*)dinv, (float2*)W);
}
int bq_zgetrs_blkinv_dev(void* stream, int n, int nrhs, const double* Ad, long lda, const double* dinv,
                         const double* W, double* Bd, long ldb)
{
  return getrs_blkinv<double>((hipStream_t)stream, n, nrhs, (const double2*)Ad, lda, (const double2*)dinv,
                              (const double2*)W, (double2*)Bd, ldb);
}
int bq_cgetrs_blkinv_dev(void* stream, int n, int nrhs, const float* Ad, long lda, const float* dinv,
                         const float* W, float* Bd, long ldb)
{
  return getrs_blkinv<float>((hipStream_t)stream, n, nrhs, (const float2*)Ad, lda, (const float2*)dinv,
                             (const float2*)W, (float2*)Bd, ldb);
}

// Host-memory convenience: solve A X = B (A n x n, B n x nrhs, both column-major host).
int bq_zgesv_host(int n, int nrhs, double* A, double* B)
{
  if (bq::ensure_device()) return -1;
  hipStream_t s = bq::stream();
  double *dA = nullptr, *dB = nullptr, *dv = nullptr;
  int* dp = nullptr;
  size_t an = sizeof(double2) * (size_t)n * n, bn = sizeof(double2) * (size_t)n * nrhs;
  BQ_HIP(hipMalloc((void**)&dA, an));
  BQ_HIP(hipMalloc((void**)&dB, bn));
  BQ_HIP(hipMalloc((void**)&dp, sizeof(int) * n));
  BQ_HIP(hipMalloc((void**)&dv, sizeof(double) * bq_lu_inv_doubles(n)));
  BQ_HIP(hipMemcpyAsync(dA, A, an, hipMemcpyHostToDevice, s));
  BQ_HIP(hipMemcpyAsync(dB, B, bn, hipMemcpyHostToDevice, s));
  int info = bq_zgetrf_dev(s, n, dA, n, dp, dv);
  if (info == 0) info = bq_zgetrs_dev(s, n, nrhs, dA, n, dp, dv, dB, n);
  if (info == 0) {
    BQ_HIP(hipMemcpyAsync(B, dB, bn, hipMemcpyDeviceToHost, s));
    BQ_HIP(hipMemcpyAsync(A, dA, an, hipMemcpyDeviceToHost, s));
  }
  BQ_HIP(hipStreamSynchronize(s));
  (void)hipFree(dA);
  (void)hipFree(dB);
  (void)hipFree(dp);
  (void)hipFree(dv);
  return info;
}

// The same in complex FP32 (A, B interleaved complex floats, host).
int bq_cgesv_host(int n, int nrhs, float* A, float* B)
{
  if (bq::ensure_device()) return -1;
  hipStream_t s = bq::stream();
  float *dA = nullptr, *dB = nullptr, *dv = nullptr;
  int* dp = nullptr;
  size_t an = sizeof(float2) * (size_t)n * n, bn = sizeof(float2) * (size_t)n * nrhs;
  BQ_HIP(hipMalloc((void**)&dA, an));
  BQ_HIP(hipMalloc((void**)&dB, bn));
  BQ_HIP(hipMalloc((void**)&dp, sizeof(int) * n));
  BQ_HIP(hipMalloc((void**)&dv, sizeof(float) * bq_lu_inv_doubles(n)));
  BQ_HIP(hipMemcpyAsync(dA, A, an, hipMemcpyHostToDevice, s));
  BQ_HIP(hipMemcpyAsync(dB, B, bn, hipMemcpyHostToDevice, s));
  int info = bq_cgetrf_dev(s, n, dA, n, dp, dv);
  if (info == 0) info = bq_cgetrs_dev(s, n, nrhs, dA, n, dp, dv, dB, n);
  if (info == 0) BQ_HIP(hipMemcpyAsync(B, dB, bn, hipMemcpyDeviceToHost, s));
  BQ_HIP(hipStreamSynchronize(s));
  (void)hipFree(dA);
  (void)hipFree(dB);
  (void)hipFree(dp);
  (void)hipFree(dv);
  return info;
}

}  // extern "C"
