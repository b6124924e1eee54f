// bq_solver.hip -- device solver session behind solve_bieq.
//
// Global system.  The reference assembles one sparse 8NN x 8NN system (unknowns X_0..X_3 =
// A_x, A_y, A_z, phi and X_4..X_7 = their normal derivatives) and solves it with PARDISO
// (d3qd2_src/d3qd2_solve_bieq.c:258-454).  Its rows have a fixed block structure: with
// Y_c = (X_c, X_{c+4}) for c = 0..3 and the rows of component c stacked in domain order,
//   M Y_c + K diag(n_c) X_3 = b_c          (c = 0, 1, 2)
//   sum_c L diag(n_c) X_c + M3 Y_3 = b_3
// where M is [H | -G] on rows of an element's main domain and [H | G] on its sub-domain
// rows, K = (k1^2 - k2^2) G and L = (1 - k1^2/k2^2) G on sub-domain rows (zero elsewhere),
// M3 = [H | -G] / [H | (k1/k2)^2 G], and n_c is the c-th component of the unit normal at
// each element node (create_matrix_csr_dac, :284-421).  Block elimination gives
//   Z = M^{-1} K,  y_c = M^{-1} b_c,
//   S = M3 - [ L (Z_top o Nn) | 0 ],  Nn_ij = n_i . n_j,
//   S Y_3 = b_3 - L (sum_c n_c o y_c,top),   Y_c = y_c - Z (n_c o X_3),
// two dense LUs of order 2NN instead of one of order 8NN (64x fewer LU flops), all
// on the device.  The result is the same linear system's solution; only the elimination
// order differs from a direct factorisation.
// Domain matrices keep 4 rows and columns per element; a linear triangle's fourth row and
// column are empty, and the global system and the dE/dH systems take the live rows only, in
// the reference's order (create_matrix_csr_dac and solve_deh_bv skip them the same way).
#include <vector>
#include <hip/hip_runtime.h>
#include "bq_runtime.hpp"
#include "bq_complex.hpp"
#include "bq_sess.hpp"
#include "../bq_solver.h"
#include "../bq_assemble.h"
#include "../bq_linalg.h"
#include "../bq_plan.h"
#include "bq_faddeeva.hpp"
#include <cmath>
#include <cstdlib>
#include <cfloat>
#include <climits>
#include <algorithm>

extern "C" int bq_qpgf_v1_dev(const bq_qplan* plan, int variant, int n, const double* d_r, double* d_out,
                              hipStream_t s);

namespace {

template <class T>
int dmalloc(T** p, size_t n)
{
  if (n == 0) {
    *p = nullptr;
    return 0;
  }
  BQ_HIP(hipMalloc((void**)p, n * sizeof(T)));
  return 0;
}

template <class T>
int upload(T** p, const T* h, size_t n, hipStream_t s)
{
  if (dmalloc(p, n)) return -1;
  if (n) BQ_HIP(hipMemcpyAsync(*p, h, n * sizeof(T), hipMemcpyHostToDevice, s));
  return 0;
}

__device__ __forceinline__ cplx ld2(const double* p, long i) { return mk(p[2 * i], p[2 * i + 1]); }
__device__ __forceinline__ void st2(double* p, long i, cplx v)
{
  p[2 * i] = v.re;
  p[2 * i + 1] = v.im;
}

// one thread per (element s of the domain, row of the domain): entries of M, M3, K, L
__global__ void k_build_blocks(int dNe, int d, int rowbase, int NN, const int* __restrict__ rowmap,
                               const int* __restrict__ sid, const int* __restrict__ eni, const int* __restrict__ md,
                               const int* __restrict__ sdm, const double* __restrict__ kn,
                               const double* __restrict__ G, const double* __restrict__ H, double* __restrict__ M,
                               double* __restrict__ M3, double* __restrict__ K, double* __restrict__ L)
{
  const long N = 4L * dNe;
  long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= N * dNe) return;
  const int sc = (int)(idx / N);
  const long rl = idx % N;
  const int rc = rowmap[rl];
  if (rc < 0) return;
  const long r = rowbase + rc, ld = 2L * NN;
  const int sd = sid[sc + 1], asd = abs(sd), mdid = md[asd], sdid = sdm[asd];
  const bool main_side = (d == mdid);
  cplx beps = mk(1, 0), delt = mk(0, 0), sigm = mk(0, 0);
  if (!main_side) {
    cplx k1 = ld2(kn, mdid), k2 = ld2(kn, sdid);
    cplx k12 = k1 * k1, k22 = k2 * k2;
    delt = k12 - k22;
    beps = cdiv(k12, k22);
    sigm = mk(1.0, 0.0) - beps;
  }
  for (int l = 0; l < 4; ++l) {
    const long e = eni[4 * asd + l];
    if (e < 0) continue;
    cplx g = ld2(G, rl * N + 4L * sc + l), h = ld2(H, rl * N + 4L * sc + l);
    st2(M, e * ld + r, h);
    st2(M3, e * ld + r, h);
    if (main_side) {
      st2(M, (NN + e) * ld + r, -g);
      st2(M3, (NN + e) * ld + r, -g);
    } else {
      st2(M, (NN + e) * ld + r, g);
      st2(M3, (NN + e) * ld + r, beps * g);
      st2(K, e * ld + r, delt * g);
      st2(L, e * ld + r, sigm * g);
    }
  }
}

// right-hand sides: b_c (c < 3) into columns NN..NN+2 of R, b_3 into b3
__global__ void k_build_rhs(int dNe, int d, int rowbase, int NN, const int* __restrict__ rowmap,
                            const int* __restrict__ sid, const int* __restrict__ md, const int* __restrict__ sdm,
                            const double* __restrict__ kn, const double* __restrict__ nrm,
                            const double* __restrict__ Ui, const double* __restrict__ dUi,
                            const double* __restrict__ G, const double* __restrict__ H, double* __restrict__ R,
                            double* __restrict__ b3)
{
  const long N = 4L * dNe;
  long rl = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (rl >= N) return;
  const int rc = rowmap[rl];
  if (rc < 0) return;
  const long r = rowbase + rc, ld = 2L * NN;
  cplx b[4] = {mk(0, 0), mk(0, 0), mk(0, 0), mk(0, 0)};
  for (int sc = 0; sc < dNe; ++sc) {
    const int sd = sid[sc + 1], asd = abs(sd), mdid = md[asd], sdid = sdm[asd];
    if (d == mdid || mdid != 0) continue;
    cplx k1 = ld2(kn, mdid), k2 = ld2(kn, sdid);
    cplx sigm = mk(1.0, 0.0) - cdiv(k1 * k1, k2 * k2);
    for (int l = 0; l < 4; ++l) {
      cplx g = ld2(G, rl * N + 4L * sc + l), h = ld2(H, rl * N + 4L * sc + l);
      const double* vn = nrm + 12L * asd + 3 * l;
      cplx ndu = mk(0, 0);
      for (int c = 0; c < 3; ++c) {
        cplx ui = ld2(Ui, 16L * asd + 4 * l + c), dui = ld2(dUi, 16L * asd + 4 * l + c);
        b[c] = b[c] - h * ui - g * dui;
        ndu += ui * vn[c];
      }
      b[3] = b[3] - sigm * g * ndu;
    }
  }
  for (int c = 0; c < 3; ++c) st2(R, (long)(NN + c) * ld + r, b[c]);
  st2(b3, r, b[3]);
}

// W = Z_top o (n_i . n_j), NN x NN column-major; Z has leading dimension 2NN
__global__ void k_hadamard_normals(int NN, const double* __restrict__ Z, const double* __restrict__ nn3,
                                   double* __restrict__ W)
{
  long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (long)NN * NN) return;
  long j = idx / NN, i = idx % NN;
  double g = nn3[3 * i] * nn3[3 * j] + nn3[3 * i + 1] * nn3[3 * j + 1] + nn3[3 * i + 2] * nn3[3 * j + 2];
  st2(W, idx, ld2(Z, j * 2L * NN + i) * g);
}

// v[e] = sum_c n_c(e) y_c[e] (top half of y_c = column NN + c of R)
__global__ void k_normal_dot(int NN, const double* __restrict__ R, const double* __restrict__ nn3, double* __restrict__ v)
{
  long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= NN) return;
  const long ld = 2L * NN;
  cplx a = mk(0, 0);
  for (int c = 0; c < 3; ++c) a += ld2(R, (long)(NN + c) * ld + e) * nn3[3 * e + c];
  st2(v, e, a);
}

// T[e][c] = n_c(e) X3[e], NN x 3 column-major
__global__ void k_scale_normals(int NN, const double* __restrict__ Y3, const double* __restrict__ nn3,
                                double* __restrict__ T)
{
  long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= NN) return;
  cplx x = ld2(Y3, e);
  for (int c = 0; c < 3; ++c) st2(T, (long)c * NN + e, x * nn3[3 * e + c]);
}

// ---- mixed-precision solve helpers (F = double or float, interleaved complex) ----------
__global__ void k_d2f(long n, const double* __restrict__ a, float* __restrict__ b)
{
  long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) b[i] = (float)a[i];
}

__global__ void k_add_f2d(long n, const float* __restrict__ d, double* __restrict__ x)
{
  long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) x[i] += (double)d[i];
}

// W[i, j] = Z[i, j] (n_i . n_j), Z the top NN rows of an n2 x NN matrix (ld 2 NN)
template <class F>
__global__ void k_hadamard_t(int NN, const F* __restrict__ Z, const double* __restrict__ nn3, F* __restrict__ W)
{
  long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (long)NN * NN) return;
  long j = idx / NN, i = idx % NN;
  const double g = nn3[3 * i] * nn3[3 * j] + nn3[3 * i + 1] * nn3[3 * j + 1] + nn3[3 * i + 2] * nn3[3 * j + 2];
  const long z = j * 2L * NN + i;
  W[2 * idx] = (F)(Z[2 * z] * g);
  W[2 * idx + 1] = (F)(Z[2 * z + 1] * g);
}

// v[e] = sum_c n_c(e) Y[e, c], Y n2 x 3 (ld 2 NN), top NN rows
template <class F>
__global__ void k_normal_dot_t(int NN, const F* __restrict__ Y, const double* __restrict__ nn3, F* __restrict__ v)
{
  long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= NN) return;
  const long ld = 2L * NN;
  double re = 0.0, im = 0.0;
  for (int c = 0; c < 3; ++c) {
    re += (double)Y[2 * ((long)c * ld + e)] * nn3[3 * e + c];
    im += (double)Y[2 * ((long)c * ld + e) + 1] * nn3[3 * e + c];
  }
  v[2 * e] = (F)re;
  v[2 * e + 1] = (F)im;
}

// T[e, c] = n_c(e) y3[e], NN x 3 column-major
template <class F>
__global__ void k_scale_normals_t(int NN, const F* __restrict__ y3, const double* __restrict__ nn3, F* __restrict__ T)
{
  long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= NN) return;
  for (int c = 0; c < 3; ++c) {
    T[2 * ((long)c * NN + e)] = (F)(y3[2 * e] * nn3[3 * e + c]);
    T[2 * ((long)c * NN + e) + 1] = (F)(y3[2 * e + 1] * nn3[3 * e + c]);
  }
}

// B (n x n, column-major) = A[live, live], A row-major N x N
__global__ void k_gather_live(long N, int n, const int* __restrict__ live, const double* __restrict__ A,
                              double* __restrict__ B)
{
  long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (long)n * n) return;
  long j = idx / n, i = idx % n;
  st2(B, idx, ld2(A, (long)live[i] * N + live[j]));
}

// B[c][i] = A[c][live[i]] (to_live) or A[c][live[i]] = B[c][i] (back), c < ncol; columns of
// length N (A) and n (B)
__global__ void k_live_rows(long N, int n, int ncol, const int* __restrict__ live, double* __restrict__ A,
                           double* __restrict__ B, int to_live)
{
  long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (long)n * ncol) return;
  long c = idx / n, i = idx % n;
  if (to_live) st2(B, idx, ld2(A, c * N + live[i]));
  else st2(A, c * N + live[i], ld2(B, idx));
}

inline unsigned nblk(long n, int b) { return (unsigned)((n + b - 1) / b); }

double asm_tolerance();

// what the matrices depend on besides the mesh: QPD2 (k, Bloch vector, lattice), every
// medium's wave number, the periodicity switch and the assembly tolerance
std::vector<double> wave_key(const bq_model* m)
{
  std::vector<double> k(m->qd, m->qd + 8);
  k.insert(k.end(), m->kn, m->kn + 2 * (m->MN + 1));
  k.push_back((double)m->ps);
  k.push_back(asm_tolerance());
  return k;
}

// Assembly tolerance: 0 = the reference's accuracy (lattice sums to exp(-40) of their leading
// term, Gauss orders bounded at 1e-15 of the integrand); tol > 0 = matrices accurate to about
// tol of their largest entries, for solves whose own accuracy target is far above the
// reference's rounding (the mixed-precision path): the Gauss-order bound becomes tol and the
// lattice-sum cut is tuned per incidence (tune_plan).  BQ_ASM_TOL sets the initial value.
double g_asm_tol = -1.0;
double asm_tolerance()
{
  if (g_asm_tol < 0.0) {
    const char* e = getenv("BQ_ASM_TOL");
    const double v = e ? atof(e) : 0.0;
    g_asm_tol = (v > 0.0 && v < 1e-3) ? v : 0.0;
  }
  return g_asm_tol;
}

// The plan's Faddeeva-series weights (bq_qplan.wtab) for the incidence's wave number: rows for
// the Ewald parameters v_j = veps0 1.02^-j down to sqrt(BQ_EW_EXPM) / k, the smallest that the
// parameter rule gives (at r = 0).  The series' relative error is Algorithm 916's parameter,
// a = pi / sqrt(-ln(relerr / 2)): 1e-16 at the reference's accuracy, and a tenth of the
// assembly tolerance (at most 1e-13) for a tuned assembly, which then also takes a single
// Newton step for the terms' reciprocals.
int build_wtab(bq_sess* S)
{
  bq_qplan& P = S->plan;
  P.wtab = nullptr;
  P.wtab_nv = 0;
  if (S->ps != 1) return 0;
  const double lrho = log(1.02), vlow = sqrt(BQ_EW_EXPM) / P.k;
  int nv = 1;
  if (P.veps0 > vlow) nv = (int)ceil(log(P.veps0 / vlow) / lrho) + 2;
  nv = std::min(nv, BQ_WTAB_MAXV);
  if (!S->wtab) {
    BQ_HIP(hipMalloc((void**)&S->wtab, sizeof(double) * BQ_WTAB_ROW * BQ_WTAB_MAXV));
    S->bytes += sizeof(double) * BQ_WTAB_ROW * BQ_WTAB_MAXV;
  }
  const double tol = S->st.asm_tol;
  const bool coarse = tol >= 1e-12;
  const double L = coarse ? -log(0.5 * std::min(0.1 * tol, 1e-13)) : 0.0;
  const double a = coarse ? M_PI / sqrt(L) : bq::kFA;
  const double span = coarse ? sqrt(L) + 0.18 : 6.3;  // a n - x beyond which the terms are negligible
  S->wtab_h.assign((size_t)BQ_WTAB_ROW * nv, 0.0);
  for (int j = 0; j < nv; ++j) {
    double* r = S->wtab_h.data() + (size_t)j * BQ_WTAB_ROW;
    const double v = P.veps0 * exp(-j * lrho), x = P.k * v;
    r[0] = v;
    r[1] = x;
    r[2] = exp(-x * x);
    r[3] = (double)std::min(BQ_WTAB_N, (int)((x + span) / a) + 2);
    r[4] = 2.0 * a / M_PI;
    r[5] = coarse ? 1.0 : 0.0;
    for (int n = 1; n <= BQ_WTAB_N; ++n) {
      const double an = a * n, em = exp(-(an + x) * (an + x)), ep = exp(-(an - x) * (an - x));
      double* c = r + 6 * n;
      c[0] = an * an;
      c[1] = exp(-an * an - x * x);
      c[2] = em;
      c[3] = an * em;
      c[4] = ep;
      c[5] = an * ep;
    }
  }
  BQ_HIP(hipMemcpyAsync(S->wtab, S->wtab_h.data(), sizeof(double) * S->wtab_h.size(), hipMemcpyHostToDevice, S->s));
  P.wtab = S->wtab;
  P.wtab_nv = nv;
  P.wtab_lrho = lrho;
  return 0;
}

// Per-incidence choice of the lattice-sum cut for tolerance tol: the QPGF (value and gradient)
// at kSamples offsets between random collocation nodes and random points of random source
// elements of the open region, with the default cut and with smaller ones; the smallest cut
// whose largest difference, relative to the largest value and gradient over the sample, stays
// below tol / 8 is taken (the differences shrink monotonically as the cut grows).
constexpr int kSamples = 4096;
int tune_plan(bq_sess* S, double tol)
{
  S->plan.cut = BQ_EXP_CUT;
  S->plan.rule_ltol = BQ_RULE_LTOL;
  S->st.asm_tol = tol;
  S->st.asm_cut = BQ_EXP_CUT;
  if (!(tol > 0.0)) return 0;
  S->plan.rule_ltol = std::max(log(tol), BQ_RULE_LTOL);
  if (S->ps != 1 || S->dNe.empty() || S->dNe[0] == 0) return 0;
  const int dNe = S->dNe[0];
  std::vector<double> R(3 * (size_t)kSamples);
  unsigned long long x = 0x9e3779b97f4a7c15ULL;  // fixed seed: the same samples for the same mesh
  auto rnd = [&]() {
    x ^= x << 13;
    x ^= x >> 7;
    x ^= x << 17;
    return (double)(x >> 11) * (1.0 / 9007199254740992.0);
  };
  for (int i = 0; i < kSamples; ++i) {
    const int t = 1 + (int)(rnd() * dNe) % dNe, sc = 1 + (int)(rnd() * dNe) % dNe;
    const int at = abs(S->sid_h[0][t]), as = abs(S->sid_h[0][sc]);
    const int tn = (int)(rnd() * 3.0);  // 0..2: a node of a triangle too
    const double z = 2.0 * rnd() - 1.0, e = 2.0 * rnd() - 1.0;
    const double* c = S->m.E + 24 * (size_t)as;  // cr[3][4]
    for (int d = 0; d < 3; ++d) {
      const double rs = c[4 * d] + c[4 * d + 1] * z + c[4 * d + 2] * e + c[4 * d + 3] * z * e;
      R[3 * i + d] = S->m.ren[12 * (size_t)at + 3 * tn + d] - rs;
    }
  }
  static const double cuts[] = {BQ_EXP_CUT, 36.0, 32.0, 29.0, 26.0, 23.0, 20.0};
  const int nc = (int)(sizeof cuts / sizeof cuts[0]);
  hipStream_t s = S->s;
  double *dR = nullptr, *dO = nullptr;
  BQ_HIP(hipMallocAsync((void**)&dR, sizeof(double) * 3 * kSamples, s));
  BQ_HIP(hipMallocAsync((void**)&dO, sizeof(double) * 8 * kSamples * nc, s));
  BQ_HIP(hipMemcpyAsync(dR, R.data(), sizeof(double) * 3 * kSamples, hipMemcpyHostToDevice, s));
  for (int j = 0; j < nc; ++j) {
    bq_qplan P = S->plan;
    P.cut = cuts[j];
    P.err = nullptr;
    if (bq_qpgf_v1_dev(&P, 1, kSamples, dR, dO + (size_t)8 * kSamples * j, s)) return -1;
  }
  std::vector<double> O((size_t)8 * kSamples * nc);
  BQ_HIP(hipMemcpyAsync(O.data(), dO, O.size() * sizeof(double), hipMemcpyDeviceToHost, s));
  BQ_HIP(hipFreeAsync(dR, s));
  BQ_HIP(hipFreeAsync(dO, s));
  BQ_HIP(hipStreamSynchronize(s));
  double sg = 0.0, sd = 0.0;
  for (int i = 0; i < kSamples; ++i) {
    const double* o = O.data() + 8 * (size_t)i;
    sg = std::max(sg, std::hypot(o[0], o[1]));
    for (int c = 0; c < 3; ++c) sd = std::max(sd, std::hypot(o[2 + 2 * c], o[3 + 2 * c]));
  }
  if (!(sg > 0.0) || !(sd > 0.0)) return 0;
  double cut = BQ_EXP_CUT;
  for (int j = 1; j < nc; ++j) {
    double err = 0.0;
    for (int i = 0; i < kSamples; ++i) {
      const double* a = O.data() + 8 * (size_t)i;
      const double* b = O.data() + (size_t)8 * kSamples * j + 8 * (size_t)i;
      err = std::max(err, std::hypot(a[0] - b[0], a[1] - b[1]) / sg);
      for (int c = 0; c < 3; ++c) err = std::max(err, std::hypot(a[2 + 2 * c] - b[2 + 2 * c], a[3 + 2 * c] - b[3 + 2 * c]) / sd);
    }
    if (!(err <= tol / 8.0)) break;
    cut = cuts[j];
  }
  S->plan.cut = cut;
  S->st.asm_cut = cut;
  return 0;
}

}  // namespace

extern "C" {

int bq_set_quadrature(const double* xl, const double* wl, const double* xh, const double* wh);

bq_sess* bq_sess_create(const bq_model* m)
{
  if (bq::ensure_device()) return nullptr;
  bq_sess* S = new bq_sess();
  S->m = *m;
  S->Ne = m->Ne;
  S->NN = m->NN;
  S->MN = m->MN;
  S->ps = m->ps;
  // each session runs on a stream of its own (default priority), so that sessions of
  // concurrent requests overlap on the device
  if (hipStreamCreateWithFlags(&S->s, hipStreamNonBlocking) != hipSuccess) {
    bq::set_error("session stream creation failed");
    delete S;
    return nullptr;
  }
  S->own_stream = true;
  hipStream_t s = S->s;
  const size_t ne1 = (size_t)m->Ne + 1;
  if (upload(&S->E, m->E, ne1 * 24, s) || upload(&S->ren, m->ren, ne1 * 12, s) ||
      upload(&S->nrm, m->nrm, ne1 * 12, s) || upload(&S->Ui, m->Ui, ne1 * 32, s) ||
      upload(&S->dUi, m->dUi, ne1 * 32, s) || upload(&S->kn, m->kn, (size_t)(m->MN + 1) * 2, s) ||
      upload(&S->eni, m->eni, ne1 * 4, s) || upload(&S->md, m->md, ne1, s) || upload(&S->sd, m->sd, ne1, s)) {
    bq_sess_destroy(S);
    return nullptr;
  }
  S->kn_h.assign(m->kn, m->kn + 2 * (m->MN + 1));
  for (int d = 0; d <= m->MN; ++d) {
    S->dNe.push_back(m->dNe[d]);
    S->sid_h.emplace_back(m->sid[d], m->sid[d] + m->dNe[d] + 1);
    int* p = nullptr;
    if (upload(&p, m->sid[d], (size_t)m->dNe[d] + 1, s)) {
      bq_sess_destroy(S);
      return nullptr;
    }
    S->sid.push_back(p);
    // live rows: all 4 per quad, 3 per triangle (shape flag at E[21])
    const long N = 4L * m->dNe[d];
    std::vector<int> map(N), lv;
    for (int t = 1; t <= m->dNe[d]; ++t) {
      const bool tri = m->E[24 * (size_t)abs(m->sid[d][t]) + 21] != 0.0;
      for (int tn = 0; tn < 4; ++tn) {
        const long r = 4L * (t - 1) + tn;
        if (tn == 3 && tri) {
          map[r] = -1;
        } else {
          map[r] = (int)lv.size();
          lv.push_back((int)r);
        }
      }
    }
    int *pm = nullptr, *pl = nullptr;
    if (upload(&pm, map.data(), map.size(), s) || upload(&pl, lv.data(), lv.size(), s)) {
      bq_sess_destroy(S);
      return nullptr;
    }
    S->rowmap.push_back(pm);
    S->live.push_back(pl);
    S->nlive.push_back((int)lv.size());
    S->two_nn += (int)lv.size();
    if (hipStreamSynchronize(s) != hipSuccess) {  // map and lv are freed at the end of the loop body
      bq::set_error("session set-up: stream synchronisation failed");
      bq_sess_destroy(S);
      return nullptr;
    }
  }
  S->G.assign(m->MN + 1, nullptr);
  S->H.assign(m->MN + 1, nullptr);
  S->dG.assign(m->MN + 1, nullptr);
  S->dH.assign(m->MN + 1, nullptr);
  S->tdF.assign(m->MN + 1, nullptr);
  if (bq_qplan_init(&S->plan, m->qd[0], m->qd + 1, m->qd + 4, m->qd + 6, 1.0e-15) && m->ps == 1) {
    bq::set_error("degenerate lattice");
    bq_sess_destroy(S);
    return nullptr;
  }
  // Gauss-Legendre tables as the host computed them for the model
  if (bq_set_quadrature(m->xli, m->wli, m->xhi, m->whi)) {
    bq_sess_destroy(S);
    return nullptr;
  }
  S->key_now = wave_key(m);
  return S;
}

int bq_sess_update(bq_sess* S, const bq_model* m)
{
  if (m->Ne != S->Ne || m->NN != S->NN || m->MN != S->MN || m->ps != S->ps) {
    bq::set_error("session update: a different mesh (Ne %d/%d, NN %d/%d)", m->Ne, S->Ne, m->NN, S->NN);
    return -2;
  }
  hipStream_t s = S->s;
  const size_t ne1 = (size_t)m->Ne + 1;
  S->m = *m;
  BQ_HIP(hipMemcpyAsync(S->Ui, m->Ui, ne1 * 32 * 8, hipMemcpyHostToDevice, s));
  BQ_HIP(hipMemcpyAsync(S->dUi, m->dUi, ne1 * 32 * 8, hipMemcpyHostToDevice, s));
  BQ_HIP(hipMemcpyAsync(S->kn, m->kn, (size_t)(m->MN + 1) * 16, hipMemcpyHostToDevice, s));
  BQ_HIP(hipStreamSynchronize(s));
  S->kn_h.assign(m->kn, m->kn + 2 * (m->MN + 1));
  if (bq_qplan_init(&S->plan, m->qd[0], m->qd + 1, m->qd + 4, m->qd + 6, 1.0e-15) && m->ps == 1) {
    bq::set_error("degenerate lattice");
    return -1;
  }
  S->key_now = wave_key(m);
  if (S->assembled && S->key_now == S->key_asm) return 1;
  S->red.prec = -1;  // the factors belong to the old matrices
  return 0;
}

int bq_sess_assembled(const bq_sess* S) { return (S->assembled && S->key_now == S->key_asm) ? 1 : 0; }

size_t bq_sess_bytes(const bq_sess* S) { return S->bytes + S->red.bytes + S->tmp_bytes; }

void bq_sess_destroy(bq_sess* S)
{
  if (!S) return;
  (void)hipStreamSynchronize(S->s);
  bq::reduced_free(S);
  if (S->work) (void)hipFree(S->work);
  if (S->qerr) (void)hipFree(S->qerr);
  if (S->wtab) (void)hipFree(S->wtab);
  auto f = [](void* p) { if (p) (void)hipFree(p); };
  f(S->E); f(S->ren); f(S->nrm); f(S->Ui); f(S->dUi); f(S->kn); f(S->eni); f(S->md); f(S->sd);
  for (auto p : S->sid) f(p);
  for (auto p : S->rowmap) f(p);
  for (auto p : S->live) f(p);
  for (auto p : S->G) f(p);
  for (auto p : S->H) f(p);
  for (auto p : S->dG) f(p);
  for (auto p : S->dH) f(p);
  for (auto p : S->tdF) f(p);
  if (S->s2) {
    (void)hipStreamSynchronize(S->s2);
    bq_lu_release_stream(S->s2);
    (void)hipStreamDestroy(S->s2);
  }
  if (S->own_stream) {
    bq_lu_release_stream(S->s);
    (void)hipStreamDestroy(S->s);
  }
  delete S;
}

const bq_stats* bq_sess_stats(const bq_sess* S) { return &S->st; }

int bq_sess_assemble(bq_sess* S)
{
  hipStream_t s = S->s;
  hipEvent_t e0, e1;
  BQ_HIP(hipEventCreate(&e0));
  BQ_HIP(hipEventCreate(&e1));
  BQ_HIP(hipEventRecord(e0, s));
  size_t wmax = 0;
  for (int d = 0; d <= S->MN; ++d) wmax = std::max(wmax, bq_assemble_work_bytes(S->dNe[d]));
  if (!S->work) {
    BQ_HIP(hipMalloc(&S->work, wmax));
    S->work_bytes = wmax;
    S->bytes += wmax;
  }
  if (tune_plan(S, asm_tolerance()) || build_wtab(S)) return -1;
  if (!S->qerr) BQ_HIP(hipMalloc((void**)&S->qerr, sizeof(int)));
  BQ_HIP(hipMemsetAsync(S->qerr, 0, sizeof(int), s));
  S->plan.err = S->qerr;
  S->st.qpgf_error = 0;
  void* work = S->work;
  S->assembled = false;
  S->red.prec = -1;
  S->st.n_gh_pairs = 0;
  S->st.n_qpgf = S->st.n_qpgf_gl = S->st.n_qpgf_gh = S->st.n_qpgf_sing = 0;
  for (int d = 0; d <= S->MN; ++d) {
    const long dNe = S->dNe[d], N = 4 * dNe;
    if (dNe == 0) continue;
    const bool periodic = (d == 0 && S->ps == 1);
    if (!S->G[d]) {
      if (dmalloc(&S->G[d], (size_t)(2 * N * N)) || dmalloc(&S->H[d], (size_t)(2 * N * N))) return -1;
      S->bytes += (size_t)32 * N * N;
      if (!periodic) {
        if (dmalloc(&S->dG[d], (size_t)(4 * N * N)) || dmalloc(&S->dH[d], (size_t)(4 * N * N)) ||
            dmalloc(&S->tdF[d], (size_t)(3 * N)))
          return -1;
        S->bytes += (size_t)64 * N * N + 24 * N;
      }
    }
    bq_adom a{};
    a.dNe = (int)dNe;
    a.periodic = periodic ? 1 : 0;
    a.open_domain = (d == 0) ? 1 : 0;
    a.kr = S->kn_h[2 * d];
    a.ki = S->kn_h[2 * d + 1];
    a.plan = S->plan;
    a.E = S->E;
    a.ren = S->ren;
    a.sid = S->sid[d];
    a.G = S->G[d];
    a.H = S->H[d];
    a.dG = S->dG[d];
    a.dH = S->dH[d];
    a.tdF = S->tdF[d];
    int ngh = 0;
    long long npts[3] = {0, 0, 0};
    a.n_gh_out = &ngh;
    a.n_points_out = npts;
    // BQ_FULL_RULE=1: the reference's 12-point rule for every regular periodic pair
    a.full_rule = (getenv("BQ_FULL_RULE") && atoi(getenv("BQ_FULL_RULE")) != 0) ? 1 : 0;
    if (bq_assemble_domain_dev(&a, work, s)) return -1;
    S->st.n_gh_pairs += ngh;
    if (periodic) {
      S->st.n_qpgf_gl += npts[0];
      S->st.n_qpgf_gh += npts[1];
      S->st.n_qpgf_sing += npts[2];
      S->st.n_qpgf += npts[0] + npts[1] + npts[2];
    }
  }
  BQ_HIP(hipEventRecord(e1, s));
  BQ_HIP(hipEventSynchronize(e1));
  float ms = 0;
  BQ_HIP(hipEventElapsedTime(&ms, e0, e1));
  S->st.t_assemble_ms = ms;
  {
    int qe = 0;
    BQ_HIP(hipMemcpyAsync(&qe, S->qerr, sizeof(int), hipMemcpyDeviceToHost, s));
    BQ_HIP(hipStreamSynchronize(s));
    if (qe) {
      // the reference's callers print the QPGF code and exit (q0_bil_dcoef_v2_cc.c:177-180)
      S->st.qpgf_error = qe;
      bq::set_error("d3hm_qpgf_d2_v1() error. err=%d (a reciprocal order has beta_n == 0)", qe);
      (void)hipEventDestroy(e0);
      (void)hipEventDestroy(e1);
      return qe;
    }
  }
  S->assembled = true;
  S->key_asm = S->key_now;
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  return 0;
}

// Precision of the global solve: 0 = FP64 LU (default), 1 = FP32 LU + FP64 refinement,
// 2 = mode 1 for block systems of order 2 NN >= kMixedMinOrder, else mode 0 (below it the LUs
// are bound by the latency of their panel chain, not by flops, and the refinement's host
// round trips cost more than the FP32 factors save).  BQ_MIXED=1 or 2 sets the initial mode.
static int g_precision = -1;
static const int kMixedMinOrder = 16384;

extern "C" double bq_assembly_tolerance(double tol)
{
  asm_tolerance();
  if (tol == 0.0 || (tol > 0.0 && tol < 1e-3)) g_asm_tol = tol;
  return g_asm_tol;
}

extern "C" int bq_solve_precision(int mode)
{
  if (g_precision < 0) {
    const char* e = getenv("BQ_MIXED");
    const int v = e ? atoi(e) : 0;
    g_precision = (v == 1 || v == 2) ? v : 0;
  }
  if (mode >= 0 && mode <= 2) g_precision = mode;
  return g_precision;
}

// rows [row0, row0 + rows) of L against its columns [col0, col0 + ncol)
struct SchurBlock {
  int row0, rows, col0, ncol;
};

// Column groups of K = R[:, :NN] for the forward solve of Z = M^-1 K.  K has the nonzero
// blocks of L (k_build_blocks writes both under the same condition), so a block's columns are
// zero outside its rows; after the factorization's row permutation they are zero above the
// first position any of those rows was moved to.  Gaps between blocks and the columns from
// ncol_blocks to ncols (right-hand sides b_c) form groups that start at row 0.  Returns false
// (no groups) if the blocks' column ranges overlap.
static bool k_groups(hipStream_t s, const int* perm_dev, int n2, int NN, int ncols, const std::vector<SchurBlock>& lblk,
                     std::vector<int>& g0, std::vector<int>& gn, std::vector<int>& gf)
{
  std::vector<int> perm(n2);
  if (hipMemcpyAsync(perm.data(), perm_dev, (size_t)n2 * sizeof(int), hipMemcpyDeviceToHost, s) != hipSuccess ||
      hipStreamSynchronize(s) != hipSuccess)
    return false;
  std::vector<SchurBlock> b = lblk;
  std::sort(b.begin(), b.end(), [](const SchurBlock& x, const SchurBlock& y) { return x.col0 < y.col0; });
  std::vector<int> first(b.size(), n2);
  for (int i = 0; i < n2; ++i)
    for (size_t q = 0; q < b.size(); ++q)
      if (perm[i] >= b[q].row0 && perm[i] < b[q].row0 + b[q].rows && i < first[q]) first[q] = i;
  int c = 0;
  for (size_t q = 0; q < b.size(); ++q) {
    if (b[q].col0 < c || b[q].col0 + b[q].ncol > NN) return false;
    if (b[q].col0 > c) {
      g0.push_back(c);
      gn.push_back(b[q].col0 - c);
      gf.push_back(0);
    }
    g0.push_back(b[q].col0);
    gn.push_back(b[q].ncol);
    gf.push_back(first[q]);
    c = b[q].col0 + b[q].ncol;
  }
  if (c < ncols) {
    g0.push_back(c);
    gn.push_back(ncols - c);
    gf.push_back(0);
  }
  return true;
}

// Y = (y_0, y_1, y_2, y_3), each of length 2 NN (unknowns, then normal derivatives) -> the
// reference's X layout: X[c NN + e] (c = 0..3) and X[(c + 4) NN + e]
static void unpack_global(int NN, const std::vector<double>& Y, double* Xh)
{
  const size_t n2 = 2 * (size_t)NN;
  for (int c = 0; c < 4; ++c)
    for (int e = 0; e < NN; ++e) {
      const double* y = Y.data() + 2 * ((size_t)c * n2);
      Xh[2 * ((size_t)c * NN + e)] = y[2 * e];
      Xh[2 * ((size_t)c * NN + e) + 1] = y[2 * e + 1];
      Xh[2 * ((size_t)(c + 4) * NN + e)] = y[2 * (NN + e)];
      Xh[2 * ((size_t)(c + 4) * NN + e) + 1] = y[2 * (NN + e) + 1];
    }
}

// The block system of bq_sess_solve_global,
//   M y_c + K T_c(y_3) = b_c (c < 3),  L sum_c n_c o (y_c)_top + M3 y_3 = b_3,
// solved by iterative refinement in FP64 with the same block elimination in complex FP32
// as the correction solver: FP32 LUs of M and of the Schur complement S = M3 - L W
// (W = Z_top o n n^T, Z = M^-1 K), residuals with the FP64 blocks (Wilkinson; Carson and
// Higham, SIAM J. Sci. Comput. 40(2), 2018).  M, M3, R = [K | b_0 b_1 b_2], L, b3 are the
// FP64 blocks as built (left unchanged); lblk: the blocks of L that hold its nonzeros.
// Returns 0 with Y (4 x 2 NN complex) once the
// correction max|dx| / max|x| is below 1e-15 or stops shrinking below 1e-10; 1 if it stops
// shrinking above that, an FP32 factor is singular or overflows (the caller falls back to
// FP64 factors); < 0 on error.
static int solve_global_mixed(bq_sess* S, const double* M, const double* M3, const double* R, const double* L,
                              const double* b3, const double* nn3, const std::vector<SchurBlock>& lblk,
                              std::vector<double>& Y)
{
  hipStream_t s = S->s;
  const int NN = S->NN, n2 = 2 * NN;
  const size_t sq = (size_t)n2 * n2 * 2, rect = (size_t)n2 * NN * 2, vec4 = (size_t)n2 * 4 * 2;
  float *M32, *S32, *R32, *L32, *W32, *r32, *v32, *T32, *inv1, *inv2;
  double *x, *r, *Tx, *vx;
  int *ip1, *ip2;
  struct Frees {
    std::vector<void*> p;
    ~Frees()
    {
      for (void* q : p) (void)hipFree(q);
    }
  } held;  // every buffer of this function is released on each return
  auto fmal = [&](float** p, size_t n) {
    *p = nullptr;
    if (hipMalloc((void**)p, n * sizeof(float)) != hipSuccess) return true;
    held.p.push_back(*p);
    return false;
  };
  auto dmal = [&](auto** p, size_t n) {
    *p = nullptr;
    if (dmalloc(p, n)) return true;
    held.p.push_back((void*)*p);
    return false;
  };
  if (fmal(&M32, sq) || fmal(&S32, sq) || fmal(&R32, rect + (size_t)n2 * 2) || fmal(&L32, rect) ||
      fmal(&W32, (size_t)NN * NN * 2) || fmal(&r32, vec4) || fmal(&v32, (size_t)NN * 2) || fmal(&T32, (size_t)NN * 6) ||
      fmal(&inv1, (size_t)bq_lu_inv_doubles(n2)) || fmal(&inv2, (size_t)bq_lu_inv_doubles(n2)) ||
      dmal(&x, vec4) || dmal(&r, vec4) || dmal(&Tx, (size_t)NN * 6) || dmal(&vx, (size_t)NN * 2) ||
      dmal(&ip1, (size_t)n2) || dmal(&ip2, (size_t)n2)) {
    bq::set_error("solve_global: device memory for the mixed-precision solve");
    return -1;
  }
  const float onef[2] = {1, 0}, monef[2] = {-1, 0};
  const double one[2] = {1, 0}, mone[2] = {-1, 0};
  // FP32 factors: M, then Z = M^-1 K and S = M3 - L W
  k_d2f<<<nblk((long)sq, 256), 256, 0, s>>>((long)sq, M, M32);
  int info = bq_cgetrf_dev(s, n2, M32, n2, ip1, inv1);
  if (info) return 1;
  k_d2f<<<nblk((long)rect, 256), 256, 0, s>>>((long)rect, R, R32);
  {
    std::vector<int> g0, gn, gf;
    const bool grp = k_groups(s, bq_lu_perm_dev(inv1, n2, 1), n2, NN, NN, lblk, g0, gn, gf);
    if (bq_cgetrs_groups_dev(s, n2, NN, M32, n2, ip1, inv1, R32, n2, grp ? (int)g0.size() : 0, g0.data(), gn.data(),
                             gf.data()))
      return -1;
  }
  k_hadamard_t<float><<<nblk((long)NN * NN, 256), 256, 0, s>>>(NN, R32, nn3, W32);
  k_d2f<<<nblk((long)rect, 256), 256, 0, s>>>((long)rect, L, L32);
  k_d2f<<<nblk((long)sq, 256), 256, 0, s>>>((long)sq, M3, S32);
  BQ_HIP(hipGetLastError());
  {
    const int kt = bq::ktime_begin("schur_cgemm", s);
    for (const SchurBlock& b : lblk)
      if (bq_cgemm_op_dev(s, 0, 0, b.rows, NN, b.ncol, monef, L32 + 2 * ((size_t)b.col0 * n2 + b.row0), n2,
                          W32 + 2 * (size_t)b.col0, NN, onef, S32 + 2 * (size_t)b.row0, n2))
        return -1;
    bq::ktime_end(kt, s);
  }
  info = bq_cgetrf_dev(s, n2, S32, n2, ip2, inv2);
  if (info) return 1;
  // refinement: x = 0, r = b
  BQ_HIP(hipMemsetAsync(x, 0, vec4 * 8, s));
  BQ_HIP(hipMemcpyAsync(r, R + rect, (size_t)n2 * 3 * 16, hipMemcpyDeviceToDevice, s));
  BQ_HIP(hipMemcpyAsync(r + (size_t)n2 * 6, b3, (size_t)n2 * 16, hipMemcpyDeviceToDevice, s));
  std::vector<double> hx(vec4), hd(vec4);
  const int kMaxIt = 12;
  const double kStall = 1e-10;
  double prev = 1e300;
  int rc = 1;
  const int kt_ir = bq::ktime_begin("refinement", s);
  for (int it = 0; it < kMaxIt; ++it) {
    // correction d = P r in FP32 (the block elimination), written over r32
    const int kt_c = bq::ktime_begin("ir_correction", s);
    k_d2f<<<nblk((long)vec4, 256), 256, 0, s>>>((long)vec4, r, r32);
    if (bq_cgetrs_dev(s, n2, 3, M32, n2, ip1, inv1, r32, n2)) return -1;
    k_normal_dot_t<float><<<nblk(NN, 256), 256, 0, s>>>(NN, r32, nn3, v32);
    if (bq_cgemm_op_dev(s, 0, 0, n2, 1, NN, monef, L32, n2, v32, NN, onef, r32 + (size_t)n2 * 6, n2)) return -1;
    if (bq_cgetrs_dev(s, n2, 1, S32, n2, ip2, inv2, r32 + (size_t)n2 * 6, n2)) return -1;
    k_scale_normals_t<float><<<nblk(NN, 256), 256, 0, s>>>(NN, r32 + (size_t)n2 * 6, nn3, T32);
    if (bq_cgemm_op_dev(s, 0, 0, n2, 3, NN, monef, R32, n2, T32, NN, onef, r32, n2)) return -1;
    k_add_f2d<<<nblk((long)vec4, 256), 256, 0, s>>>((long)vec4, r32, x);
    BQ_HIP(hipGetLastError());
    bq::ktime_end(kt_c, s);
    // sizes of the correction and of x
    std::vector<float> hdf(vec4);
    BQ_HIP(hipMemcpyAsync(hdf.data(), r32, vec4 * sizeof(float), hipMemcpyDeviceToHost, s));
    BQ_HIP(hipMemcpyAsync(hx.data(), x, vec4 * 8, hipMemcpyDeviceToHost, s));
    BQ_HIP(hipStreamSynchronize(s));
    double dmax = 0.0, xmax = 0.0;
    bool finite = true;
    for (size_t i = 0; i < vec4; ++i) {
      const double d = fabs((double)hdf[i]), a = fabs(hx[i]);
      finite = finite && d <= DBL_MAX && a <= DBL_MAX;
      dmax = d > dmax ? d : dmax;
      xmax = a > xmax ? a : xmax;
    }
    if (!finite) {  // an FP32 factor overflowed: the FP64 factors take over
      rc = 1;
      break;
    }
    const double corr = xmax > 0.0 ? dmax / xmax : 0.0;
    S->st.ir_iterations = it + 1;
    S->st.ir_correction = corr;
    if (it > 0 && corr <= 1e-15) {
      rc = 0;
      break;
    }
    if (it > 1 && corr > 0.5 * prev) {
      // Stalled: the correction stopped shrinking.  At a small level that is the rounding
      // noise of the FP64 residual (the limiting accuracy of refinement, no worse than the
      // FP64 factors' own); at a large one the FP32 factors are too inaccurate.
      rc = corr <= kStall ? 0 : 1;
      break;
    }
    prev = corr;
    // residual r = b - A x with the FP64 blocks
    const int kt_r = bq::ktime_begin("ir_residual", s);
    BQ_HIP(hipMemcpyAsync(r, R + rect, (size_t)n2 * 3 * 16, hipMemcpyDeviceToDevice, s));
    BQ_HIP(hipMemcpyAsync(r + (size_t)n2 * 6, b3, (size_t)n2 * 16, hipMemcpyDeviceToDevice, s));
    if (bq_zgemm_skinny_dev(s, n2, 3, n2, M, n2, x, n2, r, n2)) return -1;
    k_scale_normals_t<double><<<nblk(NN, 256), 256, 0, s>>>(NN, x + (size_t)n2 * 6, nn3, Tx);
    if (bq_zgemm_skinny_dev(s, n2, 3, NN, R, n2, Tx, NN, r, n2)) return -1;
    k_normal_dot_t<double><<<nblk(NN, 256), 256, 0, s>>>(NN, x, nn3, vx);
    if (bq_zgemm_skinny_dev(s, n2, 1, NN, L, n2, vx, NN, r + (size_t)n2 * 6, n2)) return -1;
    if (bq_zgemm_skinny_dev(s, n2, 1, n2, M3, n2, x + (size_t)n2 * 6, n2, r + (size_t)n2 * 6, n2)) return -1;
    BQ_HIP(hipGetLastError());
    bq::ktime_end(kt_r, s);
  }
  bq::ktime_end(kt_ir, s);
  if (rc == 1 && S->st.ir_correction <= kStall && S->st.ir_iterations == kMaxIt) rc = 0;
  if (rc == 0) Y.assign(hx.begin(), hx.end());
  return rc;
}

// Single object: the reduction of bq_reduced.hip (factors kept for later solves of the same
// matrices); FP64, or FP32 factors with FP64 GMRES refinement, falling back to FP64 factors
// when the refinement stalls above a backward error of 1e-10.
static int solve_single(bq_sess* S, double* Xh, bool mixed)
{
  hipStream_t s = S->s;
  hipEvent_t e0, e1;
  BQ_HIP(hipEventCreate(&e0));
  BQ_HIP(hipEventCreate(&e1));
  BQ_HIP(hipEventRecord(e0, s));
  S->st.path = 1;
  S->st.t_factor_ms = 0.0;
  const int want = mixed ? 1 : 0;
  int rc = 0;
  if (S->red.prec != want) rc = bq::reduced_factor(S, want);
  if (rc == 0) rc = bq::reduced_solve(S, Xh);
  if (rc > 0 && want == 1) {  // refinement stalled: FP64 factors
    rc = bq::reduced_factor(S, 0);
    if (rc == 0) rc = bq::reduced_solve(S, Xh);
  }
  S->st.mixed = S->red.prec == 1 ? 1 : 0;
  BQ_HIP(hipEventRecord(e1, s));
  BQ_HIP(hipEventSynchronize(e1));
  float ms = 0.f;
  BQ_HIP(hipEventElapsedTime(&ms, e0, e1));
  S->st.t_solve_ms = ms;
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (rc > 0) rc = -3;
  return rc;
}

int bq_sess_solve_global(bq_sess* S, double* Xh)
{
  hipStream_t s = S->s;
  const int NN = S->NN, n2 = 2 * NN;
  if (S->two_nn != n2) {
    bq::set_error("solve_global: element-node count mismatch (%d live rows vs 2 NN = %d)", S->two_nn, n2);
    return -2;
  }
  S->st.mixed = 0;
  S->st.ir_iterations = 0;
  S->st.ir_correction = 0.0;
  S->st.gmres_iterations = 0;
  S->st.backward_error = 0.0;
  {
    const int prec = bq_solve_precision(-1);
    const bool mixed = prec == 1 || (prec == 2 && n2 >= kMixedMinOrder);
    if (bq::single_object(S) && !getenv("BQ_BLOCK_ELIM")) return solve_single(S, Xh, mixed);
  }
  S->st.path = 0;
  S->st.t_factor_ms = 0.0;
  // the block system's buffers live for this call only; every return releases them
  struct Held {
    std::vector<void*> p;
    hipEvent_t e[2] = {nullptr, nullptr};
    ~Held()
    {
      for (void* q : p) (void)hipFree(q);
      for (hipEvent_t x : e)
        if (x) (void)hipEventDestroy(x);
    }
  } held;
  auto dm = [&](auto** p, size_t n) {
    if (dmalloc(p, n)) return true;
    held.p.push_back((void*)*p);
    return false;
  };
  BQ_HIP(hipEventCreate(&held.e[0]));
  BQ_HIP(hipEventCreate(&held.e[1]));
  hipEvent_t e0 = held.e[0], e1 = held.e[1];
  BQ_HIP(hipEventRecord(e0, s));
  const size_t sq = (size_t)n2 * n2 * 2, rect = (size_t)n2 * NN * 2;
  double *M, *M3, *R, *L, *b3, *W, *v, *T, *inv1, *inv2;
  int *ip1, *ip2;
  double* nn3 = nullptr;
  const size_t ninv = (size_t)bq_lu_inv_doubles(n2);
  if (dm(&M, sq) || dm(&M3, sq) || dm(&R, (size_t)n2 * (NN + 3) * 2) || dm(&L, rect) || dm(&b3, (size_t)n2 * 2) ||
      dm(&W, (size_t)NN * NN * 2) || dm(&v, (size_t)NN * 2) || dm(&T, (size_t)NN * 6) || dm(&ip1, (size_t)n2) ||
      dm(&ip2, (size_t)n2) || dm(&nn3, (size_t)NN * 3) || dm(&inv1, ninv) || dm(&inv2, ninv))
    return -1;
  {
    // peak of this call's temporaries (with the FP32 copies of the mixed solve), reported
    // with the session's resident bytes so that callers admitting sessions see the peak
    size_t t = 8 * (2 * sq + (size_t)n2 * (NN + 3) * 2 + rect + (size_t)n2 * 2 + (size_t)NN * NN * 2 +
                    (size_t)NN * 11 + 2 * ninv) + 8 * (size_t)n2;
    const int prec = bq_solve_precision(-1);
    if (prec == 1 || (prec == 2 && n2 >= kMixedMinOrder)) t += 4 * (2 * sq + 2 * rect + (size_t)NN * NN * 2);
    S->tmp_bytes = std::max(S->tmp_bytes, t);
  }
  // unit normals per element node e (main side)
  {
    std::vector<double> h((size_t)NN * 3, 0.0);
    for (int i = 1; i <= S->Ne; ++i)
      for (int l = 0; l < 4; ++l) {
        int e = S->m.eni[4 * i + l];
        if (e < 0) continue;
        for (int c = 0; c < 3; ++c) h[3 * (size_t)e + c] = S->m.nrm[12 * (size_t)i + 3 * l + c];
      }
    BQ_HIP(hipMemcpyAsync(nn3, h.data(), h.size() * sizeof(double), hipMemcpyHostToDevice, s));
    BQ_HIP(hipStreamSynchronize(s));
  }
  BQ_HIP(hipMemsetAsync(M, 0, sq * 8, s));
  BQ_HIP(hipMemsetAsync(M3, 0, sq * 8, s));
  BQ_HIP(hipMemsetAsync(R, 0, (size_t)n2 * (NN + 3) * 16, s));
  BQ_HIP(hipMemsetAsync(L, 0, rect * 8, s));
  int rowbase = 0;
  const int kt_b = bq::ktime_begin("build_blocks", s);
  for (int d = 0; d <= S->MN; ++d) {
    const long dNe = S->dNe[d], N = 4 * dNe;
    if (dNe == 0) continue;
    k_build_blocks<<<nblk(N * dNe, 256), 256, 0, s>>>((int)dNe, d, rowbase, NN, S->rowmap[d], S->sid[d], S->eni,
                                                       S->md, S->sd, S->kn, S->G[d], S->H[d], M, M3, R, L);
    BQ_HIP(hipGetLastError());
    k_build_rhs<<<nblk(N, 128), 128, 0, s>>>((int)dNe, d, rowbase, NN, S->rowmap[d], S->sid[d], S->md, S->sd, S->kn,
                                             S->nrm, S->Ui, S->dUi, S->G[d], S->H[d], R, b3);
    BQ_HIP(hipGetLastError());
    rowbase += S->nlive[d];
  }
  bq::ktime_end(kt_b, s);
  // L couples a row of domain d to X_3 of the nodes of the elements whose sub-domain is d (the
  // only entries k_build_blocks writes): per domain, its rows against one column range.  The
  // Schur product L W runs on those blocks only; the rest of L is exactly zero.
  std::vector<SchurBlock> lblk;
  {
    int rb = 0;
    for (int d = 0; d <= S->MN; ++d) {
      int cmin = INT_MAX, cmax = -1;
      for (int i = 1; i <= S->m.dNe[d]; ++i) {
        const int e = abs(S->m.sid[d][i]);
        if (S->m.md[e] == d) continue;
        for (int l = 0; l < 4; ++l) {
          const int c = S->m.eni[4 * e + l];
          if (c < 0) continue;
          cmin = std::min(cmin, c);
          cmax = std::max(cmax, c);
        }
      }
      if (cmax >= 0 && S->nlive[d] > 0) lblk.push_back(SchurBlock{rb, S->nlive[d], cmin, cmax - cmin + 1});
      rb += S->nlive[d];
    }
  }
  const double one[2] = {1, 0}, mone[2] = {-1, 0};
  S->st.mixed = 0;
  S->st.ir_iterations = 0;
  S->st.ir_correction = 0.0;
  const int prec = bq_solve_precision(-1);
  if (prec == 1 || (prec == 2 && n2 >= kMixedMinOrder)) {
    std::vector<double> Y;
    int rc = solve_global_mixed(S, M, M3, R, L, b3, nn3, lblk, Y);
    if (rc < 0) return rc;
    if (rc == 0) {
      BQ_HIP(hipEventRecord(e1, s));
      BQ_HIP(hipStreamSynchronize(s));
      float ms = 0;
      BQ_HIP(hipEventElapsedTime(&ms, e0, e1));
      S->st.t_solve_ms = ms;
      S->st.mixed = 1;
      unpack_global(NN, Y, Xh);
      return 0;
    }
    // rc > 0: the refinement did not converge; the FP64 factorization below solves it
  }
  int info = bq_zgetrf_dev(s, n2, M, n2, ip1, inv1);
  if (info) {
    bq::set_error("solve_global: block matrix M is singular (info %d)", info);
    return info > 0 ? -3 : info;
  }
  {
    std::vector<int> g0, gn, gf;
    const bool grp = k_groups(s, bq_lu_perm_dev(inv1, n2, 0), n2, NN, NN + 3, lblk, g0, gn, gf);
    if (bq_zgetrs_groups_dev(s, n2, NN + 3, M, n2, ip1, inv1, R, n2, grp ? (int)g0.size() : 0, g0.data(), gn.data(),
                             gf.data()))
      return -1;
  }
  k_hadamard_normals<<<nblk((long)NN * NN, 256), 256, 0, s>>>(NN, R, nn3, W);
  BQ_HIP(hipGetLastError());
  {
    const int kt = bq::ktime_begin("schur_zgemm", s);
    for (const SchurBlock& b : lblk)
      if (bq_zgemm_op_dev(s, 0, 0, b.rows, NN, b.ncol, mone, L + 2 * ((size_t)b.col0 * n2 + b.row0), n2,
                          W + 2 * (size_t)b.col0, NN, one, M3 + 2 * (size_t)b.row0, n2))
        return -1;
    bq::ktime_end(kt, s);
  }
  k_normal_dot<<<nblk(NN, 256), 256, 0, s>>>(NN, R, nn3, v);
  BQ_HIP(hipGetLastError());
  if (bq_zgemm_op_dev(s, 0, 0, n2, 1, NN, mone, L, n2, v, NN, one, b3, n2)) return -1;
  info = bq_zgetrf_dev(s, n2, M3, n2, ip2, inv2);
  if (info) {
    bq::set_error("solve_global: Schur complement is singular (info %d)", info);
    return info > 0 ? -3 : info;
  }
  if (bq_zgetrs_dev(s, n2, 1, M3, n2, ip2, inv2, b3, n2)) return -1;
  k_scale_normals<<<nblk(NN, 256), 256, 0, s>>>(NN, b3, nn3, T);
  BQ_HIP(hipGetLastError());
  // y_c -= Z T_c for c = 0..2 (columns NN..NN+2 of R)
  if (bq_zgemm_op_dev(s, 0, 0, n2, 3, NN, mone, R, n2, T, NN, one, R + (size_t)NN * n2 * 2, n2)) return -1;
  std::vector<double> Y((size_t)n2 * 4 * 2);
  BQ_HIP(hipMemcpyAsync(Y.data(), R + (size_t)NN * n2 * 2, (size_t)n2 * 3 * 16, hipMemcpyDeviceToHost, s));
  BQ_HIP(hipMemcpyAsync(Y.data() + (size_t)n2 * 3 * 2, b3, (size_t)n2 * 16, hipMemcpyDeviceToHost, s));
  BQ_HIP(hipEventRecord(e1, s));
  BQ_HIP(hipStreamSynchronize(s));
  float ms = 0;
  BQ_HIP(hipEventElapsedTime(&ms, e0, e1));
  S->st.t_solve_ms = ms;
  unpack_global(NN, Y, Xh);
  return 0;
}

int bq_sess_dudt(bq_sess* S, int d, const double* U, const double* dU, double* out)
{
  hipStream_t s = S->s;
  const long N = 4L * S->dNe[d];
  if (!S->dG[d]) {
    bq::set_error("dudt: domain %d has no derivative matrices", d);
    return -2;
  }
  double *du, *u, *o;
  if (upload(&du, dU, (size_t)N * 8, s) || upload(&u, U, (size_t)N * 8, s) || dmalloc(&o, (size_t)2 * N * 8))
    return -1;
  const double one[2] = {1, 0}, mone[2] = {-1, 0}, zero[2] = {0, 0};
  // dG, dH are row-major (2N x N) = column-major (N x 2N) transposed
  if (bq_zgemm_op_dev(s, 1, 0, (int)(2 * N), 4, (int)N, one, S->dG[d], N, du, N, zero, o, 2 * N)) return -1;
  if (bq_zgemm_op_dev(s, 1, 0, (int)(2 * N), 4, (int)N, mone, S->dH[d], N, u, N, one, o, 2 * N)) return -1;
  BQ_HIP(hipMemcpyAsync(out, o, (size_t)2 * N * 4 * 16, hipMemcpyDeviceToHost, s));
  BQ_HIP(hipStreamSynchronize(s));
  (void)hipFree(du);
  (void)hipFree(u);
  (void)hipFree(o);
  return 0;
}

int bq_sess_deh(bq_sess* S, int d, const double* EH, double* dEH)
{
  if (S->st.path == 1 && S->red.prec >= 0 && d <= 1) return bq::reduced_deh(S, d, EH, dEH);
  hipStream_t s = S->s;
  const long N = 4L * S->dNe[d];
  const int n = S->nlive[d];
  double *eh, *b, *bl, *gt, *inv;
  int* ip;
  if (upload(&eh, EH, (size_t)N * 12, s) || dmalloc(&b, (size_t)N * 12) || dmalloc(&bl, (size_t)n * 12) ||
      dmalloc(&gt, (size_t)n * n * 2) || dmalloc(&ip, (size_t)n) || dmalloc(&inv, (size_t)bq_lu_inv_doubles(n)))
    return -1;
  const double one[2] = {1, 0}, zero[2] = {0, 0};
  // H [E H] over all rows (empty rows and columns contribute 0), then the live rows only
  if (bq_zgemm_op_dev(s, 1, 0, (int)N, 6, (int)N, one, S->H[d], N, eh, N, zero, b, N)) return -1;
  k_live_rows<<<nblk((long)n * 6, 256), 256, 0, s>>>(N, n, 6, S->live[d], b, bl, 1);
  BQ_HIP(hipGetLastError());
  k_gather_live<<<nblk((long)n * n, 256), 256, 0, s>>>(N, n, S->live[d], S->G[d], gt);
  BQ_HIP(hipGetLastError());
  int info = bq_zgetrf_dev(s, n, gt, n, ip, inv);
  if (info) {
    bq::set_error("solve_deh: G of domain %d is singular (info %d)", d, info);
    return info > 0 ? -3 : info;
  }
  if (bq_zgetrs_dev(s, n, 6, gt, n, ip, inv, bl, n)) return -1;
  BQ_HIP(hipMemsetAsync(b, 0, (size_t)N * 6 * 16, s));
  k_live_rows<<<nblk((long)n * 6, 256), 256, 0, s>>>(N, n, 6, S->live[d], b, bl, 0);
  BQ_HIP(hipGetLastError());
  BQ_HIP(hipMemcpyAsync(dEH, b, (size_t)N * 6 * 16, hipMemcpyDeviceToHost, s));
  BQ_HIP(hipStreamSynchronize(s));
  (void)hipFree(eh);
  (void)hipFree(b);
  (void)hipFree(bl);
  (void)hipFree(gt);
  (void)hipFree(ip);
  (void)hipFree(inv);
  return 0;
}

int bq_sess_get_GH(bq_sess* S, int d, double* G, double* H)
{
  const size_t N = 4 * (size_t)S->dNe[d];
  if (G) BQ_HIP(hipMemcpyAsync(G, S->G[d], N * N * 16, hipMemcpyDeviceToHost, S->s));
  if (H) BQ_HIP(hipMemcpyAsync(H, S->H[d], N * N * 16, hipMemcpyDeviceToHost, S->s));
  BQ_HIP(hipStreamSynchronize(S->s));
  return 0;
}

int bq_sess_get_dGH(bq_sess* S, int d, double* dG, double* dH, double* tdF)
{
  const size_t N = 4 * (size_t)S->dNe[d];
  if (!S->dG[d]) return -2;
  if (dG) BQ_HIP(hipMemcpyAsync(dG, S->dG[d], 2 * N * N * 16, hipMemcpyDeviceToHost, S->s));
  if (dH) BQ_HIP(hipMemcpyAsync(dH, S->dH[d], 2 * N * N * 16, hipMemcpyDeviceToHost, S->s));
  if (tdF) BQ_HIP(hipMemcpyAsync(tdF, S->tdF[d], 3 * N * 8, hipMemcpyDeviceToHost, S->s));
  BQ_HIP(hipStreamSynchronize(S->s));
  return 0;
}

}  // extern "C"
