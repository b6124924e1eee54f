// bq_assemble.hip -- coefficient-matrix assembly for one domain on the GPU.
//
// For collocation node (t, tn) (row 4(t-1)+tn) and element s (columns 4(s-1)..+3) of a
// domain, the reference computes (d3qd2_src/d3qd2_solve_bieq.c:165-256):
//   G  = int Gk N_n |W|,  H = int dGk/dn N_n,  F = solid angle / 4 pi,
// plus for the free-space domains the tangential derivatives dG/dt, dH/dt, dF/dt.
//   domain 0, periodic:  q0_dcoef_bd_node_t2_cc (d3qd2_src/d3qd2_elem.c:830-886)
//   other domains:       dcoef_bd_node_t2       (d3qd2_src/d3qd2_elem.c:515-572)
// with the 12x12 Gauss rule (PREC_DEF_BE = 2), a 64x64 rule when |dF/dt| > 0.5 on the
// 12x12 rule (CBD_SDF), and polar singular rules when t == s.
//
// GPU organisation (one launch per phase, all on one stream):
//   k_classify   one thread per (row, s): F and dF/dt with the pair's Gauss order (no Green's
//                function needed), flags the pair GL / GH / singular / no such row (the
//                fourth slot of a linear triangle).
//   k_q0_gl      one wavefront per element pair (t, s): its 4 collocation nodes x the pair's
//                n x n Gauss points (n from 3 to 10, or 12), lattice sums uniform over the
//                wavefront; one QPGF per Gauss point feeds all four interpolation functions
//                (the reference evaluates it 4 times).  Fixed-order sums over each node's
//                lanes, one row of G and H written per node.
//   k_q0_gh      one 256-thread block per flagged pair: 4096 points.
//   k_q0_sing    one block per collocation node: polar rule (4 sub-triangles x 128 angles
//                x 12 radii), F by the polar rule and dF/dt by central differences.
//   k_fs_gl/_gh/_sing   same for the free-space Helmholtz kernel with complex k, adding
//                dG/dt and dH/dt; the singular derivatives use the principal-value split
//                of d3qd2_src/bil_dcoef.c:369-739.
//   k_row_diag   fixed-order row sums of F into the diagonal term and tdF.
// Linear triangles use the same kernels: the Gauss rule over a triangle has 3 n^2 points
// (three mapped quadrilaterals), the polar rules 3 sectors, the interpolation functions
// are linear, and the inner disk of the principal-value derivative has closed-form radial
// integrals (lit_* / q0_lit_* of the reference).  A triangle's fourth row and column stay 0.
// Every output element is written by exactly one thread: no atomics on values, so results
// do not depend on scheduling.
#include "bq_runtime.hpp"
#include "bq_qpgf.hpp"
#include "bq_elem.hpp"
#include "bq_reduce.hpp"
#include "../bq_assemble.h"

using namespace bq;

constexpr double kDEPS = 0.1;     // D_epsilon, reference d3qd2_src/d3qd2_const.h:28
constexpr double kCDFH = 1.0e-3;  // finite-difference step, d3qd2_const.h:29
constexpr double kSDF = 0.5;      // GL -> GH switch, d3qd2_const.h:78
constexpr double kTWOPI = 6.283185307179586476925286766559;

struct ColPt {
  double rt[3], vz[3], ve[3];
};

__device__ __forceinline__ void colloc(const double* __restrict__ E, const double* __restrict__ ren, int at, int tn,
                                       ColPt& c)
{
  Elem e;
  load_elem(E, at, e);
  node_tangents(e, tn, c.vz, c.ve);
  const double* p = ren + (size_t)at * 12 + 3 * tn;
  c.rt[0] = p[0];
  c.rt[1] = p[1];
  c.rt[2] = p[2];
}

// free-space solid-angle integrand at one point (F, dF/dtz, dF/dte), unscaled
__device__ __forceinline__ void f_terms(const double R[3], const double W[3], const ColPt& c, double& f, double& f0,
                                        double& f1)
{
  double aR = sqrt(dot3(R, R)), iR = 1.0 / aR;
  double rdW = dot3(R, W) * iR, rdv0 = dot3(R, c.vz) * iR, rdv1 = dot3(R, c.ve) * iR;
  double iR3 = iR * iR * iR;
  f = rdW * iR * iR;
  f0 = iR3 * (3.0 * rdv0 * rdW - dot3(W, c.vz));
  f1 = iR3 * (3.0 * rdv1 * rdW - dot3(W, c.ve));
}

// ------------------------------------------------------------------------------------
// Gauss order of a regular periodic pair (t, s).  The reference integrates every regular pair
// with the 12-point rule (d3qd2_src/q0_bil_dcoef_v2_cc.c:129-204).  Away from the target the
// integrand is analytic in a Bernstein ellipse E_rho of the source's parameter square, and the
// n-point Gauss error is at most 4.27 M(rho) rho^-2n / (rho^2 - 1) (M: the integrand's bound on
// the ellipse).  Here rho is at most half of rho_s, the ellipse through the nearest image of
// the target element's bounding sphere (D = sqrt2 (1 + gap / a), rho_s = D + sqrt(D^2 - 1),
// a the source's bounding radius), M grows like exp(k a rho / (2 sqrt2)) from the wave factor,
// and a further factor 4 covers the growth of 1/R on the ellipse.  The pair takes the
// smallest order from 3 to 10 whose bound is below 1e-15 of the integrand (under the
// rounding of the 12-point sum and the ~1e-14 truncation noise of the lattice sums
// themselves), else 12; with full != 0 every pair takes 12.
__device__ __forceinline__ void elem_sphere(const Elem& e, double (&c)[3], double& rad)
{
  const double zc[4] = {e.tri ? -0.5 : -1.0, 1.0, e.tri ? -0.5 : 1.0, -1.0};
  const double ec[4] = {e.tri ? -0.5 * kSQ3 : -1.0, e.tri ? 0.0 : -1.0, e.tri ? 0.5 * kSQ3 : 1.0, 1.0};
  const int nc = e.tri ? 3 : 4;
  double r[4][3];
  c[0] = c[1] = c[2] = 0.0;
  for (int i = 0; i < nc; ++i)
    for (int d = 0; d < 3; ++d) {
      r[i][d] = e.cr[d][0] + e.cr[d][1] * zc[i] + e.cr[d][2] * ec[i] + e.cr[d][3] * zc[i] * ec[i];
      c[d] += r[i][d] / nc;
    }
  rad = 0.0;
  for (int i = 0; i < nc; ++i) {
    const double dx = r[i][0] - c[0], dy = r[i][1] - c[1], dz = r[i][2] - c[2];
    rad = fmax(rad, sqrt(dx * dx + dy * dy + dz * dz));
  }
}

__device__ __forceinline__ int rule_from_gap(double gap, double rs, double k, double ltol)
{
  const double D = 1.4142135623730950488 * (1.0 + gap / rs), rho_s = D + sqrt(D * D - 1.0);
  const double c = k * rs * 0.35355339059327376220;  // k a / (2 sqrt 2)
  for (int n = kVruleMin; n <= kVruleMax; ++n) {
    const double rho = fmin((2.0 * n + 2.0) / fmax(c, 1e-300), 0.5 * rho_s);
    if (!(rho > 1.05)) continue;
    // log(4 x 4.27) + c rho - 2n log rho - log(rho^2 - 1) <= ltol; by default log(1e-15): below
    // the rounding of the 144-point sum the 12-point rule forms (sqrt(144) x 2^-53 = 1.3e-15)
    if (2.8379 + c * rho - 2.0 * n * log(rho) - log(rho * rho - 1.0) <= ltol) return n;
  }
  return 12;
}

__device__ __forceinline__ int pair_rule(const bq_qplan& P, const Elem& et, const Elem& es, int full)
{
  if (full) return 12;
  double ct[3], cs[3], rt, rs;
  elem_sphere(et, ct, rt);
  elem_sphere(es, cs, rs);
  double dx = ct[0] - cs[0], dy = ct[1] - cs[1];
  const double dz = ct[2] - cs[2];
  const double l1 = rint(dx * P.b1x + dy * P.b1y), l2 = rint(dx * P.b2x + dy * P.b2y);
  dx -= l1 * P.d1x + l2 * P.d2x;
  dy -= l1 * P.d1y + l2 * P.d2y;
  double d2 = 1e300;
  for (int i = -1; i <= 1; ++i)
    for (int j = -1; j <= 1; ++j) {
      const double ex = dx + i * P.d1x + j * P.d2x, ey = dy + i * P.d1y + j * P.d2y;
      d2 = fmin(d2, ex * ex + ey * ey + dz * dz);
    }
  const double gap = sqrt(d2) - rt - rs;
  if (!(gap > 0.0) || !(rs > 0.0)) return 12;
  // Horizontal coplanar pairs (flat source in a plane z = const through the target's nodes,
  // to 1e-6 of the element size): there R_z is 0 or within a few ulps of it at every Gauss
  // point, where the reference's spectral sum stops on a different test than elsewhere (its
  // r_z == 0 branch), so its integrand is not the smooth function the bound assumes.  They
  // keep the reference's 12-point rule and its Gauss points.
  const double zs = fabs(es.cr[2][1]) + fabs(es.cr[2][2]) + fabs(es.cr[2][3]);
  const double zt = fabs(et.cr[2][1]) + fabs(et.cr[2][2]) + fabs(et.cr[2][3]);
  if (zs <= 1e-6 * rs && fabs(et.cr[2][0] - es.cr[2][0]) <= zs + zt + 1e-6 * rs) return 12;
  return rule_from_gap(gap, rs, P.k, P.rule_ltol);
}

// The same for a free-space pair of a bounded domain (complex k: |k| bounds the wave
// factor's growth; no lattice images and no r_z == 0 branch in the kernel).
__device__ __forceinline__ int pair_rule_fs(double kabs, const Elem& et, const Elem& es, int full,
                                            double ltol = BQ_RULE_LTOL)
{
  if (full) return 12;
  double ct[3], cs[3], rt, rs;
  elem_sphere(et, ct, rt);
  elem_sphere(es, cs, rs);
  const double dx = ct[0] - cs[0], dy = ct[1] - cs[1], dz = ct[2] - cs[2];
  const double gap = sqrt(dx * dx + dy * dy + dz * dz) - rt - rs;
  if (!(gap > 0.0) || !(rs > 0.0)) return 12;
  return rule_from_gap(gap, rs, kabs, ltol);
}

// ------------------------------------------------------------------------------------
// F (solid angle / 4 pi) and its tangential derivatives per (row, s) take the distance-chosen
// Gauss order of the pair as the kernels do (no wave factor; the bound tightened by e because
// the derivative terms grow like 1/R^3 towards the target where the bound allows for 1/R);
// pairs the bound leaves at 12 points, which include every pair near enough for |dF/dt| to
// approach the 0.5 switch, keep the reference's 12-point rule.
__global__ void __launch_bounds__(256) k_classify(const double* __restrict__ E, const double* __restrict__ ren,
                                                  const int* __restrict__ sid, int dNe, int freespace,
                                                  int full_rule, double ltol, double* __restrict__ F3,
                                                  unsigned char* __restrict__ flag)
{
  long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
  long N = 4L * dNe;
  if (idx >= N * dNe) return;
  int row = (int)(idx / dNe), sc = (int)(idx % dNe);
  int td = sid[row / 4 + 1], tn = row & 3, sd = sid[sc + 1];
  double* o = F3 + 3 * idx;
  if (tn == 3 && elem_is_tri(E, abs(td))) {
    flag[idx] = 3;
    o[0] = o[1] = o[2] = 0.0;
    return;
  }
  if (td == sd) {
    flag[idx] = 2;
    o[0] = o[1] = o[2] = 0.0;
    return;
  }
  ColPt c;
  colloc(E, ren, abs(td), tn, c);
  Elem e;
  load_elem(E, abs(sd), e);
  int ng;
  {
    Elem et;
    load_elem(E, abs(td), et);
    ng = pair_rule_fs(0.0, et, e, full_rule, ltol - 1.0);
  }
  double F = 0, F0 = 0, F1 = 0;
  const int np = (e.tri ? 3 : 1) * ng * ng;
  for (int p = 0; p < np; ++p) {
    double z, et, w, R[3], W[3], f, f0, f1;
    gauss_point_n(e.tri, ng, p, z, et, w);
    rw_at(e, z, et, R, W);
    R[0] -= c.rt[0]; R[1] -= c.rt[1]; R[2] -= c.rt[2];
    f_terms(R, W, c, f, f0, f1);
    F += w * f;
    F0 += w * f0;
    F1 += w * f1;
  }
  double sg = (freespace && sd < 0) ? -kI4PI : kI4PI;
  o[0] = F * sg;
  o[1] = F0 * sg;
  o[2] = F1 * sg;
  flag[idx] = (fabs(F0 * kI4PI) > kSDF || fabs(F1 * kI4PI) > kSDF) ? 1 : 0;
}


// ------------------------------------------------------------------------------------
// domain 0 (periodic), regular pairs.  One wavefront per pair (t, s): the target's nodes
// share the 64 lanes (4 x 16 for a quad, 3 x 21 for a triangle, whose fourth row does not
// exist), each lane walks the source's Gauss points with stride 16 or 21.  Every lane
// evaluates the QPGF in every sweep, with lattice sums that are uniform over the wavefront
// (qpgf_v1_wave): the lanes' points lie within one element of each other, so their
// truncation disks nearly coincide, and lanes with the same loop bounds and order data stay
// converged, where lanes with their own bounds and method diverge at every row end.  A lane
// past the rule's last point or without a live node evaluates the first live lane's node at
// the rule's first point and discards the result.
// The per-lane sums live in LDS rather than registers, and the kernel is held to 128
// registers so that four wavefronts share each SIMD (the lattice sums' long dependent FP64
// chains need the other wavefronts to hide their latency more than they need registers:
// the spills this costs go to scratch that stays in cache; at C4 k_q0_gl took 2.44 s per
// point at two wavefronts, 2.06 s at three and 2.01 s at four; at five the LDS limits the
// CU to four workgroups and the 102 registers spill heavily, 4.07 s).
#ifndef BQ_Q0GL_WAVES
#define BQ_Q0GL_WAVES 4
#endif
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(BQ_Q0GL_WAVES, BQ_Q0GL_WAVES)))
k_q0_gl(bq_qplan P, const double* __restrict__ E, const double* __restrict__ ren, const int* __restrict__ sid, int dNe,
        const unsigned char* __restrict__ flag, int full_rule, double2* __restrict__ G, double2* __restrict__ H)
{
  __shared__ double sacc[16][256];
  __shared__ double selem[4][kElemStride];
  __shared__ double setab[kErfcxTab];
  erfcx_stage(setab);
  __syncthreads();
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  // one wavefront per (t, s): make the pair index wave-uniform so that the element data of
  // s are fetched with scalar loads and held in scalar registers
  const long pair = (long)blockIdx.x * 4 + __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
  if (pair >= (long)dNe * dNe) return;
  const int t = (int)(pair / dNe) + 1, s = (int)(pair % dNe) + 1;
  if (t == s) return;
  const bool ttri = elem_is_tri(E, abs(sid[t])), stri = elem_is_tri(E, abs(sid[s]));
  const int gw = ttri ? 21 : 16, nodes = ttri ? 3 : 4;
  const int tn = lane / gw, sub = lane - tn * gw;
  const long N = 4L * dNe, row = 4L * (t - 1) + (tn < nodes ? tn : 0);
  const bool active = tn < nodes && flag[row * dNe + (s - 1)] == 0;
  double* acc = &sacc[0][threadIdx.x];  // acc[k * 256]
#pragma unroll
  for (int k = 0; k < 16; ++k) acc[k * 256] = 0.0;
  // element s staged in LDS and read back per point, so it is not held across the sums
  if (lane < kElemStride) selem[wv][lane] = E[(size_t)abs(sid[s]) * kElemStride + lane];
  __builtin_amdgcn_s_waitcnt(0);
  __builtin_amdgcn_wave_barrier();
  int ng;
  {
    Elem et, es;
    load_elem(E, abs(sid[t]), et);
    load_elem(&selem[0][0], wv, es);
    ng = __builtin_amdgcn_readfirstlane(pair_rule(P, et, es, full_rule));
  }
  const unsigned long long live = __ballot(active);
  if (live != 0) {
    const int tsrc = __shfl(tn, __ffsll((long long)live) - 1, 64);
    double rt[3];
    {
      const double* p = ren + (size_t)abs(sid[t]) * 12 + 3 * (active ? tn : tsrc);
      rt[0] = p[0];
      rt[1] = p[1];
      rt[2] = p[2];
    }
    const int np = (stri ? 3 : 1) * ng * ng;
    const int nm = (np + gw - 1) / gw;
#pragma unroll 1
    for (int m = 0; m < nm; ++m) {
      const int p0 = sub + gw * m;
      const bool use = active && p0 < np;
      double z, et, w;
      gauss_point_n(stri, ng, use ? p0 : 0, z, et, w);
      double R[3], W[3], Nn[4];
      {
        Elem e;
        load_elem(&selem[0][0], wv, e);
        rw_at(e, z, et, R, W);
      }
      R[0] -= rt[0]; R[1] -= rt[1]; R[2] -= rt[2];
      QG q;
      qpgf_v1_wave(P, R[0], R[1], R[2], q, setab);
      if (use) {
        double aW = sqrt(dot3(W, W));
        cplx hn = q.d[0] * W[0] + q.d[1] * W[1] + q.d[2] * W[2];
        shapeN(stri, z, et, Nn);
#pragma unroll
        for (int n = 0; n < 4; ++n) {
          double gw2 = w * Nn[n] * aW, hw = w * Nn[n];
          acc[(2 * n) * 256] += gw2 * q.g.re;
          acc[(2 * n + 1) * 256] += gw2 * q.g.im;
          acc[(8 + 2 * n) * 256] += hw * hn.re;
          acc[(8 + 2 * n + 1) * 256] += hw * hn.im;
        }
      }
    }
  }
  __builtin_amdgcn_wave_barrier();
  // per node, lane k < 16 of the node's first 16 lanes sums value k over the node's lanes
  // in lane order (fixed order: the result does not depend on scheduling)
  if (tn < nodes && sub < 16) {
    const bool act_n = flag[row * dNe + (s - 1)] == 0;
    if (act_n) {
      const int base = wv * 64 + tn * gw;
      double v = 0.0;
      for (int j = 0; j < gw; ++j) v += sacc[sub][base + j];
      double* dst = (double*)((sub < 8 ? G : H) + row * N + 4L * (s - 1));
      dst[sub & 7] = v;
    }
  }
}

// domain 0, flagged pairs: 64x64 rule, F and dF/dt recomputed with it (wavefront-uniform
// lattice sums, accumulators in LDS, three wavefronts per SIMD)
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(3, 3)))
k_q0_gh(bq_qplan P, const double* __restrict__ E, const double* __restrict__ ren, const int* __restrict__ sid, int dNe,
        const int* __restrict__ list, double2* __restrict__ G, double2* __restrict__ H, double* __restrict__ F3)
{
  __shared__ double lds[19 * 4];
  __shared__ double sacc[16][256];
  __shared__ double setab[kErfcxTab];
  erfcx_stage(setab);
  __syncthreads();
  const long code = list[blockIdx.x];
  const long N = 4L * dNe, row = code / dNe;
  const int sc = (int)(code % dNe), t = (int)(row / 4) + 1, tn = (int)(row & 3);
  ColPt c;
  colloc(E, ren, abs(sid[t]), tn, c);
  Elem e;
  load_elem(E, abs(sid[sc + 1]), e);
  double* acc = &sacc[0][threadIdx.x];
#pragma unroll
  for (int k = 0; k < 16; ++k) acc[k * 256] = 0.0;
  double fs[3] = {0.0, 0.0, 0.0};
  const int nm = e.tri ? 48 : 16;  // 4096 or 12288 points
#pragma unroll 1
  for (int m = 0; m < nm; ++m) {
    double z, et, w;
    gauss_point(e.tri, true, threadIdx.x + 256 * m, z, et, w);
    double R[3], W[3], Nn[4], f, f0, f1;
    rw_at(e, z, et, R, W);
    R[0] -= c.rt[0]; R[1] -= c.rt[1]; R[2] -= c.rt[2];
    f_terms(R, W, c, f, f0, f1);
    fs[0] += w * f;
    fs[1] += w * f0;
    fs[2] += w * f1;
    QG q;
    qpgf_v1_wave(P, R[0], R[1], R[2], q, setab);
    double aW = sqrt(dot3(W, W));
    cplx hn = q.d[0] * W[0] + q.d[1] * W[1] + q.d[2] * W[2];
    shapeN(e.tri, z, et, Nn);
#pragma unroll
    for (int n = 0; n < 4; ++n) {
      double gw = w * Nn[n] * aW, hw = w * Nn[n];
      acc[(2 * n) * 256] += gw * q.g.re;
      acc[(2 * n + 1) * 256] += gw * q.g.im;
      acc[(8 + 2 * n) * 256] += hw * hn.re;
      acc[(8 + 2 * n + 1) * 256] += hw * hn.im;
    }
  }
  double a[19];
#pragma unroll
  for (int k = 0; k < 16; ++k) a[k] = acc[k * 256];
  a[16] = fs[0];
  a[17] = fs[1];
  a[18] = fs[2];
  block_sum256<19>(a, lds);
  if (threadIdx.x == 0) {
    double2* g = G + row * N + 4L * sc;
    double2* h = H + row * N + 4L * sc;
#pragma unroll
    for (int n = 0; n < 4; ++n) {
      g[n] = make_double2(a[2 * n], a[2 * n + 1]);
      h[n] = make_double2(a[8 + 2 * n], a[8 + 2 * n + 1]);
    }
    double* o = F3 + 3 * code;
    o[0] = a[16] * kI4PI;
    o[1] = a[17] * kI4PI;
    o[2] = a[18] * kI4PI;
  }
}

// polar F of a bilinear element e at its own parameter point (zt, et), unscaled by bdC/4pi
__device__ double polar_F_part(const Elem& e, double zt, double et)
{
  double r[5], th[5];
  corner_r_th4(zt, et, r, th);
  double F = 0.0;
  for (int q = threadIdx.x; q < 512; q += 256) {
    int j = q >> 7, m = q & 127;
    double theta, wt;
    polar_theta(th, j, m, theta, wt);
    double rm = polar_rm(r, th, j, theta), cth, sth;
    sincos(theta, &sth, &cth);
    double acc = 0.0;
    for (int i = 0; i < 12; ++i) {
      double rr = 0.5 * rm * (c_xl[i] + 1.0), K[3];
      polar_K(e, zt, et, rr, cth, sth, K);
      double aK = sqrt(dot3(K, K));
      acc += c_wl[i] * cth * sth / (aK * aK * aK);
    }
    F += wt * 0.5 * rm * acc;
  }
  return F;
}

// domain 0, singular pairs: one block per collocation node (lattice-sum accumulators in
// LDS, three wavefronts per SIMD: 108 ms per C4 point against 125 ms at two and 111 ms at
// four; wavefront-uniform lattice sums as in k_q0_gl; 128 angles per sector keep every
// wavefront wholly active or wholly idle in the angle loop)
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(3, 3)))
k_q0_sing(bq_qplan P, const double* __restrict__ E, const int* __restrict__ sid, int dNe, double2* __restrict__ G,
          double2* __restrict__ H, double* __restrict__ F3)
{
  __shared__ double lds[21 * 4];
  __shared__ double sacc[16][256];
  __shared__ double setab[kErfcxTab];
  erfcx_stage(setab);
  __syncthreads();
  const long N = 4L * dNe, row = blockIdx.x;
  const int t = (int)(row / 4) + 1, tn = (int)(row & 3), as = abs(sid[t]);
  Elem e;
  load_elem(E, as, e);
  if (e.tri && tn == 3) return;  // no fourth node (whole block)
  const double zt = elem_node_z(e.tri, tn), et = elem_node_e(e.tri, tn);
  double r[5], th[5];
  const int nsec = corner_r_th(e.tri, zt, et, r, th);
  double* sa = &sacc[0][threadIdx.x];
#pragma unroll
  for (int k = 0; k < 16; ++k) sa[k * 256] = 0.0;
  for (int q = threadIdx.x; q < 128 * nsec; q += 256) {
    int j = q >> 7, m = q & 127;
    double theta, wt;
    polar_theta(th, j, m, theta, wt);
    double rm = polar_rm(r, th, j, theta), cth, sth;
    sincos(theta, &sth, &cth);
    for (int i = 0; i < 12; ++i) {
      double rr = 0.5 * rm * (c_xl[i] + 1.0), w = wt * 0.5 * rm * c_wl[i];
      double K[3], R[3], W[3], Nn[4];
      polar_K(e, zt, et, rr, cth, sth, K);
      double z = rr * cth + zt, ee = rr * sth + et;
#pragma unroll
      for (int d = 0; d < 3; ++d) {
        R[d] = rr * K[d];
        W[d] = e.cw[d][0] + e.cw[d][1] * z + e.cw[d][2] * ee;
      }
      QG qq;
      qpgf_v1_wave(P, R[0], R[1], R[2], qq, setab);
      // the reference takes G here from d3hm_qpgf_d2_qG and the gradient from _dqG
      // (d3qd2_src/q0_bil_coef.c:366, :399, :444, :489); they differ only at R_z == 0
      if (__any(R[2] == 0.0)) {
        QG qa;
        qpgf_v1_wave(P, R[0], R[1], R[2], qa, setab, false);
        if (R[2] == 0.0) qq.g = qa.g;
      }
      double aW = sqrt(dot3(W, W));
      cplx hn = qq.d[0] * W[0] + qq.d[1] * W[1] + qq.d[2] * W[2];
      shapeN(e.tri, z, ee, Nn);
#pragma unroll
      for (int n = 0; n < 4; ++n) {
        double gw = w * Nn[n] * aW * rr, hw = w * Nn[n] * rr;
        sa[(2 * n) * 256] += gw * qq.g.re;
        sa[(2 * n + 1) * 256] += gw * qq.g.im;
        sa[(8 + 2 * n) * 256] += hw * hn.re;
        sa[(8 + 2 * n + 1) * 256] += hw * hn.im;
      }
    }
  }
  double acc[21];
#pragma unroll
  for (int k = 0; k < 16; ++k) acc[k] = sa[k * 256];
  // F at the node and at the four shifted points (central differences, CDFH); a flat
  // triangle has F = dF/dt = 0 at its own nodes (q0_lit_dcoef_v2_cc.c:331-342)
  if (e.tri) {
    acc[16] = acc[17] = acc[18] = acc[19] = acc[20] = 0.0;
  } else {
    acc[16] = polar_F_part(e, zt, et);
    acc[17] = polar_F_part(e, zt + kCDFH, et);
    acc[18] = polar_F_part(e, zt - kCDFH, et);
    acc[19] = polar_F_part(e, zt, et + kCDFH);
    acc[20] = polar_F_part(e, zt, et - kCDFH);
  }
  block_sum256<21>(acc, lds);
  if (threadIdx.x == 0) {
    double2* g = G + row * N + 4L * (t - 1);
    double2* h = H + row * N + 4L * (t - 1);
#pragma unroll
    for (int n = 0; n < 4; ++n) {
      g[n] = make_double2(acc[2 * n], acc[2 * n + 1]);
      h[n] = make_double2(acc[8 + 2 * n], acc[8 + 2 * n + 1]);
    }
    double sc = elem_plane(e) ? 0.0 : elem_bdC(e) * kI4PI;
    double Tz[3], Te[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      Tz[d] = e.cr[d][1] + e.cr[d][3] * et;
      Te[d] = e.cr[d][2] + e.cr[d][3] * zt;
    }
    double* o = F3 + 3 * (row * dNe + (t - 1));
    o[0] = acc[16] * sc;
    o[1] = (acc[17] - acc[18]) * sc / (2.0 * kCDFH) / sqrt(dot3(Tz, Tz));
    o[2] = (acc[19] - acc[20]) * sc / (2.0 * kCDFH) / sqrt(dot3(Te, Te));
  }
}

// ------------------------------------------------------------------------------------
// free-space kernel terms at one regular point; adds w-weighted contributions of the
// 6 quantities x 4 shape functions to acc (48 doubles):
//   [0..7] G, [8..15] H, [16..23] dG/dtz, [24..31] dG/dte, [32..39] dH/dtz, [40..47] dH/dte
__device__ __forceinline__ void fs_point(const double R[3], const double W[3], const double Nn[4], const ColPt& c,
                                         cplx k, double w, double (&acc)[48])
{
  double aR = sqrt(dot3(R, R)), iR = 1.0 / aR, aW = sqrt(dot3(W, W));
  cplx ca = cmul_i(k * aR);                // i k R
  cplx ce = cexp(ca);
  cplx cb = ca - mk(1.0, 0.0);
  double rdW = dot3(R, W) * iR, tD = rdW * iR * iR;
  double rdv0 = dot3(R, c.vz) * iR, rdv1 = dot3(R, c.ve) * iR, Wdv0 = dot3(W, c.vz), Wdv1 = dot3(W, c.ve);
  cplx g = ce * (iR * aW);
  cplx h = cb * ce * tD;
  cplx dg = cb * ce * (iR * iR * aW);
  cplx k2R2 = k * k * (aR * aR);
  cplx t3 = k2R2 + 3.0 * cb;
  cplx ce3 = ce * (iR * iR * iR);
  cplx dh0 = ce3 * (t3 * (rdv0 * rdW) - cb * Wdv0);
  cplx dh1 = ce3 * (t3 * (rdv1 * rdW) - cb * Wdv1);
#pragma unroll
  for (int n = 0; n < 4; ++n) {
    double wn = w * Nn[n];
    acc[2 * n] += wn * g.re;            acc[2 * n + 1] += wn * g.im;
    acc[8 + 2 * n] += wn * h.re;        acc[8 + 2 * n + 1] += wn * h.im;
    acc[16 + 2 * n] += wn * rdv0 * dg.re; acc[16 + 2 * n + 1] += wn * rdv0 * dg.im;
    acc[24 + 2 * n] += wn * rdv1 * dg.re; acc[24 + 2 * n + 1] += wn * rdv1 * dg.im;
    acc[32 + 2 * n] += wn * dh0.re;     acc[32 + 2 * n + 1] += wn * dh0.im;
    acc[40 + 2 * n] += wn * dh1.re;     acc[40 + 2 * n + 1] += wn * dh1.im;
  }
}

// scale by 1/4pi (dG by -1/4pi), flip the sign of H, dH for elements seen from their
// sub-domain side, and store G, H rows and the dG/dH derivative rows
__device__ __forceinline__ void fs_store(const double (&acc)[48], bool neg, long row, long N, long col,
                                         double2* G, double2* H, double2* dG, double2* dH)
{
  double sh = neg ? -kI4PI : kI4PI;
#pragma unroll
  for (int n = 0; n < 4; ++n) {
    G[row * N + col + n] = make_double2(acc[2 * n] * kI4PI, acc[2 * n + 1] * kI4PI);
    H[row * N + col + n] = make_double2(acc[8 + 2 * n] * sh, acc[8 + 2 * n + 1] * sh);
    dG[(2 * row) * N + col + n] = make_double2(-acc[16 + 2 * n] * kI4PI, -acc[16 + 2 * n + 1] * kI4PI);
    dG[(2 * row + 1) * N + col + n] = make_double2(-acc[24 + 2 * n] * kI4PI, -acc[24 + 2 * n + 1] * kI4PI);
    dH[(2 * row) * N + col + n] = make_double2(acc[32 + 2 * n] * sh, acc[32 + 2 * n + 1] * sh);
    dH[(2 * row + 1) * N + col + n] = make_double2(acc[40 + 2 * n] * sh, acc[40 + 2 * n + 1] * sh);
  }
}

// fs_store for one lane's 12 of the 48 sums (q-th twelve, q = 0..3): sum 12 q + 2 i + (0, 1) is
// quantity (12 q + 2 i) / 8 (G, H, dG/dtz, dG/dte, dH/dtz, dH/dte) of shape function
// ((12 q + 2 i) % 8) / 2
__device__ __forceinline__ void fs_store12(const double (&v)[12], int q, bool neg, long row, long N, long col,
                                           double2* G, double2* H, double2* dG, double2* dH)
{
  const double sh = neg ? -kI4PI : kI4PI;
#pragma unroll
  for (int i = 0; i < 6; ++i) {
    const int idx = 12 * q + 2 * i, qty = idx >> 3, n = (idx & 7) >> 1;
    double2* dst;
    double sc;
    if (qty == 0) {
      dst = G + row * N;
      sc = kI4PI;
    } else if (qty == 1) {
      dst = H + row * N;
      sc = sh;
    } else if (qty < 4) {
      dst = dG + (2 * row + (qty - 2)) * N;
      sc = -kI4PI;
    } else {
      dst = dH + (2 * row + (qty - 4)) * N;
      sc = sh;
    }
    dst[col + n] = make_double2(v[2 * i] * sc, v[2 * i + 1] * sc);
  }
}

// three wavefronts per SIMD (168 registers, no spills): 167-176 ms per C4 point against
// 197 ms at two (171 registers) and 217-226 ms at four
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(3, 3)))
k_fs_gl(double kr, double ki, const double* __restrict__ E, const double* __restrict__ ren,
        const int* __restrict__ sid, int dNe, const unsigned char* __restrict__ flag, int full_rule, double ltol,
        double2* __restrict__ G, double2* __restrict__ H, double2* __restrict__ dG, double2* __restrict__ dH)
{
  const int lane = threadIdx.x & 63;
  // one wavefront per (t, s): make the pair index wave-uniform so that the element data of
  // s are fetched with scalar loads and held in scalar registers
  const long pair = (long)blockIdx.x * 4 + __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
  if (pair >= (long)dNe * dNe) return;
  const int t = (int)(pair / dNe) + 1, s = (int)(pair % dNe) + 1;
  if (t == s) return;
  const int tn = lane >> 4, sub = lane & 15;
  const long N = 4L * dNe, row = 4L * (t - 1) + tn;
  const bool active = flag[row * dNe + (s - 1)] == 0;
  const int sd = sid[s];
  ColPt c;
  colloc(E, ren, abs(sid[t]), tn, c);
  Elem e;
  load_elem(E, abs(sd), e);
  const cplx k = mk(kr, ki);
  double acc[48];
#pragma unroll
  for (int q = 0; q < 48; ++q) acc[q] = 0.0;
  int ng;
  {
    Elem etg;
    load_elem(E, abs(sid[t]), etg);
    ng = __builtin_amdgcn_readfirstlane(pair_rule_fs(sqrt(kr * kr + ki * ki), etg, e, full_rule, ltol));
  }
  if (active) {
    const int np = (e.tri ? 3 : 1) * ng * ng;
    const int nm = (np + 15) / 16;
    for (int m = 0; m < nm; ++m) {
      const int p = sub + 16 * m;
      if (p >= np) break;
      double z, et, w;
      gauss_point_n(e.tri, ng, p, z, et, w);
      double R[3], W[3], Nn[4];
      rw_at(e, z, et, R, W);
      R[0] -= c.rt[0]; R[1] -= c.rt[1]; R[2] -= c.rt[2];
      shapeN(e.tri, z, et, Nn);
      fs_point(R, W, Nn, c, k, w, acc);
    }
  }
  // reduce-scatter over the 16 lanes of the row: halve the 48 sums at each of the first two
  // levels (each lane keeps the half its lane bit selects and adds the partner's copy of it),
  // then sum the remaining 12 over the last two; lane sub = 4q holds the 12 totals from
  // 12 q (q = 0..3)
  double h24[24], h12[12];
  const bool up8 = (sub & 8) != 0, up4 = (sub & 4) != 0;
#pragma unroll
  for (int i = 0; i < 24; ++i) {
    const double keep = up8 ? acc[24 + i] : acc[i], send = up8 ? acc[i] : acc[24 + i];
    h24[i] = keep + __shfl_xor(send, 8, 16);
  }
#pragma unroll
  for (int i = 0; i < 12; ++i) {
    const double keep = up4 ? h24[12 + i] : h24[i], send = up4 ? h24[i] : h24[12 + i];
    h12[i] = keep + __shfl_xor(send, 4, 16);
  }
#pragma unroll
  for (int i = 0; i < 12; ++i) h12[i] += __shfl_xor(h12[i], 2, 16);
#pragma unroll
  for (int i = 0; i < 12; ++i) h12[i] += __shfl_xor(h12[i], 1, 16);
  if (active && (sub & 3) == 0) fs_store12(h12, sub >> 2, sd < 0, row, N, 4L * (s - 1), G, H, dG, dH);
}

__global__ void __launch_bounds__(256) k_fs_gh(double kr, double ki, const double* __restrict__ E,
                                               const double* __restrict__ ren, const int* __restrict__ sid, int dNe,
                                               const int* __restrict__ list, double2* __restrict__ G,
                                               double2* __restrict__ H, double2* __restrict__ dG,
                                               double2* __restrict__ dH, double* __restrict__ F3)
{
  __shared__ double lds[51 * 4];
  const long code = list[blockIdx.x];
  const long N = 4L * dNe, row = code / dNe;
  const int sc = (int)(code % dNe), t = (int)(row / 4) + 1, tn = (int)(row & 3), sd = sid[sc + 1];
  ColPt c;
  colloc(E, ren, abs(sid[t]), tn, c);
  Elem e;
  load_elem(E, abs(sd), e);
  const cplx k = mk(kr, ki);
  double acc[48];
#pragma unroll
  for (int q = 0; q < 48; ++q) acc[q] = 0.0;
  double fa[3] = {0.0, 0.0, 0.0};
  const int nm = e.tri ? 48 : 16;
  for (int m = 0; m < nm; ++m) {
    double z, et, w;
    gauss_point(e.tri, true, threadIdx.x + 256 * m, z, et, w);
    double R[3], W[3], Nn[4], f, f0, f1;
    rw_at(e, z, et, R, W);
    R[0] -= c.rt[0]; R[1] -= c.rt[1]; R[2] -= c.rt[2];
    f_terms(R, W, c, f, f0, f1);
    fa[0] += w * f;
    fa[1] += w * f0;
    fa[2] += w * f1;
    shapeN(e.tri, z, et, Nn);
    fs_point(R, W, Nn, c, k, w, acc);
  }
  block_sum256<48>(acc, lds);
  block_sum256<3>(fa, lds);
  if (threadIdx.x == 0) {
    fs_store(acc, sd < 0, row, N, 4L * sc, G, H, dG, dH);
    double sg = sd < 0 ? -kI4PI : kI4PI;
    double* o = F3 + 3 * code;
    o[0] = fa[0] * sg;
    o[1] = fa[1] * sg;
    o[2] = fa[2] * sg;
  }
}

// free-space singular pair: G, H, F by the polar rule (d3qd2_src/bil_coef.c:244-342) and
// the principal-value tangential derivatives (d3qd2_src/bil_dcoef.c:369-739)
__global__ void __launch_bounds__(256) k_fs_sing(double kr, double ki, const double* __restrict__ E,
                                                 const int* __restrict__ sid, int dNe, double2* __restrict__ G,
                                                 double2* __restrict__ H, double2* __restrict__ dG,
                                                 double2* __restrict__ dH, double* __restrict__ F3)
{
  __shared__ double lds[17 * 4];
  const long N = 4L * dNe, row = blockIdx.x;
  const int t = (int)(row / 4) + 1, tn = (int)(row & 3), sd = sid[t];
  Elem e;
  load_elem(E, abs(sd), e);
  if (e.tri && tn == 3) return;  // no fourth node (whole block)
  const cplx k = mk(kr, ki);
  const double zt = elem_node_z(e.tri, tn), et = elem_node_e(e.tri, tn);
  double r[5], th[5];
  const int nsec = corner_r_th(e.tri, zt, et, r, th);
  const bool plane = elem_plane(e);
  const double bdC = elem_bdC(e);
  const double sgn = (sd < 0) ? -1.0 : 1.0;
  const long col = 4L * (t - 1);

  // (a) G, H, F
  {
    double a[17];
#pragma unroll
    for (int q = 0; q < 17; ++q) a[q] = 0.0;
    for (int q = threadIdx.x; q < 128 * nsec; q += 256) {
      int j = q >> 7, m = q & 127;
      double theta, wt;
      polar_theta(th, j, m, theta, wt);
      double rm = polar_rm(r, th, j, theta), cth, sth;
      sincos(theta, &sth, &cth);
      for (int i = 0; i < 12; ++i) {
        double rr = 0.5 * rm * (c_xl[i] + 1.0), w = wt * 0.5 * rm * c_wl[i];
        double K[3], W[3], Nn[4];
        polar_K(e, zt, et, rr, cth, sth, K);
        double z = rr * cth + zt, ee = rr * sth + et;
#pragma unroll
        for (int d = 0; d < 3; ++d) W[d] = e.cw[d][0] + e.cw[d][1] * z + e.cw[d][2] * ee;
        double aK = sqrt(dot3(K, K)), aW = sqrt(dot3(W, W));
        cplx kR = k * (rr * aK);
        cplx ce = cexp(cmul_i(kR));
        cplx g = ce * (aW / aK);
        cplx h = (cmul_i(kR) - mk(1.0, 0.0)) * ce * (cth * sth / (aK * aK * aK));
        shapeN(e.tri, z, ee, Nn);
#pragma unroll
        for (int n = 0; n < 4; ++n) {
          a[2 * n] += w * Nn[n] * g.re;
          a[2 * n + 1] += w * Nn[n] * g.im;
          a[8 + 2 * n] += w * Nn[n] * h.re;
          a[8 + 2 * n + 1] += w * Nn[n] * h.im;
        }
        a[16] += w * cth * sth / (aK * aK * aK);
      }
    }
    block_sum256<17>(a, lds);
    if (threadIdx.x == 0) {
      double sh = plane ? 0.0 : bdC * kI4PI * sgn;
#pragma unroll
      for (int n = 0; n < 4; ++n) {
        G[row * N + col + n] = make_double2(a[2 * n] * kI4PI, a[2 * n + 1] * kI4PI);
        H[row * N + col + n] = make_double2(a[8 + 2 * n] * sh, a[8 + 2 * n + 1] * sh);
      }
      F3[3 * (row * dNe + (t - 1))] = a[16] * sh;
    }
  }

  // (b) tangential derivatives, direction 0 = zeta, 1 = eta
  for (int dir = 0; dir < 2; ++dir) {
    double T[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) T[d] = (dir == 0) ? e.cr[d][1] + e.cr[d][3] * et : e.cr[d][2] + e.cr[d][3] * zt;
    const double iaT = 1.0 / sqrt(dot3(T, T));
    double a[17];  // dG (8), dH (8), dF (1)
#pragma unroll
    for (int q = 0; q < 17; ++q) a[q] = 0.0;
    // D-bar region: r in [DEPS, rm], 64 radii
    for (int q = threadIdx.x; q < 128 * nsec; q += 256) {
      int j = q >> 7, m = q & 127;
      double theta, wt;
      polar_theta(th, j, m, theta, wt);
      double rm = polar_rm(r, th, j, theta), cth, sth;
      sincos(theta, &sth, &cth);
      const double csd = dir == 0 ? cth : sth, outer = dir == 0 ? sth : cth;
      for (int i = 0; i < 64; ++i) {
        double rr = 0.5 * (rm - kDEPS) * c_xh[i] + 0.5 * (rm + kDEPS), w = wt * 0.5 * (rm - kDEPS) * c_wh[i];
        double K[3], W[3], Nn[4];
        polar_K(e, zt, et, rr, cth, sth, K);
        double z = rr * cth + zt, ee = rr * sth + et;
#pragma unroll
        for (int d = 0; d < 3; ++d) W[d] = e.cw[d][0] + e.cw[d][1] * z + e.cw[d][2] * ee;
        double aK = sqrt(dot3(K, K)), iK = 1.0 / aK, aW = sqrt(dot3(W, W)), KT = dot3(K, T);
        double iK5 = iK * iK * iK * iK * iK;
        cplx kR = k * (rr * aK);
        cplx ce = cexp(cmul_i(kR));
        cplx cb = cmul_i(kR) - mk(1.0, 0.0);
        cplx g = cb * ce * (iK * iK / rr * (KT * iK) * aW);
        cplx h = ce * ((kR * kR + 3.0 * cb) * (KT * csd) - cb * (aK * aK)) * (outer * iK5 / rr);
        shapeN(e.tri, z, ee, Nn);
#pragma unroll
        for (int n = 0; n < 4; ++n) {
          a[2 * n] += w * Nn[n] * g.re;
          a[2 * n + 1] += w * Nn[n] * g.im;
          a[8 + 2 * n] += w * Nn[n] * h.re;
          a[8 + 2 * n + 1] += w * Nn[n] * h.im;
        }
        a[16] += w * outer * (3.0 * KT * csd - aK * aK) * iK5 / rr;
      }
    }
    if (e.tri) {
      // Linear triangle: N(zt + r c, et + r s) = N(zt, et) + r tri_dN(c, s).  The constant
      // part cancels in the principal value and the radial integrals of the linear part
      // over [0, DEPS] are closed forms, leaving theta rules (lit_sdG_De_theta and
      // lit_sdG_DP_theta, d3qd2_src/lit_dcoef.c:906-952, combined as i/k De - 1/k DP).
      const double aWj = sqrt(e.cw[0][0] * e.cw[0][0] + e.cw[1][0] * e.cw[1][0] + e.cw[2][0] * e.cw[2][0]);
      const cplx ik = cdiv(mk(aWj, 0.0), k);  // |W| / k
      for (int q = threadIdx.x; q < 384; q += 256) {
        const bool disk = q < 256;
        const int qq = disk ? q : q - 256, part = qq >> 6, mm = qq & 63;
        double theta = 0.25 * bq::kPI * c_xh[mm] +
                       (disk ? (-0.75 * bq::kPI + 0.5 * bq::kPI * part) : (part == 0 ? 0.25 * bq::kPI : 0.75 * bq::kPI));
        double wt = 0.25 * bq::kPI * c_wh[mm], cth, sth;
        sincos(theta, &sth, &cth);
        double K[3];
#pragma unroll
        for (int d = 0; d < 3; ++d) K[d] = e.cr[d][1] * cth + e.cr[d][2] * sth;
        double aK = sqrt(dot3(K, K)), iK = 1.0 / aK, f = iK * iK * iK * (dot3(K, T) * iK) * wt;
        cplx keK = k * (kDEPS * aK), v;
        if (disk) {
          cplx cs = csin(keK), cc = ccos(keK);
          v = cmul_i(((keK + mk(0.0, 1.0)) * cs + (mk(2.0, 0.0) - cmul_i(keK)) * cc - mk(2.0, 0.0)) * ik) * f;
        } else {
          v = csin(keK) * ik * (-f);
        }
#pragma unroll
        for (int n = 0; n < 3; ++n) {
          double dn = tri_dN(n, cth, sth);
          a[2 * n] += dn * v.re;
          a[2 * n + 1] += dn * v.im;
        }
      }
    } else {
    // D-eps disk: theta in [-pi, pi] (4 quarters x 64), r in [0, DEPS] (12 radii)
    for (int q = threadIdx.x; q < 256; q += 256) {
      int quarter = q >> 6, mm = q & 63;
      double theta = 0.25 * bq::kPI * c_xh[mm] + (-0.75 * bq::kPI + 0.5 * bq::kPI * quarter);
      double wt = 0.25 * bq::kPI * c_wh[mm], cth, sth;
      sincos(theta, &sth, &cth);
      const double csd = dir == 0 ? cth : sth, outer = dir == 0 ? sth : cth;
      cplx gs[4], h0s[4], h1s[4];
#pragma unroll
      for (int n = 0; n < 4; ++n) gs[n] = h0s[n] = h1s[n] = mk(0, 0);
      for (int i = 0; i < 12; ++i) {
        double rr = 0.5 * kDEPS * (c_xl[i] + 1.0), wr = 0.5 * kDEPS * c_wl[i];
        double K[3], W[3], Nn[4];
        polar_K(e, zt, et, rr, cth, sth, K);
        double z = rr * cth + zt, ee = rr * sth + et;
#pragma unroll
        for (int d = 0; d < 3; ++d) W[d] = e.cw[d][0] + e.cw[d][1] * z + e.cw[d][2] * ee;
        double aK = sqrt(dot3(K, K)), iK = 1.0 / aK, aW = sqrt(dot3(W, W)), KT = dot3(K, T);
        cplx kR = k * (rr * aK);
        cplx cb = cmul_i(kR) - mk(1.0, 0.0);
        cplx cc = ccos(kR), sn = csin(kR), ce = cexp(cmul_i(kR));
        cplx g = (cb * sn + kR * cc) * (iK * iK / rr * (KT * iK) * aW);
        cplx ik = cmul_i(k);
        cplx h0 = ce * ((k * kR + 3.0 * ik) * (KT * csd) - ik * (aK * aK)) * (iK * iK * iK * iK);
        cplx h1 = sn * ((3.0 * KT * csd - aK * aK) * iK * iK * iK * iK * iK / rr);
        shapeN4(z, ee, Nn);
#pragma unroll
        for (int n = 0; n < 4; ++n) {
          gs[n] += g * (wr * Nn[n]);
          h0s[n] += h0 * (wr * Nn[n]);
          h1s[n] += h1 * (wr * Nn[n]);
        }
      }
#pragma unroll
      for (int n = 0; n < 4; ++n) {
        cplx ig = cmul_i(gs[n]) * wt;                            // + i * De
        cplx hh = (h0s[n] - cmul_i(h1s[n])) * (wt * outer);      // + (De0 - i De1)
        a[2 * n] += ig.re;
        a[2 * n + 1] += ig.im;
        a[8 + 2 * n] += hh.re;
        a[8 + 2 * n + 1] += hh.im;
      }
    }
    // D-eps principal-value pairing of +r and -r: theta in [0, pi] (2 x 64), 12 radii
    for (int q = threadIdx.x; q < 128; q += 256) {
      int h = q >> 6, mm = q & 63;
      double theta = 0.25 * bq::kPI * c_xh[mm] + (h == 0 ? 0.25 * bq::kPI : 0.75 * bq::kPI);
      double wt = 0.25 * bq::kPI * c_wh[mm], cth, sth;
      sincos(theta, &sth, &cth);
      const double csd = dir == 0 ? cth : sth, outer = dir == 0 ? sth : cth;
      for (int i = 0; i < 12; ++i) {
        double rr = 0.5 * kDEPS * (c_xl[i] + 1.0), wr = wt * 0.5 * kDEPS * c_wl[i];
        double Kp[3], Km[3], Wp[3], Wm[3], Np[4], Nm[4];
        double zp = rr * cth + zt, ep = rr * sth + et, zm = -rr * cth + zt, em = -rr * sth + et;
        polar_K(e, zt, et, rr, cth, sth, Kp);
        polar_K(e, zt, et, -rr, cth, sth, Km);
#pragma unroll
        for (int d = 0; d < 3; ++d) {
          Wp[d] = e.cw[d][0] + e.cw[d][1] * zp + e.cw[d][2] * ep;
          Wm[d] = e.cw[d][0] + e.cw[d][1] * zm + e.cw[d][2] * em;
        }
        double aKp = sqrt(dot3(Kp, Kp)), aKm = sqrt(dot3(Km, Km));
        double aWp = sqrt(dot3(Wp, Wp)), aWm = sqrt(dot3(Wm, Wm));
        double KpT = dot3(Kp, T), KmT = dot3(Km, T);
        cplx ccp = ccos(k * (aKp * rr)), ccm = ccos(k * (aKm * rr));
        double ip5 = 1.0 / (aKp * aKp * aKp * aKp * aKp), im5 = 1.0 / (aKm * aKm * aKm * aKm * aKm);
        double gp = KpT / (aKp * aKp * aKp) * aWp, gm = KmT / (aKm * aKm * aKm) * aWm;
        double hp = (3.0 * KpT * csd - aKp * aKp) * ip5, hm = (3.0 * KmT * csd - aKm * aKm) * im5;
        shapeN4(zp, ep, Np);
        shapeN4(zm, em, Nm);
        double ir = wr / rr;
#pragma unroll
        for (int n = 0; n < 4; ++n) {
          cplx g = ccp * (gp * Np[n]) - ccm * (gm * Nm[n]);
          cplx hh = ccp * (hp * Np[n]) - ccm * (hm * Nm[n]);
          a[2 * n] -= ir * g.re;
          a[2 * n + 1] -= ir * g.im;
          a[8 + 2 * n] -= ir * outer * hh.re;
          a[8 + 2 * n + 1] -= ir * outer * hh.im;
        }
        a[16] += ir * outer * (hp - hm);
      }
    }
    }  // bilinear inner disk
    block_sum256<17>(a, lds);
    if (threadIdx.x == 0) {
      double sg = -kI4PI * iaT;
      double sh = plane ? 0.0 : bdC * kI4PI * iaT * sgn;
      long dr = (2 * row + dir) * N + col;
#pragma unroll
      for (int n = 0; n < 4; ++n) {
        dG[dr + n] = make_double2(a[2 * n] * sg, a[2 * n + 1] * sg);
        dH[dr + n] = make_double2(a[8 + 2 * n] * sh, a[8 + 2 * n + 1] * sh);
      }
      F3[3 * (row * dNe + (t - 1)) + 1 + dir] = a[16] * sh;
    }
  }
}

// ------------------------------------------------------------------------------------
__global__ void k_row_diag(int dNe, int open, const double* __restrict__ E, const int* __restrict__ sid,
                           const double* __restrict__ F3, double2* __restrict__ H, double* __restrict__ tdF)
{
  long row = (long)blockIdx.x * blockDim.x + threadIdx.x, N = 4L * dNe;
  if (row >= N) return;
  if ((row & 3) == 3 && elem_is_tri(E, abs(sid[row / 4 + 1]))) {  // no such node: row stays 0
    if (tdF) tdF[3 * row] = tdF[3 * row + 1] = tdF[3 * row + 2] = 0.0;
    return;
  }
  double F = 0, F0 = 0, F1 = 0;
  const double* f = F3 + 3 * row * dNe;
  for (int s = 0; s < dNe; ++s) {
    F += f[3 * s];
    F0 += f[3 * s + 1];
    F1 += f[3 * s + 2];
  }
  double C = open ? 1.0 + F : F;
  H[row * N + row].x += C;
  if (tdF) {
    tdF[3 * row] = C;
    tdF[3 * row + 1] = F0;
    tdF[3 * row + 2] = F1;
  }
}

// algorithmic point counts per phase: [0] regular rules, [1] 64-point rule, [2] singular
// polar rules (QPGF evaluations for the periodic domain)
__global__ void __launch_bounds__(256) k_count_points(long n, int dNe, const double* __restrict__ E,
                                                      const int* __restrict__ sid,
                                                      const unsigned char* __restrict__ flag, bq_qplan P,
                                                      int periodic, int full_rule,
                                                      unsigned long long* __restrict__ cnt)
{
  __shared__ double lds[3 * 4];
  long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  double v[3] = {0.0, 0.0, 0.0};
  if (i < n) {
    const int f = flag[i];
    const int as = abs(sid[(int)(i % dNe) + 1]);
    const bool tri = elem_is_tri(E, as);
    if (f == 0) {
      int ng = 12;
      if (periodic) {
        Elem et, es;
        load_elem(E, abs(sid[(int)(i / dNe) / 4 + 1]), et);
        load_elem(E, as, es);
        ng = pair_rule(P, et, es, full_rule);
      }
      v[0] = (tri ? 3.0 : 1.0) * ng * ng;
    }
    else if (f == 1) v[1] = tri ? 12288.0 : 4096.0;
    else if (f == 2) v[2] = tri ? 4608.0 : 6144.0;  // sectors x 128 angles x 12 radii
  }
  block_sum256<3>(v, lds);
  if (threadIdx.x < 3 && v[threadIdx.x] != 0.0) atomicAdd(&cnt[threadIdx.x], (unsigned long long)v[threadIdx.x]);
}

__global__ void k_collect_gh(long n, const unsigned char* __restrict__ flag, int* __restrict__ list,
                             int* __restrict__ count)
{
  long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n || flag[i] != 1) return;
  int slot = atomicAdd(count, 1);
  list[slot] = (int)i;
}

// ------------------------------------------------------------------------------------
extern "C" {

int bq_set_quadrature(const double* xl, const double* wl, const double* xh, const double* wh)
{
  double xv[kVruleRows][10] = {{0}}, wv[kVruleRows][10] = {{0}};
  for (int r = 0; r < kVruleRows; ++r) bq_gauss_legendre(kVruleMin + r, xv[r], wv[r]);
  if (bq::set_quad_tables(xl, wl, xh, wh) || bq::set_vrule_tables(&xv[0][0], &wv[0][0])) {
    bq::set_error("quadrature table upload failed");
    return -1;
  }
  return 0;
}

// Assemble domain matrices on the device.  All pointers are device pointers; G, H are
// N x N, dG, dH are (2N) x N (rows 2*row + {0: zeta, 1: eta}), tdF is N x 3, work must
// hold bq_assemble_work_bytes(dNe).  dG/dH/tdF may be null for the periodic domain.
size_t bq_assemble_work_bytes(int dNe)
{
  size_t pairs = (size_t)4 * dNe * dNe;
  return 256 + pairs * 3 * sizeof(double) + pairs + pairs * sizeof(int);
}

int bq_assemble_domain_dev(const bq_adom* a, void* work, void* stream)
{
  hipStream_t s = (hipStream_t)stream;
  const int dNe = a->dNe;
  const long N = 4L * dNe, pairs = N * dNe;
  if (dNe <= 0) return 0;
  // work (256-byte aligned): counters, then F3 [pairs][3], flag [pairs], GH list [pairs]
  char* w = (char*)work;
  int* count = (int*)w;
  unsigned long long* npts = (unsigned long long*)(w + 64);
  double* F3 = (double*)(w + 256);
  unsigned char* flag = (unsigned char*)(w + 256 + pairs * 3 * sizeof(double));
  int* list = (int*)(w + 256 + pairs * 3 * sizeof(double) + pairs);
  double2* G = (double2*)a->G;
  double2* H = (double2*)a->H;
  BQ_HIP(hipMemsetAsync(G, 0, sizeof(double2) * N * N, s));
  BQ_HIP(hipMemsetAsync(H, 0, sizeof(double2) * N * N, s));
  BQ_HIP(hipMemsetAsync(count, 0, sizeof(int), s));
  BQ_HIP(hipMemsetAsync(npts, 0, 3 * sizeof(unsigned long long), s));
  const int freespace = a->periodic ? 0 : 1;
  BQ_KTIME("k_classify", s, k_classify<<<(unsigned)((pairs + 255) / 256), 256, 0, s>>>(a->E, a->ren, a->sid, dNe, freespace, a->full_rule, a->plan.rule_ltol, F3, flag));
  BQ_HIP(hipGetLastError());
  k_collect_gh<<<(unsigned)((pairs + 255) / 256), 256, 0, s>>>(pairs, flag, list, count);
  BQ_HIP(hipGetLastError());
  k_count_points<<<(unsigned)((pairs + 255) / 256), 256, 0, s>>>(pairs, dNe, a->E, a->sid, flag, a->plan,
                                                                  a->periodic, a->full_rule, npts);
  BQ_HIP(hipGetLastError());
  int ngh = 0;
  unsigned long long np3[3] = {0, 0, 0};
  BQ_HIP(hipMemcpyAsync(&ngh, count, sizeof(int), hipMemcpyDeviceToHost, s));
  BQ_HIP(hipMemcpyAsync(np3, npts, sizeof np3, hipMemcpyDeviceToHost, s));
  BQ_HIP(hipStreamSynchronize(s));
  const unsigned gl_blocks = (unsigned)(((long)dNe * dNe + 3) / 4);
  if (a->periodic) {
    BQ_KTIME("k_q0_gl", s,
             k_q0_gl<<<gl_blocks, 256, 0, s>>>(a->plan, a->E, a->ren, a->sid, dNe, flag, a->full_rule, G, H));
    BQ_HIP(hipGetLastError());
    if (ngh) BQ_KTIME("k_q0_gh", s, k_q0_gh<<<ngh, 256, 0, s>>>(a->plan, a->E, a->ren, a->sid, dNe, list, G, H, F3));
    BQ_HIP(hipGetLastError());
    BQ_KTIME("k_q0_sing", s, k_q0_sing<<<(unsigned)N, 256, 0, s>>>(a->plan, a->E, a->sid, dNe, G, H, F3));
    BQ_HIP(hipGetLastError());
  } else {
    double2* dG = (double2*)a->dG;
    double2* dH = (double2*)a->dH;
    BQ_HIP(hipMemsetAsync(dG, 0, sizeof(double2) * 2 * N * N, s));
    BQ_HIP(hipMemsetAsync(dH, 0, sizeof(double2) * 2 * N * N, s));
    BQ_KTIME("k_fs_gl", s, k_fs_gl<<<gl_blocks, 256, 0, s>>>(a->kr, a->ki, a->E, a->ren, a->sid, dNe, flag, a->full_rule, a->plan.rule_ltol, G, H, dG, dH));
    BQ_HIP(hipGetLastError());
    if (ngh) BQ_KTIME("k_fs_gh", s, k_fs_gh<<<ngh, 256, 0, s>>>(a->kr, a->ki, a->E, a->ren, a->sid, dNe, list, G, H, dG, dH, F3));
    BQ_HIP(hipGetLastError());
    BQ_KTIME("k_fs_sing", s, k_fs_sing<<<(unsigned)N, 256, 0, s>>>(a->kr, a->ki, a->E, a->sid, dNe, G, H, dG, dH, F3));
    BQ_HIP(hipGetLastError());
  }
  k_row_diag<<<(unsigned)((N + 255) / 256), 256, 0, s>>>(dNe, a->open_domain, a->E, a->sid, F3, H, a->tdF);
  BQ_HIP(hipGetLastError());
  if (a->n_gh_out) *a->n_gh_out = ngh;
  if (a->n_points_out)
    for (int i = 0; i < 3; ++i) a->n_points_out[i] = (long long)np3[i];
  return 0;
}

}  // extern "C"
