// bq_field.hip -- fields at arbitrary points from the boundary solution (EH_s of
// d3qd2_src/d3qd2_field.c:203-269):
//   E(r) = [sum_s sum_n G_n(r, s) dE_s,n - H_n(r, s) E_s,n] / C,  same for H,
//   C = 1 + F (open region) or F, F = sum_s F(r, s)  (solid angle / 4 pi),
// with the coefficients of q0_coef_rt / coef_rt (d3qd2_src/d3qd2_elem.c:148-185, :574-609):
// type 0 = the element nodes (a triangle's 3 nodes and centroid), 1 = 9-point (quad) or
// 7-point (triangle) rule, 2 = 12-point, 3 = 64-point Gauss rule (n^2 points on a quad,
// 3 n^2 on a triangle).  Weights are in reference-area units and |W| is the Jacobian
// (bq_elem.hpp), so one integrand serves both shapes.
// One 256-thread block per point; threads stride over (element, quadrature point) pairs
// and the block reduces E, H, F in a fixed order.
#include "bq_runtime.hpp"
#include "bq_qpgf.hpp"
#include "bq_elem.hpp"
#include "bq_reduce.hpp"
#include <cstring>
#include <vector>
#include "../bq_field.h"

using namespace bq;

namespace {

__constant__ double c_z49[9], c_e49[9], c_w49[9];
// triangle node-rule weights and the 7-point rule, weights times the reference-triangle
// area 3 sqrt3 / 4 (the reference's weights sum to 1)
__constant__ double c_w34[4], c_z37[7], c_e37[7], c_w37[7];

__device__ __forceinline__ cplx ld2(const double* p, long i) { return mk(p[2 * i], p[2 * i + 1]); }

__device__ __forceinline__ int rule_size(int type, bool tri)
{
  return type == 0 ? 4 : type == 1 ? (tri ? 7 : 9) : gauss_points(tri, type == 3);
}

__device__ __forceinline__ void rule_point(int type, bool tri, int q, double& z, double& e, double& w)
{
  if (type == 0) {
    z = elem_node_z(tri, q);
    e = elem_node_e(tri, q);
    w = tri ? c_w34[q] : 1.0;
  } else if (type == 1) {
    z = tri ? c_z37[q] : c_z49[q];
    e = tri ? c_e37[q] : c_e49[q];
    w = tri ? c_w37[q] : c_w49[q];
  } else {
    gauss_point(tri, type == 3, q, z, e, w);
  }
}

// Point q of rule `type` on element e (id as): position R, normal W, interpolation weights Nn
// and rule weight w.  The node rule reads the stored node data (d3qd2_src/q0_bil_coef.c:28-30,
// q0_lit_coef.c:29-31); a triangle's normal is constant and taken from the element, and its
// centroid point carries 1/3 of each function.
__device__ __forceinline__ void field_rule_point(const bq_fctx& C, const Elem& e, int as, int type, int q,
                                                 double (&R)[3], double (&W)[3], double (&Nn)[4], double& w)
{
  double z, et;
  rule_point(type, e.tri, q, z, et, w);
  if (type == 0) {
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      R[d] = C.ren[12L * as + 3 * q + d];
      W[d] = e.tri ? e.cw[d][0] : C.wen[12L * as + 3 * q + d];
    }
    if (e.tri && q == 3) {
      shapeN3(z, et, Nn);
    } else {
#pragma unroll
      for (int n = 0; n < 4; ++n) Nn[n] = (n == q) ? 1.0 : 0.0;
    }
  } else {
    rw_at(e, z, et, R, W);
    shapeN(e.tri, z, et, Nn);
  }
}

// Node data of domain did: KIND 0 -> E, dE, H, dH ([(dNe+1)][4][3] complex), 13 outputs per
// point (E, H, F); KIND 1 -> U, dU ([(dNe+1)][4][4] complex), 9 outputs (U, F).
template <int KIND>
struct Acc {
  static constexpr int NV = KIND == 0 ? 13 : 9;
};

// add g dX - h X over the element's nodes (weights already in g, h per node)
template <int KIND>
__device__ __forceinline__ void add_node_terms(const bq_fctx& C, int did, int sc, const cplx (&gn)[4],
                                               const cplx (&hn)[4], double (&acc)[Acc<KIND>::NV])
{
  if (KIND == 0) {
    const double *Eb = C.Eb[did], *dEb = C.dEb[did], *Hb = C.Hb[did], *dHb = C.dHb[did];
#pragma unroll
    for (int n = 0; n < 4; ++n) {
      const long o = 12L * (sc + 1) + 3 * n;
#pragma unroll
      for (int i = 0; i < 3; ++i) {
        cplx ev = gn[n] * ld2(dEb, o + i) - hn[n] * ld2(Eb, o + i);
        cplx hv = gn[n] * ld2(dHb, o + i) - hn[n] * ld2(Hb, o + i);
        acc[2 * i] += ev.re;
        acc[2 * i + 1] += ev.im;
        acc[6 + 2 * i] += hv.re;
        acc[6 + 2 * i + 1] += hv.im;
      }
    }
  } else {
    const double *Ub = C.Ub[did], *dUb = C.dUb[did];
#pragma unroll
    for (int n = 0; n < 4; ++n) {
      const long o = 16L * (sc + 1) + 4 * n;
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        cplx uv = gn[n] * ld2(dUb, o + i) - hn[n] * ld2(Ub, o + i);
        acc[2 * i] += uv.re;
        acc[2 * i + 1] += uv.im;
      }
    }
  }
}

// One 256-thread block per point.  Free points: pts[3p].  Boundary points (bd_t != null):
// the point (zeta, eta) = bd_ze[2p] of element t = bd_t[p] (position in the domain's list);
// every other element uses the rule `type` at that point, the element itself the polar rule
// around it (q0_coef_bd / coef_bd, d3qd2_src/d3qd2_elem.c:684-770, :187-230).
template <int KIND>
__global__ void __launch_bounds__(256) k_field(bq_fctx C, int npts, const double* __restrict__ pts,
                                               const int* __restrict__ did_of, const int* __restrict__ bd_t,
                                               const double* __restrict__ bd_ze, int type, double* __restrict__ out)
{
  constexpr int NV = Acc<KIND>::NV;
  __shared__ double lds[NV * 4];
  const int p = blockIdx.x;
  const int did = did_of[p];
  const int dNe = C.dNe[did];
  const int* sid = C.sid[did];
  const bool periodic = (did == 0 && C.ps == 1);
  const cplx k = mk(C.kn[2 * did], C.kn[2 * did + 1]);
  const int self = bd_t ? bd_t[p] : 0;  // 1-based position of the point's element, 0 if none
  double r0[3];
  if (self) {
    Elem et;
    load_elem(C.E, abs(sid[self]), et);
    double W[3];
    rw_at(et, bd_ze[2 * p], bd_ze[2 * p + 1], r0, W);
  } else {
    r0[0] = pts[3 * p];
    r0[1] = pts[3 * p + 1];
    r0[2] = pts[3 * p + 2];
  }
  const int nq = C.has_tri ? max(rule_size(type, false), rule_size(type, true)) : rule_size(type, false);
  double acc[NV];
#pragma unroll
  for (int i = 0; i < NV; ++i) acc[i] = 0.0;
  const long total = (long)dNe * nq;
  for (long it = threadIdx.x; it < total; it += 256) {
    const int sc = (int)(it / nq), q = (int)(it - (long)sc * nq);
    if (sc + 1 == self) continue;
    const int sd = sid[sc + 1], as = abs(sd);
    Elem e;
    load_elem(C.E, as, e);
    if (q >= rule_size(type, e.tri)) continue;
    double w, R[3], W[3], Nn[4];
    field_rule_point(C, e, as, type, q, R, W, Nn, w);
    R[0] -= r0[0]; R[1] -= r0[1]; R[2] -= r0[2];
    double aR = sqrt(dot3(R, R)), iR = 1.0 / aR, aW = sqrt(dot3(W, W));
    double tD = dot3(R, W) * iR * iR * iR;
    cplx g, h;
    double sgn = 1.0;
    if (periodic) {
      QG qq;
      qpgf_v1(C.plan, R[0], R[1], R[2], qq);
      g = qq.g * aW;
      h = qq.d[0] * W[0] + qq.d[1] * W[1] + qq.d[2] * W[2];
    } else {
      cplx ca = cmul_i(k * aR);
      cplx ce = cexp(ca);
      g = ce * (iR * aW * kI4PI);
      h = (ca - mk(1.0, 0.0)) * ce * (tD * kI4PI);
      if (sd < 0) sgn = -1.0;
    }
    acc[NV - 1] += w * tD * kI4PI * sgn;
    cplx gn[4], hn[4];
#pragma unroll
    for (int n = 0; n < 4; ++n) {
      gn[n] = g * (w * Nn[n]);
      hn[n] = h * (w * Nn[n] * sgn);
    }
    add_node_terms<KIND>(C, did, sc, gn, hn, acc);
  }
  if (self) {
    // the point's own element: polar rule around (zeta, eta), 128 angles per sector in two
    // 64-point halves, 12 radii (q0_bil_coef_bd d3qd2_src/q0_bil_coef.c:246-338,
    // q0_lit_coef_bd q0_lit_coef.c:297-371, bil_coef_bd bil_coef.c:244-342, lit_coef_bd)
    const int sd = sid[self];
    Elem e;
    load_elem(C.E, abs(sd), e);
    const double zt = bd_ze[2 * p], ett = bd_ze[2 * p + 1];
    double rr5[5], th[5];
    const int nsec = corner_r_th(e.tri, zt, ett, rr5, th);
    const bool plane = elem_plane(e);
    const double sF = plane ? 0.0 : elem_bdC(e) * kI4PI;  // polar F and free-space H factor
    const double sg = (!periodic && sd < 0) ? -1.0 : 1.0;
    for (int q = threadIdx.x; q < nsec * 1536; q += 256) {
      const int j = q / 1536, m = (q / 12) % 128, i = q % 12;
      double theta, wt;
      polar_theta(th, j, m, theta, wt);
      double rm = polar_rm(rr5, th, j, theta), cth, sth;
      sincos(theta, &sth, &cth);
      const double rr = 0.5 * rm * (c_xl[i] + 1.0), w = wt * 0.5 * rm * c_wl[i];
      double K[3], R[3], W[3], Nn[4];
      polar_K(e, zt, ett, rr, cth, sth, K);
      const double z = rr * cth + zt, ee = rr * sth + ett;
#pragma unroll
      for (int d = 0; d < 3; ++d) {
        R[d] = rr * K[d];
        W[d] = e.cw[d][0] + e.cw[d][1] * z + e.cw[d][2] * ee;
      }
      const double aK = sqrt(dot3(K, K)), aW = sqrt(dot3(W, W));
      cplx g, h;
      if (periodic) {
        QG qq;
        qpgf_v1(C.plan, R[0], R[1], R[2], qq);
        // the reference takes the value from d3hm_qpgf_d2_qG here (q0_bil_coef.c:444,
        // q0_lit_coef.c:569); it differs from _v1 only at R_z == 0
        if (R[2] == 0.0) {
          QG qa;
          qpgf_v1(C.plan, R[0], R[1], R[2], qa, false);
          qq.g = qa.g;
        }
        g = qq.g * (aW * rr);
        h = (qq.d[0] * W[0] + qq.d[1] * W[1] + qq.d[2] * W[2]) * rr;
      } else {
        const cplx kR = k * (rr * aK), ce = cexp(cmul_i(kR));
        g = ce * (aW / aK * kI4PI);
        h = (cmul_i(kR) - mk(1.0, 0.0)) * ce * (cth * sth / (aK * aK * aK) * sF * sg);
      }
      acc[NV - 1] += w * cth * sth / (aK * aK * aK) * sF * sg;
      shapeN(e.tri, z, ee, Nn);
      cplx gn[4], hn[4];
#pragma unroll
      for (int n = 0; n < 4; ++n) {
        gn[n] = g * (w * Nn[n]);
        hn[n] = h * (w * Nn[n]);
      }
      add_node_terms<KIND>(C, did, self - 1, gn, hn, acc);
    }
  }
  block_sum256<NV>(acc, lds);
  if (threadIdx.x == 0) {
    double* o = out + (long)NV * p;
    for (int i = 0; i < NV; ++i) o[i] = acc[i];
  }
}


// E, H from the potentials (EH_mEMP_s, d3qd2_src/d3qd2_field.c:90-173): one block per point,
// out[p][36] = U (4 complex), dU/dr_j (3 x 4 complex), F, dF/dr_j.  Coefficients and their
// field-point gradients as q0_dcoef_rt_grad / dcoef_rt_grad (d3qd2_src/d3qd2_elem.c:316-349,
// :650-683): the periodic kernel takes the Hessian of the lattice sum (qpgf_v2).
__global__ void __launch_bounds__(256) k_field_grad(bq_fctx C, int npts, const double* __restrict__ pts,
                                                    const int* __restrict__ did_of, int type,
                                                    double* __restrict__ out)
{
  constexpr int NV = 36;
  __shared__ double lds[NV * 4];
  const int p = blockIdx.x;
  const int did = did_of[p];
  const int dNe = C.dNe[did];
  const int* sid = C.sid[did];
  const bool periodic = (did == 0 && C.ps == 1);
  const cplx k = mk(C.kn[2 * did], C.kn[2 * did + 1]), k2 = k * k;
  const double r0[3] = {pts[3 * p], pts[3 * p + 1], pts[3 * p + 2]};
  const double *Ub = C.Ub[did], *dUb = C.dUb[did];
  const int nq = C.has_tri ? max(rule_size(type, false), rule_size(type, true)) : rule_size(type, false);
  double acc[NV];
#pragma unroll
  for (int i = 0; i < NV; ++i) acc[i] = 0.0;
  const long total = (long)dNe * nq;
  for (long it = threadIdx.x; it < total; it += 256) {
    const int sc = (int)(it / nq), q = (int)(it - (long)sc * nq);
    const int sd = sid[sc + 1], as = abs(sd);
    Elem e;
    load_elem(C.E, as, e);
    if (q >= rule_size(type, e.tri)) continue;
    double w, R[3], W[3], Nn[4];
    field_rule_point(C, e, as, type, q, R, W, Nn, w);
    R[0] -= r0[0]; R[1] -= r0[1]; R[2] -= r0[2];
    const double aR = sqrt(dot3(R, R)), iR = 1.0 / aR, aW = sqrt(dot3(W, W));
    const double rdW = dot3(R, W) * iR, tD = rdW * iR * iR;
    cplx g, h, dg[3], dh[3];
    double sgn = 1.0;
    if (periodic) {
      QG2 qq;
      qpgf_v2(C.plan, R[0], R[1], R[2], qq);
      g = qq.g * aW;
      h = qq.d[0] * W[0] + qq.d[1] * W[1] + qq.d[2] * W[2];
      // rows of the Hessian: (xx, xy, zx), (xy, yy, yz), (zx, yz, zz)
      dg[0] = qq.d[0] * (-aW);
      dg[1] = qq.d[1] * (-aW);
      dg[2] = qq.d[2] * (-aW);
      dh[0] = (qq.h[0] * W[0] + qq.h[3] * W[1] + qq.h[5] * W[2]) * -1.0;
      dh[1] = (qq.h[3] * W[0] + qq.h[1] * W[1] + qq.h[4] * W[2]) * -1.0;
      dh[2] = (qq.h[5] * W[0] + qq.h[4] * W[1] + qq.h[2] * W[2]) * -1.0;
    } else {
      const cplx ca = cmul_i(k * aR), ce = cexp(ca), cb = ca - mk(1.0, 0.0);
      g = ce * (iR * aW * kI4PI);
      h = cb * ce * (tD * kI4PI);
      const cplx gg = cb * ce * (-kI4PI * iR * iR * iR * aW);
      const cplx t1 = (k2 * (aR * aR) + 3.0 * cb) * ce * (kI4PI * iR * iR * iR * iR * rdW);
      const cplx t2 = cb * ce * (kI4PI * iR * iR * iR);
#pragma unroll
      for (int d = 0; d < 3; ++d) {
        dg[d] = gg * R[d];
        dh[d] = t1 * R[d] - t2 * W[d];
      }
      if (sd < 0) sgn = -1.0;
    }
    const double fw = w * kI4PI * sgn;
    acc[32] += fw * tD;
#pragma unroll
    for (int d = 0; d < 3; ++d) acc[33 + d] += fw * iR * iR * iR * (3.0 * R[d] * iR * rdW - W[d]);
#pragma unroll
    for (int n = 0; n < 4; ++n) {
      const double wn = w * Nn[n], wh = wn * sgn;
      const long o = 16L * (sc + 1) + 4 * n;
#pragma unroll
      for (int c = 0; c < 4; ++c) {
        const cplx u = ld2(Ub, o + c), du = ld2(dUb, o + c);
        const cplx uv = (g * du - h * sgn * u) * wn;
        acc[2 * c] += uv.re;
        acc[2 * c + 1] += uv.im;
#pragma unroll
        for (int d = 0; d < 3; ++d) {
          const cplx dv = dg[d] * du * wn - dh[d] * u * wh;
          acc[8 + 8 * d + 2 * c] += dv.re;
          acc[8 + 8 * d + 2 * c + 1] += dv.im;
        }
      }
    }
  }
  block_sum256<NV>(acc, lds);
  if (threadIdx.x == 0) {
    double* o = out + (long)NV * p;
    for (int i = 0; i < NV; ++i) o[i] = acc[i];
  }
}
}  // namespace

extern "C" int bq_field_rules_init(const double* z49, const double* e49, const double* w49, const double* w34,
                                   const double* z37, const double* e37, const double* w37)
{
  const double area = 0.75 * sqrt(3.0);
  double a34[4], a37[7];
  for (int i = 0; i < 4; ++i) a34[i] = w34[i] * area;
  for (int i = 0; i < 7; ++i) a37[i] = w37[i] * area;
  BQ_HIP(hipMemcpyToSymbol(HIP_SYMBOL(c_z49), z49, 9 * sizeof(double)));
  BQ_HIP(hipMemcpyToSymbol(HIP_SYMBOL(c_e49), e49, 9 * sizeof(double)));
  BQ_HIP(hipMemcpyToSymbol(HIP_SYMBOL(c_w49), w49, 9 * sizeof(double)));
  BQ_HIP(hipMemcpyToSymbol(HIP_SYMBOL(c_w34), a34, 4 * sizeof(double)));
  BQ_HIP(hipMemcpyToSymbol(HIP_SYMBOL(c_z37), z37, 7 * sizeof(double)));
  BQ_HIP(hipMemcpyToSymbol(HIP_SYMBOL(c_e37), e37, 7 * sizeof(double)));
  BQ_HIP(hipMemcpyToSymbol(HIP_SYMBOL(c_w37), a37, 7 * sizeof(double)));
  return 0;
}

// out[p][13] = E (3 complex), H (3 complex), F for each folded point
// kind 0: out[p][13] = E (3 complex), H (3 complex), F; kind 1: out[p][9] = U (4 complex), F;
// kind 2: out[p][36] = U, dU/dr (3 x 4 complex), F, dF/dr (free points only).
// d_bd_t / d_bd_ze (may be null): boundary queries instead of free points.
extern "C" int bq_field_eval_dev(const bq_fctx* C, int kind, int npts, const double* d_pts, const int* d_did,
                                 const int* d_bd_t, const double* d_bd_ze, int type, double* d_out, void* stream)
{
  if (npts <= 0) return 0;
  hipStream_t s = (hipStream_t)stream;
  if (kind == 0)
    BQ_KTIME("k_field", s, k_field<0><<<npts, 256, 0, s>>>(*C, npts, d_pts, d_did, d_bd_t, d_bd_ze, type, d_out));
  else if (kind == 1)
    BQ_KTIME("k_field_u", s, k_field<1><<<npts, 256, 0, s>>>(*C, npts, d_pts, d_did, d_bd_t, d_bd_ze, type, d_out));
  else
    BQ_KTIME("k_field_grad", s, k_field_grad<<<npts, 256, 0, s>>>(*C, npts, d_pts, d_did, type, d_out));
  BQ_HIP(hipGetLastError());
  return 0;
}

// ------------------------------------------------------------------------------------
// context: the solved model resident in HBM for repeated field evaluation
namespace {
struct FieldCtx {
  bq_fctx C;
  std::vector<void*> owned;
  int Ne = 0, MN = 0;
  const double* rn = nullptr;
  const int *ed = nullptr, *md = nullptr, *sd = nullptr;
};

// ---- domain of a point: sum of signed solid angles (the reference's domain_id) -----------
// Solid angle of the triangle (r0, r1, r2) seen from the origin, by the spherical excess of
// the three dihedral angles between the planes through the origin and the edges (the
// reference's fid_calc_solid_angle, d3qd2_src/d3qd2_setup.c:1351-1422), with its sign from
// the side the origin is on; flg = -1 when the origin is on the triangle's plane.
__device__ double tri_omega_dev(const double r0[3], const double r1[3], const double r2[3], int& flg, double n20o[3])
{
#pragma clang fp contract(off)
  auto crs = [](double* o, const double* a, const double* b) {
    o[0] = a[1] * b[2] - a[2] * b[1];
    o[1] = a[2] * b[0] - a[0] * b[2];
    o[2] = a[0] * b[1] - a[1] * b[0];
  };
  auto dt = [](const double* a, const double* b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; };
  auto ab = [&](const double* a) { return sqrt(dt(a, a)); };
  double n01[3], n12[3], n20[3], nt[3];
  crs(n01, r0, r1);
  crs(n12, r1, r2);
  crs(n20, r2, r0);
  crs(nt, n01, n20);
  const double a0 = atan2(ab(nt), -dt(n01, n20));
  crs(nt, n12, n01);
  const double a1 = atan2(ab(nt), -dt(n12, n01));
  crs(nt, n20, n12);
  const double a2 = atan2(ab(nt), -dt(n20, n12));
  const double D = dt(r0, n12);
  for (int i = 0; i < 3; ++i) n20o[i] = n20[i];
  if (D > 0.0) return a0 + a1 + a2 - kPI;
  if (D < 0.0) return -(a0 + a1 + a2 - kPI);
  flg = -1;
  return 0.0;
}

// element solid angle: a triangle, or a quad as the triangles (0, 1, 2) and (0, 2, 3) with the
// shared edge's plane reused, as the reference splits it
__device__ double elem_omega_dev(bool quad, double r[4][3], int& flg)
{
#pragma clang fp contract(off)
  flg = 0;
  for (int i = 0; i < (quad ? 4 : 3); ++i) {
    const double a = sqrt(r[i][0] * r[i][0] + r[i][1] * r[i][1] + r[i][2] * r[i][2]);
    if (a == 0.0) {
      flg = -1;
      return 0.0;
    }
    r[i][0] /= a;
    r[i][1] /= a;
    r[i][2] /= a;
  }
  double n20[3];
  double om = tri_omega_dev(r[0], r[1], r[2], flg, n20);
  if (flg < 0 || !quad) return om;
  auto crs = [](double* o, const double* a, const double* b) {
    o[0] = a[1] * b[2] - a[2] * b[1];
    o[1] = a[2] * b[0] - a[0] * b[2];
    o[2] = a[0] * b[1] - a[1] * b[0];
  };
  auto dt = [](const double* a, const double* b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; };
  auto ab = [&](const double* a) { return sqrt(dt(a, a)); };
  double n01[3] = {-n20[0], -n20[1], -n20[2]}, n12[3], nb[3], nt[3];
  crs(n12, r[2], r[3]);
  crs(nb, r[3], r[0]);
  crs(nt, n01, nb);
  const double a0 = atan2(ab(nt), -dt(n01, nb));
  crs(nt, n12, n01);
  const double a1 = atan2(ab(nt), -dt(n12, n01));
  crs(nt, nb, n12);
  const double a2 = atan2(ab(nt), -dt(nb, n12));
  const double D = dt(r[0], n12);
  if (D > 0.0) om += a0 + a1 + a2 - kPI;
  else if (D < 0.0) om -= a0 + a1 + a2 - kPI;
  else flg = -1;
  return flg < 0 ? 0.0 : om;
}

// One workgroup per point.  The reference walks the domains in order, adding each element's
// solid angle to its main domain and subtracting it from its sub-domain, and decides a domain
// when its sum is complete: the open region when |sum| < 2 pi, another when sum > 2 pi, and a
// domain at once when the point is on the plane of one of its elements.  So domain d's sum holds
// its own elements and those of lower-numbered domains that have d as sub-domain.
__global__ void __launch_bounds__(256) k_domain_id(int n, const double* __restrict__ pts, int Ne, int MN,
                                                   const double* __restrict__ rn, const int* __restrict__ ed,
                                                   const int* __restrict__ md, const int* __restrict__ sd,
                                                   int* __restrict__ did)
{
  __shared__ double red[4][BQ_MAXD];
  __shared__ int rfl[4];
  const int p = blockIdx.x;
  if (p >= n) return;
  const double rt[3] = {pts[3 * p], pts[3 * p + 1], pts[3 * p + 2]};
  double acc[BQ_MAXD];
  for (int d = 0; d < BQ_MAXD; ++d) acc[d] = 0.0;
  int fl = BQ_MAXD;  // lowest domain with an element whose plane holds the point
  for (int s = 1 + threadIdx.x; s <= Ne; s += 256) {
    const int m = md[s], q = sd[s];
    const bool quad = ed[4 * s + 3] != 0;
    double rv[4][3];
    for (int j = 0; j < 4; ++j)
      for (int k = 0; k < 3; ++k) rv[j][k] = rn[3 * (size_t)ed[4 * s + j] + k] - rt[k];
    int flg = 0;
    const double om = elem_omega_dev(quad, rv, flg);
    if (flg < 0) {
      fl = min(fl, m);
      continue;
    }
    for (int d = 0; d <= MN; ++d) {
      if (d == m) acc[d] += om;
      if (d == q && m < q) acc[d] -= om;
    }
  }
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  for (int d = 0; d <= MN; ++d) {
    double v = acc[d];
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
    if (lane == 0) red[wv][d] = v;
  }
  for (int o = 32; o > 0; o >>= 1) fl = min(fl, __shfl_xor(fl, o));
  if (lane == 0) rfl[wv] = fl;
  __syncthreads();
  if (threadIdx.x != 0) return;
  const int f = min(min(rfl[0], rfl[1]), min(rfl[2], rfl[3]));
  int res = -1;
  for (int d = 0; d <= MN; ++d) {
    if (f == d) {
      res = d;
      break;
    }
    const double og = red[0][d] + red[1][d] + red[2][d] + red[3][d];
    if (d == 0 ? fabs(og) < 2.0 * kPI : og > 2.0 * kPI) {
      res = d;
      break;
    }
  }
  did[p] = res;
}

template <class T>
int up(FieldCtx* F, const T* h, size_t n, const T** out, hipStream_t s)
{
  void* p = nullptr;
  if (n) {
    BQ_HIP(hipMalloc(&p, n * sizeof(T)));
    BQ_HIP(hipMemcpyAsync(p, h, n * sizeof(T), hipMemcpyHostToDevice, s));
    F->owned.push_back(p);
  }
  *out = (const T*)p;
  return 0;
}
}  // namespace

extern "C" int bq_field_ctx_create(const bq_fdata* f, void** handle)
{
  if (bq::ensure_device()) return -1;
  if (f->MN + 1 > BQ_MAXD) {
    bq::set_error("field context: %d domains exceed the supported %d", f->MN + 1, BQ_MAXD);
    return -2;
  }
  hipStream_t s = bq::stream();
  FieldCtx* F = new FieldCtx();
  bq_fctx& C = F->C;
  memset(&C, 0, sizeof C);
  C.ps = f->ps;
  C.MN = f->MN;
  if (bq_qplan_init(&C.plan, f->qd[0], f->qd + 1, f->qd + 4, f->qd + 6, 1.0e-15) && f->ps == 1) {
    bq::set_error("field context: degenerate lattice");
    delete F;
    return -3;
  }
  {
    // the QPGF error word of this context's launches
    const int zero = 0;
    const int* perr = nullptr;
    if (up(F, &zero, 1, &perr, s)) {
      bq_field_ctx_destroy(F);
      return -1;
    }
    C.plan.err = (int*)perr;
  }
  size_t n1 = (size_t)f->Ne + 1;
  for (size_t i = 1; i < n1; ++i)
    if (f->E[24 * i + 21] != 0.0) C.has_tri = 1;
  F->Ne = f->Ne;
  F->MN = f->MN;
  if (f->rn && (up(F, f->rn, (size_t)(f->Nn + 1) * 3, &F->rn, s) || up(F, f->ed, n1 * 4, &F->ed, s) ||
                up(F, f->md, n1, &F->md, s) || up(F, f->sd, n1, &F->sd, s))) {
    bq_field_ctx_destroy(F);
    return -1;
  }
  if (up(F, f->E, n1 * 24, &C.E, s) || up(F, f->ren, n1 * 12, &C.ren, s) || up(F, f->wen, n1 * 12, &C.wen, s) ||
      up(F, f->kn, (size_t)(f->MN + 1) * 2, &C.kn, s)) {
    bq_field_ctx_destroy(F);
    return -1;
  }
  for (int d = 0; d <= f->MN; ++d) {
    size_t m = (size_t)f->dNe[d] + 1;
    C.dNe[d] = f->dNe[d];
    if (up(F, f->sid[d], m, &C.sid[d], s) || up(F, f->Eb[d], m * 24, &C.Eb[d], s) ||
        up(F, f->dEb[d], m * 24, &C.dEb[d], s) || up(F, f->Hb[d], m * 24, &C.Hb[d], s) ||
        up(F, f->dHb[d], m * 24, &C.dHb[d], s) || up(F, f->Ub[d], m * 32, &C.Ub[d], s) ||
        up(F, f->dUb[d], m * 32, &C.dUb[d], s)) {
      bq_field_ctx_destroy(F);
      return -1;
    }
  }
  if (bq::set_quad_tables(f->xli, f->wli, f->xhi, f->whi) ||
      bq_field_rules_init(f->z49, f->e49, f->w49, f->w34, f->z37, f->e37, f->w37)) {
    bq_field_ctx_destroy(F);
    return -1;
  }
  BQ_HIP(hipStreamSynchronize(s));
  *handle = F;
  return 0;
}

extern "C" int bq_field_eval(void* handle, int kind, int n, const double* pts, const int* did, const int* bd_t,
                             const double* bd_ze, int type, double* out)
{
  if (n <= 0) return 0;
  if (type < 0 || type > 3) {
    bq::set_error("field: type %d not supported (0..3)", type);
    return -2;
  }
  FieldCtx* F = (FieldCtx*)handle;
  hipStream_t s = bq::stream();
  if (kind < 0 || kind > 2 || (kind == 2 && bd_t)) {
    bq::set_error("field: kind %d not supported here", kind);
    return -2;
  }
  const int nv = kind == 0 ? 13 : (kind == 1 ? 9 : 36);
  double *dp = nullptr, *dout = nullptr, *dze = nullptr;
  int *dd = nullptr, *dt = nullptr;
  BQ_HIP(hipMallocAsync((void**)&dd, sizeof(int) * n, s));
  BQ_HIP(hipMallocAsync((void**)&dout, sizeof(double) * nv * n, s));
  BQ_HIP(hipMemcpyAsync(dd, did, sizeof(int) * n, hipMemcpyHostToDevice, s));
  if (bd_t) {
    BQ_HIP(hipMallocAsync((void**)&dt, sizeof(int) * n, s));
    BQ_HIP(hipMallocAsync((void**)&dze, sizeof(double) * 2 * n, s));
    BQ_HIP(hipMemcpyAsync(dt, bd_t, sizeof(int) * n, hipMemcpyHostToDevice, s));
    BQ_HIP(hipMemcpyAsync(dze, bd_ze, sizeof(double) * 2 * n, hipMemcpyHostToDevice, s));
  } else {
    BQ_HIP(hipMallocAsync((void**)&dp, sizeof(double) * 3 * n, s));
    BQ_HIP(hipMemcpyAsync(dp, pts, sizeof(double) * 3 * n, hipMemcpyHostToDevice, s));
  }
  BQ_HIP(hipMemsetAsync(F->C.plan.err, 0, sizeof(int), s));
  if (bq_field_eval_dev(&F->C, kind, n, dp, dd, dt, dze, type, dout, s)) return -1;
  int err = 0;
  BQ_HIP(hipMemcpyAsync(out, dout, sizeof(double) * nv * n, hipMemcpyDeviceToHost, s));
  BQ_HIP(hipMemcpyAsync(&err, F->C.plan.err, sizeof(int), hipMemcpyDeviceToHost, s));
  for (void* q : {(void*)dp, (void*)dd, (void*)dout, (void*)dt, (void*)dze})
    if (q) BQ_HIP(hipFreeAsync(q, s));
  BQ_HIP(hipStreamSynchronize(s));
  if (err) {
    bq::set_error("field: quasi-periodic Green's function, a reciprocal order has beta_n == 0 (code %d)", err);
    return err;
  }
  return 0;
}

extern "C" int bq_field_domain_ids(void* handle, int n, const double* pts, int* did)
{
  if (n <= 0) return 0;
  FieldCtx* F = (FieldCtx*)handle;
  if (!F->rn) {
    bq::set_error("domain ids: the field context holds no node data");
    return -2;
  }
  hipStream_t s = bq::stream();
  double* dp = nullptr;
  int* dd = nullptr;
  BQ_HIP(hipMallocAsync((void**)&dp, sizeof(double) * 3 * (size_t)n, s));
  BQ_HIP(hipMallocAsync((void**)&dd, sizeof(int) * (size_t)n, s));
  BQ_HIP(hipMemcpyAsync(dp, pts, sizeof(double) * 3 * (size_t)n, hipMemcpyHostToDevice, s));
  BQ_KTIME("k_domain_id", s, k_domain_id<<<n, 256, 0, s>>>(n, dp, F->Ne, F->MN, F->rn, F->ed, F->md, F->sd, dd));
  BQ_HIP(hipGetLastError());
  BQ_HIP(hipMemcpyAsync(did, dd, sizeof(int) * (size_t)n, hipMemcpyDeviceToHost, s));
  BQ_HIP(hipFreeAsync(dp, s));
  BQ_HIP(hipFreeAsync(dd, s));
  BQ_HIP(hipStreamSynchronize(s));
  return 0;
}

extern "C" void bq_field_ctx_destroy(void* handle)
{
  FieldCtx* F = (FieldCtx*)handle;
  if (!F) return;
  (void)hipStreamSynchronize(bq::stream());
  for (void* p : F->owned) (void)hipFree(p);
  delete F;
}
