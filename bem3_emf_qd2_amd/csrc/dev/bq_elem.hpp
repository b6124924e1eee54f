// bq_elem.hpp -- element geometry on the device (bilinear quads and linear triangles).
//
// Element s is stored as 24 doubles: cr[3][4] (r(z,e) = c0 + c1 z + c2 e + c3 z e per
// dimension) and cw[3][3] (normal W(z,e) = w0 + w1 z + w2 e), the reference's element
// constants (d3qd2_src/d3qd2_setup.c:1150-1185), then [21] = shape (0 bilinear quad,
// 1 linear triangle) and 2 pad doubles.  For a triangle the stored cw is the reference's
// (r1-r0)x(r2-r0) times 2/(3 sqrt3), so that |W| is the surface Jacobian of the map from
// the reference triangle for both shapes and every integrand is written once.
// Quad interpolation functions N_n are centred on the +-1/sqrt(3) Gauss points
// (d3qd2_src/d3qd2_elem.c:903-918); triangle ones on the points 0.4 of the circumradius
// towards each vertex (lit_Nn, :1047-1058); collocation node n sits where N_m = delta_nm.
// Reference triangle: vertices (-1/2, -sqrt3/2), (1, 0), (-1/2, sqrt3/2).
#pragma once
#include "bq_complex.hpp"

namespace bq {

constexpr int kElemStride = 24;
constexpr double kP44 = 0.5773502691896257645091;  // 1/sqrt(3)
constexpr double kSQ3 = 1.7320508075688772935274;  // sqrt(3)

// quadrature tables (Gauss-Legendre 12 and 64 on [-1,1]).  Each translation unit that
// includes this header owns a copy (no relocatable device code); the host fills it through
// that unit's set_quad_tables() before launching.
static __constant__ double c_xl[12], c_wl[12], c_xh[64], c_wh[64];
// Gauss-Legendre of orders 3 to 10 (row n - 3, padded to 10) for the distance-adaptive
// regular rule
constexpr int kVruleMin = 3, kVruleMax = 10, kVruleRows = kVruleMax - kVruleMin + 1;
static __constant__ double c_xv[kVruleRows][10], c_wv[kVruleRows][10];

static inline int set_quad_tables(const double* xl, const double* wl, const double* xh, const double* wh)
{
  if (hipMemcpyToSymbol(HIP_SYMBOL(c_xl), xl, 12 * sizeof(double)) != hipSuccess) return -1;
  if (hipMemcpyToSymbol(HIP_SYMBOL(c_wl), wl, 12 * sizeof(double)) != hipSuccess) return -1;
  if (hipMemcpyToSymbol(HIP_SYMBOL(c_xh), xh, 64 * sizeof(double)) != hipSuccess) return -1;
  if (hipMemcpyToSymbol(HIP_SYMBOL(c_wh), wh, 64 * sizeof(double)) != hipSuccess) return -1;
  return 0;
}

// xv, wv: [kVruleRows][10], the rule of order kVruleMin + row in the first entries of each row
static inline int set_vrule_tables(const double* xv, const double* wv)
{
  if (hipMemcpyToSymbol(HIP_SYMBOL(c_xv), xv, kVruleRows * 10 * sizeof(double)) != hipSuccess) return -1;
  if (hipMemcpyToSymbol(HIP_SYMBOL(c_wv), wv, kVruleRows * 10 * sizeof(double)) != hipSuccess) return -1;
  return 0;
}

constexpr int kElemShape = 21;

struct Elem {
  double cr[3][4];
  double cw[3][3];
  bool tri;
};

__device__ __forceinline__ void load_elem(const double* __restrict__ E, int s, Elem& e)
{
  const double* p = E + (size_t)s * kElemStride;
#pragma unroll
  for (int d = 0; d < 3; ++d) {
#pragma unroll
    for (int c = 0; c < 4; ++c) e.cr[d][c] = p[4 * d + c];
#pragma unroll
    for (int c = 0; c < 3; ++c) e.cw[d][c] = p[12 + 3 * d + c];
  }
  e.tri = p[kElemShape] != 0.0;
}

__device__ __forceinline__ bool elem_is_tri(const double* __restrict__ E, int s)
{
  return E[(size_t)s * kElemStride + kElemShape] != 0.0;
}

__device__ __forceinline__ void shapeN4(double z, double t, double N[4])
{
  double zm = 1.0 - z * kSQ3, zp = 1.0 + z * kSQ3, em = 1.0 - t * kSQ3, ep = 1.0 + t * kSQ3;
  N[0] = 0.25 * zm * em;
  N[1] = 0.25 * zp * em;
  N[2] = 0.25 * zp * ep;
  N[3] = 0.25 * zm * ep;
}

// linear-triangle interpolation functions (4th slot zero)
__device__ __forceinline__ void shapeN3(double z, double t, double N[4])
{
  const double ia = 2.5;  // 0.5 / 0.2
  N[0] = (1.0 - ia * (z + kSQ3 * t)) * (1.0 / 3.0);
  N[1] = (1.0 + 2.0 * ia * z) * (1.0 / 3.0);
  N[2] = (1.0 - ia * (z - kSQ3 * t)) * (1.0 / 3.0);
  N[3] = 0.0;
}

__device__ __forceinline__ void shapeN(bool tri, double z, double t, double N[4])
{
  if (tri) shapeN3(z, t, N);
  else shapeN4(z, t, N);
}

// r dN/dr of the triangle functions along (c, s) (lit_dNn_pr_cs, d3qd2_src/lit_dcoef.c:845-856)
__device__ __forceinline__ double tri_dN(int n, double c, double s)
{
  const double ia = 2.5;
  return n == 0 ? -2.0 / 3.0 * ia * (c + kSQ3 * s) : n == 1 ? 4.0 / 3.0 * ia * c : -2.0 / 3.0 * ia * (c - kSQ3 * s);
}

// Gauss-Legendre rule on the reference triangle: three quadrilaterals (edge midpoints and
// centroid) each mapped from (a, b) in [-1,1]^2, Jacobian (sqrt3/16)(1 + a/4 + b/4)
// (d3qd2_src/lit_dcoef_v2.c:213-236).  Point p = 3 (n i + j) + part of 3 n^2.
__device__ __forceinline__ void tri_rule_point(int p, int n, const double* x, const double* wq, double& z, double& e,
                                               double& w)
{
#pragma clang fp contract(off)
  const int ij = p / 3, part = p - 3 * ij, i = ij / n, j = ij - n * i;
  const double a = x[i], b = x[j], ab = a * b;
  w = 0.0625 * kSQ3 * (1.0 + 0.25 * a + 0.25 * b) * wq[i] * wq[j];
  if (part == 0) {
    z = 0.125 * (3.0 + 2.0 * a + 2.0 * b + ab);
    e = 0.125 * kSQ3 * (-a + b);
  } else if (part == 1) {
    z = 0.0625 * (-3.0 + a - 5.0 * b - ab);
    e = 0.0625 * kSQ3 * (3.0 + 3.0 * a + b + ab);
  } else {
    z = 0.0625 * (-3.0 - 5.0 * a + b - ab);
    e = 0.0625 * kSQ3 * (-3.0 - a - 3.0 * b - ab);
  }
}

// point p of the 12-point (gh = false) or 64-point (gh = true) Gauss rule over the element:
// n^2 tensor points on the square, 3 n^2 on the triangle
__device__ __forceinline__ void gauss_point(bool tri, bool gh, int p, double& z, double& e, double& w)
{
  const int n = gh ? 64 : 12;
  const double* x = gh ? c_xh : c_xl;
  const double* wq = gh ? c_wh : c_wl;
  if (tri) {
    tri_rule_point(p, n, x, wq, z, e, w);
  } else {
    const int i = p / n, j = p - n * i;
    z = x[i];
    e = x[j];
    w = wq[i] * wq[j];
  }
}

__device__ __forceinline__ int gauss_points(bool tri, bool gh) { return (gh ? 4096 : 144) * (tri ? 3 : 1); }

// point p of the Gauss rule of order n (3 to 10, or 12) over the element: n^2 tensor points
// on the square, 3 n^2 on the triangle
__device__ __forceinline__ void gauss_point_n(bool tri, int n, int p, double& z, double& e, double& w)
{
  const double* x = (n == 12) ? c_xl : c_xv[n - kVruleMin];
  const double* wq = (n == 12) ? c_wl : c_wv[n - kVruleMin];
  if (tri) {
    tri_rule_point(p, n, x, wq, z, e, w);
  } else {
    const int i = p / n, j = p - n * i;
    z = x[i];
    e = x[j];
    w = wq[i] * wq[j];
  }
}

// Positions are evaluated without fused multiply-adds, in the reference's operation order
// (bil_r_zeta_eta, d3qd2_src/d3qd2_elem.c:949-954): whether R_z comes out exactly 0 decides
// which lattice-sum variant the reference's QPGF takes (bq_qpgf.hpp), and on a plane of
// nearly equal z that can hinge on the last bit.
__device__ __forceinline__ void rw_at(const Elem& e, double z, double t, double r[3], double w[3])
{
#pragma clang fp contract(off)
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    r[d] = e.cr[d][0] + e.cr[d][1] * z + e.cr[d][2] * t + e.cr[d][3] * z * t;
    w[d] = e.cw[d][0] + e.cw[d][1] * z + e.cw[d][2] * t;
  }
}

__device__ __forceinline__ double dot3(const double a[3], const double b[3]) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }

__device__ __forceinline__ double node_z(int n) { return (n == 1 || n == 2) ? kP44 : -kP44; }
__device__ __forceinline__ double node_e(int n) { return (n >= 2) ? kP44 : -kP44; }
// triangle collocation nodes 0..2 and the centroid (d3qd2_src/d3qd2_setup.c:1208-1210)
constexpr double kP34N1 = 0.3464101615137754587055;  // sqrt(3)/5
__device__ __forceinline__ double tri_node_z(int n) { return n == 1 ? 0.4 : n == 3 ? 0.0 : -0.2; }
__device__ __forceinline__ double tri_node_e(int n) { return n == 0 ? -kP34N1 : n == 2 ? kP34N1 : 0.0; }
__device__ __forceinline__ double elem_node_z(bool tri, int n) { return tri ? tri_node_z(n) : node_z(n); }
__device__ __forceinline__ double elem_node_e(bool tri, int n) { return tri ? tri_node_e(n) : node_e(n); }

// unit tangents at collocation node n (d3qd2_src/d3qd2_elem.c:122-143); a triangle has
// c3 = 0, so its tangents are c1 and c2 at every node
__device__ __forceinline__ void node_tangents(const Elem& e, int n, double vz[3], double ve[3])
{
  double zt = elem_node_z(e.tri, n), et = elem_node_e(e.tri, n);
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    vz[d] = e.cr[d][1] + e.cr[d][3] * et;
    ve[d] = e.cr[d][2] + e.cr[d][3] * zt;
  }
  double a = 1.0 / sqrt(dot3(vz, vz)), b = 1.0 / sqrt(dot3(ve, ve));
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    vz[d] *= a;
    ve[d] *= b;
  }
}

// the planarity test of the reference (d3qd2_src/d3qd2_elem.c:987-997); triangles are flat
__device__ __forceinline__ bool elem_plane(const Elem& e)
{
  if (e.tri) return true;
  double b[3] = {e.cr[0][1], e.cr[1][1], e.cr[2][1]}, c[3] = {e.cw[0][2], e.cw[1][2], e.cw[2][2]};
  return fabs(dot3(b, c) / (sqrt(dot3(b, b)) * sqrt(dot3(c, c)))) < 1.0e-15;
}

__device__ __forceinline__ double elem_bdC(const Elem& e)
{
  return e.cr[0][1] * e.cw[0][2] + e.cr[1][1] * e.cw[1][2] + e.cr[2][1] * e.cw[2][2];
}

// corners of the parameter square seen from (zt, et): distances and angles
// (d3qd2_src/d3qd2_elem.c:1013-1027)
__device__ __forceinline__ void corner_r_th4(double zt, double et, double r[5], double th[5])
{
  double ztp = 1.0 - zt, ztm = -1.0 - zt, etp = 1.0 - et, etm = -1.0 - et;
  r[0] = sqrt(ztm * ztm + etm * etm);
  r[1] = sqrt(ztp * ztp + etm * etm);
  r[2] = sqrt(ztp * ztp + etp * etp);
  r[3] = sqrt(ztm * ztm + etp * etp);
  r[4] = r[0];
  th[0] = atan2(etm, ztm);
  th[1] = atan2(etm, ztp);
  th[2] = atan2(etp, ztp);
  th[3] = atan2(etp, ztm);
  th[4] = th[0] + 6.283185307179586476925286766559;
}

// corners of the reference triangle seen from (zt, et) (d3qd2_src/d3qd2_elem.c:1149-1162)
__device__ __forceinline__ void corner_r_th3(double zt, double et, double r[5], double th[5])
{
  double ztp = 1.0 - zt, ztm = -0.5 - zt, etp = 0.5 * kSQ3 - et, etm = -0.5 * kSQ3 - et;
  r[0] = sqrt(ztm * ztm + etm * etm);
  r[1] = sqrt(ztp * ztp + et * et);
  r[2] = sqrt(ztm * ztm + etp * etp);
  r[3] = r[0];
  r[4] = r[0];
  th[0] = atan2(etm, ztm);
  th[1] = atan2(-et, ztp);
  th[2] = atan2(etp, ztm);
  th[3] = th[0] + 6.283185307179586476925286766559;
  th[4] = th[3];
}

// polar sectors around (zt, et): 4 for the square, 3 for the triangle
__device__ __forceinline__ int corner_r_th(bool tri, double zt, double et, double r[5], double th[5])
{
  if (tri) {
    corner_r_th3(zt, et, r, th);
    return 3;
  }
  corner_r_th4(zt, et, r, th);
  return 4;
}

// polar direction vector K(r, theta) with R = r K (exact for the bilinear map)
__device__ __forceinline__ void polar_K(const Elem& e, double zt, double et, double r, double c, double s, double K[3])
{
#pragma clang fp contract(off)
  double q = r * c * s + zt * s + et * c;
#pragma unroll
  for (int d = 0; d < 3; ++d) K[d] = e.cr[d][1] * c + e.cr[d][2] * s + e.cr[d][3] * q;
}

// theta node m (0..127) of sub-triangle j: two Gauss-64 halves of [th_j, th_{j+1}]
__device__ __forceinline__ void polar_theta(const double th[5], int j, int m, double& theta, double& wt)
{
  int h = m >> 6, mm = m & 63;
  double t0 = th[j], t1 = th[j + 1];
  theta = 0.25 * (t1 - t0) * c_xh[mm] + (h == 0 ? 0.25 * (t1 + 3.0 * t0) : 0.25 * (3.0 * t1 + t0));
  wt = 0.25 * (t1 - t0) * c_wh[mm];
}

// distance to the element edge along theta inside sub-triangle j
__device__ __forceinline__ double polar_rm(const double r[5], const double th[5], int j, double theta)
{
  double r0 = r[j], r1 = r[j + 1], t0 = th[j], t1 = th[j + 1];
  return r0 * r1 * sin(t1 - t0) / (r0 * sin(theta - t0) - r1 * sin(theta - t1));
}

}  // namespace bq
