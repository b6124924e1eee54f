// bq_elem.hpp -- bilinear-element geometry on the device.
//
// Element s is stored as 24 doubles: cr[3][4] (r(z,e) = c0 + c1 z + c2 e + c3 z e per
// dimension) and cw[3][3] (normal W(z,e) = w0 + w1 z + w2 e), the reference's element
// constants (d3qd2_src/d3qd2_setup.c:1150-1170), followed by 3 pad doubles.
// Interpolation functions N_n are centred on the +-1/sqrt(3) Gauss points
// (d3qd2_src/d3qd2_elem.c:903-918), so collocation node n sits where N_m = delta_nm.
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

static inline int set_quad_tables(const double* xl, const double* wl, const double* xh, const double* wh)
{
  if (hipMemcpyToSymbol(HIP_SYMBOL(c_xl), xl, 12 * sizeof(double)) != hipSuccess) return -1;
  if (hipMemcpyToSymbol(HIP_SYMBOL(c_wl), wl, 12 * sizeof(double)) != hipSuccess) return -1;
  if (hipMemcpyToSymbol(HIP_SYMBOL(c_xh), xh, 64 * sizeof(double)) != hipSuccess) return -1;
  if (hipMemcpyToSymbol(HIP_SYMBOL(c_wh), wh, 64 * sizeof(double)) != hipSuccess) return -1;
  return 0;
}

struct Elem {
  double cr[3][4];
  double cw[3][3];
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
}

__device__ __forceinline__ double shapeN(int n, double z, double t)
{
  double a = (n == 1 || n == 2) ? 1.0 : -1.0;
  double b = (n >= 2) ? 1.0 : -1.0;
  return 0.25 * (1.0 + a * z * kSQ3) * (1.0 + b * t * kSQ3);
}

__device__ __forceinline__ void shapeN4(double z, double t, double N[4])
{
  double zm = 1.0 - z * kSQ3, zp = 1.0 + z * kSQ3, em = 1.0 - t * kSQ3, ep = 1.0 + t * kSQ3;
  N[0] = 0.25 * zm * em;
  N[1] = 0.25 * zp * em;
  N[2] = 0.25 * zp * ep;
  N[3] = 0.25 * zm * ep;
}

__device__ __forceinline__ void rw_at(const Elem& e, double z, double t, double r[3], double w[3])
{
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    r[d] = e.cr[d][0] + e.cr[d][1] * z + e.cr[d][2] * t + e.cr[d][3] * z * t;
    w[d] = e.cw[d][0] + e.cw[d][1] * z + e.cw[d][2] * t;
  }
}

__device__ __forceinline__ double dot3(const double a[3], const double b[3]) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }

__device__ __forceinline__ double node_z(int n) { return (n == 1 || n == 2) ? kP44 : -kP44; }
__device__ __forceinline__ double node_e(int n) { return (n >= 2) ? kP44 : -kP44; }

// unit tangents at collocation node n (d3qd2_src/d3qd2_elem.c:122-143)
__device__ __forceinline__ void node_tangents(const Elem& e, int n, double vz[3], double ve[3])
{
  double zt = node_z(n), et = node_e(n);
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

// the planarity test of the reference (d3qd2_src/d3qd2_elem.c:987-997)
__device__ __forceinline__ bool elem_plane(const Elem& e)
{
  double b[3] = {e.cr[0][1], e.cr[1][1], e.cr[2][1]}, c[3] = {e.cw[0][2], e.cw[1][2], e.cw[2][2]};
  return fabs(dot3(b, c) / (sqrt(dot3(b, b)) * sqrt(dot3(c, c)))) < 1.0e-15;
}

__device__ __forceinline__ double elem_bdC(const Elem& e)
{
  return e.cr[0][1] * e.cw[0][2] + e.cr[1][1] * e.cw[1][2] + e.cr[2][1] * e.cw[2][2];
}

// corners of the parameter square seen from (zt, et): distances and angles
// (d3qd2_src/d3qd2_elem.c:1013-1027)
__device__ __forceinline__ void corner_r_th(double zt, double et, double r[5], double th[5])
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

// polar direction vector K(r, theta) with R = r K (exact for the bilinear map)
__device__ __forceinline__ void polar_K(const Elem& e, double zt, double et, double r, double c, double s, double K[3])
{
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
