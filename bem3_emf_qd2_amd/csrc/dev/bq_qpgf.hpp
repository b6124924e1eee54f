// bq_qpgf.hpp -- quasi-periodic Green's function (value + gradient) for one point, FP64.
//
// G(r) = sum_l exp(i k_t.R_l) exp(ik|r+R_l|) / (4 pi |r+R_l|) over the 2-D lattice R_l,
// evaluated like the reference (qgf2_src/d3_qpgf_d2.c:124-176, :290-438, :745-885,
// :1671-2009): the point is folded into the unit cell, then either the Fourier-domain
// (spectral) sum is used when every order within S_LIMIT of the nearest one is evanescent
// enough over |z|, or the Ewald split into a spatial sum (Faddeeva w terms) and a spectral
// sum (w for propagating, erfc for evanescent orders), with the reference's per-point
// Ewald parameter rule.
//
// GPU form: instead of the reference's adaptive stop tests (which make every lane run a
// different, data-dependent loop), each sum runs over the lattice points inside a disk
// whose radius follows from the terms' Gaussian decay, exp(-BQ_EXP_CUT) = 4e-18 of the
// leading term, row by row with the exact row bounds of the disk.  Evanescent terms use the
// overflow-free erfcx form exp(-b^2 v^2 - z^2/4v^2) erfcx(.) instead of exp(bz) erfc(.).
#pragma once
#include "bq_complex.hpp"
#include "bq_faddeeva.hpp"
#include "../bq_plan.h"

namespace bq {

constexpr double kPI = 3.14159265358979323846264338327950288;
constexpr double kI4PI = 0.0795774715459476678844418816862571882;  // 1/(4 pi)
constexpr double kSQPI = 1.7724538509055160272981674833411451828;  // sqrt(pi)

struct QG {
  cplx g;
  cplx d[3];
};

// row bounds of { t : |u + t e|^2 <= R2 } for integer t
__device__ __forceinline__ bool row_range(double ux, double uy, double ex, double ey, double e2, double R2,
                                          int& lo, int& hi)
{
  double ue = ux * ex + uy * ey, uu = ux * ux + uy * uy;
  double disc = ue * ue - e2 * (uu - R2);
  if (disc < 0.0) return false;
  double sq = sqrt(disc), ie2 = 1.0 / e2;
  lo = (int)ceil((-ue - sq) * ie2);
  hi = (int)floor((-ue + sq) * ie2);
  return lo <= hi;
}

// Fourier-domain sum for the folded point (px, py, pz), pz != 0.
__device__ void qpgf_fourier(const bq_qplan& P, double px, double py, double pz, QG& o)
{
  double az = fabs(pz), bcut = BQ_EXP_CUT / az;
  double A2 = P.k2 + bcut * bcut;
  double A = sqrt(A2) * 1.0000001;
  A2 = A * A;
  double d1n = sqrt(P.d1x * P.d1x + P.d1y * P.d1y);
  double c1 = (P.vk0 * P.d1x + P.vk1 * P.d1y) / (2.0 * kPI), w1 = A * d1n / (2.0 * kPI);
  int n1lo = (int)ceil(c1 - w1), n1hi = (int)floor(c1 + w1);
  double g22 = P.g2x * P.g2x + P.g2y * P.g2y;
  cplx S = mk(0, 0), Sx = mk(0, 0), Sy = mk(0, 0), Sz = mk(0, 0);
  for (int n1 = n1lo; n1 <= n1hi; ++n1) {
    double ux = -P.vk0 + n1 * P.g1x, uy = -P.vk1 + n1 * P.g1y;
    int lo, hi;
    if (!row_range(ux, uy, P.g2x, P.g2y, g22, A2, lo, hi)) continue;
    for (int n2 = lo; n2 <= hi; ++n2) {
      double ax = ux + n2 * P.g2x, ay = uy + n2 * P.g2y;
      double b2 = P.k2 - ax * ax - ay * ay;
      cplx dz, F;
      if (b2 > 0.0) {
        double b = sqrt(b2);
        dz = cexpi(ax * px + ay * py + b * az);
        F = dz * (1.0 / b);
      } else {
        double b = sqrt(-b2);
        dz = cexpi(ax * px + ay * py) * exp(-b * az);
        F = mk(dz.im, -dz.re) * (1.0 / b);  // -i dz / b
      }
      S += F;
      Sx += F * ax;
      Sy += F * ay;
      Sz += dz;
    }
  }
  double h = 0.5 * P.aidd, sg = (pz > 0.0) ? 1.0 : -1.0;
  o.g = cmul_i(S) * h;
  o.d[0] = Sx * (-h);
  o.d[1] = Sy * (-h);
  o.d[2] = Sz * (-h * sg);
}

// Ewald split for the folded point.
__device__ void qpgf_ewald(const bq_qplan& P, double px, double py, double pz, QG& o)
{
  double r2 = px * px + py * py + pz * pz;
  double ve = P.veps0;
  if (P.k2 * ve * ve - r2 / (4.0 * ve * ve) > BQ_EW_EXPM)
    ve = sqrt(BQ_EW_EXPM + sqrt(P.k2 * r2 + BQ_EW_EXPM * BQ_EW_EXPM)) / (1.4142135623730950488 * P.k);
  const double ive = 1.0 / ve, ke = P.k * ve, ke2 = ke * ke;

  // ---- spatial sum: |r + R_l| <= 2 ve sqrt(ke^2 + cut)
  {
    double Rmax = 2.0 * ve * sqrt(ke2 + BQ_EXP_CUT) * 1.0000001, R2 = Rmax * Rmax;
    double b1n = sqrt(P.b1x * P.b1x + P.b1y * P.b1y);
    double c1 = -(P.b1x * px + P.b1y * py), w1 = Rmax * b1n;
    int l1lo = (int)ceil(c1 - w1), l1hi = (int)floor(c1 + w1);
    double d22 = P.d2x * P.d2x + P.d2y * P.d2y;
    const double gcon = ive / kSQPI;
    cplx S = mk(0, 0), Sxy[2] = {mk(0, 0), mk(0, 0)}, Sz = mk(0, 0);
    for (int l1 = l1lo; l1 <= l1hi; ++l1) {
      double ux = px + l1 * P.d1x, uy = py + l1 * P.d1y;
      int lo, hi;
      if (!row_range(ux, uy, P.d2x, P.d2y, d22, R2 - pz * pz, lo, hi)) continue;
      for (int l2 = lo; l2 <= hi; ++l2) {
        double Rx = l1 * P.d1x + l2 * P.d2x, Ry = l1 * P.d1y + l2 * P.d2y;
        double ax = px + Rx, ay = py + Ry;
        double rl = sqrt(ax * ax + ay * ay + pz * pz), irl = 1.0 / rl;
        double y = 0.5 * rl * ive;
        cplx cef = cexpi(P.vk0 * Rx + P.vk1 * Ry) * exp(ke2 - y * y);
        cplx w = faddeeva_w(ke, y);
        S += cef * (irl * w.re);
        cplx cf = cef * (irl * irl * (irl * w.re - P.k * w.im + gcon));
        Sxy[0] += cf * ax;
        Sxy[1] += cf * ay;
        Sz += cf;
      }
    }
    o.g = S * kI4PI;
    o.d[0] = Sxy[0] * (-kI4PI);
    o.d[1] = Sxy[1] * (-kI4PI);
    o.d[2] = Sz * (-pz * kI4PI);
  }
  // ---- spectral sum: |beta| ve <= sqrt(cut)
  {
    double A2 = P.k2 + BQ_EXP_CUT * ive * ive;
    double A = sqrt(A2) * 1.0000001;
    A2 = A * A;
    double d1n = sqrt(P.d1x * P.d1x + P.d1y * P.d1y);
    double c1 = (P.vk0 * P.d1x + P.vk1 * P.d1y) / (2.0 * kPI), w1 = A * d1n / (2.0 * kPI);
    int n1lo = (int)ceil(c1 - w1), n1hi = (int)floor(c1 + w1);
    double g22 = P.g2x * P.g2x + P.g2y * P.g2y;
    double az = 0.5 * pz * ive;
    cplx S = mk(0, 0), Sx = mk(0, 0), Sy = mk(0, 0), Sz = mk(0, 0);
    for (int n1 = n1lo; n1 <= n1hi; ++n1) {
      double ux = -P.vk0 + n1 * P.g1x, uy = -P.vk1 + n1 * P.g1y;
      int lo, hi;
      if (!row_range(ux, uy, P.g2x, P.g2y, g22, A2, lo, hi)) continue;
      for (int n2 = lo; n2 <= hi; ++n2) {
        double ax = ux + n2 * P.g2x, ay = uy + n2 * P.g2y;
        double b2 = P.k2 - ax * ax - ay * ay;
        cplx ce = cexpi(ax * px + ay * py);
        cplx F, Gz;
        if (b2 > 0.0) {
          double b = sqrt(b2), ef = exp(b2 * ve * ve - az * az);
          cplx ff = cexpi(b * pz);
          cplx w = faddeeva_w(b * ve, az);
          // F = 2 ce / b (i ef Im w + ff);  Gz = 2 ce (ef Re w - ff)
          F = ce * (mk(ff.re, ef * w.im + ff.im) * (2.0 / b));
          Gz = ce * (mk(ef * w.re - ff.re, -ff.im) * 2.0);
        } else {
          double b = sqrt(-b2), bv = b * ve;
          double a1 = bv + az, a2 = bv - az;
          double e12 = exp(-bv * bv - az * az);
          double f1 = (a1 > 0.0) ? e12 * erfcx(a1) : exp(b * pz) * erfc(a1);
          double f2 = (a2 > 0.0) ? e12 * erfcx(a2) : exp(-b * pz) * erfc(a2);
          // F = ce / (i b) (f1 + f2);  Gz = ce (f1 - f2)
          F = mk(ce.im, -ce.re) * ((f1 + f2) / b);
          Gz = ce * (f1 - f2);
        }
        S += F;
        Sx += F * ax;
        Sy += F * ay;
        Sz += Gz;
      }
    }
    double cd = 0.25 * P.aidd;
    o.g += cmul_i(S) * cd;
    o.d[0] += Sx * (-cd);
    o.d[1] += Sy * (-cd);
    o.d[2] += Sz * cd;
  }
}

// Value and gradient at r (relative position, source minus field point).
__device__ void qpgf_v1(const bq_qplan& P, double rx, double ry, double rz, QG& o)
{
  double l1 = floor((rx * P.d2y - ry * P.d2x) * P.idd + 0.5);
  double l2 = floor(-(rx * P.d1y - ry * P.d1x) * P.idd + 0.5);
  double px = rx - l1 * P.d1x - l2 * P.d2x, py = ry - l1 * P.d1y - l2 * P.d2y, pz = rz;
  bool fd = false;
  if (P.fd_allowed) {
    double az = fabs(rz);
    fd = (P.fd_bN / P.fd_bNp * exp(-P.fd_bNp * az) < P.eps) && (P.fd_bN / P.fd_bNm * exp(-P.fd_bNm * az) < P.eps);
  }
  if (fd) qpgf_fourier(P, px, py, pz, o);
  else qpgf_ewald(P, px, py, pz, o);
  if (l1 != 0.0 || l2 != 0.0) {
    cplx cf = cexpi(-(P.vk0 * (l1 * P.d1x + l2 * P.d2x) + P.vk1 * (l1 * P.d1y + l2 * P.d2y)));
    o.g = o.g * cf;
    for (int i = 0; i < 3; ++i) o.d[i] = o.d[i] * cf;
  }
}

}  // namespace bq
