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
// whose radius follows from the terms' Gaussian decay, exp(-P.cut) of the
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
// value, gradient and Hessian (xx, yy, zz, xy, yz, zx), the _v2 output
struct QG2 {
  cplx g;
  cplx d[3];
  cplx h[6];
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

// Visit the integers lo..hi of a lattice row starting at c (lo <= c <= hi) and walking out
// both ways, with E = exp(i phase(n)) carried by multiplication with step = exp(i dphase)
// from E0 = exp(i phase(c)); starting at the row's centre keeps the carried chains short
// and their rounding smallest where the terms are largest.
template <class Fn>
__device__ __forceinline__ void row_walk(int lo, int hi, int c, cplx E0, cplx step, Fn&& f)
{
  cplx E = E0;
  for (int n = c; n <= hi; ++n) {
    f(n, E);
    E = E * step;
  }
  const cplx back = conj(step);
  E = E0 * back;
  for (int n = c - 1; n >= lo; --n) {
    f(n, E);
    E = E * back;
  }
}

// A reciprocal order with beta_n = 0 has no finite term: the reference returns -12 / -23 for it
// (d3_qpgf_d2.c, Fourier-domain and Ewald spectral sums).  Here the term is taken as 0 and the
// code stored in the plan's error word, which the host checks after the launch.
__device__ __forceinline__ void qpgf_flag(const bq_qplan& P, int code)
{
  if (P.err) atomicMin(P.err, code);
}

__device__ __forceinline__ int row_centre(double ue, double e2, int lo, int hi)
{
  const int c = (int)rint(-ue / e2);
  return c < lo ? lo : (c > hi ? hi : c);
}
// the same from 1 / e2 (any start within lo .. hi gives the same row sum)
__device__ __forceinline__ int row_centre_r(double ue, double ie2, int lo, int hi)
{
  const int c = (int)rint(-ue * ie2);
  return c < lo ? lo : (c > hi ? hi : c);
}

// Fourier-domain sum for the folded point (px, py, pz), pz != 0.  With HESS (Q = QG2) the
// second derivatives are summed too: each order is exp(i a.p + i b |z|) / b, so d2/dxdy brings
// -a_x a_y, d2/dz2 brings -b^2 and d2/dx dz brings i a_x (i b sgn z).
template <bool HESS, class Q>
__device__ void qpgf_fourier_t(const bq_qplan& P, double px, double py, double pz, Q& o)
{
  double az = fabs(pz), bcut = P.cut / az;
  double A2 = P.k2 + bcut * bcut;
  double A = sqrt(A2) * 1.0000001;
  A2 = A * A;
  double d1n = sqrt(P.d1x * P.d1x + P.d1y * P.d1y);
  double c1 = (P.vk0 * P.d1x + P.vk1 * P.d1y) / (2.0 * kPI), w1 = A * d1n / (2.0 * kPI);
  int n1lo = (int)ceil(c1 - w1), n1hi = (int)floor(c1 + w1);
  const double g22 = P.g2x * P.g2x + P.g2y * P.g2y;
  const double gp1 = P.g1x * px + P.g1y * py, gp2 = P.g2x * px + P.g2y * py, ap = -(P.vk0 * px + P.vk1 * py);
  const cplx step = cexpi(gp2);
  cplx S = mk(0, 0), Sx = mk(0, 0), Sy = mk(0, 0), Sz = mk(0, 0);
  cplx Hxx = mk(0, 0), Hyy = mk(0, 0), Hzz = mk(0, 0), Hxy = mk(0, 0), Hxz = mk(0, 0), Hyz = mk(0, 0);
  for (int n1 = n1lo; n1 <= n1hi; ++n1) {
    const double ux = -P.vk0 + n1 * P.g1x, uy = -P.vk1 + n1 * P.g1y;
    int lo, hi;
    if (!row_range(ux, uy, P.g2x, P.g2y, g22, A2, lo, hi)) continue;
    const int c = row_centre(ux * P.g2x + uy * P.g2y, g22, lo, hi);
    row_walk(lo, hi, c, cexpi(ap + n1 * gp1 + c * gp2), step, [&](int n2, cplx E) {
      const double ax = ux + n2 * P.g2x, ay = uy + n2 * P.g2y;
      const double b2 = P.k2 - ax * ax - ay * ay;
      cplx dz, F;
      if (b2 > 0.0) {
        const double b = sqrt(b2);
        dz = E * cexpi(b * az);
        F = dz * (1.0 / b);
      } else if (b2 < 0.0) {
        const double b = sqrt(-b2);
        dz = E * exp(-b * az);
        F = mk(dz.im, -dz.re) * (1.0 / b);  // -i dz / b
      } else {
        qpgf_flag(P, -12);
        dz = F = mk(0.0, 0.0);
      }
      S += F;
      Sx += F * ax;
      Sy += F * ay;
      Sz += dz;
      if constexpr (HESS) {
        Hxx += F * (ax * ax);
        Hyy += F * (ay * ay);
        Hxy += F * (ax * ay);
        Hzz += F * b2;
        Hxz += dz * ax;
        Hyz += dz * ay;
      }
    });
  }
  double h = 0.5 * P.aidd, sg = (pz > 0.0) ? 1.0 : -1.0;
  o.g = cmul_i(S) * h;
  o.d[0] = Sx * (-h);
  o.d[1] = Sy * (-h);
  o.d[2] = Sz * (-h * sg);
  if constexpr (HESS) {
    o.h[0] = cmul_i(Hxx) * (-h);
    o.h[1] = cmul_i(Hyy) * (-h);
    o.h[2] = cmul_i(Hzz) * (-h);
    o.h[3] = cmul_i(Hxy) * (-h);
    o.h[4] = cmul_i(Hyz) * (-h * sg);
    o.h[5] = cmul_i(Hxz) * (-h * sg);
  }
}
__device__ __forceinline__ void qpgf_fourier(const bq_qplan& P, double px, double py, double pz, QG& o)
{
  qpgf_fourier_t<false>(P, px, py, pz, o);
}

// one Ewald spectral term for the reciprocal vector (ax, ay) with ce = exp(i a . p):
// F (value) and Gz (z part, i dF/dz)
// and g0 = exp(b^2 v^2 - z^2 / 4v^2), which the second z derivative needs:
// dGz/dz = -i b^2 F - 2 ce g0 / (v sqrt(pi)).
__device__ __forceinline__ void spectral_term(const bq_qplan& P, double ax, double ay, cplx ce, double pz, double ve,
                                              cplx& F, cplx& Gz, double& g0)
{
  double b2 = P.k2 - ax * ax - ay * ay;
  double az = 0.5 * pz / ve;
  if (b2 > 0.0) {
    double b = sqrt(b2), ef = exp(b2 * ve * ve - az * az);
    g0 = ef;
    cplx ff = cexpi(b * pz);
    cplx w = faddeeva_w(b * ve, az);
    // F = 2 ce / b (i ef Im w + ff);  Gz = 2 ce (ef Re w - ff)
    F = ce * (mk(ff.re, ef * w.im + ff.im) * (2.0 / b));
    Gz = ce * (mk(ef * w.re - ff.re, -ff.im) * 2.0);
  } else if (b2 == 0.0) {
    qpgf_flag(P, -23);
    F = Gz = mk(0.0, 0.0);
    g0 = 0.0;
  } else {
    double b = sqrt(-b2), bv = b * ve;
    double a1 = bv + az, a2 = bv - az;
    double e12 = exp(-bv * bv - az * az);
    g0 = e12;
    double f1 = (a1 > 0.0) ? e12 * erfcx(a1) : exp(b * pz) * erfc(a1);
    double f2 = (a2 > 0.0) ? e12 * erfcx(a2) : exp(-b * pz) * erfc(a2);
    // F = ce / (i b) (f1 + f2);  Gz = ce (f1 - f2)
    F = mk(ce.im, -ce.re) * ((f1 + f2) / b);
    Gz = ce * (f1 - f2);
  }
}

// ---- exactly rounded building blocks for the r_z == 0 path -------------------------------
// exp(t) rounded to nearest (barring results within 1e-30 relative of a rounding midpoint):
// t = k ln2 + r with ln2 split in two doubles, then exp(r) by a 22-term Taylor-Horner
// evaluation in double-double arithmetic (Dekker / Knuth two-sum and fma two-product).
__device__ __forceinline__ void dd_two_sum(double a, double b, double& s, double& e)
{
#pragma clang fp contract(off)
  s = a + b;
  const double bb = s - a;
  e = (a - (s - bb)) + (b - bb);
}
__device__ double exp_rn(double t)
{
#pragma clang fp contract(off)
  if (!(fabs(t) < 700.0)) return exp(t);
  const double L1 = 0x1.62e42fefa39efp-1, L2 = 0x1.abc9e3b39803fp-56;  // ln 2 = L1 + L2
  const double k = rint(t * 0x1.71547652b82fep0);
  const double ph = k * L1, pl = fma(k, L1, -ph);  // k L1 = ph + pl exactly
  double rh, rl;
  dd_two_sum(t - ph, -(pl + k * L2), rh, rl);      // t - ph is exact (Sterbenz)
  double sh = 1.0, sl = 0.0;
  for (int n = 22; n >= 1; --n) {
    // (sh, sl) <- 1 + r (sh, sl) / n
    double ph2 = rh * sh, pe = fma(rh, sh, -ph2);
    pe += rh * sl + rl * sh;
    double mh, ml;
    dd_two_sum(ph2, pe, mh, ml);
    const double q1 = mh / n;
    const double q2 = (fma(-q1, (double)n, mh) + ml) / n;
    double th, tl;
    dd_two_sum(1.0, q1, th, tl);
    tl += q2;
    dd_two_sum(th, tl, sh, sl);
  }
  return ldexp(sh, (int)k);
}

// Folding of r into the unit cell and the per-point Ewald parameter, with the reference's
// operation order and no contraction (d3_qpgf_d2.c:417-423, :1680-1693), so the r_z == 0
// emulation below sees bit-identical inputs.
__device__ __forceinline__ void fold_cell(const bq_qplan& P, double rx, double ry, double& l1, double& l2,
                                          double& px, double& py)
{
#pragma clang fp contract(off)
  l1 = floor((rx * P.d2y - ry * P.d2x) * P.idd + 0.5);
  l2 = floor(-(rx * P.d1y - ry * P.d1x) * P.idd + 0.5);
  px = rx - l1 * P.d1x - l2 * P.d2x;
  py = ry - l1 * P.d1y - l2 * P.d2y;
}
__device__ __forceinline__ double ewald_param(const bq_qplan& P, double px, double py, double pz)
{
#pragma clang fp contract(off)
  const double r2 = px * px + py * py + pz * pz;
  double ve = P.veps0;
  if (P.k2 * ve * ve - r2 / (4.0 * ve * ve) > BQ_EW_EXPM)
    ve = sqrt(BQ_EW_EXPM + sqrt(P.k2 * r2 + BQ_EW_EXPM * BQ_EW_EXPM)) / (1.4142135623730950488 * P.k);
  return ve;
}

// One Ewald spectral term at z = 0, evaluated the way the reference does it
// (d3_qpgf_d2.c:1965-2009): products rounded before the subtraction (no fused multiply-add),
// so that dqG_z of a propagating order, 2 (e^{b^2 v^2} Re w(b v) - 1), is the same rounding
// residual (often exactly zero) that the reference's stop test sees.  exp is rounded to
// nearest (exp_rn) as a correctly rounding libm does; a libm off by an ulp at these
// arguments would make the reference itself stop elsewhere.
__device__ __forceinline__ void spectral_term_z0(const bq_qplan& P, double ax, double ay, double px, double py,
                                                 double ve, bool prop, cplx& F, cplx& Gz)
{
#pragma clang fp contract(off)
  const double b2 = P.k2 - ax * ax - ay * ay;
  const cplx ce = cexpi(ax * px + ay * py);
  if (b2 == 0.0) {
    qpgf_flag(P, -23);
    F = Gz = mk(0.0, 0.0);
  } else if (prop) {
    const double b = sqrt(b2), ef = exp_rn(b2 * ve * ve);
    const double x = b * ve;
    const double wre = exp_rn(-(x * x)), wim = faddeeva_w(x, 0.0).im;
    const double pr = ef * wre;
    F = ce * (mk(1.0, ef * wim) * (2.0 / b));
    Gz = ce * mk(2.0 * (pr - 1.0), 0.0);
  } else {
    const double b = sqrt(-b2), f = erfc(b * ve);
    F = mk(ce.im, -ce.re) * ((f + f) / b);
    Gz = mk(0.0, 0.0);
  }
}

// Spectral sum over exactly the terms the reference's _v1/_dqG variants add when r_z == 0
// (qgf2_src/d3_qpgf_d2.c:1840-1963): their stop test compares dqG_z, which is identically
// zero there for evanescent orders (and zero up to rounding for propagating ones), so each
// row stops at its first evanescent term past the centre (both directions), and the rows
// stop after the first row beyond the centre whose non-centre terms are all evanescent and
// whose dqG_z partial sum compares equal to the one before the row.  Kept for result parity
// with the reference.
__device__ void spectral_z0_ref(const bq_qplan& P, double px, double py, double ve, QG& o)
{
#pragma clang fp contract(off)
  const double g22 = P.g2x * P.g2x + P.g2y * P.g2y;
  const int nc = (int)round((P.vk0 * P.d1x + P.vk1 * P.d1y) / (2.0 * kPI));
  cplx S = mk(0, 0), Sx = mk(0, 0), Sy = mk(0, 0), Sz = mk(0, 0);
  for (int dir = 0; dir < 3; ++dir) {
    int n1 = (dir == 0) ? nc : (dir == 1 ? nc + 1 : nc - 1);
    const int step = (dir == 2) ? -1 : 1;
    for (int it = 0; it < 99; ++it, n1 += step) {
      const cplx Sz0 = Sz;
      const double ux = -P.vk0 + n1 * P.g1x, uy = -P.vk1 + n1 * P.g1y;
      const int n2c = (int)round(-(ux * P.g2x + uy * P.g2y) / g22);
      int nprop = 0;
      for (int side = 0; side < 3; ++side) {
        int n2 = (side == 0) ? n2c : (side == 1 ? n2c + 1 : n2c - 1);
        const int s2 = (side == 2) ? -1 : 1;
        for (int m = 0; m < 99; ++m, n2 += s2) {
          const double ax = ux + n2 * P.g2x, ay = uy + n2 * P.g2y;
          const bool prop = P.k2 - ax * ax - ay * ay > 0.0;
          cplx F, Gz;
          spectral_term_z0(P, ax, ay, px, py, ve, prop, F, Gz);
          S += F;
          Sx += F * ax;
          Sy += F * ay;
          Sz += Gz;
          if (side == 0) break;
          if (!prop) break;  // first evanescent term ends the side
          ++nprop;
        }
      }
      if (dir == 0) break;
      if (nprop == 0 && 2.0 * fabs(Sz.re - Sz0.re) <= P.eps * (fabs(Sz.re) + fabs(Sz0.re)) &&
          2.0 * fabs(Sz.im - Sz0.im) <= P.eps * (fabs(Sz.im) + fabs(Sz0.im)))
        break;
    }
  }
  const double cd = 0.25 * P.aidd;
  o.g += cmul_i(S) * cd;
  o.d[0] += Sx * (-cd);
  o.d[1] += Sy * (-cd);
}

// Ewald split for the folded point.  With HESS (Q = QG2) the second derivatives are summed
// too; z0_ref applies to the first-derivative form only.
//   spatial term f(rho) = ef Re w / rho (rho = |p + R_l|): grad = q a, Hessian = q I + a a^T q'/rho
//     with q = -ef (Re w / rho - k Im w + 1/(v sqrt(pi))) / rho^2 and
//     q'/rho = -k^2 ef Re w / rho^3 + ef / (2 v^3 sqrt(pi) rho^2) - 3 q / rho^2;
//   spectral term i F e^{i a.p}: d2/dxdy -> -a_x a_y, d2/dx dz -> i a_x Gz,
//     d2/dz2 -> -i b^2 F - 2 ce g0 / (v sqrt(pi)).
template <bool HESS, class Q>
__device__ void qpgf_ewald_t(const bq_qplan& P, double px, double py, double pz, double ve, Q& o, bool z0_ref)
{
  const double ive = 1.0 / ve, ke = P.k * ve, ke2 = ke * ke;

  // ---- spatial sum: |r + R_l| <= 2 ve sqrt(ke^2 + cut)
  {
    double Rmax = 2.0 * ve * sqrt(ke2 + P.cut) * 1.0000001, R2 = Rmax * Rmax;
    double b1n = sqrt(P.b1x * P.b1x + P.b1y * P.b1y);
    double c1 = -(P.b1x * px + P.b1y * py), w1 = Rmax * b1n;
    int l1lo = (int)ceil(c1 - w1), l1hi = (int)floor(c1 + w1);
    double d22 = P.d2x * P.d2x + P.d2y * P.d2y;
    const double gcon = ive / kSQPI;
    cplx S = mk(0, 0), Sxy[2] = {mk(0, 0), mk(0, 0)}, Sz = mk(0, 0);
    cplx Hd = mk(0, 0), Hxx = mk(0, 0), Hyy = mk(0, 0), Hzz = mk(0, 0), Hxy = mk(0, 0), Hyz = mk(0, 0),
         Hzx = mk(0, 0);
    const cplx step = mk(P.ph2re, P.ph2im);
    for (int l1 = l1lo; l1 <= l1hi; ++l1) {
      const double ux = px + l1 * P.d1x, uy = py + l1 * P.d1y;
      int lo, hi;
      if (!row_range(ux, uy, P.d2x, P.d2y, d22, R2 - pz * pz, lo, hi)) continue;
      const int c = row_centre(ux * P.d2x + uy * P.d2y, d22, lo, hi);
      const cplx E0 = cexpi(P.vk0 * (l1 * P.d1x + c * P.d2x) + P.vk1 * (l1 * P.d1y + c * P.d2y));
      row_walk(lo, hi, c, E0, step, [&](int l2, cplx Eb) {
        const double ax = ux + l2 * P.d2x, ay = uy + l2 * P.d2y;
        const double rl = sqrt(ax * ax + ay * ay + pz * pz), irl = 1.0 / rl;
        const double y = 0.5 * rl * ive;
        const cplx cef = Eb * exp(ke2 - y * y);
        const cplx w = faddeeva_w(ke, y);
        S += cef * (irl * w.re);
        const cplx cf = cef * (irl * irl * (irl * w.re - P.k * w.im + gcon));
        Sxy[0] += cf * ax;
        Sxy[1] += cf * ay;
        Sz += cf;
        if constexpr (HESS) {
          const double i2 = irl * irl;
          const cplx qr = cef * (i2 * (0.5 * gcon * ive * ive - P.k2 * w.re * irl)) + cf * (3.0 * i2);
          Hd += cf;
          Hxx += qr * (ax * ax);
          Hyy += qr * (ay * ay);
          Hzz += qr * (pz * pz);
          Hxy += qr * (ax * ay);
          Hyz += qr * (ay * pz);
          Hzx += qr * (ax * pz);
        }
      });
    }
    o.g = S * kI4PI;
    o.d[0] = Sxy[0] * (-kI4PI);
    o.d[1] = Sxy[1] * (-kI4PI);
    o.d[2] = Sz * (-pz * kI4PI);
    if constexpr (HESS) {
      o.h[0] = (Hxx - Hd) * kI4PI;
      o.h[1] = (Hyy - Hd) * kI4PI;
      o.h[2] = (Hzz - Hd) * kI4PI;
      o.h[3] = Hxy * kI4PI;
      o.h[4] = Hyz * kI4PI;
      o.h[5] = Hzx * kI4PI;
    }
  }
  // ---- spectral sum
  if constexpr (!HESS) {
    if (pz == 0.0 && z0_ref) {
      spectral_z0_ref(P, px, py, ve, o);
      return;
    }
  }
  // |beta| ve <= sqrt(cut)
  {
    double A2 = P.k2 + P.cut * ive * ive;
    double A = sqrt(A2) * 1.0000001;
    A2 = A * A;
    double d1n = sqrt(P.d1x * P.d1x + P.d1y * P.d1y);
    double c1 = (P.vk0 * P.d1x + P.vk1 * P.d1y) / (2.0 * kPI), w1 = A * d1n / (2.0 * kPI);
    int n1lo = (int)ceil(c1 - w1), n1hi = (int)floor(c1 + w1);
    double g22 = P.g2x * P.g2x + P.g2y * P.g2y;
    cplx S = mk(0, 0), Sx = mk(0, 0), Sy = mk(0, 0), Sz = mk(0, 0);
    cplx Hxx = mk(0, 0), Hyy = mk(0, 0), Hxy = mk(0, 0), Hxz = mk(0, 0), Hyz = mk(0, 0), Hb = mk(0, 0),
         Hg = mk(0, 0);
    const double gp1 = P.g1x * px + P.g1y * py, gp2 = P.g2x * px + P.g2y * py, ap = -(P.vk0 * px + P.vk1 * py);
    const cplx step = cexpi(gp2);
    for (int n1 = n1lo; n1 <= n1hi; ++n1) {
      const double ux = -P.vk0 + n1 * P.g1x, uy = -P.vk1 + n1 * P.g1y;
      int lo, hi;
      if (!row_range(ux, uy, P.g2x, P.g2y, g22, A2, lo, hi)) continue;
      const int c = row_centre(ux * P.g2x + uy * P.g2y, g22, lo, hi);
      row_walk(lo, hi, c, cexpi(ap + n1 * gp1 + c * gp2), step, [&](int n2, cplx E) {
        cplx F, Gz;
        double g0;
        const double ax = ux + n2 * P.g2x, ay = uy + n2 * P.g2y;
        spectral_term(P, ax, ay, E, pz, ve, F, Gz, g0);
        S += F;
        Sx += F * ax;
        Sy += F * ay;
        Sz += Gz;
        if constexpr (HESS) {
          Hxx += F * (ax * ax);
          Hyy += F * (ay * ay);
          Hxy += F * (ax * ay);
          Hxz += Gz * ax;
          Hyz += Gz * ay;
          Hb += F * (P.k2 - ax * ax - ay * ay);
          Hg += E * g0;
        }
      });
    }
    double cd = 0.25 * P.aidd;
    o.g += cmul_i(S) * cd;
    o.d[0] += Sx * (-cd);
    o.d[1] += Sy * (-cd);
    o.d[2] += Sz * cd;
    if constexpr (HESS) {
      o.h[0] += cmul_i(Hxx) * (-cd);
      o.h[1] += cmul_i(Hyy) * (-cd);
      o.h[2] += cmul_i(Hb) * (-cd) + Hg * (-2.0 * cd * ive / kSQPI);
      o.h[3] += cmul_i(Hxy) * (-cd);
      o.h[4] += cmul_i(Hyz) * cd;
      o.h[5] += cmul_i(Hxz) * cd;
    }
  }
}
__device__ __forceinline__ void qpgf_ewald(const bq_qplan& P, double px, double py, double pz, double ve, QG& o,
                                           bool z0_ref)
{
  qpgf_ewald_t<false>(P, px, py, pz, ve, o, z0_ref);
}

// Value and gradient at r (relative position, source minus field point).  z0_ref = true
// reproduces the reference's _v1 result at r_z == 0 (see spectral_z0_ref); false gives the
// fully converged sum there (what the reference's _qG variant returns for the value).
__device__ void qpgf_v1(const bq_qplan& P, double rx, double ry, double rz, QG& o, bool z0_ref = true)
{
  double l1, l2, px, py;
  fold_cell(P, rx, ry, l1, l2, px, py);
  const double pz = rz;
  const double ve = ewald_param(P, px, py, pz);
  // Both sums converge to the same value; take the cheaper one.  Term counts are the lattice
  // points inside each truncation disk (area / cell area).
  bool fd = false;
  if (P.fd_allowed && pz != 0.0) {
    const double bz = P.cut / fabs(pz);
    const double n_fd = kPI * (P.k2 + bz * bz) / P.rcell;
    const double n_sn = kPI * (P.k2 + P.cut / (ve * ve)) / P.rcell;
    const double rs2 = 4.0 * ve * ve * (P.k2 * ve * ve + P.cut) - pz * pz;
    const double n_sp = rs2 > 0.0 ? kPI * rs2 / P.cell : 0.0;
    fd = BQ_COST_FD * n_fd < BQ_COST_SPEC * n_sn + BQ_COST_SPAT * n_sp;
  }
  if (fd) qpgf_fourier(P, px, py, pz, o);
  else qpgf_ewald(P, px, py, pz, ve, o, z0_ref);
  if (l1 != 0.0 || l2 != 0.0) {
    cplx cf = cexpi(-(P.vk0 * (l1 * P.d1x + l2 * P.d2x) + P.vk1 * (l1 * P.d1y + l2 * P.d2y)));
    o.g = o.g * cf;
    for (int i = 0; i < 3; ++i) o.d[i] = o.d[i] * cf;
  }
}

// ---- wavefront-uniform form -----------------------------------------------------------------
// qpgf_v1_wave is qpgf_v1 for a whole wavefront (all 64 lanes must call it together, each with
// its own point).  The lanes take one method (the Fourier-domain sum only when every lane would
// take it), and the lattice sums run over the union of the lanes' truncation disks with
// wave-uniform bounds: the spectral disks are centred on -k_t for every lane, so their union
// is the largest of them, and every lane adds every order of it (the added evanescent terms are
// below each lane's own cut); the spatial disks are bounded by one disk around lane 0's point,
// and a lane adds only the lattice points inside its own disk.  The Faddeeva series take the
// wavefront's largest term count.  Loop control, order data, the propagating/evanescent choice
// and the Bloch phases of the spatial sum are then uniform and the lanes stay converged.
template <int CTRL>
__device__ __forceinline__ double wave_dpp(double x)
{
  const long long b = __double_as_longlong(x);
  const int lo = __builtin_amdgcn_update_dpp(0, (int)b, CTRL, 0xF, 0xF, false);
  const int hi = __builtin_amdgcn_update_dpp(0, (int)(b >> 32), CTRL, 0xF, 0xF, false);
  return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}
__device__ __forceinline__ double wave_lane(double x, int l)
{
  const long long b = __double_as_longlong(x);
  const int lo = __builtin_amdgcn_readlane((int)b, l), hi = __builtin_amdgcn_readlane((int)(b >> 32), l);
  return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}
// maximum over the 64 lanes: four DPP steps give each 16-lane row its maximum (quad
// permutations [1,0,3,2] and [2,3,0,1], half-row mirror, row mirror), then the rows' values
__device__ __forceinline__ double wave_max(double v)
{
  v = fmax(v, wave_dpp<0xB1>(v));
  v = fmax(v, wave_dpp<0x4E>(v));
  v = fmax(v, wave_dpp<0x141>(v));
  v = fmax(v, wave_dpp<0x140>(v));
  return fmax(fmax(wave_lane(v, 0), wave_lane(v, 16)), fmax(wave_lane(v, 32), wave_lane(v, 48)));
}
__device__ __forceinline__ double wave_min(double v) { return -wave_max(-v); }
__device__ __forceinline__ double wave_first(double v)
{
  const long long b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_readfirstlane((int)(b & 0xffffffff)), hi = __builtin_amdgcn_readfirstlane((int)(b >> 32));
  return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

// Ewald spectral term (as spectral_term) with the evanescent erfc branches folded onto one
// erfcx each: for a = bv +- z/2v <= 0, e^{+-bz} erfc(a) = 2 e^{+-bz} - e12 erfcx(-a), because
// +-bz - a^2 = -b^2 v^2 - z^2/4v^2; the Faddeeva series take nterm terms.
// ez = exp(b2 v^2 - az^2) of the order (exp(-bv^2 - az^2) when evanescent), carried by the caller
__device__ __forceinline__ void spectral_term_w(const bq_qplan& P, double ax, double ay, cplx ce, double pz, double ve,
                                                double az, int nterm, const double* et, double ez, cplx& F, cplx& Gz)
{
  const double b2 = P.k2 - ax * ax - ay * ay;
  if (b2 > 0.0) {
    const double ib = rsqrt_nr(b2), b = b2 * ib, ef = ez;
    const cplx ff = cexpi(b * pz);
    const cplx w = faddeeva_w_n(b * ve, az, nterm, et);
    F = ce * (mk(ff.re, ef * w.im + ff.im) * (2.0 * ib));
    Gz = ce * (mk(ef * w.re - ff.re, -ff.im) * 2.0);
  } else if (b2 == 0.0) {
    qpgf_flag(P, -23);
    F = Gz = mk(0.0, 0.0);
  } else {
    const double ib = rsqrt_nr(-b2), b = -b2 * ib, bv = b * ve;
    const double a1 = bv + az, a2 = bv - az;
    const double e12 = ez;
    double f1 = e12 * erfcx_tab(fabs(a1), et), f2 = e12 * erfcx_tab(fabs(a2), et);
    if (a1 <= 0.0) f1 = 2.0 * exp_fast(b * pz) - f1;
    if (a2 <= 0.0) f2 = 2.0 * exp_fast(-b * pz) - f2;
    F = mk(ce.im, -ce.re) * ((f1 + f2) * ib);
    Gz = ce * (f1 - f2);
  }
}

__device__ void qpgf_fourier_w(const bq_qplan& P, double px, double py, double pz, QG& o)
{
  const double az = fabs(pz), bcut = P.cut / az;
  const double A = wave_max(sqrt(P.k2 + bcut * bcut) * 1.0000001), A2 = A * A;
  const double d1n = sqrt(P.d1x * P.d1x + P.d1y * P.d1y);
  const double c1 = (P.vk0 * P.d1x + P.vk1 * P.d1y) / (2.0 * kPI), w1 = A * d1n / (2.0 * kPI);
  const int n1lo = (int)ceil(c1 - w1), n1hi = (int)floor(c1 + w1);
  const double g22 = P.g2x * P.g2x + P.g2y * P.g2y;
  const double gp1 = P.g1x * px + P.g1y * py, gp2 = P.g2x * px + P.g2y * py, ap = -(P.vk0 * px + P.vk1 * py);
  const cplx step = cexpi(gp2);
  cplx S = mk(0, 0), Sx = mk(0, 0), Sy = mk(0, 0), Sz = mk(0, 0);
  for (int n1 = n1lo; n1 <= n1hi; ++n1) {
    const double ux = -P.vk0 + n1 * P.g1x, uy = -P.vk1 + n1 * P.g1y;
    int lo, hi;
    if (!row_range(ux, uy, P.g2x, P.g2y, g22, A2, lo, hi)) continue;
    const int c = row_centre_r(ux * P.g2x + uy * P.g2y, 1.0 / g22, lo, hi);
    row_walk(lo, hi, c, cexpi(ap + n1 * gp1 + c * gp2), step, [&](int n2, cplx E) {
      const double ax = ux + n2 * P.g2x, ay = uy + n2 * P.g2y;
      const double b2 = P.k2 - ax * ax - ay * ay;
      cplx dz, F;
      if (b2 > 0.0) {
        const double ib = rsqrt_nr(b2), b = b2 * ib;
        dz = E * cexpi(b * az);
        F = dz * ib;
      } else if (b2 < 0.0) {
        const double ib = rsqrt_nr(-b2), b = -b2 * ib;
        dz = E * exp_fast(-b * az);
        F = mk(dz.im, -dz.re) * ib;  // -i dz / b
      } else {
        qpgf_flag(P, -12);
        dz = F = mk(0.0, 0.0);
      }
      S += F;
      Sx += F * ax;
      Sy += F * ay;
      Sz += dz;
    });
  }
  const double h = 0.5 * P.aidd, sg = (pz > 0.0) ? 1.0 : -1.0;
  o.g = cmul_i(S) * h;
  o.d[0] = Sx * (-h);
  o.d[1] = Sy * (-h);
  o.d[2] = Sz * (-h * sg);
}

__device__ void qpgf_ewald_w(const bq_qplan& P, double px, double py, double pz, double ve, QG& o, bool z0_ref,
                             const double* et, sconst_dp wrow)
{
  const double ive = 1.0 / ve, ke = P.k * ve, ke2 = ke * ke;
  // ---- spatial sum: lane's own disk |r + R_l| <= Rmax, walked over one disk around lane 0
  {
    const double Rmax = 2.0 * ve * sqrt(ke2 + P.cut) * 1.0000001, Rmax2 = Rmax * Rmax;
    const double rho2 = Rmax2 - pz * pz;
    const double px0 = wave_first(px), py0 = wave_first(py);
    const double du = sqrt((px - px0) * (px - px0) + (py - py0) * (py - py0));
    const double Ru = wave_max(rho2 > 0.0 ? du + sqrt(rho2) : 0.0) * 1.0000001, Ru2 = Ru * Ru;
    const int nterm = wrow ? (int)wrow[3] : w_series_terms(wave_max(ke));
    const double b1n = sqrt(P.b1x * P.b1x + P.b1y * P.b1y);
    const double c1 = -(P.b1x * px0 + P.b1y * py0), w1 = Ru * b1n;
    const int l1lo = (int)ceil(c1 - w1), l1hi = (int)floor(c1 + w1);
    const double d22 = P.d2x * P.d2x + P.d2y * P.d2y;
    const double gcon = ive / kSQPI;
    cplx S = mk(0, 0), Sxy[2] = {mk(0, 0), mk(0, 0)}, Sz = mk(0, 0);
    const cplx step = mk(P.ph2re, P.ph2im);
    for (int l1 = l1lo; l1 <= l1hi; ++l1) {
      const double ux0 = px0 + l1 * P.d1x, uy0 = py0 + l1 * P.d1y;
      int lo, hi;
      if (!row_range(ux0, uy0, P.d2x, P.d2y, d22, Ru2, lo, hi)) continue;
      const int c = row_centre_r(ux0 * P.d2x + uy0 * P.d2y, 1.0 / d22, lo, hi);
      const cplx E0 = cexpi(P.vk0 * (l1 * P.d1x + c * P.d2x) + P.vk1 * (l1 * P.d1y + c * P.d2y));
      const double ux = px + l1 * P.d1x, uy = py + l1 * P.d1y;
      row_walk(lo, hi, c, E0, step, [&](int l2, cplx Eb) {
        const double ax = ux + l2 * P.d2x, ay = uy + l2 * P.d2y;
        const double rl2 = ax * ax + ay * ay + pz * pz;
        if (rl2 <= Rmax2) {
          const double irl = rsqrt_nr(rl2), rl = rl2 * irl;
          const double y = 0.5 * rl * ive;
          const cplx cef = Eb * exp_fast(ke2 - y * y);
          const cplx w = wrow ? faddeeva_w_w(ke, y, nterm, wrow, et) : faddeeva_w_n(ke, y, nterm, et);
          S += cef * (irl * w.re);
          const cplx cf = cef * (irl * irl * (irl * w.re - P.k * w.im + gcon));
          Sxy[0] += cf * ax;
          Sxy[1] += cf * ay;
          Sz += cf;
        }
      });
    }
    o.g = S * kI4PI;
    o.d[0] = Sxy[0] * (-kI4PI);
    o.d[1] = Sxy[1] * (-kI4PI);
    o.d[2] = Sz * (-pz * kI4PI);
  }
  // ---- spectral sum over the union of the lanes' disks |beta| ve <= sqrt(cut)
  const double A = wave_max(sqrt(P.k2 + P.cut * ive * ive) * 1.0000001), A2 = A * A;
  const double vemax = wave_max(ve);
  if (pz == 0.0 && z0_ref) {
    spectral_z0_ref(P, px, py, ve, o);
    return;
  }
  const double d1n = sqrt(P.d1x * P.d1x + P.d1y * P.d1y);
  const double c1 = (P.vk0 * P.d1x + P.vk1 * P.d1y) / (2.0 * kPI), w1 = A * d1n / (2.0 * kPI);
  const int n1lo = (int)ceil(c1 - w1), n1hi = (int)floor(c1 + w1);
  const double g22 = P.g2x * P.g2x + P.g2y * P.g2y;
  const double az = 0.5 * pz / ve;
  cplx S = mk(0, 0), Sx = mk(0, 0), Sy = mk(0, 0), Sz = mk(0, 0);
  const double gp1 = P.g1x * px + P.g1y * py, gp2 = P.g2x * px + P.g2y * py, ap = -(P.vk0 * px + P.vk1 * py);
  const cplx step = cexpi(gp2), back = conj(step);
  // Along a row the orders' factor exp(-v^2 (|a|^2 - k^2) - az^2) is a Gaussian in n2: it is
  // carried from the row's centre by the ratio R(n2) = exp(-v^2 (2 u.g2 + (2 n2 + 1) g22)) of
  // consecutive orders, itself multiplied by Qg = exp(-2 v^2 g22) per step (three exponentials
  // per row instead of one per order)
  const double v2 = ve * ve, eaz = exp_fast(-az * az), Qg = exp_fast(-2.0 * v2 * g22);
  double ux_cur = 0.0, uy_cur = 0.0;
  auto term = [&](int n2, cplx E, double ez) {
    const double ax = ux_cur + n2 * P.g2x, ay = uy_cur + n2 * P.g2y;
    // term count of the propagating orders' Faddeeva series for the largest ve (x = b ve)
    const double bp2 = P.k2 - ax * ax - ay * ay;
    const int nt = bp2 > 0.0 ? w_series_terms(sqrt(bp2) * vemax) : 0;
    cplx F, Gz;
    spectral_term_w(P, ax, ay, E, pz, ve, az, nt, et, ez, F, Gz);
    S += F;
    Sx += F * ax;
    Sy += F * ay;
    Sz += Gz;
  };
  for (int n1 = n1lo; n1 <= n1hi; ++n1) {
    const double ux = -P.vk0 + n1 * P.g1x, uy = -P.vk1 + n1 * P.g1y;
    int lo, hi;
    if (!row_range(ux, uy, P.g2x, P.g2y, g22, A2, lo, hi)) continue;
    const double ue = ux * P.g2x + uy * P.g2y;
    const int c = row_centre_r(ue, 1.0 / g22, lo, hi);
    ux_cur = ux;
    uy_cur = uy;
    const double axc = ux + c * P.g2x, ayc = uy + c * P.g2y;
    const double Gc = exp_fast(-v2 * (axc * axc + ayc * ayc - P.k2)) * eaz;
    const double Rb = exp_fast(v2 * (2.0 * ue + (2.0 * c - 1.0) * g22));  // G(c - 1) / G(c)
    const cplx E0 = cexpi(ap + n1 * gp1 + c * gp2);
    cplx E = E0;
    double Gv = Gc, R = exp_fast(-v2 * (2.0 * ue + (2.0 * c + 1.0) * g22));  // G(c + 1) / G(c)
    for (int n2 = c; n2 <= hi; ++n2) {
      term(n2, E, Gv);
      E = E * step;
      Gv *= R;
      R *= Qg;
    }
    E = E0 * back;
    Gv = Gc * Rb;
    R = Rb * Qg;
    for (int n2 = c - 1; n2 >= lo; --n2) {
      term(n2, E, Gv);
      E = E * back;
      Gv *= R;
      R *= Qg;
    }
  }
  const double cd = 0.25 * P.aidd;
  o.g += cmul_i(S) * cd;
  o.d[0] += Sx * (-cd);
  o.d[1] += Sy * (-cd);
  o.d[2] += Sz * cd;
}

__device__ void qpgf_v1_wave(const bq_qplan& P, double rx, double ry, double rz, QG& o, const double* et,
                             bool z0_ref = true)
{
  double l1, l2, px, py;
  fold_cell(P, rx, ry, l1, l2, px, py);
  const double pz = rz;
  double ve = ewald_param(P, px, py, pz);
  // With the plan's weight table the wavefront takes one Ewald parameter, the largest grid
  // value v_j not above any lane's own (so no lane's exponential amplification grows), and
  // the spatial sums' Faddeeva series read their weights from row j.  A wavefront with a lane
  // at r_z == 0 keeps the lanes' own parameters (the reference's stop rule there depends on it).
  sconst_dp wrow = nullptr;
  if (P.wtab && __all(pz != 0.0 || !z0_ref)) {
    const double vmin = wave_min(ve);
    int j = __builtin_amdgcn_readfirstlane(max(0, (int)ceil(log(P.veps0 / vmin) / P.wtab_lrho)));
    if (j < P.wtab_nv && ((sconst_dp)P.wtab)[(long)j * BQ_WTAB_ROW] > vmin) ++j;
    if (j < P.wtab_nv) {
      wrow = (sconst_dp)P.wtab + (long)j * BQ_WTAB_ROW;
      ve = wrow[0];
    }
  }
  bool fd = false;
  if (P.fd_allowed && pz != 0.0) {
    const double bz = P.cut / fabs(pz);
    const double n_fd = kPI * (P.k2 + bz * bz) / P.rcell;
    const double n_sn = kPI * (P.k2 + P.cut / (ve * ve)) / P.rcell;
    const double rs2 = 4.0 * ve * ve * (P.k2 * ve * ve + P.cut) - pz * pz;
    const double n_sp = rs2 > 0.0 ? kPI * rs2 / P.cell : 0.0;
    fd = BQ_COST_FD_WAVE * n_fd < BQ_COST_SPEC * n_sn + BQ_COST_SPAT * n_sp;
  }
  if (__all(fd)) qpgf_fourier_w(P, px, py, pz, o);
  else qpgf_ewald_w(P, px, py, pz, ve, o, z0_ref, et, wrow);
  if (l1 != 0.0 || l2 != 0.0) {
    cplx cf = cexpi(-(P.vk0 * (l1 * P.d1x + l2 * P.d2x) + P.vk1 * (l1 * P.d1y + l2 * P.d2y)));
    o.g = o.g * cf;
    for (int i = 0; i < 3; ++i) o.d[i] = o.d[i] * cf;
  }
}

// Value, gradient and Hessian at r (the reference's _v2, qgf2_src/d3_qpgf_d2.c:178-230,
// :887-1043, :2011-2399).  Its stop tests watch the xx and yy second derivatives, which do
// not vanish at r_z == 0, so every sum here is the converged one.
__device__ void qpgf_v2(const bq_qplan& P, double rx, double ry, double rz, QG2& o)
{
  double l1, l2, px, py;
  fold_cell(P, rx, ry, l1, l2, px, py);
  const double pz = rz;
  const double ve = ewald_param(P, px, py, pz);
  bool fd = false;
  if (P.fd_allowed && pz != 0.0) {
    const double bz = P.cut / fabs(pz);
    const double n_fd = kPI * (P.k2 + bz * bz) / P.rcell;
    const double n_sn = kPI * (P.k2 + P.cut / (ve * ve)) / P.rcell;
    const double rs2 = 4.0 * ve * ve * (P.k2 * ve * ve + P.cut) - pz * pz;
    const double n_sp = rs2 > 0.0 ? kPI * rs2 / P.cell : 0.0;
    fd = BQ_COST_FD * n_fd < BQ_COST_SPEC * n_sn + BQ_COST_SPAT * n_sp;
  }
  if (fd) {
    qpgf_fourier_t<true>(P, px, py, pz, o);
  } else {
    for (int i = 0; i < 6; ++i) o.h[i] = mk(0, 0);
    qpgf_ewald_t<true>(P, px, py, pz, ve, o, false);
  }
  if (l1 != 0.0 || l2 != 0.0) {
    cplx cf = cexpi(-(P.vk0 * (l1 * P.d1x + l2 * P.d2x) + P.vk1 * (l1 * P.d1y + l2 * P.d2y)));
    o.g = o.g * cf;
    for (int i = 0; i < 3; ++i) o.d[i] = o.d[i] * cf;
    for (int i = 0; i < 6; ++i) o.h[i] = o.h[i] * cf;
  }
}

}  // namespace bq
