// bq_faddeeva.hpp -- Faddeeva function w(z) = exp(-z^2) erfc(-iz) in FP64 for device code.
//
// Methods (published):
//   * M. R. Zaghloul, A. N. Ali, "Algorithm 916", ACM TOMS 38(2):15, 2011 -- w as erfcx(y),
//     cos/sin(2xy) terms plus five series in n with weights exp(-a^2 n^2)/(a^2 n^2 + y^2),
//     a = pi / sqrt(-ln(relerr/2)).  Here relerr = 1e-16, and exp(-a^2 n^2) is advanced by
//     the exact recurrence t_{n+1} = t_n * exp(-a^2 (2n+1)), so no table and no per-term exp.
//   * the Laplace continued fraction for large |z| (Gautschi 1970), with the number of
//     levels of Poppe & Wijers (1990).
// The region split is the one used by the reference's Faddeeva package
// (qgf2_src/Faddeeva.cc:712-717): the series lose accuracy in Re w for x ~ 6, small y.
// The number of series terms is fixed per call from x (the series decays like
// exp(-(a n - x)^2)), so lanes of a wavefront run the same trip count when their x agree.
#pragma once
#include "bq_complex.hpp"

namespace bq {

constexpr double kISQPI = 0.56418958354775628694807945156;   // 1/sqrt(pi)
constexpr double kFA = 0.5127840463909251172929805;          // a for relerr = 1e-16
constexpr double kFA2 = 0.2629474782330504428098607;         // a^2
constexpr double kFC = 0.3264484628871180195202143;          // 2a/pi

__device__ __forceinline__ double rcp_nr(double d)
{
  // hardware reciprocal estimate refined by two Newton steps (full double precision)
  double r = __builtin_amdgcn_rcp(d);
  double e = fma(-d, r, 1.0);
  r = fma(r, e, r);
  e = fma(-d, r, 1.0);
  return fma(r, e, r);
}

// erfcx(x) = exp(x^2) erfc(x) for x >= 0 from a table in LDS (copied there by erfcx_stage):
// piecewise degree-7 polynomials in t = 4/(4+x) on 64 equal intervals of [4/54, 1] (x <= 50;
// coefficients from gen_erfcx_table.py), the asymptotic series 1/(x sqrt(pi)) (1 - 1/(2x^2) +
// 3/(4x^4) - ...) beyond.  Relative error about 1e-15.  The lanes of a wavefront with nearby
// arguments read the same coefficients (LDS broadcast): about 40 vector instructions, against
// about 100 for the device library's erfcx.
constexpr int kErfcxTab = 64 * 8;
__device__ const double kErfcxCoef[kErfcxTab] = {
#include "bq_erfcx_table.inc"
};
__device__ __forceinline__ void erfcx_stage(double* lds)
{
  for (int i = threadIdx.x; i < kErfcxTab; i += blockDim.x) lds[i] = kErfcxCoef[i];
}
__device__ __forceinline__ double erfcx_tab(double x, const double* tab)
{
  if (x > 50.0) {
    const double r = rcp_nr(x), r2 = 0.5 * r * r;
    return kISQPI * r * (1.0 - r2 * (1.0 - 3.0 * r2 * (1.0 - 5.0 * r2 * (1.0 - 7.0 * r2 * (1.0 - 9.0 * r2)))));
  }
  constexpr double t0 = 4.0 / 54.0, sc = 64.0 / (1.0 - t0);
  const double s = (4.0 * rcp_nr(4.0 + x) - t0) * sc;
  const int k = max(0, min(63, (int)s));
  const double u = 2.0 * (s - k) - 1.0;
  const double* c = tab + 8 * k;
  double h = c[7];
#pragma unroll
  for (int p = 6; p >= 0; --p) h = fma(h, u, c[p]);
  return h;
}
// erfcx on the real line: the table, or the device library's when tab is null
__device__ __forceinline__ double erfcx_any(double x, const double* tab)
{
  if (!tab) return erfcx(x);
  return x >= 0.0 ? erfcx_tab(x, tab) : 2.0 * exp_fast(x * x) - erfcx_tab(-x, tab);
}

__device__ __forceinline__ double rsqrt_nr(double x)
{
  // hardware reciprocal-square-root estimate refined by two Newton steps (full double precision)
  double y = __builtin_amdgcn_rsq(x);
  const double h = 0.5 * x;
  y = y * fma(-h * y, y, 1.5);
  return y * fma(-h * y, y, 1.5);
}

__device__ __forceinline__ double sinc_of(double x, double sx)
{
  return (fabs(x) < 1e-4) ? 1.0 - x * x * (1.0 / 6.0) : sx * rcp_nr(x);
}

// continued fraction for Im z >= 0
__device__ __forceinline__ cplx w_cf_upper(double x, double y)
{
  double rho2 = (x * (1.0 / 6.3)) * (x * (1.0 / 6.3)) + (y * (1.0 / 4.4)) * (y * (1.0 / 4.4));
  int nu = (int)(3.0 + 1442.0 / (26.0 * sqrt(rho2) + 77.0)) + 2;
  cplx z = mk(x, y), t = z;
  for (int k = nu; k >= 1; --k) t = z - cdiv(mk(0.5 * k, 0.0), t);
  cplx r = cdiv(mk(0.0, kISQPI), t);
  return r;
}

// Number of series terms for x = |Re z|: the terms decay like exp(-(a n - x)^2).
__device__ __forceinline__ int w_series_terms(double x) { return min(60, (int)((x + 6.3) * (1.0 / kFA)) + 2); }

// Series of Algorithm 916 for |y| <= 7 and x = |xs| < 10 outside the continued-fraction region,
// with nterm >= w_series_terms(|xs|) terms (more terms only add ones below the rounding);
// erfcx from the LDS table et (erfcx_tab), or the device library's when et is null.
__device__ __forceinline__ cplx w_series_n(double xs, double y, int nterm, const double* et = nullptr)
{
  const double x = fabs(xs);
  double s1 = 0.0, s2 = 0.0, s3 = 0.0, s4 = 0.0, s5 = 0.0;
  const double ex2 = exp_fast(-x * x);
  const double e2ax = exp_fast(2.0 * kFA * x), em2ax = 1.0 / e2ax;
  double t = exp(-kFA2), q = exp(-3.0 * kFA2);  // exp(-a^2 n^2), exp(-a^2 (2n+1))
  const double q2 = exp(-2.0 * kFA2);
  double p2 = 1.0, pm2 = 1.0, y2 = y * y;
#pragma unroll 1
  for (int n = 1; n <= nterm; ++n) {
    double an = kFA * n;
    double coef = t * ex2 * rcp_nr(fma(an, an, y2));
    p2 *= e2ax;
    pm2 *= em2ax;
    double cm = coef * pm2, cp = coef * p2;
    s1 += coef;
    s2 += cm;
    s4 = fma(cm, an, s4);
    s3 += cp;
    s5 = fma(cp, an, s5);
    t *= q;
    q *= q2;
  }
  const double ex2erfcx = (y > -6.0) ? ex2 * erfcx_any(y, et) : 2.0 * exp_fast(y2 - x * x);
  double sxy, cxy, s2xy, c2xy;
  sincos_fast(xs * y, &sxy, &cxy);
  s2xy = 2.0 * sxy * cxy;
  c2xy = fma(-2.0 * sxy, sxy, 1.0);
  const double coef1 = ex2erfcx - kFC * y * s1;
  const double coef2 = kFC * xs * ex2;
  double re = coef1 * c2xy + coef2 * sxy * sinc_of(xs * y, sxy) + 0.5 * kFC * y * (s2 + s3);
  double im = coef2 * sinc_of(2.0 * xs * y, s2xy) - coef1 * s2xy + copysign(0.5 * kFC * (s5 - s4), xs);
  return mk(re, im);
}
__device__ __forceinline__ cplx w_series(double xs, double y) { return w_series_n(xs, y, w_series_terms(fabs(xs))); }

// Read-only memory read through scalar loads when the address is wave-uniform.
typedef const __attribute__((address_space(4))) double* sconst_dp;

// w_series_n for x = xs > 0 with the x-dependent weights of the series precomputed (one row of
// the plan's weight table, T[6 n + 0..5] = (a n)^2, exp(-(a n)^2 - x^2), exp(-(a n + x)^2),
// a n exp(-(a n + x)^2), exp(-(a n - x)^2), a n exp(-(a n - x)^2), T[2] = exp(-x^2), T[4] = 2a/pi
// for the row's series parameter a, T[5] = 1 when the row's accuracy needs only one Newton step
// on a reciprocal); the row address is wave-uniform, so the weights come through scalar loads
// and a term takes one reciprocal and five fused multiply-adds.
__device__ __forceinline__ cplx w_series_w(double x, double y, int nterm, sconst_dp T, const double* et)
{
  double s1 = 0.0, s2 = 0.0, s3 = 0.0, s4 = 0.0, s5 = 0.0;
  const double ex2 = T[2], y2 = y * y, kFC = T[4];
  if (T[5] != 0.0) {
#pragma unroll 2
    for (int n = 1; n <= nterm; ++n) {
      const sconst_dp c = T + 6 * n;
      const double d = c[0] + y2;
      double r = __builtin_amdgcn_rcp(d);
      r = fma(r, fma(-d, r, 1.0), r);
      s1 = fma(c[1], r, s1);
      s2 = fma(c[2], r, s2);
      s4 = fma(c[3], r, s4);
      s3 = fma(c[4], r, s3);
      s5 = fma(c[5], r, s5);
    }
  } else {
#pragma unroll 2
    for (int n = 1; n <= nterm; ++n) {
      const sconst_dp c = T + 6 * n;
      const double r = rcp_nr(c[0] + y2);
      s1 = fma(c[1], r, s1);
      s2 = fma(c[2], r, s2);
      s4 = fma(c[3], r, s4);
      s3 = fma(c[4], r, s3);
      s5 = fma(c[5], r, s5);
    }
  }
  const double ex2erfcx = (y > -6.0) ? ex2 * erfcx_any(y, et) : 2.0 * exp_fast(y2 - x * x);
  double sxy, cxy, s2xy, c2xy;
  sincos_fast(x * y, &sxy, &cxy);
  s2xy = 2.0 * sxy * cxy;
  c2xy = fma(-2.0 * sxy, sxy, 1.0);
  const double coef1 = ex2erfcx - kFC * y * s1;
  const double coef2 = kFC * x * ex2;
  double re = coef1 * c2xy + coef2 * sxy * sinc_of(x * y, sxy) + 0.5 * kFC * y * (s2 + s3);
  double im = coef2 * sinc_of(2.0 * x * y, s2xy) - coef1 * s2xy + 0.5 * kFC * (s5 - s4);
  return mk(re, im);
}

__device__ __forceinline__ bool w_use_cf(double x, double ya)
{
  return ya > 7.0 || (x > 6.0 && (ya > 0.1 || (x > 8.0 && ya > 1e-10) || x > 28.0));
}

// The rarely taken regions (continued fraction; x >= 10 with tiny |y|; x = 0), kept out of
// line so that the common series path does not carry their registers.
__device__ __attribute__((noinline)) cplx w_cold(double xs, double y)
{
  const double x = fabs(xs), ya = fabs(y);
  if (xs == 0.0) return mk(erfcx(y), 0.0);
  if (w_use_cf(x, ya)) {
    double xr = (y < 0.0) ? -xs : xs;
    cplx w = w_cf_upper(xr, ya);
    if (y < 0.0) {
      cplx e = cexp(mk((ya - xr) * (xr + ya), 2.0 * xr * y));
      return 2.0 * e - w;
    }
    return w;
  }
  // x >= 10, |y| tiny: series centred on n0 = x/a
  double n0 = floor(x * (1.0 / kFA) + 0.5), dx = kFA * n0 - x;
  double s3 = exp(-dx * dx) / fma(kFA2 * n0, n0, y * y);
  double s5 = kFA * n0 * s3;
  for (int dn = 1; dn < 40; ++dn) {
    double np = n0 + dn, nm = n0 - dn;
    double tp = exp(-(kFA * dn + dx) * (kFA * dn + dx)) / fma(kFA2 * np, np, y * y);
    double tm = (nm > 0.0) ? exp(-(kFA * dn - dx) * (kFA * dn - dx)) / fma(kFA2 * nm, nm, y * y) : 0.0;
    s3 += tp + tm;
    s5 += kFA * (np * tp + nm * tm);
  }
  return mk(exp(-x * x) + 0.5 * kFC * y * s3, 0.5 * kFC * copysign(s5, xs));
}

__device__ __forceinline__ cplx faddeeva_w(double xs, double y)
{
  const double x = fabs(xs), ya = fabs(y);
  if (xs == 0.0 || x >= 10.0 || w_use_cf(x, ya)) return w_cold(xs, y);
  return w_series(xs, y);
}
// The same with the series' term count given by the caller (nterm >= w_series_terms(|xs|)), so
// that lanes of a wavefront whose x differ run one trip count.
__device__ __forceinline__ cplx faddeeva_w_n(double xs, double y, int nterm, const double* et = nullptr)
{
  const double x = fabs(xs), ya = fabs(y);
  if (xs == 0.0 || x >= 10.0 || w_use_cf(x, ya)) return w_cold(xs, y);
  return w_series_n(xs, y, nterm, et);
}
// The same for xs > 0 with the series weights from a row of the plan's weight table.
__device__ __forceinline__ cplx faddeeva_w_w(double xs, double y, int nterm, sconst_dp T, const double* et)
{
  const double ya = fabs(y);
  if (xs >= 10.0 || w_use_cf(xs, ya)) return w_cold(xs, y);
  return w_series_w(xs, y, nterm, T, et);
}

}  // namespace bq
