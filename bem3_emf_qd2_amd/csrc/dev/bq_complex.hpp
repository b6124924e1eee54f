// bq_complex.hpp -- minimal FP64 complex arithmetic for device code.
//
// Plain {re, im} pairs in registers; every operation is written out so the compiler can
// fuse into v_fma_f64 without a library complex type in the way.
#pragma once
#include <hip/hip_runtime.h>

struct cplx {
  double re, im;
};

__device__ __forceinline__ cplx mk(double r, double i) { return cplx{r, i}; }
__device__ __forceinline__ cplx operator+(cplx a, cplx b) { return cplx{a.re + b.re, a.im + b.im}; }
__device__ __forceinline__ cplx operator-(cplx a, cplx b) { return cplx{a.re - b.re, a.im - b.im}; }
__device__ __forceinline__ cplx operator-(cplx a) { return cplx{-a.re, -a.im}; }
__device__ __forceinline__ cplx operator*(cplx a, cplx b)
{
  return cplx{fma(a.re, b.re, -a.im * b.im), fma(a.re, b.im, a.im * b.re)};
}
__device__ __forceinline__ cplx operator*(cplx a, double s) { return cplx{a.re * s, a.im * s}; }
__device__ __forceinline__ cplx operator*(double s, cplx a) { return cplx{a.re * s, a.im * s}; }
__device__ __forceinline__ cplx& operator+=(cplx& a, cplx b)
{
  a.re += b.re;
  a.im += b.im;
  return a;
}
__device__ __forceinline__ cplx& operator-=(cplx& a, cplx b)
{
  a.re -= b.re;
  a.im -= b.im;
  return a;
}
__device__ __forceinline__ cplx cmul_i(cplx a) { return cplx{-a.im, a.re}; }  // i * a
__device__ __forceinline__ cplx conj(cplx a) { return cplx{a.re, -a.im}; }
__device__ __forceinline__ double cabs2(cplx a) { return a.re * a.re + a.im * a.im; }
__device__ __forceinline__ cplx cdiv(cplx a, cplx b)
{
  // Smith's algorithm keeps the quotient finite for widely scaled operands.
  if (fabs(b.re) >= fabs(b.im)) {
    double r = b.im / b.re, d = b.re + b.im * r;
    return cplx{(a.re + a.im * r) / d, (a.im - a.re * r) / d};
  }
  double r = b.re / b.im, d = b.im + b.re * r;
  return cplx{(a.re * r + a.im) / d, (a.im * r - a.re) / d};
}
// sin and cos: x = k pi/2 + r with pi/2 split into three doubles (P1 + P2 + P3, rounded
// from the exact value).  x - k P1 is exact under fma for 1 <= |x| < 2^52 because both
// operands are multiples of 2^-52 and the difference is below 2, so r keeps full relative
// accuracy next to multiples of pi/2.  sin r and cos r by their Taylor series to r^19 / r^18
// on |r| <= pi/4 (truncation below 1e-20).  The quadrant comes from k mod 4 in floating point.
__device__ __forceinline__ void sincos_fast(double x, double* s, double* c)
{
  const double k = rint(x * 0x1.45f306dc9c883p-1);  // x 2/pi
  double r = fma(-k, 0x1.921fb54442d18p+0, x);
  r = fma(-k, 0x1.1a62633145c07p-54, r);
  r = fma(-k, -0x1.f1976b7ed8fbcp-110, r);
  const double r2 = r * r;
  double ps = -0x1.2f49b46814157p-57;             // -1/19!
  ps = fma(ps, r2, 0x1.952c77030ad4ap-49);         //  1/17!
  ps = fma(ps, r2, -0x1.ae7f3e733b81fp-41);        // -1/15!
  ps = fma(ps, r2, 0x1.6124613a86d09p-33);         //  1/13!
  ps = fma(ps, r2, -0x1.ae64567f544e4p-26);        // -1/11!
  ps = fma(ps, r2, 0x1.71de3a556c734p-19);         //  1/9!
  ps = fma(ps, r2, -0x1.a01a01a01a01ap-13);        // -1/7!
  ps = fma(ps, r2, 0x1.1111111111111p-7);          //  1/5!
  ps = fma(ps, r2, -0x1.5555555555555p-3);         // -1/3!
  const double sr = fma(ps * r2, r, r);
  double pc = -0x1.6827863b97d97p-53;              // -1/18!
  pc = fma(pc, r2, 0x1.ae7f3e733b81fp-45);         //  1/16!
  pc = fma(pc, r2, -0x1.93974a8c07c9dp-37);        // -1/14!
  pc = fma(pc, r2, 0x1.1eed8eff8d898p-29);         //  1/12!
  pc = fma(pc, r2, -0x1.27e4fb7789f5cp-22);        // -1/10!
  pc = fma(pc, r2, 0x1.a01a01a01a01ap-16);         //  1/8!
  pc = fma(pc, r2, -0x1.6c16c16c16c17p-10);        // -1/6!
  pc = fma(pc, r2, 0x1.5555555555555p-5);          //  1/4!
  pc = fma(pc, r2, -0.5);                          // -1/2!
  const double cr = fma(pc, r2, 1.0);
  const int q = (int)(k - 4.0 * floor(k * 0.25));
  const double sv = (q & 1) ? cr : sr, cv = (q & 1) ? sr : cr;
  *s = (q & 2) ? -sv : sv;
  *c = ((q + 1) & 2) ? -cv : cv;
}

// exp(t): t = k ln2 + r with ln2 = L1 + L2 (L1 of 32 significant bits, so k L1 is exact for
// |k| < 2^21; fdlibm's split), |r| <= ln2 / 2 + ulps, exp(r) by its Taylor polynomial to r^13
// (truncation below 5e-18 relative), 2^k applied by ldexp.  t is clamped to [-746, 710] first,
// so the result underflows to 0 and overflows to inf as exp does (a NaN argument gives 0).
__device__ __forceinline__ double exp_fast(double t)
{
  t = fmin(fmax(t, -746.0), 710.0);
  const double k = rint(t * 0x1.71547652b82fep+0);
  double r = fma(-k, 0x1.62e42fee00000p-1, t);
  r = fma(-k, 0x1.a39ef35793c76p-33, r);
  double p = 0x1.6124613a86d09p-33;               // 1/13!
  p = fma(p, r, 0x1.1eed8eff8d898p-29);           // 1/12!
  p = fma(p, r, 0x1.ae64567f544e4p-26);           // 1/11!
  p = fma(p, r, 0x1.27e4fb7789f5cp-22);           // 1/10!
  p = fma(p, r, 0x1.71de3a556c734p-19);           // 1/9!
  p = fma(p, r, 0x1.a01a01a01a01ap-16);           // 1/8!
  p = fma(p, r, 0x1.a01a01a01a01ap-13);           // 1/7!
  p = fma(p, r, 0x1.6c16c16c16c17p-10);           // 1/6!
  p = fma(p, r, 0x1.1111111111111p-7);            // 1/5!
  p = fma(p, r, 0x1.5555555555555p-5);            // 1/4!
  p = fma(p, r, 0x1.5555555555555p-3);            // 1/3!
  p = fma(p, r, 0.5);
  p = fma(p, r, 1.0);
  p = fma(p, r, 1.0);
  return ldexp(p, (int)k);
}
__device__ __forceinline__ cplx cexpi(double t)
{
  double s, c;
  sincos_fast(t, &s, &c);
  return cplx{c, s};
}
__device__ __forceinline__ cplx cexp(cplx z)
{
  double e = exp(z.re), s, c;
  sincos_fast(z.im, &s, &c);
  return cplx{e * c, e * s};
}
__device__ __forceinline__ cplx ccos(cplx z)
{  // cos(x+iy) = cos x cosh y - i sin x sinh y
  double s, c;
  sincos(z.re, &s, &c);
  return cplx{c * cosh(z.im), -s * sinh(z.im)};
}
__device__ __forceinline__ cplx csin(cplx z)
{  // sin(x+iy) = sin x cosh y + i cos x sinh y
  double s, c;
  sincos(z.re, &s, &c);
  return cplx{s * cosh(z.im), c * sinh(z.im)};
}
