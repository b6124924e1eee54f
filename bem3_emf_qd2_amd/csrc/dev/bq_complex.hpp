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
__device__ __forceinline__ cplx cexpi(double t)
{
  double s, c;
  sincos(t, &s, &c);
  return cplx{c, s};
}
__device__ __forceinline__ cplx cexp(cplx z)
{
  double e = exp(z.re), s, c;
  sincos(z.im, &s, &c);
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
