/*
 * or_faddeeva.c -- Faddeeva function w(z) = exp(-z^2) erfc(-iz) for the oracle.
 *
 * TEST INFRASTRUCTURE ONLY (see oracle.h).
 *
 * The reference calls S. G. Johnson's MIT-licensed Faddeeva package
 * (qgf2_src/Faddeeva.cc:679-972).  That package implements two published methods, restated
 * here from the papers:
 *   - M. R. Zaghloul and A. N. Ali, "Algorithm 916: computing the Faddeyeva and Voigt
 *     functions", ACM TOMS 38(2) 15 (2011): w is a sum of erfcx(y), cos/sin(2xy) terms and
 *     five rapidly convergent series over n with weights exp(-a^2 n^2) / (a^2 n^2 + y^2),
 *     a = pi / sqrt(-ln(relerr/2)).
 *   - the Laplace continued fraction w(z) = (i/sqrt(pi)) / (z - (1/2)/(z - 1/(z - ...)))
 *     for large |z| (Gautschi 1970; Poppe & Wijers 1990 for the number of terms).
 * The region split between them is the one the reference package uses
 * (qgf2_src/Faddeeva.cc:712-717), because the series lose relative accuracy in Re w for
 * x ~ 6 and small y.
 */
#include <math.h>
#include <complex.h>
#include <float.h>
#include "oracle.h"

static const double ISQPI = 0.56418958354775628694807945156; /* 1/sqrt(pi) */

/* erfcx(x) = exp(x^2) erfc(x) to ~2 ulp.  For moderate x the product is formed with an
 * exact split of x^2 (hi + lo) so exp(x^2) keeps full relative accuracy; for large x the
 * asymptotic series 1/(sqrt(pi) x) * sum (-1)^k (2k-1)!! / (2x^2)^k is used. */
double or_erfcx(double x)
{
  if (x < 0.0) {
    if (x < -26.7) return HUGE_VAL;
    double hi = x * x, lo = fma(x, x, -hi);
    return 2.0 * exp(hi) * (1.0 + lo) - or_erfcx(-x);
  }
  if (x < 0.5) return exp(x * x) * erfc(x);
  if (x < 26.0) {
    double hi = x * x, lo = fma(x, x, -hi);
    return exp(hi) * (1.0 + lo) * erfc(x);
  }
  double i2x2 = 0.5 / (x * x), term = 1.0, sum = 1.0;
  for (int k = 1; k < 12; k++) {
    term *= -(2.0 * k - 1.0) * i2x2;
    sum += term;
  }
  return ISQPI / x * sum;
}

static double sinc_of(double x, double sinx)
{
  if (fabs(x) < 1e-4) return 1.0 - x * x / 6.0;
  return sinx / x;
}

/* Laplace continued fraction for Im z >= 0 (caller reflects). */
static double complex cf_upper(double x, double y)
{
  double rho = sqrt((x / 6.3) * (x / 6.3) + (y / 4.4) * (y / 4.4));
  int nu = (int)(3.0 + 1442.0 / (26.0 * rho + 77.0)) + 2; /* Poppe & Wijers count + margin */
  double complex z = x + I * y, t = z;
  for (int k = nu; k >= 1; k--) t = z - (0.5 * k) / t;
  return I * ISQPI / t;
}

double complex or_w(double complex z, double relerr)
{
  double xs = creal(z), y = cimag(z);
  if (xs == 0.0) return or_erfcx(y) + I * xs;
  double x = fabs(xs), ya = fabs(y);

  if (relerr < DBL_EPSILON) relerr = DBL_EPSILON;
  if (relerr > 0.1) relerr = 0.1;
  const double a = M_PI / sqrt(-log(relerr * 0.5));
  const double a2 = a * a, c = 2.0 * a / M_PI;

  /* large |z|: continued fraction (Faddeeva.cc:712-717 region rule) */
  if (ya > 7.0 || (x > 6.0 && (ya > 0.1 || (x > 8.0 && ya > 1e-10) || x > 28.0))) {
    double xr = (y < 0.0) ? -xs : xs;
    double complex wv = cf_upper(xr, ya);
    if (y < 0.0) {
      /* w(z) = 2 exp(-z^2) - w(-z) */
      double complex ez = cexp((ya - xr) * (xr + ya) + I * (2.0 * xr * y));
      return 2.0 * ez - wv;
    }
    return wv;
  }

  double s1 = 0, s2 = 0, s3 = 0, s4 = 0, s5 = 0;
  if (x < 10.0) {
    /* Algorithm 916, eqs. for Re w and Im w with all five series. */
    double ex2 = exp(-x * x);
    double e2ax = exp(2.0 * a * x), em2ax = 1.0 / e2ax;
    double p2 = 1.0, pm2 = 1.0;
    for (int n = 1; n < 1000; n++) {
      double an = a * n;
      double coef = exp(-a2 * n * n) * ex2 / (a2 * n * n + y * y);
      p2 *= e2ax;
      pm2 *= em2ax;
      s1 += coef;
      s2 += coef * pm2;       /* exp(-(an+x)^2)/(a^2n^2+y^2) */
      s4 += coef * pm2 * an;
      s3 += coef * p2;        /* exp(-(an-x)^2)/(a^2n^2+y^2) */
      s5 += coef * p2 * an;
      if (coef * p2 * an < relerr * s5) break;
    }
    double ex2erfcx = (y > -6.0) ? ex2 * or_erfcx(y) : 2.0 * exp(y * y - x * x);
    double sxy = sin(xs * y), s2xy = sin(2.0 * xs * y), c2xy = cos(2.0 * xs * y);
    double coef1 = ex2erfcx - c * y * s1;
    double coef2 = c * xs * ex2;
    double re = coef1 * c2xy + coef2 * sxy * sinc_of(xs * y, sxy) + 0.5 * c * y * (s2 + s3);
    double im = coef2 * sinc_of(2.0 * xs * y, s2xy) - coef1 * s2xy
              + copysign(0.5 * c * (s5 - s4), xs);
    return re + I * im;
  }

  /* x >= 10 with tiny y: only the series centred on n0 = x/a matter. */
  double complex ret = exp(-x * x);
  double n0 = floor(x / a + 0.5), dx = a * n0 - x;
  s3 = exp(-dx * dx) / (a2 * n0 * n0 + y * y);
  s5 = a * n0 * s3;
  for (int dn = 1; dn < 2000; dn++) {
    double np = n0 + dn, nm = n0 - dn;
    double tp = exp(-(a * dn + dx) * (a * dn + dx)) / (a2 * np * np + y * y);
    double tm = (nm > 0) ? exp(-(a * dn - dx) * (a * dn - dx)) / (a2 * nm * nm + y * y) : 0.0;
    s3 += tp + tm;
    s5 += a * (np * tp + (nm > 0 ? nm * tm : 0.0));
    if (a * np * tp < relerr * s5 && nm <= 0) break;
    if (a * (np * tp + nm * tm) < relerr * s5) break;
  }
  return ret + (0.5 * c) * y * (s2 + s3) + I * (0.5 * c) * copysign(s5 - s4, xs);
}

void or_faddeeva_w(const double *z, double relerr, double *w)
{
  double complex r = or_w(z[0] + I * z[1], relerr);
  w[0] = creal(r);
  w[1] = cimag(r);
}
