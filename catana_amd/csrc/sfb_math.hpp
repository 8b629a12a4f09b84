// sfb_math.hpp -- FP64 special functions shared by host and device code.
//
// j_l(x) follows the regions of GSL's gsl_sf_bessel_jl (the function the
// reference calls at sfb_raw.cpp:51): Taylor series for x^2 < 10(l+1/2)/e,
// upward recurrence from j_0, j_1 for x >= l, and for x < l the ratio
// j_l/j_{l-1} from its continued fraction followed by downward recurrence
// normalised by j_0 or j_1.  Underflow returns 0 (the reference runs with the
// GSL error handler switched off, sfb_raw.cpp:26).
//
// The Bessel-zero Newton iteration keeps the reference's stopping rule
// (SphericalBesselZeros.cpp:49-61): the loop runs while delta/z > 1e-15, a
// signed test, so it ends after the first step that moves z upward -- one or
// two Newton steps from the guess.  The converged mode tests |delta|/z.
#pragma once

#include <hip/hip_runtime.h>
#include <cmath>

#define SFB_HD __host__ __device__ inline

namespace sfb {

SFB_HD double sj0(double x) {
  if (fabs(x) < 1e-3) {
    const double x2 = x * x;
    return 1.0 - x2 / 6.0 * (1.0 - x2 / 20.0 * (1.0 - x2 / 42.0));
  }
  return sin(x) / x;
}

SFB_HD double sj1(double x) {
  if (fabs(x) < 0.25) {
    const double x2 = x * x;
    return x / 3.0 * (1.0 - x2 / 10.0 * (1.0 - x2 / 28.0 * (1.0 - x2 / 54.0 * (1.0 - x2 / 88.0))));
  }
  return (sin(x) / x - cos(x)) / x;
}

SFB_HD double sjl_series(int l, double x) {
  double pre = 1.0;
  for (int k = 1; k <= l; ++k) pre *= x / (2.0 * k + 1.0);
  if (pre == 0.0) return 0.0;
  const double h = -0.5 * x * x;
  double term = 1.0, sum = 1.0;
  for (int k = 1; k < 1000; ++k) {
    term *= h / (k * (2.0 * l + 2.0 * k + 1.0));
    sum += term;
    if (fabs(term) <= 1e-17 * fabs(sum)) break;
  }
  return pre * sum;
}

SFB_HD double sjl_downward(int l, double x) {
  const double tiny = 1e-300;
  double f = (2.0 * l + 1.0) / x, C = f, D = 0.0;
  for (int k = 1; k < 200000; ++k) {
    const double b = (2.0 * (l + k) + 1.0) / x;
    D = b - D;
    if (D == 0.0) D = tiny;
    D = 1.0 / D;
    C = b - 1.0 / C;
    if (C == 0.0) C = tiny;
    const double delta = C * D;
    f *= delta;
    if (fabs(delta - 1.0) < 1e-16) break;
  }
  const double R = 1.0 / f;
  double up = R, u = 1.0;
  int scale = 0;
  for (int k = l - 1; k >= 1; --k) {
    const double um = (2.0 * k + 1.0) / x * u - up;
    up = u;
    u = um;
    if (fabs(u) > 0x1p+600) {
      u = ldexp(u, -600);
      up = ldexp(up, -600);
      scale += 600;
    }
  }
  const double norm = (fabs(u) >= fabs(up)) ? sj0(x) / u : sj1(x) / up;
  return ldexp(R * norm, -scale);
}

SFB_HD double sbessel_jl(int l, double x) {
  if (x == 0.0) return l == 0 ? 1.0 : 0.0;
  if (l == 0) return sj0(x);
  if (x * x < 10.0 * (l + 0.5) / 2.718281828459045235) return sjl_series(l, x);
  if (x >= (double)l) {
    // upward recurrence j_{k+1} = (2k+1)/x j_k - j_{k-1}, with one reciprocal
    const double inv = 1.0 / x;
    double S, C;
    sincos(x, &S, &C);
    double jm = S * inv;             // j_0
    double j = (jm - C) * inv;       // j_1 (x >= 1 here, no cancellation issue)
    for (int k = 1; k < l; ++k) {
      const double jp = fma((2.0 * k + 1.0) * inv, j, -jm);
      jm = j;
      j = jp;
    }
    return j;
  }
  return sjl_downward(l, x);
}

enum { ZEROS_REFERENCE = 0, ZEROS_CONVERGED = 1 };

SFB_HD double newton_zero(int l, double z, int mode) {
  for (int it = 0; it < 200; ++it) {
    const double jl = sbessel_jl(l, z);
    const double d = jl / ((double)l / z * jl - sbessel_jl(l + 1, z));
    z -= d;
    const double t = (mode == ZEROS_REFERENCE) ? d / z : fabs(d) / z;
    if (!(t > 1e-15)) break;
  }
  return z;
}

// Asymptotic guesses for the first two zeros: the coefficients of
// SphericalBesselZeros.cpp:17-46, kept identical because with the reference's
// early-stopping Newton loop the converged value depends on the guess.
SFB_HD double zero_guess(int l, int which) {
  if (l == 0) return which == 0 ? 3.141592653589793238463 : 6.283185307179586476925;
  const double a = l + 0.5;
  if (which == 0)
    return a + 1.8557571 * pow(a, 1. / 3.) + 1.033150 * pow(a, -1. / 3.) - 0.00397 * pow(a, -1.) -
           0.0908 * pow(a, -5. / 3.) + 0.043 * pow(a, -7. / 3.);
  return a + 3.2446076 * pow(a, 1. / 3.) + 3.158244 * pow(a, -1. / 3.) - 0.08331 * pow(a, -1.) -
         0.8437 * pow(a, -5. / 3.) + 0.864 * pow(a, -7. / 3.);
}

// First n zeros of j_l: two guesses refined by Newton, then
// z_{k+1} = Newton(2 z_k - z_{k-1})  (SphericalBesselZeros.cpp:11-15, 63-68, 95-100).
SFB_HD void bessel_zeros(int l, int n, int mode, double *z) {
  const double z0 = newton_zero(l, zero_guess(l, 0), mode);
  const double z1 = newton_zero(l, zero_guess(l, 1), mode);
  if (n > 0) z[0] = z0;
  if (n > 1) z[1] = z1;
  for (int k = 2; k < n; ++k) z[k] = newton_zero(l, 2.0 * z[k - 1] - z[k - 2], mode);
}

}  // namespace sfb
