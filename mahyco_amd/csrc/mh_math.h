// mh_math.h -- arithmetic helpers shared by the HIP kernels and their host-side checks.
//
// mh_pow_third(x): x raised to the double nearest 1/3, correctly rounded.
//   MaHyCo's characteristic length is std::pow(V, 1./3.) (MahycoModule.cc:1219).  The
//   exponent fl(1/3) = 1/3 - d with d = 2^-54/3 ~ 1.85e-17, so
//       x^fl(1/3) = cbrt(x) * exp(-d ln x) = cbrt(x) * (1 - d ln x) + O(d^2 ln^2 x).
//   cbrt(x) is taken to double-double precision by one Newton step whose residual
//   x - c^3 is formed exactly with fma-based two-products; the d ln x term is ~1e-15
//   relative, so a plain-double log is ample.  The sum is rounded once, which yields the
//   correctly rounded power unless the exact value lies within ~1e-30 relative of a
//   rounding midpoint.  The same code on host and device gives the same bits even when
//   the library cbrt/log differ by an ulp.
// mh_pow_half(x): sqrt(x), the correctly rounded x^0.5 (pow(V, 0.5) in 2D).
#ifndef MH_MATH_H
#define MH_MATH_H

#include <cmath>

#ifdef __HIPCC__
#define MH_HD __host__ __device__ __forceinline__
#else
#define MH_HD inline
#endif

MH_HD double mh_pow_third(double x) {
  if (!(x > 0.0) || !std::isfinite(x)) return std::pow(x, 1.0 / 3.0);
  const double d = 1.8503717077085942e-17;  // 1/3 - fl(1/3) = 2^-54 / 3
  double c = std::cbrt(x);
  // c^3 = c2h*c + c2l*c, with c2h + c2l = c*c exactly
  double c2h = c * c;
  double c2l = std::fma(c, c, -c2h);
  double c3h = c2h * c;
  double c3l = std::fma(c2h, c, -c3h);
  double r = (x - c3h) - (c3l + c2l * c);  // x - c^3, accurate to ~2^-100 relative
  double lo = r / (3.0 * c2h);
  double corr = -d * std::log(x) * c;
  return c + (lo + corr);
}

MH_HD double mh_pow_half(double x) { return std::sqrt(x); }

#endif
