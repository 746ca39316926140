// gg_crtrig.hpp -- sin and cos of one argument, correctly rounded except
// with probability ~1e-5 per call.
//
// The exact FP64 right-hand sides evaluate sin(theta) and cos(theta) once
// per call.  The reference calls the host libm (f64::sin, f64::cos), whose
// results are the correctly rounded values for all but ~0.15% of arguments;
// the device library's sin and cos differ from those in the last bit for ~3%
// of arguments.  The reference controller keeps h unchanged when the error
// estimate rounds to exactly zero (geodesic.rs:318), so such last-bit
// differences decide step sequences far from the hole.  sincos_cr gives the
// correctly rounded values, so the device and the host agree wherever the
// host libm is correctly rounded.
//
// Method (table-based, in the manner of Gal's accurate tables):
//   x = k pi/2 + r, |r| <= pi/4, r as a double-double (Cody-Waite with pi/2
//     in three parts, the products' errors carried by FMA);
//   |r| = a + b with a = j/256 from a table of sin a, cos a as double-double
//     pairs, |b| <= 1/512;
//   sin(a + b) = sin a + sin a (cos b - 1) + cos a sin b, the leading sum
//     sin a + cos a b_hi formed exactly (TwoProd, TwoSum) and the small terms
//     (Taylor polynomials of degree 7 in b, error < 2^-70) added in double.
// The result before its final rounding is within ~2^-71 relative of the
// exact value; it is the correctly rounded one unless the exact value lies
// within that distance of a rounding boundary.  |x| >= 2^20 (never a polar
// angle) goes to the device library.
//
// Plain C++ on value types: the same source compiles for the GPU and for a
// host-side check (contraction off; every fused operation is an explicit
// fma).
#pragma once

#ifndef __HIPCC_RTC__
#include <math.h>
#endif

#include "gg_types.hpp"

#if defined(__HIPCC__) || defined(__HIPCC_RTC__)
#define GG_SINCOS_CONST __constant__ static const
#else
#define GG_SINCOS_CONST static const
#endif

namespace gg {

#include "gg_sincos_table.inc"

// a + b = s + e exactly (Knuth's TwoSum).
GG_HD inline void two_sum(double a, double b, double &s, double &e) {
#pragma clang fp contract(off)
  s = a + b;
  const double bb = s - a;
  e = (a - (s - bb)) + (b - bb);
}

GG_HD inline void sincos_cr(double x, double *sp, double *cp) {
#pragma clang fp contract(off)
  if (!(fabs(x) < 1048576.0)) {   // 2^20; also NaN and infinities
    *sp = sin(x);
    *cp = cos(x);
    return;
  }
  // ---- x = k pi/2 + (r_hi + r_lo) ----
  const double k = rint(x * kTwoOverPi);
  const double p1 = k * kPio2_1;
  const double e1 = fma(k, kPio2_1, -p1);          // k kPio2_1 = p1 + e1
  const double p2 = k * kPio2_2;
  const double e2 = fma(k, kPio2_2, -p2);          // k kPio2_2 = p2 + e2
  double t, te, d, de, u, ue, v, ve, r_hi, r_lo;
  two_sum(x, -p1, t, te);                          // x - p1 = t + te
  two_sum(te, -e1, d, de);                         // te - e1 = d + de
  two_sum(d, -p2, u, ue);                          // d - p2 = u + ue
  const double tail = (de + ue) - (e2 + k * kPio2_3);
  two_sum(t, u, v, ve);                            // r = v + ve + tail: exact but for tail's
  two_sum(v, ve + tail, r_hi, r_lo);               // rounding, ~2^-100 |x|
  // ---- |r| = a + b, a = j / 256 ----
  const bool neg = r_hi < 0.0;
  const double ar = neg ? -r_hi : r_hi;
  const double b_lo = neg ? -r_lo : r_lo;
  const int j = (int)rint(ar * 256.0);
  const double b = ar - (double)j * (1.0 / 256.0);  // exact (Sterbenz, or j = 0)
  const double sa_h = kSinCosTab[j][0], sa_l = kSinCosTab[j][1];
  const double ca_h = kSinCosTab[j][2], ca_l = kSinCosTab[j][3];
  // ---- sin b = b + sb_t, cos b = 1 + cb_t ----
  const double q = b * b;
  const double sb_t = b_lo + b * (q * (-1.0 / 6.0 + q * (1.0 / 120.0 + q * (-1.0 / 5040.0))));
  const double cb_t = q * (-0.5 + q * (1.0 / 24.0 + q * (-1.0 / 720.0))) - b * b_lo;
  // ---- sin(a + b) ----
  const double ps = ca_h * b;
  const double pse = fma(ca_h, b, -ps);
  double s, se;
  two_sum(sa_h, ps, s, se);
  const double s_t = se + pse + sa_l + ca_l * b + sa_h * cb_t + ca_h * sb_t;
  // ---- cos(a + b) ----
  const double pc = sa_h * b;
  const double pce = fma(sa_h, b, -pc);
  double c, ce;
  two_sum(ca_h, -pc, c, ce);
  const double c_t = ce - pce + ca_l - sa_l * b + ca_h * cb_t - sa_h * sb_t;
  double sr = s + s_t;
  const double cr = c + c_t;
  if (neg) sr = -sr;
  // ---- quadrant ----
  const int quad = ((int)k) & 3;
  double so, co;
  if (quad == 0) { so = sr; co = cr; }
  else if (quad == 1) { so = cr; co = -sr; }
  else if (quad == 2) { so = -sr; co = -cr; }
  else { so = -cr; co = sr; }
  *sp = so;
  *cp = co;
}

}  // namespace gg
