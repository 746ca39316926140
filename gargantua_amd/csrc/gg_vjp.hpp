// gg_vjp.hpp -- vector-Jacobian product of the Kerr geodesic RHS, reverse
// mode written out by hand.
//
// a^mu = -Gamma^mu_ab(r, theta; a) v^a v^b (metric.rs:37-52 with the
// Christoffel symbols of kerr.rs:100-424).  For a cotangent lam the adjoint
// needs the gradient of L = lam . a with respect to r, theta, the spin a and
// v.  With the index lowered, a^mu = -g^{mu nu} f_nu and
//   f_nu = v^r w_nu + v^theta u_nu - 1/2 delta_nu^r (w . v) - 1/2 delta_nu^theta (u . v),
//   w_nu = d_r g_{nu b} v^b,  u_nu = d_theta g_{nu b} v^b
// (the metric depends on r and theta only), so L = -Lam . f with
// Lam_nu = lam_mu g^{mu nu}.  The v-gradient is exact in closed form (f is
// quadratic in v); the (r, theta, a)-gradient comes from the cotangents of
// the five inverse-metric components and the ten metric derivatives, swept
// backwards through the metric algebra: about twice the forward algebra
// instead of three tangent passes.  Plain C++ on value types: the same
// source compiles for the GPU and for a host-side check.
#pragma once

#include <math.h>

#include "gg_types.hpp"

namespace gg {

// 1/x: on the device v_rcp_f64 and two Newton steps (~1 ulp; x is a
// positive normal quantity here: Sigma, Delta or sin^2 theta), on the host
// (the check build) the division.
GG_HD inline double vjp_rcp(double x) {
#if defined(__HIP_DEVICE_COMPILE__)
  double r = __builtin_amdgcn_rcp(x);
  double e = __builtin_fma(-x, r, 1.0);
  r = __builtin_fma(r, e, r);
  e = __builtin_fma(-x, r, 1.0);
  return __builtin_fma(r, e, r);
#else
  return 1.0 / x;
#endif
}

struct RhsBar {
  double r, th, a;   // lam^T da/dr, lam^T da/dtheta, lam^T da/dspin
  d4 v;              // lam^T da/dv
};

// The metric algebra of the VJP at (r, theta), given st = sin theta and
// ct = cos theta: the inverse metric, the metric derivatives and the
// intermediates the backward sweep reads.  It does not depend on the
// velocity or the cotangent, so one evaluation serves every product at the
// same point.
struct KerrPoint {
  double m, a, r, st, ct;
  double tm, a2, r2, sin2, sc, ct2, sigma, delta, is, id, is2, isd, dsig, rp, sm, dtt_r, dtt_t, dtp_r, U, dtp_t;
  double bp, W, s4, dpp_r, Z, dpp_t, Y, gtt, gtp, X, isn, gpp, grr, drr_t, N, id2, drr_r, two_r;
};

GG_HD inline KerrPoint kerr_point(double m, double a, double r, double st, double ct) {
  GG_CONTRACT_FAST
  const double tm = 2.0 * m;
  const double a2 = a * a, r2 = r * r, sin2 = st * st, sc = st * ct, ct2 = ct * ct;
  const double sigma = r2 + a2 * ct2;
  const double delta = r2 - tm * r + a2;
  const double is = vjp_rcp(sigma), id = vjp_rcp(delta);
  const double is2 = is * is, isd = is * id;
  const double dsig = -2.0 * a2 * sc;             // d sigma / d theta = d g_thth / d theta
  const double rp = r2 + a2;
  const double sm = sigma - 2.0 * r2;
  const double dtt_r = tm * sm * is2;             // d g_tt / dr
  const double dtt_t = -tm * r * dsig * is2;      // d g_tt / d theta
  const double dtp_r = -tm * a * sin2 * sm * is2;
  const double U = 2.0 * sc * sigma - sin2 * dsig;
  const double dtp_t = -tm * a * r * U * is2;
  const double bp = tm * a2 * r;
  const double W = tm * a2 * sm * is2;
  const double s4 = sin2 * sin2;
  const double dpp_r = 2.0 * r * sin2 + s4 * W;
  const double Z = 4.0 * sin2 * sc * sigma - s4 * dsig;
  const double dpp_t = 2.0 * rp * sc + bp * Z * is2;
  const double Y = rp * rp - a2 * delta * sin2;
  const double gtt = -Y * isd;                    // g^tt
  const double gtp = -tm * a * r * isd;           // g^t phi
  const double X = delta - a2 * sin2;
  const double isn = vjp_rcp(sin2);
  const double gpp = X * isd * isn;               // g^phi phi
  const double grr = delta * is;                  // g^rr
  const double drr_t = dsig * id;                 // d g_rr / d theta
  const double N = 2.0 * r * delta - sigma * (2.0 * r - tm);
  const double id2 = id * id;
  const double drr_r = N * id2;                   // d g_rr / dr
  const double two_r = 2.0 * r;                   // d g_thth / dr

  KerrPoint P;
  P.m = m; P.a = a; P.r = r; P.st = st; P.ct = ct;
  P.tm = tm; P.a2 = a2; P.r2 = r2; P.sin2 = sin2; P.sc = sc; P.ct2 = ct2; P.sigma = sigma; P.delta = delta;
  P.is = is; P.id = id; P.is2 = is2; P.isd = isd; P.dsig = dsig; P.rp = rp; P.sm = sm; P.dtt_r = dtt_r;
  P.dtt_t = dtt_t; P.dtp_r = dtp_r; P.U = U; P.dtp_t = dtp_t; P.bp = bp; P.W = W; P.s4 = s4; P.dpp_r = dpp_r;
  P.Z = Z; P.dpp_t = dpp_t; P.Y = Y; P.gtt = gtt; P.gtp = gtp; P.X = X; P.isn = isn; P.gpp = gpp;
  P.grr = grr; P.drr_t = drr_t; P.N = N; P.id2 = id2; P.drr_r = drr_r; P.two_r = two_r;
  return P;
}

// The VJP at a point of kerr_point.
GG_HD inline RhsBar kerr_rhs_vjp_pt(const KerrPoint &P, d4 v, d4 lam) {
  GG_CONTRACT_FAST
  const double m = P.m, a = P.a, r = P.r, st = P.st, ct = P.ct;
  const double tm = P.tm, a2 = P.a2, r2 = P.r2, sin2 = P.sin2, sc = P.sc, ct2 = P.ct2,
               sigma = P.sigma, delta = P.delta, is = P.is, id = P.id, is2 = P.is2, isd = P.isd,
               dsig = P.dsig, rp = P.rp, sm = P.sm, dtt_r = P.dtt_r, dtt_t = P.dtt_t, dtp_r = P.dtp_r,
               U = P.U, dtp_t = P.dtp_t, bp = P.bp, W = P.W, s4 = P.s4, dpp_r = P.dpp_r,
               Z = P.Z, dpp_t = P.dpp_t, Y = P.Y, gtt = P.gtt, gtp = P.gtp, X = P.X,
               isn = P.isn, gpp = P.gpp, grr = P.grr, drr_t = P.drr_t, N = P.N, id2 = P.id2,
               drr_r = P.drr_r, two_r = P.two_r;
  (void)m;
  // ---- the contraction with lowered index ----
  const double vt = v.x, vr = v.y, vth = v.z, vph = v.w;
  const double w_t = dtt_r * vt + dtp_r * vph, w_p = dtp_r * vt + dpp_r * vph;
  const double w_r = drr_r * vr, w_th = two_r * vth;
  const double u_t = dtt_t * vt + dtp_t * vph, u_p = dtp_t * vt + dpp_t * vph;
  const double u_r = drr_t * vr, u_th = dsig * vth;
  const double wv = w_t * vt + w_p * vph + w_r * vr + w_th * vth;
  const double uv = u_t * vt + u_p * vph + u_r * vr + u_th * vth;
  const double f_t = vr * w_t + vth * u_t, f_p = vr * w_p + vth * u_p;
  const double f_r = vr * w_r + vth * u_r - 0.5 * wv;
  const double f_th = vr * w_th + vth * u_th - 0.5 * uv;
  const double L_t = lam.x * gtt + lam.w * gtp, L_p = lam.x * gtp + lam.w * gpp;
  const double L_r = lam.y * grr, L_th = lam.z * is;

  // ---- v-gradient: -Lam . df/dv ----
  const double Lw = L_t * w_t + L_p * w_p + L_r * w_r + L_th * w_th;
  const double Lu = L_t * u_t + L_p * u_p + L_r * u_r + L_th * u_th;
  RhsBar o;
  o.v.x = -vr * (dtt_r * L_t + dtp_r * L_p) - vth * (dtt_t * L_t + dtp_t * L_p) + L_r * w_t + L_th * u_t;
  o.v.w = -vr * (dtp_r * L_t + dpp_r * L_p) - vth * (dtp_t * L_t + dpp_t * L_p) + L_r * w_p + L_th * u_p;
  o.v.y = -Lw - vr * (drr_r * L_r) - vth * (drr_t * L_r) + L_r * w_r + L_th * u_r;
  o.v.z = -Lu - vr * (two_r * L_th) - vth * (dsig * L_th) + L_r * w_th + L_th * u_th;

  // ---- cotangents of the inverse metric and the metric derivatives ----
  const double b_gtt = -lam.x * f_t;
  const double b_gtp = -(lam.x * f_p + lam.w * f_t);
  const double b_gpp = -lam.w * f_p;
  const double b_grr = -lam.y * f_r;
  double b_is = -lam.z * f_th;                                // g^thth = is
  const double vt2 = vt * vt, vp2 = vph * vph, vtp = vt * vph;
  const double b_dtt_r = -L_t * vr * vt + 0.5 * L_r * vt2;
  const double b_dtt_t = -L_t * vth * vt + 0.5 * L_th * vt2;
  const double b_dtp_r = -(L_t * vph + L_p * vt) * vr + L_r * vtp;
  const double b_dtp_t = -(L_t * vph + L_p * vt) * vth + L_th * vtp;
  const double b_dpp_r = -L_p * vr * vph + 0.5 * L_r * vp2;
  const double b_dpp_t = -L_p * vth * vph + 0.5 * L_th * vp2;
  const double b_drr_r = -0.5 * L_r * vr * vr;
  const double b_drr_t = -L_r * vth * vr + 0.5 * L_th * vr * vr;
  double b_r = 2.0 * (-L_th * vr * vth + 0.5 * L_r * vth * vth);      // d g_thth / dr = 2 r
  double b_dsig = -0.5 * L_th * vth * vth;                            // d g_thth / d theta = dsig

  // ---- backward through the metric algebra ----
  const double b_N = b_drr_r * id2;                           // drr_r = N id^2
  double b_id = b_drr_r * N * 2.0 * id;
  b_r += b_N * (2.0 * delta - 2.0 * sigma);                   // N = 2 r delta - sigma (2 r - 2m)
  double b_delta = b_N * 2.0 * r;
  double b_sigma = -b_N * (2.0 * r - tm);
  b_delta += b_grr * is;                                      // grr = delta is
  b_is += b_grr * delta;
  b_dsig += b_drr_t * id;                                     // drr_t = dsig id
  b_id += b_drr_t * dsig;
  const double b_X = b_gpp * isd * isn;                       // gpp = X isd isn
  double b_isd = b_gpp * X * isn;
  const double b_isn = b_gpp * X * isd;
  double b_sin2 = -b_isn * isn * isn;                         // isn = 1 / sin2
  b_delta += b_X;                                             // X = delta - a2 sin2
  double b_a2 = -b_X * sin2;
  b_sin2 += -b_X * a2;
  double b_a = -tm * r * isd * b_gtp;                         // gtp = -2m a r isd
  b_r += -tm * a * isd * b_gtp;
  b_isd += -tm * a * r * b_gtp;
  const double b_Y = -b_gtt * isd;                            // gtt = -Y isd
  b_isd += -b_gtt * Y;
  double b_rp = b_Y * 2.0 * rp;                               // Y = rp^2 - a2 delta sin2
  b_a2 += -b_Y * delta * sin2;
  b_delta += -b_Y * a2 * sin2;
  b_sin2 += -b_Y * a2 * delta;
  b_rp += b_dpp_t * 2.0 * sc;                                 // dpp_t = 2 rp sc + bp Z is2
  double b_sc = b_dpp_t * 2.0 * rp;
  const double b_bp = b_dpp_t * Z * is2;
  const double b_Z = b_dpp_t * bp * is2;
  double b_is2 = b_dpp_t * bp * Z;
  b_sin2 += b_Z * 4.0 * sc * sigma;                           // Z = 4 sin2 sc sigma - s4 dsig
  b_sc += b_Z * 4.0 * sin2 * sigma;
  b_sigma += b_Z * 4.0 * sin2 * sc;
  double b_s4 = -b_Z * dsig;
  b_dsig += -b_Z * s4;
  b_r += b_dpp_r * 2.0 * sin2;                                // dpp_r = 2 r sin2 + s4 W
  b_sin2 += b_dpp_r * 2.0 * r;
  b_s4 += b_dpp_r * W;
  const double b_W = b_dpp_r * s4;
  b_a2 += b_W * tm * sm * is2;                                // W = 2m a2 sm is2
  double b_sm = b_W * tm * a2 * is2;
  b_is2 += b_W * tm * a2 * sm;
  b_sin2 += b_s4 * 2.0 * sin2;                                // s4 = sin2^2
  b_a2 += b_bp * tm * r;                                      // bp = 2m a2 r
  b_r += b_bp * tm * a2;
  b_a += -b_dtp_t * tm * r * U * is2;                         // dtp_t = -2m a r U is2
  b_r += -b_dtp_t * tm * a * U * is2;
  const double b_U = -b_dtp_t * tm * a * r * is2;
  b_is2 += -b_dtp_t * tm * a * r * U;
  b_sc += b_U * 2.0 * sigma;                                  // U = 2 sc sigma - sin2 dsig
  b_sigma += b_U * 2.0 * sc;
  b_sin2 += -b_U * dsig;
  b_dsig += -b_U * sin2;
  b_a += -b_dtp_r * tm * sin2 * sm * is2;                     // dtp_r = -2m a sin2 sm is2
  b_sin2 += -b_dtp_r * tm * a * sm * is2;
  b_sm += -b_dtp_r * tm * a * sin2 * is2;
  b_is2 += -b_dtp_r * tm * a * sin2 * sm;
  b_r += -b_dtt_t * tm * dsig * is2;                          // dtt_t = -2m r dsig is2
  b_dsig += -b_dtt_t * tm * r * is2;
  b_is2 += -b_dtt_t * tm * r * dsig;
  b_sm += b_dtt_r * tm * is2;                                 // dtt_r = 2m sm is2
  b_is2 += b_dtt_r * tm * sm;
  b_sigma += b_sm;                                            // sm = sigma - 2 r2
  double b_r2 = -2.0 * b_sm;
  b_r2 += b_rp;                                               // rp = r2 + a2
  b_a2 += b_rp;
  b_a2 += -2.0 * sc * b_dsig;                                 // dsig = -2 a2 sc
  b_sc += -2.0 * a2 * b_dsig;
  b_is += b_isd * id;                                         // isd = is id
  b_id += b_isd * is;
  b_is += 2.0 * is * b_is2;                                   // is2 = is^2
  b_delta += -b_id * id * id;                                 // id = 1 / delta
  b_sigma += -b_is * is * is;                                 // is = 1 / sigma
  b_r2 += b_delta;                                            // delta = r2 - 2m r + a2
  b_r += -tm * b_delta;
  b_a2 += b_delta;
  b_r2 += b_sigma;                                            // sigma = r2 + a2 ct^2
  b_a2 += b_sigma * ct2;
  double b_ct = b_sigma * 2.0 * a2 * ct;
  double b_st = b_sc * ct;                                    // sc = st ct
  b_ct += b_sc * st;
  b_st += 2.0 * st * b_sin2;                                  // sin2 = st^2
  b_r += 2.0 * r * b_r2;                                      // r2 = r^2
  b_a += 2.0 * a * b_a2;                                      // a2 = a^2
  o.r = b_r;
  o.th = b_st * ct - b_ct * st;                               // st = sin th, ct = cos th
  o.a = b_a;
  return o;
}

// The VJP at (r, theta) given st = sin theta, ct = cos theta.
GG_HD inline RhsBar kerr_rhs_vjp_sc(double m, double a, double r, double st, double ct, d4 v, d4 lam) {
  return kerr_rhs_vjp_pt(kerr_point(m, a, r, st, ct), v, lam);
}

GG_HD inline RhsBar kerr_rhs_vjp(double m, double a, double r, double th, d4 v, d4 lam) {
  return kerr_rhs_vjp_sc(m, a, r, sin(th), cos(th), v, lam);
}

// sin and cos of theta0 + d from s0 = sin theta0, c0 = cos theta0 by the
// angle-addition formulas, with sin d and cos d from their Taylor series
// through d^9 and d^10 (truncation below 2^-60 relative for |d| <= 1/16);
// false when |d| > 1/16.  The adjoint's stage angles differ from the
// attempt's start angle by h a_i v^theta, nearly always far below that.
GG_HD inline bool sincos_near(double s0, double c0, double d, double &s, double &c) {
  GG_CONTRACT_FAST
  if (!(fabs(d) <= 0.0625)) return false;
  const double d2 = d * d;
  const double sd = d * (1.0 + d2 * (-1.0 / 6 + d2 * (1.0 / 120 + d2 * (-1.0 / 5040 + d2 * (1.0 / 362880)))));
  const double cm = d2 * (-0.5 + d2 * (1.0 / 24 + d2 * (-1.0 / 720 + d2 * (1.0 / 40320 + d2 * (-1.0 / 3628800)))));
  s = s0 + (s0 * cm + c0 * sd);                      // cos d = 1 + cm
  c = c0 + (c0 * cm - s0 * sd);
  return true;
}

}  // namespace gg
