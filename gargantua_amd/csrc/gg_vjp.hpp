// gg_vjp.hpp -- vector-Jacobian product of the Kerr geodesic RHS, reverse
// mode written out by hand.
//
// a^mu = -Gamma^mu_ab(r, theta; a) v^a v^b (metric.rs:37-52 with the
// Christoffel symbols of kerr.rs:100-424).  For a cotangent w the adjoint
// needs w . a and its gradient with respect to r, theta, the spin a and v.
// With C_i(v, w) the coefficient of the i-th distinct symbol in w . a, the
// value is sum_i C_i Gamma_i; the v-gradient is sum_i Gamma_i dC_i/dv
// (exact: a is quadratic in v); the (r, theta, a)-gradient comes from one
// backward sweep through the symbol algebra, seeded with Gamma_i-bar = C_i,
// which costs about twice the forward algebra instead of three tangent
// passes.  Plain C++ on value types: the same source compiles for the GPU
// and for a host-side check.
#pragma once

#include <math.h>

#include "gg_types.hpp"

namespace gg {

// Coefficient of symbol i in w . a.  Symbol order:
// 001 002 013 023 100 111 122 133 112 103 200 211 222 233 212 203 301 302 313 323
GG_HD inline double gamma_coef(int i, d4 v, d4 w) {
  GG_CONTRACT_FAST
  switch (i) {
    case 0: return -2.0 * w.x * v.x * v.y;
    case 1: return -2.0 * w.x * v.x * v.z;
    case 2: return -2.0 * w.x * v.y * v.w;
    case 3: return -2.0 * w.x * v.z * v.w;
    case 4: return -w.y * v.x * v.x;
    case 5: return -w.y * v.y * v.y;
    case 6: return -w.y * v.z * v.z;
    case 7: return -w.y * v.w * v.w;
    case 8: return -2.0 * w.y * v.y * v.z;
    case 9: return -2.0 * w.y * v.x * v.w;
    case 10: return -w.z * v.x * v.x;
    case 11: return -w.z * v.y * v.y;
    case 12: return -w.z * v.z * v.z;
    case 13: return -w.z * v.w * v.w;
    case 14: return -2.0 * w.z * v.y * v.z;
    case 15: return -2.0 * w.z * v.x * v.w;
    case 16: return -2.0 * w.w * v.x * v.y;
    case 17: return -2.0 * w.w * v.x * v.z;
    case 18: return -2.0 * w.w * v.y * v.w;
    default: return -2.0 * w.w * v.z * v.w;
  }
}

// gv += Gamma_i dC_i/dv.
GG_HD inline void gamma_coef_dv(int i, double g, d4 v, d4 w, d4 &gv) {
  GG_CONTRACT_FAST
  switch (i) {
    case 0: { double t = -2.0 * w.x * g; gv.x += t * v.y; gv.y += t * v.x; break; }
    case 1: { double t = -2.0 * w.x * g; gv.x += t * v.z; gv.z += t * v.x; break; }
    case 2: { double t = -2.0 * w.x * g; gv.y += t * v.w; gv.w += t * v.y; break; }
    case 3: { double t = -2.0 * w.x * g; gv.z += t * v.w; gv.w += t * v.z; break; }
    case 4: gv.x += -2.0 * w.y * g * v.x; break;
    case 5: gv.y += -2.0 * w.y * g * v.y; break;
    case 6: gv.z += -2.0 * w.y * g * v.z; break;
    case 7: gv.w += -2.0 * w.y * g * v.w; break;
    case 8: { double t = -2.0 * w.y * g; gv.y += t * v.z; gv.z += t * v.y; break; }
    case 9: { double t = -2.0 * w.y * g; gv.x += t * v.w; gv.w += t * v.x; break; }
    case 10: gv.x += -2.0 * w.z * g * v.x; break;
    case 11: gv.y += -2.0 * w.z * g * v.y; break;
    case 12: gv.z += -2.0 * w.z * g * v.z; break;
    case 13: gv.w += -2.0 * w.z * g * v.w; break;
    case 14: { double t = -2.0 * w.z * g; gv.y += t * v.z; gv.z += t * v.y; break; }
    case 15: { double t = -2.0 * w.z * g; gv.x += t * v.w; gv.w += t * v.x; break; }
    case 16: { double t = -2.0 * w.w * g; gv.x += t * v.y; gv.y += t * v.x; break; }
    case 17: { double t = -2.0 * w.w * g; gv.x += t * v.z; gv.z += t * v.x; break; }
    case 18: { double t = -2.0 * w.w * g; gv.y += t * v.w; gv.w += t * v.y; break; }
    default: { double t = -2.0 * w.w * g; gv.z += t * v.w; gv.w += t * v.z; break; }
  }
}

struct RhsBar {
  double r, th, a;   // w^T da/dr, w^T da/dtheta, w^T da/dspin
  d4 v;              // w^T da/dv
};

GG_HD inline RhsBar kerr_rhs_vjp(double m, double a, double r, double th, d4 v, d4 w) {
  GG_CONTRACT_FAST
  // ---- forward: the symbol algebra of rhs64f_kerr, intermediates kept ----
  const double st = sin(th), ct = cos(th);
  const double tm = 2.0 * m;
  const double a2 = a * a, r2 = r * r, sin2 = st * st, sc = st * ct, ct2 = ct * ct;
  const double sigma = r2 + a2 * ct2;
  const double delta = r2 - tm * r + a2;
  const double is = 1.0 / sigma, id = 1.0 / delta;
  const double is2 = is * is, isd = is * id;
  const double dsig = -2.0 * a2 * sc;             // d sigma / d theta
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
  const double isn = 1.0 / sin2;
  const double gpp = X * isd * isn;               // g^phi phi
  const double drr_t = dsig * id;
  const double hrr = 0.5 * delta * is;            // g^rr / 2
  const double N = 2.0 * r * delta - sigma * (2.0 * r - tm);
  const double id2 = id * id;
  const double drr_r = N * id2;
  const double hth = 0.5 * is;                    // g^theta theta / 2

  const double G[20] = {
      0.5 * (gtt * dtt_r + gtp * dtp_r), 0.5 * (gtt * dtt_t + gtp * dtp_t),
      0.5 * (gtt * dtp_r + gtp * dpp_r), 0.5 * (gtt * dtp_t + gtp * dpp_t),
      -hrr * dtt_r, hrr * drr_r, -hrr * (2.0 * r), -hrr * dpp_r, hrr * drr_t, -hrr * dtp_r,
      -hth * dtt_t, -hth * drr_t, hth * dsig, -hth * dpp_t, hth * (2.0 * r), -hth * dtp_t,
      0.5 * (gtp * dtt_r + gpp * dtp_r), 0.5 * (gtp * dtt_t + gpp * dtp_t),
      0.5 * (gtp * dtp_r + gpp * dpp_r), 0.5 * (gtp * dtp_t + gpp * dpp_t)};
  double C[20];
  RhsBar o;
  o.v = d4{0.0, 0.0, 0.0, 0.0};
  for (int i = 0; i < 20; ++i) {
    C[i] = gamma_coef(i, v, w);
    gamma_coef_dv(i, G[i], v, w, o.v);
  }

  // ---- backward: Gamma_i-bar = C_i ----
  const double b_gtt = 0.5 * (C[0] * dtt_r + C[1] * dtt_t + C[2] * dtp_r + C[3] * dtp_t);
  const double b_gtp = 0.5 * (C[0] * dtp_r + C[1] * dtp_t + C[2] * dpp_r + C[3] * dpp_t + C[16] * dtt_r +
                              C[17] * dtt_t + C[18] * dtp_r + C[19] * dtp_t);
  const double b_gpp = 0.5 * (C[16] * dtp_r + C[17] * dtp_t + C[18] * dpp_r + C[19] * dpp_t);
  const double b_dtt_r = 0.5 * (C[0] * gtt + C[16] * gtp) - C[4] * hrr;
  const double b_dtt_t = 0.5 * (C[1] * gtt + C[17] * gtp) - C[10] * hth;
  const double b_dtp_r = 0.5 * (C[0] * gtp + C[2] * gtt + C[16] * gpp + C[18] * gtp) - C[9] * hrr;
  const double b_dtp_t = 0.5 * (C[1] * gtp + C[3] * gtt + C[17] * gpp + C[19] * gtp) - C[15] * hth;
  const double b_dpp_r = 0.5 * (C[2] * gtp + C[18] * gpp) - C[7] * hrr;
  const double b_dpp_t = 0.5 * (C[3] * gtp + C[19] * gpp) - C[13] * hth;
  const double b_hrr = -C[4] * dtt_r + C[5] * drr_r - C[6] * (2.0 * r) - C[7] * dpp_r + C[8] * drr_t - C[9] * dtp_r;
  const double b_drr_r = C[5] * hrr;
  const double b_drr_t = C[8] * hrr - C[11] * hth;
  const double b_hth =
      -C[10] * dtt_t - C[11] * drr_t + C[12] * dsig - C[13] * dpp_t + C[14] * (2.0 * r) - C[15] * dtp_t;
  double b_r = -2.0 * C[6] * hrr + 2.0 * C[14] * hth;
  double b_dsig = C[12] * hth;

  double b_is = 0.5 * b_hth;                                  // hth = is / 2
  const double b_N = b_drr_r * id2;                           // drr_r = N id^2
  double b_id = b_drr_r * N * 2.0 * id;
  b_r += b_N * (2.0 * delta - 2.0 * sigma);                   // N = 2 r delta - sigma (2 r - 2m)
  double b_delta = b_N * 2.0 * r;
  double b_sigma = -b_N * (2.0 * r - tm);
  b_delta += 0.5 * b_hrr * is;                                // hrr = delta is / 2
  b_is += 0.5 * b_hrr * delta;
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

}  // namespace gg
