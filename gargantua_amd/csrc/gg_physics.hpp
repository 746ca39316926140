// gg_physics.hpp -- device-side geodesic physics for gfx950.
//
// Two right-hand sides per metric:
//   rhs64_*  FP64, no FMA contraction, each Christoffel symbol with the
//            expression tree of the reference and the contraction applied in
//            the reference's dense mu, alpha, beta order with its exact-zero
//            terms skipped (skipping x - 0*v*v changes no bit).  This is the
//            parity path: its results differ from the reference arithmetic
//            only where the device sin/cos/pow round differently from libm.
//   rhs32_*  FP32, contracted, three reciprocals instead of fifteen
//            divisions, symmetric pairs merged.  Used by the mixed kernel on
//            attempts where the step-size controller is saturated (see
//            gg_rk45.hpp), where its rounding cannot change a decision.
#pragma once

#ifndef __HIPCC_RTC__  // hipRTC supplies the runtime and fixed-width types itself
#include <hip/hip_runtime.h>
#include <stdint.h>
#endif

#include "gargantua.h"
#include "gg_types.hpp"
#include "gg_crtrig.hpp"

namespace gg {

// Per-scene constants, computed once on the host (with the host libm, so
// that tan/sin/cos of the camera are the values the reference computes).
struct DevScene {
  // metric
  int32_t metric_kind;
  int32_t disk_kind;
  int32_t err_mode;
  uint32_t max_steps;
  double mass, spin, rs, r_h;
  double r_capture;      // r_h * (1 + horizon_rel) + horizon_buffer
  double max_radius;
  // integrator
  double h0, min_step, max_step, tol, rtol, atol;
  double max_step5;       // max_step^5: the controller's clamp test
  // camera
  double cam_pos[4];
  double fwd[3], right[3], up[3];
  double half_w, half_h;
  double rhat[3], thhat[3], phhat[3];
  double g_tt, g_rr, g_thth, g_phph;            // metric at the camera
  double sqrt_g_rr, sqrt_g_thth, sqrt_g_phph;
  double sqrt_neg_gtt_obs;                      // tracer.rs:93-94
  // sub-pixel grid
  uint32_t sub_w, sub_h;
  uint32_t jitter, tiled;                        // tiled: 8x8 ray tiles
  uint64_t seed;
  // disk
  double r_in, r_out, scale_height, density_0, absorption, q, edge_width;
  // mixed precision policy: FP32 stages on outgoing rays beyond fp32_min_r,
  // on inbound rays beyond fp32_far_r, and on inbound rays beyond fp32_min_r
  // whose squared impact parameter (Q + L_z^2)/E^2 is at least fp32_b2_min
  double fp32_min_r, fp32_far_r, fp32_b2_min;
  // Hittable scene (gg_scene.hpp): primitives, BVH nodes, and -g_tt at
  // the camera for the static-emitter redshift (tracer.rs:216)
  const void *objs;
  const void *nodes;
  double g_tt_obs;
  double scene_rmin;      // no primitive reaches inside this radius (chords within it cannot hit)
  // batch bookkeeping
  uint64_t ray_begin;     // first global ray index of this scene
  uint64_t out_index;     // which output set (batch)
};

// ---------------------------------------------------------------- FP64 ----

__device__ inline double rmax1e10(double r) { return r >= 1e-10 ? r : 1e-10; }  // r.max(1e-10)

// Kerr::metric_tensor (kerr.rs:61-94), Schwarzschild (schwarzschild.rs:26-43).
struct Gmn { double tt, rr, thth, phph, tph; };

template <int METRIC>
__device__ inline Gmn metric64(const DevScene &S, double r_in, double theta) {
#pragma clang fp contract(off)
  Gmn g;
  double r = rmax1e10(r_in);
  if (METRIC == 1) {
    double m = S.mass, a = S.spin;
    double a2 = a * a;
    double st, ct;
    sincos_cr(theta, &st, &ct);   // the host libm's values (correctly rounded)
    double sigma = r * r + a * a * ct * ct;
    double delta = r * r - 2.0 * m * r + a * a;
    double sin2 = st * st;
    g.tt = -(1.0 - 2.0 * m * r / sigma);
    g.rr = sigma / delta;
    g.thth = sigma;
    g.phph = (r * r + a2 + 2.0 * m * a2 * r * sin2 / sigma) * sin2;
    g.tph = -2.0 * m * a * r * sin2 / sigma;
  } else {
    double f = 1.0 - S.rs / r;
    double st, ct;
    sincos_cr(theta, &st, &ct);
    g.tt = -f;
    g.rr = 1.0 / f;
    g.thth = r * r;
    g.phph = r * r * st * st;
    g.tph = 0.0;
  }
  return g;
}

// Kerr RHS: Christoffels of kerr.rs:100-424, contracted as metric.rs:37-52.
__device__ inline d4 rhs64_kerr(const DevScene &S, double r_in, double theta, d4 v) {
#pragma clang fp contract(off)
  double r = rmax1e10(r_in);
  const double m = S.mass, a = S.spin;
  double a2 = a * a;
  double st, cs;
  sincos_cr(theta, &st, &cs);   // the host libm's values (correctly rounded)
  double sigma = r * r + a * a * cs * cs;
  double delta = r * r - 2.0 * m * r + a * a;
  double ct = cs;  // cos(theta) evaluated twice in the reference; same value
  double sin2 = st * st;
  double sigma2 = sigma * sigma;
  double r2 = r * r;
  double dsigma_dr = 2.0 * r;
  double dsigma_dtheta = -2.0 * a2 * st * ct;
  double ddelta_dr = 2.0 * r - 2.0 * m;
  double r2pa2 = r2 + a2;
  double big_a = r2pa2 * r2pa2 - a2 * delta * sin2;
  double g_inv_tt = -big_a / (sigma * delta);
  double g_inv_tphi = -2.0 * m * a * r / (sigma * delta);
  double g_inv_rr = delta / sigma;
  double g_inv_thth = 1.0 / sigma;
  double g_inv_phph = (delta - a2 * sin2) / (sigma * delta * sin2);
  double dg_tt_dr = 2.0 * m / sigma - 2.0 * m * r * dsigma_dr / sigma2;
  double dg_tt_dth = -2.0 * m * r * dsigma_dtheta / sigma2;
  double dg_rr_dr = (dsigma_dr * delta - sigma * ddelta_dr) / (delta * delta);
  double dg_rr_dth = dsigma_dtheta / delta;
  double dg_thth_dr = dsigma_dr;
  double dg_thth_dth = dsigma_dtheta;
  double big_p = 2.0 * m * a2 * r;
  double dg_phph_dr = 2.0 * r * sin2 + sin2 * sin2 * 2.0 * m * a2 * (sigma - 2.0 * r2) / sigma2;
  double dg_phph_dth = 2.0 * r2pa2 * st * ct +
                       big_p * (4.0 * st * st * st * ct * sigma - sin2 * sin2 * dsigma_dtheta) / sigma2;
  double dg_tphi_dr = -2.0 * m * a * sin2 * (sigma - r * dsigma_dr) / sigma2;
  double dg_tphi_dth = -2.0 * m * a * r * (2.0 * st * ct * sigma - sin2 * dsigma_dtheta) / sigma2;

  double G001 = 0.5 * (g_inv_tt * dg_tt_dr + g_inv_tphi * dg_tphi_dr);
  double G002 = 0.5 * (g_inv_tt * dg_tt_dth + g_inv_tphi * dg_tphi_dth);
  double G013 = 0.5 * (g_inv_tt * dg_tphi_dr + g_inv_tphi * dg_phph_dr);
  double G023 = 0.5 * (g_inv_tt * dg_tphi_dth + g_inv_tphi * dg_phph_dth);
  double G100 = -0.5 * g_inv_rr * dg_tt_dr;
  double G111 = 0.5 * g_inv_rr * dg_rr_dr;
  double G122 = -0.5 * g_inv_rr * dg_thth_dr;
  double G133 = -0.5 * g_inv_rr * dg_phph_dr;
  double G112 = 0.5 * g_inv_rr * dg_rr_dth;
  double G103 = -0.5 * g_inv_rr * dg_tphi_dr;
  double G200 = -0.5 * g_inv_thth * dg_tt_dth;
  double G211 = -0.5 * g_inv_thth * dg_rr_dth;
  double G222 = 0.5 * g_inv_thth * dg_thth_dth;
  double G233 = -0.5 * g_inv_thth * dg_phph_dth;
  double G212 = 0.5 * g_inv_thth * dg_thth_dr;
  double G203 = -0.5 * g_inv_thth * dg_tphi_dth;
  double G301 = 0.5 * (g_inv_tphi * dg_tt_dr + g_inv_phph * dg_tphi_dr);
  double G302 = 0.5 * (g_inv_tphi * dg_tt_dth + g_inv_phph * dg_tphi_dth);
  double G313 = 0.5 * (g_inv_tphi * dg_tphi_dr + g_inv_phph * dg_phph_dr);
  double G323 = 0.5 * (g_inv_tphi * dg_tphi_dth + g_inv_phph * dg_phph_dth);

  const double v0 = v.x, v1 = v.y, v2 = v.z, v3 = v.w;
  d4 acc;
  double t;
  // mu = t: (a,b) = 01 02 10 13 20 23 31 32
  t = 0.0;
  t = t - G001 * v0 * v1; t = t - G002 * v0 * v2; t = t - G001 * v1 * v0; t = t - G013 * v1 * v3;
  t = t - G002 * v2 * v0; t = t - G023 * v2 * v3; t = t - G013 * v3 * v1; t = t - G023 * v3 * v2;
  acc.x = t;
  // mu = r: 00 03 11 12 21 22 30 33
  t = 0.0;
  t = t - G100 * v0 * v0; t = t - G103 * v0 * v3; t = t - G111 * v1 * v1; t = t - G112 * v1 * v2;
  t = t - G112 * v2 * v1; t = t - G122 * v2 * v2; t = t - G103 * v3 * v0; t = t - G133 * v3 * v3;
  acc.y = t;
  // mu = theta: 00 03 11 12 21 22 30 33
  t = 0.0;
  t = t - G200 * v0 * v0; t = t - G203 * v0 * v3; t = t - G211 * v1 * v1; t = t - G212 * v1 * v2;
  t = t - G212 * v2 * v1; t = t - G222 * v2 * v2; t = t - G203 * v3 * v0; t = t - G233 * v3 * v3;
  acc.z = t;
  // mu = phi: 01 02 10 13 20 23 31 32
  t = 0.0;
  t = t - G301 * v0 * v1; t = t - G302 * v0 * v2; t = t - G301 * v1 * v0; t = t - G313 * v1 * v3;
  t = t - G302 * v2 * v0; t = t - G323 * v2 * v3; t = t - G313 * v3 * v1; t = t - G323 * v3 * v2;
  acc.w = t;
  return acc;
}

// Schwarzschild RHS override (schwarzschild.rs:99-140).
__device__ inline d4 rhs64_schw(const DevScene &S, double r_in, double theta, d4 v) {
#pragma clang fp contract(off)
  double r = rmax1e10(r_in);
  double st, ct;
  sincos_cr(theta, &st, &ct);   // the host libm's values (correctly rounded)
  const double rs = S.rs;
  double f_denom = r * (r - rs);
  double rmrs = r - rs;
  double inv_r = 1.0 / r;
  double g_t_tr = rs / (2.0 * f_denom);
  double g_r_tt = rs * rmrs / (2.0 * r * r * r);
  double g_r_rr = -rs / (2.0 * f_denom);
  double g_r_thth = -rmrs;
  double g_r_phph = -rmrs * st * st;
  double g_th_phph = -st * ct;
  double g_ph_thph = ct / st;
  const double vt = v.x, vr = v.y, vth = v.z, vph = v.w;
  d4 acc;
  acc.x = -2.0 * g_t_tr * vt * vr;
  acc.y = -g_r_tt * vt * vt - g_r_rr * vr * vr - g_r_thth * vth * vth - g_r_phph * vph * vph;
  acc.z = -2.0 * inv_r * vr * vth - g_th_phph * vph * vph;
  acc.w = -2.0 * inv_r * vr * vph - 2.0 * g_ph_thph * vth * vph;
  return acc;
}

template <int METRIC>
__device__ inline d4 rhs64(const DevScene &S, double r, double theta, d4 v) {
  if (METRIC == 1) return rhs64_kerr(S, r, theta, v);
  return rhs64_schw(S, r, theta, v);
}

// ----------------------------------------------------------- fast FP64 ----

// Inverse metric and the metric's r and theta derivatives at one point of
// a stationary, axisymmetric metric with g_t_phi (Kerr, kerr.rs:61-94).
template <class F>
struct KerrDeriv {
  F gi_tt, gi_tph, gi_rr, gi_thth, gi_phph;
  F tt_r, tt_th, rr_r, rr_th, thth_r, thth_th, pp_r, pp_th, tp_r, tp_th;
};

// Geodesic acceleration -Gamma^mu_ab v^a v^b with the index lowered,
// -g^{mu nu} (d_a g_{nu b} v^a v^b - 1/2 d_nu g_{ab} v^a v^b): the same
// quantity as the Christoffel contraction of metric.rs:37-52 (Gamma^mu_ab =
// 1/2 g^{mu nu} (d_a g_{nu b} + d_b g_{nu a} - d_nu g_{ab})), in about half
// the operations.  With the metric depending on r and theta only,
// d_a g_{nu b} v^a v^b = v^r w_nu + v^theta u_nu for w_nu = d_r g_{nu b} v^b
// and u_nu = d_theta g_{nu b} v^b, and 1/2 d_nu g_{ab} v^a v^b is
// 1/2 w.v for nu = r, 1/2 u.v for nu = theta, 0 for t and phi.
__device__ inline float fma_t(float a, float b, float c) { return __builtin_fmaf(a, b, c); }
__device__ inline double fma_t(double a, double b, double c) { return __builtin_fma(a, b, c); }

// (Every fused operation is written out, so that each kernel instantiation
// rounds the same way: the FP64 kernels' results do not depend on whether a
// scene's constants are scalar or per lane.)
template <class F, class V>
__device__ inline V kerr_accel(const KerrDeriv<F> &D, V v) {
#pragma clang fp contract(off)
  const F vt = v.x, vr = v.y, vth = v.z, vph = v.w;
  const F w_t = fma_t(D.tt_r, vt, D.tp_r * vph), w_p = fma_t(D.tp_r, vt, D.pp_r * vph);
  const F w_r = D.rr_r * vr, w_th = D.thth_r * vth;
  const F u_t = fma_t(D.tt_th, vt, D.tp_th * vph), u_p = fma_t(D.tp_th, vt, D.pp_th * vph);
  const F u_r = D.rr_th * vr, u_th = D.thth_th * vth;
  const F wv = fma_t(w_t, vt, fma_t(w_p, vph, fma_t(w_r, vr, w_th * vth)));
  const F uv = fma_t(u_t, vt, fma_t(u_p, vph, fma_t(u_r, vr, u_th * vth)));
  const F f_t = fma_t(vr, w_t, vth * u_t), f_p = fma_t(vr, w_p, vth * u_p);
  const F f_r = fma_t(vr, w_r, fma_t(vth, u_r, F(-0.5) * wv));
  const F f_th = fma_t(vr, w_th, fma_t(vth, u_th, F(-0.5) * uv));
  V acc;
  acc.x = -fma_t(D.gi_tt, f_t, D.gi_tph * f_p);
  acc.y = -(D.gi_rr * f_r);
  acc.z = -(D.gi_thth * f_th);
  acc.w = -fma_t(D.gi_tph, f_t, D.gi_phph * f_p);
  return acc;
}

// 1/x to ~1 ulp: v_rcp_f64 refined by two Newton steps (x is a positive,
// normal quantity here: Sigma, Delta or sin^2 theta).
__device__ inline double rcp64(double x) {
  double r = __builtin_amdgcn_rcp(x);
  double e = __builtin_fma(-x, r, 1.0);
  r = __builtin_fma(r, e, r);
  e = __builtin_fma(-x, r, 1.0);
  return __builtin_fma(r, e, r);
}

// Kerr RHS in FP64 with the algebra of rhs32_kerr: the metric derivatives
// and inverse metric of kerr.rs:100-424 through 1/Sigma, 1/Delta and
// 1/sin^2 instead of fifteen divisions, explicit FMAs, and the contraction
// with lowered index (kerr_accel) instead of twenty Christoffel symbols.
// Agrees with rhs64_kerr to rounding (a few ulp of the largest term); used
// by the fast-FP64 and mixed kernels.
__device__ inline d4 rhs64f_kerr(double m, double a, double r_in, double theta, d4 v) {
#pragma clang fp contract(off)
  const double r = rmax1e10(r_in);
  double st, ct;
  sincos(theta, &st, &ct);
  const double a2 = a * a;
  const double r2 = r * r;
  const double sin2 = st * st;
  const double sc = st * ct;
  const double sigma = __builtin_fma(a2 * ct, ct, r2);
  const double twom = 2.0 * m;
  const double delta = __builtin_fma(-twom, r, r2 + a2);
  const double is = rcp64(sigma);
  const double id = rcp64(delta);
  const double is2 = is * is;
  const double isd = is * id;
  const double dsig_th = (-2.0 * a2) * sc;
  const double r2pa2 = r2 + a2;
  const double big_a = __builtin_fma(r2pa2, r2pa2, -((a2 * delta) * sin2));
  const double two_r = 2.0 * r;
  const double s4 = sin2 * sin2;
  KerrDeriv<double> D;
  D.gi_tt = -big_a * isd;
  D.gi_tph = -(twom * a * r) * isd;
  D.gi_rr = delta * is;
  D.gi_thth = is;
  D.gi_phph = __builtin_fma(-a2, sin2, delta) * isd * rcp64(sin2);
  const double sig_m2r2 = __builtin_fma(-2.0, r2, sigma);
  D.tt_r = twom * sig_m2r2 * is2;
  D.tt_th = -(twom * r) * dsig_th * is2;
  D.rr_r = __builtin_fma(two_r, delta, -(sigma * (two_r - twom))) * id * id;
  D.rr_th = dsig_th * id;
  D.thth_r = two_r;
  D.thth_th = dsig_th;
  const double big_p = twom * a2 * r;
  D.pp_r = __builtin_fma(two_r, sin2, s4 * (twom * a2) * sig_m2r2 * is2);
  D.pp_th = __builtin_fma(2.0 * r2pa2, sc, big_p * __builtin_fma(4.0 * sin2 * sc, sigma, -(s4 * dsig_th)) * is2);
  D.tp_r = -(twom * a) * sin2 * sig_m2r2 * is2;
  D.tp_th = -(twom * a * r) * __builtin_fma(2.0 * sc, sigma, -(sin2 * dsig_th)) * is2;
  return kerr_accel(D, v);
}

__device__ inline d4 rhs64f_schw(double rs, double r_in, double theta, d4 v) {
#pragma clang fp contract(off)
  const double r = rmax1e10(r_in);
  double st, ct;
  sincos(theta, &st, &ct);
  const double ir = rcp64(r);
  const double rmrs = r - rs;
  const double ifd = rcp64(r * rmrs);
  const double g_t_tr = 0.5 * rs * ifd;
  const double g_r_tt = 0.5 * rs * rmrs * ir * ir * ir;
  d4 acc;
  acc.x = -2.0 * g_t_tr * v.x * v.y;
  acc.y = __builtin_fma(-g_r_tt * v.x, v.x,
                        __builtin_fma(g_t_tr * v.y, v.y, rmrs * __builtin_fma(v.z, v.z, st * st * v.w * v.w)));
  acc.z = __builtin_fma(-2.0 * ir * v.y, v.z, st * ct * v.w * v.w);
  acc.w = __builtin_fma(-2.0 * ir * v.y, v.w, -(2.0 * (ct * rcp64(st)) * v.z * v.w));
  return acc;
}

template <int METRIC>
__device__ inline d4 rhs64f(const DevScene &S, double r, double theta, d4 v) {
  if (METRIC == 1) return rhs64f_kerr(S.mass, S.spin, r, theta, v);
  return rhs64f_schw(S.rs, r, theta, v);
}

// ---------------------------------------------------------------- FP32 ----

__device__ inline float frcp(float x) { return __builtin_amdgcn_rcpf(x); }

// Kerr geodesic acceleration in FP32: the metric derivatives and inverse
// metric with the divisions folded into 1/Sigma, 1/Delta, 1/sin^2, and the
// contraction with lowered index (kerr_accel).
// psi = theta - pi/2, formed in FP64 by the caller: near the equatorial
// plane FP32 keeps full relative precision in cos(theta) = -sin(psi), which
// a rounded FP32 theta would not (its 6e-8 absolute error feeds straight
// into the theta acceleration of rays crossing the disk plane).
__device__ inline f4 rhs32_kerr(float m, float a, float r, float psi, f4 v) {
#pragma clang fp contract(fast)
  float sp, cp;
  sincosf(psi, &sp, &cp);
  const float st = cp, ct = -sp;
  float a2 = a * a;
  float r2 = r * r;
  float sin2 = st * st;
  float sc = st * ct;
  float sigma = r2 + a2 * ct * ct;
  float delta = r2 - 2.0f * m * r + a2;
  float is = frcp(sigma);
  float id = frcp(delta);
  float is2 = is * is;
  float isd = is * id;
  float dsig_th = -2.0f * a2 * sc;
  float r2pa2 = r2 + a2;
  float big_a = r2pa2 * r2pa2 - a2 * delta * sin2;
  float twom = 2.0f * m;
  KerrDeriv<float> D;
  D.gi_tt = -big_a * isd;
  D.gi_tph = -twom * a * r * isd;
  D.gi_rr = delta * is;
  D.gi_thth = is;
  D.gi_phph = (delta - a2 * sin2) * isd * frcp(sin2);
  float sig_m2r2 = sigma - 2.0f * r2;
  D.tt_r = twom * sig_m2r2 * is2;
  D.tt_th = -twom * r * dsig_th * is2;
  D.rr_r = (2.0f * r * delta - sigma * (2.0f * r - twom)) * id * id;
  D.rr_th = dsig_th * id;
  D.thth_r = 2.0f * r;
  D.thth_th = dsig_th;
  float big_p = twom * a2 * r;
  D.pp_r = 2.0f * r * sin2 + sin2 * sin2 * twom * a2 * sig_m2r2 * is2;
  D.pp_th = 2.0f * r2pa2 * sc + big_p * (4.0f * sin2 * sc * sigma - sin2 * sin2 * dsig_th) * is2;
  D.tp_r = -twom * a * sin2 * sig_m2r2 * is2;
  D.tp_th = -twom * a * r * (2.0f * sc * sigma - sin2 * dsig_th) * is2;
  return kerr_accel(D, v);
}

__device__ inline f4 rhs32_schw(float rs, float r, float psi, f4 v) {
#pragma clang fp contract(fast)
  float sp, cp;
  sincosf(psi, &sp, &cp);
  const float st = cp, ct = -sp;
  float ir = frcp(r);
  float rmrs = r - rs;
  float ifd = frcp(r * rmrs);
  float g_t_tr = 0.5f * rs * ifd;
  float g_r_tt = 0.5f * rs * rmrs * ir * ir * ir;
  float g_r_rr = -g_t_tr;
  f4 acc;
  acc.x = -2.0f * g_t_tr * v.x * v.y;
  acc.y = -g_r_tt * v.x * v.x - g_r_rr * v.y * v.y + rmrs * (v.z * v.z + st * st * v.w * v.w);
  acc.z = -2.0f * ir * v.y * v.z + st * ct * v.w * v.w;
  acc.w = -2.0f * ir * v.y * v.w - 2.0f * (ct * frcp(st)) * v.z * v.w;
  return acc;
}

template <int METRIC>
__device__ inline f4 rhs32f(const DevScene &S, float r, float psi, f4 v) {
  if (METRIC == 1) return rhs32_kerr((float)S.mass, (float)S.spin, r, psi, v);
  return rhs32_schw((float)S.rs, r, psi, v);
}

template <int METRIC>
__device__ inline f4 rhs32(const DevScene &S, double r, double theta, f4 v) {
  const float psi = (float)(theta - 1.5707963267948966);
  if (METRIC == 1) return rhs32_kerr((float)S.mass, (float)S.spin, (float)r, psi, v);
  return rhs32_schw((float)S.rs, (float)r, psi, v);
}

// ------------------------------------------------------ redshift & disk ----

// orbital_four_velocity + circular_orbit_velocity (kerr.rs:439-449,
// schwarzschild.rs:146-153, metric.rs:82-94), then redshift_factor
// (tracer.rs:158-183).
template <int METRIC>
__device__ inline double redshift64(const DevScene &S, double r, double theta, d4 p) {
#pragma clang fp contract(off)
  Gmn g = metric64<METRIC>(S, r, theta);
  double p_t = g.tt * p.x + g.tph * p.w;
  double e_obs = -p_t / S.sqrt_neg_gtt_obs;
  e_obs = e_obs >= 0.0 ? e_obs : 0.0;
  bool have_orbit = false;
  double omega = 0.0;
  const double m = S.mass;
  if (METRIC == 1) {
    if (r > 0.0) {
      double sqrt_m = sqrt(m);
      omega = sqrt_m / (pow(r, 1.5) + S.spin * sqrt_m);
      have_orbit = true;
    }
  } else {
    if (r > 3.0 * m) {
      omega = sqrt(m / (r * r * r));
      have_orbit = true;
    }
  }
  double e_emit = 0.0;
  bool orbit_ok = false;
  if (have_orbit) {
    // circular_orbit_velocity re-evaluates the metric at the same point.
    double denom = -(g.tt + 2.0 * g.tph * omega + g.phph * omega * omega);
    if (denom > 0.0 && isfinite(denom)) {
      double u_t = 1.0 / sqrt(denom);
      double u_ph = omega * u_t;
      double p_phi = g.tph * p.x + g.phph * p.w;
      e_emit = -(p_t * u_t + p_phi * u_ph);
      e_emit = e_emit >= 0.0 ? e_emit : 0.0;
      orbit_ok = true;
    }
  }
  if (!orbit_ok) {
    double g_tt_emit = -g.tt;
    e_emit = g_tt_emit > 0.0 ? -p_t / sqrt(g_tt_emit) : 0.0;
    e_emit = e_emit >= 0.0 ? e_emit : 0.0;
  }
  return e_emit > 0.0 ? e_obs / e_emit : 0.0;
}

// AccretionDisk::density (disk.rs:40-59).
__device__ inline double disk_density(const DevScene &S, double r, double theta) {
#pragma clang fp contract(off)
  double st, ct;
  sincos_cr(theta, &st, &ct);
  double cap_r = r * fabs(st);
  double z = r * ct;
  if (cap_r < S.r_in || cap_r > S.r_out) return 0.0;
  double t = S.r_in / cap_r;
  double radial = t * t;
  double vertical = exp(-z * z / (2.0 * S.scale_height * S.scale_height));
  return S.density_0 * radial * vertical;
}

// AccretionDisk::emission given the density just computed (disk.rs:65-81).
__device__ inline double disk_emission(const DevScene &S, double rho, double r, double theta) {
#pragma clang fp contract(off)
  if (rho == 0.0) return 0.0;
  double st, ct;
  sincos_cr(theta, &st, &ct);
  double cap_r = r * fabs(st);
  if (cap_r < S.r_in) return 0.0;
  double temp_proxy = pow(S.r_in / cap_r, 0.75);
  return rho * rho * temp_proxy;
}

// Constants of motion of a geodesic at (pos, vel): E = -p_t, L_z = p_phi,
// Carter's Q = p_theta^2 + cos^2(theta) (a^2 (mu^2 - E^2) + L_z^2 / sin^2(theta))
// and the mass shell mu^2 = -g(v, v) (0 for a null ray; the camera rays of
// ray.rs:51-53 drop g_tphi, so Kerr camera rays carry mu^2 = -2 g_tphi k^t k^phi).
// Exact for the geodesic flow; the integrator's forward-Euler positions
// make them drift at O(h), which is what they diagnose.
template <int METRIC>
__device__ inline void constants_of_motion(const DevScene &S, const double pos[4], d4 v, double out[4]) {
  const Gmn g = metric64<METRIC>(S, pos[1], pos[2]);
  const double pt = g.tt * v.x + g.tph * v.w;
  const double pph = g.tph * v.x + g.phph * v.w;
  const double pth = g.thth * v.z;
  const double mu2 = -(pt * v.x + g.rr * v.y * v.y + pth * v.z + pph * v.w);
  double st, ct;
  sincos_cr(pos[2], &st, &ct);
  const double a = METRIC == 1 ? S.spin : 0.0;
  out[0] = -pt;
  out[1] = pph;
  out[2] = pth * pth + ct * ct * (a * a * (mu2 - pt * pt) + pph * pph / (st * st));
  out[3] = mu2;
}

// SplitMix64 (Vigna): n-th output (n >= 1) for `seed`.
__device__ inline uint64_t splitmix64(uint64_t seed, uint64_t n) {
  uint64_t z = seed + n * 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

// Camera::pixel_direction + GeodesicRay::from_camera (camera.rs:20-29,
// ray.rs:14-61) with the per-scene basis hoisted to the host.
__device__ inline void camera_ray(const DevScene &S, uint32_t px, uint32_t py, double pos[4], d4 &vel) {
#pragma clang fp contract(off)
  double fx = (double)px + 0.5, fy = (double)py + 0.5;
  if (S.jitter) {
    uint64_t k = (uint64_t)py * S.sub_w + px;
    fx = fx + ((double)(splitmix64(S.seed, 2 * k + 1) >> 11) * (1.0 / 9007199254740992.0) - 0.5);
    fy = fy + ((double)(splitmix64(S.seed, 2 * k + 2) >> 11) * (1.0 / 9007199254740992.0) - 0.5);
  }
  double u = fx / (double)S.sub_w * 2.0 - 1.0;
  double v = 1.0 - fy / (double)S.sub_h * 2.0;
  double uw = u * S.half_w, vh = v * S.half_h;
  double d0 = S.fwd[0] + uw * S.right[0] + vh * S.up[0];
  double d1 = S.fwd[1] + uw * S.right[1] + vh * S.up[1];
  double d2 = S.fwd[2] + uw * S.right[2] + vh * S.up[2];
  double n = sqrt(d0 * d0 + d1 * d1 + d2 * d2);
  d0 = d0 / n; d1 = d1 / n; d2 = d2 / n;
  double n_r = d0 * S.rhat[0] + d1 * S.rhat[1] + d2 * S.rhat[2];
  double n_t = d0 * S.thhat[0] + d1 * S.thhat[1] + d2 * S.thhat[2];
  double n_p = d0 * S.phhat[0] + d1 * S.phhat[1] + d2 * S.phhat[2];
  double kr = n_r / S.sqrt_g_rr;
  double kth = n_t / S.sqrt_g_thth;
  double kph = n_p / S.sqrt_g_phph;
  double spatial = S.g_rr * kr * kr + S.g_thth * kth * kth + S.g_phph * kph * kph;
  double kt = sqrt(-spatial / S.g_tt);
  for (int i = 0; i < 4; ++i) pos[i] = S.cam_pos[i];
  vel.x = kt; vel.y = kr; vel.z = kth; vel.w = kph;
}

}  // namespace gg
