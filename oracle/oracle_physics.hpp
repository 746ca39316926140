// oracle_physics.hpp -- CPU FP64 restatement of the reference geodesic path.
//
// TEST INFRASTRUCTURE ONLY.  This file is the checker that the HIP kernels of
// libgargantua are compared against; nothing in the product links it.
//
// Every function restates a reference function with the same expression
// trees (left-to-right evaluation, no FMA contraction: build with
// -ffp-contract=off), so that with T = double the arithmetic is the one the
// Rust reference performs.  The scalar type is a template parameter so the
// same code, instantiated with a forward-mode dual number, yields the FP64
// dual-number gradient (the step-size controller only ever reads values).
#pragma once

#include <cmath>
#include <cstdint>
#include <type_traits>

#include "../include/gargantua.h"

namespace oracle {

constexpr double PI = 3.14159265358979323846;
constexpr double HALF_PI = PI / 2.0;

// ---- scalar helpers: overloaded for double here, for Dual in oracle_dual.hpp
inline double val(double x) { return x; }
inline double smax(double a, double b) { return std::fmax(a, b); }  // f64::max

template <class T> struct V4 { T x[4]; T &operator[](int i) { return x[i]; } const T &operator[](int i) const { return x[i]; } };

// Metric parameters (kerr.rs:17-34 Kerr::new clamp, schwarzschild.rs:13-18).
struct MetricP {
  int kind;
  double mass;
  double spin;  // clamped
  double rs;    // 2M
};

// A metric known only by its tensor (GG_METRIC_USER): the function that
// fills g[4 mu + nu] at x, and Metric::event_horizon (has_rh false: None).
struct UserMetricO {
  void (*g)(const double x[4], double mass, double spin, double out[16]) = nullptr;
  bool has_rh = false;
  double rh = 0.0;
};
inline UserMetricO &user_metric() { static UserMetricO u; return u; }

inline MetricP make_metric(const gg_metric &g) {
  MetricP m;
  m.kind = g.kind;
  m.mass = g.mass;
  double max_spin = g.mass * 0.9999;                       // kerr.rs:31
  double s = g.spin;
  if (s < -max_spin) s = -max_spin;                        // f64::clamp
  if (s > max_spin) s = max_spin;
  m.spin = (g.kind == GG_METRIC_KERR) ? s : (g.kind == GG_METRIC_USER ? g.spin : 0.0);
  m.rs = 2.0 * g.mass;                                     // schwarzschild.rs:16
  return m;
}

// Metric tensor, 5 independent components.
template <class T> struct Gmn { T tt, rr, thth, phph, tph; };

// Kerr::metric_tensor (kerr.rs:61-94) and Schwarzschild::metric_tensor
// (schwarzschild.rs:26-43).  `a` is the spin as a scalar of type T so the
// dual-number gradient can carry d/da.
template <class T>
Gmn<T> metric_tensor(const MetricP &mp, const T &a, const V4<T> &pos) {
  using std::cos; using std::sin; using std::sqrt;
  Gmn<T> g;
  if (mp.kind == GG_METRIC_USER) {
    if constexpr (std::is_same<T, double>::value) {
      double x[4] = {pos[0], pos[1], pos[2], pos[3]}, m[16];
      user_metric().g(x, mp.mass, mp.spin, m);
      g.tt = m[0]; g.rr = m[5]; g.thth = m[10]; g.phph = m[15]; g.tph = m[3];
    }
    return g;
  }
  if (mp.kind == GG_METRIC_KERR) {
    T r = pos[1];
    if (!(val(r) >= 1e-10)) r = T(1e-10);                  // r.max(1e-10) (NaN -> 1e-10)
    T theta = pos[2];
    double m = mp.mass;
    T a2 = a * a;
    T cos_theta = cos(theta);
    T sigma = r * r + a * a * cos_theta * cos_theta;       // kerr.rs:41
    T delta = r * r - 2.0 * m * r + a * a;                 // kerr.rs:49
    T sin_theta = sin(theta);
    T sin2 = sin_theta * sin_theta;
    g.tt = -(1.0 - 2.0 * m * r / sigma);                   // kerr.rs:77
    g.rr = sigma / delta;                                  // :80
    g.thth = sigma;                                        // :83
    g.phph = (r * r + a2 + 2.0 * m * a2 * r * sin2 / sigma) * sin2;  // :86
    g.tph = -2.0 * m * a * r * sin2 / sigma;               // :89
  } else {
    T r = pos[1];
    if (!(val(r) >= 1e-10)) r = T(1e-10);                  // schwarzschild.rs:31
    T theta = pos[2];
    T f = 1.0 - mp.rs / r;                                 // :33
    T sin_theta = sin(theta);
    g.tt = -f;
    g.rr = 1.0 / f;
    g.thth = r * r;
    g.phph = r * r * sin_theta * sin_theta;
    g.tph = T(0.0);
  }
  return g;
}

// Kerr::christoffel (kerr.rs:100-424): dense Gamma[mu][alpha][beta].
template <class T>
void kerr_christoffel(const MetricP &mp, const T &a, const V4<T> &pos, T gamma[4][4][4]) {
  using std::cos; using std::sin;
  T r = pos[1];
  if (!(val(r) >= 1e-10)) r = T(1e-10);                    // :101
  T theta = pos[2];
  double m = mp.mass;
  T a2 = a * a;
  T cos_s = cos(theta);                                    // Kerr::sigma (:38-42)
  T sigma = r * r + a * a * cos_s * cos_s;
  T delta = r * r - 2.0 * m * r + a * a;                   // Kerr::delta (:46-50)
  T sin_theta = sin(theta);
  T cos_theta = cos(theta);
  T sin2 = sin_theta * sin_theta;
  T sigma2 = sigma * sigma;
  T r2 = r * r;
  T dsigma_dr = 2.0 * r;                                   // :126
  T dsigma_dtheta = -2.0 * a2 * sin_theta * cos_theta;     // :128
  T ddelta_dr = 2.0 * r - 2.0 * m;                         // :130
  T r2_plus_a2 = r2 + a2;                                  // :133
  T big_a = r2_plus_a2 * r2_plus_a2 - a2 * delta * sin2;   // :134
  T g_inv_tt = -big_a / (sigma * delta);                   // :144
  T g_inv_tphi = -2.0 * m * a * r / (sigma * delta);       // :145
  T g_inv_rr = delta / sigma;                              // :146
  T g_inv_thth = 1.0 / sigma;                              // :147
  T g_inv_phph = (delta - a2 * sin2) / (sigma * delta * sin2);  // :148
  T dg_tt_dr = 2.0 * m / sigma - 2.0 * m * r * dsigma_dr / sigma2;  // :156
  T dg_tt_dth = -2.0 * m * r * dsigma_dtheta / sigma2;     // :158
  T dg_rr_dr = (dsigma_dr * delta - sigma * ddelta_dr) / (delta * delta);  // :162
  T dg_rr_dth = dsigma_dtheta / delta;                     // :164
  T dg_thth_dr = dsigma_dr;                                // :168
  T dg_thth_dth = dsigma_dtheta;                           // :170
  T big_p = 2.0 * m * a2 * r;                              // :176
  // :187-188 (the value bound to dg_phph_dr at :208)
  T dg_phph_dr = 2.0 * r * sin2 + sin2 * sin2 * 2.0 * m * a2 * (sigma - 2.0 * r2) / sigma2;
  // :193-196
  T dg_phph_dth = 2.0 * r2_plus_a2 * sin_theta * cos_theta +
                  big_p * (4.0 * sin_theta * sin_theta * sin_theta * cos_theta * sigma -
                           sin2 * sin2 * dsigma_dtheta) / sigma2;
  T dg_tphi_dr = -2.0 * m * a * sin2 * (sigma - r * dsigma_dr) / sigma2;  // :200
  T dg_tphi_dth = -2.0 * m * a * r *                       // :202-204
                  (2.0 * sin_theta * cos_theta * sigma - sin2 * dsigma_dtheta) / sigma2;

  for (int i = 0; i < 4; ++i)
    for (int j = 0; j < 4; ++j)
      for (int k = 0; k < 4; ++k) gamma[i][j][k] = T(0.0);

  T v;
  v = 0.5 * (g_inv_tt * dg_tt_dr + g_inv_tphi * dg_tphi_dr);      gamma[0][0][1] = v; gamma[0][1][0] = v;  // :235
  v = 0.5 * (g_inv_tt * dg_tt_dth + g_inv_tphi * dg_tphi_dth);    gamma[0][0][2] = v; gamma[0][2][0] = v;  // :242
  v = 0.5 * (g_inv_tt * dg_tphi_dr + g_inv_tphi * dg_phph_dr);    gamma[0][1][3] = v; gamma[0][3][1] = v;  // :253
  v = 0.5 * (g_inv_tt * dg_tphi_dth + g_inv_tphi * dg_phph_dth);  gamma[0][2][3] = v; gamma[0][3][2] = v;  // :260
  gamma[1][0][0] = -0.5 * g_inv_rr * dg_tt_dr;                    // :272
  gamma[1][1][1] = 0.5 * g_inv_rr * dg_rr_dr;                     // :278
  gamma[1][2][2] = -0.5 * g_inv_rr * dg_thth_dr;                  // :284
  gamma[1][3][3] = -0.5 * g_inv_rr * dg_phph_dr;                  // :293
  v = 0.5 * g_inv_rr * dg_rr_dth;                                 gamma[1][1][2] = v; gamma[1][2][1] = v;  // :299
  v = -0.5 * g_inv_rr * dg_tphi_dr;                               gamma[1][0][3] = v; gamma[1][3][0] = v;  // :308
  gamma[2][0][0] = -0.5 * g_inv_thth * dg_tt_dth;                 // :318
  gamma[2][1][1] = -0.5 * g_inv_thth * dg_rr_dth;                 // :323
  gamma[2][2][2] = 0.5 * g_inv_thth * dg_thth_dth;                // :328
  gamma[2][3][3] = -0.5 * g_inv_thth * dg_phph_dth;               // :333
  v = 0.5 * g_inv_thth * dg_thth_dr;                              gamma[2][1][2] = v; gamma[2][2][1] = v;  // :339
  v = -0.5 * g_inv_thth * dg_tphi_dth;                            gamma[2][0][3] = v; gamma[2][3][0] = v;  // :347
  v = 0.5 * (g_inv_tphi * dg_tt_dr + g_inv_phph * dg_tphi_dr);    gamma[3][0][1] = v; gamma[3][1][0] = v;  // :359
  v = 0.5 * (g_inv_tphi * dg_tt_dth + g_inv_phph * dg_tphi_dth);  gamma[3][0][2] = v; gamma[3][2][0] = v;  // :366
  v = 0.5 * (g_inv_tphi * dg_tphi_dr + g_inv_phph * dg_phph_dr);  gamma[3][1][3] = v; gamma[3][3][1] = v;  // :375
  v = 0.5 * (g_inv_tphi * dg_tphi_dth + g_inv_phph * dg_phph_dth); gamma[3][2][3] = v; gamma[3][3][2] = v; // :382
  gamma[3][0][0] = T(0.0);                                        // :410
}

// Schwarzschild::christoffel (schwarzschild.rs:45-88), used by tests only.
template <class T>
void schw_christoffel(const MetricP &mp, const V4<T> &pos, T gamma[4][4][4]) {
  using std::cos; using std::sin;
  T r = pos[1];
  if (!(val(r) >= 1e-10)) r = T(1e-10);
  T theta = pos[2];
  T s = sin(theta), c = cos(theta);
  double rs = mp.rs;
  for (int i = 0; i < 4; ++i)
    for (int j = 0; j < 4; ++j)
      for (int k = 0; k < 4; ++k) gamma[i][j][k] = T(0.0);
  gamma[0][0][1] = rs / (2.0 * r * (r - rs)); gamma[0][1][0] = gamma[0][0][1];
  gamma[1][0][0] = rs * (r - rs) / (2.0 * r * r * r);
  gamma[1][1][1] = -rs / (2.0 * r * (r - rs));
  gamma[1][2][2] = -(r - rs);
  gamma[1][3][3] = -(r - rs) * s * s;
  gamma[2][1][2] = 1.0 / r; gamma[2][2][1] = 1.0 / r;
  gamma[2][3][3] = -s * c;
  gamma[3][1][3] = 1.0 / r; gamma[3][3][1] = 1.0 / r;
  gamma[3][2][3] = c / s; gamma[3][3][2] = gamma[3][2][3];
}

// Metric::geodesic_acceleration.  Kerr uses the trait default: dense 64-term
// contraction in mu, alpha, beta order (metric.rs:37-52).  Schwarzschild uses
// its inlined override (schwarzschild.rs:99-140).
// Matrix4::try_inverse (metric.rs:14-19): cofactor expansion over the
// column-major storage (the MESA gluInvertMatrix scheme that nalgebra's 4x4
// inverse follows), scaled by 1/det; false for a zero determinant.
inline bool inverse4(const double m[16], double inv[16]) {
  inv[0] = m[5] * m[10] * m[15] - m[5] * m[11] * m[14] - m[9] * m[6] * m[15] + m[9] * m[7] * m[14] +
           m[13] * m[6] * m[11] - m[13] * m[7] * m[10];
  inv[4] = -m[4] * m[10] * m[15] + m[4] * m[11] * m[14] + m[8] * m[6] * m[15] - m[8] * m[7] * m[14] -
           m[12] * m[6] * m[11] + m[12] * m[7] * m[10];
  inv[8] = m[4] * m[9] * m[15] - m[4] * m[11] * m[13] - m[8] * m[5] * m[15] + m[8] * m[7] * m[13] +
           m[12] * m[5] * m[11] - m[12] * m[7] * m[9];
  inv[12] = -m[4] * m[9] * m[14] + m[4] * m[10] * m[13] + m[8] * m[5] * m[14] - m[8] * m[6] * m[13] -
            m[12] * m[5] * m[10] + m[12] * m[6] * m[9];
  double det = m[0] * inv[0] + m[1] * inv[4] + m[2] * inv[8] + m[3] * inv[12];
  if (det == 0.0) return false;
  inv[1] = -m[1] * m[10] * m[15] + m[1] * m[11] * m[14] + m[9] * m[2] * m[15] - m[9] * m[3] * m[14] -
           m[13] * m[2] * m[11] + m[13] * m[3] * m[10];
  inv[5] = m[0] * m[10] * m[15] - m[0] * m[11] * m[14] - m[8] * m[2] * m[15] + m[8] * m[3] * m[14] +
           m[12] * m[2] * m[11] - m[12] * m[3] * m[10];
  inv[9] = -m[0] * m[9] * m[15] + m[0] * m[11] * m[13] + m[8] * m[1] * m[15] - m[8] * m[3] * m[13] -
           m[12] * m[1] * m[11] + m[12] * m[3] * m[9];
  inv[13] = m[0] * m[9] * m[14] - m[0] * m[10] * m[13] - m[8] * m[1] * m[14] + m[8] * m[2] * m[13] +
            m[12] * m[1] * m[10] - m[12] * m[2] * m[9];
  inv[2] = m[1] * m[6] * m[15] - m[1] * m[7] * m[14] - m[5] * m[2] * m[15] + m[5] * m[3] * m[14] +
           m[13] * m[2] * m[7] - m[13] * m[3] * m[6];
  inv[6] = -m[0] * m[6] * m[15] + m[0] * m[7] * m[14] + m[4] * m[2] * m[15] - m[4] * m[3] * m[14] -
           m[12] * m[2] * m[7] + m[12] * m[3] * m[6];
  inv[10] = m[0] * m[5] * m[15] - m[0] * m[7] * m[13] - m[4] * m[1] * m[15] + m[4] * m[3] * m[13] +
            m[12] * m[1] * m[7] - m[12] * m[3] * m[5];
  inv[14] = -m[0] * m[5] * m[14] + m[0] * m[6] * m[13] + m[4] * m[1] * m[14] - m[4] * m[2] * m[13] -
            m[12] * m[1] * m[6] + m[12] * m[2] * m[5];
  inv[3] = -m[1] * m[6] * m[11] + m[1] * m[7] * m[10] + m[5] * m[2] * m[11] - m[5] * m[3] * m[10] -
           m[9] * m[2] * m[7] + m[9] * m[3] * m[6];
  inv[7] = m[0] * m[6] * m[11] - m[0] * m[7] * m[10] - m[4] * m[2] * m[11] + m[4] * m[3] * m[10] +
           m[8] * m[2] * m[7] - m[8] * m[3] * m[6];
  inv[11] = -m[0] * m[5] * m[11] + m[0] * m[7] * m[9] + m[4] * m[1] * m[11] - m[4] * m[3] * m[9] -
            m[8] * m[1] * m[7] + m[8] * m[3] * m[5];
  inv[15] = m[0] * m[5] * m[10] - m[0] * m[6] * m[9] - m[4] * m[1] * m[10] + m[4] * m[2] * m[9] +
            m[8] * m[1] * m[6] - m[8] * m[2] * m[5];
  double inv_det = 1.0 / det;
  for (int i = 0; i < 16; ++i) inv[i] = inv[i] * inv_det;
  return true;
}

// numerical_christoffel (metric.rs:97-142): Gamma[mu][alpha][beta] from
// central differences (h = 1e-6) of the user tensor and its inverse.
inline void numerical_christoffel(const MetricP &mp, const double pos[4], double gamma[4][4][4]) {
  const double h = 1e-6;
  double dg[4][16];
  auto G = user_metric().g;
  for (int s = 0; s < 4; ++s) {
    double pp[4] = {pos[0], pos[1], pos[2], pos[3]}, pm[4] = {pos[0], pos[1], pos[2], pos[3]};
    pp[s] += h;
    pm[s] -= h;
    double gp[16], gm[16];
    G(pp, mp.mass, mp.spin, gp);
    G(pm, mp.mass, mp.spin, gm);
    for (int i = 0; i < 16; ++i) dg[s][i] = (gp[i] - gm[i]) / (2.0 * h);
  }
  double g[16], gi[16];
  G(pos, mp.mass, mp.spin, g);
  if (!inverse4(g, gi))
    for (int i = 0; i < 16; ++i) gi[i] = std::nan("");   // try_inverse fails: the reference panics
  for (int mu = 0; mu < 4; ++mu)
    for (int al = 0; al < 4; ++al)
      for (int be = 0; be < 4; ++be) {
        double sum = 0.0;
        for (int s = 0; s < 4; ++s)
          sum += gi[4 * s + mu] * (dg[al][4 * s + be] + dg[be][4 * s + al] - dg[s][4 * al + be]);
        gamma[mu][al][be] = 0.5 * sum;
      }
}

template <class T>
V4<T> geodesic_acceleration(const MetricP &mp, const T &a, const V4<T> &pos, const V4<T> &vel) {
  using std::cos; using std::sin;
  V4<T> acc;
  if (mp.kind == GG_METRIC_USER) {
    // the trait default (metric.rs:37-52) over numerical Christoffels
    if constexpr (std::is_same<T, double>::value) {
      double x[4] = {pos[0], pos[1], pos[2], pos[3]}, gamma[4][4][4];
      numerical_christoffel(mp, x, gamma);
      for (int mu = 0; mu < 4; ++mu) {
        acc[mu] = 0.0;
        for (int al = 0; al < 4; ++al)
          for (int be = 0; be < 4; ++be) acc[mu] = acc[mu] - gamma[mu][al][be] * vel[al] * vel[be];
      }
    }
    return acc;
  }
  if (mp.kind == GG_METRIC_KERR) {
    T gamma[4][4][4];
    kerr_christoffel(mp, a, pos, gamma);
    for (int mu = 0; mu < 4; ++mu) {
      acc[mu] = T(0.0);
      for (int al = 0; al < 4; ++al)
        for (int be = 0; be < 4; ++be) acc[mu] = acc[mu] - gamma[mu][al][be] * vel[al] * vel[be];
    }
    return acc;
  }
  T r = pos[1];
  if (!(val(r) >= 1e-10)) r = T(1e-10);                    // :104
  T theta = pos[2];
  T sin_theta = sin(theta);
  T cos_theta = cos(theta);
  double rs = mp.rs;
  T f_denom = r * (r - rs);                                // :110
  T rmrs = r - rs;
  T inv_r = 1.0 / r;
  T g_t_tr = rs / (2.0 * f_denom);                         // :115
  T g_r_tt = rs * rmrs / (2.0 * r * r * r);                // :116
  T g_r_rr = -rs / (2.0 * f_denom);                        // :117
  T g_r_thth = -rmrs;                                      // :118
  T g_r_phph = -rmrs * sin_theta * sin_theta;              // :119
  T g_th_phph = -sin_theta * cos_theta;                    // :121
  T g_ph_thph = cos_theta / sin_theta;                     // :123
  const T &vt = vel[0], &vr = vel[1], &vth = vel[2], &vph = vel[3];
  acc[0] = -2.0 * g_t_tr * vt * vr;                        // :131
  acc[1] = -g_r_tt * vt * vt - g_r_rr * vr * vr - g_r_thth * vth * vth - g_r_phph * vph * vph;  // :132-135
  acc[2] = -2.0 * inv_r * vr * vth - g_th_phph * vph * vph;  // :136
  acc[3] = -2.0 * inv_r * vr * vph - 2.0 * g_ph_thph * vth * vph;  // :137
  return acc;
}

// Metric::event_horizon (kerr.rs:427-436, schwarzschild.rs:142-144).
inline bool event_horizon(const MetricP &mp, double *rh) {
  if (mp.kind == GG_METRIC_USER) {
    *rh = user_metric().rh;
    return user_metric().has_rh;
  }
  if (mp.kind == GG_METRIC_KERR) {
    double m = mp.mass, a = mp.spin;
    double disc = m * m - a * a;
    if (disc < 0.0) return false;
    *rh = m + std::sqrt(disc);
    return true;
  }
  *rh = mp.rs;
  return true;
}

// circular_orbit_velocity (metric.rs:82-94).
template <class T>
bool circular_orbit_velocity(const MetricP &mp, const T &a, const V4<T> &pos, const T &omega, V4<T> &u) {
  using std::sqrt;
  Gmn<T> g = metric_tensor(mp, a, pos);
  T denom = -(g.tt + 2.0 * g.tph * omega + g.phph * omega * omega);
  if (val(denom) <= 0.0 || !std::isfinite(val(denom))) return false;
  T u_t = 1.0 / sqrt(denom);
  u[0] = u_t; u[1] = T(0.0); u[2] = T(0.0); u[3] = omega * u_t;
  return true;
}

// Metric::orbital_four_velocity (kerr.rs:439-449, schwarzschild.rs:146-153).
template <class T>
bool orbital_four_velocity(const MetricP &mp, const T &a, const V4<T> &pos, V4<T> &u) {
  using std::pow; using std::sqrt;
  T r = pos[1];
  double m = mp.mass;
  if (mp.kind == GG_METRIC_USER) return false;               // the trait default: None
  if (mp.kind == GG_METRIC_KERR) {
    if (val(r) <= 0.0) return false;
    double sqrt_m = std::sqrt(m);
    T omega = sqrt_m / (pow(r, 1.5) + a * sqrt_m);
    return circular_orbit_velocity(mp, a, pos, omega, u);
  }
  if (val(r) <= 3.0 * m) return false;
  T omega = sqrt(m / (r * r * r));
  return circular_orbit_velocity(mp, a, pos, omega, u);
}

// redshift_factor (crates/raytracer/src/curved/tracer.rs:158-183).
template <class T>
T redshift_factor(const MetricP &mp, const T &a, const V4<T> &pos, const V4<T> &p, const T &sqrt_neg_gtt_obs) {
  using std::sqrt;
  Gmn<T> g = metric_tensor(mp, a, pos);
  T p_t = g.tt * p[0] + g.tph * p[3];                      // :166
  T e_obs = -p_t / sqrt_neg_gtt_obs;
  if (!(val(e_obs) >= 0.0)) e_obs = T(0.0);                // .max(0.0)
  T e_emit;
  V4<T> u;
  if (orbital_four_velocity(mp, a, pos, u)) {
    T p_phi = g.tph * p[0] + g.phph * p[3];                // :172
    e_emit = -(p_t * u[0] + p_phi * u[3]);
    if (!(val(e_emit) >= 0.0)) e_emit = T(0.0);
  } else {
    T g_tt_emit = -g.tt;                                   // :177
    if (val(g_tt_emit) > 0.0) e_emit = -p_t / sqrt(g_tt_emit); else e_emit = T(0.0);
    if (!(val(e_emit) >= 0.0)) e_emit = T(0.0);
  }
  if (val(e_emit) > 0.0) return e_obs / e_emit;            // :182
  return T(0.0);
}

// ---- integrator ----------------------------------------------------------

template <class T> struct State { V4<T> pos, vel; T lambda; };

enum StepResult { STEP_CONTINUE = 0, STEP_HORIZON = 1, STEP_ESCAPED = 2, STEP_SINGULAR = 3 };

// is_valid_position (geodesic.rs:332-334).
template <class T> bool is_valid_position(const V4<T> &p) {
  for (int i = 0; i < 4; ++i)
    if (!std::isfinite(val(p[i]))) return false;
  return val(p[1]) > 0.0;
}

// f64::rem_euclid, carrying the derivative through unchanged.
inline double rem_euclid(double x, double m) {
  double r = std::fmod(x, m);
  return r < 0.0 ? r + std::fabs(m) : r;
}
inline double rem_euclid_t(double x, double m) { return rem_euclid(x, m); }

// normalize_spherical_coords (geodesic.rs:337-352).  Returns true when theta
// was reflected (the FSAL value k7 is then not the next k1).
template <class T> bool normalize_spherical_coords(V4<T> &pos) {
  bool reflected = false;
  if (val(pos[2]) < 0.0) { pos[2] = -pos[2]; pos[3] = pos[3] + PI; reflected = true; }
  if (val(pos[2]) > PI) { pos[2] = 2.0 * PI - pos[2]; pos[3] = pos[3] + PI; reflected = true; }
  pos[3] = rem_euclid_t(pos[3], 2.0 * PI);
  return reflected;
}

// Integrator parameters (RK45Integrator, geodesic.rs:155-181) plus the
// rtol/atol error mode and relative horizon margin.
struct IntegP {
  double initial_step, min_step, max_step, tolerance, max_radius, horizon_buffer;
  double horizon_rel, rtol, atol;
  int err_mode;
  uint32_t max_steps;
};

inline IntegP make_integ(const gg_integrator &g) {
  IntegP p;
  p.initial_step = g.initial_step; p.min_step = g.min_step; p.max_step = g.max_step;
  p.tolerance = g.tolerance; p.max_radius = g.max_radius; p.horizon_buffer = g.horizon_buffer;
  p.horizon_rel = g.horizon_rel; p.rtol = g.rtol; p.atol = g.atol;
  p.err_mode = g.err_mode; p.max_steps = g.max_steps;
  return p;
}

struct StepInfo {
  bool accepted = false;
  bool reflected = false;
  int rhs = 0;          // RHS evaluations made by this attempt
  double h_used = 0.0;  // h of the attempt
  // The attempt's discrete decisions (STEP_DEC_* bits): accept, err == 0,
  // a clamp of the next h, an invalid stage.  Two runs with equal decision
  // sequences differ smoothly in their parameters.
  uint32_t decisions = 0;
};
enum { STEP_DEC_ACCEPT = 1, STEP_DEC_ERR_ZERO = 2, STEP_DEC_CLAMP = 4, STEP_DEC_INVALID = 8, STEP_DEC_REFLECT = 16 };

// nalgebra Vector4::norm(): sqrt((x0^2 + x2^2) + (x1^2 + x3^2)).
inline double norm4(const double d[4]) {
  double a = d[0] * d[0], b = d[1] * d[1], c = d[2] * d[2], e = d[3] * d[3];
  a += c;
  b += e;
  return std::sqrt(0.0 + (a + b));
}

// x as the scalar type E: the value alone when E is double, else x itself.
template <class E, class T> inline E as_scalar(const T &x) {
  if constexpr (std::is_same<E, double>::value) return val(x); else return E(x);
}
// f64::max / fmax on values, carrying the tangent of the operand it returns.
inline double fmax_t(double x, double y) { return std::fmax(x, y); }
template <class T> inline T fmax_t(const T &x, const T &y) { return std::fmax(val(x), val(y)) == val(x) ? x : y; }

// RK45Integrator::step (geodesic.rs:185-328): Dormand-Prince velocity update,
// forward-Euler position, absolute error on ||new_vel - vel_err||_2.
// H is the type of the step size h.  H = double: the controller acts on
// values only, so a dual-number T gives the discretise-then-differentiate
// gradient with the step sizes frozen.  H = T: the error estimate and the
// controller (geodesic.rs:308-325) carry tangents too, so the gradient is
// that of the loss the algorithm computes, step sizes included; decisions
// (accept/reject, the clamps, err == 0) read values only, so the value part
// is the plain double run bit for bit.
// `force` >= 0 replays a recorded decision (1 accept, 0 reject) instead of
// testing the error estimate; the finite-difference gradient check uses it to
// hold the step sequence fixed.
template <class T, class H = double>
StepResult rk45_step(const MetricP &mp, const T &a, const IntegP &ip, State<T> &st, H &h,
                     StepInfo *info = nullptr, int force = -1) {
  using E = H;     // type of the error estimate
  V4<T> pos = st.pos, vel = st.vel;
  if (info) { info->accepted = false; info->reflected = false; info->rhs = 0; info->h_used = val(h); }
  if (val(pos[1]) > ip.max_radius) return STEP_ESCAPED;                   // :194-196
  double rh;
  if (event_horizon(mp, &rh)) {
    if (val(pos[1]) < rh * (1.0 + ip.horizon_rel) + ip.horizon_buffer) return STEP_HORIZON;  // :198-202
    if (val(pos[1]) < rh) return STEP_HORIZON;                            // :204-206 (metric.rs:60-66)
  }
  const double a2 = 1.0 / 5.0, a3 = 3.0 / 10.0, a4 = 4.0 / 5.0, a5 = 8.0 / 9.0;   // :210-213
  const double b21 = 1.0 / 5.0;
  const double b31 = 3.0 / 40.0, b32 = 9.0 / 40.0;
  const double b41 = 44.0 / 45.0, b42 = -56.0 / 15.0, b43 = 32.0 / 9.0;
  const double b51 = 19372.0 / 6561.0, b52 = -25360.0 / 2187.0, b53 = 64448.0 / 6561.0, b54 = -212.0 / 729.0;
  const double b61 = 9017.0 / 3168.0, b62 = -355.0 / 33.0, b63 = 46732.0 / 5247.0, b64 = 49.0 / 176.0,
               b65 = -5103.0 / 18656.0;
  const double c1 = 35.0 / 384.0, c3 = 500.0 / 1113.0, c4 = 125.0 / 192.0, c5 = -2187.0 / 6784.0,
               c6 = 11.0 / 84.0;
  const double d1 = 5179.0 / 57600.0, d3 = 7571.0 / 16695.0, d4 = 393.0 / 640.0, d5 = -92097.0 / 339200.0,
               d6 = 187.0 / 2100.0, d7 = 1.0 / 40.0;

  int nrhs = 0;
  V4<T> k1 = geodesic_acceleration(mp, a, pos, vel); ++nrhs;               // :245
  V4<T> p2, v2, p3, v3, p4, v4, p5, v5, p6, v6;
  for (int i = 0; i < 4; ++i) { p2[i] = pos[i] + h * a2 * vel[i]; v2[i] = vel[i] + h * b21 * k1[i]; }  // :248-249
  if (!is_valid_position(p2)) { h = h * 0.5; if (info) { info->rhs = nrhs; info->decisions = STEP_DEC_INVALID; } return STEP_CONTINUE; }
  V4<T> k2 = geodesic_acceleration(mp, a, p2, v2); ++nrhs;
  for (int i = 0; i < 4; ++i) { p3[i] = pos[i] + h * a3 * vel[i]; v3[i] = vel[i] + h * (b31 * k1[i] + b32 * k2[i]); }
  if (!is_valid_position(p3)) { h = h * 0.5; if (info) { info->rhs = nrhs; info->decisions = STEP_DEC_INVALID; } return STEP_CONTINUE; }
  V4<T> k3 = geodesic_acceleration(mp, a, p3, v3); ++nrhs;
  for (int i = 0; i < 4; ++i) {
    p4[i] = pos[i] + h * a4 * vel[i];
    v4[i] = vel[i] + h * (b41 * k1[i] + b42 * k2[i] + b43 * k3[i]);
  }
  if (!is_valid_position(p4)) { h = h * 0.5; if (info) { info->rhs = nrhs; info->decisions = STEP_DEC_INVALID; } return STEP_CONTINUE; }
  V4<T> k4 = geodesic_acceleration(mp, a, p4, v4); ++nrhs;
  for (int i = 0; i < 4; ++i) {
    p5[i] = pos[i] + h * a5 * vel[i];
    v5[i] = vel[i] + h * (b51 * k1[i] + b52 * k2[i] + b53 * k3[i] + b54 * k4[i]);
  }
  if (!is_valid_position(p5)) { h = h * 0.5; if (info) { info->rhs = nrhs; info->decisions = STEP_DEC_INVALID; } return STEP_CONTINUE; }
  V4<T> k5 = geodesic_acceleration(mp, a, p5, v5); ++nrhs;
  for (int i = 0; i < 4; ++i) {
    p6[i] = pos[i] + h * vel[i];
    v6[i] = vel[i] + h * (b61 * k1[i] + b62 * k2[i] + b63 * k3[i] + b64 * k4[i] + b65 * k5[i]);
  }
  if (!is_valid_position(p6)) { h = h * 0.5; if (info) { info->rhs = nrhs; info->decisions = STEP_DEC_INVALID; } return STEP_CONTINUE; }
  V4<T> k6 = geodesic_acceleration(mp, a, p6, v6); ++nrhs;
  V4<T> new_vel;
  for (int i = 0; i < 4; ++i)                                              // :294-295
    new_vel[i] = vel[i] + h * (c1 * k1[i] + c3 * k3[i] + c4 * k4[i] + c5 * k5[i] + c6 * k6[i]);
  V4<T> k7 = geodesic_acceleration(mp, a, p6, new_vel); ++nrhs;           // :298
  V4<T> vel_err;
  for (int i = 0; i < 4; ++i)                                              // :299-305
    vel_err[i] = vel[i] + h * (d1 * k1[i] + d3 * k3[i] + d4 * k4[i] + d5 * k5[i] + d6 * k6[i] + d7 * k7[i]);
  if (info) info->rhs = nrhs;

  using std::sqrt; using std::fabs; using std::pow;
  E diff[4];
  for (int i = 0; i < 4; ++i) diff[i] = as_scalar<E>(new_vel[i]) - as_scalar<E>(vel_err[i]);
  E err;            // the quantity compared against the acceptance threshold
  double thresh;    // tolerance (reference) or 1 (scaled)
  if (ip.err_mode == GG_ERR_RTOL_ATOL) {
    E s = E(0.0);
    for (int i = 0; i < 4; ++i) {
      E sc = ip.atol + ip.rtol * fmax_t(E(fabs(as_scalar<E>(vel[i]))), E(fabs(as_scalar<E>(new_vel[i]))));
      E q = diff[i] / sc;
      s = s + q * q;
    }
    err = sqrt(0.25 * s);
    thresh = 1.0;
  } else {
    // nalgebra Vector4::norm: sqrt((x0^2 + x2^2) + (x1^2 + x3^2))  (:308)
    E x0 = diff[0] * diff[0], x1 = diff[1] * diff[1], x2 = diff[2] * diff[2], x3 = diff[3] * diff[3];
    x0 = x0 + x2;
    x1 = x1 + x3;
    err = sqrt(0.0 + (x0 + x1));
    thresh = ip.tolerance;
  }

  bool accept = (val(err) < thresh || val(h) <= ip.min_step);              // :310
  if (force >= 0) accept = (force == 1);
  if (accept) {
    for (int i = 0; i < 4; ++i) st.pos[i] = pos[i] + h * vel[i];           // :312
    st.vel = new_vel;
    st.lambda = st.lambda + h;
    bool refl = normalize_spherical_coords(st.pos);
    uint32_t dec = STEP_DEC_ACCEPT | (refl ? STEP_DEC_REFLECT : 0u);
    if (val(err) > 0.0) {                                                  // :318-321
      H hn = h * 0.9 * pow(thresh / err, 0.2);
      if (val(hn) < ip.min_step) { hn = H(ip.min_step); dec |= STEP_DEC_CLAMP; }
      if (val(hn) > ip.max_step) { hn = H(ip.max_step); dec |= STEP_DEC_CLAMP; }
      h = hn;
    } else {
      dec |= STEP_DEC_ERR_ZERO;
    }
    if (info) { info->accepted = true; info->reflected = refl; info->decisions = dec; }
  } else {
    H hn = h * 0.9 * pow(thresh / err, 0.25);                              // :324
    if (info && val(hn) < ip.min_step) info->decisions = STEP_DEC_CLAMP;
    h = fmax_t(hn, H(ip.min_step));
  }
  return STEP_CONTINUE;
}

// RK4Integrator::step (geodesic.rs:73-132).  Only the reference's orbit
// tests use it; it is restated to run those tests against this RHS.
template <class T>
StepResult rk4_step(const MetricP &mp, const T &a, double h, double max_radius, double horizon_buffer,
                    State<T> &st) {
  V4<T> pos = st.pos, vel = st.vel;
  if (val(pos[1]) > max_radius) return STEP_ESCAPED;
  double rh;
  if (event_horizon(mp, &rh)) {
    if (val(pos[1]) < rh + horizon_buffer) return STEP_HORIZON;
    if (val(pos[1]) < rh) return STEP_HORIZON;
  }
  V4<T> k1p = vel, k1v = geodesic_acceleration(mp, a, pos, vel);
  V4<T> p2, v2, p3, v3, p4, v4;
  for (int i = 0; i < 4; ++i) { p2[i] = pos[i] + 0.5 * h * k1p[i]; v2[i] = vel[i] + 0.5 * h * k1v[i]; }
  if (!is_valid_position(p2)) return STEP_SINGULAR;
  V4<T> k2p = v2, k2v = geodesic_acceleration(mp, a, p2, v2);
  for (int i = 0; i < 4; ++i) { p3[i] = pos[i] + 0.5 * h * k2p[i]; v3[i] = vel[i] + 0.5 * h * k2v[i]; }
  if (!is_valid_position(p3)) return STEP_SINGULAR;
  V4<T> k3p = v3, k3v = geodesic_acceleration(mp, a, p3, v3);
  for (int i = 0; i < 4; ++i) { p4[i] = pos[i] + h * k3p[i]; v4[i] = vel[i] + h * k3v[i]; }
  if (!is_valid_position(p4)) return STEP_SINGULAR;
  V4<T> k4p = v4, k4v = geodesic_acceleration(mp, a, p4, v4);
  for (int i = 0; i < 4; ++i) {
    st.pos[i] = pos[i] + (h / 6.0) * (k1p[i] + 2.0 * k2p[i] + 2.0 * k3p[i] + k4p[i]);
    st.vel[i] = vel[i] + (h / 6.0) * (k1v[i] + 2.0 * k2v[i] + 2.0 * k3v[i] + k4v[i]);
  }
  st.lambda = st.lambda + h;
  normalize_spherical_coords(st.pos);
  return STEP_CONTINUE;
}

}  // namespace oracle
