// gg_user_metric.hpp -- METRIC 2: a metric known only by its tensor.
//
// The reference's Metric trait needs nothing but metric_tensor
// (crates/gr-core/src/metric.rs:11-12): the inverse comes from
// try_inverse (:14-19), the Christoffel symbols from central differences
// (numerical_christoffel, :97-142) and the acceleration from the dense
// contraction (:37-52); event_horizon and orbital_four_velocity default to
// None (:55-80).  Here the tensor is device code given at run time
// (gg_set_user_metric), compiled with hipRTC together with this header and
// the trace kernel; the library's own build never instantiates METRIC 2.
#pragma once

#include "gg_physics.hpp"

namespace gg {

// The caller's metric: g[4*mu + nu] = g_mu,nu at x = (t, r, theta, phi);
// mass and spin are the scene's gg_metric fields, passed through.
__device__ void gg_user_metric(const double x[4], double mass, double spin, double g[16]);

namespace {

// Matrix4::try_inverse: the cofactor expansion of a 4x4 matrix (the
// published MESA gluInvertMatrix scheme, which nalgebra's 4x4 inverse
// follows), divided by the determinant; false when it is zero.  m and inv
// are column-major like nalgebra's storage (g is symmetric, so the layout
// only fixes the rounding order).
__device__ inline bool inverse4(const double m[16], double inv[16]) {
#pragma clang fp contract(off)
  inv[0] = m[5] * m[10] * m[15] - m[5] * m[11] * m[14] - m[9] * m[6] * m[15] + m[9] * m[7] * m[14] +
           m[13] * m[6] * m[11] - m[13] * m[7] * m[10];
  inv[4] = -m[4] * m[10] * m[15] + m[4] * m[11] * m[14] + m[8] * m[6] * m[15] - m[8] * m[7] * m[14] -
           m[12] * m[6] * m[11] + m[12] * m[7] * m[10];
  inv[8] = m[4] * m[9] * m[15] - m[4] * m[11] * m[13] - m[8] * m[5] * m[15] + m[8] * m[7] * m[13] +
           m[12] * m[5] * m[11] - m[12] * m[7] * m[9];
  inv[12] = -m[4] * m[9] * m[14] + m[4] * m[10] * m[13] + m[8] * m[5] * m[14] - m[8] * m[6] * m[13] -
            m[12] * m[5] * m[10] + m[12] * m[6] * m[9];
  const double det = m[0] * inv[0] + m[1] * inv[4] + m[2] * inv[8] + m[3] * inv[12];
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
  const double inv_det = 1.0 / det;
  for (int i = 0; i < 16; ++i) inv[i] = inv[i] * inv_det;
  return true;
}

__device__ inline void user_g(const DevScene &S, const double x[4], double g[16]) {
  gg_user_metric(x, S.mass, S.spin, g);
}

// numerical_christoffel (metric.rs:97-142) contracted as
// geodesic_acceleration (metric.rs:37-52): a^mu = -Gamma^mu_ab v^a v^b with
// Gamma^mu_ab = 1/2 sum_s g^mu s (d_a g_sb + d_b g_sa - d_s g_ab) and
// d_s g = (g(x + h e_s) - g(x - h e_s)) / 2h, h = 1e-6.  A singular metric
// gives NaN (the reference panics there; NaN makes the step invalid).
// A source that defines GG_METRIC_STATIONARY (no t dependence) or
// GG_METRIC_AXISYMMETRIC (no phi dependence) skips those two tensor
// evaluations: for such a tensor g(x + h e_s) and g(x - h e_s) are equal, so
// the difference the reference computes is exactly +0 and the result is
// the same bit for bit.  The derivatives are read from the upper triangle
// (d_a g_{s b} as d_a g_{min, max}): the tensor is symmetric, so for a source
// that fills g_{mu nu} and g_{nu mu} alike this is the reference's value, and
// the 24 lower-triangle differences, and the registers they would hold, drop
// out of the kernel.
// Index of g_{i j} in the upper triangle (i <= j) of the row-major tensor.
__device__ constexpr int sym(int i, int j) { return i <= j ? 4 * i + j : 4 * j + i; }

__device__ inline d4 user_rhs(const DevScene &S, const double x[4], d4 v) {
#pragma clang fp contract(off)
  const double h = 1e-6;
  double dg[4][16];
  // unrolled, so that dg lives in registers rather than in private memory
#pragma unroll
  for (int s = 0; s < 4; ++s) {
#ifdef GG_METRIC_STATIONARY
    if (s == 0) {
      for (int i = 0; i < 16; ++i) dg[0][i] = 0.0;
      continue;
    }
#endif
#ifdef GG_METRIC_AXISYMMETRIC
    if (s == 3) {
      for (int i = 0; i < 16; ++i) dg[3][i] = 0.0;
      continue;
    }
#endif
    double xp[4] = {x[0], x[1], x[2], x[3]}, xm[4] = {x[0], x[1], x[2], x[3]};
    xp[s] = xp[s] + h;
    xm[s] = xm[s] - h;
    double gp[16], gm[16];
    user_g(S, xp, gp);
    user_g(S, xm, gm);
    for (int i = 0; i < 16; ++i) dg[s][i] = (gp[i] - gm[i]) / (2.0 * h);
  }
  double g[16], gi[16];
  user_g(S, x, g);
  if (!inverse4(g, gi)) {
    const double nan = __builtin_nan("");
    return d4{nan, nan, nan, nan};
  }
  const double vv[4] = {v.x, v.y, v.z, v.w};
  double acc[4];
  for (int mu = 0; mu < 4; ++mu) {
    acc[mu] = 0.0;
    for (int a = 0; a < 4; ++a)
      for (int b = 0; b < 4; ++b) {
        double sum = 0.0;
        for (int s = 0; s < 4; ++s)      // g^{mu s} (d_a g_{s b} + d_b g_{s a} - d_s g_{a b})
          sum += gi[4 * s + mu] * (dg[a][sym(s, b)] + dg[b][sym(s, a)] - dg[s][sym(a, b)]);
        const double gamma = 0.5 * sum;
        acc[mu] -= gamma * vv[a] * vv[b];
      }
  }
  return d4{acc[0], acc[1], acc[2], acc[3]};
}

// metric_tensor components the shading needs (g_tt, g_t phi, g_phi phi and
// the diagonal), at a full position.
__device__ inline Gmn user_gmn(const DevScene &S, const double x[4]) {
  double g[16];
  user_g(S, x, g);
  Gmn m;
  m.tt = g[0]; m.rr = g[5]; m.thth = g[10]; m.phph = g[15]; m.tph = g[3];
  return m;
}

// redshift_factor (tracer.rs:158-183) with orbital_four_velocity = None:
// the emitter is static at the emission point.
__device__ inline double user_redshift(const DevScene &S, const double x[4], d4 p) {
#pragma clang fp contract(off)
  const Gmn g = user_gmn(S, x);
  const double p_t = g.tt * p.x + g.tph * p.w;
  double e_obs = -p_t / S.sqrt_neg_gtt_obs;
  e_obs = e_obs >= 0.0 ? e_obs : 0.0;
  const double g_tt_emit = -g.tt;
  double e_emit = g_tt_emit > 0.0 ? -p_t / sqrt(g_tt_emit) : 0.0;
  e_emit = e_emit >= 0.0 ? e_emit : 0.0;
  return e_emit > 0.0 ? e_obs / e_emit : 0.0;
}

}  // namespace

}  // namespace gg
