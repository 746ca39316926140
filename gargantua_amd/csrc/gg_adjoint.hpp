// gg_adjoint.hpp -- device pieces of the hand-written adjoint (gg_grad).
//
// The adjoint differentiates the map from (spin, inclination, emissivity
// index q, inner radius r_in) to each ray's disk intensity through the camera
// ray, every attempt of the Dormand-Prince integrator (its stages, error
// estimate and step-size controller, or with the step sizes held fixed) and
// the disk crossing; the discrete decisions (accept/reject, clamps, the
// crossing step) are those of the forward render.  The per-stage RHS
// vector-Jacobian product is reverse mode (gg_vjp.hpp); the crossing and the
// camera ray, evaluated once per ray, are differentiated here in forward mode
// with one-tangent dual numbers.
#pragma once

#include "gg_physics.hpp"
#include "gg_vjp.hpp"

namespace gg {

// Value and one tangent.
struct Tn {
  double v, d;
};
__device__ inline Tn operator+(Tn a, Tn b) { return {a.v + b.v, a.d + b.d}; }
__device__ inline Tn operator-(Tn a, Tn b) { return {a.v - b.v, a.d - b.d}; }
__device__ inline Tn operator-(Tn a) { return {-a.v, -a.d}; }
__device__ inline Tn operator*(Tn a, Tn b) { return {a.v * b.v, a.d * b.v + a.v * b.d}; }
__device__ inline Tn operator/(Tn a, Tn b) {
  double q = a.v / b.v;
  return {q, (a.d - q * b.d) / b.v};
}
__device__ inline Tn operator/(double a, Tn b) {
  double q = a / b.v;
  return {q, -q * b.d / b.v};
}
__device__ inline Tn operator/(Tn a, double b) { return {a.v / b, a.d / b}; }
__device__ inline Tn operator*(double a, Tn b) { return {a * b.v, a * b.d}; }
__device__ inline Tn operator*(Tn a, double b) { return {a.v * b, a.d * b}; }
__device__ inline Tn operator+(double a, Tn b) { return {a + b.v, b.d}; }
__device__ inline Tn operator-(double a, Tn b) { return {a - b.v, -b.d}; }
__device__ inline Tn operator+(Tn a, double b) { return {a.v + b, a.d}; }
__device__ inline Tn operator-(Tn a, double b) { return {a.v - b, a.d}; }
__device__ inline Tn dsqrt(Tn a) {
  double s = sqrt(a.v);
  return {s, a.d * 0.5 / s};
}
__device__ inline Tn dpow(Tn a, double e) {
  double p = pow(a.v, e);
  return {p, a.d * e * p / a.v};
}
__device__ inline void dsincos(Tn a, Tn &s, Tn &c) {
  double sv, cv;
  sincos(a.v, &sv, &cv);
  s = {sv, a.d * cv};
  c = {cv, -a.d * sv};
}
__device__ inline Tn dconst(double x) { return {x, 0.0}; }
__device__ inline Tn dexp(Tn a) {
  double e = exp(a.v);
  return {e, a.d * e};
}
__device__ inline Tn dlog(Tn a) { return {log(a.v), a.d / a.v}; }
__device__ inline Tn dlogistic(Tn x) { return 1.0 / (1.0 + dexp(-x)); }
// value x with tangent 1 when k == seed, else 0
__device__ inline Tn seeded(double x, int k, int seed) { return {x, k == seed ? 1.0 : 0.0}; }

// Kerr metric components at (r, theta) with one tangent (kerr.rs:61-94).
struct GmnD {
  Tn tt, rr, thth, phph, tph;
};
__device__ inline GmnD kerr_metric_d(double m, Tn a, Tn r, Tn theta) {
  Tn st, ct;
  dsincos(theta, st, ct);
  Tn a2 = a * a;
  Tn sigma = r * r + a2 * ct * ct;
  Tn delta = r * r - (2.0 * m) * r + a2;
  Tn sin2 = st * st;
  GmnD g;
  g.tt = -(1.0 - (2.0 * m) * r / sigma);
  g.rr = sigma / delta;
  g.thth = sigma;
  g.phph = (r * r + a2 + (2.0 * m) * a2 * r * sin2 / sigma) * sin2;
  g.tph = -((2.0 * m) * a * r * sin2 / sigma);
  return g;
}

// redshift_factor (tracer.rs:158-183) for a Kerr emitter at (r, pi/2) and
// photon momentum (p0, ., ., p3), with one tangent.
__device__ inline Tn kerr_redshift_d(double m, Tn a, Tn r, Tn p0, Tn p3, Tn sqrt_obs) {
  GmnD g = kerr_metric_d(m, a, r, dconst(1.5707963267948966));
  Tn p_t = g.tt * p0 + g.tph * p3;
  Tn e_obs = -(p_t / sqrt_obs);
  if (!(e_obs.v >= 0.0)) e_obs = dconst(0.0);
  Tn e_emit = dconst(0.0);
  bool orbit_ok = false;
  if (r.v > 0.0) {
    double sqrt_m = sqrt(m);
    Tn omega = sqrt_m / (dpow(r, 1.5) + a * sqrt_m);
    Tn denom = -(g.tt + 2.0 * (g.tph * omega) + g.phph * omega * omega);
    if (denom.v > 0.0 && isfinite(denom.v)) {
      Tn u_t = 1.0 / dsqrt(denom);
      Tn u_ph = omega * u_t;
      Tn p_phi = g.tph * p0 + g.phph * p3;
      e_emit = -(p_t * u_t + p_phi * u_ph);
      if (!(e_emit.v >= 0.0)) e_emit = dconst(0.0);
      orbit_ok = true;
    }
  }
  if (!orbit_ok) {
    Tn g_tt_emit = -g.tt;
    e_emit = g_tt_emit.v > 0.0 ? -(p_t / dsqrt(g_tt_emit)) : dconst(0.0);
    if (!(e_emit.v >= 0.0)) e_emit = dconst(0.0);
  }
  return e_emit.v > 0.0 ? e_obs / e_emit : dconst(0.0);
}

// Camera ray of pixel (px, py) of the inclination camera
// (t, r, theta, phi) = (0, r_cam, i, 0) looking at -r_hat, up = +z, with one
// tangent through (a, i): Camera::pixel_direction and from_camera
// (camera.rs:20-29, ray.rs:14-61).  Returns (k^t, k^r, k^theta, k^phi).
struct CamD {
  Tn k[4];
};
__device__ inline CamD incl_camera_ray_d(double m, Tn a, Tn incl, double r_cam, double half_h, double half_w,
                                         uint32_t px, uint32_t py, uint32_t w, uint32_t h) {
  double u = ((double)px + 0.5) / (double)w * 2.0 - 1.0;
  double v = 1.0 - ((double)py + 0.5) / (double)h * 2.0;
  Tn si, ci;
  dsincos(incl, si, ci);
  // forward = normalize(-(sin i, 0, cos i)) = (-sin i, 0, -cos i); right = normalize(f x z)
  Tn f0 = -si, f1 = dconst(-0.0), f2 = -ci;
  Tn rx = f1 * 1.0 - f2 * 0.0, ry = f2 * 0.0 - f0 * 1.0, rz = f0 * 0.0 - f1 * 0.0;
  Tn rn = dsqrt(rx * rx + ry * ry + rz * rz);
  rx = rx / rn; ry = ry / rn; rz = rz / rn;
  // up' = normalize(right x forward)
  Tn ux = ry * f2 - rz * f1, uy = rz * f0 - rx * f2, uz = rx * f1 - ry * f0;
  Tn un = dsqrt(ux * ux + uy * uy + uz * uz);
  ux = ux / un; uy = uy / un; uz = uz / un;
  Tn d0 = f0 + (u * half_w) * rx + (v * half_h) * ux;
  Tn d1 = f1 + (u * half_w) * ry + (v * half_h) * uy;
  Tn d2 = f2 + (u * half_w) * rz + (v * half_h) * uz;
  Tn dn = dsqrt(d0 * d0 + d1 * d1 + d2 * d2);
  d0 = d0 / dn; d1 = d1 / dn; d2 = d2 / dn;
  // spherical frame at (theta = i, phi = 0)
  Tn n_r = d0 * si + d2 * ci;
  Tn n_t = d0 * ci - d2 * si;
  Tn n_p = d1;
  GmnD g = kerr_metric_d(m, a, dconst(r_cam), incl);
  CamD o;
  o.k[1] = n_r / dsqrt(g.rr);
  o.k[2] = n_t / dsqrt(g.thth);
  o.k[3] = n_p / dsqrt(g.phph);
  Tn spatial = g.rr * o.k[1] * o.k[1] + g.thth * o.k[2] * o.k[2] + g.phph * o.k[3] * o.k[3];
  o.k[0] = dsqrt(-(spatial / g.tt));
  return o;
}

}  // namespace gg

namespace gg {

// sqrt(-g_tt) at the observer (tracer.rs:93-94) for the inclination camera
// at (r_cam, theta = i), with one tangent through (a, i).
__device__ inline Tn kerr_sqrt_obs_d(double m, Tn a, double r_cam, Tn incl) {
  GmnD g = kerr_metric_d(m, a, dconst(r_cam), incl);
  Tn x = -g.tt;
  if (!(x.v >= 0.0)) x = dconst(0.0);
  return x.v > 0.0 ? dsqrt(x) : dconst(0.0);
}

// Inputs of the thin-disk crossing of one ray: parameters (spin a, q,
// r_in), sqrt(-g_tt) at the observer, and the (r, theta, p^t, p^phi) of the
// accepted step's start (0) and end (1).  Tangent k selects the input.
enum {
  CX_A = 0, CX_Q, CX_RIN, CX_SOBS, CX_R0, CX_TH0, CX_R1, CX_TH1, CX_V0T, CX_V0P, CX_V1T, CX_V1P, CX_N
};
struct CrossVals {
  double a, q, r_in, sobs, r0, th0, r1, th1, v0t, v0p, v1t, v1p;
};

// Intensity of a thin-disk hit with the tangent of input `seed`: the crossing
// of theta = pi/2 at s = (pi/2 - th0)/(th1 - th0) of the forward-Euler
// segment, r_c and the photon momentum interpolated at s, I = r_c^-q g^4
// times the logistic edge weights of width w (w > 0), g the redshift factor
// of a prograde Keplerian emitter (tracer.rs:158-183).
__device__ inline Tn crossing_intensity_d(double m, double r_out, double w, const CrossVals &x, int seed) {
  Tn a = seeded(x.a, CX_A, seed), q = seeded(x.q, CX_Q, seed), r_in = seeded(x.r_in, CX_RIN, seed);
  Tn sobs = seeded(x.sobs, CX_SOBS, seed);
  Tn r0 = seeded(x.r0, CX_R0, seed), th0 = seeded(x.th0, CX_TH0, seed);
  Tn r1 = seeded(x.r1, CX_R1, seed), th1 = seeded(x.th1, CX_TH1, seed);
  Tn v0t = seeded(x.v0t, CX_V0T, seed), v0p = seeded(x.v0p, CX_V0P, seed);
  Tn v1t = seeded(x.v1t, CX_V1T, seed), v1p = seeded(x.v1p, CX_V1P, seed);
  Tn s = (1.5707963267948966 - th0) / (th1 - th0);
  Tn rc = r0 + s * (r1 - r0);
  Tn vct = v0t + s * (v1t - v0t);
  Tn vcp = v0p + s * (v1p - v0p);
  Tn g = kerr_redshift_d(m, a, rc, vct, vcp, sobs);
  Tn g2 = g * g;
  Tn I = dexp(-(q * dlog(rc))) * (g2 * g2);
  if (w > 0.0) I = I * dlogistic((rc - r_in) / w) * dlogistic((r_out - rc) / w);
  return I;
}

}  // namespace gg
