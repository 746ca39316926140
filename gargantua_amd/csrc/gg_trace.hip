// gg_trace.hip -- persistent geodesic trace kernels for gfx950 (MI355X).
//
// One lane carries one ray with its whole state in registers.  A wavefront
// takes rays from a global atomic queue; every `refill_interval` attempts it
// ballots its retired lanes and hands each of them the next ray with a
// single atomicAdd per wavefront, so lanes whose rays finished early do not
// idle while neighbours run photon-ring orbits of 10-50x more steps.
//
// Per attempt a lane runs RK45Integrator::step (crates/gr-core/src/geodesic.rs
// :185-328): Dormand-Prince velocity stages, forward-Euler position, error on
// the velocity difference.  k1 of an attempt is k7 of the previous accepted
// attempt (pos6 = pos + h vel is the accepted position and the metric does
// not depend on t or phi) unless theta was reflected at a pole, and is kept
// across rejections; 6 RHS evaluations per attempt instead of the
// reference's 7, bit-identical results.
//
// Mixed precision: an attempt evaluates its six RHS stages in FP32 only when
// the step-size controller is saturated (h == max_step) and the previous
// attempt's error ratio and an estimate of the FP32 rounding noise leave the
// accept/reject decision and the next h certain.  The stage combinations,
// state, error norm and controller stay FP64.  If the FP32 attempt's own
// error ratio exceeds 0.4 it is discarded and redone in FP64.  Accepted
// steps therefore use the same h sequence as the FP64 kernel.
#include <hip/hip_runtime.h>

#include "gg_internal.hpp"

namespace gg {

namespace {

// Dormand-Prince 5(4) coefficients (geodesic.rs:210-242).
constexpr double A2 = 1.0 / 5.0, A3 = 3.0 / 10.0, A4 = 4.0 / 5.0, A5 = 8.0 / 9.0;
constexpr double B21 = 1.0 / 5.0;
constexpr double B31 = 3.0 / 40.0, B32 = 9.0 / 40.0;
constexpr double B41 = 44.0 / 45.0, B42 = -56.0 / 15.0, B43 = 32.0 / 9.0;
constexpr double B51 = 19372.0 / 6561.0, B52 = -25360.0 / 2187.0, B53 = 64448.0 / 6561.0, B54 = -212.0 / 729.0;
constexpr double B61 = 9017.0 / 3168.0, B62 = -355.0 / 33.0, B63 = 46732.0 / 5247.0, B64 = 49.0 / 176.0,
                 B65 = -5103.0 / 18656.0;
constexpr double C1 = 35.0 / 384.0, C3 = 500.0 / 1113.0, C4 = 125.0 / 192.0, C5 = -2187.0 / 6784.0,
                 C6 = 11.0 / 84.0;
constexpr double D1 = 5179.0 / 57600.0, D3 = 7571.0 / 16695.0, D4 = 393.0 / 640.0, D5 = -92097.0 / 339200.0,
                 D6 = 187.0 / 2100.0, D7 = 1.0 / 40.0;
constexpr double PI = 3.14159265358979323846;
constexpr double TWO_PI = 2.0 * PI;
constexpr double HALF_PI = PI / 2.0;

enum { RES_CONTINUE = 0, RES_HORIZON = 1, RES_ESCAPED = 2 };

struct Lane {
  double pos[4];
  d4 vel;
  double lambda;
  double h;
  d4 k1;
  bool k1_ok;
  uint32_t attempts, accepted;
  double intensity, transmission;
  double err_prev;   // error ratio of the previous full attempt
  uint32_t n_rhs, n_rhs32, n_full;
};

__device__ inline bool valid4(double p0, double p1, double p2, double p3) {
  return isfinite(p0) && isfinite(p1) && isfinite(p2) && isfinite(p3) && p1 > 0.0;
}

__device__ inline d4 to64(f4 a) { return d4{(double)a.x, (double)a.y, (double)a.z, (double)a.w}; }

template <int METRIC>
__device__ inline d4 eval_rhs(const DevScene &S, bool use32, double r, double th, d4 v) {
  if (use32) {
    f4 vf{(float)v.x, (float)v.y, (float)v.z, (float)v.w};
    return to64(rhs32<METRIC>(S, (float)r, (float)th, vf));
  }
  return rhs64<METRIC>(S, r, th, v);
}

// One RK45 attempt.  Returns RES_*; on acceptance sets *acc and leaves the
// pre-step position/velocity in p0/v0 for the disk hooks.
template <int METRIC, int PREC>
__device__ inline int attempt(const DevScene &S, Lane &L, bool &acc, double p0[4], d4 &v0, double &h_used) {
#pragma clang fp contract(off)
  acc = false;
  const double r = L.pos[1];
  if (r > S.max_radius) return RES_ESCAPED;          // geodesic.rs:194-196
  if (r < S.r_capture) return RES_HORIZON;           // :198-202
  if (r < S.r_h) return RES_HORIZON;                 // :204-206
  if (!L.k1_ok) {
    L.k1 = rhs64<METRIC>(S, L.pos[1], L.pos[2], L.vel);
    L.k1_ok = true;
    L.n_rhs++;
  }
  const d4 vel = L.vel;
  const d4 k1 = L.k1;
  for (int i = 0; i < 4; ++i) p0[i] = L.pos[i];
  v0 = vel;

  bool use32 = false;
  if (PREC == 1) {
    // FP32 only where the controller is pinned at max_step with margin.
    double noise;
    if (S.err_mode == 1) {
      double m0 = fabs(k1.x) / (S.atol + S.rtol * fabs(vel.x));
      double m1 = fabs(k1.y) / (S.atol + S.rtol * fabs(vel.y));
      double m2 = fabs(k1.z) / (S.atol + S.rtol * fabs(vel.z));
      double m3 = fabs(k1.w) / (S.atol + S.rtol * fabs(vel.w));
      noise = fmax(fmax(m0, m1), fmax(m2, m3));
    } else {
      noise = (fabs(k1.x) + fabs(k1.y) + fabs(k1.z) + fabs(k1.w)) / S.tol;
    }
    noise *= 1e-6 * L.h;
    const double th = L.pos[2];
    use32 = (L.h == S.max_step) && (L.err_prev < 0.3) && (noise < 0.05) && (r >= S.fp32_min_r) &&
            (th > 1e-3) && (th < PI - 1e-3);
  }

  for (int pass = 0; pass < 2; ++pass) {
    const double h = L.h;
    h_used = h;
    double q0, q1, q2, q3;
    d4 k2, k3, k4, k5, k6, k7, vs;
    int nr = 0;
    // stage 2 (geodesic.rs:248-254)
    {
      double ha = h * A2;
      q0 = p0[0] + ha * vel.x; q1 = p0[1] + ha * vel.y; q2 = p0[2] + ha * vel.z; q3 = p0[3] + ha * vel.w;
      if (!valid4(q0, q1, q2, q3)) { L.h = h * 0.5; L.n_rhs += nr; return RES_CONTINUE; }
      double hb = h * B21;
      vs.x = vel.x + hb * k1.x; vs.y = vel.y + hb * k1.y; vs.z = vel.z + hb * k1.z; vs.w = vel.w + hb * k1.w;
      k2 = eval_rhs<METRIC>(S, use32, q1, q2, vs); ++nr;
    }
    // stage 3 (:257-263)
    {
      double ha = h * A3;
      q0 = p0[0] + ha * vel.x; q1 = p0[1] + ha * vel.y; q2 = p0[2] + ha * vel.z; q3 = p0[3] + ha * vel.w;
      if (!valid4(q0, q1, q2, q3)) { L.h = h * 0.5; L.n_rhs += nr; return RES_CONTINUE; }
      vs.x = vel.x + h * (B31 * k1.x + B32 * k2.x);
      vs.y = vel.y + h * (B31 * k1.y + B32 * k2.y);
      vs.z = vel.z + h * (B31 * k1.z + B32 * k2.z);
      vs.w = vel.w + h * (B31 * k1.w + B32 * k2.w);
      k3 = eval_rhs<METRIC>(S, use32, q1, q2, vs); ++nr;
    }
    // stage 4 (:266-272)
    {
      double ha = h * A4;
      q0 = p0[0] + ha * vel.x; q1 = p0[1] + ha * vel.y; q2 = p0[2] + ha * vel.z; q3 = p0[3] + ha * vel.w;
      if (!valid4(q0, q1, q2, q3)) { L.h = h * 0.5; L.n_rhs += nr; return RES_CONTINUE; }
      vs.x = vel.x + h * (B41 * k1.x + B42 * k2.x + B43 * k3.x);
      vs.y = vel.y + h * (B41 * k1.y + B42 * k2.y + B43 * k3.y);
      vs.z = vel.z + h * (B41 * k1.z + B42 * k2.z + B43 * k3.z);
      vs.w = vel.w + h * (B41 * k1.w + B42 * k2.w + B43 * k3.w);
      k4 = eval_rhs<METRIC>(S, use32, q1, q2, vs); ++nr;
    }
    // stage 5 (:275-281)
    {
      double ha = h * A5;
      q0 = p0[0] + ha * vel.x; q1 = p0[1] + ha * vel.y; q2 = p0[2] + ha * vel.z; q3 = p0[3] + ha * vel.w;
      if (!valid4(q0, q1, q2, q3)) { L.h = h * 0.5; L.n_rhs += nr; return RES_CONTINUE; }
      vs.x = vel.x + h * (B51 * k1.x + B52 * k2.x + B53 * k3.x + B54 * k4.x);
      vs.y = vel.y + h * (B51 * k1.y + B52 * k2.y + B53 * k3.y + B54 * k4.y);
      vs.z = vel.z + h * (B51 * k1.z + B52 * k2.z + B53 * k3.z + B54 * k4.z);
      vs.w = vel.w + h * (B51 * k1.w + B52 * k2.w + B53 * k3.w + B54 * k4.w);
      k5 = eval_rhs<METRIC>(S, use32, q1, q2, vs); ++nr;
    }
    // stage 6 (:284-291): pos6 = pos + h vel is also the accepted position
    double p6[4];
    {
      p6[0] = p0[0] + h * vel.x; p6[1] = p0[1] + h * vel.y; p6[2] = p0[2] + h * vel.z; p6[3] = p0[3] + h * vel.w;
      if (!valid4(p6[0], p6[1], p6[2], p6[3])) { L.h = h * 0.5; L.n_rhs += nr; return RES_CONTINUE; }
      vs.x = vel.x + h * (B61 * k1.x + B62 * k2.x + B63 * k3.x + B64 * k4.x + B65 * k5.x);
      vs.y = vel.y + h * (B61 * k1.y + B62 * k2.y + B63 * k3.y + B64 * k4.y + B65 * k5.y);
      vs.z = vel.z + h * (B61 * k1.z + B62 * k2.z + B63 * k3.z + B64 * k4.z + B65 * k5.z);
      vs.w = vel.w + h * (B61 * k1.w + B62 * k2.w + B63 * k3.w + B64 * k4.w + B65 * k5.w);
      k6 = eval_rhs<METRIC>(S, use32, p6[1], p6[2], vs); ++nr;
    }
    // 5th-order velocity (:294-295), k7 at (pos6, new_vel) (:298)
    d4 nv;
    nv.x = vel.x + h * (C1 * k1.x + C3 * k3.x + C4 * k4.x + C5 * k5.x + C6 * k6.x);
    nv.y = vel.y + h * (C1 * k1.y + C3 * k3.y + C4 * k4.y + C5 * k5.y + C6 * k6.y);
    nv.z = vel.z + h * (C1 * k1.z + C3 * k3.z + C4 * k4.z + C5 * k5.z + C6 * k6.z);
    nv.w = vel.w + h * (C1 * k1.w + C3 * k3.w + C4 * k4.w + C5 * k5.w + C6 * k6.w);
    k7 = eval_rhs<METRIC>(S, use32, p6[1], p6[2], nv); ++nr;
    // 4th-order velocity (:299-305) and the error (:308)
    d4 ve;
    ve.x = vel.x + h * (D1 * k1.x + D3 * k3.x + D4 * k4.x + D5 * k5.x + D6 * k6.x + D7 * k7.x);
    ve.y = vel.y + h * (D1 * k1.y + D3 * k3.y + D4 * k4.y + D5 * k5.y + D6 * k6.y + D7 * k7.y);
    ve.z = vel.z + h * (D1 * k1.z + D3 * k3.z + D4 * k4.z + D5 * k5.z + D6 * k6.z + D7 * k7.z);
    ve.w = vel.w + h * (D1 * k1.w + D3 * k3.w + D4 * k4.w + D5 * k5.w + D6 * k6.w + D7 * k7.w);
    double dx = nv.x - ve.x, dy = nv.y - ve.y, dz = nv.z - ve.z, dw = nv.w - ve.w;
    double err, thresh;
    if (S.err_mode == 1) {
      double s0 = dx / (S.atol + S.rtol * fmax(fabs(vel.x), fabs(nv.x)));
      double s1 = dy / (S.atol + S.rtol * fmax(fabs(vel.y), fabs(nv.y)));
      double s2 = dz / (S.atol + S.rtol * fmax(fabs(vel.z), fabs(nv.z)));
      double s3 = dw / (S.atol + S.rtol * fmax(fabs(vel.w), fabs(nv.w)));
      double s = 0.0;
      s = s + s0 * s0; s = s + s1 * s1; s = s + s2 * s2; s = s + s3 * s3;
      err = sqrt(0.25 * s);
      thresh = 1.0;
    } else {
      // nalgebra Vector4::norm: sqrt((x0^2 + x2^2) + (x1^2 + x3^2))
      double a = dx * dx, b = dy * dy, c = dz * dz, e = dw * dw;
      a = a + c;
      b = b + e;
      err = sqrt(a + b);
      thresh = S.tol;
    }
    L.n_rhs += nr;
    if (use32) L.n_rhs32 += nr;
    double ratio = err / thresh;
    if (PREC == 1 && use32 && !(ratio < 0.4)) {  // FP32 result too close to a decision: redo in FP64
      use32 = false;
      continue;
    }
    L.n_full++;
    L.err_prev = ratio;
    if (err < thresh || h <= S.min_step) {           // accept (:310-321)
      L.pos[0] = p6[0]; L.pos[1] = p6[1]; L.pos[2] = p6[2]; L.pos[3] = p6[3];
      L.vel = nv;
      L.lambda = L.lambda + h;
      // normalize_spherical_coords (:337-352)
      bool refl = false;
      if (L.pos[2] < 0.0) { L.pos[2] = -L.pos[2]; L.pos[3] = L.pos[3] + PI; refl = true; }
      if (L.pos[2] > PI) { L.pos[2] = 2.0 * PI - L.pos[2]; L.pos[3] = L.pos[3] + PI; refl = true; }
      double ph = fmod(L.pos[3], TWO_PI);
      L.pos[3] = ph < 0.0 ? ph + TWO_PI : ph;
      L.k1 = k7;
      L.k1_ok = !refl;
      if (err > 0.0) {
        double hn = h * 0.9 * pow(thresh / err, 0.2);
        if (hn < S.min_step) hn = S.min_step;
        if (hn > S.max_step) hn = S.max_step;
        L.h = hn;
      }
      acc = true;
    } else {                                          // reject (:322-325)
      double hn = h * 0.9 * pow(thresh / err, 0.25);
      L.h = hn >= S.min_step ? hn : S.min_step;       // f64::max
    }
    return RES_CONTINUE;
  }
  return RES_CONTINUE;  // not reached
}

__device__ inline void ray_pixel(const DevScene &S, uint64_t local, uint32_t &px, uint32_t &py) {
  if (S.tiled) {
    uint64_t tile = local >> 6;
    uint32_t within = (uint32_t)(local & 63);
    uint32_t tiles_x = S.sub_w >> 3;
    px = (uint32_t)(tile % tiles_x) * 8 + (within & 7);
    py = (uint32_t)(tile / tiles_x) * 8 + (within >> 3);
  } else {
    px = (uint32_t)(local % S.sub_w);
    py = (uint32_t)(local / S.sub_w);
  }
}

__device__ inline uint32_t find_scene(const DevScene *scenes, uint32_t n, uint64_t ray) {
  uint32_t lo = 0, hi = n;  // last scene with ray_begin <= ray
  while (hi - lo > 1) {
    uint32_t mid = (lo + hi) >> 1;
    if (scenes[mid].ray_begin <= ray) lo = mid; else hi = mid;
  }
  return lo;
}

__device__ inline unsigned long long wave_sum(unsigned long long x) {
  for (int off = 32; off > 0; off >>= 1) x += __shfl_xor(x, off, 64);
  return x;
}

// Persistent trace kernel.  SRC 0: rays from the camera; 1: states_in.
// BATCH 0: one scene, its constants wavefront-uniform (scalar registers);
// BATCH 1: each lane looks up the scene of its ray.
template <int METRIC, int DISK, int PREC, int SRC, int BATCH>
__global__ __launch_bounds__(256) void trace_kernel(TraceLaunch T) {
#pragma clang fp contract(off)
  const uint32_t lane = threadIdx.x & 63;
  const uint64_t n = T.n_rays;
  constexpr bool single = (BATCH == 0);

  Lane L;
  L.n_rhs = 0; L.n_rhs32 = 0; L.n_full = 0;
  uint32_t n_att = 0;
  bool active = false;
  uint64_t ray = 0;
  uint32_t si = 0;        // scene of this lane's ray
  uint64_t out_i = 0;     // output slot of this lane's ray
  bool exhausted = false;
  // Per-lane scene copy: uniform (SGPR) for single-scene launches.
  DevScene S = T.scenes[0];

  uint32_t iter = 0;
  for (;;) {
    // ---- refill retired lanes: one atomicAdd per wavefront ----
    bool need = (iter == 0) || ((iter % T.refill_interval) == 0) || !__any(active);
    if (need && !exhausted) {
      unsigned long long idle = __ballot(!active);
      uint32_t cnt = __popcll(idle);
      if (cnt) {
        uint32_t leader = __ffsll((long long)idle) - 1;
        unsigned long long base = 0;
        if (lane == leader) base = atomicAdd(&T.counters->queue, (unsigned long long)cnt);
        base = __shfl(base, leader, 64);
        if (base >= n) exhausted = true;
        if (!active) {
          uint32_t rank = __popcll(idle & ((1ull << lane) - 1ull));
          ray = base + rank;
          if (ray < n) {
            si = single ? 0u : find_scene(T.scenes, T.n_scenes, ray);
            if (!single) S = T.scenes[si];
            uint64_t local = ray - S.ray_begin;
            if (SRC == 0) {
              uint32_t px, py;
              ray_pixel(S, local, px, py);
              out_i = (uint64_t)py * S.sub_w + px;
              camera_ray(S, px, py, L.pos, L.vel);
              L.lambda = 0.0;
            } else {
              out_i = local;
              const double *st = T.states_in + 9 * ray;
              for (int c = 0; c < 4; ++c) L.pos[c] = st[c];
              L.vel = d4{st[4], st[5], st[6], st[7]};
              L.lambda = st[8];
            }
            L.h = S.h0;
            L.k1_ok = false;
            L.attempts = 0; L.accepted = 0;
            L.intensity = 0.0; L.transmission = 1.0;
            L.err_prev = 1.0;
            active = true;
          }
        }
      }
    }
    if (!__any(active)) {
      if (exhausted) break;
      ++iter;
      continue;
    }
    ++iter;
    if (!active) continue;

    // ---- one step() call ----
    uint8_t cls = 255;
    if (L.attempts >= S.max_steps) {
      cls = GG_CLASS_MAX_STEPS;
    } else {
      bool acc;
      double p0[4];
      d4 v0;
      double h_used;
      double lam0 = L.lambda;
      int res = attempt<METRIC, PREC>(S, L, acc, p0, v0, h_used);
      L.attempts++;
      n_att++;
      if (acc) L.accepted++;
      if (DISK == 1) {
        // trace_ray_with_disk (tracer.rs:96-145)
        double dl = L.lambda - lam0;
        dl = dl >= 0.0 ? dl : 0.0;
        if (dl > 0.0) {
          double mr = 0.5 * (p0[1] + L.pos[1]);
          double mth = 0.5 * (p0[2] + L.pos[2]);
          double rho = disk_density(S, mr, mth);
          if (rho > 0.0) {
            double j = disk_emission(S, rho, mr, mth);
            d4 pm{0.5 * (v0.x + L.vel.x), 0.5 * (v0.y + L.vel.y), 0.5 * (v0.z + L.vel.z), 0.5 * (v0.w + L.vel.w)};
            double g = redshift64<METRIC>(S, mr, mth, pm);
            double g2 = g * g;
            double red = j * (g2 * g2);
            L.intensity = L.intensity + red * dl * L.transmission;
            L.transmission *= exp(-rho * dl * S.absorption);
            if (L.transmission < 1e-3) cls = GG_CLASS_DISK;
          }
        }
        if (cls == 255) {
          if (res == RES_HORIZON) cls = L.intensity > 0.0 ? GG_CLASS_DISK : GG_CLASS_HORIZON;
          else if (res == RES_ESCAPED) cls = L.intensity > 0.0 ? GG_CLASS_DISK : GG_CLASS_ESCAPED;
        }
      } else {
        if (DISK == 2 && acc) {
          // thin equatorial disk: crossing of theta = pi/2 on the segment
          bool side0 = p0[2] < HALF_PI, side1 = L.pos[2] < HALF_PI;
          if (side0 != side1) {
            double s = (HALF_PI - p0[2]) / (L.pos[2] - p0[2]);
            double rc = p0[1] + s * (L.pos[1] - p0[1]);
            double w = S.edge_width;
            bool hit = (w > 0.0) ? (rc <= S.r_out + 8.0 * w && rc > 0.0) : (rc >= S.r_in && rc <= S.r_out);
            if (hit) {
              d4 vc{v0.x + s * (L.vel.x - v0.x), v0.y + s * (L.vel.y - v0.y), v0.z + s * (L.vel.z - v0.z),
                    v0.w + s * (L.vel.w - v0.w)};
              double g = redshift64<METRIC>(S, rc, HALF_PI, vc);
              double g2 = g * g;
              double I = pow(rc, -S.q) * (g2 * g2);
              if (w > 0.0) {
                I = I * (1.0 / (1.0 + exp(-((rc - S.r_in) / w)))) * (1.0 / (1.0 + exp(-((S.r_out - rc) / w))));
              }
              L.intensity = I;
              cls = GG_CLASS_DISK;
            }
          }
        }
        if (cls == 255) {
          if (res == RES_HORIZON) cls = GG_CLASS_HORIZON;
          else if (res == RES_ESCAPED) cls = GG_CLASS_ESCAPED;
        }
      }
    }
    if (cls == GG_CLASS_MAX_STEPS && DISK == 1 && L.intensity > 0.0) cls = GG_CLASS_DISK;
    if (cls != 255) {
      // retire: structure-of-arrays stores of the final state
      const DevOut &O = T.outs[single ? 0 : S.out_index];
      if (O.cls) O.cls[out_i] = cls;
      if (O.r) O.r[out_i] = L.pos[1];
      if (O.theta) O.theta[out_i] = L.pos[2];
      if (O.phi) O.phi[out_i] = L.pos[3];
      if (O.intensity) O.intensity[out_i] = L.intensity;
      if (O.attempts) O.attempts[out_i] = L.attempts;
      if (O.accepted) O.accepted[out_i] = L.accepted;
      if (O.state) {
        const uint64_t p = O.n_rays;
        O.state[out_i] = L.pos[0];
        O.state[p + out_i] = L.pos[1];
        O.state[2 * p + out_i] = L.pos[2];
        O.state[3 * p + out_i] = L.pos[3];
        O.state[4 * p + out_i] = L.vel.x;
        O.state[5 * p + out_i] = L.vel.y;
        O.state[6 * p + out_i] = L.vel.z;
        O.state[7 * p + out_i] = L.vel.w;
        O.state[8 * p + out_i] = L.lambda;
      }
      active = false;
    }
  }
  // ---- tallies: one atomic per wavefront ----
  unsigned long long a = wave_sum(n_att), r = wave_sum(L.n_rhs), r32 = wave_sum(L.n_rhs32),
                     f = wave_sum(L.n_full);
  if (lane == 0) {
    atomicAdd(&T.counters->attempts, a);
    atomicAdd(&T.counters->rhs, r);
    atomicAdd(&T.counters->rhs32, r32);
    atomicAdd(&T.counters->attempts_full, f);
  }
}

__global__ void reset_kernel(Counters *c) {
  if (threadIdx.x < sizeof(Counters) / 8) reinterpret_cast<unsigned long long *>(c)[threadIdx.x] = 0ull;
}

template <int METRIC, int PREC>
__global__ void rhs_kernel(const DevScene *S, uint64_t n, const double *pos, const double *vel, double *acc) {
  uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  d4 v{vel[4 * i], vel[4 * i + 1], vel[4 * i + 2], vel[4 * i + 3]};
  d4 a;
  if (PREC == 1) {
    f4 vf{(float)v.x, (float)v.y, (float)v.z, (float)v.w};
    a = to64(rhs32<METRIC>(*S, (float)pos[4 * i + 1], (float)pos[4 * i + 2], vf));
  } else {
    a = rhs64<METRIC>(*S, pos[4 * i + 1], pos[4 * i + 2], v);
  }
  acc[4 * i] = a.x; acc[4 * i + 1] = a.y; acc[4 * i + 2] = a.z; acc[4 * i + 3] = a.w;
}

// ProceduralSky::sample (sky.rs:26-40).
__device__ inline void sky(double d0, double d1, double d2, double rgb[3]) {
  auto hash = [](double x) -> uint64_t {
    return (uint64_t)__double_as_longlong(x) * 6364136223846793005ull + 1442695040888963407ull;
  };
  auto rotl = [](uint64_t x, int k) -> uint64_t { return (x << k) | (x >> (64 - k)); };
  uint64_t hh = hash(d0) ^ rotl(hash(d1), 17) ^ rotl(hash(d2), 34);
  if (hh % 512 < 3) {
    double b = (double)(uint32_t)(180 + (uint32_t)(hh % 75)) / 255.0;
    rgb[0] = rgb[1] = rgb[2] = b;
  } else {
    double v = fabs(d2) * 15.0 / 255.0;
    rgb[0] = 0.0; rgb[1] = 0.0; rgb[2] = 10.0 / 255.0 + v;
  }
}

__device__ inline uint8_t to_u8(double c) {
#pragma clang fp contract(off)
  if (c < 0.0) c = 0.0;
  if (c > 1.0) c = 1.0;
  double s = (c <= 0.0031308) ? 12.92 * c : 1.055 * pow(c, 1.0 / 2.4) - 0.055;
  double r = round(s * 255.0);
  if (!(r > 0.0)) return 0;
  if (r >= 255.0) return 255;
  return (uint8_t)r;
}

// render_row (renderer.rs:309-345) shading: average of the spa^2 linear
// colours (shade_outcome_with, disk_color) then encode_srgb.
__global__ void shade_kernel(const DevScene *Sp, const DevOut *Op, uint32_t width, uint32_t height, uint32_t spa,
                             uint8_t *rgb) {
#pragma clang fp contract(off)
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= width * height) return;
  const DevOut &O = *Op;
  uint32_t px = i % width, py = i / width;
  uint32_t sw = width * spa;
  double acc[3] = {0.0, 0.0, 0.0};
  for (uint32_t sy = 0; sy < spa; ++sy)
    for (uint32_t sx = 0; sx < spa; ++sx) {
      uint64_t k = (uint64_t)(py * spa + sy) * sw + (px * spa + sx);
      double lin[3];
      uint8_t c = O.cls[k];
      if (c == GG_CLASS_HORIZON) {
        lin[0] = lin[1] = lin[2] = 0.0;
      } else if (c == GG_CLASS_ESCAPED) {
        const uint64_t p = O.n_rays;
        double v1 = O.state[5 * p + k], v2 = O.state[6 * p + k], v3 = O.state[7 * p + k];
        double nn = sqrt(v1 * v1 + v2 * v2 + v3 * v3);
        double d0 = 1.0, d1 = 0.0, d2 = 0.0;
        if (nn > 0.0) { d0 = v1 / nn; d1 = v2 / nn; d2 = v3 / nn; }
        sky(d0, d1, d2, lin);
      } else if (c == GG_CLASS_DISK) {
        double I = O.intensity[k];
        double x = I / (1.0 + I);
        lin[0] = pow(x, 0.5); lin[1] = pow(x, 1.2); lin[2] = pow(x, 2.5);
      } else {
        lin[0] = 0.0; lin[1] = 0.0; lin[2] = 10.0 / 255.0;
      }
      acc[0] += lin[0]; acc[1] += lin[1]; acc[2] += lin[2];
    }
  double inv = 1.0 / (double)(spa * spa);
  for (int c = 0; c < 3; ++c) rgb[3 * (uint64_t)i + c] = to_u8(acc[c] * inv);
}

template <int M, int D, int P>
hipError_t launch3(const TraceLaunch &t, uint32_t grid, uint32_t block, hipStream_t s) {
  if (t.states_in)
    hipLaunchKernelGGL((trace_kernel<M, D, P, 1, 0>), dim3(grid), dim3(block), 0, s, t);
  else if (t.n_scenes > 1)
    hipLaunchKernelGGL((trace_kernel<M, D, P, 0, 1>), dim3(grid), dim3(block), 0, s, t);
  else
    hipLaunchKernelGGL((trace_kernel<M, D, P, 0, 0>), dim3(grid), dim3(block), 0, s, t);
  return hipGetLastError();
}

template <int M, int D>
hipError_t launch2(int prec, const TraceLaunch &t, uint32_t grid, uint32_t block, hipStream_t s) {
  return prec == 1 ? launch3<M, D, 1>(t, grid, block, s) : launch3<M, D, 0>(t, grid, block, s);
}

template <int M>
hipError_t launch1(int disk, int prec, const TraceLaunch &t, uint32_t grid, uint32_t block, hipStream_t s) {
  switch (disk) {
    case 1: return launch2<M, 1>(prec, t, grid, block, s);
    case 2: return launch2<M, 2>(prec, t, grid, block, s);
    default: return launch2<M, 0>(prec, t, grid, block, s);
  }
}

}  // namespace

hipError_t launch_trace(int metric_kind, int disk_kind, int precision, const TraceLaunch &t, uint32_t grid,
                        uint32_t block, hipStream_t stream) {
  if (metric_kind == 1) return launch1<1>(disk_kind, precision, t, grid, block, stream);
  return launch1<0>(disk_kind, precision, t, grid, block, stream);
}

hipError_t launch_reset(Counters *c, hipStream_t stream) {
  hipLaunchKernelGGL(reset_kernel, dim3(1), dim3(64), 0, stream, c);
  return hipGetLastError();
}

hipError_t launch_rhs(int metric_kind, int precision, const DevScene *scene, uint64_t n, const double *pos,
                      const double *vel, double *acc, hipStream_t stream) {
  if (n == 0) return hipSuccess;
  uint32_t grid = (uint32_t)((n + 255) / 256);
  if (metric_kind == 1) {
    if (precision == 1) hipLaunchKernelGGL((rhs_kernel<1, 1>), dim3(grid), dim3(256), 0, stream, scene, n, pos, vel, acc);
    else hipLaunchKernelGGL((rhs_kernel<1, 0>), dim3(grid), dim3(256), 0, stream, scene, n, pos, vel, acc);
  } else {
    if (precision == 1) hipLaunchKernelGGL((rhs_kernel<0, 1>), dim3(grid), dim3(256), 0, stream, scene, n, pos, vel, acc);
    else hipLaunchKernelGGL((rhs_kernel<0, 0>), dim3(grid), dim3(256), 0, stream, scene, n, pos, vel, acc);
  }
  return hipGetLastError();
}

hipError_t launch_shade(const DevScene *scene, const DevOut *out, uint32_t width, uint32_t height, uint32_t spa,
                        uint8_t *rgb, hipStream_t stream) {
  uint32_t n = width * height;
  hipLaunchKernelGGL(shade_kernel, dim3((n + 255) / 256), dim3(256), 0, stream, scene, out, width, height, spa, rgb);
  return hipGetLastError();
}

}  // namespace gg
