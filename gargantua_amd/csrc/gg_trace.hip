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

// Stage tableau in constant memory, indexed by the (wavefront-uniform) stage
// number s = 2..7 of k_s; the values are the reference's double quotients.
__constant__ double kA[7] = {0.0, 0.0, A2, A3, A4, A5, 1.0};
__constant__ double kB[7][6] = {
    {0, 0, 0, 0, 0, 0}, {0, 0, 0, 0, 0, 0}, {0, B21, 0, 0, 0, 0}, {0, B31, B32, 0, 0, 0},
    {0, B41, B42, B43, 0, 0}, {0, B51, B52, B53, B54, 0}, {0, B61, B62, B63, B64, B65}};
__constant__ double kC[7] = {0.0, C1, 0.0, C3, C4, C5, C6};
__constant__ double kD[8] = {0.0, D1, 0.0, D3, D4, D5, D6, D7};

struct Lane {
  double pos[4];
  d4 vel;
  double lambda;
  double h;
  bool k1_ok;        // LDS slot 0 holds k1 = a(pos, vel)
  bool k1_f32;       // ... evaluated by an FP32 attempt (mixed kernel)
  uint32_t attempts, accepted;
  double intensity, transmission;
  double err_prev;   // error ratio of the previous full attempt
  uint32_t n_rhs, n_rhs32, n_full;
};

// Per-lane view of the LDS stage store: slots 0..4 hold k1..k5, structure of
// arrays [slot][component][lane] so each ds_read_b64 of a wavefront touches
// 64 consecutive doubles.
struct KStore {
  double *p;        // &lds[threadIdx.x]
  uint32_t stride;  // blockDim.x
  __device__ double &at(int slot, int c) const { return p[(slot * 4 + c) * stride]; }
  __device__ d4 get(int slot) const { return d4{at(slot, 0), at(slot, 1), at(slot, 2), at(slot, 3)}; }
  __device__ void put(int slot, d4 k) const { at(slot, 0) = k.x; at(slot, 1) = k.y; at(slot, 2) = k.z; at(slot, 3) = k.w; }
};

__device__ inline bool valid4(double p0, double p1, double p2, double p3) {
  return isfinite(p0) && isfinite(p1) && isfinite(p2) && isfinite(p3) && p1 > 0.0;
}

__device__ inline d4 to64(f4 a) { return d4{(double)a.x, (double)a.y, (double)a.z, (double)a.w}; }

// Arithmetic of one attempt.  ARITH_EXACT: FP64 RHS and combinations in the
// reference's expression order, nothing fused.  ARITH_F32: FP32 RHS, fused
// combinations.  ARITH_FAST: reciprocal-form FP64 RHS, fused combinations.
enum { ARITH_EXACT = 0, ARITH_F32 = 1, ARITH_FAST = 2 };

template <int METRIC, int ARITH>
__device__ inline d4 eval_rhs(const DevScene &S, double r, double th, d4 v) {
  if (ARITH == ARITH_EXACT) return rhs64<METRIC>(S, r, th, v);
  if (ARITH == ARITH_F32) {
    f4 vf{(float)v.x, (float)v.y, (float)v.z, (float)v.w};
    return to64(rhs32<METRIC>(S, r, th, vf));
  }
  return rhs64f<METRIC>(S, r, th, v);
}

// a*b + c: unfused (reference rounding) for ARITH_EXACT, one FMA otherwise.
template <int ARITH>
__device__ inline double madd(double a, double b, double c) {
#pragma clang fp contract(off)
  if (ARITH == ARITH_EXACT) return a * b + c;
  return __builtin_fma(a, b, c);
}

// Bound on what FP32 rounding of the six stages can add to the error ratio.
// One FP32 RHS evaluation errs by a few eps32 times the largest term of each
// component's contraction; T_c bounds those terms with the Schwarzschild-form
// symbols (the spin terms are O(a/r) of them beyond 10 M), and the error
// estimate weighs the stages by |c_i - d_i|, which sum to < 0.2.
__device__ inline float fp32_noise(const DevScene &S, const double pos[4], d4 v, double h) {
#pragma clang fp contract(fast)
  float r = (float)pos[1];
  float st, ct;
  sincosf((float)pos[2], &st, &ct);
  float ir = 1.0f / r;
  float m_r2 = (float)S.mass * ir * ir;
  float vt = fabsf((float)v.x), vr = fabsf((float)v.y), vth = fabsf((float)v.z), vph = fabsf((float)v.w);
  float T0 = 2.0f * m_r2 * vt * (vr + vph * r * (float)fabs(S.spin) * ir);
  float T1 = m_r2 * (vt * vt + vr * vr) + r * (vth * vth + st * st * vph * vph);
  float T2 = 2.0f * vr * vth * ir + fabsf(st * ct) * vph * vph;
  float T3 = 2.0f * vr * vph * ir + 2.0f * fabsf(ct / st) * vth * vph;
  const float k = 2.0e-7f * (float)h;   // 2 (safety) x 0.2 x 8 eps32
  if (S.err_mode == 1) {
    float atol = (float)S.atol, rtol = (float)S.rtol;
    float n0 = T0 / (atol + rtol * vt), n1 = T1 / (atol + rtol * vr);
    float n2 = T2 / (atol + rtol * vth), n3 = T3 / (atol + rtol * vph);
    return k * fmaxf(fmaxf(n0, n1), fmaxf(n2, n3));
  }
  return k * sqrtf(T0 * T0 + T1 * T1 + T2 * T2 + T3 * T3) / (float)S.tol;
}

struct StageOut {
  d4 nv, k7;
  double err, thresh;
  int nr;
  bool invalid;
};

// The stages of one attempt (geodesic.rs:244-308): k1 when the lane has no
// valid one (s = 1), then k2..k7, the 5th-order velocity and the error.  One
// copy of the loop per arithmetic (ARITH_EXACT, ARITH_F32, ARITH_FAST).
template <int METRIC, int ARITH>
__device__ inline StageOut stages(const DevScene &S, const KStore &K, const double p0[4], const d4 vel, double h,
                                  bool need_k1) {
#pragma clang fp contract(off)
  StageOut o;
  d4 k1{0, 0, 0, 0}, s5{0, 0, 0, 0}, s4{0, 0, 0, 0};
  o.nv = d4{0, 0, 0, 0};
  o.k7 = d4{0, 0, 0, 0};
  o.nr = 0;
  o.invalid = false;
#pragma unroll 1
  for (int s = 1; s <= 7; ++s) {
    double qr, qth;
    d4 vs;
    bool eval = true;
    if (s == 1) {
      qr = p0[1];                                                   // k1 = a(pos, vel) (:245)
      qth = p0[2];
      vs = vel;
      eval = need_k1;
    } else if (s <= 6) {
      // stage position pos + (h a_s) vel on the forward-Euler line (:248-284)
      double ha = h * kA[s];
      double q0 = madd<ARITH>(ha, vel.x, p0[0]), q1 = madd<ARITH>(ha, vel.y, p0[1]);
      double q2 = madd<ARITH>(ha, vel.z, p0[2]), q3 = madd<ARITH>(ha, vel.w, p0[3]);
      if (!valid4(q0, q1, q2, q3)) { o.invalid = true; break; }   // :250-291 halve h
      if (s == 2) {
        double hb = h * B21;                                        // vel + (h b21) k1 (:249)
        vs = d4{madd<ARITH>(hb, k1.x, vel.x), madd<ARITH>(hb, k1.y, vel.y), madd<ARITH>(hb, k1.z, vel.z),
                madd<ARITH>(hb, k1.w, vel.w)};
      } else {
        double b = kB[s][1];
        d4 sum{b * k1.x, b * k1.y, b * k1.z, b * k1.w};
#pragma unroll 1
        for (int j = 2; j < s; ++j) {
          double bj = kB[s][j];
          d4 kj = K.get(j - 1);
          sum.x = madd<ARITH>(bj, kj.x, sum.x); sum.y = madd<ARITH>(bj, kj.y, sum.y);
          sum.z = madd<ARITH>(bj, kj.z, sum.z); sum.w = madd<ARITH>(bj, kj.w, sum.w);
        }
        vs = d4{madd<ARITH>(h, sum.x, vel.x), madd<ARITH>(h, sum.y, vel.y), madd<ARITH>(h, sum.z, vel.z),
                madd<ARITH>(h, sum.w, vel.w)};
      }
      qr = q1;
      qth = q2;
    } else {
      qr = madd<ARITH>(h, vel.y, p0[1]);                            // k7 at (pos6, new_vel) (:298)
      qth = madd<ARITH>(h, vel.z, p0[2]);
      vs = o.nv;
    }
    d4 k;
    if (eval) {
      k = eval_rhs<METRIC, ARITH>(S, qr, qth, vs);
      ++o.nr;
    } else {
      k = K.get(0);
    }
    if (s <= 5) K.put(s - 1, k);
    if (s == 1) {
      k1 = k;
      s5 = d4{C1 * k1.x, C1 * k1.y, C1 * k1.z, C1 * k1.w};       // sum c_i k_i (:294-295)
      s4 = d4{D1 * k1.x, D1 * k1.y, D1 * k1.z, D1 * k1.w};       // sum d_i k_i (:299-305)
    } else if (s >= 3) {
      if (s <= 6) {
        double c = kC[s];
        s5.x = madd<ARITH>(c, k.x, s5.x); s5.y = madd<ARITH>(c, k.y, s5.y);
        s5.z = madd<ARITH>(c, k.z, s5.z); s5.w = madd<ARITH>(c, k.w, s5.w);
      }
      double d = kD[s];
      s4.x = madd<ARITH>(d, k.x, s4.x); s4.y = madd<ARITH>(d, k.y, s4.y);
      s4.z = madd<ARITH>(d, k.z, s4.z); s4.w = madd<ARITH>(d, k.w, s4.w);
    }
    if (s == 6)
      o.nv = d4{madd<ARITH>(h, s5.x, vel.x), madd<ARITH>(h, s5.y, vel.y), madd<ARITH>(h, s5.z, vel.z),
                madd<ARITH>(h, s5.w, vel.w)};
    if (s == 7) o.k7 = k;
  }
  if (o.invalid) return o;
  const d4 nv = o.nv;
  d4 ve{madd<ARITH>(h, s4.x, vel.x), madd<ARITH>(h, s4.y, vel.y), madd<ARITH>(h, s4.z, vel.z),
        madd<ARITH>(h, s4.w, vel.w)};
  double dx = nv.x - ve.x, dy = nv.y - ve.y, dz = nv.z - ve.z, dw = nv.w - ve.w;
  if (S.err_mode == 1) {
    double s0 = dx / (S.atol + S.rtol * fmax(fabs(vel.x), fabs(nv.x)));
    double s1 = dy / (S.atol + S.rtol * fmax(fabs(vel.y), fabs(nv.y)));
    double s2 = dz / (S.atol + S.rtol * fmax(fabs(vel.z), fabs(nv.z)));
    double s3 = dw / (S.atol + S.rtol * fmax(fabs(vel.w), fabs(nv.w)));
    double sq = 0.0;
    sq = sq + s0 * s0; sq = sq + s1 * s1; sq = sq + s2 * s2; sq = sq + s3 * s3;
    o.err = sqrt(0.25 * sq);
    o.thresh = 1.0;
  } else {
    // nalgebra Vector4::norm: sqrt((x0^2 + x2^2) + (x1^2 + x3^2))  (:308)
    double a = dx * dx, b = dy * dy, c = dz * dz, e = dw * dw;
    a = a + c;
    b = b + e;
    o.err = sqrt(a + b);
    o.thresh = S.tol;
  }
  return o;
}

// Whether the next attempt of this ray may run its stages in FP32: the
// controller is saturated (h == max_step, previous error ratio < 0.3), the
// ray is outgoing beyond fp32_min_r (photons outside the photon-orbit region
// never turn back, so FP32 rounding is not amplified by a later close
// approach) or inbound beyond fp32_far_r, theta is away from the poles and
// the FP32 noise bound is below 0.02 in error-ratio units.  A function of the
// ray's own state only, so every ray's arithmetic is independent of which
// lane or wavefront carries it.
__device__ inline bool fp32_eligible(const DevScene &S, const Lane &L) {
  const double r = L.pos[1], th = L.pos[2];
  bool far_enough = (L.vel.y > 0.0) ? (r >= S.fp32_min_r) : (r >= S.fp32_far_r);
  if (!(far_enough && (L.h == S.max_step) && (L.err_prev < 0.3) && (th > 1e-3) && (th < PI - 1e-3)))
    return false;
  return fp32_noise(S, L.pos, L.vel, L.h) < 0.02f;
}

// One RK45Integrator::step (geodesic.rs:185-328) on lane state L with the
// stage values in LDS.  PREC 0: exact FP64; 2: fast FP64; 1: mixed: a
// wavefront runs its stages in FP32 when every live lane's ray is
// FP32-eligible, and in exact FP64 otherwise (FP64 is always allowed), so a
// wavefront never executes both loops.  Every accept/reject decision and
// every h equal those of the FP64 kernel either way; which FP32-eligible
// steps do run in FP32 depends on the lane grouping, so mixed final states
// reproduce across launches to FP32 rounding (~1e-12 relative), not bitwise.
// Returns RES_*; `acc` reports acceptance; p0/v0 keep the pre-step state.
template <int METRIC, int PREC>
__device__ inline int attempt(const DevScene &S, Lane &L, const KStore &K, bool &acc, double p0[4], d4 &v0,
                              bool use32) {
#pragma clang fp contract(off)
  constexpr int BASE = (PREC == 2) ? ARITH_FAST : ARITH_EXACT;
  acc = false;
  const double r = L.pos[1];
  if (r > S.max_radius) return RES_ESCAPED;          // geodesic.rs:194-196
  if (r < S.r_capture) return RES_HORIZON;           // :198-202
  if (r < S.r_h) return RES_HORIZON;                 // :204-206
  const d4 vel = L.vel;
  for (int i = 0; i < 4; ++i) p0[i] = L.pos[i];
  v0 = vel;
  const double h = L.h;
  StageOut o;
  bool done = false;
  if (PREC == 1 && use32) {
    const bool need_k1 = !L.k1_ok;
    o = stages<METRIC, ARITH_F32>(S, K, p0, vel, h, need_k1);
    L.n_rhs += o.nr;
    L.n_rhs32 += o.nr;
    if (need_k1) { L.k1_ok = true; L.k1_f32 = true; }   // slot 0 now holds an FP32 k1
    // an FP32 result near a decision (or with an invalid stage) is redone in FP64
    done = !o.invalid && (o.err / o.thresh < 0.4);
    if (!done) use32 = false;
  }
  if (!done) {
    // An FP64 attempt needs an FP64 k1: after an FP32 attempt k7 (the FSAL
    // k1) carries FP32 rounding, so recompute it.
    const bool need_k1 = !L.k1_ok || L.k1_f32;
    o = stages<METRIC, BASE>(S, K, p0, vel, h, need_k1);
    L.n_rhs += o.nr;
    if (need_k1) L.k1_f32 = false;
  }
  L.k1_ok = true;
  if (o.invalid) { L.h = h * 0.5; return RES_CONTINUE; }
  const double err = o.err, thresh = o.thresh;
  L.n_full++;
  L.err_prev = err / thresh;
  if (err < thresh || h <= S.min_step) {           // accept (:310-321)
    L.pos[0] = madd<BASE>(h, vel.x, p0[0]); L.pos[1] = madd<BASE>(h, vel.y, p0[1]);
    L.pos[2] = madd<BASE>(h, vel.z, p0[2]); L.pos[3] = madd<BASE>(h, vel.w, p0[3]);
    L.vel = o.nv;
    L.lambda = L.lambda + h;
    // normalize_spherical_coords (:337-352)
    bool refl = false;
    if (L.pos[2] < 0.0) { L.pos[2] = -L.pos[2]; L.pos[3] = L.pos[3] + PI; refl = true; }
    if (L.pos[2] > PI) { L.pos[2] = 2.0 * PI - L.pos[2]; L.pos[3] = L.pos[3] + PI; refl = true; }
    double ph = fmod(L.pos[3], TWO_PI);
    L.pos[3] = ph < 0.0 ? ph + TWO_PI : ph;
    K.put(0, o.k7);                                 // FSAL: k7 is the next k1 ...
    L.k1_ok = !refl;                                // ... unless theta was reflected
    L.k1_f32 = use32;
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
#ifndef GG_MIN_WAVES
#define GG_MIN_WAVES 2
#endif
__global__ __launch_bounds__(256, GG_MIN_WAVES) void trace_kernel(TraceLaunch T) {
#pragma clang fp contract(off)
  const uint32_t lane = threadIdx.x & 63;
  const uint64_t n = T.n_rays;
  constexpr bool single = (BATCH == 0);

  extern __shared__ double lds_k[];
  const KStore K{lds_k + threadIdx.x, blockDim.x};
  Lane L;
  L.n_rhs = 0; L.n_rhs32 = 0; L.n_full = 0;
  uint32_t n_att = 0, n_retired = 0;
  bool active = false;
  uint64_t ray = 0;
  uint32_t si = 0;        // scene of this lane's ray
  uint64_t out_i = 0;     // output slot of this lane's ray
  bool exhausted = false;
  // Scene constants are read through the pointer where they are used, so
  // the compiler reloads them (scalar loads for a single scene) instead of
  // holding ~75 doubles live across the persistent loop.
  const DevScene *__restrict__ Sp = T.scenes;

  // Start ray `ray` in this lane.
  auto init_ray = [&](uint64_t rid) {
    ray = rid;
    si = single ? 0u : find_scene(T.scenes, T.n_scenes, ray);
    if (!single) Sp = T.scenes + si;
    uint64_t local = ray - Sp->ray_begin;
    if (SRC == 0) {
      uint32_t px, py;
      ray_pixel(*Sp, local, px, py);
      out_i = (uint64_t)py * Sp->sub_w + px;
      camera_ray(*Sp, px, py, L.pos, L.vel);
      L.lambda = 0.0;
    } else {
      out_i = local;
      const double *st = T.states_in + 9 * ray;
      for (int c = 0; c < 4; ++c) L.pos[c] = st[c];
      L.vel = d4{st[4], st[5], st[6], st[7]};
      L.lambda = st[8];
    }
    L.h = Sp->h0;
    L.k1_ok = false;
    L.k1_f32 = false;
    L.attempts = 0; L.accepted = 0;
    L.intensity = 0.0; L.transmission = 1.0;
    L.err_prev = 1.0;
    active = true;
  };

  __shared__ uint32_t s_cnt[2][4];
  __shared__ unsigned long long s_base;
  uint32_t iter = 0;
  for (;;) {
    {
      // ---- block regroup: compaction (and precision sorting when mixed) ----
      // Every refill_interval attempts the block's wavefronts exchange ray
      // states through LDS: live rays move to the lowest lanes (for the mixed
      // kernel FP32-eligible rays first, so that whole wavefronts can take
      // FP32 steps), retired lanes gather at the tail and take new rays with
      // one atomicAdd for the block.
      if (iter % T.refill_interval == 0) {
        const uint32_t tid = threadIdx.x, B = blockDim.x, wave = tid >> 6;
        const uint32_t key = !active ? 2u : ((PREC == 1 && fp32_eligible(*Sp, L)) ? 0u : 1u);
        const unsigned long long m0 = __ballot(key == 0), m1 = __ballot(key == 1);
        if (lane == 0) { s_cnt[0][wave] = __popcll(m0); s_cnt[1][wave] = __popcll(m1); }
        __syncthreads();
        uint32_t n0 = 0, n1 = 0, o0 = 0, o1 = 0;
        for (uint32_t w = 0; w < (B >> 6); ++w) {
          if (w < wave) { o0 += s_cnt[0][w]; o1 += s_cnt[1][w]; }
          n0 += s_cnt[0][w];
          n1 += s_cnt[1][w];
        }
        const unsigned long long below = (1ull << lane) - 1ull;
        const uint32_t dst = key == 0 ? o0 + __popcll(m0 & below) : n0 + o1 + __popcll(m1 & below);
        const uint32_t n_act = n0 + n1;
        // The exchange reuses stage slots 1..4 (free between attempts) as
        // 16 planes of B doubles; two passes of at most 14 fields.
        double *X = lds_k + 4 * B;
        d4 k1 = K.get(0);
        __syncthreads();
        if (active) {
          X[0 * B + dst] = L.pos[0]; X[1 * B + dst] = L.pos[1]; X[2 * B + dst] = L.pos[2]; X[3 * B + dst] = L.pos[3];
          X[4 * B + dst] = L.vel.x; X[5 * B + dst] = L.vel.y; X[6 * B + dst] = L.vel.z; X[7 * B + dst] = L.vel.w;
          X[8 * B + dst] = L.lambda; X[9 * B + dst] = L.h;
          X[10 * B + dst] = k1.x; X[11 * B + dst] = k1.y; X[12 * B + dst] = k1.z; X[13 * B + dst] = k1.w;
        }
        __syncthreads();
        const bool now_active = tid < n_act;
        if (now_active) {
          L.pos[0] = X[0 * B + tid]; L.pos[1] = X[1 * B + tid]; L.pos[2] = X[2 * B + tid]; L.pos[3] = X[3 * B + tid];
          L.vel = d4{X[4 * B + tid], X[5 * B + tid], X[6 * B + tid], X[7 * B + tid]};
          L.lambda = X[8 * B + tid]; L.h = X[9 * B + tid];
          k1 = d4{X[10 * B + tid], X[11 * B + tid], X[12 * B + tid], X[13 * B + tid]};
        }
        __syncthreads();
        if (active) {
          X[0 * B + dst] = L.intensity; X[1 * B + dst] = L.transmission; X[2 * B + dst] = L.err_prev;
          X[3 * B + dst] = __longlong_as_double((long long)ray);
          X[4 * B + dst] = __longlong_as_double((long long)out_i);
          X[5 * B + dst] = __longlong_as_double((long long)(((uint64_t)L.attempts << 32) | L.accepted));
          X[6 * B + dst] = __longlong_as_double(
              (long long)(((uint64_t)si << 2) | (L.k1_f32 ? 2u : 0u) | (L.k1_ok ? 1u : 0u)));
        }
        __syncthreads();
        if (now_active) {
          L.intensity = X[0 * B + tid]; L.transmission = X[1 * B + tid]; L.err_prev = X[2 * B + tid];
          ray = (uint64_t)__double_as_longlong(X[3 * B + tid]);
          out_i = (uint64_t)__double_as_longlong(X[4 * B + tid]);
          uint64_t aa = (uint64_t)__double_as_longlong(X[5 * B + tid]);
          L.attempts = (uint32_t)(aa >> 32); L.accepted = (uint32_t)aa;
          uint64_t sf = (uint64_t)__double_as_longlong(X[6 * B + tid]);
          si = (uint32_t)(sf >> 2); L.k1_ok = (sf & 1u) != 0; L.k1_f32 = (sf & 2u) != 0;
          if (!single) Sp = T.scenes + si;
          K.put(0, k1);
        }
        active = now_active;
        // refill the tail lanes (NO_FETCH: the block was full or the queue
        // already ran dry, nothing was taken this round)
        constexpr unsigned long long NO_FETCH = ~0ull;
        if (tid == 0) {
          unsigned long long base = NO_FETCH;
          if (!exhausted && n_act < B) base = atomicAdd(&T.counters->queue, (unsigned long long)(B - n_act));
          s_base = base;
        }
        __syncthreads();
        const unsigned long long base = s_base;
        if (base != NO_FETCH) {
          if (base >= n) exhausted = true;
          if (!active && base + (tid - n_act) < n) init_ray(base + (tid - n_act));
        }
        // leave together once the queue is empty and no lane of the block is live
        if (exhausted && n_act == 0 && __syncthreads_or(active) == 0) break;
      }
      ++iter;
      if (!active) continue;
    }

    // ---- one step() call ----
    uint8_t cls = 255;
    if (L.attempts >= Sp->max_steps) {
      cls = GG_CLASS_MAX_STEPS;
    } else {
      bool acc;
      double p0[4];
      d4 v0;
      double lam0 = L.lambda;
      bool use32 = false;
      if (PREC == 1) use32 = __all(fp32_eligible(*Sp, L));   // over the lanes still running
      int res = attempt<METRIC, PREC>(*Sp, L, K, acc, p0, v0, use32);
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
          double rho = disk_density(*Sp, mr, mth);
          if (rho > 0.0) {
            double j = disk_emission(*Sp, rho, mr, mth);
            d4 pm{0.5 * (v0.x + L.vel.x), 0.5 * (v0.y + L.vel.y), 0.5 * (v0.z + L.vel.z), 0.5 * (v0.w + L.vel.w)};
            double g = redshift64<METRIC>(*Sp, mr, mth, pm);
            double g2 = g * g;
            double red = j * (g2 * g2);
            L.intensity = L.intensity + red * dl * L.transmission;
            L.transmission *= exp(-rho * dl * Sp->absorption);
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
            double w = Sp->edge_width;
            bool hit = (w > 0.0) ? (rc <= Sp->r_out + 8.0 * w && rc > 0.0) : (rc >= Sp->r_in && rc <= Sp->r_out);
            if (hit) {
              d4 vc{v0.x + s * (L.vel.x - v0.x), v0.y + s * (L.vel.y - v0.y), v0.z + s * (L.vel.z - v0.z),
                    v0.w + s * (L.vel.w - v0.w)};
              double g = redshift64<METRIC>(*Sp, rc, HALF_PI, vc);
              double g2 = g * g;
              double I = pow(rc, -Sp->q) * (g2 * g2);
              if (w > 0.0) {
                I = I * (1.0 / (1.0 + exp(-((rc - Sp->r_in) / w)))) * (1.0 / (1.0 + exp(-((Sp->r_out - rc) / w))));
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
      const DevOut &O = T.outs[single ? 0 : Sp->out_index];
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
      n_retired++;
    }
  }
  // ---- tallies: one atomic per wavefront ----
  unsigned long long a = wave_sum(n_att), r = wave_sum(L.n_rhs), r32 = wave_sum(L.n_rhs32),
                     f = wave_sum(L.n_full), ret = wave_sum(n_retired);
  if (lane == 0) {
    atomicAdd(&T.counters->retired, ret);
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
  if (PREC == 1) a = eval_rhs<METRIC, ARITH_F32>(*S, pos[4 * i + 1], pos[4 * i + 2], v);
  else if (PREC == 2) a = eval_rhs<METRIC, ARITH_FAST>(*S, pos[4 * i + 1], pos[4 * i + 2], v);
  else a = eval_rhs<METRIC, ARITH_EXACT>(*S, pos[4 * i + 1], pos[4 * i + 2], v);
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

// k1..k5 of every lane: 5 slots x 4 components x 8 bytes.
inline size_t lds_bytes(uint32_t block) { return (size_t)block * 5 * 4 * sizeof(double); }

template <int M, int D, int P>
hipError_t launch3(const TraceLaunch &t, uint32_t grid, uint32_t block, hipStream_t s) {
  if (t.states_in)
    hipLaunchKernelGGL((trace_kernel<M, D, P, 1, 0>), dim3(grid), dim3(block), lds_bytes(block), s, t);
  else if (t.n_scenes > 1)
    hipLaunchKernelGGL((trace_kernel<M, D, P, 0, 1>), dim3(grid), dim3(block), lds_bytes(block), s, t);
  else
    hipLaunchKernelGGL((trace_kernel<M, D, P, 0, 0>), dim3(grid), dim3(block), lds_bytes(block), s, t);
  return hipGetLastError();
}

template <int M, int D>
hipError_t launch2(int prec, const TraceLaunch &t, uint32_t grid, uint32_t block, hipStream_t s) {
  if (prec == 1) return launch3<M, D, 1>(t, grid, block, s);
  if (prec == 2) return launch3<M, D, 2>(t, grid, block, s);
  return launch3<M, D, 0>(t, grid, block, s);
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
#define GG_RHS_LAUNCH(M, P) hipLaunchKernelGGL((rhs_kernel<M, P>), dim3(grid), dim3(256), 0, stream, scene, n, pos, vel, acc)
  if (metric_kind == 1) {
    if (precision == 1) GG_RHS_LAUNCH(1, 1); else if (precision == 2) GG_RHS_LAUNCH(1, 2); else GG_RHS_LAUNCH(1, 0);
  } else {
    if (precision == 1) GG_RHS_LAUNCH(0, 1); else if (precision == 2) GG_RHS_LAUNCH(0, 2); else GG_RHS_LAUNCH(0, 0);
  }
#undef GG_RHS_LAUNCH
  return hipGetLastError();
}

hipError_t launch_shade(const DevScene *scene, const DevOut *out, uint32_t width, uint32_t height, uint32_t spa,
                        uint8_t *rgb, hipStream_t stream) {
  uint32_t n = width * height;
  hipLaunchKernelGGL(shade_kernel, dim3((n + 255) / 256), dim3(256), 0, stream, scene, out, width, height, spa, rgb);
  return hipGetLastError();
}

}  // namespace gg
