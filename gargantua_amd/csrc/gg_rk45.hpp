// gg_rk45.hpp -- one RK45Integrator::step on a lane's ray state, shared by
// the trace kernel and the adjoint kernel (both must take bit-identical
// steps: the adjoint recomputes trajectories the trace kernel rendered).
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
#pragma once

#ifndef __HIPCC_RTC__  // hipRTC supplies the runtime and fixed-width types itself
#include <hip/hip_runtime.h>
#endif

#include "gg_internal.hpp"
#include "gg_user_metric.hpp"

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
  bool k1_fast;      // ... evaluated by a fast-FP64 attempt (mixed kernel)
  uint32_t attempts, accepted;
  double intensity, transmission;
  double err_prev;   // error ratio of the previous full attempt
  double b2;         // (Q + L_z^2) / E^2 of the ray's start state (mixed kernel)
  uint32_t n_rhs, n_rhs32, n_rhsf, n_full;
  uint32_t n_refl;   // accepted steps whose theta was reflected at a pole
};

// Per-lane view of the LDS stage store: slots 0..4 hold k1..k5.  Slot s is
// two planes of 16-byte pairs, (x, y) of every lane in plane 2s and (z, w) in
// plane 2s + 1, so a slot moves with two ds_read_b128 / ds_write_b128 per
// lane and each plane a wavefront touches is 1 KiB of consecutive LDS.  N > 0
// fixes the block size (pairs per plane) at compile time, so that every slot
// offset is an instruction immediate.
template <uint32_t N = 0>
struct KStoreT {
  double2 *p;       // &plane0[threadIdx.x] (LDS base 16-byte aligned)
  uint32_t stride;  // blockDim.x: pairs per plane (N when N > 0)
  __device__ uint32_t st() const { return N ? N : stride; }
  __device__ d4 get(int slot) const {
    const double2 a = p[(2 * slot) * st()], b = p[(2 * slot + 1) * st()];
    return d4{a.x, a.y, b.x, b.y};
  }
  __device__ void put(int slot, d4 k) const {
    p[(2 * slot) * st()] = double2{k.x, k.y};
    p[(2 * slot + 1) * st()] = double2{k.z, k.w};
  }
};
using KStore = KStoreT<0>;

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
// An estimate compared with 0.02 and carrying a factor 2 of margin, so it
// takes the hardware sine, cosine and reciprocal (errors ~1e-6 relative; the
// eligibility test runs before most attempts of FP32-candidate rays).
__device__ inline float fp32_noise(const DevScene &S, const double pos[4], d4 v, double h) {
#pragma clang fp contract(fast)
  float r = (float)pos[1];
  const float th = (float)pos[2];
  const float st = __sinf(th), ct = __cosf(th);
  float ir = frcp(r);
  float m_r2 = (float)S.mass * ir * ir;
  float vt = fabsf((float)v.x), vr = fabsf((float)v.y), vth = fabsf((float)v.z), vph = fabsf((float)v.w);
  float T0 = 2.0f * m_r2 * vt * (vr + vph * r * (float)fabs(S.spin) * ir);
  float T1 = m_r2 * (vt * vt + vr * vr) + r * (vth * vth + st * st * vph * vph);
  float T2 = 2.0f * vr * vth * ir + fabsf(st * ct) * vph * vph;
  float T3 = 2.0f * vr * vph * ir + 2.0f * fabsf(ct * frcp(st)) * vth * vph;
  const float k = 2.0e-7f * (float)h;   // 2 (safety) x 0.2 x 8 eps32
  if (S.err_mode == 1) {
    float atol = (float)S.atol, rtol = (float)S.rtol;
    float n0 = T0 * frcp(atol + rtol * vt), n1 = T1 * frcp(atol + rtol * vr);
    float n2 = T2 * frcp(atol + rtol * vth), n3 = T3 * frcp(atol + rtol * vph);
    return k * fmaxf(fmaxf(n0, n1), fmaxf(n2, n3));
  }
  return k * sqrtf(T0 * T0 + T1 * T1 + T2 * T2 + T3 * T3) * frcp((float)S.tol);
}

struct StageOut {
  d4 nv, k7;
  double err, thresh;
  int nr;
  bool invalid;
};

// Step record of one attempt for the adjoint, 32 doubles (256 bytes, two
// 128-byte lines; records are 128-byte aligned; every 4-vector 32-byte
// aligned): start r and theta (REC_R, REC_TH; t and phi enter nothing the
// adjoint differentiates), h (REC_H), the attempt's kind (REC_KIND: REC_ACCEPT,
// REC_REJECT or REC_INVALID, as a double), a velocity (REC_V: the start
// velocity of an accepted attempt, the 5th-order velocity of a rejected one,
// whose start velocity is that of the accepted attempt that follows it), the
// stage velocities v_2..v_6 (REC_VS + 4 (s - 2)) and the error-estimate
// difference new_vel - vel_err (REC_DV).  An invalid attempt (a stage
// position left the domain) records only h and its kind.
constexpr int STEP_REC = 32;
constexpr int REC_R = 0, REC_TH = 1, REC_H = 2, REC_KIND = 3, REC_V = 4, REC_VS = 8, REC_DV = 28;
enum { REC_ACCEPT = 0, REC_REJECT = 1, REC_INVALID = 2 };

// Tableau entries for a stage number known at compile time (the unrolled
// loop): the same double quotients as kA, kB, kC, kD, folded into the code.
__device__ constexpr double tabA(int s) {
  return s == 2 ? A2 : s == 3 ? A3 : s == 4 ? A4 : s == 5 ? A5 : s == 6 ? 1.0 : 0.0;
}
__device__ constexpr double tabB(int s, int j) {
  return s == 3 ? (j == 1 ? B31 : B32)
       : s == 4 ? (j == 1 ? B41 : j == 2 ? B42 : B43)
       : s == 5 ? (j == 1 ? B51 : j == 2 ? B52 : j == 3 ? B53 : B54)
       : s == 6 ? (j == 1 ? B61 : j == 2 ? B62 : j == 3 ? B63 : j == 4 ? B64 : B65)
       : 0.0;
}
__device__ constexpr double tabC(int s) {
  return s == 1 ? C1 : s == 3 ? C3 : s == 4 ? C4 : s == 5 ? C5 : s == 6 ? C6 : 0.0;
}
__device__ constexpr double tabD(int s) {
  return s == 1 ? D1 : s == 3 ? D3 : s == 4 ? D4 : s == 5 ? D5 : s == 6 ? D6 : s == 7 ? D7 : 0.0;
}

// The stages of one attempt (geodesic.rs:244-308): k1 when the lane has no
// valid one (s = 1), then k2..k7, the 5th-order velocity and the error.  One
// copy of the loop per arithmetic (ARITH_EXACT, ARITH_F32, ARITH_FAST).  The
// FP32 loop is unrolled (stage number, tableau and branches resolved at
// compile time); the FP64 loops stay rolled, their right-hand sides being
// several times larger.
// REC: write the stage velocities into rec (when non-null) as they form.
template <int METRIC, int ARITH, bool REC = false, class KS = KStore>
__device__ inline StageOut stages(const DevScene &S, const KS &K, const double p0[4], const d4 vel, double h,
                                  bool need_k1, double *rec = nullptr) {
#pragma clang fp contract(off)
  constexpr bool UNROLL = (ARITH != ARITH_EXACT);
  StageOut o;
  d4 k1{0, 0, 0, 0}, s5{0, 0, 0, 0}, s4{0, 0, 0, 0};
  o.nv = d4{0, 0, 0, 0};
  o.k7 = d4{0, 0, 0, 0};
  o.nr = 0;
  o.invalid = false;
  // one stage; false when a stage position left the domain
  auto stage = [&] __device__(const int s) -> bool {
    double qr, qth;
    double qt = 0.0, qph = 0.0;      // METRIC 2: the metric may depend on t and phi too
    d4 vs;
    bool eval = true;
    if (s == 1) {
      qr = p0[1];                                                   // k1 = a(pos, vel) (:245)
      qth = p0[2];
      if constexpr (METRIC == 2) { qt = p0[0]; qph = p0[3]; }
      vs = vel;
      eval = need_k1;
    } else if (s <= 6) {
      // stage position pos + (h a_s) vel on the forward-Euler line (:248-284)
      double ha = h * (UNROLL ? tabA(s) : kA[s]);
      double q0 = madd<ARITH>(ha, vel.x, p0[0]), q1 = madd<ARITH>(ha, vel.y, p0[1]);
      double q2 = madd<ARITH>(ha, vel.z, p0[2]), q3 = madd<ARITH>(ha, vel.w, p0[3]);
      if (!valid4(q0, q1, q2, q3)) { o.invalid = true; return false; }   // :250-291 halve h
      if (s == 2) {
        double hb = h * B21;                                        // vel + (h b21) k1 (:249)
        vs = d4{madd<ARITH>(hb, k1.x, vel.x), madd<ARITH>(hb, k1.y, vel.y), madd<ARITH>(hb, k1.z, vel.z),
                madd<ARITH>(hb, k1.w, vel.w)};
      } else {
        double b = UNROLL ? tabB(s, 1) : kB[s][1];
        d4 sum{b * k1.x, b * k1.y, b * k1.z, b * k1.w};
        auto term = [&] __device__(int j) {
          double bj = UNROLL ? tabB(s, j) : kB[s][j];
          d4 kj = K.get(j - 1);
          sum.x = madd<ARITH>(bj, kj.x, sum.x); sum.y = madd<ARITH>(bj, kj.y, sum.y);
          sum.z = madd<ARITH>(bj, kj.z, sum.z); sum.w = madd<ARITH>(bj, kj.w, sum.w);
        };
        if constexpr (UNROLL) {
#pragma unroll
          for (int j = 2; j < s; ++j) term(j);
        } else {
#pragma unroll 1
          for (int j = 2; j < s; ++j) term(j);
        }
        vs = d4{madd<ARITH>(h, sum.x, vel.x), madd<ARITH>(h, sum.y, vel.y), madd<ARITH>(h, sum.z, vel.z),
                madd<ARITH>(h, sum.w, vel.w)};
      }
      qr = q1;
      qth = q2;
      if constexpr (METRIC == 2) { qt = q0; qph = q3; }
      if (REC && rec) {
        double *o = (double *)__builtin_assume_aligned(rec + REC_VS + 4 * (s - 2), 32);
        o[0] = vs.x; o[1] = vs.y; o[2] = vs.z; o[3] = vs.w;
      }
    } else {
      qr = madd<ARITH>(h, vel.y, p0[1]);                            // k7 at (pos6, new_vel) (:298)
      qth = madd<ARITH>(h, vel.z, p0[2]);
      if constexpr (METRIC == 2) { qt = madd<ARITH>(h, vel.x, p0[0]); qph = madd<ARITH>(h, vel.w, p0[3]); }
      vs = o.nv;
    }
    d4 k;
    if (eval) {
      if constexpr (METRIC == 2) {
        const double q[4] = {qt, qr, qth, qph};
        k = user_rhs(S, q, vs);
      } else {
        k = eval_rhs<METRIC, ARITH>(S, qr, qth, vs);
      }
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
        double c = UNROLL ? tabC(s) : kC[s];
        s5.x = madd<ARITH>(c, k.x, s5.x); s5.y = madd<ARITH>(c, k.y, s5.y);
        s5.z = madd<ARITH>(c, k.z, s5.z); s5.w = madd<ARITH>(c, k.w, s5.w);
      }
      double d = UNROLL ? tabD(s) : kD[s];
      s4.x = madd<ARITH>(d, k.x, s4.x); s4.y = madd<ARITH>(d, k.y, s4.y);
      s4.z = madd<ARITH>(d, k.z, s4.z); s4.w = madd<ARITH>(d, k.w, s4.w);
    }
    if (s == 6)
      o.nv = d4{madd<ARITH>(h, s5.x, vel.x), madd<ARITH>(h, s5.y, vel.y), madd<ARITH>(h, s5.z, vel.z),
                madd<ARITH>(h, s5.w, vel.w)};
    if (s == 7) o.k7 = k;
    return true;
  };
  if constexpr (UNROLL) {
#pragma unroll
    for (int s = 1; s <= 7; ++s)
      if (!stage(s)) break;
  } else {
#pragma unroll 1
    for (int s = 1; s <= 7; ++s)
      if (!stage(s)) break;
  }
  if (o.invalid) return o;
  const d4 nv = o.nv;
  d4 ve{madd<ARITH>(h, s4.x, vel.x), madd<ARITH>(h, s4.y, vel.y), madd<ARITH>(h, s4.z, vel.z),
        madd<ARITH>(h, s4.w, vel.w)};
  double dx = nv.x - ve.x, dy = nv.y - ve.y, dz = nv.z - ve.z, dw = nv.w - ve.w;
  if (REC && rec) {
    double *dst = (double *)__builtin_assume_aligned(rec + REC_DV, 32);
    dst[0] = dx; dst[1] = dy; dst[2] = dz; dst[3] = dw;
  }
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

// The stages of an FP32 attempt: the right-hand sides, stage positions r,
// stage velocities and the sums sum c_i k_i and sum (c_i - d_i) k_i in FP32
// with the k_i in registers; each stage's psi = theta - pi/2 formed in FP64
// before its rounding (near the disk plane a rounded FP32 theta would lose
// psi's relative precision); new_vel = v + h sum c_i k_i, the error
// difference h sum (c_i - d_i) k_i and its norm in FP64.  The FP32 rounding
// of the sums is a fraction of that of the right-hand sides, which
// fp32_noise bounds.  An invalid stage (r <= 0 or not finite) sends the
// attempt to FP64 like an error ratio near a decision.
template <int METRIC, bool REC = false, class KS = KStore>
__device__ inline StageOut stages32(const DevScene &S, const KS &K, const double p0[4], const d4 vel, double h,
                                    bool need_k1, double *rec = nullptr) {
#pragma clang fp contract(fast)
  StageOut o;
  o.nr = 0;
  o.invalid = false;
  o.nv = d4{0, 0, 0, 0};
  const float hf = (float)h;
  const f4 vf{(float)vel.x, (float)vel.y, (float)vel.z, (float)vel.w};
  const float r0 = (float)p0[1];
  const double psi0 = p0[2] - HALF_PI;
  f4 k[7];
  if (need_k1) {
    k[0] = rhs32f<METRIC>(S, r0, (float)psi0, vf);
    ++o.nr;
  } else {
    const d4 kd = K.get(0);
    k[0] = f4{(float)kd.x, (float)kd.y, (float)kd.z, (float)kd.w};
  }
  const float c1 = (float)C1, e1 = (float)(C1 - D1);
  f4 s5{c1 * k[0].x, c1 * k[0].y, c1 * k[0].z, c1 * k[0].w};
  f4 se{e1 * k[0].x, e1 * k[0].y, e1 * k[0].z, e1 * k[0].w};
#pragma unroll
  for (int s = 2; s <= 7; ++s) {
    float rs, psis;
    f4 vs;
    if (s <= 6) {
      const double ha = h * tabA(s);
      rs = r0 + (float)ha * vf.y;
      psis = (float)(psi0 + ha * vel.z);
      if (!(rs > 0.0f && isfinite(rs) && isfinite(psis))) { o.invalid = true; return o; }
      f4 sum{0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
      for (int j = 1; j < s; ++j) {
        const float b = (float)tabB(s, j);
        sum.x += b * k[j - 1].x; sum.y += b * k[j - 1].y; sum.z += b * k[j - 1].z; sum.w += b * k[j - 1].w;
      }
      if (s == 2) sum = f4{(float)B21 * k[0].x, (float)B21 * k[0].y, (float)B21 * k[0].z, (float)B21 * k[0].w};
      vs = f4{vf.x + hf * sum.x, vf.y + hf * sum.y, vf.z + hf * sum.z, vf.w + hf * sum.w};
      if (REC && rec) {
        double *d = (double *)__builtin_assume_aligned(rec + REC_VS + 4 * (s - 2), 32);
        d[0] = vs.x; d[1] = vs.y; d[2] = vs.z; d[3] = vs.w;
      }
    } else {
      rs = r0 + hf * vf.y;                                          // k7 at (pos6, new_vel)
      psis = (float)(psi0 + h * vel.z);
      vs = f4{(float)o.nv.x, (float)o.nv.y, (float)o.nv.z, (float)o.nv.w};
    }
    k[s - 1] = rhs32f<METRIC>(S, rs, psis, vs);
    ++o.nr;
    const f4 &kk = k[s - 1];
    if (s >= 3 && s <= 6) {
      const float c = (float)tabC(s);
      s5.x += c * kk.x; s5.y += c * kk.y; s5.z += c * kk.z; s5.w += c * kk.w;
    }
    if (s >= 3) {
      const float e = (float)(tabC(s) - tabD(s));
      se.x += e * kk.x; se.y += e * kk.y; se.z += e * kk.z; se.w += e * kk.w;
    }
    if (s == 6)
      o.nv = d4{__builtin_fma(h, (double)s5.x, vel.x), __builtin_fma(h, (double)s5.y, vel.y),
                __builtin_fma(h, (double)s5.z, vel.z), __builtin_fma(h, (double)s5.w, vel.w)};
  }
  o.k7 = to64(k[6]);
  const double dx = h * (double)se.x, dy = h * (double)se.y, dz = h * (double)se.z, dw = h * (double)se.w;
  if (REC && rec) {
    double *dst = (double *)__builtin_assume_aligned(rec + REC_DV, 32);
    dst[0] = dx; dst[1] = dy; dst[2] = dz; dst[3] = dw;
  }
  const d4 nv = o.nv;
  if (S.err_mode == 1) {
    const double q0 = dx / (S.atol + S.rtol * fmax(fabs(vel.x), fabs(nv.x)));
    const double q1 = dy / (S.atol + S.rtol * fmax(fabs(vel.y), fabs(nv.y)));
    const double q2 = dz / (S.atol + S.rtol * fmax(fabs(vel.z), fabs(nv.z)));
    const double q3 = dw / (S.atol + S.rtol * fmax(fabs(vel.w), fabs(nv.w)));
    o.err = sqrt(0.25 * (q0 * q0 + q1 * q1 + q2 * q2 + q3 * q3));
    o.thresh = 1.0;
  } else {
    o.err = sqrt(dx * dx + dy * dy + dz * dz + dw * dw);
    o.thresh = S.tol;
  }
  return o;
}

// Whether the next attempt of this ray may run its stages in FP32: the
// controller is saturated (h == max_step, previous error ratio < 0.3), the
// ray is far enough from the photon-orbit region that FP32 rounding is not
// amplified by a later close approach -- outgoing beyond fp32_min_r
// (outgoing photons there never turn back), inbound beyond fp32_far_r, or
// inbound beyond fp32_min_r with an impact parameter b = sqrt((Q+L_z^2)/E^2)
// of at least 10 M, whose periapsis lies beyond ~8 M for any spin -- theta
// is away from the poles and the FP32 noise bound is below 0.02 in
// error-ratio units.  A function of the ray's own state only, so every ray's
// arithmetic is independent of which lane or wavefront carries it.
__device__ inline bool fp32_eligible(const DevScene &S, const Lane &L) {
  const double r = L.pos[1], th = L.pos[2];
  bool far_enough = (L.vel.y > 0.0) ? (r >= S.fp32_min_r)
                                     : (r >= S.fp32_far_r || (r >= S.fp32_min_r && L.b2 >= S.fp32_b2_min));
  if (!(far_enough && (L.h == S.max_step) && (L.err_prev < 0.3) && (th > 1e-3) && (th < PI - 1e-3)))
    return false;
  return fp32_noise(S, L.pos, L.vel, L.h) < 0.02f;
}

// One RK45Integrator::step (geodesic.rs:185-328) on lane state L with the
// stage values in LDS.  PREC 0: exact FP64; 2: fast FP64; 1: mixed, three
// tiers chosen per wavefront (use32, use_fast are wavefront-uniform):
//   FP32 stages when every live lane's ray is FP32-eligible;
//   fast-FP64 stages (reciprocal-form RHS, FMA) otherwise, unless a lane's
//     controller is unsaturated with a tiny previous error;
//   exact-order FP64 stages otherwise.
// A lane redoes its attempt one tier down when the result could steer a
// decision differently from exact FP64: an FP32 error ratio >= 0.4, a fast
// ratio within 1e-9 of the accept threshold, a fast ratio below 1e-6 with
// h < max_step (geodesic.rs:318 keeps h when the error rounds to exactly
// 0, so there the last bit decides the next h), or any invalid stage.
// Every accept/reject decision is then the FP64 kernel's; h agrees with it
// to rounding (bitwise while only FP32 and exact steps are taken), and final
// states reproduce across launches to ~1e-12, not bitwise, since which
// eligible steps take which tier depends on the lane grouping.
// Returns RES_*; `acc` reports acceptance; p0/v0 keep the pre-step state.
// REC (recording render and adjoint kernels): rec, when non-null, receives
// the step record of the attempt, whatever its outcome.
template <int METRIC, int PREC, bool REC = false, class KS = KStore>
__device__ inline int attempt(const DevScene &S, Lane &L, const KS &K, bool &acc, double p0[4], d4 &v0,
                              bool use32, bool use_fast = false, double *rec = nullptr) {
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
    o = stages32<METRIC, REC>(S, K, p0, vel, h, need_k1, rec);
    L.n_rhs += o.nr;
    L.n_rhs32 += o.nr;
    if (need_k1) { L.k1_ok = true; L.k1_f32 = true; }   // slot 0 now holds an FP32 k1
    // an FP32 result near a decision (or with an invalid stage) is redone in FP64
    done = !o.invalid && (o.err / o.thresh < 0.4);
    if (!done) use32 = false;
  }
  bool fast = false;
  if (PREC == 1 && !done && use_fast) {
    // fast FP64 needs an FP64 k1 (either arithmetic); an FP32 one is redone
    const bool need_k1 = !L.k1_ok || L.k1_f32;
    o = stages<METRIC, ARITH_FAST, REC>(S, K, p0, vel, h, need_k1, rec);
    L.n_rhs += o.nr;
    L.n_rhsf += o.nr;
    if (need_k1) { L.k1_ok = true; L.k1_f32 = false; L.k1_fast = true; }
    const double ratio = o.err / o.thresh;
    done = !o.invalid && !(h < S.max_step && ratio < 1e-6) && !(fabs(ratio - 1.0) < 1e-9);
    fast = done;
  }
  if (!done) {
    // An exact attempt needs an exact k1: after an FP32 or fast attempt k7
    // (the FSAL k1) carries that arithmetic's rounding, so recompute it.
    const bool need_k1 = !L.k1_ok || L.k1_f32 || L.k1_fast;
    o = stages<METRIC, BASE, REC>(S, K, p0, vel, h, need_k1, rec);
    L.n_rhs += o.nr;
    if (need_k1) { L.k1_f32 = false; L.k1_fast = false; }
  }
  L.k1_ok = true;
  if (o.invalid) {
    if (REC && rec) { rec[REC_H] = h; rec[REC_KIND] = (double)REC_INVALID; }
    L.h = h * 0.5;
    return RES_CONTINUE;
  }
  const double err = o.err, thresh = o.thresh;
  L.n_full++;
  L.err_prev = err / thresh;
  if (err < thresh || h <= S.min_step) {           // accept (:310-321)
    L.pos[0] = madd<BASE>(h, vel.x, p0[0]); L.pos[1] = madd<BASE>(h, vel.y, p0[1]);
    L.pos[2] = madd<BASE>(h, vel.z, p0[2]); L.pos[3] = madd<BASE>(h, vel.w, p0[3]);
    if (REC && rec) {
      double *r = (double *)__builtin_assume_aligned(rec, 32);
      r[REC_R] = p0[1]; r[REC_TH] = p0[2]; r[REC_H] = h; r[REC_KIND] = (double)REC_ACCEPT;
      r[REC_V] = vel.x; r[REC_V + 1] = vel.y; r[REC_V + 2] = vel.z; r[REC_V + 3] = vel.w;
    }
    L.vel = o.nv;
    L.lambda = L.lambda + h;
    // normalize_spherical_coords (:337-352)
    bool refl = false;
    if (L.pos[2] < 0.0) { L.pos[2] = -L.pos[2]; L.pos[3] = L.pos[3] + PI; refl = true; }
    if (L.pos[2] > PI) { L.pos[2] = 2.0 * PI - L.pos[2]; L.pos[3] = L.pos[3] + PI; refl = true; }
    L.n_refl += refl ? 1u : 0u;
    // rem_euclid(2 pi); fmod is the identity on [0, 2 pi), where phi almost always is
    const double p3 = L.pos[3];
    if (!(p3 >= 0.0 && p3 < TWO_PI)) {
      double ph = fmod(p3, TWO_PI);
      L.pos[3] = ph < 0.0 ? ph + TWO_PI : ph;
      if constexpr (METRIC == 2) refl = true;      // k7 was taken at the unwrapped phi
    }
    K.put(0, o.k7);                                 // FSAL: k7 is the next k1 ...
    L.k1_ok = !refl;                                // ... unless theta was reflected
    L.k1_f32 = use32;
    L.k1_fast = fast;
    if (err > 0.0) {
      // h 0.9 t^0.2 clamped to [min_step, max_step].  When t (0.9 h)^5
      // exceeds max_step^5 by more than rounding can move it, the clamp
      // returns max_step whatever pow rounds to: skip pow (215 instructions).
      // Not in the FP64 kernel: there it measured 10% slower on MI355X with
      // the same instruction count, the kernel running at a lower clock.
      const double t = thresh / err;
      const double u = h * 0.9, u2 = u * u;
      if (PREC != 0 && t * (u2 * u2 * u) > S.max_step5 * (1.0 + 1e-12)) {
        L.h = S.max_step;
      } else {
        double hn = u * pow(t, 0.2);
        if (hn < S.min_step) hn = S.min_step;
        if (hn > S.max_step) hn = S.max_step;
        L.h = hn;
      }
    }
    acc = true;
  } else {                                          // reject (:322-325)
    if (REC && rec) {
      double *r = (double *)__builtin_assume_aligned(rec, 32);
      r[REC_R] = p0[1]; r[REC_TH] = p0[2]; r[REC_H] = h; r[REC_KIND] = (double)REC_REJECT;
      r[REC_V] = o.nv.x; r[REC_V + 1] = o.nv.y; r[REC_V + 2] = o.nv.z; r[REC_V + 3] = o.nv.w;
    }
    double hn = h * 0.9 * pow(thresh / err, 0.25);
    L.h = hn >= S.min_step ? hn : S.min_step;       // f64::max
  }
  return RES_CONTINUE;
}

// Sum over the 64 lanes of a wavefront (every lane gets the total).
__device__ inline unsigned long long wave_sum(unsigned long long x) {
  for (int off = 32; off > 0; off >>= 1) x += __shfl_xor(x, off, 64);
  return x;
}

}  // namespace

}  // namespace gg
