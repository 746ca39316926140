// gg_grad.hip -- hand-written reverse-mode adjoint of the thin-disk render.
//
// d(loss)/d(spin, inclination, q, r_in) for L = mean_p (I_p - target_p)^2.
// By default the step sizes carry derivatives: each attempt's h is the
// controller's function of the previous h and error estimate
// (geodesic.rs:308-325), so the gradient is that of the loss the render
// computes; the discrete decisions (accept/reject, clamps, err == 0, the
// disk-crossing step) are those of the forward render.  With frozen_steps
// the step sizes are held at the render's values instead.
//
// The disk-hitting rays, sorted by attempt count (longest first, so the
// long photon-ring rays start early and neighbouring lanes stay in the same
// phase), go through these kernels, one lane per ray:
//   adjoint_fwd_kernel  re-traces the ray from the camera with the trace
//       kernel's own attempt() (bit-identical steps), storing a checkpoint
//       (pos, vel, h) every K accepted steps in a pool sized exactly from
//       the render's accepted counts, and the step record of every attempt
//       of the ray's last segment; at the render's accepted count the step
//       is the disk crossing, whose intensity is differentiated in forward
//       mode (12 inputs): the seeds of the reverse sweep.
//   adjoint_rev_kernel  back-propagates one attempt per iteration through
//       its controller, error estimate and Dormand-Prince stages (the RHS
//       vector-Jacobian products at the recorded stage points; stage
//       cotangents in LDS), segment by segment from the last; before each
//       earlier segment it re-runs that segment from its checkpoint into a
//       per-lane segment buffer (so h repeats exactly, and the end state
//       must equal the next checkpoint bit for bit).  When K is at least
//       every ray's accepted count there is one segment per ray, the
//       re-trace stores every step record, and adjoint_rev_stored_kernel
//       (the same sweep without recomputation) reads them.
//       adjoint_rev_chunked_kernel reads the records the recording render
//       wrote instead (adjoint_seed_kernel computes the crossing seeds from
//       them).
//   adjoint_cam_kernel  adds the camera ray's dependence on (a, i) and
//       writes dI/d(a, i, q, r_in) per pixel.
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include "gg_adjoint.hpp"
#include "gg_internal.hpp"
#include "gg_rk45.hpp"

namespace gg {

namespace {

enum { PH_IDLE = 0, PH_FWD = 1, PH_REV = 2, PH_RECOMP = 3 };

static_assert(STEP_REC == GRAD_STEP_REC, "step record layout");

// Reverse-sweep state of one lane between two step records.
struct Adj {
  double pr, pth;   // cotangent of the current r, theta (t and phi carry none: nothing depends on them)
  d4 vb;            // ... of the current velocity
  double hb;        // ... of the h the next attempt takes (the controller's output)
  double ga;        // spin cotangent
  d4 gv;            // start velocity of the accepted attempt after the current record
  double hn;        // h of the next attempt
};

// Block sizes of the re-trace and reverse kernels, fixed at compile time (the
// LDS slot offsets are instruction immediates); the launches use these.
constexpr uint32_t FWD_BLOCK = 256, REV_BLOCK = 64;

// One step record in registers: the reverse sweep loads the record of the
// attempt before the current one while it back-propagates the current one,
// so that the load's latency overlaps the arithmetic.
struct StepRec {
  double d[STEP_REC];
};

__device__ inline void load_rec(StepRec &R, const double *src) {
  const double *p = (const double *)__builtin_assume_aligned(src, 32);
#pragma unroll
  for (int k = 0; k < STEP_REC; ++k) R.d[k] = p[k];
}

// LDS slots 5..8 of the reverse kernels: the current record's stage
// velocities v_2..v_5, read by the rolled stage loop.
constexpr int SLOT_VS = 5;

// Adjoint of one recorded attempt (geodesic.rs:185-328), whose record holds
// its start (r, theta), h, kind, a velocity, the stage velocities v_2..v_6
// and the error difference new_vel - vel_err.  On entry A holds the
// cotangents of the attempt's outputs (state and next h); on exit those of
// its inputs.  LDS slots 0..4 hold the stage cotangents kbar_1..kbar_5,
// slots 5..8 the stage velocities v_2..v_5.
//
// Accepted: pos' = p + h v (normalised), v' = v + h sum_i c_i k_i,
//   k_i = a(p + h a_i v, v_i).  Rejected: p, v pass through.  Invalid: h/2.
// live: the next h is the controller's function of h and the error
//   estimate, h' = clamp(0.9 h err^-0.2) after an accept (h itself when err
//   is 0, geodesic.rs:318) and max(0.9 h err^-0.25, min_step) after a
//   reject; where a clamp binds h' carries no derivative.  The error's
//   cotangent goes through new_vel - vel_err, the rtol/atol scale or the
//   Euclidean norm, and k7 = a(p + h v, new_vel).  h enters the stage
//   positions (a_i v each) and the stage combinations, all of the form
//   v + h sum_j w_j k_j, whose h-derivative adds kbar_j . k_j / h; since a is
//   quadratic in its velocity, kbar . a(x, u) = (J_u^T kbar) . u / 2, the
//   VJP's own velocity output dotted with u.
// !live: the step sizes are held fixed; rejected and invalid attempts then
//   move nothing.
// Shared work: k7 and stage 6 (a_6 = 1) sit at the same point p + h v, so
// one metric algebra (kerr_point) serves both products.
template <class KS>
__device__ inline void attempt_adjoint(const KS &K, const DevScene &S, const StepRec &R, bool live, Adj &A) {
  GG_CONTRACT_FAST
  const double *rec = R.d;
  const double h = rec[REC_H];
  const int kind = (int)rec[REC_KIND];
  const double hn = A.hn;
  A.hn = h;
  if (kind == REC_INVALID) { A.hb *= 0.5; return; }   // geodesic.rs:250-291: h halved
  const bool acc = (kind == REC_ACCEPT);
  if (!acc && !live) return;
  const double m = S.mass, a = S.spin;
  const double r0 = rec[REC_R], th0 = rec[REC_TH];
  const d4 rv{rec[REC_V], rec[REC_V + 1], rec[REC_V + 2], rec[REC_V + 3]};
  const d4 v = acc ? rv : A.gv;
  const d4 nv = acc ? A.gv : rv;
  double s0, c0;                         // the stage angles are theta0 + h a_i v^theta
  sincos(th0, &s0, &c0);
  // derivative coefficients take reciprocals to ~1 ulp (v_rcp_f64 + Newton)
  const double ih = rcp64(h);

  // ---- controller (geodesic.rs:318-325): cotangents of h and of the error ----
  double hbar = 0.0, eb = 0.0;
  const d4 dv{rec[REC_DV], rec[REC_DV + 1], rec[REC_DV + 2], rec[REC_DV + 3]};
  double err = 0.0, ierr = 0.0;
  double isc[4] = {0.0, 0.0, 0.0, 0.0};   // 1 / (atol + rtol max(|v_c|, |new_vel_c|))
  if (live && A.hb != 0.0) {
    if (S.err_mode == 1) {
      isc[0] = rcp64(S.atol + S.rtol * fmax(fabs(v.x), fabs(nv.x)));
      isc[1] = rcp64(S.atol + S.rtol * fmax(fabs(v.y), fabs(nv.y)));
      isc[2] = rcp64(S.atol + S.rtol * fmax(fabs(v.z), fabs(nv.z)));
      isc[3] = rcp64(S.atol + S.rtol * fmax(fabs(v.w), fabs(nv.w)));
      const double q0 = dv.x * isc[0], q1 = dv.y * isc[1], q2 = dv.z * isc[2], q3 = dv.w * isc[3];
      err = sqrt(0.25 * (q0 * q0 + q1 * q1 + q2 * q2 + q3 * q3));
    } else {
      err = sqrt(dv.x * dv.x + dv.y * dv.y + dv.z * dv.z + dv.w * dv.w);
    }
    // err is 0 exactly when every component of new_vel - vel_err is 0
    if (err != 0.0) ierr = rcp64(err);
    if (acc) {
      if (err == 0.0) {
        hbar = A.hb;                                   // h unchanged
      } else if (hn != S.max_step && hn != S.min_step) {
        hbar = A.hb * hn * ih;                         // h' = 0.9 h err^-0.2
        eb = -0.2 * A.hb * hn * ierr;
      }
    } else if (hn != S.min_step) {
      hbar = A.hb * hn * ih;                           // h' = 0.9 h err^-0.25
      eb = -0.25 * A.hb * hn * ierr;
    }
  }
  if (!acc && eb == 0.0) { A.hb = hbar; return; }

  // the end point p + h v (before normalisation): k7's point and stage 6's
  const double dth7 = h * v.z;
  double s7, c7;
  if (!sincos_near(s0, c0, dth7, s7, c7)) sincos(th0 + dth7, &s7, &c7);
  const KerrPoint P7 = kerr_point(m, a, h * v.y + r0, s7, c7);

  double pr = A.pr, pth = A.pth;
  d4 vbar = acc ? d4{0.0, 0.0, 0.0, 0.0} : A.vb;       // a rejection passes p, v through
  d4 nvb{0.0, 0.0, 0.0, 0.0}, veb{0.0, 0.0, 0.0, 0.0}; // cotangents of new_vel, vel_err
  if (acc) {
    // normalize_spherical_coords (geodesic.rs:337-352): theta -> -theta or
    // 2 pi - theta flips the sign of its cotangent
    double th1 = h * v.z + th0;
    double sg = 1.0;
    if (th1 < 0.0) { th1 = -th1; sg = -sg; }
    if (th1 > PI) sg = -sg;
    pth *= sg;
    // pos' = p + h v (geodesic.rs:312); v' = new_vel
    vbar = d4{0.0, h * pr, h * pth, 0.0};
    hbar += pr * v.y + pth * v.z;
    nvb = A.vb;
  }
  if (eb != 0.0) {
    // err = sqrt(mean (d_c / s_c)^2), s_c = atol + rtol max(|v_c|, |new_vel_c|);
    // or the Euclidean norm of d = new_vel - vel_err (geodesic.rs:308)
    d4 db;
    if (S.err_mode == 1) {
      const double dd[4] = {dv.x, dv.y, dv.z, dv.w};
      const double vv[4] = {v.x, v.y, v.z, v.w}, nn[4] = {nv.x, nv.y, nv.z, nv.w};
      double o[4], vs_[4] = {0.0, 0.0, 0.0, 0.0}, ns_[4] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
      for (int c = 0; c < 4; ++c) {
        const bool use_v = fabs(vv[c]) >= fabs(nn[c]);
        const double q = dd[c] * isc[c];
        const double qb = 0.25 * eb * q * ierr;
        o[c] = qb * isc[c];
        const double scb = -qb * q * isc[c] * S.rtol;  // through max(|v_c|, |new_vel_c|)
        if (use_v) vs_[c] = vv[c] > 0.0 ? scb : (vv[c] < 0.0 ? -scb : 0.0);
        else ns_[c] = nn[c] > 0.0 ? scb : (nn[c] < 0.0 ? -scb : 0.0);
      }
      db = d4{o[0], o[1], o[2], o[3]};
      vbar.x += vs_[0]; vbar.y += vs_[1]; vbar.z += vs_[2]; vbar.w += vs_[3];
      nvb.x += ns_[0]; nvb.y += ns_[1]; nvb.z += ns_[2]; nvb.w += ns_[3];
    } else {
      const double k = eb * ierr;
      db = d4{k * dv.x, k * dv.y, k * dv.z, k * dv.w};
    }
    nvb.x += db.x; nvb.y += db.y; nvb.z += db.z; nvb.w += db.w;
    veb = d4{-db.x, -db.y, -db.z, -db.w};
    // k7 = a(p + h v, new_vel) (geodesic.rs:298), weight h d7 in vel_err
    const double hd7 = h * D7;
    const RhsBar jb = kerr_rhs_vjp_pt(P7, nv, d4{hd7 * veb.x, hd7 * veb.y, hd7 * veb.z, hd7 * veb.w});
    A.ga += jb.a;
    pr += jb.r;
    pth += jb.th;
    vbar.y += h * jb.r;
    vbar.z += h * jb.th;
    hbar += jb.r * v.y + jb.th * v.z + 0.5 * ih * (jb.v.x * nv.x + jb.v.y * nv.y + jb.v.z * nv.z + jb.v.w * nv.w);
    nvb.x += jb.v.x; nvb.y += jb.v.y; nvb.z += jb.v.z; nvb.w += jb.v.w;
  }
#pragma unroll
  for (int i = 2; i <= 5; ++i) {
    const double *o = rec + REC_VS + 4 * (i - 2);
    K.put(SLOT_VS + i - 2, d4{o[0], o[1], o[2], o[3]});
  }
  // new_vel = v + h sum_i c_i k_i, vel_err = v + h sum_i d_i k_i
  vbar.x += nvb.x + veb.x; vbar.y += nvb.y + veb.y; vbar.z += nvb.z + veb.z; vbar.w += nvb.w + veb.w;
  d4 w6;
#pragma unroll
  for (int i = 1; i <= 6; ++i) {
    const double hc = h * tabC(i), hd = h * tabD(i);
    const d4 wi{hc * nvb.x + hd * veb.x, hc * nvb.y + hd * veb.y, hc * nvb.z + hd * veb.z, hc * nvb.w + hd * veb.w};
    if (i == 6) w6 = wi;
    else K.put(i - 1, wi);
  }
  // one stage's product: the velocity, position and h cotangents, and
  // h b_ij vbar_i into the cotangents of the earlier stages j < i (k_j); the
  // next stage's (j = i - 1) is returned in registers, the others go back to LDS
  auto stage = [&](int i, const KerrPoint &P, double ha, const d4 vs, const d4 w) {
    const RhsBar jb = kerr_rhs_vjp_pt(P, vs, w);
    A.ga += jb.a;
    pr += jb.r;
    pth += jb.th;
    vbar.x += jb.v.x; vbar.y += jb.v.y; vbar.z += jb.v.z; vbar.w += jb.v.w;
    // the combination's h-derivative (Euler: w . a(x, u) = (J_u^T w) . u / 2)
    if (live) hbar += 0.5 * ih * (jb.v.x * vs.x + jb.v.y * vs.y + jb.v.z * vs.z + jb.v.w * vs.w);
    d4 wn{0.0, 0.0, 0.0, 0.0};
    if (i >= 2) {
      vbar.y += ha * jb.r;
      vbar.z += ha * jb.th;
      if (live) hbar += kA[i] * (jb.r * v.y + jb.th * v.z);
      d4 kb[5];
#pragma unroll
      for (int j = 1; j <= 5; ++j)
        if (j < i) kb[j - 1] = K.get(j - 1);
#pragma unroll
      for (int j = 1; j <= 5; ++j) {
        if (j < i) {
          const double hb = h * kB[i][j];
          kb[j - 1].x += hb * jb.v.x; kb[j - 1].y += hb * jb.v.y; kb[j - 1].z += hb * jb.v.z; kb[j - 1].w += hb * jb.v.w;
          if (j == i - 1) wn = kb[j - 1];
          else K.put(j - 1, kb[j - 1]);
        }
      }
    }
    return wn;
  };
  // stage 6 at the end point (a_6 = 1), its velocity from the record
  d4 w;
  {
    const double *o = rec + REC_VS + 4 * 4;
    w = stage(6, P7, h, d4{o[0], o[1], o[2], o[3]}, w6);
  }
#pragma unroll 1
  for (int i = 5; i >= 2; --i) {
    // stage position p + h a_i v, stage velocity v + h sum_j b_ij k_j
    const double ha = h * kA[i];
    double s, c;
    if (!sincos_near(s0, c0, ha * v.z, s, c)) sincos(th0 + ha * v.z, &s, &c);
    w = stage(i, kerr_point(m, a, ha * v.y + r0, s, c), ha, K.get(SLOT_VS + i - 2), w);
  }
  stage(1, kerr_point(m, a, r0, s0, c0), 0.0, v, w);
  A.pr = pr;
  A.pth = pth;
  A.vb = vbar;
  A.hb = live ? hbar : 0.0;
  if (acc) A.gv = v;
}

__device__ inline void put_state(double *dst, const double p[4], d4 v, double h) {
  dst[0] = p[0]; dst[1] = p[1]; dst[2] = p[2]; dst[3] = p[3];
  dst[4] = v.x; dst[5] = v.y; dst[6] = v.z; dst[7] = v.w;
  dst[8] = h;
}

__device__ inline void get_state(const double *src, double p[4], d4 &v, double &h) {
  p[0] = src[0]; p[1] = src[1]; p[2] = src[2]; p[3] = src[3];
  v = d4{src[4], src[5], src[6], src[7]};
  h = src[8];
}

// Per-ray record (REC doubles at list position q): the reverse sweep's seeds
// and bookkeeping, then its results.
enum {
  R_PBR = 0, R_PBTH, R_VB0, R_VB1, R_VB2, R_VB3, R_GA, R_GI, R_GQ, R_GRIN,   // cotangents
  R_XR0, R_XTH0, R_XV0T, R_XV0P,                                            // crossing's direct terms
  R_CHK, R_TAIL, R_NTAIL,                                                   // pool offsets, tail records (bits)
  R_GV0, R_GV1, R_GV2, R_GV3,                                               // velocity after the crossing step
  R_PAD0, R_PAD1, R_PAD2,
  REC
};
static_assert(REC == GRAD_REC, "record layout");

__device__ inline uint64_t as_u64(double x) { return (uint64_t)__double_as_longlong(x); }
__device__ inline double as_f64(uint64_t x) { return __longlong_as_double((long long)x); }

// Wavefront-aggregated queue: idle lanes of the wavefront take consecutive
// list positions with one atomicAdd.  Returns the position or ~0 (none).
__device__ inline uint64_t take(unsigned long long *queue, uint64_t n, bool idle, bool &exhausted) {
  const uint32_t lane = threadIdx.x & 63;
  const unsigned long long m = __ballot(idle);
  uint64_t q = ~0ull;
  if (!exhausted && m) {
    const uint32_t cnt = __popcll(m);
    const int leader = __ffsll((long long)m) - 1;
    unsigned long long base = 0;
    if ((int)lane == leader) base = atomicAdd(queue, (unsigned long long)cnt);
    base = __shfl(base, leader, 64);
    if (base + cnt >= n) exhausted = true;
    const uint64_t mine = base + __popcll(m & ((1ull << lane) - 1ull));
    if (idle && mine < n) q = mine;
  }
  return q;
}

// A ray whose re-trace left the rendered path: NaN gradient (the host
// reports the call as failed) and a NaN seed so later kernels skip it.
__device__ inline void mark_bad(const GradLaunch &G, uint64_t q, uint32_t pix, unsigned long long &n_bad) {
  double *o = G.dI + 4 * (uint64_t)pix;
  const double nan = __longlong_as_double(0x7ff8000000000000ll);
  o[0] = nan; o[1] = nan; o[2] = nan; o[3] = nan;
  G.rec[REC * q + R_PBR] = nan;
  ++n_bad;
}

// The disk crossing of a ray's last accepted step (p0, v0) -> (p1, v1),
// the trace kernel's test, and the forward-mode derivatives of its intensity
// over its 12 inputs: the reverse sweep's seeds in the ray record.  False
// when the step does not cross the disk (the ray is then not the rendered one).
__device__ inline bool crossing_seeds(const GradLaunch &G, const DevScene &S, double *rec, double p0r, double p0th,
                                      double v0t, double v0p, double p1r, double p1th, double v1t, double v1p) {
#pragma clang fp contract(off)
  const bool side0 = p0th < HALF_PI, side1 = p1th < HALF_PI;
  double rc = 0.0;
  if (side0 != side1) {
    double s = (HALF_PI - p0th) / (p1th - p0th);
    rc = p0r + s * (p1r - p0r);
  }
  const double w = S.edge_width;
  const bool hit = (side0 != side1) &&
                   ((w > 0.0) ? (rc <= S.r_out + 8.0 * w && rc > 0.0) : (rc >= S.r_in && rc <= S.r_out));
  if (!hit) return false;
  const CrossVals x{S.spin, S.q, S.r_in, S.sqrt_neg_gtt_obs, p0r, p0th, p1r, p1th, v0t, v0p, v1t, v1p};
  double d[CX_N];
#pragma unroll 1
  for (int k = 0; k < CX_N; ++k) d[k] = crossing_intensity_d(S.mass, S.r_out, w, x, k).d;
  const Tn so_a = kerr_sqrt_obs_d(S.mass, Tn{S.spin, 1.0}, G.r_cam, dconst(G.incl));
  const Tn so_i = kerr_sqrt_obs_d(S.mass, dconst(S.spin), G.r_cam, Tn{G.incl, 1.0});
  rec[R_PBR] = d[CX_R1];
  rec[R_PBTH] = d[CX_TH1];
  rec[R_VB0] = d[CX_V1T]; rec[R_VB1] = 0.0; rec[R_VB2] = 0.0; rec[R_VB3] = d[CX_V1P];
  rec[R_GA] = d[CX_A] + d[CX_SOBS] * so_a.d;
  rec[R_GI] = d[CX_SOBS] * so_i.d;
  rec[R_GQ] = d[CX_Q];
  rec[R_GRIN] = d[CX_RIN];
  rec[R_XR0] = d[CX_R0]; rec[R_XTH0] = d[CX_TH0]; rec[R_XV0T] = d[CX_V0T]; rec[R_XV0P] = d[CX_V0P];
  return true;
}

__device__ inline void put_gv(double *rec, d4 v) {
  rec[R_GV0] = v.x; rec[R_GV1] = v.y; rec[R_GV2] = v.z; rec[R_GV3] = v.w;
}

__global__ __launch_bounds__(FWD_BLOCK, 2) void adjoint_fwd_kernel(GradLaunch G) {
#pragma clang fp contract(off)
  extern __shared__ __align__(16) double lds_k[];
  const KStoreT<FWD_BLOCK> K{reinterpret_cast<double2 *>(lds_k) + threadIdx.x, FWD_BLOCK};
  const uint32_t lane = threadIdx.x & 63;
  const DevScene &S = G.scene;
  const uint32_t KK = G.K;
  Lane L;
  L.n_rhs = 0; L.n_rhs32 = 0; L.n_rhsf = 0; L.n_full = 0; L.n_refl = 0;
  bool active = false, exhausted = false;
  uint64_t q = 0, chk_base = 0, tail_base = 0;
  uint32_t pix = 0, acc_target = 0, att_target = 0, last0 = 0, att = 0, tail_cap = 0, seg_att = 0;
  unsigned long long n_att = 0, n_bad = 0;
  for (;;) {
    const uint64_t t = take(&G.cnt->queue, G.n_list, !active, exhausted);
    if (t != ~0ull) {
      q = t;
      pix = G.list[q];
      att_target = G.keys[q];
      acc_target = G.accepted[pix];
      const uint32_t nchk = (acc_target + KK - 1) / KK;
      last0 = nchk ? (nchk - 1) * KK : 0u;
      // the last segment's attempts: its accepted steps plus at most the
      // ray's rejected and invalid attempts
      tail_cap = att_target - last0;
      chk_base = atomicAdd(&G.cnt->chk_cursor, (unsigned long long)nchk);
      tail_base = atomicAdd(&G.cnt->tail_cursor, (unsigned long long)tail_cap);
      camera_ray(S, pix % G.width, pix / G.width, L.pos, L.vel);
      L.lambda = 0.0;
      L.h = S.h0;
      L.k1_ok = false; L.k1_f32 = false; L.k1_fast = false;
      L.accepted = 0;
      L.err_prev = 1.0;
      L.b2 = 0.0;
      att = 0;
      seg_att = 0;
      active = true;
      if (acc_target == 0 || att_target < acc_target || chk_base + nchk > G.chk_cap ||
          tail_base + tail_cap > G.tail_cap) {
        mark_bad(G, q, pix, n_bad);
        active = false;
      } else {
        put_state(G.chk + 9 * chk_base, L.pos, L.vel, L.h);
      }
    }
    if (exhausted && __all(!active)) break;
    if (!active) continue;

    bool acc;
    double p0[4];
    d4 v0;
    const bool in_tail = L.accepted >= last0;
    if (in_tail && seg_att >= tail_cap) {   // more attempts than the render took
      mark_bad(G, q, pix, n_bad);
      active = false;
      continue;
    }
    double *recp = in_tail ? G.tail + STEP_REC * (tail_base + seg_att) : nullptr;
    const int res = attempt<1, 0, true>(S, L, K, acc, p0, v0, false, false, recp);
    ++n_att;
    ++seg_att;
    if (res != RES_CONTINUE || ++att > S.max_steps) {   // the render reached the crossing; a re-trace must too
      mark_bad(G, q, pix, n_bad);
      active = false;
      continue;
    }
    if (!acc) continue;
    L.accepted++;
    if (L.accepted < acc_target) {
      if (L.accepted % KK == 0) {
        put_state(G.chk + 9 * (chk_base + L.accepted / KK), L.pos, L.vel, L.h);
        seg_att = 0;
      }
      continue;
    }
    // the disk crossing of the last accepted step (same test as the trace kernel)
    active = false;
    double *rec = G.rec + REC * q;
    if (att != att_target ||
        !crossing_seeds(G, S, rec, p0[1], p0[2], v0.x, v0.w, L.pos[1], L.pos[2], L.vel.x, L.vel.w)) {
      mark_bad(G, q, pix, n_bad);
      continue;
    }
    put_gv(rec, L.vel);
    rec[R_CHK] = as_f64(chk_base);
    rec[R_TAIL] = as_f64(tail_base);
    rec[R_NTAIL] = as_f64(seg_att);
  }
  unsigned long long a = wave_sum(n_att), bd = wave_sum(n_bad), r = wave_sum(L.n_rhs);
  if (lane == 0) {
    atomicAdd(&G.cnt->attempts, a);
    atomicAdd(&G.cnt->bad, bd);
    atomicAdd(&G.cnt->rhs, r);
  }
}

// The reverse sweep, one step record (attempt) per iteration, last first.
// MODE 0: some rays have more than one segment, so the earlier segments are
// recomputed from their checkpoints.  MODE 1: every step record of every ray
// is in the last-segment pool (one segment per ray).  MODE 2: the records are
// the recording render's chunks, walked back through their back-pointers.
// Every mode runs one wavefront per SIMD (64-thread blocks, 4 per CU), the
// record of the attempt before the current one loading into AGPRs.
template <int MODE>
__device__ __forceinline__ void adjoint_rev_body(const GradLaunch &G) {
#pragma clang fp contract(off)
  constexpr bool RECOMPUTE = (MODE == 0);
  extern __shared__ __align__(16) double lds_k[];
  const KStoreT<REV_BLOCK> K{reinterpret_cast<double2 *>(lds_k) + threadIdx.x, REV_BLOCK};
  const uint32_t lane = threadIdx.x & 63;
  const uint64_t gtid = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const DevScene &S = G.scene;
  const uint32_t KK = G.K;
  const bool live = G.live != 0;
  double *const seg = RECOMPUTE ? G.seg + gtid * (uint64_t)G.seg_cap * STEP_REC : nullptr;

  Lane L;
  L.n_rhs = 0; L.n_rhs32 = 0; L.n_rhsf = 0; L.n_full = 0; L.n_refl = 0;
  enum { IDLE = 0, REV = 1, RECOMP = 2 };
  int phase = IDLE;
  bool exhausted = false, first = false;
  uint64_t q = 0, chk_base = 0;
  uint32_t pix = 0, c = 0, j = 0, att = 0, chunk = 0, seg_j = 0;
  uint32_t chunk_prev = 0;   // MODE 2: the chunk before `chunk` (its back-pointer, read on entering it)
  const double *src = nullptr;
  StepRec cur;               // the record of attempt j, once cur_ok
  bool cur_ok = false;
  Adj A;
  A.pr = 0; A.pth = 0; A.vb = d4{0, 0, 0, 0}; A.hb = 0; A.ga = 0; A.gv = d4{0, 0, 0, 0}; A.hn = 0;
  unsigned long long n_att = 0, n_adj = 0, n_done = 0, n_bad = 0;

  for (;;) {
    const uint64_t t = take(&G.cnt->queue2, G.n_list, phase == IDLE, exhausted);
    if (t != ~0ull) {
      q = t;
      pix = G.list[q];
      const double *rec = G.rec + REC * q;
      if (!(rec[R_PBR] == rec[R_PBR])) {   // flagged by the forward pass
        phase = IDLE;
      } else {
        // without recomputation the whole ray is one segment
        const uint32_t acc_target = G.accepted[pix];
        const uint32_t nchk = RECOMPUTE ? (acc_target + KK - 1) / KK : 1;
        c = nchk - 1;
        if (MODE == 2) {
          chunk = (uint32_t)as_u64(rec[R_CHK]);
          chunk_prev = (uint32_t)as_u64(G.pool[(uint64_t)chunk * GRAD_CHUNK_DOUBLES]);
          j = G.keys[q] - 1;
        } else {
          chk_base = as_u64(rec[R_CHK]);
          src = G.tail + STEP_REC * as_u64(rec[R_TAIL]);
          j = (uint32_t)as_u64(rec[R_NTAIL]) - 1;
        }
        A.pr = rec[R_PBR]; A.pth = rec[R_PBTH];
        A.vb = d4{rec[R_VB0], rec[R_VB1], rec[R_VB2], rec[R_VB3]};
        A.ga = rec[R_GA];
        A.gv = d4{rec[R_GV0], rec[R_GV1], rec[R_GV2], rec[R_GV3]};
        A.hb = 0.0;       // the crossing ends the ray: its next h enters nothing
        A.hn = 0.0;
        first = true;
        cur_ok = false;
        phase = REV;
      }
    }
    if (exhausted && __all(phase == IDLE)) break;

    // ---- RECOMP: one attempt of the segment before ----
    if (RECOMPUTE && phase == RECOMP) {
      if (seg_j >= G.seg_cap) {
        mark_bad(G, q, pix, n_bad);
        phase = IDLE;
      } else {
        bool acc;
        double p0[4];
        d4 v0;
        double *recp = seg + STEP_REC * (uint64_t)seg_j;
        const int res = attempt<1, 0, true>(S, L, K, acc, p0, v0, false, false, recp);
        ++n_att;
        ++seg_j;
        if (res != RES_CONTINUE || ++att > S.max_steps) {
          mark_bad(G, q, pix, n_bad);
          phase = IDLE;
        } else if (acc) {
          L.accepted++;
          if (L.accepted == (c + 1) * KK) {
            // the recomputed segment must land bit for bit on the next checkpoint
            double qp[4], hq;
            d4 vq;
            get_state(G.chk + 9 * (chk_base + c + 1), qp, vq, hq);
            if (qp[0] != L.pos[0] || qp[1] != L.pos[1] || qp[2] != L.pos[2] || qp[3] != L.pos[3] ||
                vq.x != L.vel.x || vq.y != L.vel.y || vq.z != L.vel.z || vq.w != L.vel.w || hq != L.h) {
              mark_bad(G, q, pix, n_bad);
              phase = IDLE;
            } else {
              j = seg_j - 1;
              src = seg;
              cur_ok = false;
              phase = REV;
            }
          }
        }
      }
    }

    // ---- REV: one step record backwards ----
    if (phase == REV) {
      // the record of attempt j (loaded one attempt ahead, except the first
      // of a ray or segment) and, while it is back-propagated, that of j - 1
      auto rec_at = [&](uint32_t ch, uint32_t jj) -> const double * {
        return MODE == 2 ? G.pool + (uint64_t)ch * GRAD_CHUNK_DOUBLES + GRAD_CHUNK_HEAD + STEP_REC * (jj % GRAD_CHUNK)
                         : src + STEP_REC * (uint64_t)jj;
      };
      if (!cur_ok) load_rec(cur, rec_at(chunk, j));
      // (issued on every path, the last record loading itself again, so that
      // the wait before the current record's first use does not also wait
      // for this load)
      StepRec nxt;
      const uint32_t nchunk = (MODE == 2 && j > 0 && j % GRAD_CHUNK == 0) ? chunk_prev : chunk;
      load_rec(nxt, rec_at(nchunk, j > 0 ? j - 1 : j));
      attempt_adjoint(K, S, cur, live, A);
      ++n_adj;
      if (first) {
        const double *rec = G.rec + REC * q;
        A.pr += rec[R_XR0]; A.pth += rec[R_XTH0]; A.vb.x += rec[R_XV0T]; A.vb.w += rec[R_XV0P];
        first = false;
      }
      if (j > 0) {
        if (MODE == 2 && nchunk != chunk) {
          chunk = nchunk;
          chunk_prev = (uint32_t)as_u64(G.pool[(uint64_t)chunk * GRAD_CHUNK_DOUBLES]);
        }
        --j;
        cur = nxt;
        cur_ok = true;
      } else if (RECOMPUTE && c > 0) {
        --c;
        get_state(G.chk + 9 * (chk_base + c), L.pos, L.vel, L.h);
        L.k1_ok = false; L.k1_f32 = false; L.k1_fast = false;
        L.accepted = c * KK;
        att = 0;
        seg_j = 0;
        phase = RECOMP;
      } else {
        double *rec = G.rec + REC * q;
        rec[R_PBR] = A.pr; rec[R_PBTH] = A.pth;
        rec[R_VB0] = A.vb.x; rec[R_VB1] = A.vb.y; rec[R_VB2] = A.vb.z; rec[R_VB3] = A.vb.w;
        rec[R_GA] = A.ga;
        ++n_done;
        phase = IDLE;
      }
    }
  }
  unsigned long long a = wave_sum(n_att), s = wave_sum(n_adj), dn = wave_sum(n_done), bd = wave_sum(n_bad),
                     r = wave_sum(L.n_rhs);
  if (lane == 0) {
    atomicAdd(&G.cnt->attempts, a);
    atomicAdd(&G.cnt->steps, s);
    atomicAdd(&G.cnt->done, dn);
    atomicAdd(&G.cnt->bad, bd);
    atomicAdd(&G.cnt->rhs, r);
  }
}

__global__ __launch_bounds__(REV_BLOCK) void adjoint_rev_kernel(GradLaunch G) { adjoint_rev_body<0>(G); }
__global__ __launch_bounds__(REV_BLOCK, 1) void adjoint_rev_stored_kernel(GradLaunch G) { adjoint_rev_body<1>(G); }
__global__ __launch_bounds__(REV_BLOCK, 1) void adjoint_rev_chunked_kernel(GradLaunch G) { adjoint_rev_body<2>(G); }

// Seeds of the recorded rays (the recording render replaces the re-trace):
// the crossing step is the ray's last record, its end state the render's
// final state.
__global__ __launch_bounds__(256) void adjoint_seed_kernel(GradLaunch G) {
  const uint64_t q = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  unsigned long long n_bad = 0;
  if (q < G.n_list) {
    const uint32_t pix = G.list[q];
    const uint32_t att = G.keys[q];
    const uint32_t chunk = G.last_chunk[pix];
    double *rec = G.rec + REC * q;
    bool ok = att > 0 && chunk != 0xFFFFFFFFu && chunk < G.pool_chunks;
    if (ok) {
      const double *r = G.pool + (uint64_t)chunk * GRAD_CHUNK_DOUBLES + GRAD_CHUNK_HEAD + STEP_REC * ((att - 1) % GRAD_CHUNK);
      const uint64_t n = G.n_pix;
      const double *fs = G.final_state;
      ok = r[REC_KIND] == (double)REC_ACCEPT &&
           crossing_seeds(G, G.scene, rec, r[REC_R], r[REC_TH], r[REC_V], r[REC_V + 3], fs[n + pix],
                          fs[2 * n + pix], fs[4 * n + pix], fs[7 * n + pix]);
      put_gv(rec, d4{fs[4 * n + pix], fs[5 * n + pix], fs[6 * n + pix], fs[7 * n + pix]});
      rec[R_CHK] = as_f64(chunk);
    }
    if (!ok) mark_bad(G, q, pix, n_bad);
  }
  const unsigned long long bd = wave_sum(n_bad);
  if ((threadIdx.x & 63) == 0 && bd) atomicAdd(&G.cnt->bad, bd);
}

// The camera ray's dependence on (a, i): pos0 = (0, r_cam, i, 0), vel0 from
// Camera::pixel_direction and GeodesicRay::from_camera; dI per pixel.
__global__ __launch_bounds__(256) void adjoint_cam_kernel(GradLaunch G) {
  const uint64_t q = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= G.n_list) return;
  const uint32_t pix = G.list[q];
  const double *rec = G.rec + REC * q;
  const double *o0 = G.dI + 4 * (uint64_t)pix;
  if (!(o0[0] == o0[0])) return;   // flagged ray keeps its NaNs
  const DevScene &S = G.scene;
  const uint32_t px = pix % G.width, py = pix / G.width;
  const CamD ca = incl_camera_ray_d(S.mass, Tn{S.spin, 1.0}, dconst(G.incl), G.r_cam, S.half_h, S.half_w, px, py,
                                    G.width, G.height);
  const CamD ci = incl_camera_ray_d(S.mass, dconst(S.spin), Tn{G.incl, 1.0}, G.r_cam, S.half_h, S.half_w, px, py,
                                    G.width, G.height);
  const d4 vb{rec[R_VB0], rec[R_VB1], rec[R_VB2], rec[R_VB3]};
  double *o = G.dI + 4 * (uint64_t)pix;
  o[0] = rec[R_GA] + vb.x * ca.k[0].d + vb.y * ca.k[1].d + vb.z * ca.k[2].d + vb.w * ca.k[3].d;
  o[1] = rec[R_GI] + rec[R_PBTH] + vb.x * ci.k[0].d + vb.y * ci.k[1].d + vb.z * ci.k[2].d + vb.w * ci.k[3].d;
  o[2] = rec[R_GQ];
  o[3] = rec[R_GRIN];
}

// Residuals, zeroed gradients and the list of disk rays with their attempt
// counts (sort keys: step records), checkpoint and last-segment totals, the
// accepted and attempt totals, the largest accepted count and the most
// attempts a ray did not accept.  One atomic per block for the list and each
// total (the list order does not matter: it is sorted next).
constexpr uint32_t SELECT_BLOCK = 1024;
__global__ __launch_bounds__(SELECT_BLOCK) void select_kernel(GradSelect P) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const uint32_t lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  bool disk = false;
  uint32_t nchk = 0, tail = 0, acc = 0, att = 0;
  if (i < P.n) {
    P.resid[i] = P.intensity[i] - P.target[i];
    double *o = P.dI + 4 * i;
    o[0] = 0.0; o[1] = 0.0; o[2] = 0.0; o[3] = 0.0;
    disk = P.cls[i] == GG_CLASS_DISK;
    if (disk) {
      acc = P.accepted[i];
      att = P.attempts[i];
      if (att < acc) att = acc;
      nchk = (acc + P.K - 1) / P.K;
      // the last segment holds its accepted steps and at most every attempt
      // the ray did not accept
      tail = att - (nchk ? (nchk - 1) * P.K : 0u);
    }
  }
  const unsigned long long m = __ballot(disk);
  const unsigned long long tot = wave_sum(nchk), tl = wave_sum(tail), all = wave_sum(acc), alt = wave_sum(att);
  uint32_t mx = acc, mr = att - acc;
  for (int off = 32; off > 0; off >>= 1) {
    mx = max(mx, (uint32_t)__shfl_xor(mx, off, 64));
    mr = max(mr, (uint32_t)__shfl_xor(mr, off, 64));
  }
  constexpr uint32_t NW = SELECT_BLOCK / 64;
  __shared__ unsigned long long s_cnt[NW], s_tot[NW], s_tl[NW], s_all[NW], s_alt[NW];
  __shared__ uint32_t s_mx[NW], s_mr[NW];
  __shared__ unsigned long long s_base;
  if (lane == 0) {
    s_cnt[wave] = __popcll(m); s_tot[wave] = tot; s_tl[wave] = tl; s_all[wave] = all; s_alt[wave] = alt;
    s_mx[wave] = mx; s_mr[wave] = mr;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    unsigned long long c = 0, t = 0, l = 0, al = 0, at = 0;
    uint32_t x = 0, xr = 0;
    for (uint32_t w = 0; w < blockDim.x / 64; ++w) {
      c += s_cnt[w]; t += s_tot[w]; l += s_tl[w]; al += s_all[w]; at += s_alt[w];
      x = max(x, s_mx[w]); xr = max(xr, s_mr[w]);
    }
    unsigned long long base = 0;
    if (c) {
      base = atomicAdd(&P.cnt->n_list, c);
      atomicAdd(&P.cnt->chk_total, t);
      atomicAdd(&P.cnt->tail_total, l);
      atomicAdd(&P.cnt->acc_total, al);
      atomicAdd(&P.cnt->att_total, at);
      atomicMax(&P.cnt->max_acc, (unsigned long long)x);
      atomicMax(&P.cnt->max_rej, (unsigned long long)xr);
    }
    s_base = base;
  }
  __syncthreads();
  if (disk) {
    unsigned long long at = s_base + __popcll(m & ((1ull << lane) - 1ull));
    for (uint32_t w = 0; w < wave; ++w) at += s_cnt[w];
    P.list[at] = (uint32_t)i;
    P.keys[at] = att;
  }
}

// loss = mean r^2, grad = 2 mean r dI in two deterministic passes: block b
// sums a fixed contiguous range of pixels (fixed per-thread order, fixed
// tree), then one block sums the block partials in order.
constexpr uint32_t REDUCE_BLOCKS = 256, REDUCE_THREADS = 256;
__device__ inline void block_tree5(double acc[5], double (*part)[REDUCE_THREADS]) {
  const uint32_t t = threadIdx.x;
  for (int k = 0; k < 5; ++k) part[k][t] = acc[k];
  __syncthreads();
  for (uint32_t w = REDUCE_THREADS / 2; w > 0; w >>= 1) {
    if (t < w)
      for (int k = 0; k < 5; ++k) part[k][t] += part[k][t + w];
    __syncthreads();
  }
}

__global__ __launch_bounds__(REDUCE_THREADS) void reduce_partial_kernel(uint64_t n, const double *resid,
                                                                        const double *dI, double *partials) {
  __shared__ double part[5][REDUCE_THREADS];
  const uint64_t per = (n + gridDim.x - 1) / gridDim.x;
  const uint64_t begin = (uint64_t)blockIdx.x * per;
  const uint64_t end = begin + per < n ? begin + per : n;
  double acc[5] = {0, 0, 0, 0, 0};
  for (uint64_t i = begin + threadIdx.x; i < end; i += REDUCE_THREADS) {
    const double r = resid[i];
    acc[0] += r * r;
    const double *d = dI + 4 * i;
    acc[1] += r * d[0]; acc[2] += r * d[1]; acc[3] += r * d[2]; acc[4] += r * d[3];
  }
  block_tree5(acc, part);
  if (threadIdx.x == 0)
    for (int k = 0; k < 5; ++k) partials[5 * blockIdx.x + k] = part[k][0];
}

__global__ __launch_bounds__(REDUCE_THREADS) void reduce_final_kernel(uint64_t n, uint32_t nb, const double *partials,
                                                                      double *out5) {
  __shared__ double part[5][REDUCE_THREADS];
  double acc[5] = {0, 0, 0, 0, 0};
  for (uint32_t b = threadIdx.x; b < nb; b += REDUCE_THREADS)
    for (int k = 0; k < 5; ++k) acc[k] += partials[5 * b + k];
  block_tree5(acc, part);
  if (threadIdx.x == 0) {
    const double inv = 1.0 / (double)n;
    out5[0] = part[0][0] * inv;
    for (int k = 1; k < 5; ++k) out5[k] = 2.0 * part[k][0] * inv;
  }
}

__global__ void grad_reset_kernel(GradCounters *c) {
  if (threadIdx.x < sizeof(GradCounters) / 8) reinterpret_cast<unsigned long long *>(c)[threadIdx.x] = 0ull;
}

// w^T J of the Kerr RHS at n points: pos_bar (t, phi components 0),
// vel_bar, spin_bar.
__global__ __launch_bounds__(256) void vjp_kernel(const DevScene *S, uint64_t n, const double *pos, const double *vel, const double *cot,
                           double *pos_bar, double *vel_bar, double *spin_bar) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  d4 v{vel[4 * i], vel[4 * i + 1], vel[4 * i + 2], vel[4 * i + 3]};
  d4 w{cot[4 * i], cot[4 * i + 1], cot[4 * i + 2], cot[4 * i + 3]};
  RhsBar b = kerr_rhs_vjp(S->mass, S->spin, pos[4 * i + 1], pos[4 * i + 2], v, w);
  pos_bar[4 * i] = 0.0; pos_bar[4 * i + 1] = b.r; pos_bar[4 * i + 2] = b.th; pos_bar[4 * i + 3] = 0.0;
  vel_bar[4 * i] = b.v.x; vel_bar[4 * i + 1] = b.v.y; vel_bar[4 * i + 2] = b.v.z; vel_bar[4 * i + 3] = b.v.w;
  spin_bar[i] = b.a;
}

}  // namespace

// LDS of the reverse kernels: nine slots of 4 doubles (the recomputation's
// k_1..k_5, the reverse step's cotangents kbar_1..kbar_5 in the same slots,
// its stage velocities v_2..v_5); of the forward kernel: the trace kernel's
// k_1..k_5.
size_t adjoint_lds_bytes(uint32_t block) { return (size_t)block * 9 * 4 * sizeof(double); }
size_t adjoint_fwd_lds_bytes(uint32_t block) { return (size_t)block * 5 * 4 * sizeof(double); }

hipError_t launch_adjoint_fwd(const GradLaunch &g, uint32_t grid, uint32_t block, hipStream_t s) {
  if (block != FWD_BLOCK) return hipErrorInvalidValue;
  hipLaunchKernelGGL(adjoint_fwd_kernel, dim3(grid), dim3(block), adjoint_fwd_lds_bytes(block), s, g);
  return hipGetLastError();
}

hipError_t launch_adjoint_rev(const GradLaunch &g, int mode, uint32_t grid, uint32_t block, hipStream_t s) {
  if (block != REV_BLOCK) return hipErrorInvalidValue;
  if (mode == 0)
    hipLaunchKernelGGL(adjoint_rev_kernel, dim3(grid), dim3(block), adjoint_lds_bytes(block), s, g);
  else if (mode == 1)
    hipLaunchKernelGGL(adjoint_rev_stored_kernel, dim3(grid), dim3(block), adjoint_lds_bytes(block), s, g);
  else
    hipLaunchKernelGGL(adjoint_rev_chunked_kernel, dim3(grid), dim3(block), adjoint_lds_bytes(block), s, g);
  return hipGetLastError();
}

hipError_t launch_adjoint_seed(const GradLaunch &g, hipStream_t s) {
  if (g.n_list == 0) return hipSuccess;
  hipLaunchKernelGGL(adjoint_seed_kernel, dim3((uint32_t)((g.n_list + 255) / 256)), dim3(256), 0, s, g);
  return hipGetLastError();
}

hipError_t launch_adjoint_cam(const GradLaunch &g, hipStream_t s) {
  if (g.n_list == 0) return hipSuccess;
  hipLaunchKernelGGL(adjoint_cam_kernel, dim3((uint32_t)((g.n_list + 255) / 256)), dim3(256), 0, s, g);
  return hipGetLastError();
}

// Disk rays by attempt count, longest first (radix sort of (count, pixel)).
hipError_t sort_rays_desc(const uint32_t *keys_in, uint32_t *keys_out, const uint32_t *vals_in, uint32_t *vals_out,
                          uint64_t n, int key_bits, void *temp, size_t &temp_bytes, hipStream_t s) {
  return hipcub::DeviceRadixSort::SortPairsDescending(temp, temp_bytes, keys_in, keys_out, vals_in, vals_out,
                                                      n, 0, key_bits, s);
}

hipError_t launch_grad_select(const GradSelect &p, hipStream_t s) {
  if (p.n == 0) return hipSuccess;
  hipLaunchKernelGGL(select_kernel, dim3((uint32_t)((p.n + SELECT_BLOCK - 1) / SELECT_BLOCK)), dim3(SELECT_BLOCK), 0,
                     s, p);
  return hipGetLastError();
}

size_t grad_reduce_scratch_bytes() { return sizeof(double) * 5 * REDUCE_BLOCKS; }

hipError_t launch_grad_reduce(uint64_t n, const double *resid, const double *dI, double *partials, double *out5,
                              hipStream_t s) {
  hipLaunchKernelGGL(reduce_partial_kernel, dim3(REDUCE_BLOCKS), dim3(REDUCE_THREADS), 0, s, n, resid, dI, partials);
  hipLaunchKernelGGL(reduce_final_kernel, dim3(1), dim3(REDUCE_THREADS), 0, s, n, REDUCE_BLOCKS,
                     (const double *)partials, out5);
  return hipGetLastError();
}

hipError_t launch_grad_reset(GradCounters *c, hipStream_t s) {
  hipLaunchKernelGGL(grad_reset_kernel, dim3(1), dim3(64), 0, s, c);
  return hipGetLastError();
}

hipError_t launch_rhs_vjp(const DevScene *S, uint64_t n, const double *pos, const double *vel, const double *cot,
                          double *pos_bar, double *vel_bar, double *spin_bar, hipStream_t s) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(vjp_kernel, dim3((uint32_t)((n + 255) / 256)), dim3(256), 0, s, S, n, pos, vel, cot, pos_bar,
                     vel_bar, spin_bar);
  return hipGetLastError();
}

}  // namespace gg
