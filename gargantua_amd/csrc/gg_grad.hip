// gg_grad.hip -- hand-written reverse-mode adjoint of the thin-disk render.
//
// d(loss)/d(spin, inclination, q, r_in) for L = mean_p (I_p - target_p)^2,
// differentiating the discrete FP64 algorithm with every ray's step-size
// sequence, accept/reject pattern and disk-crossing step held at the values
// of the forward render (discretise-then-differentiate).
//
// The disk-hitting rays, sorted by accepted-step count (longest first, so
// the long photon-ring rays start early and neighbouring lanes stay in the
// same phase), go through three kernels, one lane per ray:
//   adjoint_fwd_kernel  re-traces the ray from the camera with the trace
//       kernel's own attempt() (bit-identical steps), storing a checkpoint
//       (pos, vel, h) every K accepted steps in a pool sized exactly from
//       the render's accepted counts, and the start state and h of each step
//       of the ray's last segment; at the render's accepted count the step is
//       the disk crossing, whose intensity is differentiated in forward mode
//       (12 inputs): the seeds of the reverse sweep.
//   adjoint_rev_kernel  back-propagates one accepted step per iteration
//       through its Dormand-Prince stages (the six k_i are recomputed, their
//       cotangents live in LDS), segment by segment from the last; before
//       each earlier segment it re-runs that segment from its checkpoint into
//       a per-lane segment buffer (rejections included, so h repeats exactly,
//       and the end state must equal the next checkpoint bit for bit).
//       When K is at least every ray's accepted count there is one segment
//       per ray, the re-trace stores every step record, and
//       adjoint_rev_stored_kernel (the same sweep without recomputation,
//       two wavefronts per SIMD) reads them.
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

// Adjoint of one accepted Kerr step (p, v) -> (p + h v normalised,
// v + h sum_i c_i k_i), k_i = a(p + h a_i v, v_i), from its step record
// (p, v, v_2..v_6, h).  On entry (pb_r, pb_th, vb) is the cotangent of the
// step's end state (t and phi carry none: nothing depends on them); on exit
// that of its start.  ga accumulates the spin cotangent.  LDS slots 0..5
// hold the stage cotangents kbar_1..kbar_6.
__device__ inline void step_adjoint(const KStore &K, const double *rec_in, double m, double a, double &pb_r,
                                    double &pb_th, d4 &vb, double &ga) {
  GG_CONTRACT_FAST
  const double *rec = (const double *)__builtin_assume_aligned(rec_in, 32);
  const double r0 = rec[1], th0 = rec[2];
  const d4 v{rec[4], rec[5], rec[6], rec[7]};
  const double h = rec[REC_H];
  // normalize_spherical_coords (geodesic.rs:337-352): theta -> -theta or
  // 2 pi - theta flips the sign of its cotangent
  double th1 = h * v.z + th0;
  double sg = 1.0;
  if (th1 < 0.0) { th1 = -th1; sg = -sg; }
  if (th1 > PI) sg = -sg;
  pb_th *= sg;
  // new_vel = v + h sum_i c_i k_i: kbar_i = h c_i vb; pos' = p + h v
  const d4 vbp = vb;
#pragma unroll 1
  for (int i = 1; i <= 6; ++i) {
    double hc = h * kC[i];
    K.put(i - 1, d4{hc * vbp.x, hc * vbp.y, hc * vbp.z, hc * vbp.w});
  }
  double pr = pb_r, pth = pb_th;
  d4 vbar{vbp.x, vbp.y + h * pb_r, vbp.z + h * pb_th, vbp.w};
#pragma unroll 1
  for (int i = 6; i >= 1; --i) {
    const d4 w = K.get(i - 1);
    double qr = r0, qth = th0;
    d4 vs = v;
    double ha = 0.0;
    if (i >= 2) {
      // stage position p + h a_i v, stage velocity v + h sum_j b_ij k_j
      ha = h * kA[i];
      qr = ha * v.y + r0;
      qth = ha * v.z + th0;
      const double *o = rec + REC_VS + 4 * (i - 2);
      vs = d4{o[0], o[1], o[2], o[3]};
    }
    const RhsBar jb = kerr_rhs_vjp(m, a, qr, qth, vs, w);
    ga += jb.a;
    pr += jb.r;
    pth += jb.th;
    vbar.x += jb.v.x; vbar.y += jb.v.y; vbar.z += jb.v.z; vbar.w += jb.v.w;
    if (i >= 2) {
      vbar.y += ha * jb.r;
      vbar.z += ha * jb.th;
#pragma unroll 1
      for (int j = 1; j < i; ++j) {
        double hb = h * kB[i][j];
        d4 kb = K.get(j - 1);
        kb.x += hb * jb.v.x; kb.y += hb * jb.v.y; kb.z += hb * jb.v.z; kb.w += hb * jb.v.w;
        K.put(j - 1, kb);
      }
    }
  }
  pb_r = pr;
  pb_th = pth;
  vb = vbar;
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
  R_CHK, R_TAIL,                                                            // pool offsets (bits)
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

__global__ __launch_bounds__(256, 2) void adjoint_fwd_kernel(GradLaunch G) {
#pragma clang fp contract(off)
  extern __shared__ double lds_k[];
  const KStore K{lds_k + threadIdx.x, blockDim.x};
  const uint32_t lane = threadIdx.x & 63;
  const DevScene &S = *G.scene;
  const uint32_t KK = G.K;
  Lane L;
  L.n_rhs = 0; L.n_rhs32 = 0; L.n_rhsf = 0; L.n_full = 0;
  bool active = false, exhausted = false;
  uint64_t q = 0, chk_base = 0, tail_base = 0;
  uint32_t pix = 0, acc_target = 0, last0 = 0, att = 0;
  unsigned long long n_att = 0, n_bad = 0;
  for (;;) {
    const uint64_t t = take(&G.cnt->queue, G.n_list, !active, exhausted);
    if (t != ~0ull) {
      q = t;
      pix = G.list[q];
      acc_target = G.keys[q];
      const uint32_t nchk = (acc_target + KK - 1) / KK;
      last0 = (nchk - 1) * KK;
      chk_base = atomicAdd(&G.cnt->chk_cursor, (unsigned long long)nchk);
      tail_base = atomicAdd(&G.cnt->tail_cursor, (unsigned long long)(acc_target - last0));
      camera_ray(S, pix % G.width, pix / G.width, L.pos, L.vel);
      L.lambda = 0.0;
      L.h = S.h0;
      L.k1_ok = false; L.k1_f32 = false; L.k1_fast = false;
      L.accepted = 0;
      L.err_prev = 1.0;
      L.b2 = 0.0;
      att = 0;
      active = true;
      if (acc_target == 0 || chk_base + nchk > G.chk_cap || tail_base + (acc_target - last0) > G.tail_cap) {
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
    double *recp = L.accepted >= last0 ? G.tail + STEP_REC * (tail_base + (L.accepted - last0)) : nullptr;
    const int res = attempt<1, 0, true>(S, L, K, acc, p0, v0, false, false, recp);
    ++n_att;
    if (res != RES_CONTINUE || ++att > S.max_steps) {   // the render reached the crossing; a re-trace must too
      mark_bad(G, q, pix, n_bad);
      active = false;
      continue;
    }
    if (!acc) continue;
    L.accepted++;
    if (L.accepted < acc_target) {
      if (L.accepted % KK == 0) put_state(G.chk + 9 * (chk_base + L.accepted / KK), L.pos, L.vel, L.h);
      continue;
    }
    // the disk crossing of the last accepted step (same test as the trace kernel)
    active = false;
    double *rec = G.rec + REC * q;
    if (!crossing_seeds(G, S, rec, p0[1], p0[2], v0.x, v0.w, L.pos[1], L.pos[2], L.vel.x, L.vel.w)) {
      mark_bad(G, q, pix, n_bad);
      continue;
    }
    rec[R_CHK] = as_f64(chk_base);
    rec[R_TAIL] = as_f64(tail_base);
  }
  unsigned long long a = wave_sum(n_att), bd = wave_sum(n_bad), r = wave_sum(L.n_rhs);
  if (lane == 0) {
    atomicAdd(&G.cnt->attempts, a);
    atomicAdd(&G.cnt->bad, bd);
    atomicAdd(&G.cnt->rhs, r);
  }
}

// The reverse sweep.  MODE 0: some rays have more than one segment, so the
// earlier segments are recomputed from their checkpoints.  MODE 1: every
// step record of every ray is in the last-segment pool (one segment per
// ray).  MODE 2: the records are the recording render's chunks, walked back
// through their back-pointers.  Modes 1 and 2 only read records (fewer live
// registers: two wavefronts per SIMD).
template <int MODE>
__device__ __forceinline__ void adjoint_rev_body(const GradLaunch &G) {
#pragma clang fp contract(off)
  constexpr bool RECOMPUTE = (MODE == 0);
  extern __shared__ double lds_k[];
  const KStore K{lds_k + threadIdx.x, blockDim.x};
  const uint32_t lane = threadIdx.x & 63;
  const uint64_t gtid = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const DevScene &S = *G.scene;
  const uint32_t KK = G.K;
  double *const seg = RECOMPUTE ? G.seg + gtid * (uint64_t)KK * STEP_REC : nullptr;

  Lane L;
  L.n_rhs = 0; L.n_rhs32 = 0; L.n_rhsf = 0; L.n_full = 0;
  enum { IDLE = 0, REV = 1, RECOMP = 2 };
  int phase = IDLE;
  bool exhausted = false, first = false;
  uint64_t q = 0, chk_base = 0;
  uint32_t pix = 0, c = 0, j = 0, att = 0, chunk = 0;
  const double *src = nullptr;
  double pb_r = 0, pb_th = 0, ga = 0;
  d4 vb{0, 0, 0, 0};
  unsigned long long n_att = 0, n_adj = 0, n_done = 0, n_bad = 0;

  for (;;) {
    const uint64_t t = take(&G.cnt->queue2, G.n_list, phase == IDLE, exhausted);
    if (t != ~0ull) {
      q = t;
      pix = G.list[q];
      const uint32_t acc_target = G.keys[q];
      const double *rec = G.rec + REC * q;
      if (!(rec[R_PBR] == rec[R_PBR])) {   // flagged by the forward pass
        phase = IDLE;
      } else {
        // without recomputation the whole ray is one segment
        const uint32_t nchk = RECOMPUTE ? (acc_target + KK - 1) / KK : 1;
        c = nchk - 1;
        j = acc_target - c * KK - 1;
        if (MODE == 2) {
          chunk = (uint32_t)as_u64(rec[R_CHK]);
        } else {
          chk_base = as_u64(rec[R_CHK]);
          src = G.tail + STEP_REC * as_u64(rec[R_TAIL]);
        }
        pb_r = rec[R_PBR]; pb_th = rec[R_PBTH];
        vb = d4{rec[R_VB0], rec[R_VB1], rec[R_VB2], rec[R_VB3]};
        ga = rec[R_GA];
        first = true;
        phase = REV;
      }
    }
    if (exhausted && __all(phase == IDLE)) break;

    // ---- RECOMP: one attempt of the segment before ----
    if (RECOMPUTE && phase == RECOMP) {
      bool acc;
      double p0[4];
      d4 v0;
      double *recp = seg + STEP_REC * (uint64_t)(L.accepted - c * KK);
      const int res = attempt<1, 0, true>(S, L, K, acc, p0, v0, false, false, recp);
      ++n_att;
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
            j = KK - 1;
            src = seg;
            phase = REV;
          }
        }
      }
    }

    // ---- REV: one accepted step backwards ----
    if (phase == REV) {
      const double *sp = MODE == 2 ? G.pool + (uint64_t)chunk * GRAD_CHUNK_DOUBLES + GRAD_CHUNK_HEAD + STEP_REC * (j % GRAD_CHUNK)
                                   : src + STEP_REC * (uint64_t)j;
      step_adjoint(K, sp, S.mass, S.spin, pb_r, pb_th, vb, ga);
      ++n_adj;
      if (first) {
        const double *rec = G.rec + REC * q;
        pb_r += rec[R_XR0]; pb_th += rec[R_XTH0]; vb.x += rec[R_XV0T]; vb.w += rec[R_XV0P];
        first = false;
      }
      if (j > 0) {
        if (MODE == 2 && j % GRAD_CHUNK == 0) chunk = (uint32_t)as_u64(G.pool[(uint64_t)chunk * GRAD_CHUNK_DOUBLES]);
        --j;
      } else if (RECOMPUTE && c > 0) {
        --c;
        get_state(G.chk + 9 * (chk_base + c), L.pos, L.vel, L.h);
        L.k1_ok = false; L.k1_f32 = false; L.k1_fast = false;
        L.accepted = c * KK;
        att = 0;
        phase = RECOMP;
      } else {
        double *rec = G.rec + REC * q;
        rec[R_PBR] = pb_r; rec[R_PBTH] = pb_th;
        rec[R_VB0] = vb.x; rec[R_VB1] = vb.y; rec[R_VB2] = vb.z; rec[R_VB3] = vb.w;
        rec[R_GA] = ga;
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

__global__ __launch_bounds__(64) void adjoint_rev_kernel(GradLaunch G) { adjoint_rev_body<0>(G); }
__global__ __launch_bounds__(64, 2) void adjoint_rev_stored_kernel(GradLaunch G) { adjoint_rev_body<1>(G); }
__global__ __launch_bounds__(64, 2) void adjoint_rev_chunked_kernel(GradLaunch G) { adjoint_rev_body<2>(G); }

// Seeds of the recorded rays (the recording render replaces the re-trace):
// the crossing step is the ray's last record, its end state the render's
// final state.
__global__ __launch_bounds__(256) void adjoint_seed_kernel(GradLaunch G) {
  const uint64_t q = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  unsigned long long n_bad = 0;
  if (q < G.n_list) {
    const uint32_t pix = G.list[q];
    const uint32_t acc = G.keys[q];
    const uint32_t chunk = G.last_chunk[pix];
    double *rec = G.rec + REC * q;
    bool ok = acc > 0 && chunk != 0xFFFFFFFFu && chunk < G.pool_chunks;
    if (ok) {
      const double *r = G.pool + (uint64_t)chunk * GRAD_CHUNK_DOUBLES + GRAD_CHUNK_HEAD + STEP_REC * ((acc - 1) % GRAD_CHUNK);
      const uint64_t n = G.n_pix;
      const double *fs = G.final_state;
      ok = crossing_seeds(G, *G.scene, rec, r[1], r[2], r[4], r[7], fs[n + pix], fs[2 * n + pix], fs[4 * n + pix],
                          fs[7 * n + pix]);
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
  const DevScene &S = *G.scene;
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

// Residuals, zeroed gradients and the list of disk rays with their
// accepted counts (sort keys), checkpoint and last-segment totals, the total
// and the largest accepted count.  One atomic per block for the list and
// each total (the list order does not matter: it is sorted next).
constexpr uint32_t SELECT_BLOCK = 1024;
__global__ __launch_bounds__(SELECT_BLOCK) void select_kernel(GradSelect P) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const uint32_t lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  bool disk = false;
  uint32_t nchk = 0, last = 0, acc = 0;
  if (i < P.n) {
    P.resid[i] = P.intensity[i] - P.target[i];
    double *o = P.dI + 4 * i;
    o[0] = 0.0; o[1] = 0.0; o[2] = 0.0; o[3] = 0.0;
    disk = P.cls[i] == GG_CLASS_DISK;
    if (disk) {
      acc = P.accepted[i];
      nchk = (acc + P.K - 1) / P.K;
      last = nchk ? acc - (nchk - 1) * P.K : 0;
    }
  }
  const unsigned long long m = __ballot(disk);
  const unsigned long long tot = wave_sum(nchk), tl = wave_sum(last), all = wave_sum(acc);
  uint32_t mx = acc;
  for (int off = 32; off > 0; off >>= 1) mx = max(mx, (uint32_t)__shfl_xor(mx, off, 64));
  constexpr uint32_t NW = SELECT_BLOCK / 64;
  __shared__ unsigned long long s_cnt[NW], s_tot[NW], s_tl[NW], s_all[NW];
  __shared__ uint32_t s_mx[NW];
  __shared__ unsigned long long s_base;
  if (lane == 0) { s_cnt[wave] = __popcll(m); s_tot[wave] = tot; s_tl[wave] = tl; s_all[wave] = all; s_mx[wave] = mx; }
  __syncthreads();
  if (threadIdx.x == 0) {
    unsigned long long c = 0, t = 0, l = 0, al = 0;
    uint32_t x = 0;
    for (uint32_t w = 0; w < blockDim.x / 64; ++w) {
      c += s_cnt[w]; t += s_tot[w]; l += s_tl[w]; al += s_all[w]; x = max(x, s_mx[w]);
    }
    unsigned long long base = 0;
    if (c) {
      base = atomicAdd(&P.cnt->n_list, c);
      atomicAdd(&P.cnt->chk_total, t);
      atomicAdd(&P.cnt->tail_total, l);
      atomicAdd(&P.cnt->acc_total, al);
      atomicMax(&P.cnt->max_acc, (unsigned long long)x);
    }
    s_base = base;
  }
  __syncthreads();
  if (disk) {
    unsigned long long at = s_base + __popcll(m & ((1ull << lane) - 1ull));
    for (uint32_t w = 0; w < wave; ++w) at += s_cnt[w];
    P.list[at] = (uint32_t)i;
    P.keys[at] = acc;
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

// LDS of the reverse kernel: six slots of 4 doubles (the recomputation's
// k_1..k_5, the reverse step's cotangents kbar_1..kbar_6); of the forward
// kernel: the trace kernel's k_1..k_5.
size_t adjoint_lds_bytes(uint32_t block) { return (size_t)block * 6 * 4 * sizeof(double); }
size_t adjoint_fwd_lds_bytes(uint32_t block) { return (size_t)block * 5 * 4 * sizeof(double); }

hipError_t launch_adjoint_fwd(const GradLaunch &g, uint32_t grid, uint32_t block, hipStream_t s) {
  hipLaunchKernelGGL(adjoint_fwd_kernel, dim3(grid), dim3(block), adjoint_fwd_lds_bytes(block), s, g);
  return hipGetLastError();
}

hipError_t launch_adjoint_rev(const GradLaunch &g, int mode, uint32_t grid, uint32_t block, hipStream_t s) {
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

// Disk rays by accepted count, longest first (radix sort of (count, pixel)).
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
