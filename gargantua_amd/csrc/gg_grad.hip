// gg_grad.hip -- hand-written reverse-mode adjoint of the thin-disk render.
//
// d(loss)/d(spin, inclination, q, r_in) for L = mean_p (I_p - target_p)^2,
// differentiating the discrete FP64 algorithm with every ray's step-size
// sequence, accept/reject pattern and disk-crossing step held at the values
// of the forward render (discretise-then-differentiate).
//
// Per disk-hitting ray, one lane:
//   FWD    re-traces the ray from the camera with the trace kernel's own
//          attempt() (bit-identical steps), storing a checkpoint (pos, vel,
//          h) every K accepted steps in a pool sized exactly from the
//          render's accepted counts, and every accepted step's start state
//          in a K-deep per-lane segment buffer; at the render's accepted
//          count the step is the disk crossing, whose intensity is
//          differentiated in forward mode (12 inputs, once per ray).
//   REV    back-propagates one accepted step per iteration through its
//          Dormand-Prince stages (the six k_i are recomputed, their
//          cotangents live in LDS) from the end of the current segment to
//          its start.
//   RECOMP re-runs the segment before it from its checkpoint into the
//          segment buffer (rejections included, so h repeats exactly).
// Lanes in different phases share a wavefront: a lane takes one attempt or
// one step adjoint per iteration and a new ray from the queue when done.
#include <hip/hip_runtime.h>

#include "gg_adjoint.hpp"
#include "gg_internal.hpp"
#include "gg_rk45.hpp"

namespace gg {

namespace {

enum { PH_IDLE = 0, PH_FWD = 1, PH_REV = 2, PH_RECOMP = 3 };

constexpr int KBAR = 6;   // LDS slots 6..11 hold the stage cotangents

// v + h sum_j b_sj k_j of stage s (s = 2..6) from the k_j in LDS slots
// 0..s-2, in the forward kernel's order.
__device__ inline d4 stage_vel(const KStore &K, int s, d4 v, double h) {
#pragma clang fp contract(off)
  d4 k1 = K.get(0);
  if (s == 2) {
    double hb = h * B21;
    return d4{hb * k1.x + v.x, hb * k1.y + v.y, hb * k1.z + v.z, hb * k1.w + v.w};
  }
  double b = kB[s][1];
  d4 sum{b * k1.x, b * k1.y, b * k1.z, b * k1.w};
#pragma unroll 1
  for (int j = 2; j < s; ++j) {
    double bj = kB[s][j];
    d4 kj = K.get(j - 1);
    sum.x = bj * kj.x + sum.x; sum.y = bj * kj.y + sum.y; sum.z = bj * kj.z + sum.z; sum.w = bj * kj.w + sum.w;
  }
  return d4{h * sum.x + v.x, h * sum.y + v.y, h * sum.z + v.z, h * sum.w + v.w};
}

// Adjoint of one accepted Kerr step (p, v) -> (p + h v normalised,
// v + h sum_i c_i k_i).  On entry (pb_r, pb_th, vb) is the cotangent of the
// step's end state (t and phi carry none: nothing depends on them); on exit
// that of its start.  ga accumulates the spin cotangent.
__device__ inline void step_adjoint(const DevScene &S, const KStore &K, const double p[4], d4 v, double h,
                                    double &pb_r, double &pb_th, d4 &vb, double &ga) {
#pragma clang fp contract(off)
  // normalize_spherical_coords (geodesic.rs:337-352): theta -> -theta or
  // 2 pi - theta flips the sign of its cotangent
  double th1 = h * v.z + p[2];
  double sg = 1.0;
  if (th1 < 0.0) { th1 = -th1; sg = -sg; }
  if (th1 > PI) sg = -sg;
  pb_th *= sg;
  // the six stage values k_1..k_6 (slots 0..5)
  K.put(0, rhs64<1>(S, p[1], p[2], v));
#pragma unroll 1
  for (int s = 2; s <= 6; ++s) {
    double ha = h * kA[s];
    d4 vs = stage_vel(K, s, v, h);
    K.put(s - 1, rhs64<1>(S, ha * v.y + p[1], ha * v.z + p[2], vs));
  }
  // new_vel = v + h sum_i c_i k_i: kbar_i = h c_i vb; pos' = p + h v
  const d4 vbp = vb;
#pragma unroll 1
  for (int i = 1; i <= 6; ++i) {
    double hc = h * kC[i];
    K.put(KBAR + i - 1, d4{hc * vbp.x, hc * vbp.y, hc * vbp.z, hc * vbp.w});
  }
  double pr = pb_r, pth = pb_th;
  d4 vbar{vbp.x, vbp.y + h * pb_r, vbp.z + h * pb_th, vbp.w};
  const double m = S.mass, a = S.spin;
#pragma unroll 1
  for (int i = 6; i >= 1; --i) {
    const d4 w = K.get(KBAR + i - 1);
    double qr = p[1], qth = p[2];
    d4 vs = v;
    double ha = 0.0;
    if (i >= 2) {
      ha = h * kA[i];
      qr = ha * v.y + p[1];
      qth = ha * v.z + p[2];
      vs = stage_vel(K, i, v, h);
    }
    const RhsBar jb = kerr_rhs_vjp(m, a, qr, qth, vs, w);
    ga += jb.a;
    pr += jb.r;
    pth += jb.th;
    vbar.x += jb.v.x; vbar.y += jb.v.y; vbar.z += jb.v.z; vbar.w += jb.v.w;
    if (i >= 2) {
      // stage position p + h a_i v, stage velocity v + h sum_j b_ij k_j
      vbar.y += ha * jb.r;
      vbar.z += ha * jb.th;
#pragma unroll 1
      for (int j = 1; j < i; ++j) {
        double hb = h * kB[i][j];
        d4 kb = K.get(KBAR + j - 1);
        kb.x += hb * jb.v.x; kb.y += hb * jb.v.y; kb.z += hb * jb.v.z; kb.w += hb * jb.v.w;
        K.put(KBAR + j - 1, kb);
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

__global__ __launch_bounds__(64) void adjoint_kernel(GradLaunch G) {
#pragma clang fp contract(off)
  extern __shared__ double lds_k[];
  const KStore K{lds_k + threadIdx.x, blockDim.x};
  const uint32_t lane = threadIdx.x & 63;
  const uint64_t gtid = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const DevScene &S = *G.scene;
  const uint32_t KK = G.K;
  double *const seg = G.seg + gtid * (uint64_t)KK * 9;

  Lane L;
  L.n_rhs = 0; L.n_rhs32 = 0; L.n_full = 0;
  int phase = PH_IDLE;
  uint32_t pix = 0, acc_target = 0, nchk = 0, c = 0, j = 0, att_phase = 0;
  uint64_t chk_base = 0;
  double pb_r = 0, pb_th = 0, ga = 0, gi = 0, gq = 0, grin = 0;
  d4 vb{0, 0, 0, 0};
  double x_r0 = 0, x_th0 = 0, x_v0t = 0, x_v0p = 0;   // crossing's direct cotangents on the last step's start
  bool first_rev = false;
  bool exhausted = false;
  unsigned long long n_att = 0, n_adj = 0, n_done = 0, n_bad = 0;

  auto fail_ray = [&]() {
    double *o = G.dI + 4 * (uint64_t)pix;
    const double nan = __longlong_as_double(0x7ff8000000000000ll);
    o[0] = nan; o[1] = nan; o[2] = nan; o[3] = nan;
    ++n_bad;
    phase = PH_IDLE;
  };

  for (;;) {
    // ---- hand idle lanes new rays: one atomicAdd per wavefront ----
    const unsigned long long idle = __ballot(phase == PH_IDLE);
    if (!exhausted && idle) {
      const uint32_t cnt = __popcll(idle);
      const int leader = __ffsll((long long)idle) - 1;
      unsigned long long base = 0;
      if ((int)lane == leader) base = atomicAdd(&G.cnt->queue, (unsigned long long)cnt);
      base = __shfl(base, leader, 64);
      if (base + cnt >= G.n_list) exhausted = true;
      const uint32_t rank = __popcll(idle & ((1ull << lane) - 1ull));
      if (phase == PH_IDLE && base + rank < G.n_list) {
        pix = G.list[base + rank];
        acc_target = G.accepted[pix];
        nchk = (acc_target + KK - 1) / KK;
        chk_base = atomicAdd(&G.cnt->chk_cursor, (unsigned long long)nchk);
        const uint32_t px = pix % G.width, py = pix / G.width;
        camera_ray(S, px, py, L.pos, L.vel);
        L.lambda = 0.0;
        L.h = S.h0;
        L.k1_ok = false; L.k1_f32 = false;
        L.attempts = 0; L.accepted = 0;
        L.err_prev = 1.0;
        pb_r = pb_th = ga = gi = gq = grin = 0.0;
        vb = d4{0, 0, 0, 0};
        att_phase = 0;
        phase = PH_FWD;
        if (acc_target == 0 || chk_base + nchk > G.chk_cap) fail_ray();
        else put_state(G.chk + 9 * chk_base, L.pos, L.vel, L.h);
      }
    }
    if (exhausted && __all(phase == PH_IDLE)) break;

    // ---- FWD / RECOMP: one attempt ----
    if (phase == PH_FWD || phase == PH_RECOMP) {
      const double h_used = L.h;
      bool acc;
      double p0[4];
      d4 v0;
      const int res = attempt<1, 0>(S, L, K, acc, p0, v0, false);
      ++n_att;
      ++att_phase;
      if (res != RES_CONTINUE || att_phase > S.max_steps) {
        fail_ray();   // the render reached the crossing; a re-trace must too
      } else if (acc) {
        const uint32_t g = L.accepted++;
        put_state(seg + 9 * (uint64_t)(g % KK), p0, v0, h_used);
        if (phase == PH_FWD) {
          if (L.accepted == acc_target) {
            // the disk crossing (same test as the trace kernel)
            const bool side0 = p0[2] < HALF_PI, side1 = L.pos[2] < HALF_PI;
            double rc = 0.0;
            if (side0 != side1) {
              double s = (HALF_PI - p0[2]) / (L.pos[2] - p0[2]);
              rc = p0[1] + s * (L.pos[1] - p0[1]);
            }
            const double w = S.edge_width;
            const bool hit = (side0 != side1) &&
                             ((w > 0.0) ? (rc <= S.r_out + 8.0 * w && rc > 0.0) : (rc >= S.r_in && rc <= S.r_out));
            if (!hit) {
              fail_ray();
            } else {
              CrossVals x{S.spin, S.q, S.r_in, S.sqrt_neg_gtt_obs, p0[1], p0[2], L.pos[1], L.pos[2],
                          v0.x, v0.w, L.vel.x, L.vel.w};
              double d[CX_N];
#pragma unroll 1
              for (int k = 0; k < CX_N; ++k) d[k] = crossing_intensity_d(S.mass, S.r_out, w, x, k).d;
              const Tn so_a = kerr_sqrt_obs_d(S.mass, Tn{S.spin, 1.0}, G.r_cam, dconst(G.incl));
              const Tn so_i = kerr_sqrt_obs_d(S.mass, dconst(S.spin), G.r_cam, Tn{G.incl, 1.0});
              ga = d[CX_A] + d[CX_SOBS] * so_a.d;
              gi = d[CX_SOBS] * so_i.d;
              gq = d[CX_Q];
              grin = d[CX_RIN];
              pb_r = d[CX_R1];
              pb_th = d[CX_TH1];
              vb = d4{d[CX_V1T], 0.0, 0.0, d[CX_V1P]};
              x_r0 = d[CX_R0]; x_th0 = d[CX_TH0]; x_v0t = d[CX_V0T]; x_v0p = d[CX_V0P];
              c = nchk - 1;
              j = acc_target - 1 - c * KK;
              first_rev = true;
              phase = PH_REV;
            }
          } else if (L.accepted % KK == 0) {
            put_state(G.chk + 9 * (chk_base + L.accepted / KK), L.pos, L.vel, L.h);
          }
        } else if (L.accepted == (c + 1) * KK) {
          // the recomputed segment must land bit for bit on the next checkpoint
          double q[4], hq;
          d4 vq;
          get_state(G.chk + 9 * (chk_base + c + 1), q, vq, hq);
          if (q[0] != L.pos[0] || q[1] != L.pos[1] || q[2] != L.pos[2] || q[3] != L.pos[3] || vq.x != L.vel.x ||
              vq.y != L.vel.y || vq.z != L.vel.z || vq.w != L.vel.w || hq != L.h) {
            fail_ray();
          } else {
            j = KK - 1;
            phase = PH_REV;
          }
        }
      }
    }

    // ---- REV: one accepted step backwards ----
    if (phase == PH_REV) {
      double p[4], h;
      d4 v;
      get_state(seg + 9 * (uint64_t)j, p, v, h);
      step_adjoint(S, K, p, v, h, pb_r, pb_th, vb, ga);
      ++n_adj;
      if (first_rev) {
        pb_r += x_r0; pb_th += x_th0; vb.x += x_v0t; vb.w += x_v0p;
        first_rev = false;
      }
      if (j > 0) {
        --j;
      } else if (c > 0) {
        --c;
        get_state(G.chk + 9 * (chk_base + c), L.pos, L.vel, L.h);
        L.k1_ok = false; L.k1_f32 = false;
        L.accepted = c * KK;
        att_phase = 0;
        phase = PH_RECOMP;
      } else {
        // camera: pos0 = (0, r_cam, i, 0), vel0 from Camera::pixel_direction
        // and GeodesicRay::from_camera
        const uint32_t px = pix % G.width, py = pix / G.width;
        const CamD ca = incl_camera_ray_d(S.mass, Tn{S.spin, 1.0}, dconst(G.incl), G.r_cam, S.half_h, S.half_w,
                                          px, py, G.width, G.height);
        const CamD ci = incl_camera_ray_d(S.mass, dconst(S.spin), Tn{G.incl, 1.0}, G.r_cam, S.half_h, S.half_w,
                                          px, py, G.width, G.height);
        ga += vb.x * ca.k[0].d + vb.y * ca.k[1].d + vb.z * ca.k[2].d + vb.w * ca.k[3].d;
        gi += pb_th + vb.x * ci.k[0].d + vb.y * ci.k[1].d + vb.z * ci.k[2].d + vb.w * ci.k[3].d;
        double *o = G.dI + 4 * (uint64_t)pix;
        o[0] = ga; o[1] = gi; o[2] = gq; o[3] = grin;
        ++n_done;
        phase = PH_IDLE;
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

// Residuals, zeroed gradients and the list of disk rays with their
// checkpoint count.  One atomic per wavefront for the list and the total.
__global__ void select_kernel(GradSelect P) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const uint32_t lane = threadIdx.x & 63;
  bool disk = false;
  uint32_t nchk = 0;
  if (i < P.n) {
    P.resid[i] = P.intensity[i] - P.target[i];
    double *o = P.dI + 4 * i;
    o[0] = 0.0; o[1] = 0.0; o[2] = 0.0; o[3] = 0.0;
    disk = P.cls[i] == GG_CLASS_DISK;
    if (disk) nchk = (P.accepted[i] + P.K - 1) / P.K;
  }
  const unsigned long long m = __ballot(disk);
  unsigned long long tot = wave_sum(nchk);
  if (m) {
    const int leader = __ffsll((long long)m) - 1;
    unsigned long long base = 0;
    if ((int)lane == leader) {
      base = atomicAdd(&P.cnt->n_list, (unsigned long long)__popcll(m));
      atomicAdd(&P.cnt->chk_total, tot);
    }
    base = __shfl(base, leader, 64);
    if (disk) P.list[base + __popcll(m & ((1ull << lane) - 1ull))] = (uint32_t)i;
  }
}

// loss = mean r^2, grad = 2 mean r dI: one block, fixed summation order.
__global__ __launch_bounds__(1024) void reduce_kernel(uint64_t n, const double *resid, const double *dI,
                                                      double *out5) {
  __shared__ double part[5][1024];
  const uint32_t t = threadIdx.x;
  double acc[5] = {0, 0, 0, 0, 0};
  for (uint64_t i = t; i < n; i += 1024) {
    double r = resid[i];
    acc[0] += r * r;
    const double *d = dI + 4 * i;
    acc[1] += r * d[0]; acc[2] += r * d[1]; acc[3] += r * d[2]; acc[4] += r * d[3];
  }
  for (int k = 0; k < 5; ++k) part[k][t] = acc[k];
  __syncthreads();
  for (uint32_t w = 512; w > 0; w >>= 1) {
    if (t < w)
      for (int k = 0; k < 5; ++k) part[k][t] += part[k][t + w];
    __syncthreads();
  }
  if (t == 0) {
    double inv = 1.0 / (double)n;
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

// LDS of the adjoint kernel: k_1..k_6 and their cotangents, 4 doubles each.
size_t adjoint_lds_bytes(uint32_t block) { return (size_t)block * 12 * 4 * sizeof(double); }

hipError_t launch_adjoint(const GradLaunch &g, uint32_t grid, uint32_t block, hipStream_t s) {
  hipLaunchKernelGGL(adjoint_kernel, dim3(grid), dim3(block), adjoint_lds_bytes(block), s, g);
  return hipGetLastError();
}

hipError_t launch_grad_select(const GradSelect &p, hipStream_t s) {
  if (p.n == 0) return hipSuccess;
  hipLaunchKernelGGL(select_kernel, dim3((uint32_t)((p.n + 255) / 256)), dim3(256), 0, s, p);
  return hipGetLastError();
}

hipError_t launch_grad_reduce(uint64_t n, const double *resid, const double *dI, double *out5, hipStream_t s) {
  hipLaunchKernelGGL(reduce_kernel, dim3(1), dim3(1024), 0, s, n, resid, dI, out5);
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
