// gg_trace_kernel.hpp -- the persistent geodesic trace kernel (template).
//
// One lane carries one ray with its whole state in registers.  Wavefronts
// take rays from a global atomic queue; every `refill_interval` attempts a
// block compacts its live rays into the lowest lanes through LDS and hands
// the freed tail lanes new rays with a single atomicAdd, so lanes whose rays
// finished early do not idle while neighbours run photon-ring orbits of
// 10-50x more steps.  The step itself (RK45Integrator::step, its FP32 stage
// variant for the mixed kernel) lives in gg_rk45.hpp.  gg_trace.hip
// instantiates the disk modes, gg_trace_scene.hip the scene modes.
#pragma once

#ifndef __HIPCC_RTC__  // hipRTC supplies the runtime and fixed-width types itself
#include <hip/hip_runtime.h>
#endif

#include "gg_internal.hpp"
#include "gg_rk45.hpp"
#include "gg_scene.hpp"

namespace gg {

namespace {

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

// Persistent trace kernel.  SRC 0: rays from the camera; 1: states_in;
// 2: rays from the camera, recording every attempt (the step record the
// adjoint reads: start state, h, kind, stage velocities, error difference)
// into chunks of GRAD_CHUNK records, each chunk headed by the index of the
// ray's previous chunk, and each ray's last chunk into rec_last.
// BATCH 0: one scene, its constants wavefront-uniform (scalar registers);
// BATCH 1: each lane looks up the scene of its ray.
// DISK 0 none, 1 volumetric, 2 thin; 3 the Hittable scene alone
// (trace_ray_with_scene), 4 volumetric disk and scene
// (trace_ray_with_disk_and_scene).
// GG_MIN_WAVES: waves per SIMD the register budget must allow (3: 168
// VGPRs, the disk modes; 2: 256, the Hittable scene modes; the run-time
// compiled user metrics take 1, for 512 VGPRs + AGPRs).  The recording
// kernels (SRC 2, the differentiable render's forward pass, whose grids are
// small enough to run one wavefront per SIMD) take at most 2: no scratch.
#ifndef GG_MIN_WAVES
#define GG_MIN_WAVES 2
#endif
// Doubles of LDS per wavefront: five stage slots of 64 lanes x 4 doubles.
constexpr uint32_t KREG = 5 * 64 * 4;
template <int METRIC, int DISK, int PREC, int SRC, int BATCH>
__global__ __launch_bounds__(256, (SRC == 2 && GG_MIN_WAVES > 2) ? 2 : GG_MIN_WAVES) void trace_kernel(TraceLaunch T) {
#pragma clang fp contract(off)
  const uint32_t lane = threadIdx.x & 63;
  const uint64_t n = T.n_rays;
  constexpr bool single = (BATCH == 0);

  // LDS: one region of KREG doubles per wavefront (its lanes' stage slots), so
  // the slot offsets are immediates whatever the block size
  extern __shared__ __align__(16) double lds_k[];
  const KStoreT<64> K{reinterpret_cast<double2 *>(lds_k + (threadIdx.x >> 6) * KREG) + (threadIdx.x & 63), 64};
  Lane L;
  L.n_rhs = 0; L.n_rhs32 = 0; L.n_rhsf = 0; L.n_full = 0; L.n_refl = 0;
  uint32_t n_att = 0, n_retired = 0;
  bool active = false;
  uint64_t ray = 0;
  uint32_t si = 0;        // scene of this lane's ray
  uint64_t out_i = 0;     // output slot of this lane's ray
  bool exhausted = false;
  constexpr bool record = (SRC == 2);
  constexpr uint32_t NO_CHUNK = 0xFFFFFFFFu;
  uint32_t r_chunk = NO_CHUNK;   // record chunk holding attempt `attempts`
  bool r_need = false;           // a new chunk is due before the next attempt
  // Scene constants are read through the pointer where they are used:
  // scalar loads from the kernel arguments for one scene, vector loads of
  // the lane's own scene in a batch.  The kernel-argument pointer passes
  // through an empty asm statement at the top of every iteration of the
  // persistent loop, so that the compiler reloads the constants each
  // iteration instead of hoisting ~75 invariant loads out of the loop and
  // holding (spilling) them across it.
  // (T is the kernel's only explicit argument: it starts the segment.)
  typedef const __attribute__((address_space(4))) DevScene KScene;
  typedef const __attribute__((address_space(4))) char KByte;
  KScene *scene_k = (KScene *)((KByte *)__builtin_amdgcn_kernarg_segment_ptr() + __builtin_offsetof(TraceLaunch, scene));
  const DevScene *__restrict__ Sp = single ? (const DevScene *)scene_k : T.scenes;

  // Start ray `ray` in this lane.
  auto init_ray = [&](uint64_t rid) {
    ray = (SRC != 1 && T.order) ? (uint64_t)T.order[rid] : rid;
    si = single ? 0u : find_scene(T.scenes, T.n_scenes, ray);
    if (!single) Sp = T.scenes + si;
    uint64_t local = ray - Sp->ray_begin;
    if (SRC != 1) {
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
    L.k1_fast = false;
    L.b2 = 0.0;
    if (PREC == 1) {
      double c[4];
      constants_of_motion<METRIC>(*Sp, L.pos, L.vel, c);
      L.b2 = c[0] > 0.0 ? (c[2] + c[1] * c[1]) / (c[0] * c[0]) : 0.0;
    }
    L.attempts = 0; L.accepted = 0;
    L.intensity = 0.0; L.transmission = 1.0;
    L.err_prev = 1.0;
    if (record) { r_chunk = NO_CHUNK; r_need = true; }
    active = true;
  };

  __shared__ uint32_t s_cnt[3][4];
  __shared__ unsigned long long s_base;
  uint32_t iter = 0;
  for (;;) {
    if constexpr (single) {
      asm volatile("" : "+s"(scene_k));
      Sp = (const DevScene *)scene_k;
    }
    {
      // ---- block regroup: compaction (and precision sorting when mixed) ----
      // Every refill_interval attempts the block's wavefronts exchange ray
      // states through LDS: live rays move to the lowest lanes, retired lanes
      // gather at the tail and take new rays with one atomicAdd for the
      // block.  The mixed kernel orders the live rays by the arithmetic their
      // next attempts can take -- FP32-eligible, then fast-FP64-able, then
      // those that need exact FP64 (an unsaturated controller with a tiny
      // previous error, where the error's last bit decides the next h) -- so
      // that wavefronts are uniform: a wavefront takes the costliest tier
      // any of its lanes needs.  Rays keep their class for many attempts.
      if (iter % T.refill_interval == 0) {
        const uint32_t tid = threadIdx.x, B = blockDim.x, wave = tid >> 6;
        const uint32_t key = !active ? 3u
                             : (PREC != 1 ? 1u
                                          : (fp32_eligible(*Sp, L) ? 0u
                                                                   : ((L.h < Sp->max_step && L.err_prev < 1e-6) ? 2u : 1u)));
        const unsigned long long m0 = __ballot(key == 0), m1 = __ballot(key == 1), m2 = __ballot(key == 2);
        if (lane == 0) { s_cnt[0][wave] = __popcll(m0); s_cnt[1][wave] = __popcll(m1); s_cnt[2][wave] = __popcll(m2); }
        __syncthreads();
        uint32_t n0 = 0, n1 = 0, n2 = 0, o0 = 0, o1 = 0, o2 = 0;
        for (uint32_t w = 0; w < (B >> 6); ++w) {
          if (w < wave) { o0 += s_cnt[0][w]; o1 += s_cnt[1][w]; o2 += s_cnt[2][w]; }
          n0 += s_cnt[0][w];
          n1 += s_cnt[1][w];
          n2 += s_cnt[2][w];
        }
        const unsigned long long below = (1ull << lane) - 1ull;
        const uint32_t dst = key == 0   ? o0 + __popcll(m0 & below)
                             : key == 1 ? n0 + o1 + __popcll(m1 & below)
                                        : n0 + n1 + o2 + __popcll(m2 & below);
        const uint32_t n_act = n0 + n1 + n2;
        // The exchange reuses stage slots 1..4 of every wavefront's region
        // (free between attempts) as 16 planes of 64 doubles per wavefront:
        // field f of block lane i at X(f, i); two passes of at most 14 fields.
        auto X = [&](uint32_t f, uint32_t i) -> double & {
          return lds_k[(i >> 6) * KREG + 256 + f * 64 + (i & 63)];
        };
        d4 k1 = K.get(0);
        __syncthreads();
        if (active) {
          X(0, dst) = L.pos[0]; X(1, dst) = L.pos[1]; X(2, dst) = L.pos[2]; X(3, dst) = L.pos[3];
          X(4, dst) = L.vel.x; X(5, dst) = L.vel.y; X(6, dst) = L.vel.z; X(7, dst) = L.vel.w;
          X(8, dst) = L.lambda; X(9, dst) = L.h;
          X(10, dst) = k1.x; X(11, dst) = k1.y; X(12, dst) = k1.z; X(13, dst) = k1.w;
        }
        __syncthreads();
        const bool now_active = tid < n_act;
        if (now_active) {
          L.pos[0] = X(0, tid); L.pos[1] = X(1, tid); L.pos[2] = X(2, tid); L.pos[3] = X(3, tid);
          L.vel = d4{X(4, tid), X(5, tid), X(6, tid), X(7, tid)};
          L.lambda = X(8, tid); L.h = X(9, tid);
          k1 = d4{X(10, tid), X(11, tid), X(12, tid), X(13, tid)};
        }
        __syncthreads();
        if (active) {
          X(0, dst) = L.intensity; X(1, dst) = L.transmission; X(2, dst) = L.err_prev;
          X(3, dst) = __longlong_as_double((long long)ray);
          X(4, dst) = __longlong_as_double((long long)out_i);
          X(5, dst) = __longlong_as_double((long long)(((uint64_t)L.attempts << 32) | L.accepted));
          X(6, dst) = __longlong_as_double(
              (long long)(((uint64_t)si << 3) | (L.k1_fast ? 4u : 0u) | (L.k1_f32 ? 2u : 0u) | (L.k1_ok ? 1u : 0u)));
          X(7, dst) = L.b2;
          if (record) X(8, dst) = __longlong_as_double((long long)(((uint64_t)r_need << 32) | r_chunk));
        }
        __syncthreads();
        if (now_active) {
          L.intensity = X(0, tid); L.transmission = X(1, tid); L.err_prev = X(2, tid);
          ray = (uint64_t)__double_as_longlong(X(3, tid));
          out_i = (uint64_t)__double_as_longlong(X(4, tid));
          uint64_t aa = (uint64_t)__double_as_longlong(X(5, tid));
          L.attempts = (uint32_t)(aa >> 32); L.accepted = (uint32_t)aa;
          uint64_t sf = (uint64_t)__double_as_longlong(X(6, tid));
          si = (uint32_t)(sf >> 3); L.k1_ok = (sf & 1u) != 0; L.k1_f32 = (sf & 2u) != 0; L.k1_fast = (sf & 4u) != 0;
          L.b2 = X(7, tid);
          if (record) {
            const uint64_t cw = (uint64_t)__double_as_longlong(X(8, tid));
            r_chunk = (uint32_t)cw;
            r_need = (cw >> 32) != 0;
          }
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
    double scol[3];           // linear colour of a scene hit (DISK 3, 4)
    if (L.attempts >= Sp->max_steps) {
      cls = GG_CLASS_MAX_STEPS;
    } else {
      bool acc;
      double p0[4];
      d4 v0;
      double lam0 = L.lambda;
      bool use32 = false, use_fast = false;
      if (PREC == 1) {
        // tiers per wavefront, over the lanes still running
        use32 = __all(fp32_eligible(*Sp, L));
        use_fast = __all(!(L.h < Sp->max_step && L.err_prev < 1e-6));
      }
      double *recp = nullptr;
      if (record) {
        // a new chunk every GRAD_CHUNK attempts, one atomicAdd per
        // wavefront; chunks beyond the pool are counted, not written
        const unsigned long long m = __ballot(r_need);
        if (m) {
          const int leader = __ffsll((long long)m) - 1;
          unsigned long long base = 0;
          if ((int)lane == leader) base = atomicAdd(T.rec_cursor, (unsigned long long)__popcll(m));
          base = __shfl(base, leader, 64);
          if (r_need) {
            const unsigned long long mine = base + __popcll(m & ((1ull << lane) - 1ull));
            const uint32_t prev = r_chunk;
            r_chunk = mine < T.rec_cap ? (uint32_t)mine : NO_CHUNK;
            if (r_chunk != NO_CHUNK)
              T.rec_pool[(uint64_t)r_chunk * GRAD_CHUNK_DOUBLES] = __longlong_as_double((long long)prev);
            r_need = false;
          }
        }
        if (r_chunk != NO_CHUNK)
          recp = T.rec_pool + (uint64_t)r_chunk * GRAD_CHUNK_DOUBLES + GRAD_CHUNK_HEAD + GRAD_STEP_REC * (L.attempts % GRAD_CHUNK);
      }
      int res = attempt<METRIC, PREC, record>(*Sp, L, K, acc, p0, v0, use32, use_fast, recp);
      L.attempts++;
      n_att++;
      if (record && L.attempts % GRAD_CHUNK == 0) r_need = true;
      if (acc) L.accepted++;
      // A chord whose two end points lie inside the ball r < scene_rmin stays
      // in it (the ball is convex) and no primitive reaches into it: no test.
      if (DISK >= 3 && acc && fmax(p0[1], L.pos[1]) >= Sp->scene_rmin) {
        // The accepted step's Cartesian chord against the scene, before the
        // step's own result (tracer.rs:220-238, 282-305).
        double cb[3], ca[3];
        sph2cart(p0, cb);
        sph2cart(L.pos, ca);
        const double seg[3] = {ca[0] - cb[0], ca[1] - cb[1], ca[2] - cb[2]};
        if (sqrt(sdot3(seg, seg)) > 0.0) {
          const SceneObj *objs = (const SceneObj *)Sp->objs;
          SceneHit hh;
          if (scene_hit(objs, (const BvhNode *)Sp->nodes, cb, seg, hh)) {
            double lin[3];
            shade_hit(objs[hh.obj], seg, hh.nrm, lin);
            const double g4 = g4_static<METRIC>(*Sp, hh.p);
            if (DISK == 4) {
              // scene_shaded g^4 transmission + disk_color(intensity so far)
              const double x = L.intensity / (1.0 + L.intensity);
              const double dc[3] = {pow(x, 0.5), pow(x, 1.2), pow(x, 2.5)};
              for (int c = 0; c < 3; ++c) scol[c] = lin[c] * g4 * L.transmission + dc[c];
            } else {
              for (int c = 0; c < 3; ++c) scol[c] = lin[c] * g4;
            }
            cls = GG_CLASS_SCENE;
          }
        }
      }
      if (DISK == 1 || DISK == 4) {
        // trace_ray_with_disk (tracer.rs:96-145); with the scene
        // (tracer.rs:307-347) the absorption coefficient is the constant 0.1
        double dl = L.lambda - lam0;
        dl = dl >= 0.0 ? dl : 0.0;
        if (cls == 255 && dl > 0.0) {
          double mr = 0.5 * (p0[1] + L.pos[1]);
          double mth = 0.5 * (p0[2] + L.pos[2]);
          double rho = disk_density(*Sp, mr, mth);
          if (rho > 0.0) {
            double j = disk_emission(*Sp, rho, mr, mth);
            d4 pm{0.5 * (v0.x + L.vel.x), 0.5 * (v0.y + L.vel.y), 0.5 * (v0.z + L.vel.z), 0.5 * (v0.w + L.vel.w)};
            double g;
            if constexpr (METRIC == 2) {
              const double mid[4] = {0.5 * (p0[0] + L.pos[0]), mr, mth, 0.5 * (p0[3] + L.pos[3])};
              g = user_redshift(*Sp, mid, pm);
            } else {
              g = redshift64<METRIC>(*Sp, mr, mth, pm);
            }
            double g2 = g * g;
            double red = j * (g2 * g2);
            L.intensity = L.intensity + red * dl * L.transmission;
            L.transmission *= exp(-rho * dl * (DISK == 4 ? 0.1 : Sp->absorption));
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
              double g;
              if constexpr (METRIC == 2) {
                const double pc[4] = {p0[0] + s * (L.pos[0] - p0[0]), rc, HALF_PI, p0[3]};
                g = user_redshift(*Sp, pc, vc);
              } else {
                g = redshift64<METRIC>(*Sp, rc, HALF_PI, vc);
              }
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
    if (cls == GG_CLASS_MAX_STEPS && (DISK == 1 || DISK == 4) && L.intensity > 0.0) cls = GG_CLASS_DISK;
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
      if (DISK >= 3 && O.color) {
        double *oc = O.color + 3 * out_i;
        if (cls == GG_CLASS_SCENE) { oc[0] = scol[0]; oc[1] = scol[1]; oc[2] = scol[2]; }
        else { oc[0] = 0.0; oc[1] = 0.0; oc[2] = 0.0; }
      }
      if (record) T.rec_last[out_i] = r_chunk;
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
                     rf = wave_sum(L.n_rhsf), f = wave_sum(L.n_full), ret = wave_sum(n_retired),
                     rl = wave_sum(L.n_refl);
  if (lane == 0) {
    atomicAdd(&T.counters->retired, ret);
    atomicAdd(&T.counters->attempts, a);
    atomicAdd(&T.counters->rhs, r);
    atomicAdd(&T.counters->rhs32, r32);
    atomicAdd(&T.counters->rhsf, rf);
    atomicAdd(&T.counters->attempts_full, f);
    if (rl) atomicAdd(&T.counters->reflections, rl);
  }
}

// k1..k5 of every lane: 5 slots x 4 components x 8 bytes.
inline size_t lds_bytes(uint32_t block) { return (size_t)block * 5 * 4 * sizeof(double); }

}  // namespace

}  // namespace gg
