// gg_internal.hpp -- declarations shared by the host ABI and the kernels.
#pragma once

#ifndef __HIPCC_RTC__  // hipRTC supplies the runtime and fixed-width types itself
#include <hip/hip_runtime.h>
#include <stdint.h>
#endif

#include "gg_physics.hpp"

namespace gg {

// Output pointers of one scene, device addresses.
struct DevOut {
  uint8_t *cls;
  double *r, *theta, *phi, *intensity;
  uint32_t *attempts, *accepted;
  double *state;        // 9 planes of n_rays each
  double *color;        // scene modes: linear RGB of each ray's scene hit (3 per ray, 0 when none)
  uint64_t n_rays;      // rays of this scene (plane stride of `state`)
};

// Global tallies written by the trace kernel (zeroed by reset_counters).
struct Counters {
  unsigned long long queue;          // next ray index to hand out
  unsigned long long attempts;
  unsigned long long rhs;
  unsigned long long rhs32;
  unsigned long long attempts_full;
  unsigned long long retired;
  unsigned long long rhsf;           // RHS evaluations of fast-FP64 attempts (mixed)
  unsigned long long reflections;    // accepted steps reflected at a pole
};

struct TraceLaunch {
  const DevScene *scenes;     // device array [n_scenes]
  const DevOut *outs;         // device array [n_outs]
  uint32_t n_scenes;
  uint64_t n_rays;            // total over scenes
  const double *states_in;    // gg_trace_states: 9 doubles per ray (AoS), else null
  Counters *counters;
  uint32_t refill_interval;
  const uint32_t *order;      // camera renders: queue position -> global ray, or null (ray order)
  // Recording render of gg_grad (Kerr thin disk, one scene), else null:
  // every attempt's step record in chunks (GRAD_CHUNK_DOUBLES each: the
  // index of the ray's previous chunk in a 128-byte line of its own, then
  // GRAD_CHUNK step records);
  // rec_last[output index] = the ray's last chunk.
  double *rec_pool;
  unsigned long long rec_cap;       // chunks in the pool
  unsigned long long *rec_cursor;   // chunks claimed; beyond rec_cap nothing is written
  uint32_t *rec_last;
  // A one-scene launch reads its scene from here, by value in the kernel
  // arguments: the constant kernel-argument segment lets every read of a
  // scene constant be a scalar load, where a read through `scenes` (global
  // memory the kernel also writes) is a vector load the wavefront waits on.
  DevScene scene;
};

// Tallies of gg_grad (zeroed by launch_grad_reset).
struct GradCounters {
  unsigned long long queue;        // next disk ray for the forward kernel
  unsigned long long queue2;       // next disk ray for the reverse kernel
  unsigned long long chk_cursor;   // checkpoints claimed in the pool
  unsigned long long tail_cursor;  // last-segment step records claimed
  unsigned long long attempts;     // attempts of re-traces and recomputations
  unsigned long long steps;        // step records (attempts) back-propagated
  unsigned long long done;         // rays whose reverse sweep finished
  unsigned long long bad;          // rays whose re-trace left the rendered path
  unsigned long long rhs;          // RHS evaluations inside attempts
  unsigned long long n_list;       // disk rays found by select
  unsigned long long chk_total;    // checkpoints they need
  unsigned long long tail_total;   // step records their last segments can hold
  unsigned long long acc_total;    // accepted steps of all disk rays
  unsigned long long max_acc;      // the largest accepted count
  unsigned long long rec_cursor;   // chunks claimed by the recording render
  unsigned long long att_total;    // attempts of all disk rays (their step records)
  unsigned long long max_rej;      // the most attempts a disk ray did not accept
  unsigned long long pad[1];
};

struct GradSelect {
  uint64_t n;                      // pixels
  const uint8_t *cls;
  const double *intensity, *target;
  const uint32_t *accepted, *attempts;
  uint32_t K;
  double *resid, *dI;
  uint32_t *list, *keys;           // disk pixels and their attempt counts (unsorted)
  GradCounters *cnt;
};

struct GradLaunch {
  DevScene scene;                  // by value: scene constants are scalar loads from the kernel arguments
  const uint32_t *list;            // disk pixels, longest first
  const uint32_t *keys;            // their attempt counts (step records)
  const uint32_t *accepted;        // accepted count per pixel (the render's)
  uint64_t n_list;
  double *rec;                     // per-ray records
  double *chk;                     // checkpoint pool, 9 doubles each
  unsigned long long chk_cap;
  double *tail;                    // last-segment step records, GRAD_STEP_REC doubles each
  unsigned long long tail_cap;
  double *seg;                     // per-thread segment buffers of the reverse kernel, seg_cap step records
  uint32_t seg_cap;
  uint32_t live;                   // 1: differentiate the step-size controller; 0: step sizes held fixed
  uint32_t K;
  uint32_t width, height;
  double incl, r_cam;
  double *dI;                      // [pixels][4]
  GradCounters *cnt;
  // records of the recording render (reverse mode 2), else null
  const double *pool;              // chunks, GRAD_CHUNK_DOUBLES each
  unsigned long long pool_chunks;
  const uint32_t *last_chunk;      // per pixel
  const double *final_state;       // the render's final states: 9 planes of n_pix doubles
  uint64_t n_pix;
};

constexpr uint32_t GRAD_REC = 24;       // doubles per ray record
constexpr uint32_t GRAD_STEP_REC = 32;  // doubles per step record: r, theta, h, kind, a velocity, v_2..v_6, err diff
constexpr uint32_t GRAD_CHUNK = 16;     // step records per chunk of the recording render
constexpr uint32_t GRAD_CHUNK_HEAD = 16;   // the back-pointer's 128-byte line
constexpr uint32_t GRAD_CHUNK_DOUBLES = GRAD_CHUNK_HEAD + GRAD_CHUNK * GRAD_STEP_REC;

// Equirectangular environment map on the device (HdriSky, sky.rs:43-122),
// or pixels null for the procedural starfield.
struct SkyDesc {
  const float *pixels;             // width x height x 3, row 0 = +z pole
  uint32_t width, height;
  float intensity;
};

// ---- host-side launchers (not part of a module compiled at run time) ----
#ifndef __HIPCC_RTC__
size_t adjoint_lds_bytes(uint32_t block);
size_t adjoint_fwd_lds_bytes(uint32_t block);
hipError_t launch_adjoint_fwd(const GradLaunch &g, uint32_t grid, uint32_t block, hipStream_t s);
hipError_t launch_adjoint_rev(const GradLaunch &g, int mode, uint32_t grid, uint32_t block, hipStream_t s);
hipError_t launch_adjoint_seed(const GradLaunch &g, hipStream_t s);
hipError_t launch_adjoint_cam(const GradLaunch &g, hipStream_t s);
hipError_t sort_rays_desc(const uint32_t *keys_in, uint32_t *keys_out, const uint32_t *vals_in, uint32_t *vals_out,
                          uint64_t n, int key_bits, void *temp, size_t &temp_bytes, hipStream_t s);
hipError_t launch_grad_select(const GradSelect &p, hipStream_t s);
size_t grad_reduce_scratch_bytes();
hipError_t launch_grad_reduce(uint64_t n, const double *resid, const double *dI, double *partials, double *out5,
                              hipStream_t s);
hipError_t launch_grad_reset(GradCounters *c, hipStream_t s);
hipError_t launch_rhs_vjp(const DevScene *S, uint64_t n, const double *pos, const double *vel, const double *cot,
                          double *pos_bar, double *vel_bar, double *spin_bar, hipStream_t s);

// Launch the persistent trace kernel for (metric, disk, precision); disk
// 3 and 4 are the scene modes (gg_trace_scene.hip, one scene per launch).
hipError_t launch_trace_scene(int metric_kind, int mode, int precision, const TraceLaunch &t, uint32_t grid,
                              uint32_t block, hipStream_t stream);
hipError_t launch_trace(int metric_kind, int disk_kind, int precision, const TraceLaunch &t, uint32_t grid,
                        uint32_t block, hipStream_t stream);
hipError_t launch_reset(Counters *c, hipStream_t stream);
hipError_t launch_rhs(int metric_kind, int precision, const DevScene *scene, uint64_t n, const double *pos,
                      const double *vel, double *acc, hipStream_t stream);
hipError_t launch_constants(int metric_kind, const DevScene *scene, uint64_t n, const double *states, double *out,
                            hipStream_t stream);
hipError_t launch_camera_rays(const DevScene *scene, uint64_t n, double *states, hipStream_t stream);
// keys[i] = attempts of global ray i (from its scene's attempts output),
// ids[i] = i: input of the cost-order sort.
hipError_t launch_order_keys(const DevScene *scenes, const DevOut *outs, uint32_t n_scenes, uint64_t n,
                             uint32_t *keys, uint32_t *ids, hipStream_t stream);
// RGB image (renderer.rs:44-80 shading + sRGB) from per-ray outputs.
hipError_t launch_shade(const DevScene *scene, const DevOut *out, uint32_t width, uint32_t height, uint32_t spa,
                        const SkyDesc &sky, uint8_t *rgb, hipStream_t stream);
#endif

}  // namespace gg
