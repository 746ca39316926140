// gg_internal.hpp -- declarations shared by the host ABI and the kernels.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "gg_physics.hpp"

namespace gg {

// Output pointers of one scene, device addresses.
struct DevOut {
  uint8_t *cls;
  double *r, *theta, *phi, *intensity;
  uint32_t *attempts, *accepted;
  double *state;        // 9 planes of n_rays each
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
  unsigned long long pad[2];
};

struct TraceLaunch {
  const DevScene *scenes;     // device array [n_scenes]
  const DevOut *outs;         // device array [n_outs]
  uint32_t n_scenes;
  uint64_t n_rays;            // total over scenes
  const double *states_in;    // gg_trace_states: 9 doubles per ray (AoS), else null
  Counters *counters;
  uint32_t refill_interval;
};

// Launch the persistent trace kernel for (metric, disk, precision).
hipError_t launch_trace(int metric_kind, int disk_kind, int precision, const TraceLaunch &t, uint32_t grid,
                        uint32_t block, hipStream_t stream);
hipError_t launch_reset(Counters *c, hipStream_t stream);
hipError_t launch_rhs(int metric_kind, int precision, const DevScene *scene, uint64_t n, const double *pos,
                      const double *vel, double *acc, hipStream_t stream);
// RGB image (renderer.rs:44-80 shading + sRGB) from per-ray outputs.
hipError_t launch_shade(const DevScene *scene, const DevOut *out, uint32_t width, uint32_t height, uint32_t spa,
                        uint8_t *rgb, hipStream_t stream);

}  // namespace gg
