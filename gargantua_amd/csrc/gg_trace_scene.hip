// gg_trace_scene.hip -- trace kernel launches for the Hittable scene modes:
// DISK 3 = trace_ray_with_scene, DISK 4 = trace_ray_with_disk_and_scene
// (tracer.rs:207-364).  One scene per launch (BATCH 0), rays from the
// camera or from caller states, every precision.
#include <hip/hip_runtime.h>

#include "gg_trace_kernel.hpp"

namespace gg {

namespace {

template <int M, int D, int P>
hipError_t launch_scene3(const TraceLaunch &t, uint32_t grid, uint32_t block, hipStream_t s) {
  if (t.rec_pool || t.n_scenes != 1) return hipErrorInvalidValue;
  const size_t lds = lds_bytes(block);
  if (t.states_in)
    hipLaunchKernelGGL((trace_kernel<M, D, P, 1, 0>), dim3(grid), dim3(block), lds, s, t);
  else
    hipLaunchKernelGGL((trace_kernel<M, D, P, 0, 0>), dim3(grid), dim3(block), lds, s, t);
  return hipGetLastError();
}

template <int M, int D>
hipError_t launch_scene2(int prec, const TraceLaunch &t, uint32_t grid, uint32_t block, hipStream_t s) {
  if (prec == 1) return launch_scene3<M, D, 1>(t, grid, block, s);
  if (prec == 2) return launch_scene3<M, D, 2>(t, grid, block, s);
  return launch_scene3<M, D, 0>(t, grid, block, s);
}

template <int M>
hipError_t launch_scene1(int mode, int prec, const TraceLaunch &t, uint32_t grid, uint32_t block, hipStream_t s) {
  if (mode == 4) return launch_scene2<M, 4>(prec, t, grid, block, s);
  return launch_scene2<M, 3>(prec, t, grid, block, s);
}

}  // namespace

hipError_t launch_trace_scene(int metric_kind, int mode, int precision, const TraceLaunch &t, uint32_t grid,
                              uint32_t block, hipStream_t stream) {
  if (mode != 3 && mode != 4) return hipErrorInvalidValue;
  if (metric_kind == 1) return launch_scene1<1>(mode, precision, t, grid, block, stream);
  return launch_scene1<0>(mode, precision, t, grid, block, stream);
}

}  // namespace gg
