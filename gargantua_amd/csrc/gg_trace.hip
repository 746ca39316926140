// gg_trace.hip -- trace kernel launches for the disk modes (no disk,
// volumetric, thin) and the small kernels around them: counters, RHS,
// constants of motion, camera rays, ray-order keys and shading.  The
// persistent trace kernel itself is gg_trace_kernel.hpp.
#include <hip/hip_runtime.h>

// The disk-mode trace kernels fit 168 VGPRs (a few spill slots in the FP64
// tiers): three wavefronts per SIMD.
#define GG_MIN_WAVES 3
#include "gg_trace_kernel.hpp"

namespace gg {

namespace {

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

template <int METRIC>
__global__ void com_kernel(const DevScene *S, uint64_t n, const double *states, double *out) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double *st = states + 9 * i;
  const double pos[4] = {st[0], st[1], st[2], st[3]};
  double o[4];
  constants_of_motion<METRIC>(*S, pos, d4{st[4], st[5], st[6], st[7]}, o);
  for (int k = 0; k < 4; ++k) out[4 * i + k] = o[k];
}

// GeodesicRay::from_camera for every sub-pixel ray, row-major.
__global__ void camera_rays_kernel(const DevScene *S, double *states) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const uint64_t n = (uint64_t)S->sub_w * S->sub_h;
  if (i >= n) return;
  double pos[4];
  d4 v;
  camera_ray(*S, (uint32_t)(i % S->sub_w), (uint32_t)(i / S->sub_w), pos, v);
  double *o = states + 9 * i;
  o[0] = pos[0]; o[1] = pos[1]; o[2] = pos[2]; o[3] = pos[3];
  o[4] = v.x; o[5] = v.y; o[6] = v.z; o[7] = v.w; o[8] = 0.0;
}

// Sort keys of the learned ray order: the attempt count of global ray i,
// read from its scene's attempts output at its pixel.
__global__ void order_keys_kernel(const DevScene *scenes, const DevOut *outs, uint32_t n_scenes, uint64_t n,
                                  uint32_t *keys, uint32_t *ids) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const DevScene &S = scenes[find_scene(scenes, n_scenes, i)];
  uint32_t px, py;
  ray_pixel(S, i - S.ray_begin, px, py);
  keys[i] = outs[S.out_index].attempts[(uint64_t)py * S.sub_w + px];
  ids[i] = (uint32_t)i;
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

// HdriSky::sample + fetch (sky.rs:88-122): equirectangular map with row 0
// at the +z pole, theta = acos(z), phi = atan2(y, x), u = phi/(2 pi) + 1/2,
// v = theta/pi; bilinear in f32 with wrap-x / clamp-y, times intensity.
__device__ inline void sky_hdri(const SkyDesc &K, double d0, double d1, double d2, double rgb[3]) {
#pragma clang fp contract(off)
  const double FRAC_1_PI = 0.31830988618379067154;
  double z = d2 < -1.0 ? -1.0 : (d2 > 1.0 ? 1.0 : d2);
  const double theta = acos(z);
  const double phi = atan2(d1, d0);
  const double u = phi * FRAC_1_PI * 0.5 + 0.5;
  const double v = theta * FRAC_1_PI;
  const double w = (double)K.width, h = (double)K.height;
  double fx = fmod(u * w, w);                          // rem_euclid
  if (fx < 0.0) fx = fx + fabs(w);
  double fy = v * h;
  fy = fy < 0.0 ? 0.0 : (fy > h - 1.0 ? h - 1.0 : fy);
  const uint32_t x0 = (uint32_t)floor(fx) % K.width;
  const uint32_t y0 = (uint32_t)floor(fy);
  const uint32_t x1 = (x0 + 1) % K.width;
  const uint32_t y1 = y0 + 1 < K.height ? y0 + 1 : K.height - 1;
  const float tx = (float)(fx - floor(fx)), ty = (float)(fy - floor(fy));
  const float *p00 = K.pixels + 3 * ((uint64_t)y0 * K.width + x0), *p10 = K.pixels + 3 * ((uint64_t)y0 * K.width + x1);
  const float *p01 = K.pixels + 3 * ((uint64_t)y1 * K.width + x0), *p11 = K.pixels + 3 * ((uint64_t)y1 * K.width + x1);
  for (int c = 0; c < 3; ++c) {
    const float a = p00[c] + (p10[c] - p00[c]) * tx;
    const float b = p01[c] + (p11[c] - p01[c]) * tx;
    const float px = a + (b - a) * ty;
    rgb[c] = (double)(px * K.intensity);
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
// colours (shade_outcome_with, disk_color, a scene hit's colour) then
// encode_srgb.
__global__ void shade_kernel(const DevScene *Sp, const DevOut *Op, uint32_t width, uint32_t height, uint32_t spa,
                             SkyDesc skyd, uint8_t *rgb) {
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
        if (skyd.pixels) sky_hdri(skyd, d0, d1, d2, lin); else sky(d0, d1, d2, lin);
      } else if (c == GG_CLASS_DISK) {
        double I = O.intensity[k];
        double x = I / (1.0 + I);
        lin[0] = pow(x, 0.5); lin[1] = pow(x, 1.2); lin[2] = pow(x, 2.5);
      } else if (c == GG_CLASS_SCENE) {
        lin[0] = O.color[3 * k]; lin[1] = O.color[3 * k + 1]; lin[2] = O.color[3 * k + 2];
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
  if constexpr (M == 1 && D == 2 && (P == 0 || P == 1)) {
    // the recording render of the adjoint: Kerr thin disk, FP64, one scene
    if (t.rec_pool) {
      hipLaunchKernelGGL((trace_kernel<M, D, P, 2, 0>), dim3(grid), dim3(block), lds_bytes(block), s, t);
      return hipGetLastError();
    }
  }
  if (t.rec_pool) return hipErrorInvalidValue;
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

hipError_t launch_constants(int metric_kind, const DevScene *scene, uint64_t n, const double *states, double *out,
                            hipStream_t stream) {
  if (n == 0) return hipSuccess;
  const uint32_t grid = (uint32_t)((n + 255) / 256);
  if (metric_kind == 1)
    hipLaunchKernelGGL((com_kernel<1>), dim3(grid), dim3(256), 0, stream, scene, n, states, out);
  else
    hipLaunchKernelGGL((com_kernel<0>), dim3(grid), dim3(256), 0, stream, scene, n, states, out);
  return hipGetLastError();
}

hipError_t launch_camera_rays(const DevScene *scene, uint64_t n, double *states, hipStream_t stream) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(camera_rays_kernel, dim3((uint32_t)((n + 255) / 256)), dim3(256), 0, stream, scene, states);
  return hipGetLastError();
}

hipError_t launch_order_keys(const DevScene *scenes, const DevOut *outs, uint32_t n_scenes, uint64_t n,
                             uint32_t *keys, uint32_t *ids, hipStream_t stream) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(order_keys_kernel, dim3((uint32_t)((n + 255) / 256)), dim3(256), 0, stream, scenes, outs,
                     n_scenes, n, keys, ids);
  return hipGetLastError();
}

hipError_t launch_shade(const DevScene *scene, const DevOut *out, uint32_t width, uint32_t height, uint32_t spa,
                        const SkyDesc &sky, uint8_t *rgb, hipStream_t stream) {
  uint32_t n = width * height;
  hipLaunchKernelGGL(shade_kernel, dim3((n + 255) / 256), dim3(256), 0, stream, scene, out, width, height, spa, sky,
                     rgb);
  return hipGetLastError();
}

}  // namespace gg
