// gg_api.hip -- host side of libgargantua: the C ABI of include/gargantua.h.
//
// Validates every argument before anything reaches a kernel, turns a
// gg_scene into per-scene device constants (camera basis, metric at the
// camera, horizon, thresholds) with the host libm so those values match
// what the reference computes per ray, stages outputs that live in host
// memory, launches the persistent trace kernel and times it with hipEvents
// on the context's stream.
#include <hip/hip_runtime.h>
#include <hip/hiprtc.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "gargantua.h"
#include "gg_internal.hpp"
#include "gg_scene.hpp"
// kRtcHeaderNames / kRtcHeaderSources / kRtcHeaderCount: the text of the
// device headers, written at build time, for modules compiled at run time
#include "gg_rtc_headers.inc"

using namespace gg;

namespace {

thread_local std::string g_last_error;

int fail(int code, const char *fmt, ...) {
  char buf[2048];   // room for a hipRTC log excerpt of 1500 characters
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  g_last_error = buf;
  return code;
}

#define HIP_TRY(expr)                                                                        \
  do {                                                                                       \
    hipError_t e_ = (expr);                                                                  \
    if (e_ != hipSuccess) return fail(GG_ERR_DEVICE, "%s: %s", #expr, hipGetErrorString(e_)); \
  } while (0)

// A device buffer owned by the context, grown on demand, never shrunk.
struct DevBuf {
  void *p = nullptr;
  size_t cap = 0;
  int ensure(size_t bytes) {
    if (bytes <= cap) return GG_OK;
    if (p) (void)hipFree(p);
    p = nullptr;
    cap = 0;
    hipError_t e = hipMalloc(&p, bytes);
    if (e != hipSuccess) {
      p = nullptr;
      return fail(e == hipErrorOutOfMemory ? GG_ERR_OUT_OF_MEMORY : GG_ERR_DEVICE, "hipMalloc(%zu): %s", bytes,
                  hipGetErrorString(e));
    }
    cap = bytes;
    return GG_OK;
  }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    cap = 0;
  }
};

// ---- the scratch of gg_render_batch_noalloc ----------------------------------
// One static device array per GPU: the trace kernel's counters, the scenes'
// constants and their output tables, each at a 256-byte boundary.  In host
// code `scratch` names a host-side stand-in: the device address is taken with
// hipGetSymbolAddress once per device (scratch_address) and only that address
// is handed to copies and kernels.
constexpr size_t SCRATCH_COUNTERS = 0;
constexpr size_t SCRATCH_SCENES = 256;
constexpr size_t SCRATCH_OUTS = SCRATCH_SCENES + ((sizeof(DevScene) * GG_NOALLOC_MAX_SCENES + 255) / 256) * 256;
constexpr size_t SCRATCH_END = SCRATCH_OUTS + ((sizeof(DevOut) * GG_NOALLOC_MAX_SCENES + 255) / 256) * 256;
static_assert(sizeof(Counters) <= SCRATCH_SCENES, "counters fit their slot");
static_assert(SCRATCH_END <= GG_NOALLOC_SCRATCH_BYTES, "scratch layout fits the stated maximum");
static __device__ __align__(256) unsigned char scratch[GG_NOALLOC_SCRATCH_BYTES];

constexpr int SCRATCH_MAX_DEVICES = 64;
std::mutex g_scratch_mu;
void *g_scratch_addr[SCRATCH_MAX_DEVICES] = {};

// The device address of `scratch` on `device` (which must be current), taken
// once per device.
int scratch_address(int device, void **out) {
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (device < 0 || device >= SCRATCH_MAX_DEVICES)
    return fail(GG_ERR_UNSUPPORTED, "device %d beyond the %d the scratch table holds", device, SCRATCH_MAX_DEVICES);
  if (!g_scratch_addr[device]) {
    void *p = nullptr;
    hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      return fail(GG_ERR_DEVICE, "hipGetSymbolAddress(scratch) on device %d: %s", device, hipGetErrorString(e));
    g_scratch_addr[device] = p;
  }
  *out = g_scratch_addr[device];
  return GG_OK;
}

}  // namespace

struct gg_context {
  int device = 0;
  hipStream_t stream = nullptr;
  void *scratch = nullptr;   // this device's address of `scratch` (gg_render_batch_noalloc)
  bool own_stream = false;
  uint32_t num_cus = 256;
  gg_launch launch{64, 0, 64, 1, 0};   // blocks per CU 0: by ray count (blocks_per_cu_for)
  // cached batch graph
  hipGraph_t graph = nullptr;
  hipGraphExec_t graph_exec = nullptr;
  struct GraphKey {
    uint32_t n_scenes;
    int metric, disk, precision;
    uint64_t n_rays;
    gg_launch launch;
    void *scenes, *outs, *counters;
    const void *order;
  } graph_key{};
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  DevBuf scenes, outs, counters, states_in;
  DevBuf stage[10];      // device staging for host-memory outputs
  DevBuf rhs_io, misc_io;
  // learned ray order (gg_launch.order = 1): queue position -> ray of the
  // last camera render, the signature of its grids, the sort's work buffers
  DevBuf order, order_keys, order_ids, order_keys_out, order_att, order_tmp;
  uint64_t order_sig = 0;
  std::vector<DevBuf> batch_stage;
  // gg_grad: cls, intensity, accepted, target, resid, dI, list, counters,
  // results, scene, checkpoint pool, segment buffers, keys, sorted list,
  // sorted keys, ray records, last-segment records, sort scratch,
  DevBuf grad[22];
  // record pool, last chunk per pixel and final states of the recording
  // render (grad[18..20]), attempts per pixel (grad[21]); chunks the last
  // recording needed
  uint64_t grad_chunks = 0;
  uint64_t grad_budget = 0;   // gg_set_grad_memory (0: a quarter of the free memory)
  // gg_set_sky_hdri: equirectangular environment map for rgb outputs
  // (sky_w = 0: the procedural starfield)
  DevBuf sky;
  uint32_t sky_w = 0, sky_h = 0;
  float sky_intensity = 1.0f;
  // gg_set_hittables: primitives (BVH leaf order) and BVH nodes; hit_off
  // while gg_grad renders its own scene
  DevBuf hit_objs, hit_nodes;
  uint32_t n_hit = 0;
  bool hit_off = false;
  double hit_rmin = 0.0;      // distance from the origin to the nearest primitive box
  // gg_set_user_metric: the tensor's source, its horizon, and the trace
  // modules compiled from it, one per (disk mode 0-4, ray source 0-1)
  std::string user_src;
  bool user_has_rh = false;
  double user_rh = 0.0;
  hipModule_t user_mod[10] = {};
  hipFunction_t user_trace[10] = {}, user_at[10] = {};
  // the recording render gg_grad asks render_impl for (pool null: none)
  struct {
    double *pool;
    unsigned long long cap;
    unsigned long long *cursor;
    uint32_t *last;
  } rec_req{};
};

namespace {

int user_kernels(gg_context *ctx, int mode, int src, hipFunction_t *trace, hipFunction_t *at);
int compile_user_module(const std::string &user, int mode, int src, std::vector<char> &code, std::string &name_t,
                        std::string &name_a, std::string &log);

// ---- Hittable scene: primitive constants and the BVH ----

struct HitBox { double lo[3], hi[3]; };

void cross_h(const double a[3], const double b[3], double r[3]) {
  r[0] = a[1] * b[2] - a[2] * b[1];
  r[1] = a[2] * b[0] - a[0] * b[2];
  r[2] = a[0] * b[1] - a[1] * b[0];
}
double dot_h(const double a[3], const double b[3]) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }

// Sphere::new (sphere.rs:17-27) / Quad::with_arc_material (quad.rs:24-72)
// and the materials' scatter albedo and emission (lambertian.rs:41-56,
// metal.rs:13-18,40-57, dielectric.rs:65-90, diffuse_light.rs:22-35).
int make_scene_obj(const gg_hittable &h, uint32_t i, SceneObj &o, HitBox &b) {
  std::memset(&o, 0, sizeof(o));
  auto finite3 = [](const double *v) { return std::isfinite(v[0]) && std::isfinite(v[1]) && std::isfinite(v[2]); };
  if (!finite3(h.q) || !finite3(h.color)) return fail(GG_ERR_INVALID_ARG, "hittable %u: non-finite position or colour", i);
  switch (h.material) {
    case GG_MATERIAL_LAMBERTIAN:
      for (int k = 0; k < 3; ++k) o.albedo[k] = h.color[k];
      break;
    case GG_MATERIAL_METAL: {
      double fuzz = h.fuzz < 0.0 ? 0.0 : (h.fuzz > 1.0 ? 1.0 : h.fuzz);   // fuzz.clamp(0, 1)
      if (!(fuzz == 0.0))
        return fail(GG_ERR_UNSUPPORTED, "hittable %u: a metal with fuzz > 0 shades with a random draw", i);
      for (int k = 0; k < 3; ++k) o.albedo[k] = h.color[k];
      o.metal = 1;
      break;
    }
    case GG_MATERIAL_DIELECTRIC:
      o.albedo[0] = o.albedo[1] = o.albedo[2] = 1.0;
      break;
    case GG_MATERIAL_DIFFUSE_LIGHT:
      for (int k = 0; k < 3; ++k) o.emit[k] = h.color[k];
      break;
    default:
      return fail(GG_ERR_INVALID_ARG, "hittable %u: material %d unknown", i, h.material);
  }
  if (h.shape == GG_SHAPE_SPHERE) {
    if (!std::isfinite(h.radius) || h.radius == 0.0) return fail(GG_ERR_INVALID_ARG, "hittable %u: sphere radius", i);
    o.shape = SHAPE_SPHERE;
    for (int k = 0; k < 3; ++k) o.a[k] = h.q[k];
    o.n[0] = h.radius;
    const double r = std::fabs(h.radius);
    for (int k = 0; k < 3; ++k) { b.lo[k] = h.q[k] - r; b.hi[k] = h.q[k] + r; }
  } else if (h.shape == GG_SHAPE_QUAD) {
    if (!finite3(h.u) || !finite3(h.v)) return fail(GG_ERR_INVALID_ARG, "hittable %u: non-finite quad edges", i);
    double nn[3];
    cross_h(h.u, h.v, nn);
    const double n2 = dot_h(nn, nn);
    if (!(n2 > 0.0) || !std::isfinite(n2)) return fail(GG_ERR_INVALID_ARG, "hittable %u: degenerate quad", i);
    const double len = std::sqrt(n2);
    o.shape = SHAPE_QUAD;
    for (int k = 0; k < 3; ++k) { o.a[k] = h.q[k]; o.n[k] = nn[k] / len; }
    o.d = dot_h(o.n, h.q);
    double w[3] = {nn[0] / n2, nn[1] / n2, nn[2] / n2};
    cross_h(h.v, w, o.ea);
    cross_h(w, h.u, o.eb);
    for (int k = 0; k < 3; ++k) {
      const double c[4] = {h.q[k], h.q[k] + h.u[k], h.q[k] + h.v[k], h.q[k] + h.u[k] + h.v[k]};
      b.lo[k] = std::min(std::min(c[0], c[1]), std::min(c[2], c[3])) - 0.0001;
      b.hi[k] = std::max(std::max(c[0], c[1]), std::max(c[2], c[3])) + 0.0001;
    }
  } else {
    return fail(GG_ERR_INVALID_ARG, "hittable %u: shape %d unknown", i, h.shape);
  }
  // a margin far above rounding, so that the slab test never culls a
  // primitive the exact test hits
  for (int k = 0; k < 3; ++k) {
    const double m = 1e-9 * (1.0 + std::max(std::fabs(b.lo[k]), std::fabs(b.hi[k])));
    b.lo[k] -= m;
    b.hi[k] += m;
  }
  return GG_OK;
}

// Median split on the longest axis of the centroid bounds, leaves of at
// most four primitives, nodes in depth-first order (node 0 the root, an
// internal node's first child right after it, skip = the end of its
// subtree); order[] becomes the leaf order.
int32_t build_bvh(const std::vector<HitBox> &box, std::vector<uint32_t> &order, uint32_t b, uint32_t e,
                  std::vector<BvhNode> &nodes) {
  const int32_t id = (int32_t)nodes.size();
  nodes.push_back(BvhNode{});
  BvhNode nd{};
  double lo[3], hi[3], clo[3], chi[3];
  for (int k = 0; k < 3; ++k) { lo[k] = clo[k] = INFINITY; hi[k] = chi[k] = -INFINITY; }
  for (uint32_t i = b; i < e; ++i) {
    const HitBox &x = box[order[i]];
    for (int k = 0; k < 3; ++k) {
      lo[k] = std::min(lo[k], x.lo[k]);
      hi[k] = std::max(hi[k], x.hi[k]);
      const double c = 0.5 * (x.lo[k] + x.hi[k]);
      clo[k] = std::min(clo[k], c);
      chi[k] = std::max(chi[k], c);
    }
  }
  // the FP32 box, rounded outward and padded far beyond FP32 rounding
  for (int k = 0; k < 3; ++k) {
    const double pad = 1e-5 * (1.0 + std::max(std::fabs(lo[k]), std::fabs(hi[k])));
    nd.lo[k] = std::nextafter((float)(lo[k] - pad), -INFINITY);
    nd.hi[k] = std::nextafter((float)(hi[k] + pad), INFINITY);
  }
  if (e - b <= 4) {
    nd.leaf = (int32_t)(8 * b + (e - b));
  } else {
    int axis = 0;
    for (int k = 1; k < 3; ++k)
      if (chi[k] - clo[k] > chi[axis] - clo[axis]) axis = k;
    const uint32_t mid = b + (e - b) / 2;
    std::nth_element(order.begin() + b, order.begin() + mid, order.begin() + e, [&](uint32_t p, uint32_t q) {
      return box[p].lo[axis] + box[p].hi[axis] < box[q].lo[axis] + box[q].hi[axis];
    });
    nd.leaf = 0;
    build_bvh(box, order, b, mid, nodes);
    build_bvh(box, order, mid, e, nodes);
  }
  nd.skip = (int32_t)nodes.size();
  nodes[id] = nd;
  return id;
}

// ---- host restatements needed for per-scene constants ----

// Kerr::metric_tensor / Schwarzschild::metric_tensor at the camera
// (kerr.rs:61-94, schwarzschild.rs:26-43).
void metric_host(int kind, double m, double a, double rs, const double pos[4], double &tt, double &rr,
                 double &thth, double &phph, double &tph) {
  double r = pos[1] >= 1e-10 ? pos[1] : 1e-10;
  double theta = pos[2];
  if (kind == GG_METRIC_KERR) {
    double a2 = a * a;
    double ct = std::cos(theta);
    double sigma = r * r + a * a * ct * ct;
    double delta = r * r - 2.0 * m * r + a * a;
    double st = std::sin(theta);
    double sin2 = st * st;
    tt = -(1.0 - 2.0 * m * r / sigma);
    rr = sigma / delta;
    thth = sigma;
    phph = (r * r + a2 + 2.0 * m * a2 * r * sin2 / sigma) * sin2;
    tph = -2.0 * m * a * r * sin2 / sigma;
  } else {
    double f = 1.0 - rs / r;
    double st = std::sin(theta);
    tt = -f;
    rr = 1.0 / f;
    thth = r * r;
    phph = r * r * st * st;
    tph = 0.0;
  }
}

void normalize3(double v[3]) {
  double n = std::sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
  v[0] = v[0] / n; v[1] = v[1] / n; v[2] = v[2] / n;
}
void cross3(const double a[3], const double b[3], double r[3]) {
  r[0] = a[1] * b[2] - a[2] * b[1];
  r[1] = a[2] * b[0] - a[0] * b[2];
  r[2] = a[0] * b[1] - a[1] * b[0];
}

bool finite(double x) { return std::isfinite(x); }

int validate_metric(const gg_metric &m, bool allow_user = false) {
  if (m.kind == GG_METRIC_USER && allow_user) {
    if (!finite(m.mass) || !finite(m.spin)) return fail(GG_ERR_INVALID_ARG, "metric.mass and spin must be finite");
    return GG_OK;
  }
  if (m.kind != GG_METRIC_SCHWARZSCHILD && m.kind != GG_METRIC_KERR)
    return fail(GG_ERR_INVALID_ARG, "metric.kind %d is not Schwarzschild (0) or Kerr (1)%s", m.kind,
                m.kind == GG_METRIC_USER ? " (a user metric renders through gg_render / gg_trace_states)" : "");
  if (!(m.mass > 0.0) || !finite(m.mass)) return fail(GG_ERR_INVALID_ARG, "metric.mass must be positive and finite");
  if (!finite(m.spin)) return fail(GG_ERR_INVALID_ARG, "metric.spin must be finite");
  return GG_OK;
}

int validate_integrator(const gg_integrator &I) {
  if (!(I.initial_step > 0.0) || !finite(I.initial_step)) return fail(GG_ERR_INVALID_ARG, "initial_step must be > 0");
  if (!(I.min_step > 0.0) || !finite(I.min_step)) return fail(GG_ERR_INVALID_ARG, "min_step must be > 0");
  if (!(I.max_step >= I.min_step) || !finite(I.max_step))
    return fail(GG_ERR_INVALID_ARG, "max_step must be >= min_step (f64::clamp panics otherwise)");
  if (I.err_mode == GG_ERR_REFERENCE) {
    if (!(I.tolerance > 0.0) || !finite(I.tolerance)) return fail(GG_ERR_INVALID_ARG, "tolerance must be > 0");
  } else if (I.err_mode == GG_ERR_RTOL_ATOL) {
    if (!(I.rtol >= 0.0) || !(I.atol >= 0.0) || !(I.rtol + I.atol > 0.0) || !finite(I.rtol) || !finite(I.atol))
      return fail(GG_ERR_INVALID_ARG, "rtol, atol must be >= 0 and not both 0");
  } else {
    return fail(GG_ERR_INVALID_ARG, "err_mode %d unknown", I.err_mode);
  }
  if (std::isnan(I.max_radius)) return fail(GG_ERR_INVALID_ARG, "max_radius is NaN");
  if (!finite(I.horizon_buffer) || !finite(I.horizon_rel))
    return fail(GG_ERR_INVALID_ARG, "horizon_buffer/horizon_rel must be finite");
  return GG_OK;
}

int validate_disk(const gg_disk &d) {
  if (d.kind < GG_DISK_NONE || d.kind > GG_DISK_THIN) return fail(GG_ERR_INVALID_ARG, "disk.kind %d unknown", d.kind);
  if (d.kind == GG_DISK_NONE) return GG_OK;
  if (!finite(d.r_inner) || !finite(d.r_outer)) return fail(GG_ERR_INVALID_ARG, "disk radii must be finite");
  if (d.kind == GG_DISK_VOLUMETRIC) {
    if (!(d.scale_height > 0.0)) return fail(GG_ERR_INVALID_ARG, "disk.scale_height must be > 0");
    if (!finite(d.density_0) || !finite(d.absorption)) return fail(GG_ERR_INVALID_ARG, "disk density/absorption");
  } else {
    if (!finite(d.emissivity_q)) return fail(GG_ERR_INVALID_ARG, "disk.emissivity_q must be finite");
    if (!(d.edge_width >= 0.0) || !finite(d.edge_width)) return fail(GG_ERR_INVALID_ARG, "disk.edge_width must be >= 0");
  }
  return GG_OK;
}

int validate_scene(const gg_scene *s, bool need_image) {
  if (!s) return fail(GG_ERR_INVALID_ARG, "scene is NULL");
  int rc;
  if ((rc = validate_metric(s->metric, true))) return rc;
  if ((rc = validate_integrator(s->integrator))) return rc;
  if ((rc = validate_disk(s->disk))) return rc;
  for (int i = 0; i < 4; ++i)
    if (!finite(s->camera.position[i])) return fail(GG_ERR_INVALID_ARG, "camera.position must be finite");
  if (need_image) {
    const gg_image &im = s->image;
    if (im.width == 0 || im.height == 0) return fail(GG_ERR_INVALID_ARG, "image width/height must be > 0");
    if (im.samples_per_axis == 0) return fail(GG_ERR_INVALID_ARG, "samples_per_axis must be >= 1");
    uint64_t sw = (uint64_t)im.width * im.samples_per_axis, sh = (uint64_t)im.height * im.samples_per_axis;
    if (sw > 65536 || sh > 65536) return fail(GG_ERR_INVALID_ARG, "sub-pixel grid %llux%llu too large",
                                              (unsigned long long)sw, (unsigned long long)sh);
    if (!(s->camera.fov_y_radians > 0.0) || !(s->camera.fov_y_radians < M_PI))
      return fail(GG_ERR_INVALID_ARG, "camera.fov_y_radians must be in (0, pi)");
    if (!(s->camera.aspect > 0.0) || !finite(s->camera.aspect))
      return fail(GG_ERR_INVALID_ARG, "camera.aspect must be > 0");
    double f[3] = {s->camera.look_at[0], s->camera.look_at[1], s->camera.look_at[2]};
    double u[3] = {s->camera.up[0], s->camera.up[1], s->camera.up[2]};
    double c[3];
    cross3(f, u, c);
    if (!(c[0] * c[0] + c[1] * c[1] + c[2] * c[2] > 0.0))
      return fail(GG_ERR_INVALID_ARG, "camera.look_at must be non-zero and not parallel to camera.up");
  }
  return GG_OK;
}

// Kerr::new clamp (kerr.rs:29-34).
double clamp_spin(const gg_metric &m) {
  if (m.kind != GG_METRIC_KERR) return 0.0;
  double mx = m.mass * 0.9999;
  double s = m.spin;
  if (s < -mx) s = -mx;
  if (s > mx) s = mx;
  return s;
}

// The parts of a GG_METRIC_USER scene the host cannot compute itself: the
// tensor at the camera (evaluated on the device) and the horizon.
struct UserMetricAt {
  double g[16];
  bool has_rh;
  double rh;
};

void build_dev_scene(const gg_scene &s, DevScene &D, const UserMetricAt *um = nullptr) {
  std::memset(&D, 0, sizeof(D));
  D.metric_kind = s.metric.kind;
  D.disk_kind = s.disk.kind;
  D.err_mode = s.integrator.err_mode;
  D.max_steps = s.integrator.max_steps;
  D.mass = s.metric.mass;
  D.spin = clamp_spin(s.metric);
  D.rs = 2.0 * s.metric.mass;
  if (s.metric.kind == GG_METRIC_USER && um) {
    // Metric::event_horizon: None disables both horizon tests (geodesic.rs:198-206)
    D.spin = s.metric.spin;
    D.r_h = um->has_rh ? um->rh : -INFINITY;
  } else if (s.metric.kind == GG_METRIC_KERR) {
    double disc = D.mass * D.mass - D.spin * D.spin;  // kerr.rs:427-436 (disc > 0 after the clamp)
    D.r_h = D.mass + std::sqrt(disc);
  } else {
    D.r_h = D.rs;
  }
  D.r_capture = D.r_h * (1.0 + s.integrator.horizon_rel) + s.integrator.horizon_buffer;
  if (!(D.r_h > -INFINITY)) D.r_capture = -INFINITY;
  D.max_radius = s.integrator.max_radius;
  D.h0 = s.integrator.initial_step;
  D.min_step = s.integrator.min_step;
  D.max_step = s.integrator.max_step;
  D.max_step5 = D.max_step * D.max_step * D.max_step * D.max_step * D.max_step;
  D.tol = s.integrator.tolerance;
  D.rtol = s.integrator.rtol;
  D.atol = s.integrator.atol;
  // camera basis (camera.rs:21-26) and spherical frame at the camera (ray.rs:29-35)
  const gg_camera &c = s.camera;
  for (int i = 0; i < 4; ++i) D.cam_pos[i] = c.position[i];
  double fwd[3] = {c.look_at[0], c.look_at[1], c.look_at[2]};
  normalize3(fwd);
  double right[3];
  cross3(fwd, c.up, right);
  normalize3(right);
  double up[3];
  cross3(right, fwd, up);
  normalize3(up);
  for (int i = 0; i < 3; ++i) { D.fwd[i] = fwd[i]; D.right[i] = right[i]; D.up[i] = up[i]; }
  D.half_h = std::tan(c.fov_y_radians / 2.0);
  D.half_w = D.half_h * c.aspect;
  double th = c.position[2], ph = c.position[3];
  double st = std::sin(th), ct = std::cos(th), sp = std::sin(ph), cp = std::cos(ph);
  D.rhat[0] = st * cp; D.rhat[1] = st * sp; D.rhat[2] = ct;
  D.thhat[0] = ct * cp; D.thhat[1] = ct * sp; D.thhat[2] = -st;
  D.phhat[0] = -sp; D.phhat[1] = cp; D.phhat[2] = 0.0;
  double tph;
  if (s.metric.kind == GG_METRIC_USER && um) {
    D.g_tt = um->g[0]; D.g_rr = um->g[5]; D.g_thth = um->g[10]; D.g_phph = um->g[15];
  } else {
    metric_host(s.metric.kind, D.mass, D.spin, D.rs, c.position, D.g_tt, D.g_rr, D.g_thth, D.g_phph, tph);
  }
  D.sqrt_g_rr = std::sqrt(D.g_rr);
  D.sqrt_g_thth = std::sqrt(D.g_thth);
  D.sqrt_g_phph = std::sqrt(D.g_phph);
  double g_tt_obs = -D.g_tt;                                 // tracer.rs:93-94
  D.sqrt_neg_gtt_obs = std::sqrt(g_tt_obs >= 0.0 ? g_tt_obs : 0.0);
  D.g_tt_obs = g_tt_obs;                                     // tracer.rs:216
  uint32_t spa = s.image.samples_per_axis ? s.image.samples_per_axis : 1;
  D.sub_w = s.image.width * spa;
  D.sub_h = s.image.height * spa;
  D.jitter = s.image.jitter ? 1 : 0;
  D.seed = s.image.seed;
  D.tiled = (D.sub_w % 8 == 0 && D.sub_h % 8 == 0) ? 1 : 0;
  D.r_in = s.disk.r_inner;
  D.r_out = s.disk.r_outer;
  D.scale_height = s.disk.scale_height;
  D.density_0 = s.disk.density_0;
  D.absorption = s.disk.absorption;
  D.q = s.disk.emissivity_q;
  D.edge_width = s.disk.edge_width;
  D.fp32_min_r = 10.0 * D.mass;
  D.fp32_far_r = 300.0 * D.mass;
  D.fp32_b2_min = 100.0 * D.mass * D.mass;   // b >= 10 M: periapsis beyond ~8 M for any spin
}

bool is_device_ptr(const void *p) {
  if (!p) return false;
  hipPointerAttribute_t at;
  hipError_t e = hipPointerGetAttributes(&at, p);
  if (e != hipSuccess) {
    (void)hipGetLastError();
    return false;
  }
  return at.type == hipMemoryTypeDevice || at.type == hipMemoryTypeManaged;
}

// Resolve one output pointer: device pointers are used in place, host
// pointers get a staging buffer copied back after the kernel.
struct OutPlan {
  void *host = nullptr;   // non-null: copy back to here
  void *dev = nullptr;
  size_t bytes = 0;
};

int plan_out(DevBuf &stage, void *user, size_t bytes, OutPlan &p) {
  p = OutPlan{};
  if (!user) return GG_OK;
  p.bytes = bytes;
  if (is_device_ptr(user)) {
    p.dev = user;
    return GG_OK;
  }
  int rc = stage.ensure(bytes);
  if (rc) return rc;
  p.dev = stage.p;
  p.host = user;
  return GG_OK;
}

// Blocks per CU of a launch over n_rays of a kernel built for at most
// max_waves wavefronts per SIMD (3: the disk modes; 2: the Hittable scene
// modes and the recording renders; 1: run-time user metrics).
// blocks_per_cu = 0 picks by the number of rays: max_waves wavefronts per
// SIMD when every lane at two per SIMD gets at least 3 rays, one per SIMD
// otherwise.  With a few rays per lane a render is bounded by its longest
// rays, which a wavefront alone on its SIMD advances faster (512^2 Kerr
// renders and the differentiable render's recording pass: 4-25% shorter at
// one wavefront per SIMD; 724^2 and larger: longer).
uint64_t blocks_per_cu_for(const gg_context *ctx, uint64_t n_rays, uint32_t max_waves) {
  if (ctx->launch.blocks_per_cu) return ctx->launch.blocks_per_cu;
  const uint64_t per_block = ctx->launch.block_threads / 64;
  const uint64_t lanes_two = (uint64_t)ctx->num_cus * 8 * 64;
  const uint64_t waves_per_cu = n_rays >= 3 * lanes_two ? 4 * (uint64_t)max_waves : 4;
  return waves_per_cu > per_block ? waves_per_cu / per_block : 1;
}

uint32_t grid_for(gg_context *ctx, uint64_t n_rays, uint32_t max_waves) {
  uint64_t waves_needed = (n_rays + 63) / 64;
  uint64_t per_block = ctx->launch.block_threads / 64;
  uint64_t blocks_needed = (waves_needed + per_block - 1) / per_block;
  uint64_t grid = (uint64_t)ctx->num_cus * blocks_per_cu_for(ctx, n_rays, max_waves);
  if (blocks_needed < grid) grid = blocks_needed;
  if (grid < 1) grid = 1;
  return (uint32_t)grid;
}

}  // namespace

extern "C" {

const char *gg_strerror(int status) {
  switch (status) {
    case GG_OK: return "ok";
    case GG_ERR_INVALID_ARG: return "invalid argument";
    case GG_ERR_DEVICE: return "HIP runtime error";
    case GG_ERR_NO_DEVICE: return "no HIP device";
    case GG_ERR_OUT_OF_MEMORY: return "out of device memory";
    case GG_ERR_UNSUPPORTED: return "unsupported";
    case GG_ERR_INTERNAL: return "internal error";
    default: return "unknown status";
  }
}

const char *gg_last_error(void) { return g_last_error.c_str(); }

int gg_abi_version(void) { return GG_ABI_VERSION; }

int gg_context_create(int device, void *stream, gg_context **out) {
  if (!out) return fail(GG_ERR_INVALID_ARG, "out is NULL");
  *out = nullptr;
  int ndev = 0;
  hipError_t e = hipGetDeviceCount(&ndev);
  if (e != hipSuccess || ndev == 0) return fail(GG_ERR_NO_DEVICE, "no HIP device (%s)", hipGetErrorString(e));
  if (device < 0 || device >= ndev) return fail(GG_ERR_INVALID_ARG, "device %d out of range [0, %d)", device, ndev);
  HIP_TRY(hipSetDevice(device));
  gg_context *c = new gg_context();
  c->device = device;
  int cus = 0;
  if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device) == hipSuccess && cus > 0)
    c->num_cus = (uint32_t)cus;
  if (stream) {
    c->stream = (hipStream_t)stream;
  } else {
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
      delete c;
      return fail(GG_ERR_DEVICE, "hipStreamCreate failed");
    }
    c->own_stream = true;
  }
  if (hipEventCreate(&c->ev0) != hipSuccess || hipEventCreate(&c->ev1) != hipSuccess) {
    delete c;
    return fail(GG_ERR_DEVICE, "hipEventCreate failed");
  }
  int rc = c->counters.ensure(sizeof(Counters));
  if (rc) { delete c; return rc; }
  if ((rc = scratch_address(device, &c->scratch))) { delete c; return rc; }
  *out = c;
  return GG_OK;
}

int gg_context_destroy(gg_context *ctx) {
  if (!ctx) return GG_OK;
  (void)hipSetDevice(ctx->device);
  (void)hipStreamSynchronize(ctx->stream);
  ctx->scenes.release(); ctx->outs.release(); ctx->counters.release(); ctx->states_in.release();
  ctx->rhs_io.release();
  ctx->misc_io.release();
  ctx->order.release(); ctx->order_keys.release(); ctx->order_ids.release(); ctx->order_keys_out.release();
  ctx->order_att.release(); ctx->order_tmp.release();
  for (auto &b : ctx->grad) b.release();
  for (auto &b : ctx->stage) b.release();
  for (auto &b : ctx->batch_stage) b.release();
  ctx->hit_objs.release(); ctx->hit_nodes.release(); ctx->sky.release();
  for (auto &m : ctx->user_mod)
    if (m) (void)hipModuleUnload(m);
  if (ctx->graph_exec) (void)hipGraphExecDestroy(ctx->graph_exec);
  if (ctx->graph) (void)hipGraphDestroy(ctx->graph);
  if (ctx->ev0) (void)hipEventDestroy(ctx->ev0);
  if (ctx->ev1) (void)hipEventDestroy(ctx->ev1);
  if (ctx->own_stream) (void)hipStreamDestroy(ctx->stream);
  delete ctx;
  return GG_OK;
}

int gg_set_launch(gg_context *ctx, const gg_launch *l) {
  if (!ctx || !l) return fail(GG_ERR_INVALID_ARG, "NULL argument");
  gg_launch v = *l;
  if (v.block_threads == 0) v.block_threads = 64;
  if (v.refill_interval == 0) v.refill_interval = 64;
  if (v.graph == 0) v.graph = 1;
  if (v.graph > 2) return fail(GG_ERR_INVALID_ARG, "graph must be 0, 1 or 2");
  if (v.order > 1) return fail(GG_ERR_INVALID_ARG, "order must be 0 or 1");
  if (v.block_threads % 64 || v.block_threads > 256)
    return fail(GG_ERR_INVALID_ARG, "block_threads must be a multiple of 64 in [64, 256]");
  if (v.blocks_per_cu > 64) return fail(GG_ERR_INVALID_ARG, "blocks_per_cu must be <= 64");
  ctx->launch = v;
  return GG_OK;
}

int gg_set_sky_hdri(gg_context *ctx, uint32_t width, uint32_t height, const float *pixels, float intensity) {
  if (!ctx) return fail(GG_ERR_INVALID_ARG, "NULL context");
  if (width == 0) {
    ctx->sky_w = ctx->sky_h = 0;
    return GG_OK;
  }
  if (height == 0 || !pixels) return fail(GG_ERR_INVALID_ARG, "HDRI sky needs height > 0 and pixels");
  if ((uint64_t)width * height > (1ull << 28)) return fail(GG_ERR_INVALID_ARG, "HDRI sky larger than 2^28 texels");
  if (!std::isfinite(intensity)) return fail(GG_ERR_INVALID_ARG, "HDRI sky intensity must be finite");
  HIP_TRY(hipSetDevice(ctx->device));
  // the buffer may be reallocated below: until the copy is done the context
  // has no sky, so a failure leaves the procedural one
  ctx->sky_w = ctx->sky_h = 0;
  const size_t bytes = 12ull * width * height;
  int rc = ctx->sky.ensure(bytes);
  if (rc) return rc;
  HIP_TRY(hipMemcpyAsync(ctx->sky.p, pixels, bytes, hipMemcpyDefault, ctx->stream));
  HIP_TRY(hipStreamSynchronize(ctx->stream));
  ctx->sky_w = width;
  ctx->sky_h = height;
  ctx->sky_intensity = intensity;
  return GG_OK;
}

int gg_set_hittables(gg_context *ctx, uint32_t n, const gg_hittable *objects) {
  if (!ctx) return fail(GG_ERR_INVALID_ARG, "NULL context");
  if (n == 0) {
    ctx->n_hit = 0;
    return GG_OK;
  }
  if (!objects) return fail(GG_ERR_INVALID_ARG, "hittables: NULL objects");
  if (n > (1u << 24)) return fail(GG_ERR_INVALID_ARG, "hittables: more than 2^24 primitives");
  std::vector<SceneObj> prim(n);
  std::vector<HitBox> boxes(n);
  for (uint32_t i = 0; i < n; ++i) {
    int rc = make_scene_obj(objects[i], i, prim[i], boxes[i]);
    if (rc) return rc;
  }
  std::vector<BvhNode> nodes;
  std::vector<uint32_t> order(n);
  for (uint32_t i = 0; i < n; ++i) order[i] = i;
  nodes.reserve(2 * (size_t)n);
  build_bvh(boxes, order, 0, n, nodes);
  // the nearest point of any primitive's (padded) box to the origin
  double rmin = INFINITY;
  for (const HitBox &b : boxes) {
    double d2 = 0.0;
    for (int k = 0; k < 3; ++k) {
      const double q = b.lo[k] > 0.0 ? b.lo[k] : (b.hi[k] < 0.0 ? -b.hi[k] : 0.0);
      d2 += q * q;
    }
    rmin = std::min(rmin, std::sqrt(d2) * (1.0 - 1e-9));
  }
  std::vector<SceneObj> leaf_order(n);
  for (uint32_t i = 0; i < n; ++i) leaf_order[i] = prim[order[i]];
  HIP_TRY(hipSetDevice(ctx->device));
  // the buffers may be reallocated below: until both copies are done the
  // context has no hittables, so a failure leaves none
  ctx->n_hit = 0;
  int rc;
  if ((rc = ctx->hit_objs.ensure(sizeof(SceneObj) * n))) return rc;
  if ((rc = ctx->hit_nodes.ensure(sizeof(BvhNode) * nodes.size()))) return rc;
  HIP_TRY(hipMemcpyAsync(ctx->hit_objs.p, leaf_order.data(), sizeof(SceneObj) * n, hipMemcpyHostToDevice,
                         ctx->stream));
  HIP_TRY(hipMemcpyAsync(ctx->hit_nodes.p, nodes.data(), sizeof(BvhNode) * nodes.size(), hipMemcpyHostToDevice,
                         ctx->stream));
  HIP_TRY(hipStreamSynchronize(ctx->stream));
  ctx->n_hit = n;
  ctx->hit_rmin = rmin;
  return GG_OK;
}

int gg_check_user_metric(const char *source, int32_t disk_mode, char *log, size_t log_size) {
  if (!source || !*source) return fail(GG_ERR_INVALID_ARG, "user metric source is empty");
  if (disk_mode < 0 || disk_mode > 4) return fail(GG_ERR_INVALID_ARG, "disk_mode must be 0..4");
  std::vector<char> code;
  std::string name_t, name_a, text;
  const int rc = compile_user_module(source, disk_mode, 0, code, name_t, name_a, text);
  if (log && log_size) {
    std::snprintf(log, log_size, "%s", text.c_str());
  }
  return rc;
}

int gg_set_user_metric(gg_context *ctx, const char *source, double event_horizon) {
  if (!ctx) return fail(GG_ERR_INVALID_ARG, "NULL context");
  HIP_TRY(hipSetDevice(ctx->device));
  HIP_TRY(hipStreamSynchronize(ctx->stream));
  for (int k = 0; k < 10; ++k) {
    if (ctx->user_mod[k]) (void)hipModuleUnload(ctx->user_mod[k]);
    ctx->user_mod[k] = nullptr;
    ctx->user_trace[k] = ctx->user_at[k] = nullptr;
  }
  ctx->user_src.clear();
  if (!source) return GG_OK;
  if (!*source) return fail(GG_ERR_INVALID_ARG, "user metric source is empty");
  if (!(std::isnan(event_horizon) || std::isfinite(event_horizon)))
    return fail(GG_ERR_INVALID_ARG, "event_horizon must be finite or NaN (None)");
  ctx->user_src = source;
  ctx->user_has_rh = !std::isnan(event_horizon);
  ctx->user_rh = ctx->user_has_rh ? event_horizon : 0.0;
  // compile the camera-ray, no-disk kernel now, so that a bad source fails here
  hipFunction_t t = nullptr, a = nullptr;
  int rc = user_kernels(ctx, 0, 0, &t, &a);
  if (rc) ctx->user_src.clear();
  return rc;
}

int gg_set_grad_memory(gg_context *ctx, uint64_t bytes) {
  if (!ctx) return fail(GG_ERR_INVALID_ARG, "NULL context");
  ctx->grad_budget = bytes;
  return GG_OK;
}

int gg_get_launch(gg_context *ctx, gg_launch *l) {
  if (!ctx || !l) return fail(GG_ERR_INVALID_ARG, "NULL argument");
  *l = ctx->launch;
  return GG_OK;
}

int gg_scene_rays(const gg_scene *scene, uint64_t *n_rays) {
  int rc = validate_scene(scene, true);
  if (rc) return rc;
  uint64_t spa = scene->image.samples_per_axis;
  if (n_rays) *n_rays = (uint64_t)scene->image.width * spa * scene->image.height * spa;
  return GG_OK;
}

}  // extern "C"

namespace {

// Compile the trace kernel for METRIC 2 with tensor source `user`, disk mode
// `mode` and ray source `src`, and the helper that evaluates the tensor at a
// point, with hipRTC (no device needed): the code object and both kernels'
// lowered names, or the compiler log.
int compile_user_module(const std::string &user, int mode, int src, std::vector<char> &code, std::string &name_t,
                        std::string &name_a, std::string &log) {
  // The source comes first, so that the symmetry macros it may define are
  // seen by the headers' numerical Christoffels.
  const std::string prog = "using namespace __hip_internal;\nnamespace gg {\n" + user +
                           "\n}  // namespace gg\n#include \"gg_trace_kernel.hpp\"\nnamespace gg {\n"
                           "__global__ void gg_user_metric_at(const double *x, double mass, double spin, double *g) {\n"
                           "  gg_user_metric(x, mass, spin, g);\n}\n}  // namespace gg\n";
  hiprtcProgram p;
  if (hiprtcCreateProgram(&p, prog.c_str(), "gg_user_metric.hip", kRtcHeaderCount, kRtcHeaderSources,
                          kRtcHeaderNames) != HIPRTC_SUCCESS)
    return fail(GG_ERR_INTERNAL, "hiprtcCreateProgram failed");
  char expr[96];
  std::snprintf(expr, sizeof expr, "gg::trace_kernel<2, %d, 0, %d, 0>", mode, src);
  const char *at_expr = "gg::gg_user_metric_at";
  hiprtcAddNameExpression(p, expr);
  hiprtcAddNameExpression(p, at_expr);
  // The library's own flags (gfx950, no FMA contraction, no packed FP32, no
  // machine LICM), and one wave per SIMD: the nine inlined tensor evaluations
  // of each RHS then keep their derivatives in VGPRs and AGPRs instead of
  // scratch memory.
  const char *opts[] = {"--offload-arch=gfx950", "-O3", "-std=c++17", "-ffp-contract=off",
                        "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
                        "-mllvm", "-disable-machine-licm", "-DGG_MIN_WAVES=1"};
  const hiprtcResult r = hiprtcCompileProgram(p, (int)(sizeof opts / sizeof opts[0]), opts);
  size_t n = 0;
  hiprtcGetProgramLogSize(p, &n);
  log.assign(n + 1, '\0');
  hiprtcGetProgramLog(p, &log[0]);
  if (r != HIPRTC_SUCCESS) {
    hiprtcDestroyProgram(&p);
    return fail(GG_ERR_INVALID_ARG, "user metric does not compile: %.1500s", log.c_str());
  }
  const char *lt = nullptr, *la = nullptr;
  hiprtcGetLoweredName(p, expr, &lt);
  hiprtcGetLoweredName(p, at_expr, &la);
  name_t = lt ? lt : "";
  name_a = la ? la : "";
  size_t code_n = 0;
  hiprtcGetCodeSize(p, &code_n);
  code.assign(code_n, 0);
  if (code_n) hiprtcGetCode(p, code.data());
  hiprtcDestroyProgram(&p);
  if (name_t.empty() || name_a.empty() || code_n == 0) return fail(GG_ERR_INTERNAL, "hipRTC returned no kernel");
  return GG_OK;
}

// The context's METRIC 2 kernels for (mode, src), compiled and loaded on
// first use.
int user_kernels(gg_context *ctx, int mode, int src, hipFunction_t *trace, hipFunction_t *at) {
  const int key = 2 * mode + src;
  if (ctx->user_trace[key]) {
    *trace = ctx->user_trace[key];
    *at = ctx->user_at[key];
    return GG_OK;
  }
  std::vector<char> code;
  std::string name_t, name_a, log;
  int rc = compile_user_module(ctx->user_src, mode, src, code, name_t, name_a, log);
  if (rc) return rc;
  hipModule_t mod = nullptr;
  HIP_TRY(hipModuleLoadData(&mod, code.data()));
  hipFunction_t ft = nullptr, fa = nullptr;
  hipError_t e1 = hipModuleGetFunction(&ft, mod, name_t.c_str());
  hipError_t e2 = hipModuleGetFunction(&fa, mod, name_a.c_str());
  if (e1 != hipSuccess || e2 != hipSuccess) {
    (void)hipModuleUnload(mod);
    return fail(GG_ERR_DEVICE, "hipModuleGetFunction: %s / %s", hipGetErrorString(e1), hipGetErrorString(e2));
  }
  ctx->user_mod[key] = mod;
  *trace = ctx->user_trace[key] = ft;
  *at = ctx->user_at[key] = fa;
  return GG_OK;
}

// The user tensor at x, on the device (one thread of the helper kernel).
int user_metric_at(gg_context *ctx, hipFunction_t at, const gg_metric &m, const double x[4], double g[16]) {
  int rc;
  if ((rc = ctx->misc_io.ensure(256))) return rc;
  double *dx = (double *)ctx->misc_io.p, *dg = dx + 4;
  HIP_TRY(hipMemcpyAsync(dx, x, 32, hipMemcpyHostToDevice, ctx->stream));
  double mass = m.mass, spin = m.spin;
  void *args[] = {&dx, &mass, &spin, &dg};
  HIP_TRY(hipModuleLaunchKernel(at, 1, 1, 1, 1, 1, 1, 0, ctx->stream, args, nullptr));
  HIP_TRY(hipMemcpyAsync(g, dg, 128, hipMemcpyDeviceToHost, ctx->stream));
  HIP_TRY(hipStreamSynchronize(ctx->stream));
  return GG_OK;
}

// Common body of gg_render / gg_render_batch / gg_trace_states.
int render_impl(gg_context *ctx, const gg_scene *scenes, uint32_t n_scenes, int precision, const gg_outputs *outs,
                const double *states_host, uint64_t n_states, gg_stats *stats) {
  auto t_start = std::chrono::steady_clock::now();
  if (!ctx) return fail(GG_ERR_INVALID_ARG, "context is NULL");
  if (!scenes || n_scenes == 0) return fail(GG_ERR_INVALID_ARG, "no scenes");
  if (!outs) return fail(GG_ERR_INVALID_ARG, "outputs are NULL");
  if (precision < GG_PREC_FP64 || precision > GG_PREC_FP64_FAST)
    return fail(GG_ERR_INVALID_ARG, "precision %d unknown", precision);
  const bool from_states = (states_host != nullptr);
  int rc;
  for (uint32_t i = 0; i < n_scenes; ++i)
    if ((rc = validate_scene(&scenes[i], !from_states))) return rc;
  for (uint32_t i = 1; i < n_scenes; ++i)
    if (scenes[i].metric.kind != scenes[0].metric.kind || scenes[i].disk.kind != scenes[0].disk.kind)
      return fail(GG_ERR_UNSUPPORTED, "a batch must share metric kind and disk kind (scene %u differs)", i);
  // Hittables: render_with_scene (no disk) or render_with_disk_and_scene
  // (volumetric disk), one scene per call as the reference's drivers.
  const bool with_hits = ctx->n_hit > 0 && !ctx->hit_off;
  int mode = scenes[0].disk.kind;
  if (with_hits) {
    if (n_scenes > 1) return fail(GG_ERR_UNSUPPORTED, "hittables are set: render one scene per call");
    if (mode == GG_DISK_THIN) return fail(GG_ERR_UNSUPPORTED, "hittables are set: the thin disk has no scene mode");
    mode = mode == GG_DISK_NONE ? 3 : 4;
  }
  HIP_TRY(hipSetDevice(ctx->device));
  // A user metric: its trace kernel (compiled on first use) and its tensor
  // at the camera.
  const bool user = scenes[0].metric.kind == GG_METRIC_USER;
  hipFunction_t user_trace = nullptr;
  UserMetricAt um{};
  if (user) {
    if (ctx->user_src.empty()) return fail(GG_ERR_INVALID_ARG, "metric.kind is GG_METRIC_USER but no user metric is set");
    if (n_scenes > 1) return fail(GG_ERR_UNSUPPORTED, "a user metric renders one scene per call");
    if (precision != GG_PREC_FP64) return fail(GG_ERR_UNSUPPORTED, "a user metric renders in GG_PREC_FP64 only");
    if (ctx->rec_req.pool) return fail(GG_ERR_UNSUPPORTED, "a user metric has no recording render");
    hipFunction_t at = nullptr;
    if ((rc = user_kernels(ctx, mode, from_states ? 1 : 0, &user_trace, &at))) return rc;
    if ((rc = user_metric_at(ctx, at, scenes[0].metric, scenes[0].camera.position, um.g))) return rc;
    // the trace kernel reads the tensor's derivatives from its upper
    // triangle, which equals the full read only for a symmetric tensor
    for (int i = 0; i < 4; ++i)
      for (int j = i + 1; j < 4; ++j)
        if (!(um.g[4 * i + j] == um.g[4 * j + i]))
          return fail(GG_ERR_INVALID_ARG, "the user metric tensor is not symmetric at the camera: g[%d][%d] = %.17g, "
                      "g[%d][%d] = %.17g (both triangles must hold the same value)", i, j, um.g[4 * i + j], j, i,
                      um.g[4 * j + i]);
    um.has_rh = ctx->user_has_rh;
    um.rh = ctx->user_rh;
  }

  std::vector<DevScene> ds(n_scenes);
  std::vector<DevOut> dout(n_scenes);
  std::vector<std::vector<OutPlan>> plans(n_scenes);
  uint64_t total = 0;
  if (ctx->batch_stage.size() < 10ull * n_scenes) ctx->batch_stage.resize(10ull * n_scenes);
  for (uint32_t i = 0; i < n_scenes; ++i) {
    build_dev_scene(scenes[i], ds[i], user ? &um : nullptr);
    uint64_t n = from_states ? n_states : (uint64_t)ds[i].sub_w * ds[i].sub_h;
    ds[i].ray_begin = total;
    ds[i].out_index = i;
    total += n;
    const gg_outputs &o = outs[i];
    plans[i].resize(10);
    DevBuf *st = n_scenes == 1 ? ctx->stage : &ctx->batch_stage[10ull * i];
    if (with_hits) {
      ds[i].objs = ctx->hit_objs.p;
      ds[i].nodes = ctx->hit_nodes.p;
      ds[i].scene_rmin = ctx->hit_rmin;
    }
    bool need_rgb = (o.rgb != nullptr) && !from_states;
    // rgb shading reads cls, intensity and the velocity planes: stage them
    // internally when the caller did not ask for them.
    void *cls_p = o.cls, *int_p = o.intensity, *state_p = o.state;
    if ((rc = plan_out(st[0], cls_p, n, plans[i][0]))) return rc;
    if ((rc = plan_out(st[1], o.r, 8 * n, plans[i][1]))) return rc;
    if ((rc = plan_out(st[2], o.theta, 8 * n, plans[i][2]))) return rc;
    if ((rc = plan_out(st[3], o.phi, 8 * n, plans[i][3]))) return rc;
    if ((rc = plan_out(st[4], int_p, 8 * n, plans[i][4]))) return rc;
    if ((rc = plan_out(st[5], o.attempts, 4 * n, plans[i][5]))) return rc;
    if ((rc = plan_out(st[6], o.accepted, 4 * n, plans[i][6]))) return rc;
    if ((rc = plan_out(st[7], state_p, 72 * n, plans[i][7]))) return rc;
    if (with_hits && (rc = plan_out(st[9], o.color, 24 * n, plans[i][9]))) return rc;
    if (need_rgb) {
      if (with_hits && !plans[i][9].dev) { if ((rc = st[9].ensure(24 * n))) return rc; plans[i][9].dev = st[9].p; }
      if (!plans[i][0].dev) { if ((rc = st[0].ensure(n))) return rc; plans[i][0].dev = st[0].p; }
      if (!plans[i][4].dev) { if ((rc = st[4].ensure(8 * n))) return rc; plans[i][4].dev = st[4].p; }
      if (!plans[i][7].dev) { if ((rc = st[7].ensure(72 * n))) return rc; plans[i][7].dev = st[7].p; }
      size_t rgb_bytes = 3ull * scenes[i].image.width * scenes[i].image.height;
      if ((rc = plan_out(st[8], o.rgb, rgb_bytes, plans[i][8]))) return rc;
    }

    DevOut &d = dout[i];
    d.cls = (uint8_t *)plans[i][0].dev;
    d.r = (double *)plans[i][1].dev;
    d.theta = (double *)plans[i][2].dev;
    d.phi = (double *)plans[i][3].dev;
    d.intensity = (double *)plans[i][4].dev;
    d.attempts = (uint32_t *)plans[i][5].dev;
    d.accepted = (uint32_t *)plans[i][6].dev;
    d.state = (double *)plans[i][7].dev;
    d.color = (double *)plans[i][9].dev;
    d.n_rays = n;
  }
  // learned ray order: attempts are the sort keys of the next render's queue
  const bool learn_order = !from_states && ctx->launch.order == 1 && total > 0 && total <= 0xFFFFFFFFull;
  uint64_t sig = 1469598103934665603ull;   // FNV-1a over the scenes' ray grids
  if (learn_order) {
    auto mix = [&](uint64_t v) { sig = (sig ^ v) * 1099511628211ull; };
    mix(n_scenes);
    for (uint32_t i = 0; i < n_scenes; ++i) { mix(ds[i].sub_w); mix(ds[i].sub_h); mix(ds[i].tiled); }
    bool need = false;
    for (uint32_t i = 0; i < n_scenes; ++i) need |= !plans[i][5].dev;
    if (need && (rc = ctx->order_att.ensure(4 * total))) return rc;
    for (uint32_t i = 0; i < n_scenes; ++i)
      if (!plans[i][5].dev) dout[i].attempts = (uint32_t *)ctx->order_att.p + ds[i].ray_begin;
  }
  if ((rc = ctx->scenes.ensure(sizeof(DevScene) * n_scenes))) return rc;
  if ((rc = ctx->outs.ensure(sizeof(DevOut) * n_scenes))) return rc;
  hipStream_t s = ctx->stream;
  HIP_TRY(hipMemcpyAsync(ctx->scenes.p, ds.data(), sizeof(DevScene) * n_scenes, hipMemcpyHostToDevice, s));
  HIP_TRY(hipMemcpyAsync(ctx->outs.p, dout.data(), sizeof(DevOut) * n_scenes, hipMemcpyHostToDevice, s));
  const double *states_dev = nullptr;
  if (from_states) {
    if (is_device_ptr(states_host)) {
      states_dev = states_host;
    } else {
      if ((rc = ctx->states_in.ensure(72 * n_states))) return rc;
      HIP_TRY(hipMemcpyAsync(ctx->states_in.p, states_host, 72 * n_states, hipMemcpyHostToDevice, s));
      states_dev = (const double *)ctx->states_in.p;
    }
  }

  TraceLaunch T;
  T.scene = ds[0];
  T.scenes = (const DevScene *)ctx->scenes.p;
  T.outs = (const DevOut *)ctx->outs.p;
  T.n_scenes = n_scenes;
  T.n_rays = total;
  T.states_in = states_dev;
  T.counters = (Counters *)ctx->counters.p;
  T.refill_interval = ctx->launch.refill_interval;
  T.order = (learn_order && ctx->order_sig == sig) ? (const uint32_t *)ctx->order.p : nullptr;
  T.rec_pool = ctx->rec_req.pool;
  T.rec_cap = ctx->rec_req.cap;
  T.rec_cursor = ctx->rec_req.cursor;
  T.rec_last = ctx->rec_req.last;
  // wavefronts per SIMD the launched kernel's registers allow (GG_MIN_WAVES)
  const uint32_t max_waves = user ? 1u : (mode >= 3 || ctx->rec_req.pool) ? 2u : 3u;
  uint32_t grid = grid_for(ctx, total, max_waves);
  uint32_t block = ctx->launch.block_threads;
  const bool use_graph = (n_scenes > 1) && ctx->launch.graph == 1 && total > 0;
  if (use_graph) {
    // compared with memcmp: zero the padding bytes before filling the fields
    gg_context::GraphKey key;
    std::memset(&key, 0, sizeof(key));
    key.n_scenes = n_scenes;
    key.metric = scenes[0].metric.kind;
    key.disk = scenes[0].disk.kind;
    key.precision = precision;
    key.n_rays = total;
    key.launch = ctx->launch;
    key.scenes = ctx->scenes.p;
    key.outs = ctx->outs.p;
    key.counters = ctx->counters.p;
    key.order = (const void *)T.order;
    if (!ctx->graph_exec || std::memcmp(&key, &ctx->graph_key, sizeof(key)) != 0) {
      if (ctx->graph_exec) { (void)hipGraphExecDestroy(ctx->graph_exec); ctx->graph_exec = nullptr; }
      if (ctx->graph) { (void)hipGraphDestroy(ctx->graph); ctx->graph = nullptr; }
      // The scene and output tables are device buffers rewritten before every
      // replay, so the captured kernel arguments stay valid for equal keys.
      HIP_TRY(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
      hipError_t e1 = launch_reset(T.counters, s);
      hipError_t e2 = launch_trace(scenes[0].metric.kind, scenes[0].disk.kind, precision, T, grid, block, s);
      hipGraph_t g = nullptr;
      hipError_t e3 = hipStreamEndCapture(s, &g);
      if (e1 != hipSuccess || e2 != hipSuccess || e3 != hipSuccess) {
        if (g) (void)hipGraphDestroy(g);
        return fail(GG_ERR_DEVICE, "graph capture failed: %s / %s / %s", hipGetErrorString(e1),
                    hipGetErrorString(e2), hipGetErrorString(e3));
      }
      ctx->graph = g;
      HIP_TRY(hipGraphInstantiate(&ctx->graph_exec, g, nullptr, nullptr, 0));
      std::memcpy(&ctx->graph_key, &key, sizeof(key));
    }
    HIP_TRY(hipEventRecord(ctx->ev0, s));
    HIP_TRY(hipGraphLaunch(ctx->graph_exec, s));
    HIP_TRY(hipEventRecord(ctx->ev1, s));
  } else {
    HIP_TRY(launch_reset(T.counters, s));
    HIP_TRY(hipEventRecord(ctx->ev0, s));
    if (total > 0) {
      if (user) {
        void *args[] = {&T};
        HIP_TRY(hipModuleLaunchKernel(user_trace, grid, 1, 1, block, 1, 1, (unsigned)((size_t)block * 5 * 4 * 8), s,
                                      args, nullptr));
      } else if (mode >= 3) HIP_TRY(launch_trace_scene(scenes[0].metric.kind, mode, precision, T, grid, block, s));
      else HIP_TRY(launch_trace(scenes[0].metric.kind, mode, precision, T, grid, block, s));
    }
    HIP_TRY(hipEventRecord(ctx->ev1, s));
  }
  for (uint32_t i = 0; i < n_scenes; ++i) {
    if (plans[i][8].dev) {
      const DevOut *dop = (const DevOut *)ctx->outs.p + i;
      const DevScene *dsp = (const DevScene *)ctx->scenes.p + i;
      const SkyDesc sky{ctx->sky_w ? (const float *)ctx->sky.p : nullptr, ctx->sky_w, ctx->sky_h, ctx->sky_intensity};
      HIP_TRY(launch_shade(dsp, dop, scenes[i].image.width, scenes[i].image.height,
                           scenes[i].image.samples_per_axis, sky, (uint8_t *)plans[i][8].dev, s));
    }
  }
  if (learn_order) {
    // the next render's queue: rays by this render's attempt counts, longest
    // first.  The stored order is invalid until the sort below has run: a
    // failure on the way (a reallocation leaves the buffer uninitialised)
    // must not leave the previous grid's signature pointing at it.
    ctx->order_sig = 0;
    if ((rc = ctx->order.ensure(4 * total)) || (rc = ctx->order_keys.ensure(4 * total)) ||
        (rc = ctx->order_ids.ensure(4 * total)) || (rc = ctx->order_keys_out.ensure(4 * total)))
      return rc;
    uint32_t max_steps = 0;
    for (uint32_t i = 0; i < n_scenes; ++i) max_steps = std::max(max_steps, scenes[i].integrator.max_steps);
    int key_bits = 1;
    while (key_bits < 32 && (max_steps >> key_bits) != 0) ++key_bits;
    size_t temp_bytes = 0;
    HIP_TRY(sort_rays_desc(nullptr, nullptr, nullptr, nullptr, total, key_bits, nullptr, temp_bytes, s));
    if ((rc = ctx->order_tmp.ensure(temp_bytes ? temp_bytes : 1))) return rc;
    HIP_TRY(launch_order_keys((const DevScene *)ctx->scenes.p, (const DevOut *)ctx->outs.p, n_scenes, total,
                              (uint32_t *)ctx->order_keys.p, (uint32_t *)ctx->order_ids.p, s));
    HIP_TRY(sort_rays_desc((const uint32_t *)ctx->order_keys.p, (uint32_t *)ctx->order_keys_out.p,
                           (const uint32_t *)ctx->order_ids.p, (uint32_t *)ctx->order.p, total, key_bits,
                           ctx->order_tmp.p, temp_bytes, s));
    ctx->order_sig = sig;
  }
  Counters cnt;
  HIP_TRY(hipMemcpyAsync(&cnt, T.counters, sizeof(Counters), hipMemcpyDeviceToHost, s));
  for (uint32_t i = 0; i < n_scenes; ++i)
    for (auto &p : plans[i])
      if (p.host && p.dev) HIP_TRY(hipMemcpyAsync(p.host, p.dev, p.bytes, hipMemcpyDeviceToHost, s));
  HIP_TRY(hipStreamSynchronize(s));
  // exact-once coverage: every ray handed out by the queue retired once
  if (cnt.retired != total)
    return fail(GG_ERR_INTERNAL, "ray coverage: %llu of %llu rays retired", (unsigned long long)cnt.retired,
                (unsigned long long)total);
  if (stats) {
    float ms = 0.0f;
    HIP_TRY(hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
    stats->kernel_ms = ms;
    stats->total_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t_start).count();
    stats->rays = total;
    stats->attempts = cnt.attempts;
    stats->rhs_evals = cnt.rhs;
    stats->rhs_evals_fp32 = cnt.rhs32;
    stats->rhs_evals_fast = cnt.rhsf;
    stats->attempts_full = cnt.attempts_full;
    stats->reflections = cnt.reflections;
    stats->grid_blocks = grid;
    stats->block_threads = block;
  }
  return GG_OK;
}

}  // namespace

extern "C" {

int gg_render(gg_context *ctx, const gg_scene *scene, int precision, const gg_outputs *out, gg_stats *stats) {
  return render_impl(ctx, scene, 1, precision, out, nullptr, 0, stats);
}

int gg_render_batch(gg_context *ctx, const gg_scene *scenes, uint32_t n, int precision, const gg_outputs *out,
                    gg_stats *stats) {
  return render_impl(ctx, scenes, n, precision, out, nullptr, 0, stats);
}

// gg_render_batch without allocating: the per-call device data go to this
// GPU's `scratch` (its address taken by gg_context_create), host-side tables
// live on the stack, the outputs are the caller's device buffers, and the
// launch is direct.  No hipMalloc, hipFree or allocating library call.
int gg_render_batch_noalloc(gg_context *ctx, const gg_scene *scenes, uint32_t n, int precision,
                            const gg_outputs *outs, gg_stats *stats) {
  auto t_start = std::chrono::steady_clock::now();
  if (!ctx) return fail(GG_ERR_INVALID_ARG, "context is NULL");
  if (!scenes || n == 0) return fail(GG_ERR_INVALID_ARG, "no scenes");
  if (!outs) return fail(GG_ERR_INVALID_ARG, "outputs are NULL");
  if (n > GG_NOALLOC_MAX_SCENES)
    return fail(GG_ERR_INVALID_ARG, "gg_render_batch_noalloc takes at most %d scenes (%u given): its scratch holds "
                                    "%d scenes' tables", GG_NOALLOC_MAX_SCENES, n, GG_NOALLOC_MAX_SCENES);
  if (precision < GG_PREC_FP64 || precision > GG_PREC_FP64_FAST)
    return fail(GG_ERR_INVALID_ARG, "precision %d unknown", precision);
  if (!ctx->scratch) return fail(GG_ERR_INTERNAL, "no scratch address for device %d", ctx->device);
  int rc;
  for (uint32_t i = 0; i < n; ++i) {
    if ((rc = validate_scene(&scenes[i], true))) return rc;
    if (scenes[i].metric.kind == GG_METRIC_USER)
      return fail(GG_ERR_UNSUPPORTED, "gg_render_batch_noalloc: a user metric compiles its kernel (allocates)");
    if (scenes[i].metric.kind != scenes[0].metric.kind || scenes[i].disk.kind != scenes[0].disk.kind)
      return fail(GG_ERR_UNSUPPORTED, "a batch must share metric kind and disk kind (scene %u differs)", i);
  }
  if (ctx->n_hit > 0 && !ctx->hit_off)
    return fail(GG_ERR_UNSUPPORTED, "gg_render_batch_noalloc: hittables are set (scene mode renders one scene per "
                                    "gg_render call)");
  HIP_TRY(hipSetDevice(ctx->device));
  DevScene ds[GG_NOALLOC_MAX_SCENES];
  DevOut dout[GG_NOALLOC_MAX_SCENES];
  uint64_t total = 0;
  for (uint32_t i = 0; i < n; ++i) {
    const gg_outputs &o = outs[i];
    if (o.rgb || o.color)
      return fail(GG_ERR_UNSUPPORTED, "gg_render_batch_noalloc: rgb and color outputs need staging buffers");
    const void *ptrs[8] = {o.cls, o.r, o.theta, o.phi, o.intensity, o.attempts, o.accepted, o.state};
    for (const void *p : ptrs)
      if (p && !is_device_ptr(p))
        return fail(GG_ERR_INVALID_ARG, "gg_render_batch_noalloc: scene %u has a host output pointer (device "
                                        "memory only)", i);
    build_dev_scene(scenes[i], ds[i]);
    const uint64_t rays = (uint64_t)ds[i].sub_w * ds[i].sub_h;
    ds[i].ray_begin = total;
    ds[i].out_index = i;
    total += rays;
    DevOut &d = dout[i];
    d.cls = o.cls;
    d.r = o.r;
    d.theta = o.theta;
    d.phi = o.phi;
    d.intensity = o.intensity;
    d.attempts = o.attempts;
    d.accepted = o.accepted;
    d.state = o.state;
    d.color = nullptr;
    d.n_rays = rays;
  }
  char *base = (char *)ctx->scratch;
  Counters *cnt_d = (Counters *)(base + SCRATCH_COUNTERS);
  DevScene *ds_d = (DevScene *)(base + SCRATCH_SCENES);
  DevOut *do_d = (DevOut *)(base + SCRATCH_OUTS);
  hipStream_t s = ctx->stream;
  HIP_TRY(hipMemcpyAsync(ds_d, ds, sizeof(DevScene) * n, hipMemcpyHostToDevice, s));
  HIP_TRY(hipMemcpyAsync(do_d, dout, sizeof(DevOut) * n, hipMemcpyHostToDevice, s));
  TraceLaunch T;
  std::memset(&T, 0, sizeof(T));
  T.scene = ds[0];
  T.scenes = ds_d;
  T.outs = do_d;
  T.n_scenes = n;
  T.n_rays = total;
  T.counters = cnt_d;
  T.refill_interval = ctx->launch.refill_interval;
  const uint32_t grid = grid_for(ctx, total, 3);   // disk modes only (checked above)
  const uint32_t block = ctx->launch.block_threads;
  HIP_TRY(launch_reset(cnt_d, s));       // the counters start at zero (a kernel, not a memset)
  HIP_TRY(hipEventRecord(ctx->ev0, s));
  if (total > 0) HIP_TRY(launch_trace(scenes[0].metric.kind, scenes[0].disk.kind, precision, T, grid, block, s));
  HIP_TRY(hipEventRecord(ctx->ev1, s));
  Counters cnt;
  HIP_TRY(hipMemcpyAsync(&cnt, cnt_d, sizeof(Counters), hipMemcpyDeviceToHost, s));
  HIP_TRY(hipStreamSynchronize(s));
  // exact-once coverage: every ray handed out by the queue retired once
  if (cnt.retired != total)
    return fail(GG_ERR_INTERNAL, "ray coverage: %llu of %llu rays retired", (unsigned long long)cnt.retired,
                (unsigned long long)total);
  if (stats) {
    float ms = 0.0f;
    HIP_TRY(hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
    stats->kernel_ms = ms;
    stats->total_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t_start).count();
    stats->rays = total;
    stats->attempts = cnt.attempts;
    stats->rhs_evals = cnt.rhs;
    stats->rhs_evals_fp32 = cnt.rhs32;
    stats->rhs_evals_fast = cnt.rhsf;
    stats->attempts_full = cnt.attempts_full;
    stats->reflections = cnt.reflections;
    stats->grid_blocks = grid;
    stats->block_threads = block;
  }
  return GG_OK;
}

int gg_autotune(gg_context *ctx, const gg_scene *scenes, uint32_t n, int precision, const gg_launch *candidates,
                uint32_t n_candidates, gg_launch *best, double *best_kernel_ms) {
  if (!ctx || !scenes || n == 0) return fail(GG_ERR_INVALID_ARG, "NULL argument");
  std::vector<gg_launch> cands;
  if (candidates && n_candidates) {
    cands.assign(candidates, candidates + n_candidates);
  } else {
    // 4, 8 or 12 resident wavefronts per CU (one to three per SIMD)
    for (uint32_t bt : {64u, 128u, 256u})
      for (uint32_t waves : {4u, 8u, 12u})
        for (uint32_t ri : {8u, 32u}) cands.push_back(gg_launch{bt, waves * 64u / bt, ri, 2, 0});
  }
  const gg_launch saved = ctx->launch;
  std::vector<gg_outputs> none(n);
  std::memset(none.data(), 0, sizeof(gg_outputs) * n);
  double best_ms = 1e300;
  gg_launch best_l = saved;
  // Each candidate renders twice and the second launch is timed: graphs are
  // built and, with order = 1, the ray order is learned by the first.
  auto trial = [&](const gg_launch &c) -> int {
    int rc = gg_set_launch(ctx, &c);
    if (rc) return rc;
    gg_stats st;
    rc = render_impl(ctx, scenes, n, precision, none.data(), nullptr, 0, &st);
    if (!rc) rc = render_impl(ctx, scenes, n, precision, none.data(), nullptr, 0, &st);
    if (rc) return rc;
    if (st.kernel_ms < best_ms) { best_ms = st.kernel_ms; best_l = ctx->launch; }
    return GG_OK;
  };
  for (const gg_launch &c : cands) {
    int rc = trial(c);
    if (rc) { ctx->launch = saved; return rc; }
  }
  if (!(candidates && n_candidates)) {
    // second stage of the built-in search: the best shape with the learned order
    gg_launch c = best_l;
    c.order = 1;
    int rc = trial(c);
    if (rc) { ctx->launch = saved; return rc; }
  }
  best_l.graph = saved.graph;
  ctx->launch = best_l;
  if (best) *best = best_l;
  if (best_kernel_ms) *best_kernel_ms = best_ms;
  return GG_OK;
}

int gg_trace_states(gg_context *ctx, const gg_scene *scene, int precision, uint64_t n, const double *states_in,
                    const gg_outputs *out, gg_stats *stats) {
  if (!states_in && n > 0) return fail(GG_ERR_INVALID_ARG, "states_in is NULL");
  if (n == 0) return GG_OK;
  return render_impl(ctx, scene, 1, precision, out, states_in, n, stats);
}

int gg_geodesic_acceleration(gg_context *ctx, const gg_metric *metric, int precision, uint64_t n, const double *pos,
                             const double *vel, double *acc) {
  if (!ctx || !metric) return fail(GG_ERR_INVALID_ARG, "NULL argument");
  if (n == 0) return GG_OK;
  if (!pos || !vel || !acc) return fail(GG_ERR_INVALID_ARG, "NULL pos/vel/acc");
  if (precision < GG_PREC_FP64 || precision > GG_PREC_FP64_FAST)
    return fail(GG_ERR_INVALID_ARG, "precision %d unknown", precision);
  int rc = validate_metric(*metric);
  if (rc) return rc;
  HIP_TRY(hipSetDevice(ctx->device));
  gg_scene sc;
  std::memset(&sc, 0, sizeof(sc));
  sc.metric = *metric;
  sc.camera.position[1] = 10.0 * metric->mass;
  sc.camera.position[2] = M_PI / 2.0;
  sc.camera.look_at[0] = -1.0;
  sc.camera.up[2] = 1.0;
  sc.camera.fov_y_radians = 1.0;
  sc.camera.aspect = 1.0;
  sc.image.width = sc.image.height = sc.image.samples_per_axis = 1;
  DevScene D;
  build_dev_scene(sc, D);
  size_t bytes = 32 * n;
  if ((rc = ctx->rhs_io.ensure(3 * bytes + sizeof(DevScene) + 256))) return rc;
  char *base = (char *)ctx->rhs_io.p;
  DevScene *dS = (DevScene *)base;
  double *dp = (double *)(base + 256 * ((sizeof(DevScene) + 255) / 256));
  double *dv = dp + 4 * n, *da = dv + 4 * n;
  hipStream_t s = ctx->stream;
  HIP_TRY(hipMemcpyAsync(dS, &D, sizeof(D), hipMemcpyHostToDevice, s));
  HIP_TRY(hipMemcpyAsync(dp, pos, bytes, hipMemcpyDefault, s));
  HIP_TRY(hipMemcpyAsync(dv, vel, bytes, hipMemcpyDefault, s));
  HIP_TRY(launch_rhs(metric->kind, precision, dS, n, dp, dv, da, s));
  HIP_TRY(hipMemcpyAsync(acc, da, bytes, hipMemcpyDefault, s));
  HIP_TRY(hipStreamSynchronize(s));
  return GG_OK;
}

int gg_camera_rays(gg_context *ctx, const gg_scene *scene, double *states) {
  if (!ctx || !states) return fail(GG_ERR_INVALID_ARG, "NULL argument");
  int rc = validate_scene(scene, true);
  if (rc) return rc;
  HIP_TRY(hipSetDevice(ctx->device));
  DevScene D;
  build_dev_scene(*scene, D);
  const uint64_t n = (uint64_t)D.sub_w * D.sub_h;
  const size_t head = 256 * ((sizeof(DevScene) + 255) / 256);
  const bool dev = is_device_ptr(states);
  if ((rc = ctx->misc_io.ensure(head + (dev ? 0 : 72 * n)))) return rc;
  char *base = (char *)ctx->misc_io.p;
  double *out = dev ? states : (double *)(base + head);
  hipStream_t s = ctx->stream;
  HIP_TRY(hipMemcpyAsync(base, &D, sizeof(D), hipMemcpyHostToDevice, s));
  HIP_TRY(launch_camera_rays((const DevScene *)base, n, out, s));
  if (!dev) HIP_TRY(hipMemcpyAsync(states, out, 72 * n, hipMemcpyDeviceToHost, s));
  HIP_TRY(hipStreamSynchronize(s));
  return GG_OK;
}

int gg_constants_of_motion(gg_context *ctx, const gg_metric *metric, uint64_t n, const double *states, double *out) {
  if (!ctx || !metric) return fail(GG_ERR_INVALID_ARG, "NULL argument");
  if (n == 0) return GG_OK;
  if (!states || !out) return fail(GG_ERR_INVALID_ARG, "NULL states/out");
  int rc = validate_metric(*metric);
  if (rc) return rc;
  HIP_TRY(hipSetDevice(ctx->device));
  gg_scene sc;
  std::memset(&sc, 0, sizeof(sc));
  sc.metric = *metric;
  sc.camera.position[1] = 10.0 * metric->mass;
  sc.camera.position[2] = M_PI / 2.0;
  sc.camera.look_at[0] = -1.0;
  sc.camera.up[2] = 1.0;
  sc.camera.fov_y_radians = 1.0;
  sc.camera.aspect = 1.0;
  DevScene D;
  build_dev_scene(sc, D);
  const size_t head = 256 * ((sizeof(DevScene) + 255) / 256);
  const bool din = is_device_ptr(states), dout = is_device_ptr(out);
  if ((rc = ctx->misc_io.ensure(head + (din ? 0 : 72 * n) + (dout ? 0 : 32 * n)))) return rc;
  char *base = (char *)ctx->misc_io.p;
  const double *sin_ = din ? states : (const double *)(base + head);
  double *o = dout ? out : (double *)(base + head + (din ? 0 : 72 * n));
  hipStream_t s = ctx->stream;
  HIP_TRY(hipMemcpyAsync(base, &D, sizeof(D), hipMemcpyHostToDevice, s));
  if (!din) HIP_TRY(hipMemcpyAsync((void *)sin_, states, 72 * n, hipMemcpyHostToDevice, s));
  HIP_TRY(launch_constants(metric->kind, (const DevScene *)base, n, sin_, o, s));
  if (!dout) HIP_TRY(hipMemcpyAsync(out, o, 32 * n, hipMemcpyDeviceToHost, s));
  HIP_TRY(hipStreamSynchronize(s));
  return GG_OK;
}

int gg_grad_scene_to_scene(const gg_grad_scene *g, gg_scene *s) {
  if (!g || !s) return fail(GG_ERR_INVALID_ARG, "NULL argument");
  if (!(g->mass > 0.0) || !std::isfinite(g->mass)) return fail(GG_ERR_INVALID_ARG, "mass must be positive and finite");
  if (!std::isfinite(g->spin) || !(std::fabs(g->spin) < 0.9999 * g->mass))
    return fail(GG_ERR_INVALID_ARG, "|spin| must be < 0.9999 mass (Kerr::new clamps beyond, without derivative)");
  if (!(g->inclination > 0.0) || !(g->inclination < M_PI))
    return fail(GG_ERR_INVALID_ARG, "inclination must be in (0, pi)");
  double r_h = g->mass + std::sqrt(g->mass * g->mass - g->spin * g->spin);
  if (!(g->r_camera > r_h) || !std::isfinite(g->r_camera))
    return fail(GG_ERR_INVALID_ARG, "r_camera must be outside the horizon");
  if (!std::isfinite(g->emissivity_q) || !std::isfinite(g->r_inner) || !(g->r_outer > 0.0) || !std::isfinite(g->r_outer))
    return fail(GG_ERR_INVALID_ARG, "disk q, r_inner, r_outer must be finite (r_outer > 0)");
  if (!(g->edge_width >= 0.0) || !std::isfinite(g->edge_width)) return fail(GG_ERR_INVALID_ARG, "edge_width must be >= 0");
  if (g->width == 0 || g->height == 0 || g->width > 65536 || g->height > 65536)
    return fail(GG_ERR_INVALID_ARG, "width/height must be in [1, 65536]");
  std::memset(s, 0, sizeof(*s));
  s->metric.kind = GG_METRIC_KERR;
  s->metric.mass = g->mass;
  s->metric.spin = g->spin;
  const double st = std::sin(g->inclination), ct = std::cos(g->inclination);
  s->camera.position[1] = g->r_camera;
  s->camera.position[2] = g->inclination;
  s->camera.look_at[0] = -(st * 1.0);
  s->camera.look_at[1] = -(st * 0.0);
  s->camera.look_at[2] = -ct;
  s->camera.up[2] = 1.0;
  s->camera.fov_y_radians = g->fov_y_radians;
  s->camera.aspect = (double)g->width / (double)g->height;
  s->disk.kind = GG_DISK_THIN;
  s->disk.r_inner = g->r_inner;
  s->disk.r_outer = g->r_outer;
  s->disk.emissivity_q = g->emissivity_q;
  s->disk.edge_width = g->edge_width;
  s->disk.scale_height = 1.0;
  s->disk.density_0 = 1.0;
  s->disk.temperature_0 = 1.0;
  s->disk.absorption = 0.1;
  s->integrator = g->integrator;
  s->image.width = g->width;
  s->image.height = g->height;
  s->image.samples_per_axis = 1;
  return validate_scene(s, true);
}

int gg_grad(gg_context *ctx, const gg_grad_scene *scene, const double *target, uint32_t checkpoint_every,
            double *loss, double grad[4], double *intensity, double *d_intensity, gg_grad_stats *stats) {
  auto t_start = std::chrono::steady_clock::now();
  if (!ctx || !scene || !target) return fail(GG_ERR_INVALID_ARG, "NULL context, scene or target");
  if (checkpoint_every > 65536) return fail(GG_ERR_INVALID_ARG, "checkpoint_every must be in [0, 65536]");
  if (scene->precision != GG_PREC_FP64 && scene->precision != GG_PREC_MIXED)
    return fail(GG_ERR_INVALID_ARG, "gg_grad_scene.precision must be GG_PREC_FP64 or GG_PREC_MIXED");
  if (scene->frozen_steps > 1) return fail(GG_ERR_INVALID_ARG, "gg_grad_scene.frozen_steps must be 0 or 1");
  const bool mixed = scene->precision == GG_PREC_MIXED;
  if (mixed && checkpoint_every != 0)
    return fail(GG_ERR_UNSUPPORTED, "the mixed forward render needs checkpoint_every = 0 (recorded steps): its FP32 "
                                    "steps depend on the lane grouping, so checkpoints cannot be recomputed");
  gg_scene sc;
  int rc = gg_grad_scene_to_scene(scene, &sc);
  if (rc) return rc;
  HIP_TRY(hipSetDevice(ctx->device));
  const uint64_t n = (uint64_t)scene->width * scene->height;
  DevBuf *B = ctx->grad;
  if ((rc = B[0].ensure(n)) || (rc = B[1].ensure(8 * n)) || (rc = B[2].ensure(4 * n)) ||
      (rc = B[3].ensure(8 * n)) || (rc = B[4].ensure(8 * n)) || (rc = B[5].ensure(32 * n)) ||
      (rc = B[6].ensure(4 * n)) || (rc = B[7].ensure(sizeof(GradCounters))) || (rc = B[8].ensure(64 + grad_reduce_scratch_bytes())) ||
      (rc = B[12].ensure(4 * n)) || (rc = B[13].ensure(4 * n)) ||
      (rc = B[14].ensure(4 * n)) || (rc = B[21].ensure(4 * n)))
    return rc;
  uint8_t *d_cls = (uint8_t *)B[0].p;
  double *d_int = (double *)B[1].p, *d_target = (double *)B[3].p, *d_resid = (double *)B[4].p;
  double *d_dI = (double *)B[5].p, *d_out5 = (double *)B[8].p;
  uint32_t *d_acc = (uint32_t *)B[2].p, *d_list_in = (uint32_t *)B[6].p, *d_keys_in = (uint32_t *)B[12].p;
  uint32_t *d_att = (uint32_t *)B[21].p;
  uint32_t *d_list = (uint32_t *)B[13].p, *d_keys = (uint32_t *)B[14].p;
  GradCounters *d_cnt = (GradCounters *)B[7].p;
  hipStream_t s = ctx->stream;

  // Modes.  2 (checkpoint_every = 0, when the records fit): the render
  // records every attempt and the reverse sweep reads those records.
  // 1: a re-trace of the disk rays stores every step record (checkpoint_every
  // at least the longest ray's accepted count, or automatic when the render's
  // records did not fit but the disk rays' do).  0: the re-trace stores a
  // checkpoint every K accepted steps and the reverse sweep recomputes
  // segments.
  // Budget of the automatic choice: gg_set_grad_memory, else a quarter of the
  // free device memory, counting the record pool this context already holds.
  const bool automatic = (checkpoint_every == 0);
  uint32_t K = automatic ? 32 : checkpoint_every;
  size_t free_b = 0, total_b = 0;
  if (automatic && !ctx->grad_budget) HIP_TRY(hipMemGetInfo(&free_b, &total_b));
  const uint64_t budget = ctx->grad_budget ? ctx->grad_budget : (uint64_t)(free_b + B[18].cap) / 4;
  const uint64_t chunk_bytes = 8ull * GRAD_CHUNK_DOUBLES;
  int mode = 0;
  uint64_t pool_chunks = 0;
  gg_stats rs;
  GradCounters cnt;
  for (int pass = 0;; ++pass) {
    bool record = false;
    if (automatic) {
      // pool size: the last recording's need with a margin, else 8 chunks
      // (128 steps) per ray; at least 2 chunks per ray, within the budget
      uint64_t want = ctx->grad_chunks ? ctx->grad_chunks + ctx->grad_chunks / 4 + 64 : 8 * n;
      if (want < 2 * n) want = 2 * n;
      if (want > budget / chunk_bytes) want = budget / chunk_bytes;
      if (want > 0xFFFFFFFEull) want = 0xFFFFFFFEull;
      if (want >= n) {
        if ((rc = B[18].ensure(want * chunk_bytes)) || (rc = B[19].ensure(4 * n)) || (rc = B[20].ensure(72 * n)))
          return rc;
        pool_chunks = want;
        record = true;
      }
    }
    HIP_TRY(launch_grad_reset(d_cnt, s));
    // the FP64 render: classes, intensities, attempt and accepted counts
    // (and, recording, the step records and final states)
    gg_outputs o;
    std::memset(&o, 0, sizeof(o));
    o.cls = d_cls;
    o.intensity = d_int;
    o.accepted = d_acc;
    o.attempts = d_att;
    if (record) {
      o.state = (double *)B[20].p;
      ctx->rec_req = {(double *)B[18].p, pool_chunks, &d_cnt->rec_cursor, (uint32_t *)B[19].p};
    }
    ctx->hit_off = true;      // the gradient's scene is its own: no hittables
    rc = render_impl(ctx, &sc, 1, record && mixed ? GG_PREC_MIXED : GG_PREC_FP64, &o, nullptr, 0, &rs);
    ctx->hit_off = false;
    ctx->rec_req = {};
    if (rc) return rc;

    HIP_TRY(hipMemcpyAsync(d_target, target, 8 * n, hipMemcpyDefault, s));
    GradSelect sel{n, d_cls, d_int, d_target, d_acc, d_att, K, d_resid, d_dI, d_list_in, d_keys_in, d_cnt};
    HIP_TRY(launch_grad_select(sel, s));
    HIP_TRY(hipMemcpyAsync(&cnt, d_cnt, sizeof(cnt), hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
    if (!record) break;
    if (cnt.rec_cursor <= pool_chunks) {
      ctx->grad_chunks = cnt.rec_cursor;
      mode = 2;
      break;
    }
    // the pool was too small: once more with the measured size if it fits
    ctx->grad_chunks = cnt.rec_cursor;
    if (pass > 0 || cnt.rec_cursor * chunk_bytes > budget) {
      pool_chunks = 0;
      break;
    }
  }
  if (mixed && mode != 2)
    return fail(GG_ERR_UNSUPPORTED, "the mixed forward render needs room for the recorded steps (%llu bytes of "
                                    "chunks; budget %llu): raise gg_set_grad_memory or use GG_PREC_FP64",
                (unsigned long long)(chunk_bytes * cnt.rec_cursor), (unsigned long long)budget);
  const uint64_t n_list = cnt.n_list;
  uint64_t chk_total = cnt.chk_total, tail_total = cnt.tail_total;
  if (mode == 2) {
    chk_total = 0;
    tail_total = 0;
  } else if (K >= cnt.max_acc) {
    mode = 1;
  } else if (automatic && 8 * GRAD_STEP_REC * cnt.att_total + 72 * n_list <= budget) {
    K = (uint32_t)cnt.max_acc;
    chk_total = n_list;
    tail_total = cnt.att_total;
    mode = 1;
  }
  // a recomputed segment holds K accepted steps and at most every attempt
  // its ray did not accept
  const uint64_t seg_cap = (uint64_t)K + cnt.max_rej;
  if (mode == 0 && seg_cap > 0xFFFFFFFFull) return fail(GG_ERR_INVALID_ARG, "segment buffer too large");

  // launch shapes: forward 256-thread blocks, 2 per CU; reverse 64-thread
  // blocks, 4 per CU (one wavefront per SIMD: its registers hold the record
  // it is loading ahead); never more wavefronts than disk rays
  const uint64_t waves = (n_list + 63) / 64;
  uint64_t grid_f = (uint64_t)ctx->num_cus * 2, grid_r = (uint64_t)ctx->num_cus * 4;
  if ((waves + 3) / 4 < grid_f) grid_f = (waves + 3) / 4;
  if (waves < grid_r) grid_r = waves;
  if (grid_f < 1) grid_f = 1;
  if (grid_r < 1) grid_r = 1;
  const uint64_t n_threads = grid_r * 64;
  const uint64_t chk_bytes = 72 * (chk_total ? chk_total : 1);
  const uint64_t tail_bytes = 8 * GRAD_STEP_REC * (tail_total ? tail_total : 1);
  const uint64_t seg_bytes = mode == 0 ? 8 * GRAD_STEP_REC * seg_cap * n_threads : 0;
  if ((rc = B[10].ensure(chk_bytes)) || (rc = B[11].ensure(seg_bytes ? seg_bytes : 1)) ||
      (rc = B[15].ensure(8 * GRAD_REC * (n_list + 1))) || (rc = B[16].ensure(tail_bytes)))
    return rc;

  float adj_ms = 0.0f;
  if (n_list > 0) {
    int key_bits = 1;
    while (key_bits < 32 && (sc.integrator.max_steps >> key_bits) != 0) ++key_bits;
    size_t temp_bytes = 0;
    HIP_TRY(sort_rays_desc(d_keys_in, d_keys, d_list_in, d_list, n_list, key_bits, nullptr, temp_bytes, s));
    if ((rc = B[17].ensure(temp_bytes ? temp_bytes : 1))) return rc;
    GradLaunch G;
    build_dev_scene(sc, G.scene);
    G.list = d_list;
    G.keys = d_keys;
    G.accepted = d_acc;
    G.n_list = n_list;
    G.rec = (double *)B[15].p;
    G.chk = (double *)B[10].p;
    G.chk_cap = chk_total;
    G.tail = (double *)B[16].p;
    G.tail_cap = tail_total;
    G.seg = (double *)B[11].p;
    G.seg_cap = mode == 0 ? (uint32_t)seg_cap : 0u;
    G.live = scene->frozen_steps ? 0u : 1u;
    G.K = K;
    G.width = scene->width;
    G.height = scene->height;
    G.incl = scene->inclination;
    G.r_cam = scene->r_camera;
    G.dI = d_dI;
    G.cnt = d_cnt;
    G.pool = mode == 2 ? (const double *)B[18].p : nullptr;
    G.pool_chunks = mode == 2 ? pool_chunks : 0;
    G.last_chunk = mode == 2 ? (const uint32_t *)B[19].p : nullptr;
    G.final_state = mode == 2 ? (const double *)B[20].p : nullptr;
    G.n_pix = n;
    HIP_TRY(hipEventRecord(ctx->ev0, s));
    HIP_TRY(sort_rays_desc(d_keys_in, d_keys, d_list_in, d_list, n_list, key_bits, B[17].p, temp_bytes, s));
    if (mode == 2)
      HIP_TRY(launch_adjoint_seed(G, s));
    else
      HIP_TRY(launch_adjoint_fwd(G, (uint32_t)grid_f, 256, s));
    HIP_TRY(launch_adjoint_rev(G, mode, (uint32_t)grid_r, 64, s));
    HIP_TRY(launch_adjoint_cam(G, s));
    HIP_TRY(hipEventRecord(ctx->ev1, s));
  }
  HIP_TRY(launch_grad_reduce(n, d_resid, d_dI, d_out5 + 8, d_out5, s));
  double out5[5];
  HIP_TRY(hipMemcpyAsync(out5, d_out5, sizeof(out5), hipMemcpyDeviceToHost, s));
  HIP_TRY(hipMemcpyAsync(&cnt, d_cnt, sizeof(cnt), hipMemcpyDeviceToHost, s));
  if (intensity) HIP_TRY(hipMemcpyAsync(intensity, d_int, 8 * n, hipMemcpyDefault, s));
  if (d_intensity) HIP_TRY(hipMemcpyAsync(d_intensity, d_dI, 32 * n, hipMemcpyDefault, s));
  HIP_TRY(hipStreamSynchronize(s));
  if (n_list > 0) HIP_TRY(hipEventElapsedTime(&adj_ms, ctx->ev0, ctx->ev1));
  // every disk ray finished, and every attempt of each was back-propagated
  // exactly once
  if (cnt.bad != 0 || cnt.done != n_list || cnt.steps != cnt.att_total || cnt.chk_cursor != chk_total ||
      cnt.tail_cursor != tail_total)
    return fail(GG_ERR_INTERNAL,
                "adjoint: %llu of %llu rays done, %llu left the rendered path, %llu of %llu steps, %llu/%llu "
                "checkpoints, %llu/%llu step records",
                (unsigned long long)cnt.done, (unsigned long long)n_list, (unsigned long long)cnt.bad,
                (unsigned long long)cnt.steps, (unsigned long long)cnt.att_total,
                (unsigned long long)cnt.chk_cursor, (unsigned long long)chk_total,
                (unsigned long long)cnt.tail_cursor, (unsigned long long)tail_total);
  if (loss) *loss = out5[0];
  if (grad)
    for (int k = 0; k < 4; ++k) grad[k] = out5[1 + k];
  if (stats) {
    stats->forward_ms = rs.kernel_ms;
    stats->adjoint_ms = adj_ms;
    stats->total_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t_start).count();
    stats->rays = n;
    stats->disk_rays = n_list;
    stats->steps = cnt.steps;
    stats->attempts = cnt.attempts;
    stats->checkpoints = chk_total;
    stats->checkpoint_bytes = 72 * chk_total;
    stats->segment_bytes = seg_bytes;
    stats->record_bytes = mode == 2 ? chunk_bytes * cnt.rec_cursor : 8 * GRAD_STEP_REC * tail_total;
    stats->grid_blocks = (uint32_t)grid_r;
    stats->block_threads = 64;
    stats->checkpoint_every = mode == 2 ? 0 : K;
    stats->mode = (uint32_t)mode;
  }
  return GG_OK;
}

int gg_geodesic_acceleration_vjp(gg_context *ctx, const gg_metric *metric, uint64_t n, const double *pos,
                                 const double *vel, const double *cot, double *pos_bar, double *vel_bar,
                                 double *spin_bar) {
  if (!ctx || !metric) return fail(GG_ERR_INVALID_ARG, "NULL argument");
  if (n == 0) return GG_OK;
  if (!pos || !vel || !cot || !pos_bar || !vel_bar || !spin_bar) return fail(GG_ERR_INVALID_ARG, "NULL array");
  int rc = validate_metric(*metric);
  if (rc) return rc;
  if (metric->kind != GG_METRIC_KERR) return fail(GG_ERR_UNSUPPORTED, "the VJP is implemented for Kerr only");
  HIP_TRY(hipSetDevice(ctx->device));
  gg_scene sc;
  std::memset(&sc, 0, sizeof(sc));
  sc.metric = *metric;
  sc.camera.position[1] = 10.0 * metric->mass;
  sc.camera.position[2] = M_PI / 2.0;
  sc.camera.look_at[0] = -1.0;
  sc.camera.up[2] = 1.0;
  sc.camera.fov_y_radians = 1.0;
  sc.camera.aspect = 1.0;
  DevScene D;
  build_dev_scene(sc, D);
  const size_t bytes = 32 * n;
  if ((rc = ctx->rhs_io.ensure(6 * bytes + 8 * n + 256 * ((sizeof(DevScene) + 255) / 256)))) return rc;
  char *base = (char *)ctx->rhs_io.p;
  DevScene *dS = (DevScene *)base;
  double *dp = (double *)(base + 256 * ((sizeof(DevScene) + 255) / 256));
  double *dv = dp + 4 * n, *dw = dv + 4 * n, *dpb = dw + 4 * n, *dvb = dpb + 4 * n, *dsb = dvb + 4 * n;
  hipStream_t s = ctx->stream;
  HIP_TRY(hipMemcpyAsync(dS, &D, sizeof(D), hipMemcpyHostToDevice, s));
  HIP_TRY(hipMemcpyAsync(dp, pos, bytes, hipMemcpyDefault, s));
  HIP_TRY(hipMemcpyAsync(dv, vel, bytes, hipMemcpyDefault, s));
  HIP_TRY(hipMemcpyAsync(dw, cot, bytes, hipMemcpyDefault, s));
  HIP_TRY(launch_rhs_vjp(dS, n, dp, dv, dw, dpb, dvb, dsb, s));
  HIP_TRY(hipMemcpyAsync(pos_bar, dpb, bytes, hipMemcpyDefault, s));
  HIP_TRY(hipMemcpyAsync(vel_bar, dvb, bytes, hipMemcpyDefault, s));
  HIP_TRY(hipMemcpyAsync(spin_bar, dsb, 8 * n, hipMemcpyDefault, s));
  HIP_TRY(hipStreamSynchronize(s));
  return GG_OK;
}

}  // extern "C"
