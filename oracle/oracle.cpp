// oracle.cpp -- C entry points of the FP64 CPU oracle (liboracle.so).
//
// TEST INFRASTRUCTURE ONLY: the checker the HIP kernels are compared against,
// and the CPU baseline timed beside them.  Build with -O2 -ffp-contract=off
// (Rust never contracts a*b+c into an FMA).  Nothing in libgargantua links it.
//
// Reference files restated here (paths relative to the reference root):
//   crates/raytracer/src/curved/camera.rs:20-29     pixel_direction
//   crates/raytracer/src/curved/ray.rs:14-61        GeodesicRay::from_camera
//   crates/raytracer/src/curved/tracer.rs:54-152    trace_ray, trace_ray_with_disk
//   crates/raytracer/src/curved/tracer.rs:185-189   spatial_dir
//   crates/raytracer/src/curved/disk.rs:40-81     density, emission
//   crates/raytracer/src/curved/renderer.rs:44-80   shade_outcome_with, disk_color, encode_srgb
//   crates/raytracer/src/curved/renderer.rs:256-345 render_inner, render_row
//   crates/raytracer/src/curved/sky.rs:25-41        ProceduralSky::sample
// and, through oracle_physics.hpp, gr-core's metric.rs, kerr.rs,
// schwarzschild.rs and geodesic.rs.
#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstring>
#include <thread>
#include <type_traits>
#include <vector>

#include "oracle_physics.hpp"
#include "oracle_dual.hpp"

using namespace oracle;

namespace {

template <class T> struct V3 { T x[3]; };

template <class T> T dot3(const V3<T> &a, const V3<T> &b) {  // nalgebra dot, U3: (a + b) + c
  return a.x[0] * b.x[0] + a.x[1] * b.x[1] + a.x[2] * b.x[2];
}
template <class T> V3<T> normalize3(const V3<T> &v) {  // v / v.norm()
  using std::sqrt;
  T n = sqrt(dot3(v, v));
  V3<T> r;
  for (int i = 0; i < 3; ++i) r.x[i] = v.x[i] / n;
  return r;
}
template <class T> V3<T> cross3(const V3<T> &a, const V3<T> &b) {
  V3<T> r;
  r.x[0] = a.x[1] * b.x[2] - a.x[2] * b.x[1];
  r.x[1] = a.x[2] * b.x[0] - a.x[0] * b.x[2];
  r.x[2] = a.x[0] * b.x[1] - a.x[1] * b.x[0];
  return r;
}

// Camera with scalar type T (the inclination camera differentiates it).
template <class T> struct CameraT {
  V4<T> pos;
  V3<T> look_at, up;
  double fov_y, aspect;
};

template <class T> CameraT<T> camera_from(const gg_camera &c) {
  CameraT<T> k;
  for (int i = 0; i < 4; ++i) k.pos[i] = T(c.position[i]);
  for (int i = 0; i < 3; ++i) { k.look_at.x[i] = T(c.look_at[i]); k.up.x[i] = T(c.up[i]); }
  k.fov_y = c.fov_y_radians;
  k.aspect = c.aspect;
  return k;
}

// Camera::pixel_direction (camera.rs:20-29).
template <class T> V3<T> pixel_direction(const CameraT<T> &cam, double u, double v) {
  V3<T> forward = normalize3(cam.look_at);
  V3<T> right = normalize3(cross3(forward, cam.up));
  V3<T> up = normalize3(cross3(right, forward));
  double half_h = std::tan(cam.fov_y / 2.0);
  double half_w = half_h * cam.aspect;
  V3<T> d;
  for (int i = 0; i < 3; ++i) d.x[i] = forward.x[i] + u * half_w * right.x[i] + v * half_h * up.x[i];
  return normalize3(d);
}

// SplitMix64 (Vigna): n-th output (n >= 1) of the generator seeded with `seed`.
inline uint64_t splitmix64(uint64_t seed, uint64_t n) {
  uint64_t z = seed + n * 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}
inline double unit53(uint64_t x) { return (double)(x >> 11) * (1.0 / 9007199254740992.0); }

// GeodesicRay::from_camera (ray.rs:14-61); with jitter the pixel centre
// offset +0.5 becomes +0.5 + j, j = U[-0.5, 0.5) from SplitMix64(seed, ray).
template <class T>
State<T> from_camera(const MetricP &mp, const T &a, const CameraT<T> &cam, const gg_image &img, uint32_t px,
                     uint32_t py, uint32_t width, uint32_t height) {
  using std::cos; using std::sin; using std::sqrt;
  double fx = (double)px + 0.5, fy = (double)py + 0.5;
  if (img.jitter) {
    uint64_t k = (uint64_t)py * width + px;
    fx = fx + (unit53(splitmix64(img.seed, 2 * k + 1)) - 0.5);
    fy = fy + (unit53(splitmix64(img.seed, 2 * k + 2)) - 0.5);
  }
  double u = fx / (double)width * 2.0 - 1.0;   // ray.rs:22
  double v = 1.0 - fy / (double)height * 2.0;  // ray.rs:23
  V3<T> n = pixel_direction(cam, u, v);
  V4<T> pos = cam.pos;
  T theta = pos[2], phi = pos[3];
  T st = sin(theta), ct = cos(theta), sp = sin(phi), cp = cos(phi);
  V3<T> r_hat{{st * cp, st * sp, ct}};
  V3<T> theta_hat{{ct * cp, ct * sp, -st}};
  V3<T> phi_hat{{-sp, cp, T(0.0)}};
  T n_r = dot3(n, r_hat), n_t = dot3(n, theta_hat), n_p = dot3(n, phi_hat);
  Gmn<T> g = metric_tensor(mp, a, pos);
  T kr = n_r / sqrt(g.rr);
  T kth = n_t / sqrt(g.thth);
  T kph = n_p / sqrt(g.phph);
  T spatial = g.rr * kr * kr + g.thth * kth * kth + g.phph * kph * kph;  // ray.rs:51-52
  T kt = sqrt(-spatial / g.tt);                                          // ray.rs:53
  State<T> s;
  s.pos = pos;
  s.vel[0] = kt; s.vel[1] = kr; s.vel[2] = kth; s.vel[3] = kph;
  s.lambda = T(0.0);
  return s;
}

template <class T> struct DiskT {
  int kind;
  T r_inner, q;
  double r_outer, scale_height, density_0, temperature_0, absorption, edge_width;
};

template <class T> DiskT<T> disk_from(const gg_disk &d) {
  DiskT<T> k;
  k.kind = d.kind;
  k.r_inner = T(d.r_inner); k.q = T(d.emissivity_q);
  k.r_outer = d.r_outer; k.scale_height = d.scale_height; k.density_0 = d.density_0;
  k.temperature_0 = d.temperature_0; k.absorption = d.absorption; k.edge_width = d.edge_width;
  return k;
}

// AccretionDisk::density (disk.rs:40-59).
double disk_density(const DiskT<double> &d, const V4<double> &pos) {
  double r = pos[1], theta = pos[2];
  double s = std::sin(theta), c = std::cos(theta);
  double cap_r = r * std::fabs(s);
  double z = r * c;
  if (cap_r < d.r_inner || cap_r > d.r_outer) return 0.0;
  double t = d.r_inner / cap_r;
  double radial = t * t;                                                  // powi(2)
  double vertical = std::exp(-z * z / (2.0 * d.scale_height * d.scale_height));
  return d.density_0 * radial * vertical;
}

// AccretionDisk::emission (disk.rs:65-81).
double disk_emission(const DiskT<double> &d, const V4<double> &pos) {
  double rho = disk_density(d, pos);
  if (rho == 0.0) return 0.0;
  double r = pos[1], theta = pos[2];
  double cap_r = r * std::fabs(std::sin(theta));
  if (cap_r < d.r_inner) return 0.0;
  double temp_proxy = std::pow(d.r_inner / cap_r, 0.75);
  return rho * rho * temp_proxy;
}

template <class T> struct TraceOut {
  uint8_t cls;
  uint64_t decisions = 14695981039346656037ull;   // FNV-1a of every attempt's StepInfo::decisions
  T intensity;
  double color[3] = {0.0, 0.0, 0.0};   // RayOutcome::Scene { color }
  uint32_t attempts, accepted;
  uint64_t rhs;
  uint32_t reflections = 0;   // accepted steps whose theta was reflected at a pole
  State<T> st;
};

// Recorded controller decisions of one ray, for frozen-step replays.
struct StepLog {
  std::vector<double> h;
  std::vector<uint8_t> acc;
  bool recording = false;
  bool replaying = false;
};

template <class T> T logistic(const T &x) { using std::exp; return 1.0 / (1.0 + exp(-x)); }

void disk_color(double intensity, double rgb[3]);

// ---- Hittable scene: World::hit over Sphere / Quad, shade_hit, static g^4 ----
struct HitObj {
  int32_t shape, material;
  double q[3], u[3], v[3], radius, color[3], fuzz;
  // Quad::with_arc_material (quad.rs:24-72)
  double normal[3], d, e_alpha[3], e_beta[3];
};
std::vector<HitObj> g_objs;

V3<double> v3(const double *p) { return V3<double>{{p[0], p[1], p[2]}}; }

// Sphere::hit (sphere.rs:53-88) / Quad::hit (quad.rs:75-127) on the
// interval (t_min, t_max): root, point and outward normal.
bool hit_one(const HitObj &h, const V3<double> &o, const V3<double> &dir, double t_min, double t_max, double &t,
             V3<double> &pt, V3<double> &outward) {
  auto surrounds = [&](double x) { return t_min < x && x < t_max; };   // Interval::surrounds
  if (h.shape == GG_SHAPE_SPHERE) {
    V3<double> oc{{o.x[0] - h.q[0], o.x[1] - h.q[1], o.x[2] - h.q[2]}};
    double a = dot3(dir, dir);
    double half_b = dot3(oc, dir);
    double c = dot3(oc, oc) - h.radius * h.radius;
    double disc = half_b * half_b - a * c;
    if (disc < 0.0) return false;
    double sq = std::sqrt(disc);
    double root = (-half_b - sq) / a;
    if (!surrounds(root)) {
      root = (-half_b + sq) / a;
      if (!surrounds(root)) return false;
    }
    t = root;
    for (int k = 0; k < 3; ++k) pt.x[k] = o.x[k] + dir.x[k] * root;                 // ray.at
    for (int k = 0; k < 3; ++k) outward.x[k] = (pt.x[k] - h.q[k]) / h.radius;
    return true;
  }
  V3<double> nrm = v3(h.normal);
  double denom = dot3(nrm, dir);
  if (std::fabs(denom) < 1e-8) return false;
  double tq = (h.d - dot3(nrm, o)) / denom;
  if (!surrounds(tq)) return false;
  for (int k = 0; k < 3; ++k) pt.x[k] = o.x[k] + dir.x[k] * tq;
  V3<double> planar{{pt.x[0] - h.q[0], pt.x[1] - h.q[1], pt.x[2] - h.q[2]}};
  double alpha = dot3(planar, v3(h.e_alpha)), beta = dot3(planar, v3(h.e_beta));
  if (alpha < 0.0 || alpha > 1.0 || beta < 0.0 || beta > 1.0) return false;
  t = tq;
  outward = nrm;
  return true;
}

// World::hit (scene/world.rs:67-80) on (0, 1), then HitRecord::new
// (core/hit.rs:26-48) orients the normal against the ray.
bool world_hit(const V3<double> &o, const V3<double> &dir, int &obj, V3<double> &pt, V3<double> &normal) {
  double closest = 1.0;
  bool any = false;
  for (size_t i = 0; i < g_objs.size(); ++i) {
    double t;
    V3<double> p, out;
    if (hit_one(g_objs[i], o, dir, 0.0, closest, t, p, out)) {
      closest = t; any = true; obj = (int)i; pt = p; normal = out;
    }
  }
  if (any && !(dot3(dir, normal) < 0.0))
    for (int k = 0; k < 3; ++k) normal.x[k] = -normal.x[k];
  return any;
}

// shade_hit (tracer.rs:369-384) with each material's scatter attenuation and
// emission: Lambertian albedo (lambertian.rs:41-56), Metal albedo when the
// mirror direction leaves the surface (metal.rs:40-57, fuzz 0), Dielectric
// white (dielectric.rs:65-90), DiffuseLight no scatter and its emission
// (diffuse_light.rs:22-35).
void shade_hit(const HitObj &h, const V3<double> &dir, const V3<double> &normal, double lin[3]) {
  double albedo[3] = {0.0, 0.0, 0.0}, emission[3] = {0.0, 0.0, 0.0};
  V3<double> unit = normalize3(dir);
  switch (h.material) {
    case GG_MATERIAL_LAMBERTIAN: for (int k = 0; k < 3; ++k) albedo[k] = h.color[k]; break;
    case GG_MATERIAL_METAL: {
      double dn = dot3(unit, normal);
      V3<double> refl;
      for (int k = 0; k < 3; ++k) refl.x[k] = unit.x[k] - 2.0 * dn * normal.x[k];  // v - 2 v.n n
      if (dot3(refl, normal) > 0.0) for (int k = 0; k < 3; ++k) albedo[k] = h.color[k];
      break;
    }
    case GG_MATERIAL_DIELECTRIC: albedo[0] = albedo[1] = albedo[2] = 1.0; break;
    default: for (int k = 0; k < 3; ++k) emission[k] = h.color[k]; break;
  }
  V3<double> neg{{-unit.x[0], -unit.x[1], -unit.x[2]}};
  double nd = std::fmax(dot3(neg, normal), 0.0);
  double facing = 0.5 + 0.5 * nd;
  for (int k = 0; k < 3; ++k) lin[k] = albedo[k] * facing + emission[k];
}

// spherical_to_cartesian (tracer.rs:12-18).
V3<double> sph_to_cart(const V4<double> &p) {
  double r = p[1], st = std::sin(p[2]);
  return V3<double>{{r * st * std::cos(p[3]), r * st * std::sin(p[3]), r * std::cos(p[2])}};
}

// cartesian_to_spherical (tracer.rs:23-28) and redshift_g4_static (:40-52).
double g4_static(const MetricP &mp, double a, const V3<double> &p, double g_tt_obs) {
  double r = std::fmax(std::sqrt(p.x[0] * p.x[0] + p.x[1] * p.x[1] + p.x[2] * p.x[2]), 1e-10);
  double z = p.x[2] / r;
  z = z < -1.0 ? -1.0 : (z > 1.0 ? 1.0 : z);
  V4<double> pe;
  pe[0] = 0.0; pe[1] = r; pe[2] = std::acos(z); pe[3] = std::atan2(p.x[1], p.x[0]);
  double g_tt_emit = -metric_tensor(mp, a, pe).tt;
  if (g_tt_emit <= 0.0 || g_tt_obs <= 0.0) return 0.0;
  double g = std::sqrt(g_tt_emit / g_tt_obs);
  double g2 = g * g;
  return g2 * g2;
}

// Per-ray loop.  GG_DISK_NONE restates trace_ray (tracer.rs:54-70),
// GG_DISK_VOLUMETRIC restates trace_ray_with_disk (tracer.rs:80-152) and
// GG_DISK_THIN is the thin equatorial disk: the accepted step is the straight
// segment pos0 -> pos0 + h vel0 (forward-Euler position), so the crossing of
// theta = pi/2 sits at s = (pi/2 - th0)/(th1 - th0) of it; r_c and the photon
// momentum are interpolated linearly at s, and a hit gives
// I = r_c^-q g^4 (x logistic edge weights when edge_width > 0).
// H: the type of the step size (double, or T to differentiate the
// step-size controller as well; see rk45_step).
template <class T, class H = double>
TraceOut<T> trace(const MetricP &mp, const T &a, const IntegP &ip, const DiskT<T> &disk, const V4<T> &observer,
                  State<T> st, StepLog *log = nullptr) {
  using std::sqrt; using std::exp; using std::pow;
  TraceOut<T> o;
  o.intensity = T(0.0);
  o.accepted = 0;
  o.rhs = 0;
  H h = H(ip.initial_step);
  T intensity = T(0.0);
  double transmission = 1.0;
  T sqrt_neg_gtt_obs = T(0.0);
  if (disk.kind != GG_DISK_NONE) {
    T g_tt_obs = -metric_tensor(mp, a, observer).tt;                     // tracer.rs:93
    if (!(val(g_tt_obs) >= 0.0)) g_tt_obs = T(0.0);
    sqrt_neg_gtt_obs = sqrt(g_tt_obs);                                    // tracer.rs:94
  }
  // trace_ray_with_scene (tracer.rs:207-248) without a disk and
  // trace_ray_with_disk_and_scene (tracer.rs:258-364) with a volumetric one
  const bool scene = !g_objs.empty() && (disk.kind == GG_DISK_NONE || disk.kind == GG_DISK_VOLUMETRIC);
  double g_tt_obs_scene = 0.0;
  double absorption = disk.absorption;
  if constexpr (std::is_same<T, double>::value) {
    if (scene) {
      g_tt_obs_scene = -metric_tensor(mp, a, observer).tt;              // tracer.rs:216, 270
      absorption = 0.1;                                                  // tracer.rs:268
    }
  }
  uint8_t cls = GG_CLASS_MAX_STEPS;
  uint32_t k = 0;
  for (; k < ip.max_steps; ++k) {
    State<T> before = st;
    StepInfo info;
    int force = -1;
    if constexpr (std::is_same<H, double>::value) {
      if (log && log->replaying && k < log->h.size()) { h = log->h[k]; force = log->acc[k]; }
      if (log && log->recording) log->h.push_back(h);
    }
    StepResult res = rk45_step(mp, a, ip, st, h, &info, force);
    o.decisions = (o.decisions ^ info.decisions) * 1099511628211ull;
    if (log && log->recording) log->acc.push_back(info.accepted ? 1 : 0);
    o.rhs += info.rhs;
    if (info.reflected) ++o.reflections;
    if (info.accepted) ++o.accepted;

    if constexpr (std::is_same<T, double>::value) {
      if (scene) {
        V3<double> cb = sph_to_cart(before.pos), ca = sph_to_cart(st.pos);
        V3<double> seg{{ca.x[0] - cb.x[0], ca.x[1] - cb.x[1], ca.x[2] - cb.x[2]}};
        if (std::sqrt(dot3(seg, seg)) > 0.0) {
          int obj;
          V3<double> pt, nrm;
          if (world_hit(cb, seg, obj, pt, nrm)) {
            double lin[3];
            shade_hit(g_objs[obj], seg, nrm, lin);
            double g4 = g4_static(mp, a, pt, g_tt_obs_scene);
            if (disk.kind == GG_DISK_VOLUMETRIC) {
              double dl[3];
              disk_color(intensity, dl);
              for (int c = 0; c < 3; ++c) o.color[c] = lin[c] * g4 * transmission + dl[c];
            } else {
              for (int c = 0; c < 3; ++c) o.color[c] = lin[c] * g4;
            }
            cls = GG_CLASS_SCENE; ++k; break;
          }
        }
      }
    }

    if (disk.kind == GG_DISK_VOLUMETRIC) {
      if constexpr (std::is_same<T, double>::value) {
        double dlambda = std::fmax(st.lambda - before.lambda, 0.0);      // tracer.rs:101
        if (dlambda > 0.0) {
          V4<double> mid;
          for (int i = 0; i < 4; ++i) mid[i] = 0.5 * (before.pos[i] + st.pos[i]);
          double rho = disk_density(disk, mid);
          if (rho > 0.0) {
            double j = disk_emission(disk, mid);
            V4<double> pm;
            for (int i = 0; i < 4; ++i) pm[i] = 0.5 * (before.vel[i] + st.vel[i]);
            double g = redshift_factor(mp, a, mid, pm, sqrt_neg_gtt_obs);
            double g2 = g * g;
            double redshifted = j * (g2 * g2);                           // powi(4)
            intensity = intensity + redshifted * dlambda * transmission;
            transmission *= std::exp(-rho * dlambda * absorption);
            if (transmission < 1e-3) { cls = GG_CLASS_DISK; ++k; break; }
          }
        }
        if (res == STEP_HORIZON || res == STEP_SINGULAR) {
          cls = val(intensity) > 0.0 ? GG_CLASS_DISK : GG_CLASS_HORIZON; ++k; break;
        }
        if (res == STEP_ESCAPED) { cls = val(intensity) > 0.0 ? GG_CLASS_DISK : GG_CLASS_ESCAPED; ++k; break; }
        continue;
      }
    }
    if (disk.kind == GG_DISK_THIN && info.accepted) {
      double th0 = val(before.pos[2]), th1 = val(st.pos[2]);
      bool side0 = th0 < HALF_PI, side1 = th1 < HALF_PI;
      if (side0 != side1) {
        T s = (HALF_PI - before.pos[2]) / (st.pos[2] - before.pos[2]);
        T rc = before.pos[1] + s * (st.pos[1] - before.pos[1]);
        double w = disk.edge_width;
        bool hit = (w > 0.0) ? (val(rc) <= disk.r_outer + 8.0 * w && val(rc) > 0.0)
                             : (val(rc) >= val(disk.r_inner) && val(rc) <= disk.r_outer);
        if (hit) {
          V4<T> pc, vc;
          pc[0] = before.pos[0] + s * (st.pos[0] - before.pos[0]);
          pc[1] = rc;
          pc[2] = T(HALF_PI);
          pc[3] = before.pos[3];
          for (int i = 0; i < 4; ++i) vc[i] = before.vel[i] + s * (st.vel[i] - before.vel[i]);
          T g = redshift_factor(mp, a, pc, vc, sqrt_neg_gtt_obs);
          T g2 = g * g;
          T I = pow(rc, -disk.q) * (g2 * g2);
          if (w > 0.0) I = I * logistic((rc - disk.r_inner) / w) * logistic((disk.r_outer - rc) / w);
          intensity = I;
          cls = GG_CLASS_DISK;
          ++k;
          break;
        }
      }
    }
    if (res == STEP_HORIZON || res == STEP_SINGULAR) { cls = GG_CLASS_HORIZON; ++k; break; }  // tracer.rs:63-65
    if (res == STEP_ESCAPED) { cls = GG_CLASS_ESCAPED; ++k; break; }
  }
  if (cls == GG_CLASS_MAX_STEPS && disk.kind == GG_DISK_VOLUMETRIC && val(intensity) > 0.0) cls = GG_CLASS_DISK;
  o.cls = cls;
  o.intensity = intensity;
  o.attempts = k;
  o.st = st;
  return o;
}

// ProceduralSky::sample (sky.rs:26-40).
void sky_sample(const double dir[3], double rgb[3]) {
  auto hash = [](double x) -> uint64_t {
    uint64_t bits;
    std::memcpy(&bits, &x, 8);
    return bits * 6364136223846793005ull + 1442695040888963407ull;
  };
  auto rotl = [](uint64_t x, int k) -> uint64_t { return (x << k) | (x >> (64 - k)); };
  uint64_t h = hash(dir[0]) ^ rotl(hash(dir[1]), 17) ^ rotl(hash(dir[2]), 34);
  if (h % 512 < 3) {
    double b = (double)(uint32_t)(180 + (uint32_t)(h % 75)) / 255.0;
    rgb[0] = rgb[1] = rgb[2] = b;
  } else {
    double v = std::fabs(dir[2]) * 15.0 / 255.0;
    rgb[0] = 0.0; rgb[1] = 0.0; rgb[2] = 10.0 / 255.0 + v;
  }
}

// HdriSky (sky.rs:43-122): equirectangular map, row 0 the +z pole, columns
// wrapping in azimuth, bilinear in f32 with wrap-x / clamp-y (fetch,
// :88-106), times intensity; sample (:108-122) maps theta = acos(z) to
// v = theta / pi and phi = atan2(y, x) to u = phi / (2 pi) + 1/2.  Empty:
// the procedural sky.
struct HdriSkyT {
  uint32_t width = 0, height = 0;
  std::vector<float> pixels;   // width * height * 3
  float intensity = 1.0f;
};
HdriSkyT g_hdri;

void hdri_sample(const HdriSkyT &k, const double dir[3], double rgb[3]) {
  const double FRAC_1_PI = 0.31830988618379067154;
  double z = dir[2] < -1.0 ? -1.0 : (dir[2] > 1.0 ? 1.0 : dir[2]);   // f64::clamp
  double theta = std::acos(z);
  double phi = std::atan2(dir[1], dir[0]);
  double u = phi * FRAC_1_PI * 0.5 + 0.5;
  double v = theta * FRAC_1_PI;
  double w = (double)k.width, h = (double)k.height;
  double fx = std::fmod(u * w, w);                                     // rem_euclid
  if (fx < 0.0) fx = fx + std::fabs(w);
  double fy = v * h;
  fy = fy < 0.0 ? 0.0 : (fy > h - 1.0 ? h - 1.0 : fy);
  size_t x0 = (size_t)std::floor(fx) % k.width;
  size_t y0 = (size_t)std::floor(fy);
  size_t x1 = (x0 + 1) % k.width;
  size_t y1 = std::min<size_t>(y0 + 1, k.height - 1);
  float tx = (float)(fx - std::floor(fx)), ty = (float)(fy - std::floor(fy));
  auto at = [&](size_t y, size_t x, int c) { return k.pixels[3 * (y * k.width + x) + c]; };
  for (int c = 0; c < 3; ++c) {
    float a = at(y0, x0, c) + (at(y0, x1, c) - at(y0, x0, c)) * tx;
    float b = at(y1, x0, c) + (at(y1, x1, c) - at(y1, x0, c)) * tx;
    float p = a + (b - a) * ty;
    rgb[c] = (double)(p * k.intensity);
  }
}

// spatial_dir (tracer.rs:185-189).
void spatial_dir(const double vel[4], double out[3]) {
  double s[3] = {vel[1], vel[2], vel[3]};
  double n = std::sqrt(s[0] * s[0] + s[1] * s[1] + s[2] * s[2]);
  if (n > 0.0) { for (int i = 0; i < 3; ++i) out[i] = s[i] / n; }
  else { out[0] = 1.0; out[1] = 0.0; out[2] = 0.0; }
}

// disk_color (renderer.rs:62-66).
void disk_color(double intensity, double rgb[3]) {
  double x = intensity / (1.0 + intensity);
  rgb[0] = std::pow(x, 0.5); rgb[1] = std::pow(x, 1.2); rgb[2] = std::pow(x, 2.5);
}

// encode_srgb (renderer.rs:69-80): clamp, sRGB curve, round, saturating u8.
uint8_t to_u8(double c) {
  if (c < 0.0) c = 0.0;
  if (c > 1.0) c = 1.0;
  double s = (c <= 0.0031308) ? 12.92 * c : 1.055 * std::pow(c, 1.0 / 2.4) - 0.055;
  double r = std::round(s * 255.0);
  if (!(r > 0.0)) return 0;  // NaN and negatives saturate to 0
  if (r >= 255.0) return 255;
  return (uint8_t)r;
}

// shade_outcome_with (renderer.rs:44-53) with the HDRI sky when one is set,
// else the procedural sky.
void shade(uint8_t cls, double intensity, const double vel[4], const double color[3], double lin[3]) {
  switch (cls) {
    case GG_CLASS_SCENE: for (int c = 0; c < 3; ++c) lin[c] = color[c]; break;
    case GG_CLASS_HORIZON: lin[0] = lin[1] = lin[2] = 0.0; break;
    case GG_CLASS_ESCAPED: {
      double d[3];
      spatial_dir(vel, d);
      if (g_hdri.width) hdri_sample(g_hdri, d, lin); else sky_sample(d, lin);
      break;
    }
    case GG_CLASS_DISK: disk_color(intensity, lin); break;
    default: lin[0] = 0.0; lin[1] = 0.0; lin[2] = 10.0 / 255.0; break;
  }
}

void store(const gg_outputs *out, uint64_t i, const TraceOut<double> &o) {
  if (out->cls) out->cls[i] = o.cls;
  if (out->color && !g_objs.empty())
    for (int c = 0; c < 3; ++c) out->color[3 * i + c] = o.color[c];
  if (out->r) out->r[i] = o.st.pos[1];
  if (out->theta) out->theta[i] = o.st.pos[2];
  if (out->phi) out->phi[i] = o.st.pos[3];
  if (out->intensity) out->intensity[i] = o.intensity;
  if (out->attempts) out->attempts[i] = o.attempts;
  if (out->accepted) out->accepted[i] = o.accepted;
}

void store_state(double *state, uint64_t n, uint64_t i, const State<double> &s) {
  if (!state) return;
  for (int c = 0; c < 4; ++c) state[c * n + i] = s.pos[c];
  for (int c = 0; c < 4; ++c) state[(4 + c) * n + i] = s.vel[c];
  state[8 * n + i] = s.lambda;
}

template <class F> void parallel_for(uint64_t n, int nthreads, uint64_t chunk, F f) {
  if (nthreads <= 0) nthreads = (int)std::max(1u, std::thread::hardware_concurrency());
  std::atomic<uint64_t> next{0};
  auto worker = [&]() {
    for (;;) {
      uint64_t b = next.fetch_add(chunk);
      if (b >= n) break;
      uint64_t e = std::min(n, b + chunk);
      for (uint64_t i = b; i < e; ++i) f(i);
    }
  };
  if (nthreads == 1) { worker(); return; }
  std::vector<std::thread> ts;
  for (int t = 0; t < nthreads; ++t) ts.emplace_back(worker);
  for (auto &t : ts) t.join();
}

bool scene_ok(const gg_scene *s) {
  return s && s->metric.mass > 0.0 && s->image.width > 0 && s->image.height > 0 &&
         s->image.samples_per_axis > 0;
}

thread_local uint64_t g_last_rhs = 0;
thread_local uint64_t g_last_reflections = 0;   // pole reflections of the last render

}  // namespace

extern "C" {

int oracle_metric_tensor(const gg_metric *m, const double pos[4], double g[16]) {
  MetricP mp = make_metric(*m);
  V4<double> p{{pos[0], pos[1], pos[2], pos[3]}};
  Gmn<double> G = metric_tensor(mp, mp.spin, p);
  for (int i = 0; i < 16; ++i) g[i] = 0.0;
  g[0] = G.tt; g[5] = G.rr; g[10] = G.thth; g[15] = G.phph; g[3] = G.tph; g[12] = G.tph;
  return 0;
}

// The tensor of GG_METRIC_USER scenes (fn null: none) and its horizon
// (has_rh 0: Metric::event_horizon is None).
void oracle_set_user_metric(void (*fn)(const double *, double, double, double *), int has_rh, double rh) {
  user_metric().g = fn;
  user_metric().has_rh = has_rh != 0;
  user_metric().rh = rh;
}

int oracle_inverse4(const double m[16], double inv[16]) { return inverse4(m, inv) ? 1 : 0; }

int oracle_christoffel(const gg_metric *m, const double pos[4], double gamma[64]) {
  MetricP mp = make_metric(*m);
  V4<double> p{{pos[0], pos[1], pos[2], pos[3]}};
  double G[4][4][4];
  if (mp.kind == GG_METRIC_USER) numerical_christoffel(mp, pos, G);
  else if (mp.kind == GG_METRIC_KERR) kerr_christoffel(mp, mp.spin, p, G);
  else schw_christoffel(mp, p, G);
  std::memcpy(gamma, G, sizeof(G));
  return 0;
}

int oracle_geodesic_acceleration(const gg_metric *m, uint64_t n, const double *pos, const double *vel, double *acc) {
  MetricP mp = make_metric(*m);
  for (uint64_t i = 0; i < n; ++i) {
    V4<double> p{{pos[4 * i], pos[4 * i + 1], pos[4 * i + 2], pos[4 * i + 3]}};
    V4<double> v{{vel[4 * i], vel[4 * i + 1], vel[4 * i + 2], vel[4 * i + 3]}};
    V4<double> a = geodesic_acceleration(mp, mp.spin, p, v);
    for (int c = 0; c < 4; ++c) acc[4 * i + c] = a[c];
  }
  return 0;
}

// E = -p_t, L_z = p_phi, Carter Q and mu^2 = -g(v, v) of n states (9
// doubles each); the textbook Kerr constants in Boyer-Lindquist coordinates.
int oracle_constants_of_motion(const gg_metric *m, uint64_t n, const double *states, double *out) {
  MetricP mp = make_metric(*m);
  const double a = mp.kind == GG_METRIC_KERR ? mp.spin : 0.0;
  for (uint64_t i = 0; i < n; ++i) {
    const double *s = states + 9 * i;
    V4<double> p{{s[0], s[1], s[2], s[3]}};
    Gmn<double> g = metric_tensor(mp, mp.spin, p);
    double pt = g.tt * s[4] + g.tph * s[7];
    double pph = g.tph * s[4] + g.phph * s[7];
    double pth = g.thth * s[6];
    double pr = g.rr * s[5];
    double mu2 = -(pt * s[4] + pr * s[5] + pth * s[6] + pph * s[7]);
    double st = std::sin(s[2]), ct = std::cos(s[2]);
    out[4 * i] = -pt;
    out[4 * i + 1] = pph;
    out[4 * i + 2] = pth * pth + ct * ct * (a * a * (mu2 - pt * pt) + pph * pph / (st * st));
    out[4 * i + 3] = mu2;
  }
  return 0;
}

// d a^mu / d(r, theta, v^t, v^r, v^theta, v^phi, spin) of
// geodesic_acceleration by forward-mode dual numbers; jac[28 i + 7 mu + k].
int oracle_rhs_jacobian(const gg_metric *m, uint64_t n, const double *pos, const double *vel, double *jac) {
  using D = Dual<7>;
  MetricP mp = make_metric(*m);
  for (uint64_t i = 0; i < n; ++i) {
    V4<D> p, v;
    p[0] = D(pos[4 * i]);
    p[1] = D::seed(pos[4 * i + 1], 0);
    p[2] = D::seed(pos[4 * i + 2], 1);
    p[3] = D(pos[4 * i + 3]);
    for (int c = 0; c < 4; ++c) v[c] = D::seed(vel[4 * i + c], 2 + c);
    D a = D::seed(mp.spin, 6);
    V4<D> acc = geodesic_acceleration(mp, a, p, v);
    for (int mu = 0; mu < 4; ++mu)
      for (int k = 0; k < 7; ++k) jac[28 * i + 7 * mu + k] = acc[mu].d[k];
  }
  return 0;
}

int oracle_event_horizon(const gg_metric *m, double *rh) {
  MetricP mp = make_metric(*m);
  return event_horizon(mp, rh) ? 1 : 0;
}

int oracle_orbital_four_velocity(const gg_metric *m, const double pos[4], double u[4]) {
  MetricP mp = make_metric(*m);
  V4<double> p{{pos[0], pos[1], pos[2], pos[3]}}, uu;
  if (!orbital_four_velocity(mp, mp.spin, p, uu)) return 0;
  for (int i = 0; i < 4; ++i) u[i] = uu[i];
  return 1;
}

double oracle_redshift_factor(const gg_metric *m, const double pos[4], const double p[4], double sqrt_neg_gtt_obs) {
  MetricP mp = make_metric(*m);
  V4<double> x{{pos[0], pos[1], pos[2], pos[3]}}, k{{p[0], p[1], p[2], p[3]}};
  return redshift_factor(mp, mp.spin, x, k, sqrt_neg_gtt_obs);
}

// One RK45Integrator::step on state[9]; returns the StepResult, updates
// state and *h, and reports acceptance and RHS count.
int oracle_rk45_step(const gg_metric *m, const gg_integrator *ig, double state[9], double *h, int *accepted,
                     int *rhs) {
  MetricP mp = make_metric(*m);
  IntegP ip = make_integ(*ig);
  State<double> s;
  for (int i = 0; i < 4; ++i) { s.pos[i] = state[i]; s.vel[i] = state[4 + i]; }
  s.lambda = state[8];
  StepInfo info;
  int r = rk45_step(mp, mp.spin, ip, s, *h, &info);
  for (int i = 0; i < 4; ++i) { state[i] = s.pos[i]; state[4 + i] = s.vel[i]; }
  state[8] = s.lambda;
  if (accepted) *accepted = info.accepted ? 1 : 0;
  if (rhs) *rhs = info.rhs;
  return r;
}

// RK4Integrator::step (geodesic.rs:73-132), for the reference's orbit tests.
int oracle_rk4_step(const gg_metric *m, double step, double max_radius, double horizon_buffer, double state[9]) {
  MetricP mp = make_metric(*m);
  State<double> s;
  for (int i = 0; i < 4; ++i) { s.pos[i] = state[i]; s.vel[i] = state[4 + i]; }
  s.lambda = state[8];
  int r = rk4_step(mp, mp.spin, step, max_radius, horizon_buffer, s);
  for (int i = 0; i < 4; ++i) { state[i] = s.pos[i]; state[4 + i] = s.vel[i]; }
  state[8] = s.lambda;
  return r;
}

void oracle_pixel_direction(const gg_camera *c, double u, double v, double out[3]) {
  CameraT<double> cam = camera_from<double>(*c);
  V3<double> d = pixel_direction(cam, u, v);
  for (int i = 0; i < 3; ++i) out[i] = d.x[i];
}

// Initial state of ray (px, py) of a (width x height) sub-pixel grid.
int oracle_from_camera(const gg_scene *s, uint32_t px, uint32_t py, uint32_t width, uint32_t height,
                       double state[9]) {
  MetricP mp = make_metric(s->metric);
  CameraT<double> cam = camera_from<double>(s->camera);
  State<double> st = from_camera(mp, mp.spin, cam, s->image, px, py, width, height);
  for (int i = 0; i < 4; ++i) { state[i] = st.pos[i]; state[4 + i] = st.vel[i]; }
  state[8] = st.lambda;
  return 0;
}

double oracle_disk_density(const gg_disk *d, const double pos[4]) {
  V4<double> p{{pos[0], pos[1], pos[2], pos[3]}};
  return disk_density(disk_from<double>(*d), p);
}
double oracle_disk_emission(const gg_disk *d, const double pos[4]) {
  V4<double> p{{pos[0], pos[1], pos[2], pos[3]}};
  return disk_emission(disk_from<double>(*d), p);
}
void oracle_disk_color(double intensity, double rgb[3]) { disk_color(intensity, rgb); }
void oracle_encode_srgb(const double lin[3], uint8_t out[3]) {
  for (int i = 0; i < 3; ++i) out[i] = to_u8(lin[i]);
}
void oracle_sky(const double dir[3], double rgb[3]) { sky_sample(dir, rgb); }

// The Hittable scene of oracle_render / oracle_trace_states (n = 0: none).
int oracle_set_hittables(uint32_t n, const gg_hittable *objs) {
  std::vector<HitObj> v(n);
  for (uint32_t i = 0; i < n; ++i) {
    const gg_hittable &g = objs[i];
    HitObj &h = v[i];
    h.shape = g.shape; h.material = g.material;
    for (int k = 0; k < 3; ++k) { h.q[k] = g.q[k]; h.u[k] = g.u[k]; h.v[k] = g.v[k]; h.color[k] = g.color[k]; }
    h.radius = g.radius;
    h.fuzz = std::fmin(std::fmax(g.fuzz, 0.0), 1.0);                 // Metal::new clamps
    if (g.material == GG_MATERIAL_METAL && h.fuzz != 0.0) return GG_ERR_UNSUPPORTED;
    if (g.material < GG_MATERIAL_LAMBERTIAN || g.material > GG_MATERIAL_DIFFUSE_LIGHT) return GG_ERR_INVALID_ARG;
    if (g.shape == GG_SHAPE_QUAD) {
      V3<double> u = v3(g.u), w0 = v3(g.v);
      V3<double> n = cross3(u, w0);
      V3<double> nn = normalize3(n);
      for (int k = 0; k < 3; ++k) h.normal[k] = nn.x[k];
      h.d = dot3(nn, v3(g.q));
      double n2 = dot3(n, n);
      V3<double> w{{n.x[0] / n2, n.x[1] / n2, n.x[2] / n2}};
      V3<double> ea = cross3(w0, w), eb = cross3(w, u);
      for (int k = 0; k < 3; ++k) { h.e_alpha[k] = ea.x[k]; h.e_beta[k] = eb.x[k]; }
    } else if (g.shape != GG_SHAPE_SPHERE) {
      return GG_ERR_INVALID_ARG;
    }
  }
  g_objs = std::move(v);
  return 0;
}

// One chord against the scene: hit flag, object, point, normal and the
// shade_hit colour (out[0..11]).
int oracle_scene_hit(const double origin[3], const double dir[3], double out[12]) {
  int obj = -1;
  V3<double> pt, nrm;
  if (!world_hit(v3(origin), v3(dir), obj, pt, nrm)) { out[0] = 0.0; return 0; }
  double lin[3];
  shade_hit(g_objs[obj], v3(dir), nrm, lin);
  out[0] = 1.0; out[1] = obj;
  for (int k = 0; k < 3; ++k) { out[2 + k] = pt.x[k]; out[5 + k] = nrm.x[k]; out[8 + k] = lin[k]; }
  return 1;
}

double oracle_g4_static(const gg_metric *m, const double point[3], double g_tt_obs) {
  MetricP mp = make_metric(*m);
  return g4_static(mp, mp.spin, v3(point), g_tt_obs);
}
// The sky of oracle_render (width = 0: the procedural sky).
int oracle_set_sky_hdri(uint32_t width, uint32_t height, const float *pixels, float intensity) {
  if (width == 0) { g_hdri = HdriSkyT{}; return 0; }
  if (height == 0 || !pixels) return GG_ERR_INVALID_ARG;
  g_hdri.width = width;
  g_hdri.height = height;
  g_hdri.pixels.assign(pixels, pixels + 3ull * width * height);
  g_hdri.intensity = intensity;
  return 0;
}
void oracle_hdri_sample(uint32_t width, uint32_t height, const float *pixels, float intensity, const double dir[3],
                        double rgb[3]) {
  HdriSkyT k;
  k.width = width; k.height = height; k.intensity = intensity;
  k.pixels.assign(pixels, pixels + 3ull * width * height);
  hdri_sample(k, dir, rgb);
}
uint64_t oracle_splitmix64(uint64_t seed, uint64_t n) { return splitmix64(seed, n); }

// Trace the rays of scene s listed in idx[0..n) (ray index = sub_py *
// sub_w + sub_px); output slot i belongs to idx[i].
int oracle_render_rays(const gg_scene *s, uint64_t n, const uint64_t *idx, const gg_outputs *out, int nthreads) {
  if (!scene_ok(s) || !out) return GG_ERR_INVALID_ARG;
  MetricP mp = make_metric(s->metric);
  IntegP ip = make_integ(s->integrator);
  CameraT<double> cam = camera_from<double>(s->camera);
  DiskT<double> disk = disk_from<double>(s->disk);
  uint32_t spa = s->image.samples_per_axis;
  uint32_t sw = s->image.width * spa, sh = s->image.height * spa;
  std::atomic<uint64_t> rhs{0}, refl{0};
  parallel_for(n, nthreads, 16, [&](uint64_t i) {
    uint64_t k = idx ? idx[i] : i;
    uint32_t px = (uint32_t)(k % sw), py = (uint32_t)(k / sw);
    State<double> st = from_camera(mp, mp.spin, cam, s->image, px, py, sw, sh);
    TraceOut<double> o = trace(mp, mp.spin, ip, disk, cam.pos, st);
    store(out, i, o);
    store_state(out->state, n, i, o.st);
    rhs.fetch_add(o.rhs, std::memory_order_relaxed);
    refl.fetch_add(o.reflections, std::memory_order_relaxed);
  });
  g_last_rhs = rhs.load();
  g_last_reflections = refl.load();
  return 0;
}

// Whole-image render: per-ray SoA outputs over the sub-pixel grid and the
// reference's RgbImage (renderer.rs:256-345) in out->rgb.
int oracle_render(const gg_scene *s, const gg_outputs *out, int nthreads) {
  if (!scene_ok(s) || !out) return GG_ERR_INVALID_ARG;
  MetricP mp = make_metric(s->metric);
  IntegP ip = make_integ(s->integrator);
  CameraT<double> cam = camera_from<double>(s->camera);
  DiskT<double> disk = disk_from<double>(s->disk);
  uint32_t spa = s->image.samples_per_axis;
  uint32_t w = s->image.width, hgt = s->image.height;
  uint32_t sw = w * spa, sh = hgt * spa;
  uint64_t n = (uint64_t)sw * sh;
  double inv_spp = 1.0 / (double)(spa * spa);
  std::atomic<uint64_t> rhs{0}, refl{0};
  parallel_for(hgt, nthreads, 1, [&](uint64_t py) {
    for (uint32_t px = 0; px < w; ++px) {
      double acc[3] = {0.0, 0.0, 0.0};
      for (uint32_t sy = 0; sy < spa; ++sy)
        for (uint32_t sx = 0; sx < spa; ++sx) {
          uint32_t spx = px * spa + sx, spy = (uint32_t)py * spa + sy;
          State<double> st = from_camera(mp, mp.spin, cam, s->image, spx, spy, sw, sh);
          TraceOut<double> o = trace(mp, mp.spin, ip, disk, cam.pos, st);
          uint64_t i = (uint64_t)spy * sw + spx;
          store(out, i, o);
          store_state(out->state, n, i, o.st);
          rhs.fetch_add(o.rhs, std::memory_order_relaxed);
          refl.fetch_add(o.reflections, std::memory_order_relaxed);
          double lin[3];
          double vel[4] = {o.st.vel[0], o.st.vel[1], o.st.vel[2], o.st.vel[3]};
          shade(o.cls, o.intensity, vel, o.color, lin);
          acc[0] += lin[0]; acc[1] += lin[1]; acc[2] += lin[2];
        }
      if (out->rgb) {
        uint8_t *p = out->rgb + 3 * ((uint64_t)py * w + px);
        double lin[3] = {acc[0] * inv_spp, acc[1] * inv_spp, acc[2] * inv_spp};
        for (int c = 0; c < 3; ++c) p[c] = to_u8(lin[c]);
      }
    }
  });
  g_last_rhs = rhs.load();
  g_last_reflections = refl.load();
  return 0;
}

// Trace caller-supplied states (9 doubles each, AoS).
int oracle_trace_states(const gg_scene *s, uint64_t n, const double *states, const gg_outputs *out, int nthreads) {
  if (!s || !out) return GG_ERR_INVALID_ARG;
  MetricP mp = make_metric(s->metric);
  IntegP ip = make_integ(s->integrator);
  CameraT<double> cam = camera_from<double>(s->camera);
  DiskT<double> disk = disk_from<double>(s->disk);
  std::atomic<uint64_t> rhs{0}, refl{0};
  parallel_for(n, nthreads, 16, [&](uint64_t i) {
    State<double> st;
    for (int c = 0; c < 4; ++c) { st.pos[c] = states[9 * i + c]; st.vel[c] = states[9 * i + 4 + c]; }
    st.lambda = states[9 * i + 8];
    TraceOut<double> o = trace(mp, mp.spin, ip, disk, cam.pos, st);
    store(out, i, o);
    store_state(out->state, n, i, o.st);
    rhs.fetch_add(o.rhs, std::memory_order_relaxed);
    refl.fetch_add(o.reflections, std::memory_order_relaxed);
  });
  g_last_rhs = rhs.load();
  g_last_reflections = refl.load();
  return 0;
}

uint64_t oracle_last_rhs(void) { return g_last_rhs; }
uint64_t oracle_last_reflections(void) { return g_last_reflections; }

// Prograde ISCO radius (Bardeen, Press & Teukolsky 1972, eq. 2.21).
double oracle_r_isco(double mass, double spin) {
  double a = spin / mass;
  double z1 = 1.0 + std::cbrt(1.0 - a * a) * (std::cbrt(1.0 + a) + std::cbrt(1.0 - a));
  double z2 = std::sqrt(3.0 * a * a + z1 * z1);
  double s = a >= 0.0 ? 1.0 : -1.0;
  return mass * (3.0 + z2 - s * std::sqrt((3.0 - z1) * (3.0 + z1 + 2.0 * z2)));
}

}  // extern "C"

// ---- differentiable thin-disk render: FP64 forward-mode dual numbers ------
//
// Tangent directions 0..3 = (spin, inclination, q, r_inner).  oracle_grad:
// L and dL/dtheta by dual numbers, the step sizes carrying tangents through
// the error estimate and the controller (the gradient of the loss the render
// computes), or, with frozen_steps set, held at their values (the gradient of
// the discrete algorithm with its step sequence fixed).  oracle_loss_frozen:
// L only, with the step sequence replayed from the unperturbed run of `ref`
// (frozen-step finite differences).
namespace {
struct GradCommon {
  MetricP mp;
  IntegP ip;
  DiskT<double> disk_d;
};

template <class T>
CameraT<T> incl_camera(const T &incl, double r_cam, double fov, double aspect) {
  using std::cos; using std::sin;
  CameraT<T> c;
  c.pos[0] = T(0.0); c.pos[1] = T(r_cam); c.pos[2] = incl; c.pos[3] = T(0.0);
  T st = sin(incl), ct = cos(incl);
  c.look_at.x[0] = -(st * 1.0); c.look_at.x[1] = -(st * 0.0); c.look_at.x[2] = -ct;
  c.up.x[0] = T(0.0); c.up.x[1] = T(0.0); c.up.x[2] = T(1.0);
  c.fov_y = fov;
  c.aspect = aspect;
  return c;
}
}  // namespace

extern "C" {

int oracle_grad_scene_to_scene(const gg_grad_scene *g, gg_scene *s) {
  std::memset(s, 0, sizeof(*s));
  s->metric.kind = GG_METRIC_KERR; s->metric.mass = g->mass; s->metric.spin = g->spin;
  CameraT<double> c = incl_camera<double>(g->inclination, g->r_camera, g->fov_y_radians,
                                          (double)g->width / (double)g->height);
  for (int i = 0; i < 4; ++i) s->camera.position[i] = c.pos[i];
  for (int i = 0; i < 3; ++i) { s->camera.look_at[i] = c.look_at.x[i]; s->camera.up[i] = c.up.x[i]; }
  s->camera.fov_y_radians = g->fov_y_radians; s->camera.aspect = c.aspect;
  s->disk.kind = GG_DISK_THIN; s->disk.r_inner = g->r_inner; s->disk.r_outer = g->r_outer;
  s->disk.emissivity_q = g->emissivity_q; s->disk.edge_width = g->edge_width;
  s->disk.scale_height = 1.0; s->disk.density_0 = 1.0; s->disk.temperature_0 = 1.0; s->disk.absorption = 0.1;
  s->integrator = g->integrator;
  s->image.width = g->width; s->image.height = g->height; s->image.samples_per_axis = 1;
  return 0;
}

// Dual-number loss gradient over all width*height pixel-centre rays.
// intensity (optional) gets the image, dI (optional, [4*N]) the per-pixel
// tangents.
int oracle_grad(const gg_grad_scene *g, const double *target, double *loss, double grad[4], double *intensity,
                double *dI, int nthreads) {
  using D = Dual<4>;
  if (!g || !target || g->width == 0 || g->height == 0 || g->mass <= 0.0) return GG_ERR_INVALID_ARG;
  gg_metric gm{GG_METRIC_KERR, 0, g->mass, g->spin};
  MetricP mp = make_metric(gm);
  IntegP ip = make_integ(g->integrator);
  D a = D::seed(mp.spin, 0), incl = D::seed(g->inclination, 1);
  double aspect = (double)g->width / (double)g->height;
  CameraT<D> cam = incl_camera<D>(incl, g->r_camera, g->fov_y_radians, aspect);
  DiskT<D> disk;
  disk.kind = GG_DISK_THIN; disk.r_inner = D::seed(g->r_inner, 3); disk.q = D::seed(g->emissivity_q, 2);
  disk.r_outer = g->r_outer; disk.edge_width = g->edge_width;
  disk.scale_height = 1.0; disk.density_0 = 1.0; disk.temperature_0 = 1.0; disk.absorption = 0.1;
  gg_image img{g->width, g->height, 1, 0, 0};
  uint64_t n = (uint64_t)g->width * g->height;
  std::vector<double> Ls(n), dLs(4 * n);
  parallel_for(n, nthreads, 16, [&](uint64_t i) {
    uint32_t px = (uint32_t)(i % g->width), py = (uint32_t)(i / g->width);
    State<D> st = from_camera(mp, a, cam, img, px, py, g->width, g->height);
    TraceOut<D> o = g->frozen_steps ? trace<D, double>(mp, a, ip, disk, cam.pos, st)
                                    : trace<D, D>(mp, a, ip, disk, cam.pos, st);
    D I = o.intensity;
    if (intensity) intensity[i] = I.v;
    if (dI) for (int k = 0; k < 4; ++k) dI[4 * i + k] = I.d[k];
    double r = I.v - target[i];
    Ls[i] = r * r;
    for (int k = 0; k < 4; ++k) dLs[4 * i + k] = 2.0 * r * I.d[k];
  });
  double L = 0.0, G[4] = {0, 0, 0, 0};
  for (uint64_t i = 0; i < n; ++i) {
    L += Ls[i];
    for (int k = 0; k < 4; ++k) G[k] += dLs[4 * i + k];
  }
  if (loss) *loss = L / (double)n;
  for (int k = 0; k < 4; ++k) grad[k] = G[k] / (double)n;
  return 0;
}

// Per pixel, a hash of the sequence of discrete decisions (accept/reject,
// err == 0, clamps of h, invalid stages, pole reflections) of the FP64 render
// oracle_grad differentiates.  Where two parameter sets give equal hashes, the
// loss is smooth between them.
int oracle_grad_decisions(const gg_grad_scene *g, uint64_t *out, int nthreads) {
  if (!g || !out || g->width == 0 || g->height == 0 || g->mass <= 0.0) return GG_ERR_INVALID_ARG;
  gg_metric gm{GG_METRIC_KERR, 0, g->mass, g->spin};
  MetricP mp = make_metric(gm);
  IntegP ip = make_integ(g->integrator);
  CameraT<double> cam = incl_camera<double>(g->inclination, g->r_camera, g->fov_y_radians,
                                            (double)g->width / (double)g->height);
  DiskT<double> disk;
  disk.kind = GG_DISK_THIN; disk.r_inner = g->r_inner; disk.q = g->emissivity_q; disk.r_outer = g->r_outer;
  disk.edge_width = g->edge_width; disk.scale_height = 1.0; disk.density_0 = 1.0; disk.temperature_0 = 1.0;
  disk.absorption = 0.1;
  gg_image img{g->width, g->height, 1, 0, 0};
  uint64_t n = (uint64_t)g->width * g->height;
  parallel_for(n, nthreads, 16, [&](uint64_t i) {
    uint32_t px = (uint32_t)(i % g->width), py = (uint32_t)(i / g->width);
    State<double> st = from_camera(mp, mp.spin, cam, img, px, py, g->width, g->height);
    out[i] = trace(mp, mp.spin, ip, disk, cam.pos, st).decisions;
  });
  return 0;
}

// Loss at parameters `g` with each ray's step sequence replayed from its run
// at parameters `ref` (frozen-step finite differences).  With ref == g this
// is the plain loss.
int oracle_loss_frozen(const gg_grad_scene *ref, const gg_grad_scene *g, const double *target, double *loss,
                       double *intensity, int nthreads) {
  if (!ref || !g || !target) return GG_ERR_INVALID_ARG;
  gg_metric gm0{GG_METRIC_KERR, 0, ref->mass, ref->spin}, gm1{GG_METRIC_KERR, 0, g->mass, g->spin};
  MetricP mp0 = make_metric(gm0), mp1 = make_metric(gm1);
  IntegP ip = make_integ(g->integrator);
  double aspect = (double)g->width / (double)g->height;
  CameraT<double> cam0 = incl_camera<double>(ref->inclination, ref->r_camera, ref->fov_y_radians, aspect);
  CameraT<double> cam1 = incl_camera<double>(g->inclination, g->r_camera, g->fov_y_radians, aspect);
  auto mkdisk = [](const gg_grad_scene *x) {
    DiskT<double> d;
    d.kind = GG_DISK_THIN; d.r_inner = x->r_inner; d.q = x->emissivity_q; d.r_outer = x->r_outer;
    d.edge_width = x->edge_width; d.scale_height = 1.0; d.density_0 = 1.0; d.temperature_0 = 1.0;
    d.absorption = 0.1;
    return d;
  };
  DiskT<double> d0 = mkdisk(ref), d1 = mkdisk(g);
  gg_image img{g->width, g->height, 1, 0, 0};
  uint64_t n = (uint64_t)g->width * g->height;
  std::vector<double> Ls(n);
  parallel_for(n, nthreads, 16, [&](uint64_t i) {
    uint32_t px = (uint32_t)(i % g->width), py = (uint32_t)(i / g->width);
    StepLog log;
    log.recording = true;
    State<double> s0 = from_camera(mp0, mp0.spin, cam0, img, px, py, g->width, g->height);
    trace(mp0, mp0.spin, ip, d0, cam0.pos, s0, &log);
    log.recording = false;
    log.replaying = true;
    State<double> s1 = from_camera(mp1, mp1.spin, cam1, img, px, py, g->width, g->height);
    TraceOut<double> o = trace(mp1, mp1.spin, ip, d1, cam1.pos, s1, &log);
    if (intensity) intensity[i] = o.intensity;
    double r = o.intensity - target[i];
    Ls[i] = r * r;
  });
  double L = 0.0;
  for (uint64_t i = 0; i < n; ++i) L += Ls[i];
  *loss = L / (double)n;
  return 0;
}

}  // extern "C"
