// Driver for the sanitizer run of the oracle (tests/test_oracle_sanitize.py):
// built with oracle/oracle.cpp under AddressSanitizer and
// UndefinedBehaviorSanitizer, it takes every entry point the tests use
// through small inputs: each disk mode and error mode, supersampled RGB, the
// Hittable scene, the HDRI sky, a run-time tensor metric, the spin clamp,
// caller states that are invalid, captured, escaped or run out of attempts,
// and the forward-mode gradient with its decision hashes and frozen loss.
// Exit status 0 and "oracle sanitize: ok" when every call succeeded and every
// ray retired with a valid class; the sanitizers abort on any finding.
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <vector>

#include "gargantua.h"

extern "C" {
int oracle_render(const gg_scene *s, const gg_outputs *out, int nthreads);
int oracle_trace_states(const gg_scene *s, uint64_t n, const double *states, const gg_outputs *out, int nthreads);
int oracle_set_hittables(uint32_t n, const gg_hittable *objs);
int oracle_set_sky_hdri(uint32_t width, uint32_t height, const float *pixels, float intensity);
void oracle_set_user_metric(void (*fn)(const double *, double, double, double *), int has_rh, double rh);
int oracle_christoffel(const gg_metric *m, const double pos[4], double gamma[64]);
int oracle_geodesic_acceleration(const gg_metric *m, uint64_t n, const double *pos, const double *vel, double *acc);
int oracle_rhs_jacobian(const gg_metric *m, uint64_t n, const double *pos, const double *vel, double *jac);
int oracle_constants_of_motion(const gg_metric *m, uint64_t n, const double *states, double *out);
int oracle_grad(const gg_grad_scene *g, const double *target, double *loss, double grad[4], double *intensity,
                double *dI, int nthreads);
int oracle_grad_decisions(const gg_grad_scene *g, uint64_t *out, int nthreads);
int oracle_loss_frozen(const gg_grad_scene *ref, const gg_grad_scene *g, const double *target, double *loss,
                       double *intensity, int nthreads);
double oracle_r_isco(double mass, double spin);
}

namespace {

int g_fail = 0;

void expect(bool ok, const char *what) {
  if (!ok) {
    std::fprintf(stderr, "FAILED: %s\n", what);
    g_fail = 1;
  }
}

gg_integrator rtol_integrator(double max_radius, uint32_t max_steps) {
  gg_integrator ig{};
  ig.initial_step = 0.1; ig.min_step = 1e-8; ig.max_step = 1.0; ig.tolerance = 1e-6;
  ig.max_radius = max_radius; ig.horizon_buffer = 0.0; ig.horizon_rel = 0.01;
  ig.rtol = 1e-8; ig.atol = 1e-10; ig.err_mode = GG_ERR_RTOL_ATOL; ig.max_steps = max_steps;
  return ig;
}

gg_scene make_scene(int metric, double spin, double r_cam, double incl_deg, double fov_deg, uint32_t w, uint32_t h) {
  gg_scene s;
  std::memset(&s, 0, sizeof s);
  s.metric.kind = metric; s.metric.mass = 1.0; s.metric.spin = spin;
  const double inc = incl_deg * M_PI / 180.0;
  const double cam[4] = {0.0, r_cam, inc, 0.0};
  std::memcpy(s.camera.position, cam, sizeof cam);
  s.camera.look_at[0] = -std::sin(inc); s.camera.look_at[1] = 0.0; s.camera.look_at[2] = -std::cos(inc);
  s.camera.up[2] = 1.0;
  s.camera.fov_y_radians = fov_deg * M_PI / 180.0;
  s.camera.aspect = (double)w / h;
  s.disk.kind = GG_DISK_THIN;
  s.disk.r_inner = metric == GG_METRIC_KERR ? oracle_r_isco(1.0, spin) : 6.0;
  s.disk.r_outer = 20.0; s.disk.emissivity_q = 3.0;
  s.integrator = rtol_integrator(1000.0, 20000);
  s.image.width = w; s.image.height = h; s.image.samples_per_axis = 1;
  return s;
}

// Every output plane of one render; checks that each ray retired with a
// class and a positive attempt count.
struct Out {
  std::vector<uint8_t> cls, rgb;
  std::vector<double> r, th, ph, inten, state, color;
  std::vector<uint32_t> att, acc;
  gg_outputs o;
  Out(uint64_t n, uint64_t pixels) : cls(n), rgb(3 * pixels), r(n), th(n), ph(n), inten(n), state(9 * n), color(3 * n),
                                     att(n), acc(n) {
    o = gg_outputs{cls.data(), r.data(), th.data(), ph.data(), inten.data(), att.data(), acc.data(), state.data(),
                   pixels ? rgb.data() : nullptr, color.data()};
  }
  int count(uint8_t c) const {
    int k = 0;
    for (uint8_t x : cls) k += x == c;
    return k;
  }
  bool retired() const {
    for (size_t i = 0; i < cls.size(); ++i)
      if (cls[i] > GG_CLASS_SCENE || att[i] == 0) return false;
    return true;
  }
};

uint64_t n_rays(const gg_scene &s) {
  const uint64_t k = s.image.samples_per_axis;
  return (uint64_t)s.image.width * k * s.image.height * k;
}

void render(const gg_scene &s, const char *what, bool rgb, int min_disk, int min_horizon) {
  Out o(n_rays(s), rgb ? (uint64_t)s.image.width * s.image.height : 0);
  expect(oracle_render(&s, &o.o, 2) == 0, what);
  expect(o.retired(), what);
  expect(o.count(GG_CLASS_DISK) >= min_disk, what);
  expect(o.count(GG_CLASS_HORIZON) >= min_horizon, what);
  std::printf("%s: %d disk, %d horizon, %d escaped, %d max-steps, %d scene\n", what, o.count(GG_CLASS_DISK),
              o.count(GG_CLASS_HORIZON), o.count(GG_CLASS_ESCAPED), o.count(GG_CLASS_MAX_STEPS),
              o.count(GG_CLASS_SCENE));
}

// Reissner-Nordstrom with Q = 0.5 M, as a tensor source.
void rn_metric(const double *pos, double mass, double /*spin*/, double *g) {
  const double r = pos[1], st = std::sin(pos[2]), q = 0.5 * mass;
  const double f = 1.0 - 2.0 * mass / r + q * q / (r * r);
  for (int i = 0; i < 16; ++i) g[i] = 0.0;
  g[0] = -f; g[5] = 1.0 / f; g[10] = r * r; g[15] = r * r * st * st;
}

}  // namespace

int main() {
  // thin disk: Schwarzschild (reference error norm) and Kerr (rtol/atol, jittered, 2x2 supersampled RGB)
  gg_scene s = make_scene(GG_METRIC_SCHWARZSCHILD, 0.0, 50.0, 80.0, 30.0, 24, 16);
  s.integrator.err_mode = GG_ERR_REFERENCE;
  render(s, "schwarzschild thin disk, reference error norm", false, 1, 1);
  s = make_scene(GG_METRIC_KERR, 0.9, 100.0, 60.0, 25.0, 12, 10);
  s.image.jitter = 1; s.image.seed = 2026; s.image.samples_per_axis = 2;
  render(s, "kerr thin disk, jitter, 2x2 samples, rgb", true, 1, 1);
  // volumetric disk and no disk
  s = make_scene(GG_METRIC_KERR, 0.5, 30.0, 75.0, 40.0, 12, 12);
  s.disk.kind = GG_DISK_VOLUMETRIC; s.disk.r_inner = 6.0; s.disk.scale_height = 1.0; s.disk.density_0 = 1.0;
  s.disk.temperature_0 = 1.0; s.disk.absorption = 0.1;
  s.integrator.max_radius = 200.0; s.integrator.max_steps = 5000;
  render(s, "kerr volumetric disk, rgb", true, 1, 0);
  s.disk.kind = GG_DISK_NONE;
  render(s, "kerr without disk, rgb", true, 0, 1);
  // spin beyond extremal (clamped) and near the poles
  s = make_scene(GG_METRIC_KERR, 1.5, 20.0, 2.0, 20.0, 8, 8);
  render(s, "spin 1.5, camera 2 deg from the pole", false, 0, 0);
  // Hittable scene, with and without the volumetric disk
  gg_hittable objs[3];
  std::memset(objs, 0, sizeof objs);
  objs[0].shape = GG_SHAPE_SPHERE; objs[0].material = GG_MATERIAL_DIFFUSE_LIGHT;
  objs[0].q[0] = -40.0; objs[0].radius = 4.0; objs[0].color[0] = 2.0; objs[0].color[1] = 1.6; objs[0].color[2] = 0.8;
  objs[1].shape = GG_SHAPE_SPHERE; objs[1].material = GG_MATERIAL_LAMBERTIAN;
  objs[1].q[0] = -30.0; objs[1].q[1] = 8.0; objs[1].radius = 3.0; objs[1].color[0] = 0.8;
  objs[2].shape = GG_SHAPE_QUAD; objs[2].material = GG_MATERIAL_METAL;
  objs[2].q[0] = -50.0; objs[2].q[1] = -20.0; objs[2].q[2] = -20.0; objs[2].u[1] = 40.0; objs[2].v[2] = 40.0;
  objs[2].color[1] = 0.9;
  expect(oracle_set_hittables(3, objs) == 0, "set hittables");
  s = make_scene(GG_METRIC_SCHWARZSCHILD, 0.0, 50.0, 90.0, 60.0, 16, 12);
  s.disk.kind = GG_DISK_NONE; s.integrator.max_radius = 200.0; s.integrator.max_steps = 5000;
  render(s, "hittable scene", true, 0, 0);
  s.disk.kind = GG_DISK_VOLUMETRIC; s.disk.r_inner = 6.0; s.disk.scale_height = 1.0; s.disk.density_0 = 1.0;
  s.disk.temperature_0 = 1.0; s.disk.absorption = 0.1;
  render(s, "hittable scene with volumetric disk", true, 0, 0);
  expect(oracle_set_hittables(0, nullptr) == 0, "clear hittables");
  // HDRI sky
  std::vector<float> sky(8 * 4 * 3);
  for (size_t i = 0; i < sky.size(); ++i) sky[i] = 0.01f * (float)(i % 37);
  expect(oracle_set_sky_hdri(8, 4, sky.data(), 1.5f) == 0, "set hdri");
  s = make_scene(GG_METRIC_KERR, 0.7, 40.0, 70.0, 30.0, 10, 8);
  render(s, "hdri sky", true, 1, 0);
  expect(oracle_set_sky_hdri(0, 0, nullptr, 1.0f) == 0, "clear hdri");
  // a metric known only by its tensor
  oracle_set_user_metric(rn_metric, 1, 1.0 + std::sqrt(1.0 - 0.25));
  s = make_scene(GG_METRIC_USER, 0.0, 40.0, 80.0, 30.0, 8, 8);
  s.disk.r_inner = 6.0; s.integrator.max_radius = 200.0;
  render(s, "tensor metric (Reissner-Nordstrom)", false, 1, 1);
  oracle_set_user_metric(nullptr, 0, 0.0);
  // caller states: valid, NaN, r <= 0, inside the horizon, beyond escape, and out of attempts
  s = make_scene(GG_METRIC_KERR, 0.9, 100.0, 60.0, 25.0, 1, 1);
  const double nan = std::nan("");
  const double states[6][9] = {{0, 30.0, 1.2, 0.3, 1.0, -0.9, 0.01, 0.002, 0},
                               {0, nan, 1.2, 0.3, 1.0, -0.9, 0.01, 0.002, 0},
                               {0, -1.0, 1.2, 0.3, 1.0, -0.9, 0.01, 0.002, 0},
                               {0, 1.2, 1.2, 0.3, 1.0, -0.9, 0.01, 0.002, 0},
                               {0, 2000.0, 1.2, 0.3, 1.0, 0.9, 0.0, 0.0, 0},
                               {0, 80.0, 1.5, 0.0, 1.0, -0.5, 0.0, 0.001, 0}};
  std::vector<double> flat(6 * 9);
  for (int i = 0; i < 6; ++i)
    for (int c = 0; c < 9; ++c) flat[9 * i + c] = states[i][c];
  {
    Out o(6, 0);
    expect(oracle_trace_states(&s, 6, flat.data(), &o.o, 2) == 0, "trace_states");
    expect(o.retired(), "trace_states retired");
    s.integrator.max_steps = 3;
    Out o3(6, 0);
    expect(oracle_trace_states(&s, 6, flat.data(), &o3.o, 1) == 0, "trace_states, 3 attempts");
    expect(o3.cls[5] == GG_CLASS_MAX_STEPS, "trace_states runs out of attempts");
  }
  // metric functions at a few points
  gg_metric km{GG_METRIC_KERR, 0, 1.0, 0.7};
  const double pos[8] = {0, 10.0, M_PI / 3, 0.2, 0, 4.0, 0.4, 1.0};
  const double vel[8] = {1.0, -0.5, 0.01, 0.003, 1.2, 0.3, -0.02, 0.05};
  double gam[64], acc[8], jac[2 * 4 * 7], com[6];
  expect(oracle_christoffel(&km, pos, gam) == 0, "christoffel");
  expect(oracle_geodesic_acceleration(&km, 2, pos, vel, acc) == 0, "geodesic acceleration");
  expect(oracle_rhs_jacobian(&km, 2, pos, vel, jac) == 0, "rhs jacobian");
  expect(oracle_constants_of_motion(&km, 1, flat.data(), com) == 0, "constants of motion");
  // the differentiable render: dual-number gradient, decision hashes, frozen-step loss
  gg_grad_scene g;
  std::memset(&g, 0, sizeof g);
  g.mass = 1.0; g.spin = 0.5; g.inclination = 45.0 * M_PI / 180.0; g.emissivity_q = 3.0;
  g.r_inner = oracle_r_isco(1.0, 0.5) + 0.5; g.r_outer = 20.0; g.edge_width = 0.25; g.r_camera = 50.0;
  g.fov_y_radians = 30.0 * M_PI / 180.0; g.integrator = rtol_integrator(200.0, 20000); g.width = 8; g.height = 8;
  std::vector<double> target(64, 0.01), inten(64), dI(64 * 4);
  std::vector<uint64_t> dec(64);
  double loss = 0.0, grad[4] = {0, 0, 0, 0};
  expect(oracle_grad(&g, target.data(), &loss, grad, inten.data(), dI.data(), 2) == 0, "grad");
  expect(std::isfinite(loss) && loss > 0.0, "grad loss");
  expect(oracle_grad_decisions(&g, dec.data(), 2) == 0, "grad decisions");
  gg_grad_scene g1 = g;
  g1.spin = 0.55;
  expect(oracle_loss_frozen(&g, &g1, target.data(), &loss, inten.data(), 2) == 0, "frozen loss");
  g.frozen_steps = 1;
  expect(oracle_grad(&g, target.data(), &loss, grad, inten.data(), nullptr, 1) == 0, "grad, frozen steps");

  if (g_fail) return 1;
  std::printf("oracle sanitize: ok\n");
  return 0;
}
