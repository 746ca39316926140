// gg_scene.hpp -- Hittable geometry met by the chords of a curved photon path.
//
// trace_ray_with_scene / trace_ray_with_disk_and_scene (tracer.rs:207-364)
// map every accepted step to the straight Cartesian chord between its end
// points and ask the scene for the closest hit with chord parameter in
// (0, 1).  Primitives are the reference's Sphere (shape/sphere.rs:53-88)
// and Quad (shape/quad.rs:75-127); materials enter only through the albedo
// of one `scatter` call and their emission (shade_hit, tracer.rs:369-384).
// The host builds a bounding-volume hierarchy over the primitives (flat
// array in depth-first order, leaves of up to four primitives), walked
// without a stack; the closest hit it returns is the minimum over
// primitives of each primitive's own first root in the open interval, so it
// equals the reference's BVH answer whatever the tree.
#pragma once

#include "gg_physics.hpp"
#include "gg_user_metric.hpp"

namespace gg {

enum : uint32_t { SHAPE_SPHERE = 0, SHAPE_QUAD = 1 };

// One primitive as the kernels read it (host-precomputed, 160 bytes).
struct SceneObj {
  double a[3];        // sphere centre | quad corner Q
  double n[3];        // quad unit normal; sphere: n[0] = radius
  double ea[3], eb[3];// quad e_alpha = v x w, e_beta = w x u (quad.rs:29-31)
  double d;           // quad plane offset normal . Q
  double albedo[3];   // attenuation of the material's scatter
  double emit[3];     // material emission
  uint32_t shape;
  uint32_t metal;     // mirror: the albedo only when the reflection leaves the surface (metal.rs:40-57)
};

// BVH node (32 bytes) in depth-first order: an internal node's first child
// is the next node; `skip` is the first node after the node's subtree (the
// root's skip is the node count).  Leaf: primitives [first, first + count)
// with leaf = 8 first + count, count 1..4; an internal node has leaf = 0.
// The box is FP32, padded outward by 1e-5 (1 + |coordinate|): far more than
// the FP32 rounding of the slab test, so the test is conservative (it never
// culls a primitive the FP64 hit test would find) at half the FP64 cost.
struct BvhNode {
  float lo[3], hi[3];
  int32_t leaf;
  int32_t skip;
};

namespace {

__device__ inline double sdot3(const double a[3], const double b[3]) {
#pragma clang fp contract(off)
  return a[0] * b[0] + a[1] * b[1] + a[2] * b[2];           // nalgebra dot, U3: (a + b) + c
}

// spherical_to_cartesian (tracer.rs:12-18).  sincos shares one argument
// reduction between the sine and the cosine and returns the values sin and
// cos return.
__device__ inline void sph2cart(const double p[4], double c[3]) {
#pragma clang fp contract(off)
  double st, ct, sp, cp;
  sincos_cr(p[2], &st, &ct);   // the host libm's values (correctly rounded)
  sincos_cr(p[3], &sp, &cp);
  const double r = p[1];
  c[0] = r * st * cp;
  c[1] = r * st * sp;
  c[2] = r * ct;
}

// Slab test of [lo, hi] against origin + t dir, t in [0, t_max], in FP32.
// A chord parallel to a slab whose origin lies on its plane gives 0 * inf:
// that axis then does not constrain t.
__device__ inline bool box_hit(const BvhNode &b, const float o[3], const float inv[3], float t_max) {
  float t0 = 0.0f, t1 = t_max;
  for (int k = 0; k < 3; ++k) {
    const float ta = (b.lo[k] - o[k]) * inv[k], tb = (b.hi[k] - o[k]) * inv[k];
    float tl = fminf(ta, tb), th = fmaxf(ta, tb);
    if (ta != ta || tb != tb) { tl = -__builtin_inff(); th = __builtin_inff(); }
    t0 = fmaxf(t0, tl);
    t1 = fminf(t1, th);
  }
  return t0 <= t1;
}

// Sphere::hit (sphere.rs:53-88) / Quad::hit (quad.rs:75-127) for t in
// (0, t_max): the root, the hit point and the outward normal.
__device__ inline bool hit_object(const SceneObj &s, const double o[3], const double dir[3], double t_max, double &t,
                                  double p[3], double nrm[3]) {
#pragma clang fp contract(off)
  if (s.shape == SHAPE_SPHERE) {
    const double r = s.n[0];
    const double oc[3] = {o[0] - s.a[0], o[1] - s.a[1], o[2] - s.a[2]};
    const double a = sdot3(dir, dir);
    const double half_b = sdot3(oc, dir);
    const double c = sdot3(oc, oc) - r * r;
    const double disc = half_b * half_b - a * c;
    if (disc < 0.0) return false;
    const double sq = sqrt(disc);
    double root = (-half_b - sq) / a;
    if (!(0.0 < root && root < t_max)) {
      root = (-half_b + sq) / a;
      if (!(0.0 < root && root < t_max)) return false;
    }
    t = root;
    for (int k = 0; k < 3; ++k) p[k] = o[k] + dir[k] * root;
    for (int k = 0; k < 3; ++k) nrm[k] = (p[k] - s.a[k]) / r;
    return true;
  }
  const double denom = sdot3(s.n, dir);
  if (fabs(denom) < 1e-8) return false;
  const double tq = (s.d - sdot3(s.n, o)) / denom;
  if (!(0.0 < tq && tq < t_max)) return false;
  double q[3];
  for (int k = 0; k < 3; ++k) q[k] = o[k] + dir[k] * tq;
  const double pl[3] = {q[0] - s.a[0], q[1] - s.a[1], q[2] - s.a[2]};
  const double al = sdot3(pl, s.ea), be = sdot3(pl, s.eb);
  if (al < 0.0 || al > 1.0 || be < 0.0 || be > 1.0) return false;
  t = tq;
  for (int k = 0; k < 3; ++k) { p[k] = q[k]; nrm[k] = s.n[k]; }
  return true;
}

struct SceneHit {
  double t;
  double p[3];
  double nrm[3];     // facing against the chord (HitRecord::new, core/hit.rs:26-48)
  uint32_t obj;
};

// Closest primitive hit by the chord origin + t dir, t in (0, 1): a
// stackless depth-first walk (every step moves to a later node, so it ends
// after at most the node count).
__device__ inline bool scene_hit(const SceneObj *objs, const BvhNode *nodes, const double o[3], const double dir[3],
                                 SceneHit &h) {
#pragma clang fp contract(off)
  const float of[3] = {(float)o[0], (float)o[1], (float)o[2]};
  const float inv[3] = {1.0f / (float)dir[0], 1.0f / (float)dir[1], 1.0f / (float)dir[2]};
  double t_max = 1.0;
  float t_max_f = 1.0001f;                 // above every t_max the boxes can be asked about
  bool any = false;
  const int32_t end = nodes[0].skip;
  int32_t i = 0;
  while (i < end) {
    const BvhNode &nd = nodes[i];
    if (!box_hit(nd, of, inv, t_max_f)) {
      i = nd.skip;
      continue;
    }
    if (nd.leaf == 0) {
      ++i;                                   // descend: the first child follows
      continue;
    }
    const int32_t first = nd.leaf >> 3, last = first + (nd.leaf & 7);
    for (int32_t k = first; k < last; ++k) {
      double t, p[3], nr[3];
      if (hit_object(objs[k], o, dir, t_max, t, p, nr)) {
        t_max = t;
        t_max_f = (float)t * 1.0001f + 1e-6f;  // rounded up: the FP32 box test stays conservative
        any = true;
        h.t = t;
        h.obj = (uint32_t)k;
        for (int c = 0; c < 3; ++c) { h.p[c] = p[c]; h.nrm[c] = nr[c]; }
      }
    }
    i = nd.skip;
  }
  if (any) {
    // front_face = dir . outward < 0; the normal faces against the chord
    if (!(sdot3(dir, h.nrm) < 0.0))
      for (int k = 0; k < 3; ++k) h.nrm[k] = -h.nrm[k];
  }
  return any;
}

// shade_hit (tracer.rs:369-384): albedo x (0.5 + 0.5 max(-d.n, 0)) + emission.
__device__ inline void shade_hit(const SceneObj &s, const double dir[3], const double nrm[3], double lin[3]) {
#pragma clang fp contract(off)
  const double len = sqrt(sdot3(dir, dir));
  const double ud[3] = {dir[0] / len, dir[1] / len, dir[2] / len};
  double alb[3] = {s.albedo[0], s.albedo[1], s.albedo[2]};
  if (s.metal) {
    // reflect(unit d, n) = d - (2 d.n) n, scattered only above the surface
    const double dn = sdot3(ud, nrm);
    const double rf[3] = {ud[0] - 2.0 * dn * nrm[0], ud[1] - 2.0 * dn * nrm[1], ud[2] - 2.0 * dn * nrm[2]};
    if (!(sdot3(rf, nrm) > 0.0)) alb[0] = alb[1] = alb[2] = 0.0;
  }
  const double mud[3] = {-ud[0], -ud[1], -ud[2]};
  double nd = sdot3(mud, nrm);
  nd = nd > 0.0 ? nd : 0.0;                                     // f64::max(nd, 0.0)
  const double facing = 0.5 + 0.5 * nd;
  for (int k = 0; k < 3; ++k) lin[k] = alb[k] * facing + s.emit[k];
}

// redshift_g4_static (tracer.rs:40-52) at cartesian_to_spherical(point)
// (tracer.rs:23-28): (g_tt(emit) / g_tt(obs))^2, 0 where either is <= 0.
template <int METRIC>
__device__ inline double g4_static(const DevScene &S, const double pt[3]) {
#pragma clang fp contract(off)
  double r = sqrt(pt[0] * pt[0] + pt[1] * pt[1] + pt[2] * pt[2]);
  r = r >= 1e-10 ? r : 1e-10;
  double zr = pt[2] / r;
  zr = zr < -1.0 ? -1.0 : (zr > 1.0 ? 1.0 : zr);
  const double th = acos(zr);
  double gte;
  if constexpr (METRIC == 2) {
    const double x[4] = {0.0, r, th, atan2(pt[1], pt[0])};
    gte = -user_gmn(S, x).tt;
  } else {
    gte = -metric64<METRIC>(S, r, th).tt;
  }
  const double gto = S.g_tt_obs;
  if (gte <= 0.0 || gto <= 0.0) return 0.0;
  const double g = sqrt(gte / gto);
  const double g2 = g * g;
  return g2 * g2;
}

}  // namespace

}  // namespace gg
