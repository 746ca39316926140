/*
 * gargantua.h -- C ABI of the MI355X geodesic renderer (libgargantua.so).
 *
 * The boundary replaces the per-pixel null-geodesic hot path of the reference
 * renderer (Rust crates gr-core + raytracer::curved).  Every struct is plain
 * old data; every pointer is caller-owned and is never retained after a call
 * returns.  No C++ or torch type crosses the boundary.
 *
 * Reference interfaces each entry point replaces:
 *   gg_render            raytracer::curved::render / render_with_disk
 *                        (crates/raytracer/src/curved/renderer.rs:82-101, driver
 *                        render_inner/render_row :256-345) and the per-ray loops
 *                        trace_ray / trace_ray_with_disk
 *                        (crates/raytracer/src/curved/tracer.rs:54-152).
 *   gg_trace_states      trace_ray / trace_ray_with_disk on caller-supplied
 *                        GeodesicState values (tracer.rs:54-87; state type
 *                        crates/gr-core/src/geodesic.rs:13-20).
 *   gg_geodesic_acceleration
 *                        Metric::geodesic_acceleration, batched
 *                        (crates/gr-core/src/metric.rs:37-52, Kerr via
 *                        kerr.rs:100-424, Schwarzschild override
 *                        schwarzschild.rs:99-140).
 *   gg_render_batch      many independent render() calls (one per scene) fused
 *                        into one persistent launch; no single reference
 *                        counterpart, the reference calls render() per image.
 *   gg_grad              no reference counterpart: gradient of an image loss
 *                        through the thin-disk render (reverse mode, hand
 *                        written adjoint with checkpoints).
 *   gg_grad_scene_to_scene, gg_geodesic_acceleration_vjp
 *                        no reference counterpart: the scene gg_grad renders
 *                        and the RHS vector-Jacobian product it uses.
 *   gg_camera_rays       GeodesicRay::from_camera (ray.rs:14-61) per pixel.
 *   gg_constants_of_motion
 *                        no reference counterpart: E, L_z, Carter Q, mu^2.
 *
 * Errors: every entry point returns GG_OK (0) or a negative gg_status; the
 * message of the last failure on the calling thread is gg_last_error().
 * Invalid arguments never reach a kernel.  Where the reference panics
 * (Kerr::new with mass <= 0, kerr.rs:30) the ABI returns GG_ERR_INVALID_ARG.
 *
 * Threading: calls on distinct contexts may run concurrently.  Calls that
 * share one context are serialised by the caller.
 */
#ifndef GARGANTUA_H
#define GARGANTUA_H

#ifndef __HIPCC_RTC__   /* hipRTC supplies the fixed-width types itself */
#include <stddef.h>
#include <stdint.h>
#endif

#ifdef __cplusplus
extern "C" {
#endif

#define GG_ABI_VERSION 1

typedef enum gg_status {
  GG_OK = 0,
  GG_ERR_INVALID_ARG = -1,
  GG_ERR_DEVICE = -2,        /* HIP runtime error (message in gg_last_error) */
  GG_ERR_NO_DEVICE = -3,     /* no HIP device visible                          */
  GG_ERR_OUT_OF_MEMORY = -4,
  GG_ERR_UNSUPPORTED = -5,
  GG_ERR_INTERNAL = -6
} gg_status;

/* ---- metric (crates/gr-core/src/kerr.rs:17-34, schwarzschild.rs:7-18) ---- */
typedef enum gg_metric_kind {
  GG_METRIC_SCHWARZSCHILD = 0,
  GG_METRIC_KERR = 1,
  GG_METRIC_USER = 2        /* the context's gg_set_user_metric tensor; mass and
                               spin are passed to it unchanged               */
} gg_metric_kind;

typedef struct gg_metric {
  int32_t kind;     /* gg_metric_kind                                         */
  int32_t _pad;
  double mass;      /* M > 0                                                  */
  double spin;      /* a; clamped to [-0.9999 M, 0.9999 M] like Kerr::new    */
} gg_metric;

/* ---- camera (crates/raytracer/src/curved/camera.rs:9-15) ---- */
typedef struct gg_camera {
  double position[4];   /* (t, r, theta, phi) in the metric chart             */
  double look_at[3];    /* camera-local Cartesian frame                       */
  double up[3];
  double fov_y_radians;
  double aspect;
} gg_camera;

/* ---- disk (crates/raytracer/src/curved/disk.rs:8-31) ---- */
typedef enum gg_disk_kind {
  GG_DISK_NONE = 0,        /* trace_ray: sky / horizon only (tracer.rs:54-70)  */
  GG_DISK_VOLUMETRIC = 1,  /* trace_ray_with_disk (tracer.rs:80-152)           */
  GG_DISK_THIN = 2         /* thin equatorial disk, I = r^-q g^4, terminate on
                              crossing (no reference counterpart)              */
} gg_disk_kind;

typedef struct gg_disk {
  int32_t kind;            /* gg_disk_kind                                    */
  int32_t _pad;
  double r_inner;          /* AccretionDisk::r_inner (default 6)              */
  double r_outer;          /* AccretionDisk::r_outer (default 20)             */
  double scale_height;     /* volumetric: Gaussian H (default 1)              */
  double density_0;        /* volumetric: rho_0 (default 1)                   */
  double temperature_0;    /* volumetric: T_0 (default 1)                     */
  double absorption;       /* volumetric: 0.1 in tracer.rs:91                 */
  double emissivity_q;     /* thin: I = r_c^-q g^4                            */
  double edge_width;       /* thin: 0 = hard [r_inner, r_outer]; w > 0 = smooth
                              logistic edges of width w (differentiable mode) */
} gg_disk;

/* ---- integrator (crates/gr-core/src/geodesic.rs:155-181) ---- */
typedef enum gg_err_mode {
  GG_ERR_REFERENCE = 0,    /* accept iff ||dv||_2 < tolerance (geodesic.rs:308-325) */
  GG_ERR_RTOL_ATOL = 1     /* accept iff RMS(dv_i / (atol + rtol max|v_i|)) < 1    */
} gg_err_mode;

typedef struct gg_integrator {
  double initial_step;     /* 0.1                                             */
  double min_step;         /* 1e-8                                            */
  double max_step;         /* 1.0                                             */
  double tolerance;        /* 1e-6 (GG_ERR_REFERENCE)                          */
  double max_radius;       /* escape: r > max_radius (renderer uses 200)      */
  double horizon_buffer;   /* capture: r < r_h (1 + horizon_rel) + buffer     */
  double horizon_rel;      /* 0 in the reference                              */
  double rtol;             /* GG_ERR_RTOL_ATOL                                 */
  double atol;
  int32_t err_mode;        /* gg_err_mode                                     */
  uint32_t max_steps;      /* attempts per ray, as tracer.rs:61 counts them   */
} gg_integrator;

/* ---- image ---- */
typedef struct gg_image {
  uint32_t width;
  uint32_t height;
  uint32_t samples_per_axis; /* RenderOptions::samples_per_axis (>= 1)        */
  uint32_t jitter;           /* 0: pixel centres (ray.rs:22-23); 1: centre +
                                U[-0.5,0.5)^2 from SplitMix64(seed, ray)      */
  uint64_t seed;
} gg_image;

/* A scene is everything one render() call needs. */
typedef struct gg_scene {
  gg_metric metric;
  gg_camera camera;
  gg_disk disk;
  gg_integrator integrator;
  gg_image image;
} gg_scene;

typedef enum gg_precision {
  GG_PREC_FP64 = 0,       /* every operation in FP64, reference expression order */
  GG_PREC_MIXED = 1,      /* RHS stages in FP32 where the step-size controller is
                             saturated and FP32 rounding cannot change a decision,
                             fast FP64 where its rounding cannot change one, exact
                             FP64 elsewhere (tiers chosen per wavefront)          */
  GG_PREC_FP64_FAST = 2   /* FP64 with reciprocal-based RHS algebra and FMA      */
} gg_precision;

/* Ray classes (RayOutcome, crates/raytracer/src/curved/outcome.rs:3-22;
 * StepResult::Singular folds into Horizon as tracer.rs:65 does). */
typedef enum gg_ray_class {
  GG_CLASS_HORIZON = 0,
  GG_CLASS_ESCAPED = 1,
  GG_CLASS_DISK = 2,
  GG_CLASS_MAX_STEPS = 3,
  GG_CLASS_SCENE = 4      /* RayOutcome::Scene: a chord met a gg_hittable      */
} gg_ray_class;

/* Per-ray outputs, structure of arrays.  N = (width*spa) * (height*spa),
 * ray index = sub_py * (width*spa) + sub_px.  Any pointer may be NULL.
 * Pointers may address host memory or device memory of the context's device
 * (gg_render detects which). */
typedef struct gg_outputs {
  uint8_t *cls;         /* [N] gg_ray_class                                   */
  double *r;            /* [N] final r                                        */
  double *theta;        /* [N] final theta                                    */
  double *phi;          /* [N] final phi                                      */
  double *intensity;    /* [N] disk intensity (0 if none)                     */
  uint32_t *attempts;   /* [N] RK45 step() calls                              */
  uint32_t *accepted;   /* [N] accepted steps                                 */
  double *state;        /* [9*N] SoA planes: t r th ph vt vr vth vph lambda   */
  uint8_t *rgb;         /* [3*width*height] reference shading + sRGB          */
  double *color;        /* [3*N] linear RGB of GG_CLASS_SCENE rays (0 for
                           other classes); written only while hittables are
                           set (gg_set_hittables)                             */
} gg_outputs;

/* ---- Hittable scene (crates/raytracer/src/{shape,material}) ---- */
typedef enum gg_shape {
  GG_SHAPE_SPHERE = 0,    /* Sphere::new(center = q, radius, material)        */
  GG_SHAPE_QUAD = 1       /* Quad::new(q, u, v, material)                      */
} gg_shape;

typedef enum gg_material {
  GG_MATERIAL_LAMBERTIAN = 0,    /* Lambertian::new(albedo = color)            */
  GG_MATERIAL_METAL = 1,         /* Metal::new(albedo = color, fuzz); fuzz 0 only */
  GG_MATERIAL_DIELECTRIC = 2,    /* Dielectric::new(ior): white attenuation     */
  GG_MATERIAL_DIFFUSE_LIGHT = 3  /* DiffuseLight::new(emit = color)             */
} gg_material;

typedef struct gg_hittable {
  int32_t shape;          /* gg_shape                                          */
  int32_t material;       /* gg_material                                       */
  double q[3];            /* sphere centre | quad corner                       */
  double u[3];            /* quad edge u                                       */
  double v[3];            /* quad edge v                                       */
  double radius;          /* sphere radius (> 0)                               */
  double color[3];        /* albedo, or the emission of a diffuse light        */
  double fuzz;            /* metal (clamped to [0, 1] as Metal::new does)      */
  double ior;             /* dielectric (does not enter the one-bounce shading) */
} gg_hittable;

/* Launch statistics (filled when the pointer is non-NULL). */
typedef struct gg_stats {
  double kernel_ms;         /* hipEvent time of the trace kernel(s)           */
  double total_ms;          /* whole call incl. copies                        */
  uint64_t rays;
  uint64_t attempts;        /* sum of step() calls                            */
  uint64_t rhs_evals;       /* RHS evaluations actually made                  */
  uint64_t rhs_evals_fp32;  /* of which evaluated in FP32 (mixed)             */
  uint64_t rhs_evals_fast;  /* of which evaluated in fast FP64 (mixed)        */
  uint64_t attempts_full;   /* attempts that reached the error estimate       */
  uint64_t reflections;     /* accepted steps whose theta crossed a pole and
                               was reflected (normalize_spherical_coords,
                               geodesic.rs:337-352)                           */
  uint32_t grid_blocks;
  uint32_t block_threads;
} gg_stats;

/* Launch shape of the persistent trace kernel. 0 = library default. */
typedef struct gg_launch {
  uint32_t block_threads;     /* 64..256, multiple of 64 (default 64)         */
  uint32_t blocks_per_cu;     /* blocks launched per CU; 0 (default) = by ray
                                 count: when every lane at 8 wavefronts per CU
                                 gets at least 3 rays, as many wavefronts as
                                 the kernel's registers allow (12 per CU for
                                 the disk modes, 8 for the scene modes and
                                 recording renders, 4 for user metrics),
                                 else 4 (one per SIMD)                         */
  uint32_t refill_interval;   /* attempts between retired-lane refills (64)   */
  uint32_t graph;             /* gg_render_batch: 1 = capture the launch in a
                                 hipGraph and replay it on later calls with the
                                 same shape (default), 2 = launch directly     */
  uint32_t order;             /* gg_render of one scene: 0 = rays in 8x8-tile
                                 order (default); 1 = longest first, by the
                                 per-ray attempt counts of the previous render
                                 with the same sub-pixel grid (tile order for
                                 the first).  Every ray is traced in full
                                 either way; FP64 outputs are identical, mixed
                                 outputs agree to FP32 rounding.              */
} gg_launch;

typedef struct gg_context gg_context;

const char *gg_strerror(int status);
const char *gg_last_error(void);
int gg_abi_version(void);

/* stream: a hipStream_t of the device, or NULL for a library-owned stream. */
int gg_context_create(int device, void *stream, gg_context **out);
int gg_context_destroy(gg_context *ctx);
int gg_set_launch(gg_context *ctx, const gg_launch *launch);
int gg_get_launch(gg_context *ctx, gg_launch *launch);

/* Rays traced by gg_render for a scene (validates the scene). */
int gg_scene_rays(const gg_scene *scene, uint64_t *n_rays);

int gg_render(gg_context *ctx, const gg_scene *scene, int precision,
              const gg_outputs *out, gg_stats *stats);

/* n scenes in one persistent launch; out[i] belongs to scenes[i].  All scenes
 * share metric kind and disk kind.  The launch (queue reset + trace kernel) is
 * captured in a hipGraph cached on the context and replayed by later calls
 * with the same number of scenes, kinds, precision and launch shape. */
int gg_render_batch(gg_context *ctx, const gg_scene *scenes, uint32_t n,
                    int precision, const gg_outputs *out, gg_stats *stats);

/* No reference counterpart: for callers that cannot allocate device memory
 * (code under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  gg_render_batch for at most GG_NOALLOC_MAX_SCENES scenes
 * that allocates nothing: the scenes' constants, their output tables and
 * the trace kernel's counters live in one static device array per GPU
 * (GG_NOALLOC_SCRATCH_BYTES, held until the process ends) instead of
 * context buffers that grow on demand, and the launch is direct (no graph,
 * tile order: gg_launch.order and .graph are not used).  Every output
 * pointer must be device memory (host outputs would need staging); rgb and
 * color outputs, user metrics and hittables are not supported
 * (GG_ERR_UNSUPPORTED); more than GG_NOALLOC_MAX_SCENES scenes is
 * GG_ERR_INVALID_ARG.  The counters are cleared by a kernel at the start of
 * each call.  All calls on one GPU share its array: calls that overlap on
 * one GPU (from several threads or contexts) race and must be serialised by
 * the caller.  The results are those of gg_render_batch. */
#define GG_NOALLOC_MAX_SCENES 64
#define GG_NOALLOC_SCRATCH_BYTES (64u * 1024u)
int gg_render_batch_noalloc(gg_context *ctx, const gg_scene *scenes, uint32_t n, int precision,
                            const gg_outputs *out, gg_stats *stats);

/* Times the batch (outputs discarded) under each candidate launch shape and
 * installs the fastest on the context (each candidate renders twice, the
 * second launch is timed).  candidates == NULL: a built-in grid of block
 * sizes, blocks per CU and refill intervals in tile order, then the best of
 * them with the learned ray order (order = 1).  The graph setting of the
 * context is kept.  No reference counterpart (the reference schedules rows
 * with rayon). */
int gg_autotune(gg_context *ctx, const gg_scene *scenes, uint32_t n, int precision,
                const gg_launch *candidates, uint32_t n_candidates, gg_launch *best,
                double *best_kernel_ms);

/* Trace n caller-supplied states (9 doubles each, AoS: t r th ph vt vr vth vph
 * lambda).  Camera and image fields of the scene are ignored except for the
 * observer position used by disk redshift (camera.position). */
int gg_trace_states(gg_context *ctx, const gg_scene *scene, int precision,
                    uint64_t n, const double *states_in, const gg_outputs *out,
                    gg_stats *stats);

/* acc[4*i..] = -Gamma^mu_ab v^a v^b at (pos[4*i..], vel[4*i..]); AoS. */
int gg_geodesic_acceleration(gg_context *ctx, const gg_metric *metric,
                             int precision, uint64_t n, const double *pos,
                             const double *vel, double *acc);

/* GeodesicRay::from_camera (crates/raytracer/src/curved/ray.rs:14-61) for
 * every sub-pixel ray of the scene: states[9*i..] = (t r th ph vt vr vth vph
 * lambda=0), i = sub_py*(width*spa) + sub_px (with jitter when set). */
int gg_camera_rays(gg_context *ctx, const gg_scene *scene, double *states);

/* Constants of motion of n states (9 doubles each, AoS, as gg_trace_states
 * takes them): out[4*i..] = (E = -p_t, L_z = p_phi, Carter Q, mu^2 =
 * -g(v,v)).  No reference counterpart: diagnostics of the integrator's
 * drift (exact along geodesics; forward-Euler positions drift at O(h)). */
int gg_constants_of_motion(gg_context *ctx, const gg_metric *metric, uint64_t n,
                           const double *states, double *out);

/* ---- differentiable thin-disk render (no reference counterpart) ---- */
typedef struct gg_grad_scene {
  double mass;
  double spin;             /* a      (differentiated)                        */
  double inclination;      /* i, rad (differentiated)                        */
  double emissivity_q;     /* q      (differentiated)                        */
  double r_inner;          /* r_in   (differentiated)                        */
  double r_outer;
  double edge_width;       /* > 0: logistic edge width                       */
  double r_camera;
  double fov_y_radians;
  gg_integrator integrator;
  uint32_t width;
  uint32_t height;
  int32_t precision;       /* forward render: GG_PREC_FP64 (0, default) or
                              GG_PREC_MIXED, which needs checkpoint_every = 0
                              and room for the recorded steps (the mixed
                              kernel's steps depend on the lane grouping, so
                              only the render's own records are its path) */
  uint32_t frozen_steps;   /* 0 (default): the step sizes carry derivatives
                              through the error estimate and the step-size
                              controller (geodesic.rs:308-325): the gradient
                              of the loss the render computes.  1: each
                              ray's step sizes are held at the render's
                              values (the discrete map with its step
                              sequence fixed)                                */
} gg_grad_scene;

/* The gg_scene gg_grad renders: Kerr(mass, spin), camera at
 * (t, r, theta, phi) = (0, r_camera, inclination, 0) looking at -r_hat with
 * up = +z, aspect width/height, the thin disk (r_inner, r_outer, q,
 * edge_width), one ray per pixel centre.  Render it with gg_render at
 * GG_PREC_FP64 for the image gg_grad differentiates. */
int gg_grad_scene_to_scene(const gg_grad_scene *grad_scene, gg_scene *scene);

/* Statistics of one gg_grad call. */
typedef struct gg_grad_stats {
  double forward_ms;        /* trace kernel of the loss render                */
  double adjoint_ms;        /* adjoint kernel                                 */
  double total_ms;          /* whole call                                     */
  uint64_t rays;            /* width*height                                   */
  uint64_t disk_rays;       /* rays differentiated (class GG_CLASS_DISK)      */
  uint64_t steps;           /* step records (attempts) back-propagated        */
  uint64_t attempts;        /* attempts re-run by the adjoint (re-trace +
                               segment recomputation)                         */
  uint64_t checkpoints;     /* checkpoints stored                             */
  uint64_t checkpoint_bytes;
  uint64_t segment_bytes;   /* per-thread segment buffers of the reverse pass */
  uint64_t record_bytes;    /* step records stored (256 bytes each, one per
                               attempt): room for each disk ray's last segment
                               by the re-trace (modes 0, 1), or every ray's
                               attempts in 4224-byte chunks of 16 by the
                               render (mode 2)                                */
  uint32_t grid_blocks;      /* reverse kernel launch                        */
  uint32_t block_threads;
  uint32_t checkpoint_every; /* interval used (mode 1: at least the longest
                               ray's accepted count; mode 2: 0)               */
  uint32_t mode;             /* 0: checkpoints, segments recomputed; 1: every
                               step record stored by a re-trace of the disk
                               rays; 2: recorded by the render itself         */
} gg_grad_stats;

/* Loss L = mean_p (I_p - target_p)^2 over the width*height pixel-centre rays
 * of gg_grad_scene_to_scene(scene), rendered in FP64.  grad[0..3] =
 * dL/d(spin, inclination, emissivity_q, r_inner) of that discrete algorithm
 * (reverse mode): each attempt's step size is the controller's function of
 * the previous step size and error estimate (geodesic.rs:308-325) and
 * carries its derivative, or, with scene->frozen_steps, is held at the
 * render's value; the accept/reject pattern, the clamps of h and the
 * disk-crossing step are the render's.  Checkpoints every checkpoint_every
 * accepted steps, segments recomputed between them; the result does not
 * depend on checkpoint_every.  checkpoint_every = 0: the render records
 * every attempt of every ray (256 bytes each, in chunks of 16) and nothing
 * is re-traced or recomputed, when the records fit in the memory of
 * gg_set_grad_memory (default a quarter of the free device memory); else a
 * re-trace stores every disk ray's step records when those fit, else
 * checkpoints every 32 accepted steps.  An interval at least the longest
 * ray's accepted count also stores every step record (re-trace).
 * Only rays of class GG_CLASS_DISK carry a gradient.
 * target: [N] host or device.  Optional outputs: intensity [N] (the image),
 * d_intensity [4*N] (dI_p/d(spin, inclination, q, r_inner), AoS).
 * Requires |spin| < 0.9999 mass (the clamp of Kerr::new has no derivative),
 * 0 < inclination < pi, checkpoint_every <= 65536. */
int gg_grad(gg_context *ctx, const gg_grad_scene *scene, const double *target,
            uint32_t checkpoint_every, double *loss, double grad[4],
            double *intensity, double *d_intensity, gg_grad_stats *stats);

/* Equirectangular environment map for the escaped rays of every following
 * render with rgb output on this context (RenderOptions.sky = HdriSky,
 * sky.rs:43-122): width x height linear RGB f32 texels, row-major, row 0 the
 * +z pole, columns wrapping in azimuth; bilinear with wrap-x / clamp-y,
 * scaled by intensity.  pixels (host or device) are copied.  width = 0
 * restores the procedural starfield (ProceduralSky, the reference's
 * default); this call is the library's counterpart of RenderOptions.sky.
 * An invalid argument leaves the previous sky; a device error (allocation,
 * copy) leaves the procedural one. */
int gg_set_sky_hdri(gg_context *ctx, uint32_t width, uint32_t height,
                    const float *pixels, float intensity);

/* Hittable geometry for every following gg_render / gg_trace_states on this
 * context: render_with_scene (disk kind NONE) and render_with_disk_and_scene
 * (disk kind VOLUMETRIC), crates/raytracer/src/curved/renderer.rs:103-254,
 * tracer.rs:207-364.  Each accepted step's Cartesian chord is tested against
 * the n primitives (the library builds its own BVH; the closest hit is what
 * the reference's BvhNode/World returns); a hit ends the ray as
 * GG_CLASS_SCENE with its linear colour in gg_outputs.color (one-bounce
 * shade_hit x static g^4, plus the disk emission in front of it).  n = 0
 * clears the scene.  GG_ERR_UNSUPPORTED for a thin disk or a batch while
 * hittables are set, and for a metal with fuzz > 0 (its reference shading
 * draws a random number).  An invalid primitive leaves the previous
 * geometry; a device error (allocation, copy) leaves none. */
int gg_set_hittables(gg_context *ctx, uint32_t n, const gg_hittable *objects);

/* A metric known only by its tensor, for scenes of kind GG_METRIC_USER on
 * this context: a Metric that implements metric_tensor alone
 * (crates/gr-core/src/metric.rs:11-80).  `source` is HIP device code that
 * defines
 *     __device__ void gg_user_metric(const double x[4], double mass,
 *                                    double spin, double g[16]);
 * filling g[4*mu + nu] = g_mu,nu at x = (t, r, theta, phi) (any helper
 * functions it needs may precede it).  The tensor is symmetric: the
 * Christoffel symbols read its derivatives from the upper triangle
 * (mu <= nu), so g_mu,nu and g_nu,mu must hold the same value; a render
 * whose tensor is not symmetric at the camera fails with GG_ERR_INVALID_ARG.  A source may also
 * `#define GG_METRIC_STATIONARY` (g does not depend on t) and/or
 * `#define GG_METRIC_AXISYMMETRIC` (g does not depend on phi): the two
 * tensor evaluations of that coordinate's central difference, which are
 * then exactly zero, are skipped, with bit-identical results.  The library compiles it at run time
 * (hipRTC) into its trace kernel; the Christoffel symbols are the reference's
 * central differences (numerical_christoffel, metric.rs:97-142, h = 1e-6) of
 * this tensor and its 4x4 inverse.  event_horizon is Metric::event_horizon:
 * a radius, or NaN for None (no horizon test).  The emitter of disk light is
 * static (orbital_four_velocity = None).  Renders of such scenes take
 * precision GG_PREC_FP64 and one scene per call.  A compile error returns
 * GG_ERR_INVALID_ARG with the compiler log in gg_last_error().  NULL source
 * clears the metric. */
int gg_set_user_metric(gg_context *ctx, const char *source, double event_horizon);

/* Compile a gg_set_user_metric source for disk mode 0-4 (none, volumetric,
 * thin, hittables, volumetric + hittables) without loading it: no device is
 * needed.  The compiler log (warnings too) is copied into log (log_size
 * bytes, NUL-terminated) when log is non-NULL.  No counterpart in the
 * reference; for build-time checks of a tensor source. */
int gg_check_user_metric(const char *source, int32_t disk_mode, char *log, size_t log_size);

/* Device memory gg_grad's automatic mode (checkpoint_every = 0) may use for
 * step records; 0 (the default): a quarter of the free device memory.  The
 * choice of mode follows it; the gradient does not depend on it. */
int gg_set_grad_memory(gg_context *ctx, uint64_t bytes);

/* Vector-Jacobian product of gg_geodesic_acceleration for Kerr (FP64):
 * for each point, cot[4*i..] = w, pos_bar[4*i..] = w^T da/dpos (t and phi
 * components 0), vel_bar[4*i..] = w^T da/dvel, spin_bar[i] = w^T da/dspin.
 * The building block of gg_grad's adjoint. */
int gg_geodesic_acceleration_vjp(gg_context *ctx, const gg_metric *metric,
                                 uint64_t n, const double *pos,
                                 const double *vel, const double *cot,
                                 double *pos_bar, double *vel_bar,
                                 double *spin_bar);

#ifdef __cplusplus
}
#endif
#endif /* GARGANTUA_H */
