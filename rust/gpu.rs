//! libgargantua (MI355X) backend for the curved-spacetime renderer.
//!
//! The reference's side of the C ABI in `include/gargantua.h`: a module a
//! maintainer drops into `crates/raytracer/src/curved/gpu.rs` behind a `gpu`
//! feature, with `pub mod gpu;` in `curved/mod.rs` and the link line
//! `println!("cargo:rustc-link-search=native=<dir of libgargantua.so>")` in a
//! `build.rs`.  `render` and `render_with_disk` (renderer.rs:82-101) keep
//! their signatures: `render_inner` (renderer.rs:256-306) takes the GPU path
//! through `render_rgb` when the metric answers `gpu_descriptor()` (a default
//! method added to `Metric`, metric.rs:11; Kerr and Schwarzschild override
//! it), and the CPU path otherwise.  `render_gradient` is the gradient entry
//! point of the differentiable render, which has no counterpart in the
//! reference.
//!
//! Every struct is `#[repr(C)]` and mirrors its `gg_*` typedef field by field
//! (a CPU test of this repository checks names, order and offsets against
//! gcc's layout of the header); every `extern "C"` function is one the
//! header declares.
use std::ffi::{CStr, CString};
use std::os::raw::{c_char, c_int, c_void};

// ---- data structures (gargantua.h) ----------------------------------------

/// gg_metric: 0 Schwarzschild, 1 Kerr, 2 user tensor (kerr.rs:17-34, schwarzschild.rs:7-18).
#[repr(C)]
#[derive(Clone, Copy, Default, Debug)]
pub struct GgMetric { pub kind: i32, pub _pad: i32, pub mass: f64, pub spin: f64 }

/// gg_camera (camera.rs:9-15).
#[repr(C)]
#[derive(Clone, Copy, Default, Debug)]
pub struct GgCamera { pub position: [f64; 4], pub look_at: [f64; 3], pub up: [f64; 3], pub fov_y_radians: f64,
                      pub aspect: f64 }

/// gg_disk: 0 none, 1 volumetric (disk.rs:8-31), 2 thin.
#[repr(C)]
#[derive(Clone, Copy, Default, Debug)]
pub struct GgDisk { pub kind: i32, pub _pad: i32, pub r_inner: f64, pub r_outer: f64, pub scale_height: f64,
                    pub density_0: f64, pub temperature_0: f64, pub absorption: f64, pub emissivity_q: f64,
                    pub edge_width: f64 }

/// gg_integrator: RK45Integrator (geodesic.rs:155-181) plus the rtol/atol error mode.
#[repr(C)]
#[derive(Clone, Copy, Default, Debug)]
pub struct GgIntegrator { pub initial_step: f64, pub min_step: f64, pub max_step: f64, pub tolerance: f64,
                          pub max_radius: f64, pub horizon_buffer: f64, pub horizon_rel: f64, pub rtol: f64,
                          pub atol: f64, pub err_mode: i32, pub max_steps: u32 }

/// gg_image: pixels, spa^2 sub-pixels (renderer.rs:326-342), optional jitter.
#[repr(C)]
#[derive(Clone, Copy, Default, Debug)]
pub struct GgImage { pub width: u32, pub height: u32, pub samples_per_axis: u32, pub jitter: u32, pub seed: u64 }

/// gg_scene: everything one render call needs.
#[repr(C)]
#[derive(Clone, Copy, Default, Debug)]
pub struct GgScene { pub metric: GgMetric, pub camera: GgCamera, pub disk: GgDisk, pub integrator: GgIntegrator,
                     pub image: GgImage }

/// gg_outputs: caller-owned structure-of-arrays outputs, host or device; any may be null.
#[repr(C)]
#[derive(Clone, Copy, Debug)]
pub struct GgOutputs { pub cls: *mut u8, pub r: *mut f64, pub theta: *mut f64, pub phi: *mut f64,
                       pub intensity: *mut f64, pub attempts: *mut u32, pub accepted: *mut u32,
                       pub state: *mut f64, pub rgb: *mut u8, pub color: *mut f64 }

impl Default for GgOutputs {
    fn default() -> Self {
        let n = std::ptr::null_mut();
        GgOutputs { cls: n as *mut u8, r: n as *mut f64, theta: n as *mut f64, phi: n as *mut f64,
                    intensity: n as *mut f64, attempts: n as *mut u32, accepted: n as *mut u32,
                    state: n as *mut f64, rgb: n as *mut u8, color: n as *mut f64 }
    }
}

/// gg_hittable: one Sphere (sphere.rs) or Quad (quad.rs) with its material.
#[repr(C)]
#[derive(Clone, Copy, Default, Debug)]
pub struct GgHittable { pub shape: i32, pub material: i32, pub q: [f64; 3], pub u: [f64; 3], pub v: [f64; 3],
                        pub radius: f64, pub color: [f64; 3], pub fuzz: f64, pub ior: f64 }

/// gg_stats: statistics of one render call.
#[repr(C)]
#[derive(Clone, Copy, Default, Debug)]
pub struct GgStats { pub kernel_ms: f64, pub total_ms: f64, pub rays: u64, pub attempts: u64, pub rhs_evals: u64,
                     pub rhs_evals_fp32: u64, pub rhs_evals_fast: u64, pub attempts_full: u64,
                     pub reflections: u64, pub grid_blocks: u32, pub block_threads: u32 }

/// gg_launch: launch shape of the persistent trace kernel (0 = library default).
#[repr(C)]
#[derive(Clone, Copy, Default, Debug)]
pub struct GgLaunch { pub block_threads: u32, pub blocks_per_cu: u32, pub refill_interval: u32, pub graph: u32,
                      pub order: u32 }

/// gg_grad_scene: the differentiable thin-disk render.
#[repr(C)]
#[derive(Clone, Copy, Default, Debug)]
pub struct GgGradScene { pub mass: f64, pub spin: f64, pub inclination: f64, pub emissivity_q: f64,
                         pub r_inner: f64, pub r_outer: f64, pub edge_width: f64, pub r_camera: f64,
                         pub fov_y_radians: f64, pub integrator: GgIntegrator, pub width: u32, pub height: u32,
                         pub precision: i32, pub frozen_steps: u32 }

/// gg_grad_stats: statistics of one gradient call.
#[repr(C)]
#[derive(Clone, Copy, Default, Debug)]
pub struct GgGradStats { pub forward_ms: f64, pub adjoint_ms: f64, pub total_ms: f64, pub rays: u64,
                         pub disk_rays: u64, pub steps: u64, pub attempts: u64, pub checkpoints: u64,
                         pub checkpoint_bytes: u64, pub segment_bytes: u64, pub record_bytes: u64,
                         pub grid_blocks: u32, pub block_threads: u32, pub checkpoint_every: u32, pub mode: u32 }

/// Opaque gg_context.
pub enum GgContext {}

pub const GG_PREC_FP64: c_int = 0;
pub const GG_PREC_MIXED: c_int = 1;
pub const GG_PREC_FP64_FAST: c_int = 2;
pub const GG_ABI_VERSION: c_int = 1;

// ---- entry points ------------------------------------------------------------

#[link(name = "gargantua")]
extern "C" {
    pub fn gg_strerror(status: c_int) -> *const c_char;
    pub fn gg_last_error() -> *const c_char;
    pub fn gg_abi_version() -> c_int;
    pub fn gg_context_create(device: c_int, stream: *mut c_void, out: *mut *mut GgContext) -> c_int;
    pub fn gg_context_destroy(ctx: *mut GgContext) -> c_int;
    pub fn gg_set_launch(ctx: *mut GgContext, launch: *const GgLaunch) -> c_int;
    pub fn gg_get_launch(ctx: *mut GgContext, launch: *mut GgLaunch) -> c_int;
    pub fn gg_scene_rays(scene: *const GgScene, n_rays: *mut u64) -> c_int;
    pub fn gg_render(ctx: *mut GgContext, scene: *const GgScene, precision: c_int, out: *const GgOutputs,
                     stats: *mut GgStats) -> c_int;
    pub fn gg_render_batch(ctx: *mut GgContext, scenes: *const GgScene, n: u32, precision: c_int,
                           out: *const GgOutputs, stats: *mut GgStats) -> c_int;
    pub fn gg_render_batch_noalloc(ctx: *mut GgContext, scenes: *const GgScene, n: u32, precision: c_int,
                                   out: *const GgOutputs, stats: *mut GgStats) -> c_int;
    pub fn gg_autotune(ctx: *mut GgContext, scenes: *const GgScene, n: u32, precision: c_int,
                       candidates: *const GgLaunch, n_candidates: u32, best: *mut GgLaunch,
                       best_kernel_ms: *mut f64) -> c_int;
    pub fn gg_trace_states(ctx: *mut GgContext, scene: *const GgScene, precision: c_int, n: u64,
                           states_in: *const f64, out: *const GgOutputs, stats: *mut GgStats) -> c_int;
    pub fn gg_geodesic_acceleration(ctx: *mut GgContext, metric: *const GgMetric, precision: c_int, n: u64,
                                    pos: *const f64, vel: *const f64, acc: *mut f64) -> c_int;
    pub fn gg_camera_rays(ctx: *mut GgContext, scene: *const GgScene, states: *mut f64) -> c_int;
    pub fn gg_constants_of_motion(ctx: *mut GgContext, metric: *const GgMetric, n: u64, states: *const f64,
                                  out: *mut f64) -> c_int;
    pub fn gg_grad_scene_to_scene(grad_scene: *const GgGradScene, scene: *mut GgScene) -> c_int;
    pub fn gg_grad(ctx: *mut GgContext, scene: *const GgGradScene, target: *const f64, checkpoint_every: u32,
                   loss: *mut f64, grad: *mut f64, intensity: *mut f64, d_intensity: *mut f64,
                   stats: *mut GgGradStats) -> c_int;
    pub fn gg_set_sky_hdri(ctx: *mut GgContext, width: u32, height: u32, pixels: *const f32,
                           intensity: f32) -> c_int;
    pub fn gg_set_hittables(ctx: *mut GgContext, n: u32, objects: *const GgHittable) -> c_int;
    pub fn gg_set_user_metric(ctx: *mut GgContext, source: *const c_char, event_horizon: f64) -> c_int;
    pub fn gg_check_user_metric(source: *const c_char, disk_mode: i32, log: *mut c_char, log_size: usize) -> c_int;
    pub fn gg_set_grad_memory(ctx: *mut GgContext, bytes: u64) -> c_int;
    pub fn gg_geodesic_acceleration_vjp(ctx: *mut GgContext, metric: *const GgMetric, n: u64, pos: *const f64,
                                        vel: *const f64, cot: *const f64, pos_bar: *mut f64, vel_bar: *mut f64,
                                        spin_bar: *mut f64) -> c_int;
}

// ---- safe wrappers -------------------------------------------------------------

/// A non-zero status as `Err(message)`: gg_strerror(status): gg_last_error().
pub fn check(rc: c_int) -> Result<(), String> {
    if rc == 0 {
        return Ok(());
    }
    unsafe {
        Err(format!("{}: {}", CStr::from_ptr(gg_strerror(rc)).to_string_lossy(),
                    CStr::from_ptr(gg_last_error()).to_string_lossy()))
    }
}

thread_local! {
    // One library context per thread: the reference renders from rayon's
    // pool, and contexts are independent.
    static CTX: std::cell::RefCell<Option<*mut GgContext>> = std::cell::RefCell::new(None);
}

/// Runs f with this thread's context (created on device 0 on first use).
pub fn with_ctx<T>(f: impl FnOnce(*mut GgContext) -> Result<T, String>) -> Result<T, String> {
    CTX.with(|c| {
        let mut c = c.borrow_mut();
        if c.is_none() {
            if unsafe { gg_abi_version() } != GG_ABI_VERSION {
                return Err("libgargantua ABI mismatch".into());
            }
            let mut h = std::ptr::null_mut();
            check(unsafe { gg_context_create(0, std::ptr::null_mut(), &mut h) })?;
            *c = Some(h);
        }
        f(c.unwrap())
    })
}

/// The scene `render_inner` (renderer.rs:256-306) traces: the reference
/// integrator defaults with max_radius 200 and 5000 attempts
/// (renderer.rs:264-265), the volumetric disk when one is given.
pub fn scene_for(metric: GgMetric, camera: &crate::curved::Camera, disk: Option<&crate::curved::AccretionDisk>,
                 width: u32, height: u32, samples_per_axis: u32) -> GgScene {
    let p = &camera.position;
    let mut s = GgScene::default();
    s.metric = metric;
    s.camera = GgCamera { position: [p[0], p[1], p[2], p[3]],
                          look_at: [camera.look_at.x, camera.look_at.y, camera.look_at.z],
                          up: [camera.up.x, camera.up.y, camera.up.z],
                          fov_y_radians: camera.fov_y_radians, aspect: camera.aspect };
    if let Some(d) = disk {
        s.disk = GgDisk { kind: 1, _pad: 0, r_inner: d.r_inner, r_outer: d.r_outer, scale_height: d.scale_height,
                          density_0: d.density_0, temperature_0: d.temperature_0, absorption: 0.1,
                          emissivity_q: 0.0, edge_width: 0.0 };
    }
    let i = gr_core::RK45Integrator { max_radius: 200.0, ..gr_core::RK45Integrator::default() };
    s.integrator = GgIntegrator { initial_step: i.initial_step, min_step: i.min_step, max_step: i.max_step,
                                  tolerance: i.tolerance, max_radius: i.max_radius,
                                  horizon_buffer: i.horizon_buffer, horizon_rel: 0.0, rtol: 0.0, atol: 0.0,
                                  err_mode: 0, max_steps: 5000 };
    s.image = GgImage { width, height, samples_per_axis: samples_per_axis.max(1), jitter: 0, seed: 0 };
    s
}

/// The RgbImage of render_row (renderer.rs:309-345): the library traces and
/// shades (shade_outcome_with, disk_color, encode_srgb) on the GPU, with the
/// procedural sky or the map of set_sky_hdri.
pub fn render_rgb(scene: &GgScene, precision: c_int) -> Result<image::RgbImage, String> {
    let (w, h) = (scene.image.width, scene.image.height);
    let mut rgb = vec![0u8; (w as usize) * (h as usize) * 3];
    let out = GgOutputs { rgb: rgb.as_mut_ptr(), ..GgOutputs::default() };
    with_ctx(|ctx| check(unsafe { gg_render(ctx, scene, precision, &out, std::ptr::null_mut()) }))?;
    Ok(image::RgbImage::from_raw(w, h, rgb).expect("w*h*3 bytes"))
}

/// RenderOptions.sky = HdriSky (sky.rs:43-122) for the following renders of
/// this thread's context; None restores the procedural sky.
pub fn set_sky_hdri(texels: Option<(u32, u32, &[[f32; 3]], f32)>) -> Result<(), String> {
    with_ctx(|ctx| match texels {
        Some((w, h, px, k)) => check(unsafe { gg_set_sky_hdri(ctx, w, h, px.as_ptr() as *const f32, k) }),
        None => check(unsafe { gg_set_sky_hdri(ctx, 0, 0, std::ptr::null(), 1.0) }),
    })
}

/// The primitives of a Hittable scene (render_with_scene,
/// renderer.rs:103-254) for the following renders; an empty slice clears them.
pub fn set_hittables(objects: &[GgHittable]) -> Result<(), String> {
    with_ctx(|ctx| check(unsafe { gg_set_hittables(ctx, objects.len() as u32, objects.as_ptr()) }))
}

/// A metric known only by its tensor, as device source (metric kind 2).
pub fn set_user_metric(source: &str, event_horizon: Option<f64>) -> Result<(), String> {
    let c = CString::new(source).map_err(|e| e.to_string())?;
    with_ctx(|ctx| check(unsafe { gg_set_user_metric(ctx, c.as_ptr(), event_horizon.unwrap_or(f64::NAN)) }))
}

/// Loss L = mean (I - target)^2 of the FP64 thin-disk image of `scene` and
/// dL/d(spin, inclination, emissivity q, r_in) by the library's reverse-mode
/// adjoint (step sizes differentiated through the controller unless
/// scene.frozen_steps is set).  `target` holds width*height intensities,
/// row-major; `checkpoint_every = 0` lets the library keep the whole
/// trajectory when it fits and checkpoint otherwise.
pub fn render_gradient(scene: &GgGradScene, target: &[f64], checkpoint_every: u32)
                       -> Result<(f64, [f64; 4], GgGradStats), String> {
    let n = (scene.width as usize) * (scene.height as usize);
    if target.len() != n {
        return Err(format!("target has {} pixels, scene {}", target.len(), n));
    }
    let (mut loss, mut grad, mut st) = (0.0f64, [0.0f64; 4], GgGradStats::default());
    with_ctx(|ctx| check(unsafe {
        gg_grad(ctx, scene, target.as_ptr(), checkpoint_every, &mut loss, grad.as_mut_ptr(), std::ptr::null_mut(),
                std::ptr::null_mut(), &mut st)
    }))?;
    Ok((loss, grad, st))
}
