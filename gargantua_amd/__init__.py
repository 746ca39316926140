"""gargantua_amd -- MI355X geodesic renderer behind the reference's render API.

Host-side mirror of the reference interface on top of libgargantua.so (the
C ABI of include/gargantua.h):

  Schwarzschild, Kerr            crates/gr-core/src/{schwarzschild,kerr}.rs
  RK45Integrator                 crates/gr-core/src/geodesic.rs:155-181
  Camera                         crates/raytracer/src/curved/camera.rs:9-15
  AccretionDisk                  crates/raytracer/src/curved/disk.rs:8-31
  RenderOptions                  crates/raytracer/src/curved/renderer.rs:17-26
  HdriSky                        crates/raytracer/src/curved/sky.rs:43-122
  Sphere, Quad, World            crates/raytracer/src/{shape/sphere.rs,shape/quad.rs,scene/world.rs}
  Lambertian, Metal, Dielectric, crates/raytracer/src/material/*.rs
  DiffuseLight
  render_with_scene,             crates/raytracer/src/curved/renderer.rs:103-254
  render_with_disk_and_scene
  render, render_with_disk       crates/raytracer/src/curved/renderer.rs:82-101
  Renderer.trace_states          trace_ray / trace_ray_with_disk (tracer.rs:54-152)
  Renderer.geodesic_acceleration Metric::geodesic_acceleration (metric.rs:37-52)

Every computation runs in the HIP library; this module only builds the POD
structs and numpy buffers.  If libgargantua.so is missing it raises.
"""
from __future__ import annotations

import ctypes as C
import math
from dataclasses import dataclass, field
from typing import Optional, Sequence

import numpy as np

from . import _abi
from . import configs  # noqa: F401  (scene builders for the benchmark configurations)
from . import hdr, png  # noqa: F401  (Radiance .hdr input, PNG output)
from ._abi import (GG_CLASS_DISK, GG_CLASS_ESCAPED, GG_CLASS_HORIZON, GG_CLASS_MAX_STEPS, GG_CLASS_SCENE,  # noqa: F401
                   GG_DISK_NONE, GG_DISK_THIN, GG_DISK_VOLUMETRIC, GG_ERR_INVALID_ARG, GG_ERR_REFERENCE,
                   GG_ERR_RTOL_ATOL, GG_ERR_UNSUPPORTED, GG_METRIC_KERR, GG_METRIC_SCHWARZSCHILD, GG_OK)

PRECISIONS = {"fp64": _abi.GG_PREC_FP64, "mixed": _abi.GG_PREC_MIXED, "fp64fast": _abi.GG_PREC_FP64_FAST}


class GargantuaError(RuntimeError):
    """A libgargantua call returned a negative status."""


# ---------------------------------------------------------------- metrics --

@dataclass
class Schwarzschild:
    """Schwarzschild::new(mass) (schwarzschild.rs:13-18)."""
    mass: float = 1.0

    def to_c(self) -> _abi.gg_metric:
        return _abi.gg_metric(kind=GG_METRIC_SCHWARZSCHILD, mass=self.mass, spin=0.0)

    def event_horizon(self) -> float:
        return 2.0 * self.mass


@dataclass
class Kerr:
    """Kerr::new(mass, spin); the library clamps |spin| to 0.9999 M (kerr.rs:29-34)."""
    mass: float = 1.0
    spin: float = 0.9

    def __post_init__(self):
        if not self.mass > 0.0:
            raise ValueError("Kerr: mass must be positive")  # kerr.rs:30 panics

    def to_c(self) -> _abi.gg_metric:
        return _abi.gg_metric(kind=GG_METRIC_KERR, mass=self.mass, spin=self.spin)

    def event_horizon(self) -> float:
        a = max(-0.9999 * self.mass, min(0.9999 * self.mass, self.spin))
        return self.mass + math.sqrt(self.mass * self.mass - a * a)


# ------------------------------------------------------- scene components --

@dataclass
class UserMetric:
    """A metric known only by its tensor (a Metric implementing metric_tensor alone,
    crates/gr-core/src/metric.rs:11-80).

    `source` is HIP device code defining
        __device__ void gg_user_metric(const double x[4], double mass, double spin, double g[16])
    that fills g[4*mu + nu] = g_mu,nu at x = (t, r, theta, phi); mass and spin are passed through.  The
    library compiles it at run time into its trace kernel; Christoffel symbols are central differences
    (metric.rs:97-142).  event_horizon is Metric::event_horizon (None: no horizon test).
    """
    source: str
    mass: float = 1.0
    spin: float = 0.0
    event_horizon: Optional[float] = None

    def to_c(self) -> _abi.gg_metric:
        return _abi.gg_metric(kind=_abi.GG_METRIC_USER, mass=self.mass, spin=self.spin)


@dataclass
class RK45Integrator:
    """RK45Integrator defaults (geodesic.rs:170-181) plus the rtol/atol mode."""
    initial_step: float = 0.1
    min_step: float = 1e-8
    max_step: float = 1.0
    tolerance: float = 1e-6
    max_radius: float = 100.0
    horizon_buffer: float = 0.01
    horizon_rel: float = 0.0
    rtol: float = 0.0
    atol: float = 0.0
    err_mode: int = GG_ERR_REFERENCE
    max_steps: int = 5000

    def to_c(self) -> _abi.gg_integrator:
        return _abi.gg_integrator(initial_step=self.initial_step, min_step=self.min_step, max_step=self.max_step,
                                  tolerance=self.tolerance, max_radius=self.max_radius,
                                  horizon_buffer=self.horizon_buffer, horizon_rel=self.horizon_rel,
                                  rtol=self.rtol, atol=self.atol, err_mode=self.err_mode,
                                  max_steps=self.max_steps)


@dataclass
class Camera:
    """Camera (camera.rs:9-15): position (t, r, theta, phi), look_at, up, fov_y, aspect."""
    position: Sequence[float] = (0.0, 30.0, math.pi / 2.0, 0.0)
    look_at: Sequence[float] = (-1.0, 0.0, 0.0)
    up: Sequence[float] = (0.0, 0.0, 1.0)
    fov_y_radians: float = math.radians(60.0)
    aspect: float = 1.0

    def to_c(self) -> _abi.gg_camera:
        c = _abi.gg_camera()
        c.position[:] = [float(x) for x in self.position]
        c.look_at[:] = [float(x) for x in self.look_at]
        c.up[:] = [float(x) for x in self.up]
        c.fov_y_radians = self.fov_y_radians
        c.aspect = self.aspect
        return c


@dataclass
class AccretionDisk:
    """AccretionDisk defaults (disk.rs:21-31); kind selects the disk model."""
    r_inner: float = 6.0
    r_outer: float = 20.0
    scale_height: float = 1.0
    density_0: float = 1.0
    temperature_0: float = 1.0
    absorption: float = 0.1
    emissivity_q: float = 3.0
    edge_width: float = 0.0
    kind: int = GG_DISK_VOLUMETRIC

    def to_c(self) -> _abi.gg_disk:
        return _abi.gg_disk(kind=self.kind, r_inner=self.r_inner, r_outer=self.r_outer,
                            scale_height=self.scale_height, density_0=self.density_0,
                            temperature_0=self.temperature_0, absorption=self.absorption,
                            emissivity_q=self.emissivity_q, edge_width=self.edge_width)


class HdriSky:
    """HdriSky (sky.rs:43-122): equirectangular linear-RGB environment map.

    Texel (0, 0) is the +z pole; rows run to the -z pole and columns wrap in
    azimuth.  The map is sampled on the GPU (bilinear in f32, wrap-x /
    clamp-y, times intensity) for every escaped ray of a render with rgb
    output once it is set on a Renderer or passed in RenderOptions.sky.
    """

    def __init__(self, pixels: np.ndarray, intensity: float = 1.0):
        px = np.ascontiguousarray(pixels, dtype=np.float32)
        if px.ndim != 3 or px.shape[2] != 3 or px.shape[0] == 0 or px.shape[1] == 0:
            raise ValueError(f"HdriSky pixels must be (height, width, 3), got {px.shape}")
        self.pixels = px
        self.intensity = float(intensity)

    @property
    def width(self) -> int:
        return self.pixels.shape[1]

    @property
    def height(self) -> int:
        return self.pixels.shape[0]

    @staticmethod
    def from_pixels(width: int, height: int, pixels) -> "HdriSky":
        """HdriSky::from_pixels (sky.rs:72-77): row-major [r, g, b] texels, row 0 = +z pole."""
        px = np.asarray(pixels, dtype=np.float32)
        if px.size != width * height * 3:
            raise ValueError("pixel buffer size mismatch")
        return HdriSky(px.reshape(height, width, 3))

    @staticmethod
    def from_file(path) -> "HdriSky":
        """HdriSky::from_file (sky.rs:57-70) for Radiance .hdr files."""
        from . import hdr
        return HdriSky(hdr.read(path))

    def with_intensity(self, intensity: float) -> "HdriSky":
        """HdriSky::with_intensity (sky.rs:79-83)."""
        return HdriSky(self.pixels, intensity)


# ---- Hittable scene: materials, shapes, world ----

@dataclass
class Lambertian:
    """Lambertian (material/lambertian.rs:9-18): scatters with attenuation `albedo`."""
    albedo: Sequence[float]


@dataclass
class Metal:
    """Metal (material/metal.rs:7-19); fuzz is clamped to [0, 1] and only 0 is supported."""
    albedo: Sequence[float]
    fuzz: float = 0.0


@dataclass
class Dielectric:
    """Dielectric (material/dielectric.rs): white attenuation; ior does not enter one-bounce shading."""
    ior: float = 1.5


@dataclass
class DiffuseLight:
    """DiffuseLight (material/diffuse_light.rs:6-20): emits `emit`, scatters nothing."""
    emit: Sequence[float]

    @staticmethod
    def white(intensity: float) -> "DiffuseLight":
        return DiffuseLight((intensity, intensity, intensity))


def _material_fields(m) -> tuple:
    if isinstance(m, Lambertian):
        return _abi.GG_MATERIAL_LAMBERTIAN, m.albedo, 0.0, 1.0
    if isinstance(m, Metal):
        return _abi.GG_MATERIAL_METAL, m.albedo, m.fuzz, 1.0
    if isinstance(m, Dielectric):
        return _abi.GG_MATERIAL_DIELECTRIC, (1.0, 1.0, 1.0), 0.0, m.ior
    if isinstance(m, DiffuseLight):
        return _abi.GG_MATERIAL_DIFFUSE_LIGHT, m.emit, 0.0, 1.0
    raise TypeError(f"unsupported material {type(m).__name__}")


@dataclass
class Sphere:
    """Sphere::new(center, radius, material) (shape/sphere.rs:17-27)."""
    center: Sequence[float]
    radius: float
    material: object

    def to_c(self) -> _abi.gg_hittable:
        mk, col, fuzz, ior = _material_fields(self.material)
        h = _abi.gg_hittable(shape=_abi.GG_SHAPE_SPHERE, material=mk, radius=float(self.radius), fuzz=fuzz, ior=ior)
        h.q[:] = [float(x) for x in self.center]
        h.color[:] = [float(x) for x in col]
        return h


@dataclass
class Quad:
    """Quad::new(q, u, v, material) (shape/quad.rs:19-72): the parallelogram q + a u + b v, a, b in [0, 1]."""
    q: Sequence[float]
    u: Sequence[float]
    v: Sequence[float]
    material: object

    def to_c(self) -> _abi.gg_hittable:
        mk, col, fuzz, ior = _material_fields(self.material)
        h = _abi.gg_hittable(shape=_abi.GG_SHAPE_QUAD, material=mk, fuzz=fuzz, ior=ior)
        h.q[:] = [float(x) for x in self.q]
        h.u[:] = [float(x) for x in self.u]
        h.v[:] = [float(x) for x in self.v]
        h.color[:] = [float(x) for x in col]
        return h


class World:
    """World (scene/world.rs:5-60): a list of primitives.  build_bvh() returns the world itself, since
    the library builds its own hierarchy when the world is set on a Renderer."""

    def __init__(self, objects=None):
        self.objects = list(objects or [])

    def add(self, obj) -> "World":
        self.objects.append(obj)
        return self

    def __len__(self) -> int:
        return len(self.objects)

    def build_bvh(self) -> "World":
        return self

    def to_c(self):
        arr = (_abi.gg_hittable * len(self.objects))(*[o.to_c() for o in self.objects])
        return arr


@dataclass
class RenderOptions:
    """RenderOptions (renderer.rs:17-26); sky None is the procedural starfield."""
    samples_per_axis: int = 1
    show_progress: bool = False
    sky: Optional[HdriSky] = None


@dataclass
class Scene:
    """Everything one render call needs (gg_scene)."""
    metric: object = field(default_factory=Schwarzschild)
    camera: Camera = field(default_factory=Camera)
    disk: Optional[AccretionDisk] = None
    integrator: RK45Integrator = field(default_factory=RK45Integrator)
    width: int = 32
    height: int = 24
    samples_per_axis: int = 1
    jitter: bool = False
    seed: int = 0

    def to_c(self) -> _abi.gg_scene:
        s = _abi.gg_scene()
        s.metric = self.metric.to_c()
        s.camera = self.camera.to_c()
        s.disk = self.disk.to_c() if self.disk is not None else _abi.gg_disk(kind=GG_DISK_NONE)
        s.integrator = self.integrator.to_c()
        s.image = _abi.gg_image(width=self.width, height=self.height, samples_per_axis=self.samples_per_axis,
                                jitter=1 if self.jitter else 0, seed=self.seed)
        return s

    @property
    def n_rays(self) -> int:
        return self.width * self.samples_per_axis * self.height * self.samples_per_axis

    def to_dict(self) -> dict:
        """Plain-JSON form of the scene (see from_dict)."""
        from dataclasses import asdict
        m = self.metric
        kind = "kerr" if isinstance(m, Kerr) else ("user" if isinstance(m, UserMetric) else "schwarzschild")
        return {"metric": {"kind": kind, **asdict(m)},
                "camera": {k: list(v) if isinstance(v, (tuple, list)) else v for k, v in asdict(self.camera).items()},
                "disk": asdict(self.disk) if self.disk is not None else None,
                "integrator": asdict(self.integrator), "width": self.width, "height": self.height,
                "samples_per_axis": self.samples_per_axis, "jitter": self.jitter, "seed": self.seed}

    @staticmethod
    def from_dict(d: dict) -> "Scene":
        md = dict(d["metric"])
        kind = md.pop("kind")
        metric = Kerr(**md) if kind == "kerr" else (UserMetric(**md) if kind == "user" else Schwarzschild(**md))
        cam = d["camera"]
        return Scene(metric=metric,
                     camera=Camera(position=tuple(cam["position"]), look_at=tuple(cam["look_at"]), up=tuple(cam["up"]),
                                   fov_y_radians=cam["fov_y_radians"], aspect=cam["aspect"]),
                     disk=AccretionDisk(**d["disk"]) if d["disk"] is not None else None,
                     integrator=RK45Integrator(**d["integrator"]), width=d["width"], height=d["height"],
                     samples_per_axis=d["samples_per_axis"], jitter=d["jitter"], seed=d["seed"])


# --------------------------------------------------------------- renderer --

def _check(lib, rc: int, what: str):
    if rc != 0:
        raise GargantuaError(f"{what}: {lib.gg_strerror(rc).decode()} ({rc}): {lib.gg_last_error().decode()}")


def alloc_outputs(n: int, *, state: bool = True, rgb_pixels: int = 0, color: bool = False) -> dict:
    """Host numpy buffers for one scene's per-ray outputs."""
    o = {
        "cls": np.zeros(n, np.uint8), "r": np.zeros(n), "theta": np.zeros(n), "phi": np.zeros(n),
        "intensity": np.zeros(n), "attempts": np.zeros(n, np.uint32), "accepted": np.zeros(n, np.uint32),
    }
    if state:
        o["state"] = np.zeros((9, n))
    if rgb_pixels:
        o["rgb"] = np.zeros((rgb_pixels, 3), np.uint8)
    if color:
        o["color"] = np.zeros((n, 3))
    return o


def outputs_struct(o: dict) -> _abi.gg_outputs:
    """gg_outputs pointing at numpy arrays or at (device) integer addresses."""
    def p(k):
        v = o.get(k)
        if v is None:
            return None
        if isinstance(v, int):
            return v
        assert v.flags["C_CONTIGUOUS"]
        return v.ctypes.data
    return _abi.gg_outputs(cls=p("cls"), r=p("r"), theta=p("theta"), phi=p("phi"), intensity=p("intensity"),
                           attempts=p("attempts"), accepted=p("accepted"), state=p("state"), rgb=p("rgb"),
                           color=p("color"))


def stats_dict(st: _abi.gg_stats) -> dict:
    return {k: getattr(st, k) for k, _ in _abi.gg_stats._fields_}


class Renderer:
    """A libgargantua context (device, stream, device buffers)."""

    def __init__(self, device: int = 0, stream: Optional[int] = None):
        self.lib = _abi.load()
        h = C.c_void_p()
        _check(self.lib, self.lib.gg_context_create(device, C.c_void_p(stream) if stream else None, C.byref(h)),
               "gg_context_create")
        self.ctx = h
        self.n_hittables = 0

    def close(self):
        if self.ctx:
            self.lib.gg_context_destroy(self.ctx)
            self.ctx = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_launch(self, block_threads=0, blocks_per_cu=0, refill_interval=0, graph=0, order=0):
        L = _abi.gg_launch(block_threads, blocks_per_cu, refill_interval, graph, order)
        _check(self.lib, self.lib.gg_set_launch(self.ctx, C.byref(L)), "gg_set_launch")

    def get_launch(self) -> dict:
        L = _abi.gg_launch()
        _check(self.lib, self.lib.gg_get_launch(self.ctx, C.byref(L)), "gg_get_launch")
        return {k: getattr(L, k) for k, _ in _abi.gg_launch._fields_}

    def render(self, scene: Scene, precision: str = "fp64", outputs: Optional[dict] = None, rgb: bool = False,
               state: bool = True) -> dict:
        """Trace every ray of `scene`; returns the per-ray arrays and stats."""
        self._ensure_metric(scene)
        sc = scene.to_c()
        o = outputs if outputs is not None else alloc_outputs(
            scene.n_rays, state=state, rgb_pixels=scene.width * scene.height if rgb else 0,
            color=self.n_hittables > 0)
        st = _abi.gg_stats()
        _check(self.lib, self.lib.gg_render(self.ctx, C.byref(sc), PRECISIONS[precision],
                                            C.byref(outputs_struct(o)), C.byref(st)), "gg_render")
        o["stats"] = stats_dict(st)
        return o

    def render_batch(self, scenes: Sequence[Scene], precision: str = "mixed",
                     outputs: Optional[Sequence[dict]] = None) -> tuple:
        arr = (_abi.gg_scene * len(scenes))(*[s.to_c() for s in scenes])
        outs = list(outputs) if outputs is not None else [alloc_outputs(s.n_rays, state=False) for s in scenes]
        oarr = (_abi.gg_outputs * len(scenes))(*[outputs_struct(o) for o in outs])
        st = _abi.gg_stats()
        _check(self.lib, self.lib.gg_render_batch(self.ctx, arr, len(scenes), PRECISIONS[precision], oarr,
                                                  C.byref(st)), "gg_render_batch")
        return outs, stats_dict(st)

    def autotune(self, scenes: Sequence[Scene], precision: str = "mixed", candidates=None) -> tuple:
        """Time the batch under candidate launch shapes; install and return the fastest."""
        arr = (_abi.gg_scene * len(scenes))(*[s.to_c() for s in scenes])
        if candidates:
            cands = (_abi.gg_launch * len(candidates))(*[_abi.gg_launch(*c) for c in candidates])
            nc = len(candidates)
        else:
            cands, nc = None, 0
        best = _abi.gg_launch()
        ms = C.c_double()
        _check(self.lib, self.lib.gg_autotune(self.ctx, arr, len(scenes), PRECISIONS[precision], cands, nc,
                                              C.byref(best), C.byref(ms)), "gg_autotune")
        return {k: getattr(best, k) for k, _ in _abi.gg_launch._fields_}, ms.value

    def trace_states(self, scene: Scene, states: np.ndarray, precision: str = "fp64") -> dict:
        """trace_ray / trace_ray_with_disk from given states [n, 9] (t r th ph vt vr vth vph lambda)."""
        self._ensure_metric(scene)
        states = np.ascontiguousarray(states, dtype=np.float64)
        n = states.shape[0]
        o = alloc_outputs(n, color=self.n_hittables > 0)
        st = _abi.gg_stats()
        sc = scene.to_c()
        _check(self.lib, self.lib.gg_trace_states(self.ctx, C.byref(sc), PRECISIONS[precision], n,
                                                  states.ctypes.data, C.byref(outputs_struct(o)), C.byref(st)),
               "gg_trace_states")
        o["stats"] = stats_dict(st)
        return o

    def geodesic_acceleration(self, metric, pos: np.ndarray, vel: np.ndarray, precision: str = "fp64"):
        pos = np.ascontiguousarray(pos, dtype=np.float64).reshape(-1, 4)
        vel = np.ascontiguousarray(vel, dtype=np.float64).reshape(-1, 4)
        acc = np.zeros_like(pos)
        m = metric.to_c()
        _check(self.lib, self.lib.gg_geodesic_acceleration(self.ctx, C.byref(m), PRECISIONS[precision], len(pos),
                                                           pos.ctypes.data, vel.ctypes.data, acc.ctypes.data),
               "gg_geodesic_acceleration")
        return acc

    def camera_rays(self, scene: Scene) -> np.ndarray:
        """GeodesicRay::from_camera for every sub-pixel ray: [n, 9] (t r th ph vt vr vth vph lambda)."""
        out = np.zeros((scene.n_rays, 9))
        sc = scene.to_c()
        _check(self.lib, self.lib.gg_camera_rays(self.ctx, C.byref(sc), out.ctypes.data), "gg_camera_rays")
        return out

    def constants_of_motion(self, metric, states: np.ndarray) -> np.ndarray:
        """[n, 4]: E, L_z, Carter Q, mu^2 of states [n, 9] (or [n, 8])."""
        states = np.ascontiguousarray(states, dtype=np.float64)
        if states.shape[1] == 8:
            states = np.ascontiguousarray(np.concatenate([states, np.zeros((len(states), 1))], 1))
        out = np.zeros((len(states), 4))
        m = metric.to_c()
        _check(self.lib, self.lib.gg_constants_of_motion(self.ctx, C.byref(m), len(states), states.ctypes.data,
                                                         out.ctypes.data), "gg_constants_of_motion")
        return out

    def geodesic_acceleration_vjp(self, metric, pos: np.ndarray, vel: np.ndarray, cot: np.ndarray) -> dict:
        """w^T J of the Kerr RHS: {"pos_bar": [n,4], "vel_bar": [n,4], "spin_bar": [n]}."""
        pos = np.ascontiguousarray(pos, dtype=np.float64).reshape(-1, 4)
        vel = np.ascontiguousarray(vel, dtype=np.float64).reshape(-1, 4)
        cot = np.ascontiguousarray(cot, dtype=np.float64).reshape(-1, 4)
        n = len(pos)
        pb, vb, sb = np.zeros((n, 4)), np.zeros((n, 4)), np.zeros(n)
        m = metric.to_c()
        _check(self.lib, self.lib.gg_geodesic_acceleration_vjp(self.ctx, C.byref(m), n, pos.ctypes.data,
                                                               vel.ctypes.data, cot.ctypes.data, pb.ctypes.data,
                                                               vb.ctypes.data, sb.ctypes.data),
               "gg_geodesic_acceleration_vjp")
        return {"pos_bar": pb, "vel_bar": vb, "spin_bar": sb}

    def set_sky_hdri(self, sky: Optional[HdriSky]) -> None:
        """Sky of every following render with rgb output (None: the procedural starfield)."""
        if sky is None:
            _check(self.lib, self.lib.gg_set_sky_hdri(self.ctx, 0, 0, None, 1.0), "gg_set_sky_hdri")
            return
        px = np.ascontiguousarray(sky.pixels, dtype=np.float32)
        _check(self.lib, self.lib.gg_set_sky_hdri(self.ctx, px.shape[1], px.shape[0], px.ctypes.data,
                                                  sky.intensity), "gg_set_sky_hdri")

    def set_hittables(self, world: Optional[World]) -> None:
        """Scene geometry of every following render / trace_states (None or empty: none)."""
        if world is None or len(world) == 0:
            _check(self.lib, self.lib.gg_set_hittables(self.ctx, 0, None), "gg_set_hittables")
            self.n_hittables = 0
            return
        arr = world.to_c()
        rc = self.lib.gg_set_hittables(self.ctx, len(world), C.cast(arr, C.c_void_p))
        if rc not in (GG_OK, GG_ERR_INVALID_ARG, GG_ERR_UNSUPPORTED):
            self.n_hittables = 0            # a device error leaves the context without geometry
        _check(self.lib, rc, "gg_set_hittables")
        self.n_hittables = len(world)

    def set_user_metric(self, metric: Optional[UserMetric]) -> None:
        """Compile `metric.source` for scenes whose metric is a UserMetric (None: clear)."""
        if metric is None:
            _check(self.lib, self.lib.gg_set_user_metric(self.ctx, None, 0.0), "gg_set_user_metric")
            self._user_metric = None
            return
        rh = float("nan") if metric.event_horizon is None else float(metric.event_horizon)
        self._user_metric = None            # the library drops the previous source before compiling
        _check(self.lib, self.lib.gg_set_user_metric(self.ctx, metric.source.encode(), rh), "gg_set_user_metric")
        self._user_metric = (metric.source, metric.event_horizon)

    def _ensure_metric(self, scene) -> None:
        m = scene.metric
        if isinstance(m, UserMetric) and getattr(self, "_user_metric", None) != (m.source, m.event_horizon):
            self.set_user_metric(m)

    def set_grad_memory(self, nbytes: int = 0) -> None:
        """Device memory grad()'s automatic mode may use for step records (0: a quarter of the free memory)."""
        _check(self.lib, self.lib.gg_set_grad_memory(self.ctx, int(nbytes)), "gg_set_grad_memory")

    def grad(self, gscene, target, checkpoint_every: int = 0, want_dI: bool = False,
             want_image: bool = True) -> dict:
        """Loss, dL/d(spin, inclination, q, r_inner) and the image of a gg_grad_scene (see configs.grad_512_params).

        target: [width*height] float64, a numpy array or a contiguous device tensor (anything with
        data_ptr(), numel() and a float64 dtype, e.g. a torch CUDA tensor), which is read in place.
        checkpoint_every = 0: the render records every accepted step and nothing is re-traced (stats mode 2)
        when the records fit the memory of set_grad_memory; else a re-trace stores the disk rays' steps (mode 1)
        or checkpoints every 32 accepted steps (mode 0).  The gradient is the same in every mode."""
        n = gscene.width * gscene.height
        if hasattr(target, "data_ptr"):
            if str(target.dtype) not in ("torch.float64", "float64") or target.numel() != n or \
                    not target.is_contiguous():
                raise ValueError(f"device target must be {n} contiguous float64 values")
            tptr = target.data_ptr()
        else:
            target = np.ascontiguousarray(target, dtype=np.float64).reshape(-1)
            if target.size != n:
                raise ValueError(f"target has {target.size} pixels, scene {n}")
            tptr = target.ctypes.data
        loss = C.c_double()
        g = np.zeros(4)
        inten = np.zeros(n) if want_image else None
        dI = np.zeros((n, 4)) if want_dI else None
        st = _abi.gg_grad_stats()
        _check(self.lib, self.lib.gg_grad(self.ctx, C.byref(gscene), tptr, checkpoint_every,
                                          C.byref(loss), g.ctypes.data, inten.ctypes.data if want_image else None,
                                          dI.ctypes.data if want_dI else None, C.byref(st)), "gg_grad")
        out = {"loss": loss.value, "grad": g, "intensity": inten,
               "stats": {k: getattr(st, k) for k, _ in _abi.gg_grad_stats._fields_}}
        if want_dI:
            out["dI"] = dI
        return out

    def grad_image(self, gscene) -> np.ndarray:
        """The FP64 image gg_grad differentiates (gg_render of gg_grad_scene_to_scene)."""
        sc = _abi.gg_scene()
        _check(self.lib, self.lib.gg_grad_scene_to_scene(C.byref(gscene), C.byref(sc)), "gg_grad_scene_to_scene")
        n = gscene.width * gscene.height
        o = {"intensity": np.zeros(n)}
        st = _abi.gg_stats()
        _check(self.lib, self.lib.gg_render(self.ctx, C.byref(sc), PRECISIONS["fp64"], C.byref(outputs_struct(o)),
                                            C.byref(st)), "gg_render")
        return o["intensity"]

    def fit(self, gscene, target, steps: int, lr: float = 0.02, betas=(0.9, 0.999), log=None) -> dict:
        """Gradient descent on theta = (spin, inclination, q, r_inner) of `gscene` towards the image `target`
        (Adam on the raw parameters, bias-corrected; each step one grad() with records kept by the render).
        The iterate stays inside |spin| <= 0.99, inclination in [0.05, pi - 0.05] and r_inner >= 1.
        Returns theta after each step and the loss before it; `log(it, loss, theta)` is called on steps 1, 20, 40..."""
        import copy
        b1, b2 = betas
        theta = np.array([gscene.spin, gscene.inclination, gscene.emissivity_q, gscene.r_inner], float)
        m = np.zeros(4)
        v = np.zeros(4)
        thetas, losses = [], []
        g = copy.copy(gscene)
        for it in range(1, steps + 1):
            g.spin, g.inclination, g.emissivity_q, g.r_inner = (float(x) for x in theta)
            out = self.grad(g, target, checkpoint_every=0, want_image=False)
            grad = out["grad"]
            m = b1 * m + (1 - b1) * grad
            v = b2 * v + (1 - b2) * grad * grad
            theta = theta - lr * (m / (1 - b1 ** it)) / (np.sqrt(v / (1 - b2 ** it)) + 1e-12)
            theta[0] = min(max(theta[0], -0.99), 0.99)
            theta[1] = min(max(theta[1], 0.05), math.pi - 0.05)
            theta[3] = max(theta[3], 1.0)
            thetas.append(theta.copy())
            losses.append(out["loss"])
            if log is not None and (it == 1 or it % 20 == 0):
                log(it, out["loss"], theta)
        return {"theta": np.array(thetas), "loss": np.array(losses)}


_default_renderer: Optional[Renderer] = None


def _renderer() -> Renderer:
    global _default_renderer
    if _default_renderer is None:
        _default_renderer = Renderer(0)
    return _default_renderer


def _reference_render_scene(metric, camera, disk, width, height, opts) -> Scene:
    # render_inner hard-codes max_radius 200 and 5000 attempts (renderer.rs:264-265).
    integ = RK45Integrator(max_radius=200.0, max_steps=5000)
    return Scene(metric=metric, camera=camera, disk=disk, integrator=integ, width=width, height=height,
                 samples_per_axis=max(1, opts.samples_per_axis if opts else 1))


def render(metric, camera: Camera, width: int, height: int, opts: Optional[RenderOptions] = None) -> np.ndarray:
    """raytracer::curved::render (renderer.rs:82-90): returns an (height, width, 3) uint8 image."""
    opts = opts or RenderOptions()
    sc = _reference_render_scene(metric, camera, None, width, height, opts)
    r = _renderer()
    r.set_sky_hdri(opts.sky)
    o = r.render(sc, "fp64", rgb=True)
    return o["rgb"].reshape(height, width, 3)


def render_with_disk(metric, camera: Camera, disk: AccretionDisk, width: int, height: int,
                     opts: Optional[RenderOptions] = None) -> np.ndarray:
    """raytracer::curved::render_with_disk (renderer.rs:92-101)."""
    opts = opts or RenderOptions()
    sc = _reference_render_scene(metric, camera, disk, width, height, opts)
    r = _renderer()
    r.set_sky_hdri(opts.sky)
    o = r.render(sc, "fp64", rgb=True)
    return o["rgb"].reshape(height, width, 3)


def render_with_scene(metric, camera: Camera, scene: World, width: int, height: int,
                      opts: Optional[RenderOptions] = None) -> np.ndarray:
    """raytracer::curved::render_with_scene (renderer.rs:103-174)."""
    return _render_scene(metric, camera, None, scene, width, height, opts)


def render_with_disk_and_scene(metric, camera: Camera, disk: AccretionDisk, scene: World, width: int, height: int,
                               opts: Optional[RenderOptions] = None) -> np.ndarray:
    """raytracer::curved::render_with_disk_and_scene (renderer.rs:176-254); the disk is volumetric."""
    return _render_scene(metric, camera, disk, scene, width, height, opts)


def _render_scene(metric, camera, disk, scene, width, height, opts) -> np.ndarray:
    opts = opts or RenderOptions()
    sc = _reference_render_scene(metric, camera, disk, width, height, opts)
    r = _renderer()
    r.set_sky_hdri(opts.sky)
    r.set_hittables(scene)
    try:
        o = r.render(sc, "fp64", rgb=True)
    finally:
        r.set_hittables(None)
    return o["rgb"].reshape(height, width, 3)
