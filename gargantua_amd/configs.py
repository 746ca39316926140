"""Scene builders for the benchmark workloads.

Each function returns a `Scene` whose parameters are the ones the workload
names; parameters the workload leaves open are fixed here, once (and listed
in the design notes):

  schwarzschild_256   Schwarzschild M=1, 256x256, camera r=50M, inclination 80 deg,
                      FOV 30 deg, thin disk 6..20M, I = r^-3 g^4, rtol 1e-8 / atol 1e-10,
                      capture r < 1.01 r_h, escape r > 1000M, 20000 attempts.
  kerr_1024           Kerr a=0.9, 1024x1024, camera r=100M, inclination 60 deg, FOV 25 deg,
                      disk r_ISCO..30M, q=3, jitter seed 2026, escape 1000M, 20000 attempts.
  kerr_4096           Kerr a=0.998, 4096x4096, camera r=1000M, inclination 85 deg,
                      FOV 20 deg, disk r_ISCO..30M, q=3, escape 1100M, 50000 attempts.
  batch_64            64 seeded scenes (generator seed 2026), FOV 30 deg, escape 1000M.
  grad_512            differentiable thin disk 512x512, camera r=50M, FOV 30 deg,
                      r_out 20M, logistic edges of width 0.25M.
"""
from __future__ import annotations

import math

import numpy as np


def r_isco(mass: float, spin: float) -> float:
    """Prograde ISCO radius (Bardeen, Press & Teukolsky 1972, eq. 2.21)."""
    a = spin / mass
    z1 = 1.0 + (1.0 - a * a) ** (1.0 / 3.0) * ((1.0 + a) ** (1.0 / 3.0) + (1.0 - a) ** (1.0 / 3.0))
    z2 = math.sqrt(3.0 * a * a + z1 * z1)
    s = 1.0 if a >= 0 else -1.0
    return mass * (3.0 + z2 - s * math.sqrt((3.0 - z1) * (3.0 + z1 + 2.0 * z2)))


def inclination_camera(r_camera: float, inclination: float, fov_y: float, aspect: float = 1.0):
    """Camera at (t, r, theta, phi) = (0, r, i, 0) looking at the hole (-r_hat), up = +z."""
    from . import Camera
    st, ct = math.sin(inclination), math.cos(inclination)
    return Camera(position=(0.0, r_camera, inclination, 0.0), look_at=(-(st * 1.0), -(st * 0.0), -ct),
                  up=(0.0, 0.0, 1.0), fov_y_radians=fov_y, aspect=aspect)


def rtol_integrator(max_radius: float, max_steps: int, rtol: float = 1e-8, atol: float = 1e-10):
    """h0/min/max = 0.1/1e-8/1.0 (reference defaults), capture r < 1.01 r_h."""
    from . import GG_ERR_RTOL_ATOL, RK45Integrator
    return RK45Integrator(initial_step=0.1, min_step=1e-8, max_step=1.0, tolerance=1e-6, max_radius=max_radius,
                          horizon_buffer=0.0, horizon_rel=0.01, rtol=rtol, atol=atol, err_mode=GG_ERR_RTOL_ATOL,
                          max_steps=max_steps)


def thin_disk(r_inner: float, r_outer: float, q: float = 3.0, edge_width: float = 0.0):
    from . import GG_DISK_THIN, AccretionDisk
    return AccretionDisk(r_inner=r_inner, r_outer=r_outer, emissivity_q=q, edge_width=edge_width, kind=GG_DISK_THIN)


def kerr_thin_scene(width: int, height: int, spin: float, r_camera: float, inclination_deg: float, fov_deg: float,
                    max_radius: float = 1000.0, max_steps: int = 20000, r_outer: float = 30.0, q: float = 3.0,
                    jitter: bool = False, seed: int = 0, mass: float = 1.0):
    from . import Kerr, Scene
    return Scene(metric=Kerr(mass, spin),
                 camera=inclination_camera(r_camera, math.radians(inclination_deg), math.radians(fov_deg),
                                           width / height),
                 disk=thin_disk(r_isco(mass, spin), r_outer, q), integrator=rtol_integrator(max_radius, max_steps),
                 width=width, height=height, jitter=jitter, seed=seed)


def schwarzschild_256(width: int = 256, height: int = 256):
    from . import Schwarzschild, Scene
    return Scene(metric=Schwarzschild(1.0),
                 camera=inclination_camera(50.0, math.radians(80.0), math.radians(30.0), width / height),
                 disk=thin_disk(6.0, 20.0, 3.0), integrator=rtol_integrator(1000.0, 20000),
                 width=width, height=height)


def kerr_1024(width: int = 1024, height: int = 1024, seed: int = 2026):
    return kerr_thin_scene(width, height, spin=0.9, r_camera=100.0, inclination_deg=60.0, fov_deg=25.0,
                           max_radius=1000.0, max_steps=20000, r_outer=30.0, q=3.0, jitter=True, seed=seed)


def kerr_4096(width: int = 4096, height: int = 4096):
    return kerr_thin_scene(width, height, spin=0.998, r_camera=1000.0, inclination_deg=85.0, fov_deg=20.0,
                           max_radius=1100.0, max_steps=50000, r_outer=30.0, q=3.0)


def batch_64(seed: int = 2026, n: int = 64, resolutions=(128, 256, 512, 1024)):
    """Seeded multi-scene list: a~U[0,0.998], i~U[5,89] deg, r_cam~U[30,500], res from `resolutions`,
    r_out~U[15,40], q~U[2,3.5], r_in = r_ISCO(a)."""
    rng = np.random.default_rng(seed)
    scenes = []
    for _ in range(n):
        a = float(rng.uniform(0.0, 0.998))
        inc = float(rng.uniform(5.0, 89.0))
        rc = float(rng.uniform(30.0, 500.0))
        res = int(rng.choice(resolutions))
        r_out = float(rng.uniform(15.0, 40.0))
        q = float(rng.uniform(2.0, 3.5))
        scenes.append(kerr_thin_scene(res, res, spin=a, r_camera=rc, inclination_deg=inc, fov_deg=30.0,
                                      max_radius=1000.0, max_steps=20000, r_outer=r_out, q=q))
    return scenes


def grad_512_params(width: int = 512, height: int = 512, spin=0.5, inclination_deg=45.0, q=3.0, r_in=None,
                    precision: str = "fp64"):
    """gg_grad_scene fields for the differentiable render (r_in defaults to r_ISCO(a) + 0.5M); precision of the
    forward render, "fp64" or "mixed"."""
    from . import _abi
    if r_in is None:
        r_in = r_isco(1.0, spin) + 0.5
    g = _abi.gg_grad_scene()
    g.mass = 1.0
    g.spin = spin
    g.inclination = math.radians(inclination_deg)
    g.emissivity_q = q
    g.r_inner = r_in
    g.r_outer = 20.0
    g.edge_width = 0.25
    g.r_camera = 50.0
    g.fov_y_radians = math.radians(30.0)
    g.integrator = rtol_integrator(200.0, 20000).to_c()
    g.width = width
    g.height = height
    g.precision = {"fp64": _abi.GG_PREC_FP64, "mixed": _abi.GG_PREC_MIXED}[precision]
    return g


def lensed_spheres_world():
    """build_lensed_spheres (crates/gr-renderer/src/main.rs:173-202): a light behind the hole and four
    Lambertian spheres in a cross around the line of sight, at r = 40-60M."""
    from . import DiffuseLight, Lambertian, Sphere, World
    return (World()
            .add(Sphere((-60.0, 0.0, 0.0), 4.0, DiffuseLight((2.0, 1.6, 0.8))))
            .add(Sphere((-40.0, 14.0, 0.0), 3.5, Lambertian((0.85, 0.10, 0.10))))
            .add(Sphere((-40.0, -14.0, 0.0), 3.5, Lambertian((0.10, 0.75, 0.20))))
            .add(Sphere((-40.0, 0.0, 14.0), 3.5, Lambertian((0.15, 0.30, 0.90))))
            .add(Sphere((-40.0, 0.0, -14.0), 3.5, Lambertian((0.95, 0.85, 0.20))))
            .build_bvh())
