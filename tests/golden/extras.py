"""The extras of the widening-row golden cases (make_golden_extras.py): each
case key names a world, a sky texture or a tensor source, rebuilt here."""
import contextlib
import math

import numpy as np

import gargantua_amd as gg

# Reissner-Nordstrom, charge Q in the spin slot (the same tensor as tests/test_user_metric.py)
REISSNER_NORDSTROM = """
__device__ void gg_user_metric(const double x[4], double m, double q, double g[16]) {
  for (int i = 0; i < 16; ++i) g[i] = 0.0;
  const double r = x[1], s = sin(x[2]);
  const double f = 1.0 - 2.0 * m / r + q * q / (r * r);
  g[0] = -f; g[5] = 1.0 / f; g[10] = r * r; g[15] = r * r * s * s;
}
"""


def sky_texture():
    """A smooth, asymmetric 64x32 equirectangular map, intensity 0.8."""
    th = (np.arange(32) + 0.5) / 32 * math.pi
    ph = (np.arange(64) + 0.5) / 64 * 2 * math.pi - math.pi
    px = np.zeros((32, 64, 3), np.float32)
    px[..., 0] = 0.6 + 0.4 * np.cos(th)[:, None]
    px[..., 1] = 0.5 + 0.3 * np.sin(ph)[None, :] * np.sin(th)[:, None]
    px[..., 2] = 0.25 + 0.2 * np.cos(2 * ph)[None, :]
    return gg.HdriSky(px, intensity=0.8)


def CASES():
    ref = gg.RK45Integrator(max_radius=200.0, max_steps=5000)
    cli_cam = gg.Camera(position=(0.0, 30.0, math.pi / 2, 0.0), look_at=(-1.0, 0.0, 0.0), up=(0.0, 0.0, 1.0),
                        fov_y_radians=math.radians(60.0), aspect=16 / 12)
    cam = gg.configs.inclination_camera(30.0, math.radians(80.0), math.radians(50.0))
    thin = gg.AccretionDisk(kind=gg.GG_DISK_THIN, r_inner=5.0, r_outer=20.0)
    rn = gg.UserMetric(REISSNER_NORDSTROM, mass=1.0, spin=0.6, event_horizon=1.8)
    return {
        "lensed_spheres_schw": (gg.Scene(metric=gg.Schwarzschild(1.0), camera=cli_cam, integrator=ref, width=16,
                                         height=12), {"world": "lensed-spheres"}),
        "lensed_spheres_kerr_disk": (gg.Scene(metric=gg.Kerr(1.0, 0.9), camera=cli_cam, disk=gg.AccretionDisk(),
                                              integrator=ref, width=16, height=12), {"world": "lensed-spheres"}),
        "hdri_kerr_thin_spa2": (gg.Scene(metric=gg.Kerr(1.0, 0.9), camera=cam, disk=thin, integrator=ref, width=12,
                                         height=8, samples_per_axis=2), {"sky": "smooth-64x32"}),
        "reissner_nordstrom_thin": (gg.Scene(metric=rn, camera=cam, disk=thin, integrator=ref, width=16, height=16),
                                    {"metric": "reissner-nordstrom-0.6"}),
    }


@contextlib.contextmanager
def applied(target, extra):
    """Set a case's world / sky / tensor on an Oracle or a Renderer for the duration of the block."""
    world = gg.configs.lensed_spheres_world() if extra.get("world") == "lensed-spheres" else None
    sky = sky_texture() if extra.get("sky") == "smooth-64x32" else None
    is_oracle = not isinstance(target, gg.Renderer)
    if world is not None:
        target.set_hittables(world)
    if sky is not None:
        if is_oracle:
            target.set_sky_hdri(sky.pixels, sky.intensity)
        else:
            target.set_sky_hdri(sky)
    if extra.get("metric") and is_oracle:
        target.set_user_metric(CASES()["reissner_nordstrom_thin"][0].metric)
    try:
        yield
    finally:
        if world is not None:
            target.set_hittables(None)
        if sky is not None:
            target.set_sky_hdri(None)
        if extra.get("metric") and is_oracle:
            target.set_user_metric(None)
