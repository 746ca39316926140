"""Closed-form pins of the oracle's two disk models.

Volumetric disk (trace_ray_with_disk, tracer.rs:96-123; AccretionDisk,
disk.rs:40-81).  A radial ray in the equatorial plane of Schwarzschild
(v^theta = v^phi = 0, v^r = -E) has dr/dlambda = -E exactly, so the
reference's forward-Euler positions are exact and r_k = r_0 - k h E.  With
E = 1, start r_0 = 21 and h = 2^-n the grid meets the disk edges (6, 20)
exactly.  On the equator z = 0, so rho = rho_0 (r_in/r)^2, j = rho^2
(r_in/r)^(3/4), and the redshift of a Keplerian emitter seen by the static
camera is g = sqrt(1 - 3M/r) / sqrt(1 - 2M/r_obs) whatever the photon's
energy (p_phi = 0, so p_t cancels).  The transmission is closed form,
T(r) = exp(-0.1 rho_0 r_in^2 (1/r - 1/r_out)), and the reference's sum
I_h = sum_k j(m_k) g(m_k)^4 h T_(k-1) (midpoint sample, transmission before
the step) converges to I = int j g^4 T dr at first order, with the leading
error (h/2) int j g^4 (0.1 rho) T dr: its sign pins that the transmission
is taken before the step's absorption.  With rho_0 = 50 the ray stops where
T falls below 1e-3 (tau = ln 1000).

Thin disk (the crossing intensity I = r_c^-q g^4 with g from
redshift_factor, tracer.rs:158-183).  In Schwarzschild the centre column of
an odd-width image has n_phi = 0 exactly, so k^phi = 0, v^phi stays 0 and
p_phi = 0 at the crossing: g = sqrt(1 - 3M/r_c) / sqrt(1 - 2M/r_obs)
exactly, with r_c interpolated on the crossing step between its end states.
"""
import math

import numpy as np
import pytest

import gargantua_amd as gg
from gargantua_amd import configs

R0, R_IN, R_OUT = 21.0, 6.0, 20.0


def _f(r):
    return 1.0 - 2.0 / r


def _radial_intensity(oracle, rho0, h):
    disk = gg.AccretionDisk(r_inner=R_IN, r_outer=R_OUT, density_0=rho0)
    integ = gg.RK45Integrator(initial_step=h, min_step=1e-8, max_step=h, max_radius=100.0, horizon_buffer=0.01,
                              rtol=1e-8, atol=1e-10, err_mode=gg.GG_ERR_RTOL_ATOL, max_steps=100000)
    cam = gg.Camera(position=(0.0, R0, math.pi / 2, 0.0), look_at=(-1.0, 0.0, 0.0))
    sc = gg.Scene(metric=gg.Schwarzschild(1.0), camera=cam, disk=disk, integrator=integ, width=1, height=1)
    st = np.array([[0.0, R0, math.pi / 2, 0.0, 1.0 / _f(R0), -1.0, 0.0, 0.0, 0.0]])
    o = oracle.trace_states(sc, st)
    assert o["cls"][0] == gg.GG_CLASS_DISK
    return o["intensity"][0]


def _emission_integral(rho0):
    from scipy.integrate import quad
    tau = lambda r: 0.1 * rho0 * R_IN ** 2 * (1.0 / r - 1.0 / R_OUT)
    j = lambda r: (rho0 * (R_IN / r) ** 2) ** 2 * (R_IN / r) ** 0.75
    g4 = lambda r: ((1.0 - 3.0 / r) / _f(R0)) ** 2
    r_x = R_IN
    if tau(R_IN) > math.log(1000.0):
        r_x = 1.0 / (1.0 / R_OUT + math.log(1000.0) / (0.1 * rho0 * R_IN ** 2))
    kw = dict(epsabs=0.0, epsrel=1e-13, limit=200)
    I = quad(lambda r: j(r) * g4(r) * math.exp(-tau(r)), r_x, R_OUT, **kw)[0]
    C1 = 0.5 * quad(lambda r: j(r) * g4(r) * 0.1 * rho0 * (R_IN / r) ** 2 * math.exp(-tau(r)), r_x, R_OUT, **kw)[0]
    return I, C1


def test_volumetric_radial_ray_converges_to_the_emission_integral(oracle):
    I, C1 = _emission_integral(1.0)
    hs = [0.25, 0.125, 0.0625, 0.03125]
    dev = []
    for h in hs:
        In = _radial_intensity(oracle, 1.0, h)
        dev.append((In - I) / h - C1)
        print(f"h {h}: I_h {In:.12e}, I {I:.12e}, (I_h - I)/h {(In - I) / h:.6e}, leading term {C1:.6e}")
    dev = np.array(dev)
    # (I_h - I)/h -> C1 > 0, the remainder itself first order in h
    assert abs(dev[-1]) < 0.01 * C1
    ratios = dev[:-1] / dev[1:]
    assert np.all((ratios > 1.7) & (ratios < 2.3)), ratios


def test_volumetric_radial_ray_stops_at_the_transmission_floor(oracle):
    I, C1 = _emission_integral(50.0)
    errs = []
    for h in (0.125, 0.0625, 0.03125):
        In = _radial_intensity(oracle, 50.0, h)
        errs.append(In - I)
        print(f"h {h}: I_h {In:.10e}, I {I:.10e}, (I_h - I)/h {(In - I) / h:.5e}, leading term {C1:.5e}")
    errs = np.array(errs)
    assert np.all(errs > 0) and np.all(np.diff(errs) < 0)
    assert abs(errs[-1] / 0.03125 - C1) < 0.02 * C1
    assert errs[-1] < 0.01 * I


@pytest.mark.parametrize("incl_deg", [30.0, 60.0, 80.0])
def test_thin_disk_intensity_of_zero_angular_momentum_photons(oracle, incl_deg):
    w = 65
    cam = configs.inclination_camera(50.0, math.radians(incl_deg), math.radians(30.0))
    integ = configs.rtol_integrator(1000.0, 20000)
    scene = gg.Scene(metric=gg.Schwarzschild(1.0), camera=cam, disk=configs.thin_disk(6.0, 20.0, 3.0),
                     integrator=integ, width=w, height=w)
    o = oracle.render(scene)
    col = np.arange(w) * w + w // 2
    disk = col[o["cls"][col] == gg.GG_CLASS_DISK]
    assert len(disk) >= 4
    f_obs = _f(50.0)
    checked = 0
    for i in disk:
        start = oracle.from_camera(scene, i % w, i // w)
        assert start[7] == 0.0                       # k^phi of the centre column
        # the state before the crossing attempt: the same ray stopped one attempt early
        pre_scene = gg.Scene(metric=scene.metric, camera=cam, disk=None, width=1, height=1,
                             integrator=configs.rtol_integrator(1000.0, int(o["attempts"][i]) - 1))
        pre = oracle.trace_states(pre_scene, start[None, :])
        assert pre["cls"][0] == gg.GG_CLASS_MAX_STEPS
        r0, th0 = pre["state"][1, 0], pre["state"][2, 0]
        r1, th1 = o["state"][1, i], o["state"][2, i]
        assert o["state"][7, i] == 0.0 and pre["state"][7, 0] == 0.0   # v^phi stays 0
        assert (th0 - math.pi / 2) * (th1 - math.pi / 2) < 0          # the last step crossed the equator
        s = (math.pi / 2 - th0) / (th1 - th0)
        rc = r0 + s * (r1 - r0)
        want = rc ** -3.0 * ((1.0 - 3.0 / rc) / f_obs) ** 2
        assert o["intensity"][i] == pytest.approx(want, rel=1e-12)
        checked += 1
    print(f"i = {incl_deg} deg: {checked} centre-column disk pixels match r_c^-3 g^4")
