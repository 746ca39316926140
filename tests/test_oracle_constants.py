"""Pins of the constants of motion (E, L_z, Carter Q, mu^2) the oracle and the
library compute: they must be what the geodesic flow conserves, which the
integrator only approximates (forward-Euler positions, O(h) drift)."""
import math

import numpy as np
import pytest

import gargantua_amd as gg


def _scene(spin, max_step, n=6):
    sc = gg.configs.kerr_thin_scene(n, n, spin=spin, r_camera=40.0, inclination_deg=60.0, fov_deg=40.0,
                                    max_radius=200.0)
    sc.disk = None
    sc.integrator.max_step = max_step
    sc.integrator.rtol, sc.integrator.atol = 1e-12, 1e-14
    return sc


def _drift(oracle, scene):
    o = oracle.render(scene)
    s0 = np.stack([oracle.from_camera(scene, k % scene.width, k // scene.width) for k in range(scene.n_rays)])
    s0 = np.concatenate([s0[:, :8], np.zeros((len(s0), 1))], 1) if s0.shape[1] == 8 else s0
    s1 = o["state"].T
    c0 = oracle.constants_of_motion(scene.metric, s0)
    c1 = oracle.constants_of_motion(scene.metric, s1)
    scale = np.array([np.abs(c0[:, 0]), np.abs(c0[:, 1]) + 1e-3, np.abs(c0[:, 2]) + np.abs(c0[:, 1]) ** 2]).T
    return np.median(np.abs(c1[:, :3] - c0[:, :3]) / scale, axis=0), c0, c1


@pytest.mark.parametrize("spin", [0.0, 0.9])
def test_constants_drift_shrinks_with_step(oracle, spin):
    """A wrong Q (a dropped or sign-flipped a^2 term, a missing cos^2) is not
    conserved by the exact flow: its drift would not fall with max_step."""
    d_big, _, _ = _drift(oracle, _scene(spin, 1.0))
    d_small, _, _ = _drift(oracle, _scene(spin, 0.01))
    print(f"a={spin}: median drift (E, Lz, Q) at max_step 1: {d_big}, 0.01: {d_small}")
    assert np.all(d_small < 0.2 * d_big)


def test_equatorial_ray_has_zero_carter_q(oracle):
    m = gg.Kerr(1.0, 0.7)
    st = np.array([[0.0, 12.0, math.pi / 2, 0.3, 1.3, -0.8, 0.0, 0.02, 0.0]])
    assert abs(oracle.constants_of_motion(m, st)[0, 2]) < 1e-15


def test_camera_ray_mass_shell(oracle):
    """ray.rs:51-53 sets k^t from the diagonal metric only: null at a = 0,
    mu^2 = -2 g_tphi k^t k^phi for Kerr."""
    for spin in (0.0, 0.9):
        sc = _scene(spin, 1.0)
        for k in range(sc.n_rays):
            s0 = oracle.from_camera(sc, k % sc.width, k // sc.width)
            s0 = np.concatenate([s0[:8], [0.0]])
            mu2 = oracle.constants_of_motion(sc.metric, s0[None])[0, 3]
            g = oracle.metric_tensor(sc.metric, s0[:4])
            want = -2.0 * g[0, 3] * s0[4] * s0[7]
            assert abs(mu2 - want) <= 1e-12 * (abs(s0[4]) ** 2), (spin, k, mu2, want)


@pytest.mark.parametrize("spin,pix", [(0.9, (1, 1)), (0.5, (0, 2))])
def test_constants_conserved_along_accurate_geodesic(oracle, spin, pix):
    """Integrated accurately (DOP853, rtol 1e-12, on the oracle's own RHS), a
    ray keeps E, L_z, Q and mu^2 to ~1e-12.  A sign-flipped a^2 term or a
    dropped cos^2 in Q varies by 2e-3..2e-2 along the same rays, so the
    forms are pinned, not just restated."""
    from scipy.integrate import solve_ivp
    m = gg.Kerr(1.0, spin)
    sc = _scene(spin, 1.0)
    s0 = oracle.from_camera(sc, *pix)

    def f(lam, y):
        a = oracle.geodesic_acceleration(m, y[None, :4], y[None, 4:8])[0]
        return np.concatenate([y[4:8], a])

    def near(lam, y):
        return y[1] - 2.5
    near.terminal = True
    sol = solve_ivp(f, (0, 150), s0[:8], method="DOP853", rtol=1e-12, atol=1e-12, max_step=0.5, events=near)
    ys = sol.y.T
    assert np.ptp(ys[:, 2]) > 0.5           # the ray swings in theta, so Q's cos^2 terms matter
    c = oracle.constants_of_motion(m, np.concatenate([ys, np.zeros((len(ys), 1))], 1))
    var = np.ptp(c, axis=0) / np.array([abs(c[0, 0]), abs(c[0, 1]), abs(c[0, 2]), abs(c[0, 0]) ** 2])
    print(f"a={spin} pixel {pix}: relative variation (E, Lz, Q, mu2) {var}")
    assert np.all(var < 1e-10)
