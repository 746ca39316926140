"""The reference's own unit tests, restated against the FP64 oracle.

Each test names the reference test it restates (file:line).  These pin the
oracle's metric, Christoffel symbols, RHS, horizon and orbit functions with
the reference's own assertions and tolerances.
"""
import math

import numpy as np

import gargantua_amd as gg

EPS = 1e-12


def test_kerr_reduces_to_schwarzschild_when_a_is_zero(oracle):
    # crates/gr-core/src/kerr.rs:507-537
    kerr, schw = gg.Kerr(1.0, 0.0), gg.Schwarzschild(1.0)
    for pos in [(0.0, 10.0, math.pi / 3.0, 0.0), (0.0, 6.0, math.pi / 2.0, 1.0), (1.0, 20.0, math.pi / 4.0, 2.5),
                (0.0, 5.0, 2.0, 0.5)]:
        gk, gs = oracle.metric_tensor(kerr, pos), oracle.metric_tensor(schw, pos)
        assert np.max(np.abs(gk - gs)) < EPS


def test_kerr_horizon(oracle):
    # kerr.rs:540-550
    assert abs(oracle.event_horizon(gg.Kerr(1.0, 0.5)) - (1.0 + math.sqrt(0.75))) < EPS


def _numerical_christoffel(oracle, metric, pos, h=1e-5):
    # the FD helper of kerr.rs:471-504: central differences of g, then
    # Gamma^mu_ab = 1/2 g^mu s (d_a g_sb + d_b g_sa - d_s g_ab)
    pos = np.asarray(pos, float)
    dg = np.zeros((4, 4, 4))
    for s in range(4):
        p, m = pos.copy(), pos.copy()
        p[s] += h
        m[s] -= h
        dg[s] = (oracle.metric_tensor(metric, p) - oracle.metric_tensor(metric, m)) / (2.0 * h)
    ginv = np.linalg.inv(oracle.metric_tensor(metric, pos))
    gam = np.zeros((4, 4, 4))
    for mu in range(4):
        for a in range(4):
            for b in range(4):
                gam[mu, a, b] = 0.5 * sum(ginv[mu, s] * (dg[a, s, b] + dg[b, s, a] - dg[s, a, b]) for s in range(4))
    return gam


def test_kerr_christoffel_matches_numerical(oracle):
    # kerr.rs:553-575
    kerr = gg.Kerr(1.0, 0.7)
    pos = (0.0, 10.0, math.pi / 3.0, 0.7)
    diff = np.abs(oracle.christoffel(kerr, pos) - _numerical_christoffel(oracle, kerr, pos))
    assert diff.max() < 1e-4


def test_kerr_christoffel_matches_numerical_more_points(oracle):
    # same check over spins and radii the benchmark scenes visit, relative to
    # the symbol scale so a dropped term anywhere fails
    rng = np.random.default_rng(7)
    for a in (0.0, 0.3, 0.9, 0.998):
        kerr = gg.Kerr(1.0, a)
        for _ in range(6):
            pos = (0.0, float(rng.uniform(2.5, 60.0)), float(rng.uniform(0.3, math.pi - 0.3)), 0.4)
            an = oracle.christoffel(kerr, pos)
            nu = _numerical_christoffel(oracle, kerr, pos, h=1e-6)
            scale = np.max(np.abs(an))
            assert np.max(np.abs(an - nu)) < 1e-6 * max(scale, 1.0), (a, pos)


def test_schwarzschild_christoffel_matches_numerical(oracle):
    schw = gg.Schwarzschild(1.0)
    for pos in [(0.0, 10.0, 1.0, 0.0), (0.0, 3.5, 2.2, 1.0)]:
        an = oracle.christoffel(schw, pos)
        nu = _numerical_christoffel(oracle, schw, pos, h=1e-6)
        assert np.max(np.abs(an - nu)) < 1e-6


def _rk4_orbit(oracle, metric, r0, omega, step):
    state = np.array([0.0, r0, math.pi / 2.0, 0.0, 1.0, 0.0, 0.0, omega, 0.0])
    steps = int((2.0 * math.pi / omega) / step)
    for _ in range(steps):
        _, state = oracle.rk4_step(metric, step, state)
    return state


def test_kerr_circular_orbit_equatorial(oracle):
    # kerr.rs:578-610 (RK4 h = 0.05 over one period)
    m, a, r0 = 1.0, 0.5, 10.0
    omega = math.sqrt(m) / (r0 ** 1.5 + a * math.sqrt(m))
    st = _rk4_orbit(oracle, gg.Kerr(m, a), r0, omega, 0.05)
    assert abs(st[1] - r0) < 0.5


def test_circular_orbit(oracle):
    # crates/gr-core/src/geodesic.rs:360-389 (Schwarzschild, RK4 h = 0.01, r = 6)
    r = 6.0
    st = _rk4_orbit(oracle, gg.Schwarzschild(1.0), r, math.sqrt(1.0 / r ** 3), 0.01)
    assert abs(st[1] - r) < 0.1


def test_kerr_metric_signature_and_symmetry(oracle):
    # kerr.rs:613-629
    g = oracle.metric_tensor(gg.Kerr(1.0, 0.9), (0.0, 10.0, math.pi / 2.0, 0.0))
    assert g[0, 0] < 0 and g[1, 1] > 0 and g[2, 2] > 0 and g[3, 3] > 0
    assert abs(g[0, 3] - g[3, 0]) < EPS
    assert abs(g[0, 3]) > 1e-10


def test_keplerian_orbit_is_normalized_timelike(oracle):
    # crates/gr-core/src/schwarzschild.rs:162-171
    schw = gg.Schwarzschild(1.0)
    pos = (0.0, 10.0, math.pi / 2.0, 0.0)
    u = oracle.orbital_four_velocity(schw, pos)
    g = oracle.metric_tensor(schw, pos)
    assert abs(u @ g @ u + 1.0) < 1e-9
    assert u[3] > 0.0


def test_kerr_keplerian_orbit_is_normalized_timelike(oracle):
    # the Kerr branch of the same contract (kerr.rs:439-449)
    for a in (0.5, 0.9, 0.998):
        kerr = gg.Kerr(1.0, a)
        for r in (3.0, 6.0, 25.0):
            pos = (0.0, r, math.pi / 2.0, 0.0)
            u = oracle.orbital_four_velocity(kerr, pos)
            g = oracle.metric_tensor(kerr, pos)
            assert abs(u @ g @ u + 1.0) < 1e-9
            assert abs(u[3] / u[0] - 1.0 / (r ** 1.5 + a)) < 1e-12  # Omega = sqrt(M)/(r^1.5 + a sqrt(M))


def test_no_circular_orbit_inside_photon_sphere(oracle):
    # schwarzschild.rs:173-178
    assert oracle.orbital_four_velocity(gg.Schwarzschild(1.0), (0.0, 2.5, math.pi / 2.0, 0.0)) is None


def test_schwarzschild_flat_at_infinity(oracle):
    # schwarzschild.rs:180-188
    g = oracle.metric_tensor(gg.Schwarzschild(1.0), (0.0, 1000.0, math.pi / 2.0, 0.0))
    assert abs(g[0, 0] + 1.0) < 0.01 and abs(g[1, 1] - 1.0) < 0.01


def test_horizon_shadow_and_escape(oracle):
    # crates/gr-renderer/tests/render_sanity.rs:16-27 (32x24 Schwarzschild render)
    cam = gg.Camera(position=(0.0, 30.0, math.pi / 2.0, 0.0), look_at=(-1.0, 0.0, 0.0), up=(0.0, 0.0, 1.0),
                    fov_y_radians=math.radians(60.0), aspect=32 / 24)
    scene = gg.Scene(metric=gg.Schwarzschild(1.0), camera=cam,
                     integrator=gg.RK45Integrator(max_radius=200.0, max_steps=5000), width=32, height=24)
    rgb = oracle.render(scene, rgb=True)["rgb"]
    assert np.any(np.all(rgb == 0, axis=1)), "expected at least one black pixel (horizon shadow)"
    assert np.any(np.any(rgb != 0, axis=1)), "expected at least one non-black pixel (escaped rays)"


def test_kerr_spin_clamp(oracle):
    # kerr.rs:27-32: |spin| > mass is clamped to +-0.9999 mass, so the
    # horizon stays M + sqrt(M^2 - a^2) with a = 0.9999 M, and every metric
    # quantity equals that of the clamped spin.
    a_max = 0.9999
    assert oracle.event_horizon(gg.Kerr(1.0, 1.5)) == 1.0 + math.sqrt(1.0 - a_max * a_max)
    assert oracle.event_horizon(gg.Kerr(2.0, -7.0)) == 2.0 + math.sqrt(4.0 - (2.0 * a_max) ** 2)
    rng = np.random.default_rng(11)
    for sign in (1.0, -1.0):
        over, at = gg.Kerr(1.0, sign * 1.5), gg.Kerr(1.0, sign * a_max)
        for _ in range(8):
            pos = (0.0, rng.uniform(2.0, 50.0), rng.uniform(0.1, math.pi - 0.1), rng.uniform(0.0, 6.0))
            assert np.array_equal(oracle.metric_tensor(over, pos), oracle.metric_tensor(at, pos))
        pos = np.array([[0.0, 7.0, 1.1, 0.3]])
        vel = np.array([[1.0, -0.4, 0.01, 0.02]])
        assert np.array_equal(oracle.geodesic_acceleration(over, pos, vel),
                              oracle.geodesic_acceleration(at, pos, vel))
    # inside the bound the spin is kept as given
    assert oracle.event_horizon(gg.Kerr(1.0, 0.5)) == 1.0 + math.sqrt(0.75)
