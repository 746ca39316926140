"""Hittable scenes through the curved renderer: render_with_scene and
render_with_disk_and_scene (crates/raytracer/src/curved/renderer.rs:103-254,
tracer.rs:207-384).

The oracle's chord test, shading and static redshift are pinned by closed
forms (ray-sphere and ray-quad geometry, the facing ratio, g^4 of static
emitters in Schwarzschild and on the Kerr axis) and by the flat-space limit,
where the photon paths are straight lines whose hits are known analytically.
The GPU (BVH traversal in the trace kernel) is held to the oracle (linear
World::hit).  The reference has no tests of its own for this path."""
import ctypes
import math

import numpy as np
import pytest

import gargantua_amd as gg


def lensed_spheres():
    return gg.configs.lensed_spheres_world()


def scene_camera(r=30.0, fov_deg=60.0):
    return gg.Camera(position=(0.0, r, math.pi / 2, 0.0), look_at=(-1.0, 0.0, 0.0), up=(0.0, 0.0, 1.0),
                     fov_y_radians=math.radians(fov_deg), aspect=1.0)


def scene_of(metric, disk=None, w=32, h=32, integ=None, spa=1):
    return gg.Scene(metric=metric, camera=scene_camera(), disk=disk,
                    integrator=integ or gg.RK45Integrator(max_radius=200.0, max_steps=5000), width=w, height=h,
                    samples_per_axis=spa)


@pytest.fixture
def with_world(oracle):
    """Set a world on the oracle for one test and clear it afterwards."""
    def set_(world):
        assert oracle.set_hittables(world) == 0
    yield set_
    oracle.set_hittables(None)


# --- closed forms of one chord ----------------------------------------------------------

def test_sphere_chord_closed_forms(oracle, with_world):
    alb = np.array([0.2, 0.4, 0.6])
    with_world(gg.World().add(gg.Sphere((5.0, 0.0, 0.0), 1.0, gg.Lambertian(tuple(alb)))))
    # Head-on: t = (5 - 1) / 10, point (4, 0, 0), normal against the chord, facing ratio 1.
    obj, p, n, lin = oracle.scene_hit([0.0, 0.0, 0.0], [10.0, 0.0, 0.0])
    np.testing.assert_array_equal(p, [4.0, 0.0, 0.0])
    np.testing.assert_array_equal(n, [-1.0, 0.0, 0.0])
    np.testing.assert_array_equal(lin, alb)
    # Offset by b = 0.5: the hit sits at x = 5 - sqrt(1 - b^2) and -d.n = sqrt(1 - b^2).
    obj, p, n, lin = oracle.scene_hit([0.0, 0.5, 0.0], [10.0, 0.0, 0.0])
    c = math.sqrt(0.75)
    assert abs(p[0] - (5.0 - c)) < 1e-14 and abs(p[1] - 0.5) < 1e-15
    np.testing.assert_allclose(lin, alb * (0.5 + 0.5 * c), rtol=1e-14)
    # From inside: the far root, a back face, its normal flipped toward the chord.
    obj, p, n, lin = oracle.scene_hit([5.0, 0.0, 0.0], [10.0, 0.0, 0.0])
    np.testing.assert_array_equal(p, [6.0, 0.0, 0.0])
    np.testing.assert_array_equal(n, [-1.0, 0.0, 0.0])
    # The chord ends before the sphere (t in (0, 1) only), or passes beside it.
    assert oracle.scene_hit([0.0, 0.0, 0.0], [3.0, 0.0, 0.0]) is None
    assert oracle.scene_hit([0.0, 1.5, 0.0], [10.0, 0.0, 0.0]) is None
    # Grazing: -d.n = 0 gives the facing floor 0.5.
    obj, p, n, lin = oracle.scene_hit([0.0, 1.0, 0.0], [10.0, 0.0, 0.0])
    np.testing.assert_allclose(lin, alb * 0.5, atol=1e-7)


def test_quad_chord_closed_forms(oracle, with_world):
    emit = (2.0, 1.0, 0.5)
    with_world(gg.World().add(gg.Quad((5.0, -1.0, -1.0), (0.0, 2.0, 0.0), (0.0, 0.0, 2.0), gg.DiffuseLight(emit))))
    obj, p, n, lin = oracle.scene_hit([0.0, 0.2, -0.3], [10.0, 0.0, 0.0])
    np.testing.assert_allclose(p, [5.0, 0.2, -0.3], atol=1e-15)
    np.testing.assert_array_equal(n, [-1.0, 0.0, 0.0])
    np.testing.assert_array_equal(lin, emit)                 # no scatter: albedo 0, plus emission
    assert oracle.scene_hit([0.0, 1.5, 0.0], [10.0, 0.0, 0.0]) is None      # alpha = 1.25
    assert oracle.scene_hit([0.0, 0.0, -1.2], [10.0, 0.0, 0.0]) is None     # beta = -0.1
    assert oracle.scene_hit([0.0, 0.0, 0.0], [0.0, 10.0, 0.0]) is None      # parallel to the plane
    # From behind (+x side) the front face flips.
    obj, p, n, lin = oracle.scene_hit([9.0, 0.0, 0.0], [-10.0, 0.0, 0.0])
    np.testing.assert_array_equal(n, [1.0, 0.0, 0.0])


def test_materials_and_closest_hit(oracle, with_world):
    a = gg.Sphere((5.0, 0.0, 0.0), 1.0, gg.Metal((0.9, 0.8, 0.7)))
    b = gg.Sphere((7.0, 0.0, 0.0), 1.5, gg.Dielectric(1.5))
    for world in (gg.World([a, b]), gg.World([b, a])):       # the closest hit, whatever the order
        with_world(world)
        obj, p, n, lin = oracle.scene_hit([0.0, 0.0, 0.0], [10.0, 0.0, 0.0])
        np.testing.assert_array_equal(p, [4.0, 0.0, 0.0])
        np.testing.assert_array_equal(lin, [0.9, 0.8, 0.7])   # mirror reflection leaves: the albedo
    with_world(gg.World([b]))
    obj, p, n, lin = oracle.scene_hit([0.0, 0.0, 0.0], [10.0, 0.0, 0.0])
    np.testing.assert_array_equal(lin, [1.0, 1.0, 1.0])       # dielectric: white attenuation
    assert oracle.set_hittables(gg.World([gg.Sphere((0, 0, 0), 1.0, gg.Metal((1, 1, 1), fuzz=0.3))])) != 0


def test_static_redshift_closed_forms(oracle):
    # Schwarzschild: g^4 = ((1 - 2M/r_e) / (1 - 2M/r_o))^2 for static emitter and observer.
    m = gg.Schwarzschild(1.0)
    g_obs = 1.0 - 2.0 / 30.0
    for r_e in (3.0, 10.0, 45.0):
        got = oracle.g4_static(m, [0.0, r_e, 0.0], g_obs)
        assert abs(got - ((1.0 - 2.0 / r_e) / g_obs) ** 2) < 1e-14
    assert oracle.g4_static(m, [1.5, 0.0, 0.0], g_obs) == 0.0      # inside the horizon: g_tt >= 0
    # Kerr: on the axis Sigma = r^2 + a^2, in the equator Sigma = r^2.
    k = gg.Kerr(1.0, 0.9)
    for r in (4.0, 20.0):
        ax = oracle.g4_static(k, [0.0, 0.0, r], g_obs)
        assert abs(ax - ((1.0 - 2.0 * r / (r * r + 0.81)) / g_obs) ** 2) < 1e-14
        eq = oracle.g4_static(k, [r, 0.0, 0.0], g_obs)
        assert abs(eq - ((1.0 - 2.0 / r) / g_obs) ** 2) < 1e-14
    assert oracle.g4_static(k, [1.95, 0.0, 0.0], g_obs) == 0.0    # ergoregion: g_tt > 0


# --- renders through the oracle ---------------------------------------------------------------

def _cartesian_dir(state):
    """Spatial direction of a camera ray: the Jacobian of (r, th, ph) -> (x, y, z) applied to its velocity."""
    r, th, ph = state[1:4]
    vr, vth, vph = state[5:8]
    st, ct, sp, cp = math.sin(th), math.cos(th), math.sin(ph), math.cos(ph)
    return np.array([st * cp * vr + r * ct * cp * vth - r * st * sp * vph,
                     st * sp * vr + r * ct * sp * vth + r * st * cp * vph,
                     ct * vr - r * st * vth])


def _flat_render_errors(oracle, world, max_step):
    """Classification mismatches and colour errors of a flat-space render against straight rays."""
    sc = scene_of(gg.Schwarzschild(1e-9), w=40, h=40,
                  integ=gg.RK45Integrator(max_radius=200.0, max_step=max_step, max_steps=40000))
    got = oracle.render(sc)
    o0 = np.array([30.0, 0.0, 0.0])
    mismatched, errs = 0, []
    for i in range(sc.n_rays):
        d = _cartesian_dir(oracle.from_camera(sc, i % sc.width, i // sc.width))
        d /= np.linalg.norm(d)
        best, margin = None, np.inf
        for obj in world.objects:
            c, rad = np.array(obj.center), obj.radius
            tc = np.dot(c - o0, d)
            b = np.linalg.norm(c - o0 - tc * d)                  # impact parameter
            margin = min(margin, abs(b - rad) / rad)
            if tc > 0 and b < rad and (best is None or tc < best[0]):
                best = (tc - math.sqrt(rad * rad - b * b), c, rad, obj)
        if margin < 0.03:
            continue
        if (best is not None) != (got["cls"][i] == gg.GG_CLASS_SCENE):
            mismatched += 1
            continue
        if best is not None:
            t, c, rad, obj = best
            nrm = (o0 + t * d - c) / rad
            facing = 0.5 + 0.5 * max(-np.dot(d, nrm), 0.0)
            errs.append(np.abs(got["color"][i] - np.array(obj.material.albedo) * facing).max())
    return mismatched, np.array(errs)


def test_flat_space_limit_matches_straight_rays(oracle, with_world):
    # M -> 0: geodesics are straight lines, so each camera ray hits what its
    # straight line hits, with the facing ratio of its analytic normal.  The
    # reference's forward-Euler positions bend the paths by O(h), so rays
    # within 3% of a silhouette are left out and the colour error shrinks
    # linearly with the step.
    world = gg.World([gg.Sphere((12.0, 5.0, 0.0), 2.0, gg.Lambertian((0.8, 0.2, 0.1))),
                      gg.Sphere((15.0, -6.0, 3.0), 2.5, gg.Lambertian((0.1, 0.7, 0.3))),
                      gg.Sphere((8.0, 0.0, -7.0), 2.0, gg.Lambertian((0.2, 0.3, 0.9)))])
    with_world(world)
    bad_fine, err_fine = _flat_render_errors(oracle, world, 0.05)
    bad_coarse, err_coarse = _flat_render_errors(oracle, world, 0.2)
    assert bad_fine == 0 and len(err_fine) > 150
    assert err_fine.max() < 0.015
    assert err_coarse.max() / err_fine.max() > 2.5


def test_disk_and_scene_reduces_to_scene_without_emission(oracle, with_world):
    # trace_ray_with_disk_and_scene with a disk of zero density is
    # trace_ray_with_scene: transmission 1 and disk_color(0) = 0.  Its
    # absorption coefficient is the constant 0.1 (tracer.rs:268), so the
    # disk's own absorption field changes nothing.
    with_world(lensed_spheres())
    base = oracle.render(scene_of(gg.Schwarzschild(1.0), w=24, h=24), rgb=True)
    for absorption in (0.1, 5.0):
        dk = gg.AccretionDisk(density_0=0.0, absorption=absorption)
        got = oracle.render(scene_of(gg.Schwarzschild(1.0), disk=dk, w=24, h=24), rgb=True)
        np.testing.assert_array_equal(got["cls"], base["cls"])
        np.testing.assert_array_equal(got["color"], base["color"])
        np.testing.assert_array_equal(got["rgb"], base["rgb"])
    dk1 = oracle.render(scene_of(gg.Schwarzschild(1.0), disk=gg.AccretionDisk(absorption=0.1), w=24, h=24))
    dk5 = oracle.render(scene_of(gg.Schwarzschild(1.0), disk=gg.AccretionDisk(absorption=5.0), w=24, h=24))
    np.testing.assert_array_equal(dk1["color"], dk5["color"])
    np.testing.assert_array_equal(dk1["intensity"], dk5["intensity"])


def test_scene_hits_are_lensed(oracle, with_world):
    # The backdrop light sits straight behind the hole: in flat space it
    # covers the image centre; the hole replaces that with its shadow and
    # moves the light into an Einstein ring around it.
    with_world(lensed_spheres())
    curved = oracle.render(scene_of(gg.Schwarzschild(1.0), w=32, h=32))
    cls = curved["cls"].reshape(32, 32)
    assert cls[16, 16] == gg.GG_CLASS_HORIZON
    ring = curved["color"].reshape(32, 32, 3)
    lit = (cls == gg.GG_CLASS_SCENE) & (ring[..., 0] > 1.0)  # the DiffuseLight (emission 2.0 red)
    assert lit.sum() > 8
    yy, xx = np.nonzero(lit)
    rr = np.hypot(yy - 15.5, xx - 15.5)
    assert rr.min() > 2.0                                     # a ring around the shadow, not a disk


# --- GPU against the oracle -------------------------------------------------------------------

def _compare(got, want, min_agree=0.999):
    agree = got["cls"] == want["cls"]
    assert agree.mean() >= min_agree, agree.mean()
    sc = agree & (want["cls"] == gg.GG_CLASS_SCENE)
    assert sc.sum() > 0
    d = np.abs(got["color"][sc] - want["color"][sc]).max(axis=1) / np.maximum(np.abs(want["color"][sc]).max(axis=1),
                                                                                1e-300)
    # colours of the same hit agree to 1e-9 on nearly every ray, and to 1e-6 on all
    assert np.mean(d <= 1e-9) >= 0.99, np.sort(d)[-5:]
    assert d.max() <= 1e-6, np.sort(d)[-5:]
    if "rgb" in want:
        # Escaped rays meet the procedural sky, which hashes the raw bits of
        # the escape direction (sky.rs:26-40): a last-bit difference can add
        # or remove a star, so those pixels have a looser bound (one ray per
        # pixel here).
        px = np.abs(got["rgb"].astype(int) - want["rgb"].astype(int)).max(axis=1)
        esc = want["cls"] == gg.GG_CLASS_ESCAPED
        assert np.mean(px[~esc] <= 1) >= 0.999
        if esc.any():
            assert np.mean(px[esc] <= 1) >= 0.9


@pytest.mark.gpu
@pytest.mark.parametrize("metric", [gg.Schwarzschild(1.0), gg.Kerr(1.0, 0.9)], ids=["schw", "kerr09"])
@pytest.mark.parametrize("with_disk", [False, True], ids=["scene", "disk_scene"])
def test_gpu_scene_matches_oracle(renderer, oracle, with_world, metric, with_disk):
    world = lensed_spheres()
    with_world(world)
    sc = scene_of(metric, disk=gg.AccretionDisk(r_inner=6.0, r_outer=20.0) if with_disk else None, w=48, h=48)
    want = oracle.render(sc, rgb=True)
    renderer.set_hittables(world)
    try:
        got = renderer.render(sc, "fp64", rgb=True)
        mixed = renderer.render(sc, "mixed", rgb=True)
    finally:
        renderer.set_hittables(None)
    print(metric, with_disk, np.bincount(want["cls"], minlength=5), got["stats"]["kernel_ms"])
    _compare(got, want)
    # mixed precision: near a silhouette an FP32 step can move the hit, elsewhere the same colours
    agree = mixed["cls"] == want["cls"]
    assert agree.mean() >= 0.99
    sc_ = agree & (want["cls"] == gg.GG_CLASS_SCENE)
    rel = np.abs(mixed["color"][sc_] - want["color"][sc_]).max(axis=1)
    assert np.mean(rel <= 1e-4) >= 0.98


@pytest.mark.gpu
def test_gpu_scene_bvh_many_primitives(renderer, oracle, with_world):
    # 400 spheres and quads scattered around the hole: the BVH traversal on
    # the GPU against the linear World::hit of the oracle.
    rng = np.random.default_rng(5)
    world = gg.World()
    mats = [gg.Lambertian((0.8, 0.3, 0.2)), gg.Metal((0.7, 0.7, 0.9)), gg.Dielectric(1.5),
            gg.DiffuseLight((1.5, 1.2, 0.6))]
    for i in range(400):
        c = rng.uniform(-80, 80, 3)
        if np.linalg.norm(c) < 12:
            c *= 12 / np.linalg.norm(c)
        m = mats[i % 4]
        if i % 3 == 0:
            world.add(gg.Quad(tuple(c), tuple(rng.uniform(-4, 4, 3)), tuple(rng.uniform(-4, 4, 3)), m))
        else:
            world.add(gg.Sphere(tuple(c), float(rng.uniform(0.5, 3.0)), m))
    with_world(world)
    sc = scene_of(gg.Kerr(1.0, 0.7), w=40, h=40)
    sc.camera = gg.Camera(position=(0.0, 35.0, 1.2, 0.3), look_at=(-0.9, -0.2, -0.35), up=(0.0, 0.0, 1.0),
                          fov_y_radians=math.radians(70.0), aspect=1.0)
    want = oracle.render(sc, rgb=True)
    renderer.set_hittables(world)
    try:
        got = renderer.render(sc, "fp64", rgb=True)
    finally:
        renderer.set_hittables(None)
    print(np.bincount(want["cls"], minlength=5))
    assert np.sum(want["cls"] == gg.GG_CLASS_SCENE) > 100
    _compare(got, want)


@pytest.mark.gpu
def test_gpu_scene_trace_states_and_module_drivers(renderer, oracle, with_world):
    world = lensed_spheres()
    with_world(world)
    sc = scene_of(gg.Schwarzschild(1.0), w=20, h=20)
    states = np.stack([oracle.from_camera(sc, i % 20, i // 20) for i in range(400)])
    want = oracle.trace_states(sc, states)
    renderer.set_hittables(world)
    try:
        got = renderer.trace_states(sc, states, "fp64")
    finally:
        renderer.set_hittables(None)
    _compare(got, want)
    img = gg.render_with_scene(sc.metric, sc.camera, world, 20, 20)
    ref = oracle.render(scene_of(gg.Schwarzschild(1.0), w=20, h=20), rgb=True)["rgb"].reshape(20, 20, 3)
    assert np.mean(np.abs(img.astype(int) - ref.astype(int)).max(axis=2) <= 1) >= 0.99
    dk = gg.AccretionDisk()
    img2 = gg.render_with_disk_and_scene(sc.metric, sc.camera, dk, world, 20, 20)
    ref2 = oracle.render(scene_of(gg.Schwarzschild(1.0), disk=dk, w=20, h=20), rgb=True)["rgb"].reshape(20, 20, 3)
    assert np.mean(np.abs(img2.astype(int) - ref2.astype(int)).max(axis=2) <= 1) >= 0.99


@pytest.mark.gpu
def test_gpu_scene_unsupported_and_reset(renderer):
    world = lensed_spheres()
    sc = scene_of(gg.Schwarzschild(1.0), w=16, h=16)
    base = renderer.render(sc, "fp64", rgb=True)
    lib, ctx = renderer.lib, renderer.ctx
    bad = gg.World([gg.Sphere((0, 0, 0), 1.0, gg.Metal((1, 1, 1), fuzz=0.2))]).to_c()
    assert lib.gg_set_hittables(ctx, 1, ctypes.cast(bad, ctypes.c_void_p)) != 0
    renderer.set_hittables(world)
    try:
        thin = scene_of(gg.Schwarzschild(1.0), disk=gg.AccretionDisk(kind=gg.GG_DISK_THIN), w=16, h=16)
        with pytest.raises(gg.GargantuaError):
            renderer.render(thin, "fp64")
        with pytest.raises(gg.GargantuaError):
            renderer.render_batch([sc, sc], "fp64")
        lit = renderer.render(sc, "fp64", rgb=True)
        assert np.any(lit["cls"] == gg.GG_CLASS_SCENE)
        # an invalid primitive leaves the geometry already set (gargantua.h, gg_set_hittables)
        with pytest.raises(gg.GargantuaError):
            renderer.set_hittables(gg.World([gg.Sphere((0, 0, 0), 0.0, gg.Lambertian((1, 1, 1)))]))
        kept = renderer.render(sc, "fp64", rgb=True)
        np.testing.assert_array_equal(kept["cls"], lit["cls"])
        np.testing.assert_array_equal(kept["rgb"], lit["rgb"])
    finally:
        renderer.set_hittables(None)
    again = renderer.render(sc, "fp64", rgb=True)
    np.testing.assert_array_equal(again["cls"], base["cls"])
    np.testing.assert_array_equal(again["rgb"], base["rgb"])
