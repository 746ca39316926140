"""Edge cases of the persistent launch against the oracle: images smaller
than one wavefront and widths or heights that are not multiples of the 8x8
ray tiles (untiled numbering), a batch of such ragged scenes in one graph
launch against single renders, and caller states that are invalid (NaN,
r <= 0), start captured or escaped, or run out of attempts."""
import math

import numpy as np
import pytest

import gargantua_amd as gg
from tests.parity import state_rel_err
from tests.test_gpu_batch import assert_same

pytestmark = pytest.mark.gpu

SIZES = [(1, 1), (3, 2), (7, 5), (9, 8), (8, 8), (16, 3), (65, 1)]


@pytest.mark.parametrize("size", SIZES, ids=[f"{w}x{h}" for w, h in SIZES])
def test_tiny_and_ragged_images_match_oracle(renderer, oracle, size):
    w, h = size
    # jittered sub-pixel positions: no ray sits exactly in the camera's meridional
    # plane (L_z = 0), where the reference's pole reflection keeps v^theta and a
    # last-bit difference can add a reflection (phi + pi; covered in test_gpu_configs)
    schw = gg.Scene(metric=gg.Schwarzschild(1.0),
                    camera=gg.configs.inclination_camera(50.0, math.radians(80.0), math.radians(30.0), w / h),
                    disk=gg.configs.thin_disk(6.0, 20.0, 3.0), integrator=gg.configs.rtol_integrator(1000.0, 20000),
                    width=w, height=h, jitter=True, seed=11)
    for scene in (gg.configs.kerr_1024(w, h), schw):
        g = renderer.render(scene, "fp64")
        o = oracle.render(scene)
        np.testing.assert_array_equal(g["cls"], o["cls"])
        same = g["attempts"] == o["attempts"]
        assert np.count_nonzero(~same) <= 1, f"{w}x{h}: attempts differ on {np.count_nonzero(~same)} rays"
        assert np.all(state_rel_err(g, o)[same] <= 1e-6)
        m = renderer.render(scene, "mixed")
        np.testing.assert_array_equal(m["cls"], g["cls"])
        assert g["stats"]["rays"] == scene.n_rays == w * h


def test_ragged_batch_equals_single_renders(renderer):
    scenes = [gg.configs.kerr_thin_scene(w, h, spin=a, r_camera=rc, inclination_deg=inc, fov_deg=30.0)
              for (w, h), a, rc, inc in zip([(1, 1), (7, 5), (64, 64), (9, 8), (3, 130)],
                                            [0.1, 0.5, 0.9, 0.998, 0.3], [40.0, 80.0, 100.0, 60.0, 200.0],
                                            [10.0, 45.0, 60.0, 85.0, 30.0])]
    for precision in ("fp64", "mixed"):
        outs, st = renderer.render_batch(scenes, precision)
        assert st["rays"] == sum(s.n_rays for s in scenes)
        for sc, o in zip(scenes, outs):
            assert_same(o, renderer.render(sc, precision, state=False), precision, f"{sc.width}x{sc.height} ")


def _edge_states():
    nan = math.nan
    return np.array([
        [0, 30.0, 1.2, 0.3, 1.0, -0.9, 0.01, 0.002, 0],    # an ordinary inbound ray
        [0, nan, 1.2, 0.3, 1.0, -0.9, 0.01, 0.002, 0],     # NaN position: halves h until out of attempts
        [0, -1.0, 1.2, 0.3, 1.0, -0.9, 0.01, 0.002, 0],    # r <= 0
        [0, 1.2, 1.2, 0.3, 1.0, -0.9, 0.01, 0.002, 0],     # inside the horizon
        [0, 2000.0, 1.2, 0.3, 1.0, 0.9, 0.0, 0.0, 0],      # beyond the escape radius
        [0, 80.0, 1.5, 0.0, 1.0, -0.5, 0.0, 0.001, 0],     # a long orbit, cut by the attempt cap below
        [0, 50.0, 1e-9, 0.0, 1.0, -0.7, 0.001, 0.0, 0],    # on the polar axis
    ])


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
@pytest.mark.parametrize("max_steps", [20000, 3])
def test_edge_states_match_oracle(renderer, oracle, precision, max_steps):
    scene = gg.configs.kerr_1024(1, 1)
    scene.integrator.max_steps = max_steps
    states = _edge_states()
    g = renderer.trace_states(scene, states, precision)
    o = oracle.trace_states(scene, states)
    np.testing.assert_array_equal(g["cls"], o["cls"])
    np.testing.assert_array_equal(g["attempts"], o["attempts"])
    for k in ("r", "theta", "phi"):
        np.testing.assert_array_equal(np.isnan(g[k]), np.isnan(o[k]), err_msg=k)
    fin = np.isfinite(o["r"])
    assert np.all(state_rel_err({k: g[k][fin] for k in ("r", "theta", "phi")},
                                {k: o[k][fin] for k in ("r", "theta", "phi")}) <= 1e-6)
    if max_steps == 3:
        assert g["cls"][5] == gg.GG_CLASS_MAX_STEPS
