"""Batched scenes in one persistent launch (gg_render_batch), its hipGraph
replay and the launch-shape autotuner.  Every ray's arithmetic is the same
in a batch as alone, so batch outputs equal single-scene outputs bit for bit
in the FP64 precisions.  The mixed kernel takes FP32 steps only in wavefronts
whose lanes are all FP32-eligible, so which eligible steps run in FP32 depends
on the lane grouping (and its fast-FP64 steps round h differently from exact
FP64): final states agree across launches to ~1e-12 except where a photon-orbit
approach amplifies the difference, and a rare ray can move a late decision (the
step that crosses the escape radius: one attempt more or less, 1 ray in ~4e5
observed).  So classes must agree, attempt counts on all but 1e-4 of the rays
(by at most one attempt), and final states within 1e-6 (the mixed-mode bar)
on all but 1e-4."""
import numpy as np
import pytest

import gargantua_amd as gg
from tests.parity import state_rel_err

pytestmark = pytest.mark.gpu

KEYS = ("cls", "r", "theta", "phi", "intensity", "attempts", "accepted")


def assert_same(a, b, precision, what=""):
    if precision != "mixed":
        for k in KEYS:
            np.testing.assert_array_equal(a[k], b[k], err_msg=f"{what}{k}")
        return
    np.testing.assert_array_equal(a["cls"], b["cls"], err_msg=f"{what}cls")
    same = (a["attempts"] == b["attempts"]) & (a["accepted"] == b["accepted"])
    assert np.mean(~same) <= 1e-4, f"{what}attempts differ on {np.count_nonzero(~same)} rays"
    d = np.abs(a["attempts"].astype(np.int64) - b["attempts"].astype(np.int64))
    assert d.max() <= 1, f"{what}attempts differ by {d.max()}"
    err = state_rel_err(a, b)[same]
    ierr = np.abs(a["intensity"] - b["intensity"])[same] / np.maximum(np.abs(b["intensity"][same]), 1e-300)
    ok = (err <= 1e-6) & ((ierr <= 1e-6) | (b["intensity"][same] == 0))
    assert np.mean(~ok) <= 1e-4, f"{what}final states differ by more than 1e-6 on {np.count_nonzero(~ok)} rays"


def _scenes():
    return gg.configs.batch_64(seed=7, n=6, resolutions=(24, 40, 64))


@pytest.mark.parametrize("precision", ["fp64", "mixed", "fp64fast"])
def test_batch_equals_single_renders(renderer, precision):
    scenes = _scenes()
    outs, st = renderer.render_batch(scenes, precision)
    assert st["rays"] == sum(s.n_rays for s in scenes)
    for sc, o in zip(scenes, outs):
        ref = renderer.render(sc, precision, state=False)
        assert_same(o, ref, precision)


def test_graph_replays_match_eager(renderer):
    scenes = _scenes()
    renderer.set_launch(graph=2)
    eager, _ = renderer.render_batch(scenes, "mixed")
    renderer.set_launch(graph=1)
    for rep in range(3):
        outs, _ = renderer.render_batch(scenes, "mixed")
        for a, b in zip(outs, eager):
            assert_same(a, b, "mixed", f"replay {rep}: ")
        renderer.render(scenes[rep], "fp64", state=False)   # eager work between replays
    # new scene contents through the same graph (same shapes, different parameters)
    other = gg.configs.batch_64(seed=8, n=6, resolutions=(24, 40, 64))
    for s, t in zip(other, scenes):
        s.width = s.height = t.width
        s.camera.aspect = 1.0
    outs, _ = renderer.render_batch(other, "mixed")
    for sc, o in zip(other, outs):
        ref = renderer.render(sc, "mixed", state=False)
        assert_same(o, ref, "mixed")


def test_autotune_picks_a_valid_shape(renderer):
    scenes = _scenes()
    best, ms = renderer.autotune(scenes, "mixed",
                                 candidates=[(256, 4, 16, 2), (128, 8, 8, 2), (64, 16, 32, 2)])
    assert ms > 0 and best["block_threads"] in (64, 128, 256)
    assert renderer.get_launch()["block_threads"] == best["block_threads"]
    outs, _ = renderer.render_batch(scenes, "mixed")
    ref = renderer.render(scenes[0], "mixed", state=False)
    assert_same(outs[0], ref, "mixed")
    renderer.set_launch(256, 2, 16, 1)


def test_batch_rejects_mixed_kinds(renderer):
    scenes = _scenes()[:2]
    scenes[1].disk = None
    with pytest.raises(gg.GargantuaError, match="share metric kind"):
        renderer.render_batch(scenes, "fp64")


@pytest.mark.parametrize("precision", ["fp64", "mixed", "fp64fast"])
def test_every_ray_retired_once_with_refills(renderer, precision):
    """More rays than resident lanes, so lanes are refilled many times: every
    output slot is written (sentinels gone), the library's own retirement
    count matches, and a short refill interval changes no bit."""
    scene = gg.configs.kerr_1024(768, 512)
    out = gg.alloc_outputs(scene.n_rays, state=False)
    out["cls"][:] = 255
    out["attempts"][:] = 0xFFFFFFFF
    renderer.set_launch(256, 1, 16, 1)
    a = renderer.render(scene, precision, outputs=out)
    assert a["stats"]["rays"] == scene.n_rays
    assert not np.any(a["cls"] == 255)
    assert a["stats"]["attempts"] == int(a["attempts"].astype(np.uint64).sum())
    renderer.set_launch(128, 2, 4, 1)
    b = renderer.render(scene, precision, state=False)
    renderer.set_launch(256, 2, 16, 1)
    assert_same(a, b, precision)


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_learned_ray_order_changes_no_result(renderer, precision):
    """gg_launch.order = 1 queues rays longest first by the previous render's
    attempt counts: only the schedule changes (FP64 bitwise, mixed to FP32
    rounding), and every ray is still traced."""
    scene = gg.configs.kerr_1024(256, 256)
    renderer.set_launch(256, 2, 16, 1, 0)
    ref = renderer.render(scene, precision, state=False)
    renderer.set_launch(256, 2, 16, 1, 1)
    first = renderer.render(scene, precision, state=False)     # learns the order
    second = renderer.render(scene, precision, state=False)    # uses it
    renderer.set_launch(256, 2, 16, 1, 0)
    for o in (first, second):
        assert o["stats"]["rays"] == scene.n_rays
        assert o["stats"]["attempts"] == int(o["attempts"].astype(np.uint64).sum())
        assert_same(o, ref, precision)


def test_learned_ray_order_in_batches(renderer):
    scenes = _scenes()
    renderer.set_launch(256, 2, 16, 1, 0)
    ref, _ = renderer.render_batch(scenes, "fp64")
    renderer.set_launch(256, 2, 16, 1, 1)
    for rep in range(3):                      # learn, then replay the graph with the order
        outs, st = renderer.render_batch(scenes, "fp64")
        assert st["rays"] == sum(s.n_rays for s in scenes)
        for a, b in zip(outs, ref):
            assert_same(a, b, "fp64")
    renderer.set_launch(256, 2, 16, 1, 0)


def test_default_launch_shape_by_ray_count(renderer):
    """blocks_per_cu = 0 (the default) launches one wavefront per SIMD below
    3 rays per lane at two (a 512^2 image: 4 blocks of 64 threads per CU)
    and three per SIMD above (1024^2: 12 per CU); the FP64 outputs are those
    of an explicit launch shape, bit for bit."""
    small = gg.configs.kerr_1024(512, 512)
    large = gg.configs.kerr_1024(1024, 1024)
    renderer.set_launch()
    assert renderer.get_launch()["blocks_per_cu"] == 0
    a_small = renderer.render(small, "fp64", state=False)
    a_large = renderer.render(large, "fp64", state=False)
    gs, gl = a_small["stats"]["grid_blocks"], a_large["stats"]["grid_blocks"]
    assert a_small["stats"]["block_threads"] == 64 and gs % 4 == 0 and gl == 3 * gs, (gs, gl)
    renderer.set_launch(64, 8, 64, 1, 0)
    for sc, a in ((small, a_small), (large, a_large)):
        b = renderer.render(sc, "fp64", state=False)
        assert b["stats"]["grid_blocks"] == 2 * gs
        assert_same(a, b, "fp64")
    renderer.set_launch(256, 2, 16, 1, 0)
