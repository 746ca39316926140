"""The command-line renderer (gargantua_amd/cli.py) against the reference CLI's
arguments and dispatch (crates/gr-renderer/src/main.rs:17-169)."""
import math

import numpy as np
import pytest

import gargantua_amd as gg
from gargantua_amd import cli, png


def test_defaults_are_the_reference_cli_defaults():
    a = cli.parse_args([])
    assert (a.metric, a.mass, a.spin, a.width, a.height) == ("schwarzschild", 1.0, 0.9, 800, 600)
    assert (a.disk, a.scene, a.r_inner, a.r_outer, a.scale_height) == (False, "none", 6.0, 20.0, 1.0)
    assert (a.samples, a.quiet, a.hdri, a.hdri_intensity, a.output) == (1, False, None, 1.0, "out.png")


@pytest.mark.parametrize("argv", [["--metric", "minkowski"], ["--scene", "cornell"],
                                  ["--hdri", "/nonexistent/sky.hdr"]])
def test_unknown_inputs_exit_1(argv, tmp_path):
    # main.rs:91-99, 113-120, 131-134 exit(1) before rendering
    with pytest.raises(SystemExit) as e:
        cli.run(cli.parse_args(argv + ["--output", str(tmp_path / "x.png")]))
    assert e.value.code == 1


@pytest.mark.gpu
@pytest.mark.parametrize("argv", [["--metric", "kerr", "--disk", "--scene", "lensed-spheres"],
                                  ["--metric", "schwarzschild", "--samples", "2"]])
def test_gpu_cli_image_matches_oracle(oracle, tmp_path, argv):
    out = tmp_path / "out.png"
    args = cli.parse_args(argv + ["--width", "40", "--height", "30", "--quiet", "--output", str(out)])
    img = cli.run(args)
    np.testing.assert_array_equal(png.read(out), img)
    metric = gg.Kerr(args.mass, args.spin) if args.metric == "kerr" else gg.Schwarzschild(args.mass)
    cam = gg.Camera(position=(0.0, 30.0, math.pi / 2, 0.0), look_at=(-1.0, 0.0, 0.0), up=(0.0, 0.0, 1.0),
                    fov_y_radians=math.radians(60.0), aspect=40 / 30)
    scene = gg.Scene(metric=metric, camera=cam, disk=gg.AccretionDisk() if args.disk else None,
                     integrator=gg.RK45Integrator(max_radius=200.0, max_steps=5000), width=40, height=30,
                     samples_per_axis=args.samples)
    world = gg.configs.lensed_spheres_world() if args.scene == "lensed-spheres" else None
    assert oracle.set_hittables(world) == 0
    try:
        ref = oracle.render(scene, rgb=True)
    finally:
        oracle.set_hittables(None)
    want = ref["rgb"].reshape(30, 40, 3)
    close = np.abs(img.astype(int) - want.astype(int)).max(axis=2) <= 1
    # Escaped rays see the bit-hashed procedural sky (sky.rs:26-40): a last-bit
    # difference in the escape direction can add or remove a star, so pixels
    # with an escaped sub-ray are held to a looser bound.
    spa = args.samples
    esc = (ref["cls"].reshape(30, spa, 40, spa) == gg.GG_CLASS_ESCAPED).any(axis=(1, 3))
    assert close[~esc].all()
    if esc.any():
        assert close[esc].mean() >= 0.9


@pytest.mark.gpu
def test_gpu_cli_hdri_file(oracle, tmp_path):
    # --hdri / --hdri-intensity (main.rs:66-72, 91-99): a Radiance file through HdriSky.from_file
    from gargantua_amd import hdr
    th = (np.arange(16) + 0.5) / 16 * math.pi
    ph = (np.arange(32) + 0.5) / 32 * 2 * math.pi
    sky = np.stack([0.6 + 0.4 * np.cos(th)[:, None] + 0 * ph[None, :],
                    0.5 + 0.3 * np.sin(ph)[None, :] * np.sin(th)[:, None],
                    0.3 + 0.2 * np.cos(2 * ph)[None, :] + 0 * th[:, None]], axis=-1)
    path = tmp_path / "sky.hdr"
    hdr.write(path, sky)
    out = tmp_path / "o.png"
    args = cli.parse_args(["--metric", "schwarzschild", "--width", "32", "--height", "24", "--hdri", str(path),
                           "--hdri-intensity", "1.5", "--quiet", "--output", str(out)])
    img = cli.run(args)
    cam = gg.Camera(position=(0.0, 30.0, math.pi / 2, 0.0), look_at=(-1.0, 0.0, 0.0), up=(0.0, 0.0, 1.0),
                    fov_y_radians=math.radians(60.0), aspect=32 / 24)
    scene = gg.Scene(metric=gg.Schwarzschild(1.0), camera=cam,
                     integrator=gg.RK45Integrator(max_radius=200.0, max_steps=5000), width=32, height=24)
    oracle.set_sky_hdri(hdr.read(path), 1.5)
    try:
        want = oracle.render(scene, rgb=True)["rgb"].reshape(24, 32, 3)
    finally:
        oracle.set_sky_hdri(None)
    assert np.abs(img.astype(int) - want.astype(int)).max() <= 1
