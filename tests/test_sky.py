"""HdriSky (crates/raytracer/src/curved/sky.rs:43-122): the oracle's sample
pinned by the reference's own sky tests (sky.rs:131-163) and by closed forms
of the equirectangular bilinear map; the Radiance reader by round trips; the
GPU shading of escaped rays against the oracle's render with the same sky."""
import math

import numpy as np
import pytest

import gargantua_amd as gg
from gargantua_amd import hdr


def _dir(theta, phi):
    return np.array([math.sin(theta) * math.cos(phi), math.sin(theta) * math.sin(phi), math.cos(theta)])


# --- the reference's tests (sky.rs), restated -------------------------------

def test_hdri_sky_constant_color(oracle):
    # sky.rs:131-139
    px = np.tile(np.array([0.4, 0.6, 0.8], np.float32), (2, 4, 1))
    c = oracle.hdri_sample(px, [1.0, 0.0, 0.0])
    assert np.all(np.abs(c - [0.4, 0.6, 0.8]) < 1e-4)


def test_hdri_sky_pole_is_top_row(oracle):
    # sky.rs:141-152
    px = np.zeros((2, 4, 3), np.float32)
    px[0] = [1.0, 0.0, 0.0]
    px[1] = [0.0, 0.0, 1.0]
    top = oracle.hdri_sample(px, [0.0, 0.0, 1.0])
    assert top[0] > 0.9 and top[2] < 0.1
    bot = oracle.hdri_sample(px, [0.0, 0.0, -1.0])
    assert bot[2] > 0.9 and bot[0] < 0.1


def test_hdri_sky_intensity_scales(oracle):
    # sky.rs:154-160
    px = np.tile(np.array([0.5, 0.25, 0.1], np.float32), (2, 2, 1))
    c = oracle.hdri_sample(px, [1.0, 0.0, 0.0], intensity=2.0)
    assert abs(c[0] - 1.0) < 1e-4 and abs(c[1] - 0.5) < 1e-4


# --- closed forms of the map --------------------------------------------------

W, H = 64, 32


def _ramps():
    """r = column index, g = row index, b = 1: bilinear reproduces (fx, fy) exactly."""
    px = np.zeros((H, W, 3), np.float32)
    px[..., 0] = np.arange(W, dtype=np.float32)[None, :]
    px[..., 1] = np.arange(H, dtype=np.float32)[:, None]
    px[..., 2] = 1.0
    return px


def test_hdri_bilinear_reproduces_texel_coordinates(oracle):
    # u = phi / 2pi + 1/2 -> fx = u W; v = theta / pi -> fy = clamp(v H, 0, H - 1).
    # A swapped u/v, a lost 1/2 offset, a mirrored row order or tx/ty
    # exchanged each moves (r, g) away from (fx, fy).
    px = _ramps()
    rng = np.random.default_rng(7)
    for _ in range(200):
        theta = rng.uniform(0.05, math.pi - 0.05)
        phi = rng.uniform(-math.pi + 0.2, math.pi - 0.2)       # away from the seam column
        fx = (phi / (2 * math.pi) + 0.5) * W
        fy = min(theta / math.pi * H, H - 1.0)
        if fx >= W - 1:
            continue
        c = oracle.hdri_sample(px, _dir(theta, phi))
        assert abs(c[0] - fx) < 1e-4 * W and abs(c[1] - fy) < 1e-4 * H, (theta, phi, c, fx, fy)
        assert c[2] == 1.0


def test_hdri_wraps_in_x_and_clamps_in_y(oracle):
    px = _ramps()
    # Seam: fx in (W - 1, W) blends column W - 1 with column 0.
    phi = math.pi - 0.25 * (2 * math.pi / W)                    # fx = W - 0.25
    c = oracle.hdri_sample(px, _dir(math.pi / 2, phi))
    assert abs(c[0] - 0.25 * (W - 1)) < 1e-3
    # -z pole: v = 1, fy = H clamped to H - 1: the last row, no blending past it.
    c = oracle.hdri_sample(px, [0.0, 0.0, -1.0])
    assert c[1] == H - 1
    # phi = atan2(+-0, x < 0) = +-pi is u = 0 or 1: both land on column 0.
    a = oracle.hdri_sample(px, [-1.0, 0.0, 0.0])
    b = oracle.hdri_sample(px, [-1.0, -0.0, 0.0])
    assert a[0] == 0.0 and b[0] == 0.0


def test_hdri_texel_centres_are_exact(oracle):
    rng = np.random.default_rng(11)
    px = rng.uniform(0, 4, (H, W, 3)).astype(np.float32)
    # +z: u = 1/2, v = 0 -> texel (0, W/2) exactly; -z -> (H-1, W/2).
    np.testing.assert_array_equal(oracle.hdri_sample(px, [0.0, 0.0, 1.0]), px[0, W // 2].astype(np.float64))
    np.testing.assert_array_equal(oracle.hdri_sample(px, [0.0, 0.0, -1.0]), px[H - 1, W // 2].astype(np.float64))
    # Equator along +x: u = 1/2, v = 1/2 -> texel (H/2, W/2).
    np.testing.assert_array_equal(oracle.hdri_sample(px, [1.0, 0.0, 0.0]), px[H // 2, W // 2].astype(np.float64))


def test_hdri_intensity_is_an_f32_product(oracle):
    px = np.full((2, 2, 3), 0.1, np.float32)
    c = oracle.hdri_sample(px, [1.0, 0.0, 0.0], intensity=3.0)
    assert c[0] == float(np.float32(0.1) * np.float32(3.0))


# --- Radiance files ---------------------------------------------------------------

def test_rgbe_decode_closed_form():
    # c * 2^(e - 136); e = 0 is black.
    t = np.array([[128, 64, 1, 129], [255, 0, 0, 136], [7, 7, 7, 0]], np.uint8)
    got = hdr.rgbe_to_float(t)
    np.testing.assert_array_equal(got[0], [1.0, 0.5, 1.0 / 128.0])
    np.testing.assert_array_equal(got[1], [255.0, 0.0, 0.0])
    np.testing.assert_array_equal(got[2], [0.0, 0.0, 0.0])


@pytest.mark.parametrize("width,rle", [(5, True), (40, True), (40, False), (300, True)])
def test_radiance_round_trip(tmp_path, width, rle):
    rng = np.random.default_rng(width)
    px = rng.uniform(0, 50, (7, width, 3))
    px[:, : width // 2] = px[:, :1]                          # long runs for the run-length coder
    px[0, 0] = 0.0
    p = tmp_path / "sky.hdr"
    hdr.write(p, px, rle=rle)
    got = hdr.read(p)
    assert got.shape == px.shape and got.dtype == np.float32
    # One RGBE quantum: the shared exponent keeps 8 bits under the largest channel.
    err = np.abs(got - px)
    assert np.all(err <= px.max(axis=-1, keepdims=True) / 128.0 + 1e-30)
    # Decoding what was decoded and re-encoded is exact (RGBE values are fixed points).
    hdr.write(p, got, rle=rle)
    np.testing.assert_array_equal(hdr.read(p), got)
    sky = gg.HdriSky.from_file(p).with_intensity(2.0)
    assert sky.width == width and sky.height == 7 and sky.intensity == 2.0


def test_radiance_rejects_bad_files(tmp_path):
    p = tmp_path / "bad.hdr"
    p.write_bytes(b"P6\n1 1\n255\n\x00\x00\x00")
    with pytest.raises(ValueError):
        hdr.read(p)
    p.write_bytes(b"#?RADIANCE\nFORMAT=32-bit_rle_xyze\n\n-Y 1 +X 1\n\x00\x00\x00\x00")
    with pytest.raises(ValueError):
        hdr.read(p)
    p.write_bytes(b"#?RADIANCE\n\n+Y 1 +X 1\n\x00\x00\x00\x00")
    with pytest.raises(ValueError):
        hdr.read(p)


def test_hdri_sky_shape_checks():
    with pytest.raises(ValueError):
        gg.HdriSky(np.zeros((4, 3), np.float32))
    with pytest.raises(ValueError):
        gg.HdriSky.from_pixels(4, 2, np.zeros(7 * 3, np.float32))
    s = gg.HdriSky.from_pixels(4, 2, [[0.4, 0.6, 0.8]] * 8)
    assert s.pixels.shape == (2, 4, 3)


# --- GPU: escaped rays shaded from the map -----------------------------------------------

def _sky_texture():
    # Smooth and non-symmetric, so that a wrong mapping shows in the image.
    th = (np.arange(H) + 0.5) / H * math.pi
    ph = (np.arange(W) + 0.5) / W * 2 * math.pi - math.pi
    px = np.zeros((H, W, 3), np.float32)
    px[..., 0] = 0.6 + 0.4 * np.cos(th)[:, None] * np.ones(W)[None]
    px[..., 1] = 0.5 + 0.3 * np.sin(ph)[None, :] * np.sin(th)[:, None]
    px[..., 2] = 0.25 + 0.2 * np.cos(2 * ph)[None, :]
    return gg.HdriSky(px, intensity=0.8)


def _scene(metric, disk, spa):
    th = math.radians(80.0)
    cam = gg.Camera(position=(0.0, 30.0, th, 0.0), look_at=(-math.sin(th), 0.0, -math.cos(th)), up=(0.0, 0.0, 1.0),
                    fov_y_radians=math.radians(40.0), aspect=4.0 / 3.0)
    return gg.Scene(metric=metric, camera=cam, disk=disk, integrator=gg.RK45Integrator(max_radius=200.0,
                                                                                    max_steps=5000),
                    width=24, height=18, samples_per_axis=spa)


@pytest.mark.gpu
@pytest.mark.parametrize("metric,disk,spa", [
    (gg.Schwarzschild(1.0), None, 1),
    (gg.Kerr(1.0, 0.9), gg.AccretionDisk(kind=gg.GG_DISK_THIN), 2),
])
def test_gpu_hdri_render_matches_oracle(renderer, oracle, metric, disk, spa):
    sky = _sky_texture()
    sc = _scene(metric, disk, spa)
    oracle.set_sky_hdri(sky.pixels, sky.intensity)
    try:
        want = oracle.render(sc, rgb=True)
    finally:
        oracle.set_sky_hdri(None)
    assert np.any(want["cls"] == gg.GG_CLASS_ESCAPED)
    renderer.set_sky_hdri(sky)
    try:
        got = renderer.render(sc, "fp64", rgb=True)
    finally:
        renderer.set_sky_hdri(None)
    np.testing.assert_array_equal(got["cls"], want["cls"])
    d = np.abs(got["rgb"].astype(int) - want["rgb"].astype(int))
    assert d.max() <= 1
    assert np.mean(d == 0) >= 0.999


@pytest.mark.gpu
def test_gpu_hdri_reset_restores_the_procedural_sky(renderer):
    sc = _scene(gg.Schwarzschild(1.0), None, 1)
    base = renderer.render(sc, "fp64", rgb=True)["rgb"].copy()
    renderer.set_sky_hdri(_sky_texture())
    lit = renderer.render(sc, "fp64", rgb=True)["rgb"].copy()
    renderer.set_sky_hdri(None)
    again = renderer.render(sc, "fp64", rgb=True)["rgb"]
    assert not np.array_equal(base, lit)
    np.testing.assert_array_equal(base, again)


@pytest.mark.gpu
def test_gpu_hdri_constant_sky_gives_its_colour(renderer):
    # Every escaped pixel of a constant sky is encode_srgb of that colour.
    sc = _scene(gg.Schwarzschild(1.0), None, 1)
    renderer.set_sky_hdri(gg.HdriSky(np.tile(np.array([0.4, 0.6, 0.8], np.float32), (3, 5, 1))))
    try:
        o = renderer.render(sc, "fp64", rgb=True)
    finally:
        renderer.set_sky_hdri(None)
    esc = o["cls"] == gg.GG_CLASS_ESCAPED

    def srgb(c):
        s = 12.92 * c if c <= 0.0031308 else 1.055 * c ** (1 / 2.4) - 0.055
        return int(round(s * 255.0))
    want = [srgb(float(np.float32(v))) for v in (0.4, 0.6, 0.8)]
    assert esc.any()
    np.testing.assert_array_equal(o["rgb"][esc], np.tile(want, (int(esc.sum()), 1)))


@pytest.mark.gpu
def test_gpu_hdri_module_render_takes_options_sky(renderer, oracle):
    sky = _sky_texture()
    sc = _scene(gg.Schwarzschild(1.0), None, 1)
    img = gg.render(sc.metric, sc.camera, sc.width, sc.height, gg.RenderOptions(sky=sky))
    oracle.set_sky_hdri(sky.pixels, sky.intensity)
    try:
        want = oracle.render(sc, rgb=True)["rgb"].reshape(sc.height, sc.width, 3)
    finally:
        oracle.set_sky_hdri(None)
    assert np.abs(img.astype(int) - want.astype(int)).max() <= 1


@pytest.mark.gpu
def test_gpu_hdri_bad_args(renderer):
    lib, ctx = renderer.lib, renderer.ctx
    px = np.zeros(12, np.float32)
    assert lib.gg_set_sky_hdri(ctx, 2, 0, px.ctypes.data, 1.0) != 0
    assert lib.gg_set_sky_hdri(ctx, 2, 2, None, 1.0) != 0
    assert lib.gg_set_sky_hdri(ctx, 2, 2, px.ctypes.data, float("inf")) != 0
    assert lib.gg_set_sky_hdri(None, 0, 0, None, 1.0) != 0


# --- PNG output of the RgbImage ----------------------------------------------------------

def test_png_known_answers():
    from gargantua_amd import png
    b = png.encode(np.zeros((1, 1, 3), np.uint8))
    assert b[:8] == b"\x89PNG\r\n\x1a\n"
    # The IEND chunk of every PNG: zero length, type, CRC-32 of "IEND" = AE 42 60 82.
    assert b[-12:] == b"\x00\x00\x00\x00IEND\xaeB`\x82"
    # IHDR of a 1x1 8-bit RGB image, with its CRC-32 90 77 53 DE.
    assert b[8:33] == bytes.fromhex("0000000d494844520000000100000001080200000090775" "3de")


def test_png_round_trip_and_filters(tmp_path):
    import struct
    import zlib
    from gargantua_amd import png
    rng = np.random.default_rng(3)
    img = rng.integers(0, 256, (9, 13, 3), dtype=np.uint8)
    p = tmp_path / "x.png"
    png.write(p, img)
    np.testing.assert_array_equal(png.read(p), img)
    # Scanlines stored with filters 1..4 (sub, up, average, Paeth) decode to the same image.
    h, w = img.shape[:2]
    flat = img.reshape(h, 3 * w).astype(np.int32)
    rows = []
    for y in range(h):
        f = 1 + y % 4
        prev = flat[y - 1] if y else np.zeros(3 * w, np.int32)
        enc = np.zeros(3 * w, np.int32)
        for x in range(3 * w):
            a = flat[y, x - 3] if x >= 3 else 0
            c = prev[x - 3] if x >= 3 else 0
            b = prev[x]
            if f == 1:
                pred = a
            elif f == 2:
                pred = b
            elif f == 3:
                pred = (a + b) // 2
            else:
                pp = a + b - c
                pa, pb, pc = abs(pp - a), abs(pp - b), abs(pp - c)
                pred = a if (pa <= pb and pa <= pc) else (b if pb <= pc else c)
            enc[x] = (flat[y, x] - pred) & 0xFF
        rows.append(bytes([f]) + enc.astype(np.uint8).tobytes())
    ihdr = struct.pack(">IIBBBBB", w, h, 8, 2, 0, 0, 0)
    body = png._chunk(b"IHDR", ihdr) + png._chunk(b"IDAT", zlib.compress(b"".join(rows))) + png._chunk(b"IEND", b"")
    np.testing.assert_array_equal(png.decode(png.SIGNATURE + body), img)
    with pytest.raises(ValueError):
        png.decode(png.SIGNATURE + body[:-1] + b"\x00")          # IEND CRC broken
