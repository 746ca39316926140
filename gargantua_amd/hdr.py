"""Radiance RGBE (.hdr) files for HdriSky.from_file.

HdriSky::from_file (crates/raytracer/src/curved/sky.rs:57-70) decodes with
the `image` crate (0.25.10, Cargo.lock:626-629) and keeps the f32 RGB
values it yields.  That crate's Radiance decoder maps an RGBE texel to
(c_i * 2^(e - 136)) for e != 0 and to black for e = 0, and reads the
header (magic line, variables, blank line, resolution line "-Y H +X W"),
then each scanline either flat or in the adaptive run-length form (2, 2,
W_hi, W_lo, then the four channels one after another, each as runs: a
count byte above 128 repeats the next byte count - 128 times, any other
count copies that many literal bytes).  The same is restated here in
numpy; `write` is its inverse for fixtures and for callers that build a
sky in memory.
"""
from __future__ import annotations

import numpy as np


def rgbe_to_float(rgbe: np.ndarray) -> np.ndarray:
    """[..., 4] uint8 RGBE texels -> [..., 3] float32 linear RGB."""
    rgbe = np.asarray(rgbe, dtype=np.uint8)
    e = rgbe[..., 3].astype(np.float32)
    scale = np.where(rgbe[..., 3] == 0, np.float32(0.0), np.exp2(e - np.float32(136.0))).astype(np.float32)
    return (rgbe[..., :3].astype(np.float32) * scale[..., None]).astype(np.float32)


def float_to_rgbe(rgb: np.ndarray) -> np.ndarray:
    """[..., 3] float -> [..., 4] uint8 RGBE (shared exponent of the largest channel, mantissas truncated)."""
    rgb = np.asarray(rgb, dtype=np.float64)
    m = rgb.max(axis=-1)
    out = np.zeros(rgb.shape[:-1] + (4,), dtype=np.uint8)
    pos = m > 1e-38
    mant, ex = np.frexp(np.where(pos, m, 1.0))            # m = mant * 2^ex, mant in [0.5, 1)
    scale = np.where(pos, 256.0 / np.exp2(ex.astype(np.float64)), 0.0)
    q = np.clip(np.floor(rgb * scale[..., None]), 0, 255)
    out[..., :3] = q.astype(np.uint8)
    out[..., 3] = np.where(pos, np.clip(ex + 128, 0, 255), 0).astype(np.uint8)
    return out


def _read_line(buf: bytes, pos: int):
    end = buf.find(b"\n", pos)
    if end < 0:
        raise ValueError("truncated Radiance header")
    return buf[pos:end].decode("latin-1"), end + 1


def read(path) -> np.ndarray:
    """Decode a Radiance .hdr file into a (height, width, 3) float32 array, row 0 the first scanline."""
    with open(path, "rb") as f:
        buf = f.read()
    line, pos = _read_line(buf, 0)
    if not line.startswith("#?"):
        raise ValueError("not a Radiance file (no #? magic line)")
    while True:
        line, pos = _read_line(buf, pos)
        if line.strip() == "":
            break
        if line.startswith("FORMAT=") and line.strip() != "FORMAT=32-bit_rle_rgbe":
            raise ValueError(f"unsupported Radiance format {line.strip()!r}")
    line, pos = _read_line(buf, pos)
    parts = line.split()
    if len(parts) != 4 or parts[0] != "-Y" or parts[2] != "+X":
        raise ValueError(f"unsupported Radiance resolution line {line!r} (only -Y H +X W)")
    height, width = int(parts[1]), int(parts[3])
    if height <= 0 or width <= 0:
        raise ValueError("empty Radiance image")
    data = np.frombuffer(buf, dtype=np.uint8, offset=pos)
    rgbe = np.empty((height, width, 4), dtype=np.uint8)
    p = 0
    for y in range(height):
        head = data[p:p + 4]
        if len(head) < 4:
            raise ValueError("truncated Radiance scanline")
        adaptive = (8 <= width <= 0x7FFF and head[0] == 2 and head[1] == 2 and head[2] < 128)
        if adaptive:
            if (int(head[2]) << 8 | int(head[3])) != width:
                raise ValueError("Radiance scanline width mismatch")
            p += 4
            for c in range(4):
                x = 0
                while x < width:
                    if p >= len(data):
                        raise ValueError("truncated Radiance run")
                    n = int(data[p])
                    p += 1
                    if n > 128:
                        n -= 128
                        if x + n > width:
                            raise ValueError("Radiance run overflows its scanline")
                        rgbe[y, x:x + n, c] = data[p]
                        p += 1
                    else:
                        if n == 0 or x + n > width:
                            raise ValueError("bad Radiance literal run")
                        rgbe[y, x:x + n, c] = data[p:p + n]
                        p += n
                    x += n
        else:
            row = data[p:p + 4 * width]
            if len(row) < 4 * width:
                raise ValueError("truncated Radiance scanline")
            rgbe[y] = row.reshape(width, 4)
            p += 4 * width
    return rgbe_to_float(rgbe)


def write(path, pixels: np.ndarray, rle: bool = True) -> None:
    """Encode a (height, width, 3) float array as a Radiance .hdr file."""
    px = np.asarray(pixels, dtype=np.float64)
    height, width, _ = px.shape
    rgbe = float_to_rgbe(px)
    out = bytearray(b"#?RADIANCE\nFORMAT=32-bit_rle_rgbe\n\n" + f"-Y {height} +X {width}\n".encode())
    use_rle = rle and 8 <= width <= 0x7FFF
    for y in range(height):
        if not use_rle:
            out += rgbe[y].tobytes()
            continue
        out += bytes([2, 2, width >> 8, width & 0xFF])
        for c in range(4):
            ch = rgbe[y, :, c]
            x = 0
            while x < width:
                n = 1
                while x + n < width and n < 127 and ch[x + n] == ch[x]:
                    n += 1
                if n >= 3:
                    out += bytes([128 + n, ch[x]])
                    x += n
                    continue
                s = x
                while x < width and x - s < 128 and not (x + 2 < width and ch[x] == ch[x + 1] == ch[x + 2]):
                    x += 1
                out += bytes([x - s]) + ch[s:x].tobytes()
    with open(path, "wb") as f:
        f.write(bytes(out))
