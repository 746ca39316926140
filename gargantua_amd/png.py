"""PNG output for rendered images (RgbImage::save, crates/gr-renderer/src/main.rs:137).

The reference saves its 8-bit sRGB RgbImage through the `image` crate's
PNG encoder.  The bytes of a PNG depend on the encoder's filter and
deflate choices, so the contract is the decoded pixels: `write` stores an
(height, width, 3) uint8 image as colour type 2, bit depth 8,
non-interlaced, scanline filter 0; `read` decodes that form and the other
four scanline filters of the PNG specification (ISO/IEC 15948 sec. 9).
"""
from __future__ import annotations

import struct
import zlib

import numpy as np

SIGNATURE = b"\x89PNG\r\n\x1a\n"


def _chunk(kind: bytes, data: bytes) -> bytes:
    return struct.pack(">I", len(data)) + kind + data + struct.pack(">I", zlib.crc32(kind + data) & 0xFFFFFFFF)


def encode(rgb: np.ndarray, level: int = 6) -> bytes:
    img = np.ascontiguousarray(rgb, dtype=np.uint8)
    if img.ndim != 3 or img.shape[2] != 3 or img.shape[0] == 0 or img.shape[1] == 0:
        raise ValueError(f"PNG needs an (height, width, 3) uint8 image, got {img.shape}")
    h, w, _ = img.shape
    raw = np.zeros((h, 1 + 3 * w), dtype=np.uint8)           # filter byte 0 per scanline
    raw[:, 1:] = img.reshape(h, 3 * w)
    ihdr = struct.pack(">IIBBBBB", w, h, 8, 2, 0, 0, 0)
    return SIGNATURE + _chunk(b"IHDR", ihdr) + _chunk(b"IDAT", zlib.compress(raw.tobytes(), level)) + \
        _chunk(b"IEND", b"")


def write(path, rgb: np.ndarray) -> None:
    """Save an (height, width, 3) uint8 image as PNG."""
    with open(path, "wb") as f:
        f.write(encode(rgb))


def _paeth(a: int, b: int, c: int) -> int:
    p = a + b - c
    pa, pb, pc = abs(p - a), abs(p - b), abs(p - c)
    if pa <= pb and pa <= pc:
        return a
    return b if pb <= pc else c


def decode(buf: bytes) -> np.ndarray:
    if buf[:8] != SIGNATURE:
        raise ValueError("not a PNG file")
    pos, idat, ihdr = 8, b"", None
    while pos < len(buf):
        n, kind = struct.unpack(">I4s", buf[pos:pos + 8])
        data = buf[pos + 8:pos + 8 + n]
        (crc,) = struct.unpack(">I", buf[pos + 8 + n:pos + 12 + n])
        if zlib.crc32(kind + data) & 0xFFFFFFFF != crc:
            raise ValueError(f"PNG chunk {kind!r} fails its CRC")
        if kind == b"IHDR":
            ihdr = struct.unpack(">IIBBBBB", data)
        elif kind == b"IDAT":
            idat += data
        elif kind == b"IEND":
            break
        pos += 12 + n
    if ihdr is None:
        raise ValueError("PNG without IHDR")
    w, h, depth, ctype, _, _, interlace = ihdr
    if depth != 8 or ctype != 2 or interlace != 0:
        raise ValueError("only 8-bit RGB non-interlaced PNG is read")
    raw = np.frombuffer(zlib.decompress(idat), dtype=np.uint8).reshape(h, 1 + 3 * w)
    out = np.zeros((h, 3 * w), dtype=np.uint8)
    bpp = 3
    for y in range(h):
        f, line = raw[y, 0], raw[y, 1:].astype(np.int32)
        prev = out[y - 1].astype(np.int32) if y else np.zeros(3 * w, np.int32)
        if f == 0:
            cur = line
        elif f == 2:
            cur = (line + prev) & 0xFF
        elif f in (1, 3, 4):
            cur = np.zeros(3 * w, np.int32)
            for x in range(3 * w):
                a = cur[x - bpp] if x >= bpp else 0
                c = prev[x - bpp] if x >= bpp else 0
                pred = a if f == 1 else ((a + prev[x]) >> 1 if f == 3 else _paeth(a, prev[x], c))
                cur[x] = (line[x] + pred) & 0xFF
        else:
            raise ValueError(f"bad PNG filter type {f}")
        out[y] = cur.astype(np.uint8)
    return out.reshape(h, w, 3)


def read(path) -> np.ndarray:
    with open(path, "rb") as f:
        return decode(f.read())
