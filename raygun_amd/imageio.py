"""Image writers for frames read back from the ray-tracing core.

The reference presents its result through the swapchain (the blit of
raygun/render/render_system.cpp:102-147) and never writes files; these
writers are the other end of that readback for callers without a window:

- ``write_png``  8-bit RGBA or RGB (zlib-compressed, filter type 0)
- ``write_ppm``  binary PPM (P6), 8-bit RGB
- ``write_pfm``  little-endian PFM (Pf / PF), float32, rows bottom to top as
  the format defines
- ``read_pfm``   the inverse of ``write_pfm``

Inputs are numpy arrays (height, width, channels); 8-bit images come from
``Context.read_image_rgba8`` and float images from ``half_to_float`` of an
rgba16f image.
"""
import struct
import zlib

import numpy as np


def _chunk(kind, data):
    body = kind + data
    return struct.pack(">I", len(data)) + body + struct.pack(">I", zlib.crc32(body) & 0xFFFFFFFF)


def write_png(path, img):
    """img: (H, W, 4) or (H, W, 3) uint8."""
    a = np.ascontiguousarray(img, np.uint8)
    if a.ndim != 3 or a.shape[2] not in (3, 4):
        raise ValueError("write_png: expected (H, W, 3|4) uint8")
    h, w, ch = a.shape
    color_type = 6 if ch == 4 else 2
    raw = np.zeros((h, 1 + w * ch), np.uint8)  # filter byte 0 per row
    raw[:, 1:] = a.reshape(h, w * ch)
    data = b"\x89PNG\r\n\x1a\n"
    data += _chunk(b"IHDR", struct.pack(">IIBBBBB", w, h, 8, color_type, 0, 0, 0))
    data += _chunk(b"IDAT", zlib.compress(raw.tobytes(), 6))
    data += _chunk(b"IEND", b"")
    with open(path, "wb") as f:
        f.write(data)


def write_ppm(path, img):
    """img: (H, W, 3|4) uint8; alpha is dropped."""
    a = np.ascontiguousarray(np.asarray(img, np.uint8)[:, :, :3])
    h, w, _ = a.shape
    with open(path, "wb") as f:
        f.write(b"P6\n%d %d\n255\n" % (w, h))
        f.write(a.tobytes())


def write_pfm(path, img):
    """img: (H, W) grey or (H, W, 3|4) float (alpha dropped)."""
    a = np.asarray(img, np.float32)
    if a.ndim == 3:
        a = a[:, :, :3]
    h, w = a.shape[:2]
    kind = b"PF" if a.ndim == 3 else b"Pf"
    with open(path, "wb") as f:
        f.write(b"%s\n%d %d\n-1.0\n" % (kind, w, h))  # negative scale: little-endian
        f.write(np.ascontiguousarray(a[::-1]).astype("<f4").tobytes())


def read_pfm(path):
    with open(path, "rb") as f:
        kind = f.readline().strip()
        w, h = (int(x) for x in f.readline().split())
        scale = float(f.readline())
        data = np.frombuffer(f.read(), "<f4" if scale < 0 else ">f4")
    shape = (h, w, 3) if kind == b"PF" else (h, w)
    return data.reshape(shape)[::-1].astype(np.float32)
