"""Frame output on the host: a rendered frame (packed 8-bit sRGB pixels, red
in the low byte, as pack_sRGB packs them, RayTracer.cu:262-266) written as
an RGBA PNG, row by row with the frame's stride, like the reference's
stbi_write_png of rt::Frame (MainWindow_UI.cpp:165-174)."""
from __future__ import annotations

import struct
import zlib

import numpy as np


def frame_rgba(frame: np.ndarray, width: int, height: int, stride: int | None = None) -> np.ndarray:
    """(height, width, 4) uint8 RGBA of a packed frame whose rows are `stride`
    pixels apart (default: width)."""
    stride = width if stride is None else int(stride)
    if stride < width:
        raise ValueError("stride smaller than the width")
    f = np.ascontiguousarray(frame, np.uint32).reshape(-1)
    if f.size < stride * (height - 1) + width:
        raise ValueError("frame smaller than height rows of the stride")
    rows = np.lib.stride_tricks.as_strided(f, (height, width), (4 * stride, 4)) if height else f[:0].reshape(0, width)
    return np.ascontiguousarray(rows).view(np.uint8).reshape(height, width, 4)


def png_bytes(rgba: np.ndarray) -> bytes:
    """PNG (8-bit RGBA, no interlace, filter 0 per row) of an (h, w, 4) image."""
    rgba = np.ascontiguousarray(rgba, np.uint8)
    h, w, c = rgba.shape
    if c != 4:
        raise ValueError("expected 4 channels")
    raw = b"".join(b"\x00" + rgba[y].tobytes() for y in range(h))

    def chunk(tag: bytes, data: bytes) -> bytes:
        return struct.pack(">I", len(data)) + tag + data + struct.pack(">I", zlib.crc32(tag + data) & 0xFFFFFFFF)

    return (b"\x89PNG\r\n\x1a\n" + chunk(b"IHDR", struct.pack(">IIBBBBB", w, h, 8, 6, 0, 0, 0))
            + chunk(b"IDAT", zlib.compress(raw, 6)) + chunk(b"IEND", b""))


def write_png(path, frame: np.ndarray, width: int, height: int, stride: int | None = None) -> None:
    """Writes a packed frame (see frame_rgba) to `path` as a PNG."""
    with open(path, "wb") as f:
        f.write(png_bytes(frame_rgba(frame, width, height, stride)))
