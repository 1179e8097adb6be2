"""Frame output (cgspline_amd.image): packed pixels to RGBA PNG with a row
stride, decoded again with Pillow, on random frames and on a frame rendered by
the oracle."""
from __future__ import annotations

import io

import numpy as np
import pytest

PIL_Image = pytest.importorskip("PIL.Image")


def decode(data: bytes) -> np.ndarray:
    im = PIL_Image.open(io.BytesIO(data))
    assert im.mode == "RGBA"
    return np.asarray(im)


def test_png_round_trip_with_stride(tmp_path):
    from cgspline_amd import image
    r = np.random.default_rng(3)
    w, h, stride = 37, 11, 48
    frame = r.integers(0, 2 ** 32, size=stride * h, dtype=np.uint64).astype(np.uint32)
    p = tmp_path / "f.png"
    image.write_png(p, frame, w, h, stride)
    got = decode(p.read_bytes())
    rows = frame.reshape(h, stride)[:, :w]
    assert got.shape == (h, w, 4)
    # red is the low byte of a packed pixel, alpha the high byte
    assert np.array_equal(got[..., 0], rows & 0xFF) and np.array_equal(got[..., 3], rows >> 24)
    assert np.array_equal(np.ascontiguousarray(got).view(np.uint32).reshape(h, w), rows)
    with pytest.raises(ValueError):
        image.frame_rgba(frame, w, h, w - 1)
    with pytest.raises(ValueError):
        image.frame_rgba(frame[:stride * (h - 1)], w, h, stride)


def test_png_of_rendered_frame(oracle, c1):
    """The oracle's frame of the height-field workload written and decoded."""
    from cgspline_amd import image
    wl = c1
    _, frame, _, _ = oracle.OracleScene(wl.scene).render(wl.camera, rows=(0, 32))
    w, h = wl.camera.width, 32
    got = decode(image.png_bytes(image.frame_rgba(frame.reshape(-1), w, h)))
    assert np.array_equal(np.ascontiguousarray(got).view(np.uint32).reshape(h, w), frame[:h])
    assert np.all(got[..., 3] == 255)
