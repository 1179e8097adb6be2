"""The `.be` model reader (Surface::load_be, Surface.cpp:149-213) on the CPU:
cgs_load_be against the oracle's restatement on the reference's four models
(tests/golden/assets/bezier, copied by tests/golden/make_assets.py), the reference's
error messages, separators, and the indices pinned by the teapot's
rotational symmetry."""
from __future__ import annotations

import hashlib
import json
import os

import numpy as np
import pytest

ASSETS = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "assets")
NAMES = ("teapot", "teacup", "teaspoon", "longwave")


def meta():
    with open(os.path.join(ASSETS, "assets.json")) as f:
        return json.load(f)


def path(name) -> str:
    return os.path.join(ASSETS, "bezier", name)


def read(name) -> bytes:
    with open(path(name), "rb") as f:
        return f.read()


def test_fixtures_intact():
    for rel, m in meta().items():
        with open(os.path.join(ASSETS, rel), "rb") as f:
            data = f.read()
        assert len(data) == m["bytes"] and hashlib.sha256(data).hexdigest() == m["sha256"], rel


@pytest.mark.parametrize("name", NAMES)
def test_load_be_matches_oracle(oracle, name):
    import cgspline_amd as cgs
    v, i = cgs.load_be(path(name))
    ov, oi = oracle.load_be(read(name))
    m = meta()["bezier/" + name]
    assert v.shape == (m["points"], 4) and i.shape == (m["patches"], 16)
    assert np.array_equal(v.view(np.uint32), ov.view(np.uint32))
    assert np.array_equal(i, oi)
    assert np.all(v[:, 3] == 1.0)


def test_teapot_first_points():
    """The Utah teapot's published first points (Newell's data set)."""
    import cgspline_amd as cgs
    v, i = cgs.load_be(read("teapot"))
    assert np.array_equal(v[0, :3], np.float32([1.4, 0.0, 2.4]))
    assert np.array_equal(v[1, :3], np.float32([1.4, -0.784, 2.4]))
    assert np.array_equal(i[0], np.arange(16))


def quadrant_gaps(v, idx):
    """The teapot's rim, body and lid (patch groups 0-2, 5, 6) and bottom
    (group 7) are surfaces of revolution about z in four quadrant patches:
    patch 4g+k+1 is patch 4g+k turned by −90° (bottom: +90°), point for
    point.  Returns the largest coordinate gap per group."""
    P = v[:, :3].astype(np.float64)
    gaps = []
    for g, s in ((0, -1), (1, -1), (2, -1), (5, -1), (6, -1), (7, 1)):
        R = np.array([[0, -s, 0], [s, 0, 0], [0, 0, 1.0]])
        gaps.append(max(np.abs(P[idx[4 * g + k]] @ R.T - P[idx[4 * g + (k + 1) % 4]]).max() for k in range(4)))
    return np.array(gaps)


def test_teapot_quadrants_are_rotations():
    """Pins the indices against the geometry they name: exact 90° symmetry
    of the read teapot; with one-based indices kept (a missing `- 1`) every
    group breaks."""
    import cgspline_amd as cgs
    v, idx = cgs.load_be(read("teapot"))
    assert quadrant_gaps(v, idx).max() == 0.0
    one_based = np.minimum(idx + 1, v.shape[0] - 1)
    assert quadrant_gaps(v, one_based).min() > 0.01


def test_separators_and_spacing():
    """One separator after each number, whitespace before numbers skipped:
    commas, newlines, ", " and blank lines all read the same."""
    import cgspline_amd as cgs
    ids = ",".join(str(k) for k in range(1, 17))
    pts = "\n".join(f"{k}.5,{-k}e-1,{k * 0.25}" for k in range(16))
    a = cgs.load_be(f"1\n{ids}\n16\n{pts}\n".encode())
    b = cgs.load_be(("  1\n\n" + ", ".join(str(k) for k in range(1, 17)) + "\n 16\n" +
                     "\n".join(f" {k}.5, {-k}e-1, {k * 0.25}" for k in range(16))).encode())
    c = cgs.load_be(f"1\n{ids.replace(',', ' ')}\n16\n{pts.replace(',', ' ')}".encode())
    for x in (b, c):
        assert np.array_equal(a[0], x[0]) and np.array_equal(a[1], x[1])
    assert np.array_equal(a[0][3, :3], np.float32([3.5, -0.3, 0.75]))


@pytest.mark.parametrize("text,message", [
    (b"", "Could not read the no. patches"),
    (b"x", "Could not read the no. patches"),
    (b"1\n1,2,3", "Parsing patch failed"),
    (b"1\n" + b",".join(b"1" for _ in range(16)) + b"\n", "Could not read the no. vertices"),
    (b"1\n" + b",".join(b"1" for _ in range(16)) + b"\n2\n1,2,3\n4,5", "Parsing point failed"),
    (b"1\n" + b",".join(b"1" for _ in range(16)) + b"\n1\n1,2,x", "Parsing point failed"),
    (b"1\n" + b",".join(b"0" for _ in range(16)) + b"\n1\n1,2,3", "Invalid/corrupted file. Indices out of range"),
    (b"1\n" + b",".join(b"3" for _ in range(16)) + b"\n1\n1,2,3", "Invalid/corrupted file. Indices out of range"),
    # index == point count: out of range here (the reference lets it through)
    (b"1\n" + b",".join(b"2" for _ in range(16)) + b"\n1\n1,2,3", "Invalid/corrupted file. Indices out of range"),
    # a patch count the text cannot hold fails without allocating it
    (b"4000000000\n1,2", "Parsing patch failed"),
])
def test_errors(oracle, text, message):
    import cgspline_amd as cgs
    with pytest.raises(cgs.CgsError, match=message.replace(".", r"\.")):
        cgs.load_be(text)
    if not text.startswith(b"4000000000"):
        with pytest.raises(oracle.BeError, match=message.replace(".", r"\.")):
            oracle.load_be(text)


def test_teaset_scene_builds(oracle):
    """The teaset workload (Teaset01.dl restated) over the four models: the
    camera sees the floor and all three objects."""
    from cgspline_amd import scenes
    wl = scenes.teaset(ASSETS, width=96, height=72)
    assert wl.scene.npatch == 32 + 26 + 16 + 4
    _, _, obj, _ = oracle.OracleScene(wl.scene).trace(scenes.camera_rays(wl.camera))
    assert set(np.unique(obj)) == {-1, 0, 1, 2, 3}
