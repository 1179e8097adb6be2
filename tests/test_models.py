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


# ---------------------------------------------------------------------------
# `.se` element surfaces (Surface::load_se) and cg::PatchBVH's boxes
# ---------------------------------------------------------------------------

SURFACES = {
    # name: (nodes, elements, {(type, node count, boundary/matrix): count})
    "quad16": (25, 16, {(1, 9, True): 4, (1, 12, True): 8, (1, 16, False): 4}),
    "tube": (600, 600, {(1, 12, True): 240, (1, 16, False): 280, (1, 8, True): 80}),
    "ball0": (8, 6, {(2, 8, True): 6}),
    "dice16": (98, 96, {(1, 16, False): 72, (2, 14, True): 24}),
    "cube16": (98, 96, {(1, 9, True): 24, (1, 12, True): 48, (1, 16, False): 24}),
}


def se_text(name) -> bytes:
    with open(os.path.join(ASSETS, "surfaces", name + ".se"), "rb") as f:
        return f.read()


@pytest.mark.parametrize("name", sorted(SURFACES))
def test_load_se_matches_oracle(oracle, name):
    """The library's reader = the oracle's token-by-token restatement: points,
    grouped indices, transposed matrices, patch order; and the element boxes
    bit for bit.  Counts per element kind as the files hold them."""
    import cgspline_amd as cgs
    t = cgs.load_se(se_text(name))
    pts, idx, mat, refs = oracle.load_se(se_text(name))
    assert np.array_equal(t.points.view(np.uint32), pts.view(np.uint32))
    assert np.array_equal(t.indices, idx) and np.array_equal(t.patches, refs)
    assert np.array_equal(t.matrices.view(np.uint32), mat.view(np.uint32))
    nodes, elements, kinds = SURFACES[name]
    assert t.points.shape[0] == nodes and t.patches.shape[0] == elements
    got = {}
    for off, moff, size, typ in t.patches:
        got[(int(typ), int(size), bool(moff >= 0))] = got.get((int(typ), int(size), bool(moff >= 0)), 0) + 1
    assert got == kinds
    # patches of one group are contiguous, groups in order of first appearance
    keys = [(int(r[3]), int(r[2]), bool(r[1] >= 0)) for r in t.patches]
    assert sum(1 for a, b in zip(keys, keys[1:]) if a != b) == len(kinds) - 1
    e = t.elements()
    assert np.array_equal(e.view(np.uint32), oracle.se_elements(pts, idx, mat, refs).view(np.uint32))


@pytest.mark.parametrize("name", sorted(SURFACES))
def test_se_basis_rows_are_affine(name):
    """Every basis-matrix row (one output control point) sums to 1 over its
    nodes, as a change of polynomial basis must; a matrix read without the
    reference's transposition fails this."""
    import cgspline_amd as cgs
    t = cgs.load_se(se_text(name))
    for off, moff, size, typ in t.patches:
        if moff < 0:
            continue
        n = 20 if typ == 2 else 16
        M = t.matrices[moff:moff + n * size].reshape(size, n).T.astype(np.float64)   # rows × nodes
        assert np.allclose(M.sum(axis=1), 1.0, atol=1e-5), (name, off)
        assert not np.allclose(M.reshape(size, n).sum(axis=1), 1.0, atol=1e-5) or size == n


def test_quad_boxes_lie_in_its_plane():
    """quad16.se is the flat quad y = 0: every element box has y = 0, and the
    boxes together span the quad x in [-0.5, 0.5], z in [-1, 1]."""
    import cgspline_amd as cgs
    e = cgs.load_se(se_text("quad16")).elements()
    assert np.abs(e[:, [1, 4]]).max() < 1e-6
    assert e[:, 0].min() <= -0.5 + 1e-6 and e[:, 3].max() >= 0.5 - 1e-6
    assert e[:, 2].min() <= -1.0 + 1e-6 and e[:, 5].max() >= 1.0 - 1e-6


def test_se_leaf_reads_and_overruns():
    """The leaf test's 16 indices run on into the next patch's for a patch of
    fewer nodes; the last patch of ball0 would read past the table (undefined
    in the reference): it gets 16 zeros and the empty box."""
    import cgspline_amd as cgs
    t = cgs.load_se(se_text("ball0"))
    p = t.patch_indices()
    assert np.array_equal(p[0], t.indices[0:16])           # 8 nodes of its own, 8 of the next
    assert np.all(p[-1] == 0)
    e = t.elements()
    assert e[-1, 0] > e[-1, 3] and np.all(e[:-1, 0] <= e[:-1, 3])


SE_ERRORS = [
    (b"", "could not read no. points / nodes"),
    (b"2\n0 0 0 0 1\n3 0 0 0 1\n", "unexpected point id"),
    (b"1\n0 0 0 0 1\n1\n0 1 1 2 1\n0\n1\n", "degree value unsupported"),
    (b"1\n0 0 0 0 1\n1\n0 3 1 3 1\n0\n1\n", "unknown patch type"),
    (b"1\n0 0 0 0 1\n1\n0 1 1 3 0\n5\n1\n", "Invalid/corrupted file. Indices out of range"),
    (b"1\n0 0 0 0 1\n1\n0 1 1 3 1\n0\n1\n0.5\n", "read failed during element data"),
    (b"# a comment\n1 # nodes\n0 0 0 0 1\n0\n", None),
]


@pytest.mark.parametrize("text, message", SE_ERRORS)
def test_se_errors(oracle, text, message):
    import cgspline_amd as cgs
    if message is None:
        t = cgs.load_se(text)
        assert t.points.shape == (1, 4) and t.patches.shape[0] == 0
        return
    with pytest.raises(cgs.CgsError, match=message.replace(".", r"\.")):
        cgs.load_se(text)
    with pytest.raises(oracle.SeError, match=message.replace(".", r"\.")):
        oracle.load_se(text)


def _se_rays(v, n, seed):
    r = np.random.default_rng(seed)
    xyz = v[:, :3].astype(np.float64)
    c = xyz.mean(axis=0)
    rad = float(np.linalg.norm(xyz - c, axis=1).max())
    d = r.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    o = c + 2.5 * rad * d
    target = c + r.uniform(-0.6, 0.6, (n, 3)) * rad
    rays = np.zeros((n, 8), np.float32)
    rays[:, 0:3], rays[:, 3], rays[:, 4:7] = o, np.inf, target - o
    return rays


@pytest.mark.parametrize("name", ["tube", "ball0", "dice16"])
def test_se_pick_equals_clipping_the_entered_boxes(oracle, name):
    """The oracle's PatchBVH pick over a `.se` table against brute force: every
    ray clipped against every patch whose element box it enters (slab test in
    float64 on the normalized direction).  Boxes the ray only grazes, within
    1e-4 of the box size, are left out of the comparison both ways; boxes
    flat in one axis are never entered.  An any pick is set exactly where one
    of those clips hits; a closest pick is one of them and no box entered
    before it holds a nearer hit."""
    import cgspline_amd as cgs
    from cgspline_amd import scenes
    t = cgs.load_se(se_text(name))
    el, pi = t.elements(), t.patch_indices()
    rays = _se_rays(t.points, 160, 91)
    sc = scenes.assemble([(t.points, pi)], [(0, scenes.IDENTITY16, scenes.IDENTITY16,
                                             np.full(8, 0.5, np.float32))], np.zeros((0, 12), np.float32))
    osc = oracle.OracleScene(sc, surface0_elements=el)
    tuv, patch = osc.pick(rays)
    occ = osc.pick(rays, any_hit=True)
    # slab test per (ray, box), float64
    d = rays[:, 4:7].astype(np.float64)
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    o = rays[:, 0:3].astype(np.float64)
    lo, hi = el[:, 0:3].astype(np.float64), el[:, 3:6].astype(np.float64)
    with np.errstate(divide="ignore", invalid="ignore"):
        t1 = (lo[None] - o[:, None]) / d[:, None]
        t2 = (hi[None] - o[:, None]) / d[:, None]
    tmin = np.nanmax(np.minimum(t1, t2), axis=2)
    tmax = np.nanmin(np.maximum(t1, t2), axis=2)
    size = np.maximum(np.linalg.norm(hi - lo, axis=1), 1e-6)[None]
    entered = (tmax - tmin > 1e-4 * size) & (tmax > 1e-4 * size)
    grazed = ~entered & (tmin - tmax <= 1e-4 * size) & (tmax > -1e-4 * size)
    # a box flat in one axis (the tube's end caps) gives tmin = tmax exactly
    # wherever the ray crosses it: the strict slab test never enters it
    grazed &= ~np.any(lo == hi, axis=1)[None]
    P = t.points[pi.reshape(-1)].reshape(-1, 16, 4)
    m = P.shape[0]
    res, _ = oracle.clip_batch(np.repeat(rays, m, axis=0), np.tile(P, (rays.shape[0], 1, 1)))
    res = res.reshape(rays.shape[0], m, 4)
    found = res[:, :, 3] == 1
    tt = np.where(found & entered, res[:, :, 0], np.inf)
    clean = ~(found & grazed).any(axis=1)
    assert clean.sum() >= 0.9 * rays.shape[0]
    hit = patch != 0xFFFFFFFF
    # any: every entered leaf is clipped from distance infinity
    assert np.array_equal(occ[clean] != 0, np.isfinite(tt[clean]).any(axis=1))
    assert np.array_equal(hit[clean], np.isfinite(tt[clean]).any(axis=1))
    # closest: the hit is a clip of an entered box; the clipped nodes need not
    # lie in the box, so a box entered beyond the best distance so far is
    # skipped even where its patch would hit nearer.  Every box entered before
    # the final hit was visited: none of them holds a nearer hit.
    margin = 1e-4 * size[0]
    for k in np.flatnonzero(hit & clean):
        j = int(patch[k])
        assert entered[k, j] and tuv[k, 0] == res[k, j, 0], k
        assert tuv[k, 1] == res[k, j, 1] and tuv[k, 2] == res[k, j, 2], k
        before = entered[k] & (tmin[k] < tuv[k, 0] - margin)
        assert np.all(tt[k, before] >= tuv[k, 0]), k
    assert hit[clean].sum() >= 10
