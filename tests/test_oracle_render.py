"""Pins of the oracle's renderer and scene by closed forms and invariants
(CPU only): the camera (RayTracer.cu:59-71,141-163), the shading sum with a
shadow ray and one reflection bounce (RayTracer.cu:398-544, PBR.h, PBR.cu)
and the per-object rescaling of t by |M d| (RayTracer.cu:296-350).  Each pin
also runs against an oracle copy with a planted mistake (tests/planted.py),
which must fail it."""
from __future__ import annotations

import math

import numpy as np
import pytest

import planted
from conftest import rng


def make_cam(flip=False, w=101, h=61, fovy_deg=40.0):
    from cgspline_amd import scenes
    cam = scenes.make_camera((1.0, -2.0, 3.0), (0.2, 0.5, -0.4), (0.0, 0.0, 1.0), fovy_deg, w, h)
    cam.flip_y = flip
    return cam


def angle(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return math.atan2(np.linalg.norm(np.cross(a, b)), a @ b)


def check_camera(oracle):
    """The centre pixel's ray is the view axis; rays through the frame's top
    and bottom edges make fovy/2 with it, through the left edge
    atan(tan(fovy/2) w/h) (square pixels); flip_y mirrors the rows."""
    for flip in (False, True):
        cam = make_cam(flip)
        w, h = cam.width, cam.height
        M = cam.to_world.reshape(4, 4).astype(np.float64)     # rows are columns
        right, up, back = M[0, :3], M[1, :3], M[2, :3]
        axis = -back
        cx, cy = (w - 1) / 2, (h - 1) / 2
        xy = np.array([[cx, cy], [cx, -0.5], [cx, h - 0.5], [-0.5, cy], [w - 0.5, cy]], np.float32)
        d = oracle.OracleScene.pixel_rays(cam, xy).astype(np.float64)
        assert np.abs(np.linalg.norm(d, axis=1) - 1).max() < 1e-6
        assert angle(d[0], axis) < 2e-6, "centre ray off the view axis"
        half = cam.fovy / 2
        for k in (1, 2):
            assert abs(angle(d[k], axis) - half) < 2e-6, f"edge ray {k} at {angle(d[k], axis)} rad, not fovy/2"
        sign = -1 if flip else 1
        assert sign * (d[1] @ up) > 0 and sign * (d[2] @ up) < 0, "top edge not above the axis"
        hx = math.atan(math.tan(half) * w / h)
        for k in (3, 4):
            assert abs(angle(d[k], axis) - hx) < 2e-6, "left/right edge ray not at atan(tan(fovy/2) w/h)"
        assert d[3] @ right < 0 < d[4] @ right


def test_camera_closed_form(oracle):
    check_camera(oracle)


def test_python_camera_rays_equal_the_oracle(oracle):
    """scenes.camera_rays (the rays traced on the GPU) = the oracle's camera
    bit for bit at every pixel centre."""
    from cgspline_amd import scenes
    # 60 degrees: the C library's tanf of pi/6 in float32 is not the
    # correctly rounded tangent, and the camera must use the former
    for flip, fovy in ((False, 40.0), (True, 40.0), (False, 60.0)):
        cam = make_cam(flip, w=64, h=48, fovy_deg=fovy)
        rays = scenes.camera_rays(cam)
        jj, ii = np.meshgrid(np.arange(cam.height), np.arange(cam.width), indexing="ij")
        xy = np.stack([ii.ravel(), jj.ravel()], 1).astype(np.float32)
        d = oracle.OracleScene.pixel_rays(cam, xy)
        assert np.array_equal(rays[:, 4:7].view(np.uint32), d.view(np.uint32))
        assert np.array_equal(rays[:, 0:3], np.broadcast_to(cam.position, (d.shape[0], 3)))


# ---------------------------------------------------------------------------
# shading composition: one flat square, one point light, one bounce
# ---------------------------------------------------------------------------

MAT = np.array([0.6, 0.4, 0.2, 0.04, 0.5, 0.9, 0.3, 0.25], np.float32)   # diffuse, specular, metalness, roughness
LIGHT_POS = np.array([0.3, -0.2, 2.0])
LIGHT_COL = np.array([2.0, 3.0, 4.0])
BG = (0.2, 0.3, 0.4)
AMB = (0.05, 0.1, 0.15)


def flat_square_scene(l2w=None, w2l=None):
    """A bilinear square [-1,1]^2 x {0} as one bicubic patch (control points
    on the i/3 grid: an affine parameterisation) under the given object
    transform."""
    from cgspline_amd import scenes
    v = np.zeros((16, 4), np.float32)
    for j in range(4):
        for i in range(4):
            v[4 * j + i] = (-1 + 2 * i / 3, -1 + 2 * j / 3, 0.0, 1.0)
    idx = np.arange(16, dtype=np.uint32)[None]
    light = scenes.point_light(LIGHT_POS, LIGHT_COL)
    l2w = scenes.IDENTITY16 if l2w is None else l2w
    w2l = scenes.IDENTITY16 if w2l is None else w2l
    return scenes.assemble([(v, idx)], [(0, l2w, w2l, MAT)], [light], name="flat", background=BG, ambient=AMB)


def expected_colour(o, d, depth):
    """FP64 closed form of the renderer's colour for a ray from o along d
    against the square z = 0, |x|, |y| <= 1 (RayTracer.cu:398-544)."""
    o = np.asarray(o, np.float64)
    d = np.asarray(d, np.float64)
    d = d / np.linalg.norm(d)
    bg, amb = np.array(BG), np.array(AMB)
    if d[2] == 0:
        return bg
    t = -o[2] / d[2]
    P = o + t * d
    if t <= 0 or abs(P[0]) > 1 or abs(P[1]) > 1:
        return bg
    diffuse, f0 = MAT[0:3].astype(np.float64), MAT[3:6].astype(np.float64)
    metal, r = float(MAT[6]), float(MAT[7])
    dif = (1 - metal) * diffuse / math.pi
    N = np.array([0.0, 0.0, 1.0])
    V = -d
    dotNV = N @ V
    if dotNV < 0:
        N, dotNV = -N, -dotNV
    col = amb * dif
    Lv = LIGHT_POS - P
    dist = np.linalg.norm(Lv)
    L = Lv / dist
    dotNL = N @ L
    if dotNL >= 1e-14:
        H = (L + V) / np.linalg.norm(L + V)
        F = lambda x: f0 + (1 - f0) * (1 - x) ** 5
        k = (r + 1) ** 2 / 8
        G1 = lambda x: x / (x * (1 - k) + k)
        a2 = (max(r, 1e-3) ** 2) ** 2
        D = a2 / (math.pi * ((N @ H) ** 2 * (a2 - 1) + 1) ** 2)
        spec = F(H @ L) * G1(dotNL) * G1(dotNV) * D / (4 * dotNL * dotNV)
        col = col + (LIGHT_COL / dist ** 2) * dotNL * (dif + metal * spec)
    out = math.pi * col
    if depth > 0:
        k = (r + 1) ** 2 / 8
        g = (dotNV / (dotNV * (1 - k) + k)) ** 2
        Fv = f0 + (1 - f0) * (1 - dotNV) ** 5
        a = (1 - r) ** 2 * dotNV * (dif + metal * g * Fv)
        if a.max() > 2.0 ** -8:
            out = out + bg * a      # the mirror ray leaves the plane upward: background
    return out


def check_shading(oracle, depth=1):
    """Every pixel of a view of the square equals the FP64 closed form: ambient
    and Lambert + GGX under one point light with its 1/d^2 falloff, times pi,
    plus the background seen along the mirror bounce, attenuated by
    (N.V)(diffuse + metalness F G)(1 - roughness)^2."""
    from cgspline_amd import scenes
    sc = flat_square_scene()
    cam = scenes.make_camera((0.4, -1.8, 2.2), (0.0, 0.1, 0.0), (0.0, 0.0, 1.0), 50.0, 48, 40)
    color, _, _, counts = oracle.OracleScene(sc).render(cam, eps=1e-4, depth=depth)
    rays = scenes.camera_rays(cam)
    want = np.array([expected_colour(r[0:3], r[4:7], depth) for r in rays]).reshape(color.shape)
    o, d = rays[:, 0:3].astype(np.float64), rays[:, 4:7].astype(np.float64)
    t = -o[:, 2] / d[:, 2]
    P = o + t[:, None] * d
    hits = (t > 0) & (np.abs(P[:, 0]) < 1) & (np.abs(P[:, 1]) < 1)
    assert 0.2 < hits.mean() < 0.95, "the view should show the square and the background"
    # FP32 against FP64: the GGX term near its peak, (N.H)^2 (a^2 - 1) + 1
    # with a^2 = r^4 = 0.004, amplifies the rounding of N.H about 50-fold
    rel = (np.abs(color - want) / np.maximum(np.abs(want), 1e-3)).max()
    assert rel < 2e-4, f"colour off the closed form by {rel:.2e} relative"
    assert counts["shadow"] > 0 and (counts["secondary"] > 0) == (depth > 0)


def test_shading_closed_form(oracle):
    check_shading(oracle)


def test_shading_closed_form_without_bounce(oracle):
    check_shading(oracle, depth=0)


def test_shaded_rays_closed_form_and_render(oracle):
    """shadeRay from arbitrary rays (the shaded queries' oracle): rays from
    random points above and below the square in random directions, some with
    a range that ends before the plane (background), equal the FP64 closed
    form; from the render's own camera rays it equals the render bit for bit."""
    from cgspline_amd import scenes
    sc = flat_square_scene()
    osc = oracle.OracleScene(sc)
    r = rng(91)
    n = 4000
    o = np.column_stack([r.uniform(-1.5, 1.5, n), r.uniform(-1.5, 1.5, n), r.choice([-1, 1], n) * r.uniform(0.3, 2, n)])
    tgt = np.column_stack([r.uniform(-1.2, 1.2, n), r.uniform(-1.2, 1.2, n), np.zeros(n)])
    d = tgt - o
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    rays = np.zeros((n, 8), np.float32)
    rays[:, 0:3], rays[:, 3], rays[:, 4:7] = o, np.inf, d
    t = -rays[:, 2].astype(np.float64) / rays[:, 6]
    short = np.arange(n) % 5 == 0
    rays[short, 3] = (0.5 * t[short]).astype(np.float32)
    color, _, hits, patch, obj, counts = osc.shade(rays, eps=1e-4, depth=1)
    for k in range(n):
        want = np.array(BG) if short[k] else expected_colour(rays[k, 0:3], rays[k, 4:7], 1)
        rel = np.abs(color[k] - want) / np.maximum(np.abs(want), 1e-3)
        assert rel.max() < 2e-4, (k, color[k], want)
    assert np.all(obj[short] == -1) and np.all(patch[short] == 0xFFFFFFFF)
    assert np.all(hits[short, 0] == rays[short, 3])
    hit = obj == 0
    assert hit.sum() > n // 2
    P = rays[hit, 0:3].astype(np.float64) + hits[hit, 0:1] * rays[hit, 4:7]
    assert np.abs(P[:, 2]).max() < 1e-5
    assert np.abs(hits[hit, 1] - (P[:, 0] + 1) / 2).max() < 1e-4 and np.abs(hits[hit, 2] - (P[:, 1] + 1) / 2).max() < 1e-4
    assert counts["primary"] == n and counts["shadow"] > 0
    # the render's camera rays, shaded one by one, give the render's colours
    cam = scenes.make_camera((0.4, -1.8, 2.2), (0.0, 0.1, 0.0), (0.0, 0.0, 1.0), 60.0, 40, 30)
    rcolor, rframe, _, rcounts = osc.render(cam, eps=1e-4, depth=1)
    crays = scenes.camera_rays(cam)
    scolor, spx, _, _, _, scounts = osc.shade(crays, eps=1e-4, depth=1)
    assert np.array_equal(scolor.view(np.uint32), rcolor.reshape(-1, 3).view(np.uint32))
    assert np.array_equal(spx, rframe.reshape(-1)) and scounts == rcounts


# ---------------------------------------------------------------------------
# two-level scenes: t rescaled by |M d|
# ---------------------------------------------------------------------------

def check_t_rescale(oracle, seed=50):
    """The square under a rotation, a non-uniform scale and a translation:
    world t equals the exact ray-plane distance of the transformed square,
    and (u,v) its inverse parameterisation."""
    from cgspline_amd import scenes
    r = rng(seed)
    R = scenes.random_rotation(r)
    S = np.array([2.0, 0.5, 1.5])
    T = np.array([0.5, -1.0, 2.0])
    m = np.eye(4)
    m[:3, :3] = R * S
    m[:3, 3] = T
    l2w = m.T.reshape(16).astype(np.float32)
    w2l = np.linalg.inv(m).T.reshape(16).astype(np.float32)
    sc = flat_square_scene(l2w, w2l)
    n = 400
    uv = r.uniform(0.05, 0.95, (n, 2))
    local = np.stack([-1 + 2 * uv[:, 0], -1 + 2 * uv[:, 1], np.zeros(n), np.ones(n)], 1)
    X = (m @ local.T).T[:, :3]
    Nw = np.linalg.inv(m[:3, :3]).T @ np.array([0.0, 0.0, 1.0])
    Nw /= np.linalg.norm(Nw)
    o = X + Nw * r.uniform(1.0, 4.0, (n, 1)) * np.sign(r.normal(size=(n, 1))) + r.normal(0, 0.5, (n, 3))
    rays = np.zeros((n, 8), np.float32)
    # unit world directions, as the renderer's rays are: t is then a distance
    rays[:, 0:3], rays[:, 3], rays[:, 4:7] = o, np.inf, (X - o) / np.linalg.norm(X - o, axis=1, keepdims=True)
    hits, patch, obj, _ = oracle.OracleScene(sc).trace(rays)
    assert (obj == 0).mean() > 0.97
    ok = obj == 0
    o64 = rays[:, 0:3].astype(np.float64)
    d64 = rays[:, 4:7].astype(np.float64)
    d64 /= np.linalg.norm(d64, axis=1, keepdims=True)
    t = ((T - o64) @ Nw) / (d64 @ Nw)
    rel = np.abs(hits[ok, 0] - t[ok]) / t[ok]
    assert rel.max() < 1e-4 and np.median(rel) < 1e-5, f"world t off by {rel.max():.2e} relative"
    P = o64 + t[:, None] * d64
    lp = (np.linalg.inv(m) @ np.c_[P, np.ones(n)].T).T
    uv_true = np.stack([(lp[:, 0] + 1) / 2, (lp[:, 1] + 1) / 2], 1)
    assert np.abs(hits[ok, 1:3] - uv_true[ok]).max() < 2e-3


def test_t_rescaled_by_object_transform(oracle):
    check_t_rescale(oracle)


def test_pretransformed_surface_gives_the_same_hits(oracle):
    """A curved surface under a TRS object transform and the same surface
    with its control points transformed to world space (identity object):
    the same hits, t within the clipper's tolerance."""
    from cgspline_amd import scenes
    wl = scenes.heightfield_c1(width=48, height=48)
    base = wl.scene
    r = rng(51)
    R = scenes.random_rotation(r)
    m = np.eye(4)
    m[:3, :3] = R * np.array([1.7, 0.6, 1.2])
    m[:3, 3] = (0.3, 0.2, -0.5)
    l2w = m.T.reshape(16).astype(np.float32)
    w2l = np.linalg.inv(m).T.reshape(16).astype(np.float32)
    a = scenes.assemble([(base.vertices, base.indices)], [(0, l2w, w2l, base.obj_material[0])], base.lights)
    vw = base.vertices.astype(np.float64) @ m.T
    b = scenes.assemble([(vw.astype(np.float32), base.indices)],
                        [(0, scenes.IDENTITY16, scenes.IDENTITY16, base.obj_material[0])], base.lights)
    centre = (m @ np.array([0.5, 0.5, 0.0, 1.0]))[:3]
    cam = scenes.make_camera(centre + R @ np.array([0.2, -1.5, 1.4]), centre, R @ np.array([0, 0, 1.0]), 45.0, 48, 48)
    rays = scenes.camera_rays(cam)
    ha, pa, oa, _ = oracle.OracleScene(a).trace(rays)
    hb, pb, ob, _ = oracle.OracleScene(b).trace(rays)
    both = (oa == 0) & (ob == 0)
    assert both.sum() > 200 and ((oa == 0) != (ob == 0)).mean() < 0.01
    rel = np.abs(ha[both, 0] - hb[both, 0]) / hb[both, 0]
    assert np.median(rel) < 1e-6 and rel.max() < 1e-3
    assert (pa[both] == pb[both]).mean() > 0.99


@pytest.mark.parametrize("plant, func", [(3, "check_camera"), (4, "check_camera"), (5, "check_shading"),
                                         (6, "check_shading"), (7, "check_t_rescale")])
def test_planted_render_mistakes_fail_the_pins(plant, func):
    """tan(fovy) for tan(fovy/2) (3) and half-pixel centres (4) fail the
    camera pin; a dropped pi (5) and a reflection without (1 - r)^2 (6) fail
    the shading closed form; an unrescaled local t (7) fails the plane
    distance under a scaled transform."""
    planted.check_fails(plant, "test_oracle_render", func)
