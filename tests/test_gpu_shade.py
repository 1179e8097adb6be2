"""Shaded ray queries (cgs_shade, cgs_server_submit_shaded): each ray is
traced and shaded as the render shades a camera ray (RayTracer.cu:194-294,
385-544).  Camera rays give the render's colours bit for bit; other rays
(from inside the scene, with finite ranges, towards the lights) give the
oracle's shadeRay bit for bit; served shaded requests equal cgs_shade alone,
mixed with closest-hit and occlusion requests in the same launches."""
from __future__ import annotations

import os

import numpy as np
import pytest

from conftest import rng, say
from test_gpu_parity import assert_bits_equal

pytestmark = pytest.mark.gpu

ASSETS = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "assets")


@pytest.fixture(scope="module")
def teaset():
    from cgspline_amd import scenes
    return scenes.teaset(ASSETS, width=160, height=120, depth=3)


def world_bounds(sc):
    """The box of every object's world-space control points."""
    lo, hi = np.full(3, np.inf), np.full(3, -np.inf)
    for k in range(sc.nobj):
        s = int(sc.obj_surface[k])
        v = sc.vertices[int(sc.surf_vbase[s]):int(sc.surf_vbase[s + 1])].astype(np.float64)
        M = sc.obj_local2world[k].reshape(4, 4).T.astype(np.float64)   # column-major
        w = v @ M.T
        p = w[:, :3] / w[:, 3:4]
        lo, hi = np.minimum(lo, p.min(axis=0)), np.maximum(hi, p.max(axis=0))
    return lo, hi


def scene_rays(sc, n, seed):
    """Rays from random points of the scene's bounding box in random
    directions, a quarter of them with a finite range, and rays from the
    lights towards random points of the box."""
    r = rng(seed)
    lo, hi = world_bounds(sc)
    o = r.uniform(lo, hi, (n, 3))
    d = r.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    rays = np.zeros((n, 8), np.float32)
    rays[:, 0:3], rays[:, 3], rays[:, 4:7] = o, np.inf, d
    rays[::4, 3] = 0.5 * float(np.linalg.norm(np.asarray(hi) - np.asarray(lo)))
    k = n // 8
    lp = sc.lights[r.integers(0, sc.lights.shape[0], k), 4:7]
    tgt = r.uniform(lo, hi, (k, 3))
    dd = tgt - lp
    dd /= np.linalg.norm(dd, axis=1, keepdims=True)
    rays[-k:, 0:3], rays[-k:, 3], rays[-k:, 4:7] = lp, np.inf, dd
    return rays


def check_against_oracle(got, want):
    color, px, tuv, patch, obj, cnt = got
    ocolor, opx, otuv, opatch, oobj, ocnt = want
    assert cnt["overflow"] == 0
    assert (cnt["primary"], cnt["shadow"], cnt["secondary"]) == (ocnt["primary"], ocnt["shadow"], ocnt["secondary"])
    assert np.array_equal(obj, oobj) and np.array_equal(patch, opatch)
    assert_bits_equal(tuv, otuv, "first hit t,u,v")
    assert_bits_equal(color, ocolor, "color")
    ch = lambda f: np.stack([(f >> s) & 0xFF for s in (0, 8, 16, 24)], -1).astype(np.int32)
    assert np.abs(ch(px) - ch(opx)).max() <= 1      # powf in the sRGB encode


def render_rays(oracle, cam):
    """The render's own camera rays, bit for bit: directions from the
    oracle's pixelRayDirection at the integer pixel centres, row by row."""
    xy = np.stack(np.meshgrid(np.arange(cam.width), np.arange(cam.height), indexing="xy"), -1)
    d = oracle.OracleScene.pixel_rays(cam, xy.reshape(-1, 2).astype(np.float32))
    rays = np.zeros((d.shape[0], 8), np.float32)
    rays[:, 0:3], rays[:, 3], rays[:, 4:7] = cam.position, np.inf, d
    return rays


def test_camera_rays_shade_as_the_render(ctx, oracle, teaset):
    """cgs_shade of the render's own camera rays equals cgs_render's colours
    and frame bit for bit, and its ray counts."""
    wl = teaset
    ds = ctx.scene(wl.scene)
    rays = render_rays(oracle, wl.camera)
    say("cgs_shade camera rays", rays.shape[0], "depth", wl.depth)
    color, px, tuv, patch, obj, cnt = ds.shade(rays, eps=wl.eps, depth=wl.depth)
    rcolor, frame, rcnt = ds.render(wl.camera, eps=wl.eps, depth=wl.depth)
    assert_bits_equal(color, rcolor.reshape(-1, 3), "color vs render")
    assert np.array_equal(px, frame.reshape(-1))
    assert cnt == rcnt
    assert rcnt["secondary"] > 0
    # the first segment's hit is the closest hit of the camera ray
    g_tuv, g_patch, g_obj = ds.trace_closest(rays)
    assert np.array_equal(obj, g_obj) and np.array_equal(patch, g_patch)
    assert_bits_equal(tuv, g_tuv, "first hit vs trace_closest")


@pytest.mark.parametrize("depth", [0, 3])
def test_scene_rays_shade_as_the_oracle(ctx, oracle, teaset, depth):
    wl = teaset
    ds = ctx.scene(wl.scene)
    rays = scene_rays(wl.scene, 6000, seed=81 + depth)
    say("cgs_shade scene rays", rays.shape[0], "depth", depth)
    got = ds.shade(rays, eps=wl.eps, depth=depth)
    want = oracle.OracleScene(wl.scene).shade(rays, eps=wl.eps, depth=depth)
    check_against_oracle(got, want)
    hit = got[4] >= 0
    assert 1000 < hit.sum() < rays.shape[0]
    assert got[5]["shadow"] > 0


def test_closed_surfaces_shade_as_the_oracle(ctx, oracle, c2_small):
    """64 transformed objects (per-object t rescale on every segment) at
    depth 2, camera rays and scene rays."""
    wl = c2_small
    ds = ctx.scene(wl.scene)
    rays = np.concatenate([render_rays(oracle, wl.camera)[::7], scene_rays(wl.scene, 3000, seed=83)])
    say("cgs_shade closed surfaces", rays.shape[0])
    got = ds.shade(rays, eps=wl.eps, depth=2)
    want = oracle.OracleScene(wl.scene).shade(rays, eps=wl.eps, depth=2)
    check_against_oracle(got, want)


def test_served_shaded_requests_equal_shade_alone(ctx, oracle, teaset, c2_small):
    """Shaded, closest-hit and occlusion requests on two scenes in one server
    (a budget that forces several launches): every shaded request equals
    cgs_shade of its rays alone, every other request its query alone."""
    from cgspline_amd.serving import Server
    ds = [ctx.scene(teaset.scene), ctx.scene(c2_small.scene)]
    pools = [scene_rays(teaset.scene, 20000, seed=84), scene_rays(c2_small.scene, 20000, seed=85)]
    r = rng(86)
    reqs = []
    for k in range(30):
        s = int(r.integers(0, 2))
        n = int(r.integers(1, 4000))
        a = int(r.integers(0, pools[s].shape[0] - n + 1))
        kind = ("shade", "closest", "any")[k % 3]
        reqs.append((s, pools[s][a:a + n].copy(), kind, int(r.integers(0, 4))))
    sv = Server(ctx, ds, hbm_budget=8 * 4000 * 68, slots=2)
    say("serving shaded", len(reqs), "requests")
    ids = []
    for s, rays, kind, depth in reqs:
        if kind == "shade":
            ids.append(sv.submit_shaded(s, rays, eps=1e-4, depth=depth))
        else:
            ids.append(sv.submit(s, rays, kind == "any"))
    for rid, (s, rays, kind, depth) in zip(ids, reqs):
        rec, lat = sv.wait(rid)
        if kind == "shade":
            color, px, tuv, patch, obj, _ = ds[s].shade(rays, eps=1e-4, depth=depth)
            assert_bits_equal(rec["color"], color, "served color")
            assert np.array_equal(rec["pixels"], px)
            hb = rec["hits"].view(np.uint32)
            assert np.array_equal(hb[:, :3], tuv.view(np.uint32)) and np.array_equal(hb[:, 3], patch)
            assert np.array_equal(rec["objects"], obj)
        elif kind == "any":
            assert np.array_equal(rec["occluded"], ds[s].trace_any(rays))
        else:
            h, p, o = ds[s].trace_closest(rays)
            hb = rec["hits"].view(np.uint32)
            assert np.array_equal(hb[:, :3], h.view(np.uint32)) and np.array_equal(hb[:, 3], p)
            assert np.array_equal(rec["objects"], o)
    sv.close()
    # one served shaded request against the oracle
    s, rays, _, depth = next(q for q in reqs if q[2] == "shade")
    want = oracle.OracleScene([teaset.scene, c2_small.scene][s]).shade(rays, eps=1e-4, depth=depth)
    color = ds[s].shade(rays, eps=1e-4, depth=depth)[0]
    assert_bits_equal(color, want[0], "served vs oracle")


def trough_scene():
    """One bicubic patch bent into a trough along y (x, z from the curve
    (-1, 1), (-1, -1), (1, -1), (1, 1)), lit from beyond its right rim: the
    shadow rays of the floor and left wall pass through the right wall, the
    same patch they start on."""
    from cgspline_amd import scenes
    xz = [(-1.0, 1.0), (-1.0, -1.0), (1.0, -1.0), (1.0, 1.0)]
    v = np.zeros((16, 4), np.float32)
    for j in range(4):
        for i in range(4):
            v[4 * j + i] = (xz[i][0], -1.0 + 2.0 * j / 3.0, xz[i][1], 1.0)
    light = scenes.point_light((3.0, 0.2, 2.0), (6.0, 6.0, 6.0))
    mat = np.array([0.7, 0.6, 0.5, 0.04, 0.04, 0.04, 0.2, 0.4], np.float32)
    return scenes.assemble([(v, np.arange(16, dtype=np.uint32)[None])],
                           [(0, scenes.IDENTITY16, scenes.IDENTITY16, mat)], [light], name="trough")


def test_self_shadowed_patch_renders_as_the_oracle(ctx, oracle):
    """Shadow rays occluded by nothing but the patch they start on: the
    render clips a shadow ray's own patch after the rest of the scene, and
    with one patch in the scene that clip alone decides.  The oracle (its
    traversal order) shows such rays exist; the frame and counts equal it."""
    from cgspline_amd import scenes
    sc = trough_scene()
    cam = scenes.make_camera((0.3, -2.6, 3.2), (0.0, 0.0, -0.3), (0.0, 0.0, 1.0), 45.0, 96, 72)
    osc = oracle.OracleScene(sc)
    rays = scenes.camera_rays(cam)
    tuv, patch, obj, _ = osc.trace(rays)
    hit = obj >= 0
    P = rays[hit, 0:3].astype(np.float64) + tuv[hit, 0:1] * rays[hit, 4:7]
    L = np.array([3.0, 0.2, 2.0]) - P
    d = np.linalg.norm(L, axis=1)
    L /= d[:, None]
    sh = np.zeros((P.shape[0], 8), np.float32)
    sh[:, 0:3], sh[:, 3], sh[:, 4:7] = P + 1e-4 * L, d, L
    occ, _ = osc.trace(sh, any_hit=True)
    say("trough:", int(hit.sum()), "hits,", int(occ.sum()), "self-occluded shadow rays")
    assert hit.sum() > 1000 and 100 < occ.sum() < hit.sum()
    ds = ctx.scene(sc)
    color, frame, counts = ds.render(cam, eps=1e-4, depth=1)
    ocolor, oframe, _, orc = osc.render(cam, eps=1e-4, depth=1)
    assert counts["overflow"] == 0
    assert (counts["primary"], counts["shadow"], counts["secondary"]) == (orc["primary"], orc["shadow"], orc["secondary"])
    assert_bits_equal(color.reshape(-1, 3), ocolor.reshape(-1, 3), "color")
    ds.destroy()
