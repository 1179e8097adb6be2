"""The cull tree (cgs_cull.cpp): scenes of two or more objects find the
objects a ray reaches through an SAH BVH over the reference's top-level leaf
boxes, and fall back to the reference's top-level traversal where the visit
order could decide the answer.  Every result equals the oracle's (which
walks the reference's tree) and the same scene's results with the tree off,
bit for bit, including scenes built to reach the fallbacks: objects
duplicated in place (every hit an exact tie between two objects), and thin
grids with rays at their corners (a hit at or near the box entry)."""
from __future__ import annotations

import contextlib
import os

import numpy as np
import pytest

from conftest import say
from test_gpu_parity import assert_bits_equal

pytestmark = pytest.mark.gpu


@contextlib.contextmanager
def cull_off():
    old = os.environ.get("CGS_CULL")
    os.environ["CGS_CULL"] = "0"
    try:
        yield
    finally:
        if old is None:
            del os.environ["CGS_CULL"]
        else:
            os.environ["CGS_CULL"] = old


def _compare(ctx, oracle, sc, rays, what, render=None):
    """Closest and any hits with the tree on, off and in the oracle."""
    on = ctx.scene(sc)
    with cull_off():
        off = ctx.scene(sc)
    assert on.culled and not off.culled, what
    osc = oracle.OracleScene(sc)
    say("cull tree", what, rays.shape[0], "rays")
    a_tuv, a_p, a_o = on.trace_closest(rays)
    b_tuv, b_p, b_o = off.trace_closest(rays)
    o_tuv, o_p, o_o, _ = osc.trace(rays)
    for tuv, p, o, who in ((a_tuv, a_p, a_o, "tree"), (b_tuv, b_p, b_o, "no tree")):
        assert np.array_equal(o, o_o), f"{what} ({who}): objects differ at {np.flatnonzero(o != o_o)[:5]}"
        assert np.array_equal(p, o_p), f"{what} ({who}): patches differ at {np.flatnonzero(p != o_p)[:5]}"
        assert_bits_equal(tuv, o_tuv, f"{what} ({who})")
    occ = on.trace_any(rays)
    o_occ, _ = osc.trace(rays, any_hit=True)
    assert np.array_equal(occ == 1, o_occ) and np.array_equal(occ, off.trace_any(rays)), what
    if render is not None:
        cam, eps, depth = render
        c1, f1, n1 = on.render(cam, eps=eps, depth=depth)
        c2, f2, n2 = off.render(cam, eps=eps, depth=depth)
        oc, of, _, _ = osc.render(cam, eps=eps, depth=depth)
        assert np.array_equal(c1.view(np.uint32), oc.view(np.uint32)), what
        assert np.array_equal(c2.view(np.uint32), oc.view(np.uint32)), what
        assert n1 == n2
    return a_o


def test_tree_is_built_where_allowed(ctx, c1, c2_small, c3_small):
    """Multi-object scenes of weight-1 patches under affine transforms get the
    tree; a one-object scene, a rational patch and CGS_CULL=0 do not."""
    from test_oracle_polish import rational_case
    say("cull tree flags")
    assert ctx.scene(c2_small.scene).culled and ctx.scene(c3_small.scene).culled
    assert not ctx.scene(c1.scene).culled
    sc, _, _, _ = rational_case()
    two = _duplicate(sc)
    assert not ctx.scene(two).culled
    with cull_off():
        assert not ctx.scene(c3_small.scene).culled


def test_tree_equals_reference_traversal(ctx, oracle, c3_small):
    from cgspline_amd import scenes
    rays = scenes.camera_rays(c3_small.camera)[::3]
    obj = _compare(ctx, oracle, c3_small.scene, rays, "displaced grids",
                   render=(c3_small.camera, c3_small.eps, c3_small.depth))
    assert (obj >= 0).sum() > 500


def _duplicate(sc, shift=None):
    """The scene with every object listed twice (same surface, same
    transform): each hit is an exact tie between the two copies."""
    import dataclasses
    n = sc.nobj
    order = np.repeat(np.arange(n), 2)
    return dataclasses.replace(sc, obj_surface=sc.obj_surface[order], obj_local2world=sc.obj_local2world[order],
                               obj_world2local=sc.obj_world2local[order], obj_material=sc.obj_material[order])


def test_exact_ties_between_objects(ctx, oracle, c2_small):
    """Duplicated objects: every hit ties; the reported object is the one the
    reference's traversal visits first."""
    from cgspline_amd import scenes
    sc = _duplicate(c2_small.scene)
    rays = scenes.camera_rays(c2_small.camera)[::5]
    obj = _compare(ctx, oracle, sc, rays, "duplicated objects",
                   render=(c2_small.camera, c2_small.eps, c2_small.depth))
    hit = obj >= 0
    assert hit.sum() > 500
    # both copies of some objects are reported: the order, not the id, decides
    assert (obj[hit] % 2 == 0).any() and (obj[hit] % 2 == 1).any()


def test_hits_near_the_box_entry(ctx, oracle):
    """Thin grids (height 1e-4 of their size, so a hit lies close to the box
    entry) tilted against the rays, and rays aimed at the grids' corner
    control points, where a hit lies on the box boundary.  (An exactly flat
    grid is never hit: its zero-thickness boxes fail the reference's strict
    tmin < tmax, rt/BVH.h:147-176.)"""
    from cgspline_amd import scenes
    rng = np.random.default_rng(77)
    surfaces, objects = [], []
    for k in range(24):
        v, i = scenes.heightfield_grid(rng, amplitude=1e-4, n=4)
        surfaces.append((v, i))
        tilt = scenes.rotation_xyz([rng.uniform(-20, 20), rng.uniform(-20, 20), rng.uniform(0, 360)])[:3, :3]
        l2w, w2l = scenes.trs(np.array([rng.uniform(0, 4), rng.uniform(0, 4), 0.5 * k]), tilt, rng.uniform(0.5, 2))
        objects.append((k, l2w, w2l, np.array([0.5, 0.5, 0.5, 0.04, 0.04, 0.04, 0.0, 0.5], np.float32)))
    sc = scenes.assemble(surfaces, objects, [scenes.point_light((2.0, 2.0, 30.0), (900.0, 900.0, 900.0))],
                         name="planar grids")
    cam = scenes.make_camera((2.3, 1.7, 25.0), (2.0, 2.0, 0.0), (0.0, 1.0, 0.0), 40.0, 96, 96)
    rays = scenes.camera_rays(cam)
    # rays from the camera at each grid's four corner control points (world)
    corners = []
    for k, (v, _) in enumerate(surfaces):
        L = objects[k][1].reshape(4, 4).T.astype(np.float64)
        for c in (0, 12, v.shape[0] - 13, v.shape[0] - 1):
            corners.append(L[:3, :3] @ v[c, :3].astype(np.float64) + L[:3, 3])
    corners = np.array(corners)
    org = np.array([2.3, 1.7, 25.0])
    d = corners - org
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    extra = np.zeros((d.shape[0], 8), np.float32)
    extra[:, 0:3], extra[:, 3], extra[:, 4:7] = org, np.inf, d
    rays = np.concatenate([rays, extra])
    obj = _compare(ctx, oracle, sc, rays, "thin grids", render=(cam, 1e-4, 1))
    assert (obj >= 0).sum() > 1000


# ---------------------------------------------------------------------------
# cull trees over the patches of one large surface
# ---------------------------------------------------------------------------

def one_surface(sc, repeat=1):
    """Every object's patches mapped to world space as one surface under one
    identity object (each patch listed `repeat` times in a row: with 2 every
    hit is an exact tie between two patches)."""
    from cgspline_amd import scenes
    parts, idx, nv = [], [], 0
    for k in range(sc.nobj):
        s = int(sc.obj_surface[k])
        v0, v1 = int(sc.surf_vbase[s]), int(sc.surf_vbase[s + 1])
        p0, p1 = int(sc.surf_pbase[s]), int(sc.surf_pbase[s + 1])
        M = sc.obj_local2world[k].reshape(4, 4).T.astype(np.float64)
        w = sc.vertices[v0:v1].astype(np.float64) @ M.T
        parts.append(np.column_stack([w[:, :3] / w[:, 3:4], np.ones(v1 - v0)]).astype(np.float32))
        idx.append(np.repeat(sc.indices[p0:p1], repeat, axis=0) + nv)
        nv += v1 - v0
    return scenes.assemble([(np.concatenate(parts), np.concatenate(idx))],
                           [(0, scenes.IDENTITY16, scenes.IDENTITY16, sc.obj_material[0])], sc.lights,
                           name="one surface")


def _compare_surface(ctx, oracle, sc, rays, what, render=None):
    """As _compare, for a one-object scene whose surface has the tree."""
    on = ctx.scene(sc)
    with cull_off():
        off = ctx.scene(sc)
    assert on.bvh_culled(0) and not off.bvh_culled(0) and not on.culled, what
    osc = oracle.OracleScene(sc)
    say("surface cull tree", what, sc.npatch, "patches", rays.shape[0], "rays")
    o_tuv, o_p, o_o, _ = osc.trace(rays)
    for ds, who in ((on, "tree"), (off, "no tree")):
        tuv, p, o = ds.trace_closest(rays)
        assert np.array_equal(o, o_o), f"{what} ({who}): objects differ at {np.flatnonzero(o != o_o)[:5]}"
        assert np.array_equal(p, o_p), f"{what} ({who}): patches differ at {np.flatnonzero(p != o_p)[:5]}"
        assert_bits_equal(tuv, o_tuv, f"{what} ({who})")
    occ = on.trace_any(rays)
    o_occ, _ = osc.trace(rays, any_hit=True)
    assert np.array_equal(occ == 1, o_occ) and np.array_equal(occ, off.trace_any(rays)), what
    if render is not None:
        cam, eps, depth = render
        c1, _, n1 = on.render(cam, eps=eps, depth=depth)
        c2, _, n2 = off.render(cam, eps=eps, depth=depth)
        oc, _, _, _ = osc.render(cam, eps=eps, depth=depth)
        assert np.array_equal(c1.view(np.uint32), oc.view(np.uint32)), what
        assert np.array_equal(c2.view(np.uint32), oc.view(np.uint32)), what
        assert n1 == n2
    return o_p


def test_surface_tree_equals_reference_traversal(ctx, oracle, c3_small):
    """The 96 displaced grids (24,576 patches) as one surface: closest and
    any hits and the frame with the surface's tree equal the reference's
    traversal and the oracle."""
    from cgspline_amd import scenes
    sc = one_surface(c3_small.scene)
    cam = scenes.make_camera(*_view(c3_small.camera), 128, 128)
    rays = scenes.camera_rays(c3_small.camera)[::5]
    patch = _compare_surface(ctx, oracle, sc, rays, "grids as one surface", render=(cam, c3_small.eps, c3_small.depth))
    assert (patch != 0xFFFFFFFF).sum() > 300


def _view(cam):
    """(position, target, up, fovy in degrees) of a look-at camera."""
    import math
    M = cam.to_world.reshape(4, 4).astype(np.float64)
    pos = M[3, :3]
    return pos, pos - M[2, :3], M[1, :3], math.degrees(cam.fovy)


def test_exact_ties_between_patches(ctx, oracle, c3_small):
    """Every patch listed twice: each hit ties between the two copies, and
    the reported patch is the one the reference's traversal reaches first."""
    from cgspline_amd import scenes
    sc = one_surface(c3_small.scene, repeat=2)
    rays = scenes.camera_rays(c3_small.camera)[::7]
    patch = _compare_surface(ctx, oracle, sc, rays, "duplicated patches")
    hit = patch != 0xFFFFFFFF
    assert hit.sum() > 200
    assert (patch[hit] % 2 == 0).any() and (patch[hit] % 2 == 1).any()


def test_patch_hits_near_the_box_entry(ctx, oracle):
    """Thin grids (height 1e-4 of their size) as one surface of 4096 patches,
    tilted against the rays, and rays aimed at every patch corner of a few
    grids, where a hit lies on its patch's box boundary."""
    from cgspline_amd import scenes
    rng = np.random.default_rng(78)
    parts, idx, nv = [], [], 0
    for k in range(16):
        v, i = scenes.heightfield_grid(rng, amplitude=1e-4)
        tilt = scenes.rotation_xyz([rng.uniform(-20, 20), rng.uniform(-20, 20), rng.uniform(0, 360)])[:3, :3]
        p = (v[:, :3].astype(np.float64) @ tilt.T) * rng.uniform(0.5, 2) + [rng.uniform(0, 4), rng.uniform(0, 4), 0.5 * k]
        parts.append(np.column_stack([p, np.ones(v.shape[0])]).astype(np.float32))
        idx.append(i + nv)
        nv += v.shape[0]
    V = np.concatenate(parts)
    I = np.concatenate(idx)
    sc = scenes.assemble([(V, I)], [(0, scenes.IDENTITY16, scenes.IDENTITY16,
                                     np.array([0.5, 0.5, 0.5, 0.04, 0.04, 0.04, 0.0, 0.5], np.float32))],
                         [scenes.point_light((2.0, 2.0, 30.0), (900.0, 900.0, 900.0))], name="thin grids, one surface")
    assert sc.npatch == 4096
    cam = scenes.make_camera((2.3, 1.7, 25.0), (2.0, 2.0, 0.0), (0.0, 1.0, 0.0), 40.0, 96, 96)
    org = np.array([2.3, 1.7, 25.0])
    corners = V[np.unique(I[::37][:, [0, 3, 12, 15]]), :3].astype(np.float64)
    d = corners - org
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    extra = np.zeros((d.shape[0], 8), np.float32)
    extra[:, 0:3], extra[:, 3], extra[:, 4:7] = org, np.inf, d
    rays = np.concatenate([scenes.camera_rays(cam), extra])
    patch = _compare_surface(ctx, oracle, sc, rays, "thin grids, one surface", render=(cam, 1e-4, 1))
    assert (patch != 0xFFFFFFFF).sum() > 1000
