"""GPU parity at the edges of the ray/box/patch arithmetic: axis-aligned and
signed-zero directions (infinite reciprocals, and 0·∞ = NaN in the slab test
when an origin lies on a box plane), rays along z (the reference's clipping
projection degenerates and misses), zero-length ray extents, and empty
inputs.  Every result must equal the oracle's bit for bit."""
from __future__ import annotations

import numpy as np
import pytest

from conftest import rng, say
from test_gpu_parity import assert_bits_equal

pytestmark = pytest.mark.gpu


def _rays(o, d, tmax=np.inf):
    n = o.shape[0]
    r = np.zeros((n, 8), np.float32)
    r[:, 0:3], r[:, 3], r[:, 4:7] = o, tmax, d
    return r


def _compare_closest(ds, osc, rays, what):
    say("trace_closest", what, rays.shape[0], "rays")
    g_tuv, g_p, g_o = ds.trace_closest(rays)
    o_tuv, o_p, o_o, _ = osc.trace(rays)
    assert np.array_equal(g_o, o_o), f"{what}: object ids differ at {np.flatnonzero(g_o != o_o)[:5]}"
    assert np.array_equal(g_p, o_p), what
    assert_bits_equal(g_tuv, o_tuv, what)
    return g_o


def _compare_any(ds, osc, rays, what):
    say("trace_any", what, rays.shape[0], "rays")
    g = ds.trace_any(rays)
    o, _ = osc.trace(rays, any_hit=True)
    assert not np.any(g == 2), what
    assert np.array_equal(g == 1, o), f"{what}: occlusion differs at {np.flatnonzero((g == 1) != o)[:5]}"
    return o


def _box_planes(osc):
    """Every coordinate of the top-level BVH's child boxes, per axis."""
    nodes = osc.bvh(-1)["nodes"]
    f = nodes[:, :12].copy().view(np.float32)
    inner = nodes[:, 12] != 0xFFFFFFFF
    boxes = f[inner].reshape(-1, 4, 3)           # lmin, lmax, rmin, rmax
    return [np.unique(boxes[:, :, a]) for a in range(3)]


def test_axis_aligned_rays_bitexact(ctx, oracle, c2_small):
    """Rays along ±x, ±y, ±z (two infinite reciprocals each), some with a
    negative zero, and some whose origin lies exactly on a top-level box
    plane, where the slab test meets 0·∞."""
    sc = c2_small.scene
    ds = ctx.scene(sc)
    osc = oracle.OracleScene(sc)
    planes = _box_planes(osc)
    r = rng(41)
    g = np.linspace(0.25, 9.75, 24, dtype=np.float32)
    A, B = np.meshgrid(g, g, indexing="xy")
    A, B = A.reshape(-1), B.reshape(-1)
    total_hits = 0
    for axis in range(3):
        a1, a2 = [k for k in range(3) if k != axis]
        for sign in (1.0, -1.0):
            o = np.zeros((A.size, 3), np.float32)
            o[:, axis] = -2.0 if sign > 0 else 12.0
            o[:, a1], o[:, a2] = A, B
            # a quarter of the rays start on a box plane of the other two axes
            on = r.random(A.size) < 0.25
            o[on, a1] = r.choice(planes[a1], on.sum())
            on2 = r.random(A.size) < 0.25
            o[on2, a2] = r.choice(planes[a2], on2.sum())
            d = np.zeros((A.size, 3), np.float32)
            d[:, axis] = sign
            neg = r.random(A.size) < 0.5                  # -0.0 in the other components
            d[neg, a1] = -0.0
            d[~neg, a2] = -0.0
            rays = _rays(o, d)
            obj = _compare_closest(ds, osc, rays, f"axis {axis} sign {sign:+.0f}")
            total_hits += int((obj >= 0).sum())
            _compare_any(ds, osc, rays, f"axis {axis} sign {sign:+.0f}")
    assert total_hits > 100


def test_signed_zero_directions_share_a_wave(ctx, oracle, c2_small):
    """Rays of one 64-ray chunk that differ only in the sign of a zero
    direction component fall in different direction octants, so the
    traversal cannot use the single-octant loop for them; the results
    still equal the oracle's."""
    sc = c2_small.scene
    ds = ctx.scene(sc)
    osc = oracle.OracleScene(sc)
    g = np.linspace(0.5, 9.5, 32, dtype=np.float32)
    A, B = np.meshgrid(g, g, indexing="xy")
    n = A.size
    o = np.stack([np.full(n, -2.0, np.float32), A.reshape(-1), B.reshape(-1)], -1)
    d = np.zeros((n, 3), np.float32)
    d[:, 0] = 1.0
    d[:, 1] = np.where(np.arange(n) % 2 == 0, 0.0, -0.0).astype(np.float32)
    d[:, 2] = np.where(np.arange(n) % 3 == 0, -0.0, 0.0).astype(np.float32)
    obj = _compare_closest(ds, osc, _rays(o, d), "signed zeros")
    assert (obj >= 0).sum() > 10


def test_rays_along_z_miss_every_patch(ctx, oracle, c1):
    """In object space a direction along ±z makes the clipping projection
    d×(d+(0,0,8)) vanish (Spline_BezierClipping.cu:59-60), so such rays
    never hit, straight down onto the height field included; the same rays
    tilted slightly do hit."""
    ds = ctx.scene(c1.scene)
    osc = oracle.OracleScene(c1.scene)
    g = np.linspace(0.05, 0.95, 16, dtype=np.float32)
    A, B = np.meshgrid(g, g, indexing="xy")
    n = A.size
    o = np.stack([A.reshape(-1), B.reshape(-1), np.full(n, 2.0, np.float32)], -1)
    down = np.tile(np.array([0.0, 0.0, -1.0], np.float32), (n, 1))
    obj = _compare_closest(ds, osc, _rays(o, down), "straight down")
    assert (obj < 0).all()
    up = _rays(o * np.float32([1, 1, -1]), -down)
    assert (_compare_closest(ds, osc, up, "straight up") < 0).all()
    tilted = down.copy()
    tilted[:, 0] = 1e-2
    tilted /= np.linalg.norm(tilted, axis=1, keepdims=True)
    assert (_compare_closest(ds, osc, _rays(o, tilted.astype(np.float32)), "tilted") >= 0).sum() > n // 2


def test_clip_along_z_and_axis_directions(ctx, oracle):
    """The per-patch clipper on directions along ±z (always a miss) and
    along ±x, ±y (infinite reciprocals do not enter the clipper)."""
    from test_gpu_parity import gpu_clip, random_patch_pairs
    rays, P = random_patch_pairs(512, seed=43)
    target = P[:, :, :3].mean(axis=1)
    outs = []
    for d in ([0, 0, 1], [0, 0, -1], [1, 0, 0], [0, -1, 0], [-0.0, 0.0, -1.0]):
        d = np.asarray(d, np.float32)
        o = (target - 3.0 * d[None, :]).astype(np.float32)
        rr = _rays(o, np.tile(d, (o.shape[0], 1)))
        say("bezier_clip direction", d.tolist())
        g_h, g_f = gpu_clip(ctx, rr, P)
        o_h, _ = oracle.clip_batch(rr, P)
        o_f = o_h[:, 3] == 1
        assert np.array_equal(g_f, o_f), d
        assert_bits_equal(g_h[g_f], o_h[o_f, :3], f"clip along {d.tolist()}")
        outs.append(g_f)
    assert not outs[0].any() and not outs[1].any() and not outs[4].any()
    assert outs[2].sum() + outs[3].sum() > 50


def test_zero_extent_rays_hit_nothing(ctx, oracle, c2_small):
    """tmax = 0 culls every box (entry distance < tmax fails): no hits, no
    occlusion; tmax = smallest denormal behaves the same."""
    from cgspline_amd import scenes
    ds = ctx.scene(c2_small.scene)
    osc = oracle.OracleScene(c2_small.scene)
    rays = scenes.camera_rays(c2_small.camera)[::97].copy()
    for tmax in (0.0, np.float32(1e-45)):
        rays[:, 3] = tmax
        assert (_compare_closest(ds, osc, rays, f"tmax {tmax}") < 0).all()
        assert not _compare_any(ds, osc, rays, f"tmax {tmax}").any()


def test_empty_inputs(ctx, c2_small):
    """Zero rays or zero patches: every entry point returns at once with
    empty results."""
    ds = ctx.scene(c2_small.scene)
    none = np.zeros((0, 8), np.float32)
    say("empty inputs")
    tuv, p, o = ds.trace_closest(none)
    assert tuv.shape == (0, 3) and p.shape == (0,) and o.shape == (0,)
    assert ds.trace_any(none).shape == (0,)
    assert ctx.bezier_clip(none, np.zeros((0, 16, 4), np.float32)).shape == (0, 4)
    assert ctx.subdivision(none, np.zeros((0, 16, 4), np.float32)).shape == (0, 4)


def test_scene_outlives_the_context_that_made_it(ctx, oracle, c1):
    """A scene's device arrays are the scene's, not its creating context's:
    after that context is closed the scene still traces and renders through
    another one, with the oracle's results."""
    import cgspline_amd as cgs
    from cgspline_amd import scenes
    other = cgs.Context(0)
    ds = other.scene(c1.scene)
    say("close the creating context, then trace and render through another")
    other.close()
    v = ds.via(ctx)
    rays = scenes.camera_rays(c1.camera)[::7]
    h, p, o = v.trace_closest(rays)
    oh, op, oo, _ = oracle.OracleScene(c1.scene).trace(rays)
    assert np.array_equal(p, op) and np.array_equal(o, oo)
    assert np.array_equal(h.view(np.uint32), oh.view(np.uint32))
    color, _, _ = v.render(c1.camera, eps=c1.eps, depth=c1.depth)
    ocolor, _, _, _ = oracle.OracleScene(c1.scene).render(c1.camera, eps=c1.eps, depth=c1.depth)
    assert np.array_equal(color.view(np.uint32), ocolor.view(np.uint32))
    lib = cgs.lib()
    lib.cgs_scene_destroy(ds.h)
    ds.h = None


def test_degenerate_rays_bitexact(ctx, oracle, c2_small):
    """Rays with no usable axis: a zero direction (all three reciprocals
    infinite) from origins on top-level box corners, NaN directions and
    infinite origins.  For these the octant-specialised node loop would differ
    from the reference's seeded slab test, so the wave falls back to the
    generic loop; every result still equals the oracle's."""
    sc = c2_small.scene
    ds = ctx.scene(sc)
    osc = oracle.OracleScene(sc)
    planes = _box_planes(osc)
    r = rng(47)
    n = 512
    o = np.stack([r.choice(planes[a], n) for a in range(3)], -1).astype(np.float32)
    d = np.zeros((n, 3), np.float32)
    d[r.random((n, 3)) < 0.5] = -0.0
    rays = _rays(o, d)
    _compare_closest(ds, osc, rays, "zero directions from box corners")
    _compare_any(ds, osc, rays, "zero directions from box corners")
    mixed = _rays(o, np.tile(np.float32([0.3, 0.4, 0.5]), (n, 1)))
    mixed[::3, 4:7] = np.nan                      # NaN directions in every wave
    mixed[1::5, 0] = np.inf                       # infinite origins
    _compare_closest(ds, osc, mixed, "NaN directions and infinite origins")
    _compare_any(ds, osc, mixed, "NaN directions and infinite origins")


def test_non_finite_geometry_bitexact(ctx, oracle, c1):
    """A control point at +inf makes the top-level box infinite: the
    octant-specialised loop is not used for that BVH, and hits equal the
    oracle's."""
    import dataclasses
    from cgspline_amd import scenes
    v = c1.scene.vertices.copy()
    v[c1.scene.indices[17, 5], 2] = np.inf
    sc = dataclasses.replace(c1.scene, vertices=v)
    ds = ctx.scene(sc)
    osc = oracle.OracleScene(sc)
    rays = scenes.camera_rays(c1.camera)[::7]
    obj = _compare_closest(ds, osc, rays, "infinite control point")
    assert (obj >= 0).sum() > 100
