"""GPU parity: the HIP path (through the C ABI) against the CPU oracle.

Integer and index results (BVH nodes, keys, hash tables, hit patch/object
ids) must be identical.  FP32 results (t, u, v, shading) are computed with the
same operation order, no FMA contraction and IEEE division/sqrt on both sides,
so they must be identical bit for bit as well; the only tolerance is ±1 in an
8-bit channel of the packed frame (sRGB encoding uses powf, whose last bit
differs between libm and the device library).
"""
from __future__ import annotations

import numpy as np
import pytest

from conftest import rng, say

pytestmark = pytest.mark.gpu

EMPTY = 0xFFFFFFFF


def bits(a):
    return np.ascontiguousarray(a, np.float32).view(np.uint32)


def assert_bits_equal(a, b, what):
    ba, bb = bits(a), bits(b)
    bad = np.flatnonzero((ba != bb).reshape(ba.shape[0], int(np.prod(ba.shape[1:]))).any(axis=1)) if ba.ndim > 1 \
        else np.flatnonzero(ba != bb)
    assert bad.size == 0, f"{what}: {bad.size} mismatches, first {bad[:5]}: {np.asarray(a)[bad[:3]]} vs {np.asarray(b)[bad[:3]]}"


def pairs_c1(oracle, wl, step=5, extra=2, seed=3):
    """(ray, patch) pairs from the height-field workload: every `step`-th
    camera ray against the patch it hits plus `extra` random patches."""
    from cgspline_amd import scenes
    sc = wl.scene
    rays = scenes.camera_rays(wl.camera)[::step]
    osc = oracle.OracleScene(sc)
    _, patch, obj, _ = osc.trace(rays)
    P = sc.surface_patches(0)
    r = rng(seed)
    ids = [np.where(obj >= 0, patch, r.integers(0, P.shape[0], rays.shape[0]))]
    for _ in range(extra):
        ids.append(r.integers(0, P.shape[0], rays.shape[0]))
    ids = np.concatenate(ids).astype(np.int64)
    R = np.concatenate([rays] * (extra + 1))
    return R, P[ids]


def random_patch_pairs(n, seed):
    """Random control nets (wavy, folded, degenerate rows) and rays aimed at
    random points near them."""
    r = rng(seed)
    base = np.stack(np.meshgrid(np.linspace(0, 1, 4), np.linspace(0, 1, 4), indexing="xy"), -1).reshape(16, 2)
    P = np.zeros((n, 16, 4), np.float32)
    scale = np.exp(r.uniform(np.log(1e-2), np.log(10.0), n))[:, None]
    P[:, :, 0] = base[None, :, 0] * scale + r.normal(0, 0.15, (n, 16)) * scale
    P[:, :, 1] = base[None, :, 1] * scale + r.normal(0, 0.15, (n, 16)) * scale
    P[:, :, 2] = r.normal(0, 0.4, (n, 16)) * scale
    P[:, :, 3] = 1.0
    deg = r.random(n) < 0.15                       # collapse a row to a point
    P[deg, 0:4, :3] = P[deg, 0:1, :3]
    target = P[:, :, :3].mean(axis=1) + r.normal(0, 0.3, (n, 3)) * scale
    o = target + r.normal(0, 1, (n, 3)) * 3 * scale
    d = target - o
    rays = np.zeros((n, 8), np.float32)
    rays[:, 0:3], rays[:, 3], rays[:, 4:7] = o, np.inf, d
    return rays, P


def gpu_clip(ctx, rays, P):
    out = ctx.bezier_clip(rays, P)
    found = out.view(np.uint32)[:, 3]
    assert not np.any(found == 0xFFFFFFFE), "clip stack overflow"
    return out[:, :3], found == 1


def test_clip_batch_heightfield_bitexact(ctx, oracle, c1):
    rays, P = pairs_c1(oracle, c1)
    say("bezier_clip heightfield pairs", rays.shape[0])
    g, gf = gpu_clip(ctx, rays, P)
    o, _ = oracle.clip_batch(rays, P)
    assert np.array_equal(gf, o[:, 3] == 1)
    assert gf.sum() > 1000
    assert_bits_equal(g, o[:, :3], "t,u,v")


@pytest.mark.parametrize("seed", [11, 12])
def test_clip_batch_random_bitexact(ctx, oracle, seed):
    rays, P = random_patch_pairs(20000, seed)
    say("bezier_clip random patches seed", seed)
    g, gf = gpu_clip(ctx, rays, P)
    o, w = oracle.clip_batch(rays, P)
    assert np.array_equal(gf, o[:, 3] == 1)
    assert gf.sum() > 2000 and w["subdivs"] > 0 and w["prunes"] > 0
    assert_bits_equal(g, o[:, :3], "t,u,v")


def test_subdivision_batch_bitexact(ctx, oracle, c1):
    rays, P = pairs_c1(oracle, c1, step=11, extra=1)
    for thr in (2.0 ** -4, 2.0 ** -7):
        say("subdivision threshold", thr)
        g = ctx.subdivision(rays, P, thr)
        o, _ = oracle.clip_batch(rays, P, method=1, threshold=thr)
        gf = g.view(np.uint32)[:, 3]
        assert not np.any(gf == 0xFFFFFFFE)
        assert np.array_equal(gf == 1, o[:, 3] == 1)
        assert_bits_equal(g[:, :3], o[:, :3], "t,u,v")


def heap_keys(npatch):
    keys, stack = [], [(1, npatch)]
    while stack:
        k, c = stack.pop()
        keys.append(k)
        if c > 1:
            m = c // 2
            stack.append(((k << 1) ^ 1, c - m))
            stack.append((k << 1, m))
    return np.array(keys, np.uint32)


@pytest.mark.parametrize("keys", [
    "heap1", "heap2", "heap3", "heap256", "heap4096", "heap33000", "shuffled4096", "random3000",
    "heap524288", "heap1048576"])
def test_phf_build_matches_oracle(ctx, oracle, keys):
    """The GPU builder against the oracle, up to the largest BVHs: 2^19
    patches (1,048,575 keys) hash after 7 attempts; 2^20 patches (2,097,151
    keys) fail all 16, and both report the same status and parameters."""
    r = rng(5)
    if keys.startswith("heap"):
        k = heap_keys(int(keys[4:]))
    elif keys.startswith("shuffled"):
        k = heap_keys(int(keys[8:]))
        r.shuffle(k)
    else:
        k = r.choice(np.uint32(2 ** 31), int(keys[6:]), replace=False).astype(np.uint32)
    say("phf build", keys, k.shape[0])
    st, p, D = ctx.phf_build(k)
    ost, op, oD = oracle.phf_build(k)
    code = {0: 0, 1: 4, 2: 5}[ost]
    assert st == code
    assert p == op
    assert np.array_equal(D, oD)
    if st == 0:
        h = ctx.phf_hash(p, D, k)
        assert np.unique(h).size == k.size and h.max() < p["table_size"]


def test_scene_bvh_and_hash_match_oracle(ctx, oracle, c2_small):
    sc = c2_small.scene
    say("scene create closed surfaces", sc.npatch)
    ds = ctx.scene(sc)
    osc = oracle.OracleScene(sc)
    for s in [0, 1, 17, sc.nsurf - 1, -1]:
        g, o = ds.bvh(s), osc.bvh(s)
        for key in ("nodes", "keys", "indices", "displacement", "table"):
            assert np.array_equal(g[key], o[key]), (s, key)
        for key in ("A", "B", "mask", "table_size", "attempts"):
            assert g[key] == o[key], (s, key)


def _trace_compare(ctx, oracle, wl, rows=None, label=""):
    from cgspline_amd import scenes
    sc = wl.scene
    rays = scenes.camera_rays(wl.camera, rows)
    say("trace_closest", label, rays.shape[0], "rays")
    ds = ctx.scene(sc)
    g_tuv, g_p, g_o = ds.trace_closest(rays)
    osc = oracle.OracleScene(sc)
    o_tuv, o_p, o_o, w = osc.trace(rays)
    assert np.array_equal(g_o, o_o), f"object ids differ at {np.flatnonzero(g_o != o_o)[:5]}"
    assert np.array_equal(g_p, o_p)
    assert_bits_equal(g_tuv, o_tuv, "t,u,v")
    assert (g_o >= 0).sum() > 0
    return ds, osc, rays, g_tuv, g_o


def test_trace_closest_heightfield_bitexact(ctx, oracle, c1):
    _trace_compare(ctx, oracle, c1, label="heightfield")


def test_trace_closest_closed_surfaces_bitexact(ctx, oracle, c2_small):
    _trace_compare(ctx, oracle, c2_small, label="closed surfaces")


def test_trace_closest_displaced_grids_bitexact(ctx, oracle, c3_small):
    _trace_compare(ctx, oracle, c3_small, label="displaced grids")


def _shadow_rays(scene, rays, tuv, obj):
    """Rays from each hit point towards the first light, ending at it."""
    light = scene.lights[0, 4:7]
    hit = obj >= 0
    P = rays[hit, 0:3] + tuv[hit, 0:1] * rays[hit, 4:7]
    L = light[None, :] - P
    d = np.linalg.norm(L, axis=1).astype(np.float32)
    L = (L / d[:, None]).astype(np.float32)
    sh = np.zeros((P.shape[0], 8), np.float32)
    sh[:, 0:3] = P + 1e-4 * L
    sh[:, 3] = d
    sh[:, 4:7] = L
    return sh


def _any_compare(ds, osc, sh):
    say("trace_any", sh.shape[0], "shadow rays")
    g = ds.trace_any(sh)
    o, w = osc.trace(sh, any_hit=True)
    assert not np.any(g == 2)
    assert np.array_equal(g == 1, o)
    return o, w


def test_trace_any_bitexact(ctx, oracle, c2_small):
    ds, osc, rays, tuv, obj = _trace_compare(ctx, oracle, c2_small, label="closed surfaces (shadow set)")
    o, _ = _any_compare(ds, osc, _shadow_rays(c2_small.scene, rays, tuv, obj))
    assert 0 < o.sum() < o.size


def test_trace_past_the_sibling_ring_bitexact(ctx, oracle, grazing):
    """Traversals that postpone more siblings at once than the device keeps
    in its per-lane LDS ring (the 8 most recent) find the older ones through
    the perfect hash, as the reference finds every sibling (rt/BVH.h:
    178-266, 268-350): closest hits and shadow answers still equal the
    oracle's.  The oracle's count of postponed siblings shows that the ring
    overflowed."""
    ds, osc, rays, tuv, obj = _trace_compare(ctx, oracle, grazing, label="grazing height field")
    _, _, _, w = osc.trace(rays)
    assert w["maxPostponed"] > 8, w
    o, w_any = _any_compare(ds, osc, _shadow_rays(grazing.scene, rays, tuv, obj))
    assert w_any["maxPostponed"] > 8, w_any
    assert 0 < o.sum() < o.size


def test_trace_deep_top_level_bitexact(ctx, oracle, stacked):
    """180,000 stacked objects: rays postpone 17 top-level siblings at once,
    so past the ring's 8 the traversal reads a sibling whose heap key is 512
    or more (not staged in LDS) through the top level's perfect hash.  Hits
    and occlusion equal the oracle's."""
    sc, rays = stacked
    say("scene create: stacked arcs", sc.nobj, "objects")
    ds = ctx.scene(sc)
    osc = oracle.OracleScene(sc)
    say("trace_closest through", sc.nobj, "objects")
    g_tuv, g_p, g_o = ds.trace_closest(rays)
    o_tuv, o_p, o_o, w = osc.trace(rays)
    assert w["maxPostponed"] >= 17
    assert np.array_equal(g_o, o_o) and np.array_equal(g_p, o_p)
    assert_bits_equal(g_tuv, o_tuv, "t,u,v")
    assert (g_o == sc.nobj - 1).all()
    say("trace_any through", sc.nobj, "objects")
    _any_compare(ds, osc, rays)


def test_bench_frame_matches_oracle(ctx, oracle):
    """The benchmark frame itself: 4,096 displaced grids (1,048,576 patches),
    2048×2048, depth 1.  Rendered twice (the second render takes the tiles
    longest first), both equal the oracle's frame, colours and ray counts."""
    from cgspline_amd import scenes
    wl = scenes.displaced_grids()
    say("scene create: 1M-patch bench scene")
    ds = ctx.scene(wl.scene)
    say("oracle render of the bench frame")
    ocolor, oframe, _, orc = oracle.OracleScene(wl.scene).render(wl.camera, eps=wl.eps, depth=wl.depth)
    ch = lambda f: np.stack([(f >> s) & 0xFF for s in (0, 8, 16, 24)], -1).astype(np.int32)
    for k in range(2):
        say("render the bench frame, pass", k)
        color, frame, counts = ds.render(wl.camera, eps=wl.eps, depth=wl.depth)
        assert counts["overflow"] == 0
        assert (counts["primary"], counts["shadow"], counts["secondary"]) == \
            (orc["primary"], orc["shadow"], orc["secondary"])
        assert_bits_equal(color.reshape(-1, 3), ocolor.reshape(-1, 3), f"color, pass {k}")
        assert np.abs(ch(frame) - ch(oframe)).max() <= 1


@pytest.mark.parametrize("name", ["c1", "c2_small", "c3_small", "grazing"])
def test_render_matches_oracle(ctx, oracle, request, name):
    wl = request.getfixturevalue(name)
    say("render", name, wl.camera.width, wl.camera.height, "depth", wl.depth)
    ds = ctx.scene(wl.scene)
    color, frame, counts = ds.render(wl.camera, eps=wl.eps, depth=wl.depth)
    osc = oracle.OracleScene(wl.scene)
    ocolor, oframe, _, orc = osc.render(wl.camera, eps=wl.eps, depth=wl.depth)
    assert counts["overflow"] == 0
    assert (counts["primary"], counts["shadow"], counts["secondary"]) == (orc["primary"], orc["shadow"], orc["secondary"])
    assert_bits_equal(color.reshape(-1, 3), ocolor.reshape(-1, 3), "color")
    ch = lambda f: np.stack([(f >> s) & 0xFF for s in (0, 8, 16, 24)], -1).astype(np.int32)
    assert np.abs(ch(frame) - ch(oframe)).max() <= 1


def test_geometryless_objects_match_oracle(ctx, oracle, c2_small):
    """A scene mixing surfaces with objects that have no geometry (the
    reference's mesh and sphere primitives: empty bounds, never hit): the
    top-level BVH and hash, the hits and the frame equal the oracle's."""
    from test_oracle_bvh import with_geometryless_objects
    from cgspline_amd import scenes
    sc2, keep = with_geometryless_objects(c2_small.scene)
    say("scene with geometry-less objects", sc2.nobj, "objects")
    ds = ctx.scene(sc2)
    osc = oracle.OracleScene(sc2)
    g, o = ds.bvh(-1), osc.bvh(-1)
    for key in ("nodes", "keys", "indices", "displacement", "table"):
        assert np.array_equal(g[key], o[key]), key
    rays = scenes.camera_rays(c2_small.camera)[::3]
    say("trace_closest with geometry-less objects", rays.shape[0])
    tuv, patch, obj = ds.trace_closest(rays)
    otuv, opatch, oobj, _ = osc.trace(rays)
    assert np.array_equal(obj, oobj) and np.array_equal(patch, opatch)
    assert_bits_equal(tuv, otuv, "t,u,v")
    assert not np.isin(obj, np.flatnonzero(~keep)).any()
    wl = c2_small
    say("render with geometry-less objects")
    color, frame, counts = ds.render(wl.camera, eps=wl.eps, depth=wl.depth)
    ocolor, oframe, _, orc = osc.render(wl.camera, eps=wl.eps, depth=wl.depth)
    assert (counts["primary"], counts["shadow"], counts["secondary"]) == \
        (orc["primary"], orc["shadow"], orc["secondary"])
    assert_bits_equal(color.reshape(-1, 3), ocolor.reshape(-1, 3), "color")


@pytest.mark.parametrize("name", ["c3_small", "grazing"])
def test_repeated_renders_equal_the_first(ctx, oracle, request, name):
    """Each render records its tiles' times and the next render of a frame
    of the same size takes the tiles longest first.  The frame, the colours
    and the ray counts do not depend on that order: renders 2 and 3
    (reordered), renders of another size and of another view of the same size
    in between, and the renders after them all equal the oracle's."""
    import dataclasses
    wl = request.getfixturevalue(name)
    ds = ctx.scene(wl.scene)
    osc = oracle.OracleScene(wl.scene)
    ocolor, oframe, _, orc = osc.render(wl.camera, eps=wl.eps, depth=wl.depth)
    other = dataclasses.replace(wl.camera, width=wl.camera.width // 2 + 3)
    moved = dataclasses.replace(wl.camera, position=wl.camera.position + np.float32(0.25))
    first = {}
    seq = [wl.camera, wl.camera, wl.camera, other, other, wl.camera, moved, wl.camera, wl.camera]
    for k, cam in enumerate(seq):
        say("render", name, "pass", k, cam.width, cam.height)
        color, frame, counts = ds.render(cam, eps=wl.eps, depth=wl.depth)
        if cam is other or cam is moved:
            # a frame whose width is not a multiple of 8 (edge tiles), reordered
            # on its second render: equal to its first render
            if id(cam) in first:
                c0, f0, n0 = first[id(cam)]
                assert n0 == counts and np.array_equal(f0, frame), k
                assert np.array_equal(c0.view(np.uint32), color.view(np.uint32)), k
            first[id(cam)] = (color, frame, counts)
            continue
        assert (counts["primary"], counts["shadow"], counts["secondary"]) == \
            (orc["primary"], orc["shadow"], orc["secondary"]), k
        assert_bits_equal(color.reshape(-1, 3), ocolor.reshape(-1, 3), f"color, pass {k}")
        ch = lambda f: np.stack([(f >> s) & 0xFF for s in (0, 8, 16, 24)], -1).astype(np.int32)
        assert np.abs(ch(frame) - ch(oframe)).max() <= 1, k


@pytest.mark.parametrize("pad", [13, 64])
def test_render_into_padded_frame(ctx, oracle, c1, pad):
    """rt::Frame rows padded to a stride (Frame.h:29-94): pixel (x, y) goes to
    frame[stride*y + x]; the rows equal the dense render's, the padding is
    left as it was."""
    import cgspline_amd as cgs
    wl = c1
    w, h = wl.camera.width, wl.camera.height
    stride = w + pad
    say("render into a padded frame, stride", stride)
    ds = ctx.scene(wl.scene)
    _, dense, _ = ds.render(wl.camera, eps=wl.eps, depth=wl.depth)
    sentinel = np.full(stride * h, 0x5A5A5A5A, np.uint32)
    df = cgs.DeviceArray.from_host(ctx, sentinel)
    cgs.check(ds.render_dev(wl.camera, df, eps=wl.eps, depth=wl.depth, stride=stride), "cgs_render")
    ctx.sync()
    out = df.to_host(np.uint32, (h, stride))
    assert np.array_equal(out[:, :w], dense.reshape(h, w))
    assert np.all(out[:, w:] == 0x5A5A5A5A)
    with pytest.raises(cgs.CgsError, match="stride"):
        cgs.check(ds.render_dev(wl.camera, df, eps=wl.eps, depth=wl.depth, stride=w - 1), "cgs_render")


def test_trace_equals_golden_c1_hits(ctx, c1):
    """The plain CPU workload's committed golden hits
    (tests/golden/c1_heightfield_hits.npz), bit for bit."""
    import os
    import conftest
    from cgspline_amd import scenes
    g = np.load(os.path.join(conftest.ROOT, "tests", "golden", "c1_heightfield_hits.npz"))
    rays = scenes.camera_rays(c1.camera)
    say("golden c1 hits", rays.shape[0], "rays")
    tuv, patch, obj = ctx.scene(c1.scene).trace_closest(rays)
    assert np.array_equal(np.where(obj >= 0, patch, np.uint32(EMPTY)), g["patch"])
    for k, name in enumerate("tuv"):
        hit = obj >= 0
        assert np.array_equal(bits(tuv[hit, k]), g[name][hit].view(np.uint32)), name
