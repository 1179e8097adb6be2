"""GPU parity on the reference's own models (tests/golden/assets: teapot,
teacup, teaspoon, longwave, read with cgs_load_be) in the Teaset01.dl
arrangement: closest hits, shadow rays and the rendered frame against the
oracle, bit for bit (±1 in an 8-bit frame channel, from powf in the sRGB
encode).  The teapot's lid and bottom patches collapse a row of control
points to a point; the spoon is a thin, strongly curved surface."""
from __future__ import annotations

import os

import numpy as np
import pytest

from conftest import say
from test_gpu_parity import assert_bits_equal

pytestmark = pytest.mark.gpu

ASSETS = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "assets")


@pytest.fixture(scope="module")
def teaset():
    from cgspline_amd import scenes
    return scenes.teaset(ASSETS, width=320, height=240, depth=2)


def test_teaset_bvh_and_hash_match_oracle(ctx, oracle, teaset):
    sc = teaset.scene
    say("teaset scene create", sc.npatch, "patches")
    ds = ctx.scene(sc)
    osc = oracle.OracleScene(sc)
    for s in [0, 1, 2, 3, -1]:
        g, o = ds.bvh(s), osc.bvh(s)
        for key in ("nodes", "keys", "indices", "displacement", "table"):
            assert np.array_equal(g[key], o[key]), (s, key)


def test_teaset_trace_closest_and_any_bitexact(ctx, oracle, teaset):
    from cgspline_amd import scenes
    sc = teaset.scene
    rays = scenes.camera_rays(teaset.camera)
    say("teaset trace_closest", rays.shape[0], "rays")
    ds = ctx.scene(sc)
    tuv, patch, obj = ds.trace_closest(rays)
    osc = oracle.OracleScene(sc)
    otuv, opatch, oobj, _ = osc.trace(rays)
    assert np.array_equal(obj, oobj) and np.array_equal(patch, opatch)
    assert_bits_equal(tuv, otuv, "t,u,v")
    assert set(np.unique(obj)) == {-1, 0, 1, 2, 3}
    # shadow rays toward the light from every hit
    hit = obj >= 0
    P = rays[hit, 0:3] + tuv[hit, 0:1] * rays[hit, 4:7]
    L = sc.lights[0, 4:7][None, :] - P
    d = np.linalg.norm(L, axis=1).astype(np.float32)
    L = (L / d[:, None]).astype(np.float32)
    sh = np.zeros((P.shape[0], 8), np.float32)
    sh[:, 0:3], sh[:, 3], sh[:, 4:7] = P + 1e-4 * L, d, L
    say("teaset trace_any", sh.shape[0], "shadow rays")
    g = ds.trace_any(sh)
    o, _ = osc.trace(sh, any_hit=True)
    assert not np.any(g == 2)
    assert np.array_equal(g == 1, o)
    assert 0 < o.sum() < o.size


def test_teaset_render_matches_oracle(ctx, oracle, teaset):
    wl = teaset
    say("teaset render", wl.camera.width, wl.camera.height, "depth", wl.depth)
    ds = ctx.scene(wl.scene)
    color, frame, counts = ds.render(wl.camera, eps=wl.eps, depth=wl.depth)
    ocolor, oframe, _, orc = oracle.OracleScene(wl.scene).render(wl.camera, eps=wl.eps, depth=wl.depth)
    assert counts["overflow"] == 0
    assert (counts["primary"], counts["shadow"], counts["secondary"]) == \
        (orc["primary"], orc["shadow"], orc["secondary"])
    assert orc["secondary"] > 0
    assert_bits_equal(color.reshape(-1, 3), ocolor.reshape(-1, 3), "color")
    ch = lambda f: np.stack([(f >> s) & 0xFF for s in (0, 8, 16, 24)], -1).astype(np.int32)
    assert np.abs(ch(frame) - ch(oframe)).max() <= 1


@pytest.mark.parametrize("name", ["Teaset01", "Teaset02", "teaset", "Test32", "Test56", "Test80", "Test104",
                                  "Test128", "Test152"])
def test_reference_scene_file_render_matches_oracle(ctx, oracle, name):
    """The reference's own scene files (read with cgs_dl_read): Test152 is
    152 objects, 51 point lights; the frame, colours and ray counts equal
    the oracle's."""
    from cgspline_amd import scenes
    wl = scenes.from_dl(os.path.join(ASSETS, "scenes", name + ".dl"), width=160, height=120, depth=2)
    say("scene file render", name, wl.scene.nobj, "objects", wl.scene.lights.shape[0], "lights")
    ds = ctx.scene(wl.scene)
    color, frame, counts = ds.render(wl.camera, eps=wl.eps, depth=wl.depth)
    ocolor, oframe, _, orc = oracle.OracleScene(wl.scene).render(wl.camera, eps=wl.eps, depth=wl.depth)
    assert counts["overflow"] == 0
    assert (counts["primary"], counts["shadow"], counts["secondary"]) == \
        (orc["primary"], orc["shadow"], orc["secondary"])
    assert orc["shadow"] > 0 and orc["secondary"] > 0
    assert_bits_equal(color.reshape(-1, 3), ocolor.reshape(-1, 3), "color")
    ch = lambda f: np.stack([(f >> s) & 0xFF for s in (0, 8, 16, 24)], -1).astype(np.int32)
    assert np.abs(ch(frame) - ch(oframe)).max() <= 1


def test_test152_benchmark_frame_matches_oracle(ctx, oracle):
    """The Test152 benchmark frame itself: 1920×1080, recursion depth 6, 91.6 M
    rays.  Rendered twice (the second render takes the tiles longest first),
    both equal the oracle's frame, colours and ray counts."""
    from cgspline_amd import scenes
    wl = scenes.from_dl(os.path.join(ASSETS, "scenes", "Test152.dl"), width=1920, height=1080, depth=6)
    ds = ctx.scene(wl.scene)
    say("oracle render of Test152 at 1920x1080, depth 6")
    ocolor, oframe, _, orc = oracle.OracleScene(wl.scene).render(wl.camera, eps=wl.eps, depth=wl.depth)
    ch = lambda f: np.stack([(f >> s) & 0xFF for s in (0, 8, 16, 24)], -1).astype(np.int32)
    for k in range(2):
        say("render Test152 at 1920x1080, depth 6, pass", k)
        color, frame, counts = ds.render(wl.camera, eps=wl.eps, depth=wl.depth)
        assert counts["overflow"] == 0
        assert (counts["primary"], counts["shadow"], counts["secondary"]) == \
            (orc["primary"], orc["shadow"], orc["secondary"])
        assert_bits_equal(color.reshape(-1, 3), ocolor.reshape(-1, 3), f"color, pass {k}")
        assert np.abs(ch(frame) - ch(oframe)).max() <= 1


@pytest.mark.parametrize("name", ["tube", "dice16"])
def test_se_surface_scene_render_matches_oracle(ctx, oracle, name):
    """A `.se` surface read through a scene file, as the reference's renderer
    receives it (its grouped index list as patches of 16 consecutive
    indices), lit by one point light: BVH, hits and frame equal the
    oracle's."""
    import cgspline_amd as cgs
    from cgspline_amd import scenes
    se_dir = os.path.join(ASSETS, "surfaces")
    text = f'Surface {{ name:"s" file:"{name}.se" }}\nScene {{ objects:[ {{ surface:{{ model:Get("s") }} }} ] }}'
    d = cgs.read_dl(text.encode(), name="se.dl", working_dir=se_dir)
    v, ix = d.scene.vertices, d.scene.indices
    c = v[:, :3].astype(np.float64).mean(axis=0)
    r = float(np.linalg.norm(v[:, :3] - c, axis=1).max())
    mat = np.array([0.7, 0.6, 0.5, 0.04, 0.04, 0.04, 0.3, 0.4], np.float32)
    sc = scenes.assemble([(v, ix)], [(0, scenes.IDENTITY16, scenes.IDENTITY16, mat)],
                         [scenes.point_light(c + 3 * r * np.array([0.3, -0.5, 0.8]), (8.0, 8.0, 8.0))], name=name)
    cam = scenes.make_camera(c + 2.5 * r * np.array([0.4, -0.8, 0.45]), c, (0.0, 0.0, 1.0), 50.0, 96, 72)
    say("se surface scene render", name, ix.shape[0], "patches")
    ds = ctx.scene(sc)
    color, frame, counts = ds.render(cam, eps=1e-4, depth=2)
    ocolor, oframe, _, orc = oracle.OracleScene(sc).render(cam, eps=1e-4, depth=2)
    assert counts["overflow"] == 0
    assert (counts["primary"], counts["shadow"], counts["secondary"]) == (orc["primary"], orc["shadow"], orc["secondary"])
    assert_bits_equal(color.reshape(-1, 3), ocolor.reshape(-1, 3), "color")
    assert orc["shadow"] > 0
    ds.destroy()
