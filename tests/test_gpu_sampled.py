"""The sampled (Monte Carlo) render on the GPU (cgs_render_sampled) against
the oracle's sampled render, bit for bit in the colours and the ray counts:
the flat square of the closed-form pins, the height field, the closed
surfaces, the reference's teaset and Test152 scenes at recursion depth 6,
and rows of the displaced-grid frame, at depth 0 too; padded frames; the
same seed gives the same frame, another seed another; the sample count is
checked."""
from __future__ import annotations

import os

import numpy as np
import pytest

from conftest import say
from test_gpu_parity import assert_bits_equal

pytestmark = pytest.mark.gpu

ASSETS = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "assets")


def _compare(ctx, oracle, scene, cam, samples, seed, eps, depth, rows=None):
    ds = ctx.scene(scene)
    color, frame, counts = ds.render(cam, eps=eps, depth=depth, samples=samples, seed=seed)
    ocolor, oframe, orc = oracle.OracleScene(scene).render_sampled(cam, samples=samples, seed=seed, eps=eps, depth=depth,
                                                                   rows=rows)
    assert counts["overflow"] == 0
    if rows is None:
        assert (counts["primary"], counts["shadow"], counts["secondary"]) == \
            (orc["primary"], orc["shadow"], orc["secondary"])
        r0, r1 = 0, cam.height
    else:
        r0, r1 = rows
    assert_bits_equal(color[r0:r1].reshape(-1, 3), ocolor[r0:r1].reshape(-1, 3), "color")
    ch = lambda f: np.stack([(f >> s) & 0xFF for s in (0, 8, 16, 24)], -1).astype(np.int32)
    assert np.abs(ch(frame[r0:r1]) - ch(oframe[r0:r1])).max() <= 1
    ds.destroy()
    return color, orc


def test_sampled_flat_square_equals_oracle(ctx, oracle):
    from cgspline_amd import scenes
    from test_oracle_render import flat_square_scene
    cam = scenes.make_camera((0.4, -1.8, 2.2), (0.0, 0.1, 0.0), (0.0, 0.0, 1.0), 50.0, 64, 48)
    say("sampled render: flat square 64x48, 6 samples, depth 1")
    _, orc = _compare(ctx, oracle, flat_square_scene(), cam, 6, 42, 1e-4, 1)
    assert orc["secondary"] > 0


def test_sampled_heightfield_equals_oracle(ctx, oracle, c1):
    from cgspline_amd import scenes
    cam = scenes.make_camera(*_view(c1.camera), 96, 72)
    say("sampled render: height field 96x72, 4 samples, depth 3")
    _compare(ctx, oracle, c1.scene, cam, 4, 42, c1.eps, 3)


def test_sampled_closed_surfaces_equal_oracle(ctx, oracle, c2_small):
    from cgspline_amd import scenes
    cam = scenes.make_camera(*_view(c2_small.camera), 80, 45)
    say("sampled render: closed surfaces 80x45, 3 samples, depth 2")
    _compare(ctx, oracle, c2_small.scene, cam, 3, 1234, c2_small.eps, 2)


@pytest.mark.parametrize("name", ["Teaset01", "Test152"])
def test_sampled_reference_scene_equals_oracle(ctx, oracle, name):
    """The reference's scene files at its default recursion depth 6, where
    sampled bounces run through several objects."""
    from cgspline_amd import scenes
    wl = scenes.from_dl(os.path.join(ASSETS, "scenes", name + ".dl"), width=96, height=64, depth=6)
    say("sampled render:", name, "96x64, 2 samples, depth 6")
    _, orc = _compare(ctx, oracle, wl.scene, wl.camera, 2, 42, wl.eps, 6)
    assert orc["secondary"] > 0 and orc["shadow"] > 0


def test_sampled_displaced_grid_rows_equal_oracle(ctx, oracle, c3_small):
    """Rows of the 256² displaced-grid frame (24,576 patches), 2 samples."""
    wl = c3_small
    say("sampled render: displaced grids 256x256 rows 100-116, 2 samples")
    _compare(ctx, oracle, wl.scene, wl.camera, 2, 42, wl.eps, wl.depth, rows=(100, 116))


def test_sampled_surface_cull_tree_equals_oracle(ctx, oracle, c3_small):
    """A scene whose one surface has a cull tree (the displaced grids' 24,576
    patches as one surface): the sampled kernel that carries the wide-hash and
    surface-cull paths."""
    from test_gpu_cull import one_surface
    wl = c3_small
    sc = one_surface(wl.scene)
    ds = ctx.scene(sc)
    assert ds.bvh_culled(0)
    ds.destroy()
    say("sampled render: one surface with a cull tree, rows 96-111, 2 samples")
    _compare(ctx, oracle, sc, wl.camera, 2, 42, wl.eps, wl.depth, rows=(96, 112))


def test_sampled_flat_square_without_bounce_equals_oracle(ctx, oracle):
    """Depth 0: each sample is its hit's direct light without the ambient
    term, or the background; no draws are taken."""
    from cgspline_amd import scenes
    from test_oracle_render import flat_square_scene
    cam = scenes.make_camera((0.4, -1.8, 2.2), (0.0, 0.1, 0.0), (0.0, 0.0, 1.0), 50.0, 48, 40)
    say("sampled render: flat square 48x40, 5 samples, depth 0")
    _, orc = _compare(ctx, oracle, flat_square_scene(), cam, 5, 7, 1e-4, 0)
    assert orc["secondary"] == 0 and orc["primary"] == 48 * 40 * 5


def test_sampled_render_into_padded_frame(ctx, c1):
    """A frame of rows padded to a stride: the rows equal the dense sampled
    render's, the padding is left as it was."""
    import cgspline_amd as cgs
    from cgspline_amd import scenes
    cam = scenes.make_camera(*_view(c1.camera), 64, 40)
    w, h, stride = cam.width, cam.height, cam.width + 13
    ds = ctx.scene(c1.scene)
    _, dense, _ = ds.render(cam, eps=c1.eps, depth=2, samples=3, seed=5)
    df = cgs.DeviceArray.from_host(ctx, np.full(stride * h, 0x5A5A5A5A, np.uint32))
    cgs.check(ds.render_sampled_dev(cam, df, samples=3, seed=5, eps=c1.eps, depth=2, stride=stride),
              "cgs_render_sampled")
    ctx.sync()
    out = df.to_host(np.uint32, (h, stride))
    assert np.array_equal(out[:, :w], dense.reshape(h, w))
    assert np.all(out[:, w:] == 0x5A5A5A5A)
    ds.destroy()


def test_sampled_render_is_seeded(ctx, oracle, c1):
    from cgspline_amd import scenes
    cam = scenes.make_camera(*_view(c1.camera), 64, 48)
    ds = ctx.scene(c1.scene)
    a, fa, ca = ds.render(cam, eps=c1.eps, depth=2, samples=3, seed=9)
    b, fb, cb = ds.render(cam, eps=c1.eps, depth=2, samples=3, seed=9)
    c, _, _ = ds.render(cam, eps=c1.eps, depth=2, samples=3, seed=10)
    assert np.array_equal(a.view(np.uint32), b.view(np.uint32)) and np.array_equal(fa, fb) and ca == cb
    assert not np.array_equal(a, c)
    # one sample looks through the pixel centre: the first segment of the plain render
    one, _, c1c = ds.render(cam, eps=c1.eps, depth=0, samples=1)
    plain, _, pc = ds.render(cam, eps=c1.eps, depth=0)
    assert c1c["primary"] == pc["primary"] == cam.width * cam.height
    ds.destroy()


def test_sample_count_is_checked(ctx, c1):
    import cgspline_amd as cgs
    from cgspline_amd import scenes
    cam = scenes.make_camera(*_view(c1.camera), 8, 8)
    ds = ctx.scene(c1.scene)
    for bad in (0, 65537):
        with pytest.raises(cgs.CgsError, match="samples must lie in"):
            ds.render(cam, samples=bad)
    ds.destroy()


def _view(cam):
    import math
    M = cam.to_world.reshape(4, 4).astype(np.float64)
    pos = M[3, :3]
    return pos, pos - M[2, :3], M[1, :3], math.degrees(cam.fovy)
