"""Error paths of the boundary on the GPU, each forced from the host:

* CGS_ERR_BUSY: a second cgs_render while the context's previous render is
  still in flight (RayTracer.cu:109-115: the reference warns and does
  nothing).
* Clip-stack overflow: a ray lying in the plane of a large planar patch
  projects the patch onto a segment through the origin, so every region
  keeps the origin and the host algorithm splits without end; the GPU's
  32-entry region stack overflows after 33 iterations and the result is
  reported as CGS_PATCH_OVERFLOW (kClipBatch, trace closest), 2 (trace any)
  and in the render's overflow count.  The oracle with the same cap finds
  the same overflow at the same iteration; slightly tilted rays do not
  overflow and equal the oracle bit for bit."""
from __future__ import annotations

import ctypes as C

import numpy as np
import pytest

from conftest import say

pytestmark = pytest.mark.gpu

OVERFLOW = 0xFFFFFFFE


def test_second_render_while_busy_is_refused(ctx, oracle, grazing):
    """A render of the grazing height field (its rays cost many clip
    iterations) is enqueued; the next cgs_render of the context, issued
    microseconds later, finds it running and is refused with CGS_ERR_BUSY.
    Should the machine ever finish the first render before the second call,
    the frame is enlarged and the pair issued again."""
    import dataclasses
    import cgspline_amd as cgs
    ds = ctx.scene(grazing.scene)
    lib = cgs.lib()
    s = ctx.create_stream()
    busy_seen = False
    try:
        for size in (512, 1024, 2048):
            cam = dataclasses.replace(grazing.camera, width=size, height=size)
            frame = cgs.DeviceArray(ctx, 4 * size * size)
            say("busy check: render", size, "x", size, "then a second render")
            assert ds.render_dev(cam, frame, None, None, eps=grazing.eps, depth=grazing.depth, stream=s) == 0
            running = lib.cgs_render_running(ctx.h)
            second = ds.render_dev(cam, frame, None, None, eps=grazing.eps, depth=grazing.depth, stream=s)
            ctx.sync(s)
            assert lib.cgs_render_running(ctx.h) == 0
            if second == 6:
                assert b"not ready" in lib.cgs_last_error()
                busy_seen = True
                break
            assert second == 0 and running == 0, (second, running)
            ctx.sync(s)
        assert busy_seen
        # the refused call enqueued nothing: the frame is the first render's
        w = h = size
        got = frame.to_host(np.uint32, (h, w))
        cam_small = dataclasses.replace(grazing.camera, width=size, height=size)
        _, want, _, _ = oracle.OracleScene(grazing.scene).render(cam_small, eps=grazing.eps, depth=grazing.depth,
                                                                 rows=(0, 64))
        ch = lambda fr: np.stack([(fr >> k) & 0xFF for k in (0, 8, 16, 24)], -1).astype(np.int32)
        assert np.abs(ch(got[:64]) - ch(want[:64])).max() <= 1
        # and the next render goes through
        assert ds.render_dev(cam_small, frame, None, None, eps=grazing.eps, depth=grazing.depth, stream=s) == 0
        ctx.sync(s)
    finally:
        ctx.destroy_stream(s)


def planar_patch(scale):
    P = np.zeros((16, 4), np.float32)
    for j in range(4):
        for i in range(4):
            P[4 * j + i] = (scale * i / 3, scale * j / 3, 0.0, 1.0)
    return P


def in_plane_rays(scale, n, tilt, seed=7):
    r = np.random.default_rng(seed)
    rays = np.zeros((n, 8), np.float32)
    rays[:, 0] = -0.5 * scale
    rays[:, 1] = r.uniform(0.1, 0.9, n) * scale
    rays[:, 3] = np.inf
    d = np.stack([np.ones(n), r.uniform(-0.2, 0.2, n), np.full(n, tilt)], 1)
    rays[:, 4:7] = d / np.linalg.norm(d, axis=1, keepdims=True)
    rays[:, 2] = -rays[:, 6] / rays[:, 4] * scale     # tilted rays cross the plane at x = scale / 2
    return rays


def test_clip_stack_overflow_reported(ctx, oracle):
    scale = 16.0
    P = planar_patch(scale)
    flat = in_plane_rays(scale, 64, 0.0)
    tilted = in_plane_rays(scale, 64, 0.05, seed=8)
    rays = np.concatenate([flat, tilted])
    pats = np.repeat(P[None], rays.shape[0], 0)
    oracle.set_clip_stack_cap(32)
    try:
        out, work = oracle.clip_batch(flat, pats[:64])
        assert work["overflows"] == 64 and work["clipIters"] == 64 * 33
        oout, owork = oracle.clip_batch(tilted, pats[64:])
        assert owork["overflows"] == 0 and (oout[:, 3] == 1).mean() > 0.9
    finally:
        oracle.set_clip_stack_cap(0)
    say("kClipBatch: in-plane rays overflow")
    g = ctx.bezier_clip(rays, pats)
    gp = g.view(np.uint32)[:, 3]
    assert (gp[:64] == OVERFLOW).all()
    found = oout[:, 3] == 1
    assert np.array_equal(gp[64:] == 1, found) and not (gp[64:] == OVERFLOW).any()
    assert np.array_equal(g[64:][found, :3].view(np.uint32), oout[found, :3].view(np.uint32))


def test_scene_overflow_reported(ctx, oracle):
    """The same through a one-patch scene: closest-hit patch id
    CGS_PATCH_OVERFLOW, occlusion 2, and the render's overflow counter."""
    from cgspline_amd import scenes
    scale = 16.0
    P = planar_patch(scale)
    sc = scenes.assemble([(P, np.arange(16, dtype=np.uint32)[None])],
                         [(0, scenes.IDENTITY16, scenes.IDENTITY16, np.array([0.5] * 6 + [0.0, 0.5], np.float32))],
                         [scenes.point_light((8.0, 8.0, 20.0), (1.0, 1.0, 1.0))])
    ds = ctx.scene(sc)
    flat = in_plane_rays(scale, 64, 0.0)
    tilted = in_plane_rays(scale, 64, 0.05, seed=8)
    say("scene trace: in-plane rays overflow")
    tuv, patch, obj = ds.trace_closest(np.concatenate([flat, tilted]))
    assert (patch[:64] == OVERFLOW).all()
    osc = oracle.OracleScene(sc)
    oh, op, oo, _ = osc.trace(tilted)
    assert np.array_equal(obj[64:], oo) and np.array_equal(patch[64:][oo >= 0], op[oo >= 0])
    assert np.array_equal(tuv[64:].view(np.uint32), oh.view(np.uint32))
    occ = ds.trace_any(np.concatenate([flat, tilted]))
    oocc, _ = osc.trace(tilted, any_hit=True)
    assert (occ[:64] == 2).all() and np.array_equal(occ[64:] == 1, oocc) and (occ[64:] <= 1).all()
