"""Concurrency at the boundary: calls of one context issued on different
streams, scenes shared between contexts, and the host-side size guards of
the persistent launches.  Every result is checked against the oracle."""
from __future__ import annotations

import ctypes as C

import numpy as np
import pytest

from conftest import say

pytestmark = pytest.mark.gpu


def bits(a):
    return np.ascontiguousarray(a, np.float32).view(np.uint32)


@pytest.fixture(scope="module")
def c2_rays(c2_small):
    from cgspline_amd import scenes
    return scenes.camera_rays(c2_small.camera)


def test_one_context_two_streams_equal_oracle(ctx, oracle, c2_small, c2_rays):
    """Closest-hit, any-hit and a render of one context, issued on two
    streams without waiting in between, equal the oracle: the context's
    region stacks are never used by two launches at once."""
    import cgspline_amd as cgs
    sc = ctx.scene(c2_small.scene)
    rays = c2_rays
    n = rays.shape[0]
    half = n // 2
    say("two streams", n, "rays")
    s1, s2 = ctx.create_stream(), ctx.create_stream()
    try:
        dr = cgs.DeviceArray.from_host(ctx, rays)
        dh = cgs.DeviceArray(ctx, 16 * n)
        do = cgs.DeviceArray(ctx, 4 * n)
        docc = cgs.DeviceArray(ctx, 4 * n)
        w, h = c2_small.camera.width, c2_small.camera.height
        df = cgs.DeviceArray(ctx, 4 * w * h)
        dc = cgs.DeviceArray(ctx, 12 * w * h)
        for rep in range(3):
            sc.trace_closest_dev(dr.dptr, half, dh.dptr, do.dptr, stream=s1)
            sc.trace_closest_dev(dr.dptr + 32 * half, n - half, dh.dptr + 16 * half, do.dptr + 4 * half, stream=s2)
            sc.trace_any_dev(dr.dptr, n, docc.dptr, stream=s1)
            cgs.check(sc.render_dev(c2_small.camera, df, dc, eps=c2_small.eps, depth=c2_small.depth, stream=s2),
                      "cgs_render")
            ctx.sync(s1)
            ctx.sync(s2)
        hits = dh.to_host(np.float32, (n, 4))
        obj = do.to_host(np.int32, (n,))
        occ = docc.to_host(np.uint32, (n,))
        color = dc.to_host(np.float32, (h, w, 3))
    finally:
        ctx.destroy_stream(s1)
        ctx.destroy_stream(s2)
    osc = oracle.OracleScene(c2_small.scene)
    ohits, opatch, oobj, _ = osc.trace(rays)
    assert np.array_equal(hits.view(np.uint32)[:, 3], opatch) and np.array_equal(obj, oobj)
    assert np.array_equal(bits(hits[:, :3]), bits(ohits))
    oocc, _ = osc.trace(rays, any_hit=True)
    assert np.array_equal(occ.astype(bool), oocc) and occ.max() <= 1
    ocolor, _, _, _ = osc.render(c2_small.camera, eps=c2_small.eps, depth=c2_small.depth)
    assert np.array_equal(bits(color), bits(ocolor))


def test_scene_shared_between_contexts(ctx, oracle, c2_small, c2_rays):
    """A scene made resident by one context is traced and rendered through
    two other contexts at the same time, each on its own stream."""
    import cgspline_amd as cgs
    say("shared scene, three contexts")
    owner = cgs.Context(0)
    try:
        sc = owner.scene(c2_small.scene)
        a, b = sc.via(ctx), sc.via(cgs.Context(0))
        rays = c2_rays
        n = rays.shape[0]
        sa, sb = a.ctx.create_stream(), b.ctx.create_stream()
        try:
            outs = []
            for v, s in ((a, sa), (b, sb)):
                dr = cgs.DeviceArray.from_host(v.ctx, rays)
                dh = cgs.DeviceArray(v.ctx, 16 * n)
                do = cgs.DeviceArray(v.ctx, 4 * n)
                v.trace_closest_dev(dr.dptr, n, dh.dptr, do.dptr, stream=s)
                outs.append((dr, dh, do, v.ctx, s))
            res = []
            for dr, dh, do, c, s in outs:
                c.sync(s)
                res.append((dh.to_host(np.float32, (n, 4)), do.to_host(np.int32, (n,))))
            ca, _, _ = a.render(c2_small.camera, eps=c2_small.eps, depth=c2_small.depth)
            cb, _, _ = b.render(c2_small.camera, eps=c2_small.eps, depth=c2_small.depth)
        finally:
            a.ctx.destroy_stream(sa)
            b.ctx.destroy_stream(sb)
        osc = oracle.OracleScene(c2_small.scene)
        ohits, opatch, oobj, _ = osc.trace(rays)
        for hits, obj in res:
            assert np.array_equal(hits.view(np.uint32)[:, 3], opatch) and np.array_equal(obj, oobj)
            assert np.array_equal(bits(hits[:, :3]), bits(ohits))
        ocolor, _, _, _ = osc.render(c2_small.camera, eps=c2_small.eps, depth=c2_small.depth)
        assert np.array_equal(bits(ca), bits(ocolor)) and np.array_equal(bits(cb), bits(ocolor))
        b.ctx.close()
        sc.destroy()
    finally:
        owner.close()


def test_sizes_past_the_work_counter_are_refused(ctx, c1):
    """Launches whose 32-bit work counter would wrap are refused on the host
    before anything runs: a 65535×65535 render, a trace and a clip of 2^32-1
    items, a batch of that many rays and an oversized serving request."""
    import cgspline_amd as cgs
    from cgspline_amd import _abi, serving
    say("size guards")
    sc = ctx.scene(c1.scene)
    d = cgs.DeviceArray(ctx, 4096)
    cam = _abi.Camera((C.c_float * 16)(*c1.camera.to_world), (C.c_float * 3)(*c1.camera.position), c1.camera.fovy, 0)
    op = _abi.RenderOptions(1e-4, 0, 0)
    lib = cgs.lib()
    assert lib.cgs_render(ctx.h, sc.h, C.byref(cam), C.byref(op), 65535, 65535, d.dptr, None, None, None) == 1
    big = 0xFFFFFFFF
    assert lib.cgs_trace_closest(ctx.h, sc.h, d.dptr, big, d.dptr, None, None) == 1
    assert lib.cgs_trace_any(ctx.h, sc.h, d.dptr, big, d.dptr, None) == 1
    assert lib.cgs_bezier_clip(ctx.h, d.dptr, d.dptr, big, 2.0 ** -12, d.dptr, None) == 1
    jobs = ctx.batch_jobs([dict(scene=sc, rays=d.dptr, n=big, hits=d.dptr)])
    assert lib.cgs_trace_batch(ctx.h, jobs, 1, None) == 1
    srv = serving.Server(ctx, [sc], hbm_budget=1 << 62)
    try:
        rid = C.c_uint64()
        hbuf = np.zeros(4, np.float32)
        assert lib.cgs_server_submit(srv.h, 0, hbuf.ctypes.data, big, 0, hbuf.ctypes.data, None, None,
                                     C.byref(rid)) == 1
        # a request that fits is still served, and a collected id is forgotten
        rays = np.zeros((3, 8), np.float32)
        rays[:, 3] = np.inf
        rays[:, 6] = -1.0
        rid = srv.submit(0, rays)
        srv.wait(rid)
        assert lib.cgs_server_wait(srv.h, rid, None) == 1
        assert lib.cgs_server_poll(srv.h, rid, None) == -1
    finally:
        srv.close()
    # nothing ran: the context still traces normally
    h, p, o = sc.trace_closest(np.array([[0.5, 0.5, 5.0, np.inf, 0.0, 0.0, -1.0, 0.0]], np.float32))
    assert p.shape == (1,)
