"""The alternative GPU schedule gives the default path's results bit for bit
(the default path is checked against the oracle elsewhere): CGS_STAGED=1
selects the staged wavefront tracer (traverse kernel, compacted clip queue,
finish kernel, separate shading passes)."""
from __future__ import annotations

import os

import numpy as np
import pytest

from conftest import say

pytestmark = pytest.mark.gpu

VARIANTS = {"staged": {"CGS_STAGED": "1"}}


def open_with_env(env):
    import cgspline_amd as cgs
    old = {k: os.environ.get(k) for k in env}
    os.environ.update(env)
    try:
        return cgs.Context(0)
    finally:
        for k, v in old.items():
            if v is None:
                del os.environ[k]
            else:
                os.environ[k] = v


@pytest.fixture(scope="module", params=sorted(VARIANTS))
def variant_ctx(request):
    say("cgs_open(0)", request.param)
    c = open_with_env(VARIANTS[request.param])
    yield request.param, c
    c.close()


@pytest.mark.parametrize("name", ["c2_small", "c3_small", "grazing"])
def test_variant_trace_equals_default(ctx, variant_ctx, request, name):
    from cgspline_amd import scenes
    vname, vctx = variant_ctx
    wl = request.getfixturevalue(name)
    say("trace", vname, name)
    rays = scenes.camera_rays(wl.camera)
    a = ctx.scene(wl.scene)
    b = vctx.scene(wl.scene)
    ha, pa, oa = a.trace_closest(rays)
    hb, pb, ob = b.trace_closest(rays)
    assert np.array_equal(pa, pb) and np.array_equal(oa, ob)
    assert np.array_equal(ha.view(np.uint32), hb.view(np.uint32))
    assert np.array_equal(a.trace_any(rays), b.trace_any(rays))


def test_variant_deep_top_level_equals_default(ctx, variant_ctx, stacked):
    """Rays through 180,000 stacked objects (17 top-level siblings postponed
    at once) give the default tracer's hits and occlusion."""
    vname, vctx = variant_ctx
    sc, rays = stacked
    say("trace", vname, "stacked arcs")
    a, b = ctx.scene(sc), vctx.scene(sc)
    ha, pa, oa = a.trace_closest(rays)
    hb, pb, ob = b.trace_closest(rays)
    assert np.array_equal(pa, pb) and np.array_equal(oa, ob)
    assert np.array_equal(ha.view(np.uint32), hb.view(np.uint32))
    assert np.array_equal(a.trace_any(rays), b.trace_any(rays))


@pytest.mark.parametrize("name", ["c1", "c2_small", "c3_small", "grazing"])
def test_variant_render_equals_default(ctx, variant_ctx, request, name):
    vname, vctx = variant_ctx
    wl = request.getfixturevalue(name)
    say("render", vname, name)
    ca, fa, na = ctx.scene(wl.scene).render(wl.camera, eps=wl.eps, depth=wl.depth)
    cb, fb, nb = vctx.scene(wl.scene).render(wl.camera, eps=wl.eps, depth=wl.depth)
    assert na == nb
    assert np.array_equal(ca.view(np.uint32), cb.view(np.uint32)) and np.array_equal(fa, fb)


def test_variant_render_into_padded_frame(ctx, variant_ctx, c1):
    """The variant writes padded frames like the default tracer."""
    import cgspline_amd as cgs
    vname, vctx = variant_ctx
    wl = c1
    w, h = wl.camera.width, wl.camera.height
    stride = w + 7
    say("padded render", vname, stride)
    _, dense, _ = ctx.scene(wl.scene).render(wl.camera, eps=wl.eps, depth=wl.depth)
    ds = vctx.scene(wl.scene)
    df = cgs.DeviceArray.from_host(vctx, np.full(stride * h, 0x5A5A5A5A, np.uint32))
    cgs.check(ds.render_dev(wl.camera, df, eps=wl.eps, depth=wl.depth, stride=stride), "cgs_render")
    vctx.sync()
    out = df.to_host(np.uint32, (h, stride))
    assert np.array_equal(out[:, :w], dense.reshape(h, w)) and np.all(out[:, w:] == 0x5A5A5A5A)
