"""The staged wavefront tracer (CGS_STAGED=1: traverse kernel, compacted
clip queue, finish kernel) gives the default persistent path's results bit
for bit; the persistent path is checked against the oracle elsewhere."""
from __future__ import annotations

import os

import numpy as np
import pytest

from conftest import say

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def staged_ctx():
    import cgspline_amd as cgs
    say("cgs_open(0) staged")
    old = os.environ.get("CGS_STAGED")
    os.environ["CGS_STAGED"] = "1"
    try:
        c = cgs.Context(0)
    finally:
        if old is None:
            del os.environ["CGS_STAGED"]
        else:
            os.environ["CGS_STAGED"] = old
    yield c
    c.close()


@pytest.mark.parametrize("name", ["c2_small", "c3_small"])
def test_staged_trace_equals_persistent(ctx, staged_ctx, request, name):
    from cgspline_amd import scenes
    wl = request.getfixturevalue(name)
    say("staged trace", name)
    rays = scenes.camera_rays(wl.camera)
    a = ctx.scene(wl.scene)
    b = staged_ctx.scene(wl.scene)
    ha, pa, oa = a.trace_closest(rays)
    hb, pb, ob = b.trace_closest(rays)
    assert np.array_equal(pa, pb) and np.array_equal(oa, ob)
    assert np.array_equal(ha.view(np.uint32), hb.view(np.uint32))
    assert np.array_equal(a.trace_any(rays), b.trace_any(rays))


@pytest.mark.parametrize("name", ["c1", "c3_small"])
def test_staged_render_equals_persistent(ctx, staged_ctx, request, name):
    wl = request.getfixturevalue(name)
    say("staged render", name)
    ca, fa, na = ctx.scene(wl.scene).render(wl.camera, eps=wl.eps, depth=wl.depth)
    cb, fb, nb = staged_ctx.scene(wl.scene).render(wl.camera, eps=wl.eps, depth=wl.depth)
    assert na == nb
    assert np.array_equal(ca.view(np.uint32), cb.view(np.uint32)) and np.array_equal(fa, fb)
