"""GPU parity of the opt-in leaf methods of cgs_trace_closest_ex /
cgs_trace_any_ex against the oracle run with the same options: subdivision
at the reference's threshold and at other depths (the commented-out leaf
call, Primitive.cu:236,263), a coarser clipping tolerance, a cap on clipping
passes, and the Newton polish of closest hits.  Every result equals the
oracle's bit for bit; the default record equals the plain entry points."""
from __future__ import annotations

import numpy as np
import pytest

from conftest import say
from test_gpu_parity import assert_bits_equal

pytestmark = pytest.mark.gpu

CLOSEST_MODES = [
    dict(method=1),                                   # subdivision, threshold 2^-4
    dict(method=1, subdiv_depth=8),
    dict(method=1, subdiv_depth=12),
    dict(clip_tol=2.0 ** -8),
    dict(clip_tol=2.0 ** -6),
    dict(max_clip_iters=5),
    dict(polish=True),
    dict(method=1, subdiv_depth=10, polish=True),
    dict(max_clip_iters=6, polish=True),
]
ANY_MODES = [dict(method=1), dict(method=1, subdiv_depth=11), dict(max_clip_iters=5), dict(clip_tol=2.0 ** -6)]


def _workload(request, name, step):
    from cgspline_amd import scenes
    wl = request.getfixturevalue(name)
    return wl.scene, scenes.camera_rays(wl.camera)[::step]


@pytest.mark.parametrize("name,step", [("c2_small", 5), ("c3_small", 7)])
def test_closest_modes_bitexact(request, ctx, oracle, name, step):
    sc, rays = _workload(request, name, step)
    ds = ctx.scene(sc)
    osc = oracle.OracleScene(sc)
    for opts in CLOSEST_MODES:
        say("trace_closest_ex", name, rays.shape[0], "rays", opts)
        g_tuv, g_p, g_o = ds.trace_closest(rays, **opts)
        o_tuv, o_p, o_o, _ = osc.trace(rays, **opts)
        assert np.array_equal(g_o, o_o), f"{opts}: objects differ at {np.flatnonzero(g_o != o_o)[:5]}"
        assert np.array_equal(g_p, o_p), f"{opts}: patches differ at {np.flatnonzero(g_p != o_p)[:5]}"
        assert_bits_equal(g_tuv, o_tuv, f"{name} {opts}")
        assert (g_o >= 0).sum() > 50


@pytest.mark.parametrize("name,step", [("c2_small", 5), ("c3_small", 7)])
def test_any_modes_bitexact(request, ctx, oracle, name, step):
    sc, rays = _workload(request, name, step)
    ds = ctx.scene(sc)
    osc = oracle.OracleScene(sc)
    for opts in ANY_MODES:
        say("trace_any_ex", name, rays.shape[0], "rays", opts)
        g = ds.trace_any(rays, **opts)
        o, _ = osc.trace(rays, any_hit=True, **opts)
        assert not np.any(g == 2), opts
        assert np.array_equal(g == 1, o), f"{opts}: occlusion differs at {np.flatnonzero((g == 1) != o)[:5]}"
        assert o.sum() > 50


def test_default_record_is_the_plain_call(ctx, c2_small):
    """A zeroed cgs_trace_options and a null pointer give the plain entry
    points' results."""
    import ctypes as C
    import cgspline_amd as cgs
    from cgspline_amd import scenes
    ds = ctx.scene(c2_small.scene)
    rays = scenes.camera_rays(c2_small.camera)[::9]
    n = rays.shape[0]
    plain = ds.trace_closest(rays)
    dr = cgs.DeviceArray.from_host(ctx, rays)
    for opts in (C.byref(cgs.trace_options()), None):
        dh = cgs.DeviceArray(ctx, 16 * n)
        do = cgs.DeviceArray(ctx, 4 * n)
        say("cgs_trace_closest_ex with", "zeroed record" if opts is not None else "null")
        cgs.check(cgs.lib().cgs_trace_closest_ex(ctx.h, ds.h, dr.dptr, n, dh.dptr, do.dptr, opts, None))
        ctx.sync()
        h = dh.to_host(np.float32, (n, 4))
        assert np.array_equal(h[:, :3].view(np.uint32), plain[0].view(np.uint32))
        assert np.array_equal(h.view(np.uint32)[:, 3], plain[1])
        assert np.array_equal(do.to_host(np.int32, (n,)), plain[2])
    assert np.array_equal(ds.trace_any(rays, method=0), ds.trace_any(rays))


def test_invalid_options_are_refused(ctx, c2_small):
    """A subdivision depth above 16, a negative or non-finite tolerance and
    an unknown method are CGS_ERR_INVALID, with nothing launched."""
    import ctypes as C
    import cgspline_amd as cgs
    from cgspline_amd import scenes
    ds = ctx.scene(c2_small.scene)
    rays = scenes.camera_rays(c2_small.camera)[:64]
    for bad in (dict(subdiv_depth=17), dict(clip_tol=-1.0), dict(clip_tol=float("inf")),
                dict(clip_tol=float("nan"))):
        say("refused options", bad)
        with pytest.raises(cgs.CgsError, match="INVALID"):
            ds.trace_closest(rays, **bad)
        with pytest.raises(cgs.CgsError, match="INVALID"):
            ds.trace_any(rays, **bad)
    with pytest.raises(ValueError):
        cgs.trace_options(method=2)
    raw = cgs.trace_options()
    raw.method = 2
    dr = cgs.DeviceArray.from_host(ctx, rays)
    dh = cgs.DeviceArray(ctx, 16 * 64)
    st = cgs.lib().cgs_trace_closest_ex(ctx.h, ds.h, dr.dptr, 64, dh.dptr, None, C.byref(raw), None)
    assert st == 1 and b"unknown method" in cgs.lib().cgs_last_error()


def test_rational_patch_polish_bitexact(ctx, oracle):
    """The polish on a rational patch (weights 0.5..2, where the quotient
    rule matters) equals the oracle's bit for bit and lies within 1e-6 of the
    FP64 root."""
    from test_oracle_polish import BAR, rational_case, root
    sc, H, rays, seed_uvt = rational_case()
    ds = ctx.scene(sc)
    say("rational patch, polish", rays.shape[0], "rays")
    g_tuv, g_p, g_o = ds.trace_closest(rays, polish=True)
    o_tuv, o_p, o_o, _ = oracle.OracleScene(sc).trace(rays, polish=True)
    assert np.array_equal(g_o, o_o) and np.array_equal(g_p, o_p)
    assert_bits_equal(g_tuv, o_tuv, "rational polish")
    rx = root(H, rays, seed_uvt)
    hit = g_o >= 0
    assert hit.sum() >= 0.98 * rays.shape[0]
    assert np.abs(g_tuv[hit, 1:3] - rx[hit, 0:2]).max() <= BAR
