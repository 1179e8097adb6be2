"""The wide hash family on the GPU (cgs_phf_build_wide / _hash_wide /
_fill_table_wide) against the oracle's: parameters, displacements, hashes
and tables bit for bit, on 64-bit keys and on the 2,097,151 heap keys of one
BVH over 2^20 patches; and the scene fallback: the displaced grids as one
BVH over 1,048,576 patches, whose reference-family hash fails, hashed with
the wide family, its BVH, table and hits equal to the oracle's."""
from __future__ import annotations

import numpy as np
import pytest

from conftest import say
from test_gpu_parity import assert_bits_equal
from test_oracle_bvh import heap_keys

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("kind", ["heap_2e20", "lifted_64bit", "small"])
def test_wide_build_equals_oracle(ctx, oracle, kind):
    if kind == "heap_2e20":
        keys = heap_keys(1 << 20).astype(np.uint64)
    elif kind == "lifted_64bit":
        keys = heap_keys(1 << 17).astype(np.uint64) + (np.uint64(0x3C3) << np.uint64(37))
    else:
        keys = heap_keys(37).astype(np.uint64)
    say("cgs_phf_build_wide", kind, keys.shape[0], "keys")
    st, p, D = ctx.phf_build_wide(keys)
    ost, op, oD = oracle.phf_build_wide(keys)
    assert st == ost == 0
    assert p == op and np.array_equal(D, oD)
    h = ctx.phf_hash_wide(p, D, keys)
    assert np.array_equal(h, oracle.phf_hash_wide(op, oD, keys))
    assert np.unique(h).shape[0] == keys.shape[0]
    t = ctx.phf_fill_table_wide(p, D, keys)
    want = np.full(p["table_size"], 0xFFFFFFFF, np.uint32)
    want[h] = np.arange(keys.shape[0], dtype=np.uint32)
    assert np.array_equal(t, want)


def test_wide_bucket_overflow_reported(ctx, oracle):
    keys = np.array([7 + (k << 33) for k in range(5)], np.uint64)
    st, _, _ = ctx.phf_build_wide(keys)
    ost, _, _ = oracle.phf_build_wide(keys)
    assert ost == 2 and st == 5     # CGS_ERR_BUCKET_OVERFLOW


@pytest.fixture(scope="module")
def single_bvh():
    from cgspline_amd import scenes
    return scenes.displaced_grids_single()


def test_single_bvh_scene_hashes_wide_and_matches_oracle(ctx, oracle, single_bvh):
    from cgspline_amd import scenes
    wl = single_bvh
    say("single-BVH scene", wl.scene.npatch, "patches")
    ds = ctx.scene(wl.scene)
    osc = oracle.OracleScene(wl.scene)
    assert ds.bvh_wide(0) and not ds.bvh_wide(-1)
    g, o = ds.bvh(0), osc.bvh(0)
    assert g["wide"] and o["wide"] and g["keys"].shape[0] == 2_097_151
    for key in ("nodes", "keys", "indices", "displacement", "table", "A", "B", "mask", "table_size", "attempts"):
        assert np.array_equal(g[key], o[key]), key
    rays = scenes.camera_rays(wl.camera)[::997]
    say("single-BVH trace_closest", rays.shape[0], "rays")
    tuv, patch, obj = ds.trace_closest(rays)
    otuv, opatch, oobj, work = osc.trace(rays)
    assert np.array_equal(obj, oobj) and np.array_equal(patch, opatch)
    assert_bits_equal(tuv, otuv, "t,u,v")
    assert (obj >= 0).sum() > 300 and work["hashProbes"] > 0
    occ = ds.trace_any(rays)
    oocc, _ = osc.trace(rays, any_hit=True)
    assert np.array_equal(occ == 1, oocc)
    ds.destroy()


def test_single_bvh_render_matches_oracle(ctx, oracle, single_bvh):
    """A 40×30 render of the single-BVH scene (shadow rays and a reflection
    bounce, through the surface's cull tree) equals the oracle's."""
    from cgspline_amd import scenes
    wl = single_bvh
    cam = scenes.make_camera(*_view(wl.camera), 40, 30)
    ds = ctx.scene(wl.scene)
    assert ds.bvh_culled(0)
    say("single-BVH render 40x30")
    color, frame, counts = ds.render(cam, eps=wl.eps, depth=wl.depth)
    ocolor, oframe, _, orc = oracle.OracleScene(wl.scene).render(cam, eps=wl.eps, depth=wl.depth)
    assert counts["overflow"] == 0
    assert (counts["primary"], counts["shadow"], counts["secondary"]) == (orc["primary"], orc["shadow"], orc["secondary"])
    assert_bits_equal(color.reshape(-1, 3), ocolor.reshape(-1, 3), "color")
    assert orc["shadow"] > 0
    ds.destroy()


def _view(cam):
    import math
    M = cam.to_world.reshape(4, 4).astype(np.float64)
    pos = M[3, :3]
    return pos, pos - M[2, :3], M[1, :3], math.degrees(cam.fovy)
