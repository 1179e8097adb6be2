"""GPU parity of the reference's per-patch header functions (Spline.h,
SplineMat.h: interpolate, derivatives, normal, subpatch; cgs_patch_eval,
cgs_patch_subpatch) and of cg::PatchBVH's queries (PatchBVH.cpp:84-121;
cgs_patchbvh_*) against the oracle: evaluations and subpatches bit for bit;
PatchBVH hits equal to the nearest of the oracle's clips of every patch (an
exact tie between two patches may report either), any hits equal to whether
some patch clips with a hit in front."""
from __future__ import annotations

import os

import numpy as np
import pytest

from conftest import rng, say
from test_gpu_parity import assert_bits_equal

pytestmark = pytest.mark.gpu

ASSETS = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "assets", "bezier")
KINDS = ["interp_dc", "interp_mat", "du_dc", "dv_dc", "du_mat", "dv_mat", "normal_dc", "normal_mat"]


def _teapot():
    import cgspline_amd as cgs
    return cgs.load_be(os.path.join(ASSETS, "teapot"))


def _patches(v, i):
    return np.ascontiguousarray(v[i], np.float32)       # (np, 16, 4)


def test_patch_eval_every_kind_bitexact(ctx, oracle):
    """All eight kinds on the teapot's patches (degenerate lid and bottom
    patches included, where the normal's 1e-4 step applies) at random (u,v)
    and at the corners and edges."""
    v, i = _teapot()
    P = _patches(v, i)
    r = rng(71)
    uv = np.concatenate([r.random((P.shape[0] * 6, 2)),
                         np.tile(np.array([[0, 0], [1, 0], [0, 1], [1, 1], [0.5, 0], [0, 0.5]]), (P.shape[0], 1))])
    uv = uv.astype(np.float32)
    pp = np.concatenate([np.repeat(P, 6, axis=0), np.repeat(P, 6, axis=0)])
    for kind, name in enumerate(KINDS):
        say("cgs_patch_eval", name, pp.shape[0])
        g = ctx.patch_eval(kind, pp, uv)
        o = np.stack([oracle.evaluate(pp[k], uv[k, 0], uv[k, 1], name) for k in range(pp.shape[0])]).astype(np.float32)
        if name.startswith("normal"):
            o[:, 3] = 0.0
        assert_bits_equal(g, o, name)


def test_subpatch_bitexact(ctx, oracle):
    v, i = _teapot()
    P = _patches(v, i)
    r = rng(72)
    a = np.sort(r.random((P.shape[0], 2, 2)), axis=2).astype(np.float32)
    ranges = np.stack([a[:, 0, 0], a[:, 0, 1], a[:, 1, 0], a[:, 1, 1]], 1)
    ranges[::5] = [0, 1, 0, 1]
    say("cgs_patch_subpatch", P.shape[0])
    g = ctx.subpatch(P, ranges)
    o = np.stack([oracle.subpatch(P[k], *ranges[k]) for k in range(P.shape[0])])
    assert_bits_equal(g.reshape(-1, 64), o.reshape(-1, 64), "subpatch")


def _pick_rays(v, n, seed, tmax=np.inf):
    """Rays from a sphere around the model towards points near it."""
    r = rng(seed)
    xyz = v[:, :3]
    c = xyz.mean(axis=0)
    rad = float(np.linalg.norm(xyz - c, axis=1).max())
    d = r.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    o = c + 2.5 * rad * d
    target = c + r.uniform(-0.6, 0.6, (n, 3)) * rad
    dirs = (target - o) * r.uniform(0.5, 3.0, (n, 1))     # not unit: the query normalizes once
    rays = np.zeros((n, 8), np.float32)
    rays[:, 0:3], rays[:, 3], rays[:, 4:7] = o, tmax, dirs
    return rays


def _brute_force(oracle, rays, P):
    """Every (ray, patch) clip by the oracle: (n, np) arrays of t (inf where
    not found), u, v."""
    n, m = rays.shape[0], P.shape[0]
    R = np.repeat(rays, m, axis=0)
    Q = np.tile(P, (n, 1, 1))
    res, _ = oracle.clip_batch(R, Q)
    found = (res[:, 3] == 1).reshape(n, m)
    t = np.where(found, res[:, 0].reshape(n, m), np.inf)
    return t, res[:, 1].reshape(n, m), res[:, 2].reshape(n, m), found


def test_patchbvh_closest_equals_nearest_patch(ctx, oracle):
    v, i = _teapot()
    P = _patches(v, i)
    pb = ctx.patch_bvh(v, i)
    rays = _pick_rays(v, 2000, 73)
    rays[1::4, 3] = 4.0                                     # some with a finite range
    say("cgs_patchbvh_intersect", rays.shape[0], "rays,", P.shape[0], "patches")
    g_tuv, g_p = pb.intersect(rays)
    t, u, w, found = _brute_force(oracle, rays, P)
    best = t.min(axis=1)
    hit = np.isfinite(best)
    assert hit.sum() > 600 and (~hit).sum() > 100
    assert np.array_equal(g_p == 0xFFFFFFFF, ~hit)
    assert np.array_equal(g_tuv[hit, 0].view(np.uint32), best[hit].astype(np.float32).view(np.uint32))
    for k in np.nonzero(hit)[0]:
        ties = np.nonzero(t[k] == best[k])[0]
        assert g_p[k] in ties, k
        j = int(g_p[k])
        assert g_tuv[k, 1] == u[k, j] and g_tuv[k, 2] == w[k, j], k
    assert np.all(g_tuv[~hit, 0] == rays[~hit, 3])
    pb.destroy()


def test_patchbvh_any(ctx, oracle):
    v, i = _teapot()
    P = _patches(v, i)
    pb = ctx.patch_bvh(v, i)
    rays = _pick_rays(v, 2000, 74)
    say("cgs_patchbvh_intersect_any", rays.shape[0])
    g = pb.intersect_any(rays)
    r_inf = rays.copy()
    r_inf[:, 3] = np.inf
    _, _, _, found = _brute_force(oracle, r_inf, P)
    want = found.any(axis=1)
    assert not np.any(g == 2)
    assert np.array_equal(g == 1, want)
    assert want.sum() > 600 and (~want).sum() > 100
    pb.destroy()


SE_ASSETS = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "assets", "surfaces")


@pytest.mark.parametrize("name, min_hits", [("quad16", 0), ("tube", 1000), ("ball0", 1000), ("dice16", 2000),
                                             ("cube16", 0)])
def test_patchbvh_over_se_tables_matches_oracle(ctx, oracle, name, min_hits):
    """cg::PatchBVH of the reference's `.se` surfaces: the tree over the boxes
    of the basis-transformed points, the leaf test on each patch's 16 indices
    (PatchBVH.cpp:40-121).  Closest and any picks equal the oracle's
    PatchBVH queries on the same tree, bit for bit.  The flat quad and the
    cube's faces have boxes of zero thickness, which the reference's strict
    slab test never enters: no ray hits them, there as here."""
    import cgspline_amd as cgs
    from cgspline_amd import scenes
    t = cgs.load_se(os.path.join(SE_ASSETS, name + ".se"))
    el, pi = t.elements(), t.patch_indices()
    pb = cgs.PatchBVH.from_table(ctx, t)
    rays = _pick_rays(t.points, 3000, 75)
    rays[1::3, 3] = 2.0
    say("PatchBVH over", name, pi.shape[0], "patches,", rays.shape[0], "rays")
    g_tuv, g_p = pb.intersect(rays)
    sc = scenes.assemble([(t.points, pi)], [(0, scenes.IDENTITY16, scenes.IDENTITY16,
                                             np.full(8, 0.5, np.float32))], np.zeros((0, 12), np.float32))
    osc = oracle.OracleScene(sc, surface0_elements=el)
    o_tuv, o_p = osc.pick(rays)
    assert np.array_equal(g_p, o_p), np.flatnonzero(g_p != o_p)[:5]
    hit = o_p != 0xFFFFFFFF
    assert_bits_equal(g_tuv, o_tuv, "t,u,v")
    occ = pb.intersect_any(rays)
    o_occ = osc.pick(rays, any_hit=True)
    assert not np.any(occ == 2) and np.array_equal(occ == 1, o_occ)
    assert hit.sum() >= min_hits and (min_hits > 0 or hit.sum() == 0)
    pb.destroy()
