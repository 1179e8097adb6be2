"""Pins of the oracle's ray–patch intersection (CPU only).

Closed forms: planar (affine) patches have an exact ray–plane answer and an
exact inverse parameterisation.  Curved patches: an FP64 Newton solve of
S(u,v) − O − t·d = 0 seeded by brute-force sampling gives the true root; the
clipper (FP32 and FP64 instantiations) must land within its stopping
tolerance of it.  Misses, the tmax cut-off and the documented degenerate
direction (d ∥ z) are checked too.
"""
from __future__ import annotations

import numpy as np
import pytest

from conftest import rng
from test_oracle_spline import surface

TOL = 2.0 ** -12


def affine_patch(r):
    """Control net of an affine map (u,v) ↦ Q + u·E1 + v·E2 (planar).  Q and
    E/3 are multiples of 1/64, so every control point is exact in float32."""
    while True:
        Q = r.integers(-64, 65, 3) / 64.0
        E1 = 3 * r.integers(-32, 33, 3) / 64.0
        E2 = 3 * r.integers(-32, 33, 3) / 64.0
        if np.linalg.norm(np.cross(E1, E2)) > 0.5:
            break
    P = np.zeros((16, 4))
    for j in range(4):
        for i in range(4):
            P[4 * j + i, :3] = Q + (i / 3.0) * E1 + (j / 3.0) * E2
            P[4 * j + i, 3] = 1.0
    return P.astype(np.float32), Q, E1, E2


def ray_at(o, target):
    d = target - o
    ray = np.zeros(8, np.float32)
    ray[0:3], ray[3], ray[4:7] = o, np.inf, d
    return ray


def test_affine_patch_closed_form(oracle):
    r = rng(20)
    rays, pats, want = [], [], []
    for _ in range(300):
        P, Q, E1, E2 = affine_patch(r)
        uv = r.integers(3, 62, 2) / 64.0
        X = Q + uv[0] * E1 + uv[1] * E2
        o = X + r.integers(-48, 49, 3) / 16.0
        rays.append(ray_at(o, X))
        pats.append(P)
        want.append((np.linalg.norm(X - o), *uv))
    rays, pats, want = np.array(rays), np.array(pats), np.array(want)
    got32, _ = oracle.clip_batch(rays, pats)
    got64, _ = oracle.clip_batch(rays, pats, fp64=True)
    assert (got32[:, 3] == 1).all() and (got64[:, 3] == 1).all()
    assert np.max(np.abs(got64[:, 0] - want[:, 0]) / want[:, 0]) < 1e-12
    assert np.max(np.abs(got64[:, 1:3] - want[:, 1:3])) < 1e-10
    assert np.max(np.abs(got32[:, 0] - want[:, 0]) / want[:, 0]) < 2e-5
    assert np.max(np.abs(got32[:, 1:3] - want[:, 1:3])) < 2e-3


def curved_patch(r, scale=1.0):
    g = np.stack(np.meshgrid(np.linspace(0, 1, 4), np.linspace(0, 1, 4), indexing="xy"), -1).reshape(16, 2)
    P = np.zeros((16, 4))
    P[:, 0:2] = g * scale
    P[:, 2] = r.normal(0, 0.25, 16) * scale
    P[:, 3] = 1.0
    return P.astype(np.float32)


def test_curved_patch_against_newton_root(oracle):
    """Rays down onto a bumpy patch: the hit is within the clipper's tolerance
    of the FP64 Newton root, in FP32 and FP64."""
    r = rng(21)
    rays, pats, roots = [], [], []
    while len(rays) < 200:
        P = curved_patch(r)
        u, v = r.uniform(0.1, 0.9, 2)
        X = surface(P, u, v)[:3]
        o = X + np.array([r.normal() * 0.3, r.normal() * 0.3, 3.0])
        d = (X - o) / np.linalg.norm(X - o)
        ok, uvt = oracle.newton(P, o, d, [u, v, np.linalg.norm(X - o)])
        if not ok:
            continue
        # keep the nearest root only: no other sample of the patch is closer along the ray
        rays.append(ray_at(o.astype(np.float32), X.astype(np.float32)))
        pats.append(P)
        roots.append(uvt)
    rays, pats, roots = np.array(rays), np.array(pats), np.array(roots)
    g32, _ = oracle.clip_batch(rays, pats)
    g64, _ = oracle.clip_batch(rays, pats, fp64=True)
    for g in (g32, g64):
        found = g[:, 3] == 1
        assert found.mean() > 0.97
        dt = np.abs(g[found, 0] - roots[found, 2]) / roots[found, 2]
        assert np.median(dt) < 1e-5 and dt.max() < 2e-3
        duv = np.abs(g[found, 1:3] - roots[found, 0:2]).max(axis=1)
        assert np.median(duv) < 1e-3


def test_misses_and_tmax_cutoff(oracle):
    r = rng(22)
    P, Q, E1, E2 = affine_patch(r)
    X = Q + 0.5 * E1 + 0.5 * E2
    o = X + np.cross(E1, E2) * 2
    toward = ray_at(o, X)
    away = ray_at(o, 2 * o - X)
    beside = ray_at(o, X + 5 * E1)
    dist = np.linalg.norm(X - o)
    short = toward.copy()
    short[3] = dist * 0.9
    longer = toward.copy()
    longer[3] = dist * 1.1
    out, _ = oracle.clip_batch(np.array([toward, away, beside, short, longer]), np.array([P] * 5))
    assert list(out[:, 3]) == [1, 0, 0, 0, 1]
    assert out[3, 0] == np.float32(short[3])     # distance untouched on a miss


def test_direction_parallel_to_z_is_a_miss(oracle):
    """axis0 = d × (d + (0,0,8)) vanishes for d ∥ z (Spline_BezierClipping.cu:59):
    the projection degenerates and the reference reports no hit."""
    P = curved_patch(rng(23))
    ray = np.zeros(8, np.float32)
    ray[0:3], ray[3], ray[4:7] = (0.5, 0.5, 5.0), np.inf, (0.0, 0.0, -1.0)
    out, _ = oracle.clip_batch(ray[None], P[None])
    assert out[0, 3] == 0


def test_brute_force_sampling_agrees(oracle):
    """Tiny patches, rays in random directions: sample the surface on a 128²
    grid, polish the sample nearest to the ray with FP64 Newton.  A root inside
    the patch must be reported (t within 1e-4 relative); a ray far from every
    sample must miss."""
    from test_oracle_spline import bern
    r = rng(24)
    uu, vv = np.meshgrid(np.linspace(0, 1, 128), np.linspace(0, 1, 128))
    Bu, Bv = bern(uu.ravel()), bern(vv.ravel())
    n_hit = n_miss = 0
    for _ in range(150):
        P = curved_patch(r, scale=0.5)
        o = np.array([0.25, 0.25, 0.0]) + np.array([r.normal() * 0.3, r.normal() * 0.3, 1.5])
        target = np.array([r.uniform(-0.1, 0.6), r.uniform(-0.1, 0.6), 0.0])
        ray = ray_at(o.astype(np.float32), target.astype(np.float32))
        o64 = ray[0:3].astype(np.float64)
        d = ray[4:7].astype(np.float64) / np.linalg.norm(ray[4:7].astype(np.float64))
        pts = np.einsum("nj,ni,jid->nd", Bv, Bu, P.astype(np.float64).reshape(4, 4, 4)[:, :, :3])
        w = pts - o64
        lateral = np.linalg.norm(w - (w @ d)[:, None] * d, axis=1)
        out, _ = oracle.clip_batch(ray[None], P[None])
        k = int(np.argmin(lateral))
        if lateral[k] > 0.05:
            assert out[0, 3] == 0
            n_miss += 1
            continue
        ok, uvt = oracle.newton(P, o64, d, [uu.ravel()[k], vv.ravel()[k], w[k] @ d])
        if ok and 0.02 < uvt[0] < 0.98 and 0.02 < uvt[1] < 0.98 and uvt[2] > 0:
            # only single-root cases: no other sample near the ray far from this root
            near = lateral < 0.02
            if np.max(np.abs((w[near] @ d) - uvt[2])) > 0.05:
                continue
            assert out[0, 3] == 1
            assert abs(out[0, 0] - uvt[2]) / uvt[2] < 1e-4
            n_hit += 1
    assert n_hit > 20 and n_miss > 5
