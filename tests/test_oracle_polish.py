"""The oracle's Newton polish of closest hits (polishHit64, the opt-in
cgs_trace_options.polish), pinned on a rational bicubic patch against an
independent FP64 root: numpy Bernstein evaluation of
S(u,v) = sum B_i(u) B_j(v) w_ij P_ij / sum B_i(u) B_j(v) w_ij with its exact
quotient-rule derivatives, and Newton on S(u,v) - o - t d = 0 for each FP32
ray.  Polished hits lie within 1e-6 of that root in u, v and relative t;
the clipped (unpolished) hits of the same rays do not, so the bar tells the
two apart.  Planted mistakes in the polish (a residual sign, the quotient
rule's -p w_u term dropped) fail the pin."""
from __future__ import annotations

import math

import numpy as np
import pytest

import planted
from conftest import rng

BAR = 1e-6


def _bernstein(t):
    t = np.asarray(t, np.float64)[..., None]
    k = np.arange(4)
    c = np.array([1.0, 3.0, 3.0, 1.0])
    b = c * t ** k * (1 - t) ** (3 - k)
    # d/dt of C(3,k) t^k (1-t)^(3-k)
    db = c * (k * t ** np.maximum(k - 1, 0) * (1 - t) ** (3 - k)
              - (3 - k) * t ** k * (1 - t) ** np.maximum(2 - k, 0))
    return b, db


def rational_eval(H, u, v):
    """Point and partial derivatives of the rational patch with homogeneous
    control points H (16, 4), row j = v index, column i = u index."""
    G = H.reshape(4, 4, 4).astype(np.float64)                  # [j, i, xyzw]
    bu, dbu = _bernstein(u)
    bv, dbv = _bernstein(v)
    X = np.einsum("ni,nj,jic->nc", bu, bv, G)
    Xu = np.einsum("ni,nj,jic->nc", dbu, bv, G)
    Xv = np.einsum("ni,nj,jic->nc", bu, dbv, G)
    w = X[:, 3:4]
    p = X[:, :3] / w
    return p, (Xu[:, :3] - p * Xu[:, 3:4]) / w, (Xv[:, :3] - p * Xv[:, 3:4]) / w


def root(H, rays, uvt, steps=40):
    """FP64 Newton on S(u,v) - o - t d = 0 for the FP32 rays (n, 8), from
    (u, v, t) seeds; returns the converged (n, 3)."""
    o = rays[:, 0:3].astype(np.float64)
    d = rays[:, 4:7].astype(np.float64)
    x = np.array(uvt, np.float64)
    for _ in range(steps):
        p, pu, pv = rational_eval(H, x[:, 0], x[:, 1])
        F = p - o - x[:, 2:3] * d
        J = np.stack([pu, pv, -d], axis=2)
        x = x - np.linalg.solve(J, F[..., None])[..., 0]
    p, _, _ = rational_eval(H, x[:, 0], x[:, 1])
    assert np.abs(p - o - x[:, 2:3] * d).max() < 1e-12
    return x


def rational_case(seed=60, n=400):
    from cgspline_amd import scenes
    r = rng(seed)
    j, i = np.meshgrid(np.arange(4), np.arange(4), indexing="ij")
    P = np.stack([i / 3.0, j / 3.0, r.uniform(-0.05, 0.05, (4, 4))], -1).reshape(16, 3)
    w = r.uniform(0.5, 2.0, 16)
    H = np.concatenate([P * w[:, None], w[:, None]], 1).astype(np.float32)
    eye = np.eye(4, dtype=np.float32).reshape(16)
    mat = np.array([0.8, 0.8, 0.8, 0.0, 0.04, 0.04, 0.04, 0.5], np.float32)
    sc = scenes.assemble([(H, np.arange(16, dtype=np.uint32)[None])], [(0, eye, eye, mat)], [], name="rational patch")
    uv0 = r.uniform(0.1, 0.9, (n, 2))
    target, _, _ = rational_eval(H, uv0[:, 0], uv0[:, 1])
    d = np.array([0.25, 0.15, -1.0])
    d /= np.linalg.norm(d)
    rays = np.zeros((n, 8), np.float32)
    rays[:, 0:3] = target - 3.0 * d
    rays[:, 3] = np.inf
    rays[:, 4:7] = d
    return sc, H, rays, np.concatenate([uv0, np.full((n, 1), 3.0)], 1)


def _errors(h, o, rx):
    hit = o >= 0
    duv = np.abs(h[hit, 1:3].astype(np.float64) - rx[hit, 0:2]).max()
    dt = (np.abs(h[hit, 0].astype(np.float64) - rx[hit, 2]) / rx[hit, 2]).max()
    return hit, duv, dt


def check_polish_rational(oracle, seed=60):
    sc, H, rays, seed_uvt = rational_case(seed)
    rx = root(H, rays, seed_uvt)
    osc = oracle.OracleScene(sc)
    h, p, o, _ = osc.trace(rays, polish=True)
    hit, duv, dt = _errors(h, o, rx)
    assert hit.sum() >= 0.98 * rays.shape[0], f"{hit.sum()} hits of {rays.shape[0]}"
    assert duv <= BAR and dt <= BAR, f"polished: uv {duv:.2e}, rel t {dt:.2e}"
    return duv, dt


def test_polished_rational_hits_within_1e6_of_the_root(oracle):
    duv, dt = check_polish_rational(oracle)
    print(f"\npolished: max |duv| {duv:.2e}, max rel t {dt:.2e}")


def test_clipped_rational_hits_are_outside_the_bar(oracle):
    """Without the polish the clipper's 2^-12 stopping rule leaves (u, v)
    farther from the root than the bar: the bar tells polished from clipped."""
    sc, H, rays, seed_uvt = rational_case()
    rx = root(H, rays, seed_uvt)
    h, p, o, _ = oracle.OracleScene(sc).trace(rays)
    hit, duv, dt = _errors(h, o, rx)
    assert hit.sum() >= 0.98 * rays.shape[0]
    assert duv > 10 * BAR, duv


def test_root_matches_a_closed_form():
    """The numpy root itself, on a bilinear-in-disguise patch (all weights 1,
    control points on the plane z = 0 at (i/3, j/3)): S(u,v) = (u, v, 0), so a
    ray through (a, b, 0) has the root (a, b, |o - (a,b,0)|)."""
    j, i = np.meshgrid(np.arange(4), np.arange(4), indexing="ij")
    H = np.stack([i / 3.0, j / 3.0, np.zeros((4, 4)), np.ones((4, 4))], -1).reshape(16, 4)
    a, b = 0.3125, 0.71875
    d = np.array([0.25, 0.15, -1.0], np.float32)
    d /= np.linalg.norm(d)
    o = (np.array([a, b, 0.0]) - 3.0 * d).astype(np.float32)
    rays = np.zeros((1, 8), np.float32)
    rays[0, 0:3], rays[0, 4:7] = o, d
    x = root(H, rays, [[0.5, 0.5, 1.0]])[0]
    ex_t = -float(o[2]) / float(d[2])
    assert math.isclose(x[2], ex_t, rel_tol=1e-14)
    assert math.isclose(x[0], float(o[0]) + ex_t * float(d[0]), abs_tol=1e-14)
    assert math.isclose(x[1], float(o[1]) + ex_t * float(d[1]), abs_tol=1e-14)


@pytest.mark.parametrize("plant", [8, 9])
def test_planted_polish_mistakes_fail_the_pin(plant):
    planted.check_fails(plant, "test_oracle_polish", "check_polish_rational")
