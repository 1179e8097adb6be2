"""Pins of the oracle's subdivision intersection (doSubdivision host
semantics, Spline_Subdivision.cu:168-288) by what the mathematics fixes
(CPU only).

The method halves the parameter square, alternating u and v, until both
sides of a region are below the threshold, then intersects the two corner
triangles of that region and reports the region's centre.  So:

* The final regions are the cells of a square grid whose side h is the
  largest power of two strictly below the threshold.  (Which side is cut
  first changes only the order in which cells are visited, and so only
  exact ties in t; the grid is the same.)
* On a planar patch with an affine parameterisation the two triangles of a
  cell tile it exactly, so a ray through the point S(u*, v*) is reported at
  the centre of the cell holding (u*, v*), with the plane's exact t.
* On a curved patch the triangles deviate from the surface by O(h^2 |S''|):
  t lies within that deviation (over the cosine of incidence) of the FP64
  Newton root, and the reported centre lies within the root's cell or a
  neighbour.

Each pin is also run against copies of the oracle with a planted mistake
(tests/planted.py): reporting a corner instead of the centre, and comparing
the sides with <= instead of <; both must fail."""
from __future__ import annotations

import numpy as np
import pytest

import planted
from conftest import rng
from test_oracle_clip import affine_patch, curved_patch, ray_at
from test_oracle_spline import surface


def cell_size(threshold):
    h = 1.0
    while not h < threshold:
        h /= 2
    return h


def check_affine_cells(oracle, threshold=2.0 ** -4, n=240, seed=40):
    """Planar patches: (u,v) is exactly the centre of the root's cell and t
    the plane's distance."""
    h = cell_size(threshold)
    m = int(round(1 / h))
    r = rng(seed)
    rays, pats, want = [], [], []
    for _ in range(n):
        P, Q, E1, E2 = affine_patch(r)
        cell = r.integers(0, m, 2)
        frac = r.uniform(0.15, 0.85, 2)
        uv = (cell + frac) * h
        X = Q + uv[0] * E1 + uv[1] * E2
        N = np.cross(E1, E2)
        N /= np.linalg.norm(N)
        o = X + N * r.uniform(1.0, 3.0) + r.normal(0, 0.4, 3)
        rays.append(ray_at(o.astype(np.float32), X.astype(np.float32)))
        pats.append(P)
        ray = rays[-1]
        o32, d32 = ray[0:3].astype(np.float64), ray[4:7].astype(np.float64)
        d32 /= np.linalg.norm(d32)
        # the float32 ray's own intersection with the plane
        t = (Q - o32) @ N / (d32 @ N)
        Xr = o32 + t * d32
        A = np.stack([E1, E2], 1)
        uvr = np.linalg.lstsq(A, Xr - Q, rcond=None)[0]
        want.append((t, *((np.floor(uvr / h) + 0.5) * h)))
    rays, pats, want = np.array(rays), np.array(pats), np.array(want)
    got, _ = oracle.clip_batch(rays, pats, method=1, threshold=threshold)
    assert (got[:, 3] == 1).all(), f"{int((got[:, 3] != 1).sum())} planar hits missed"
    bad = np.flatnonzero((got[:, 1] != want[:, 1]) | (got[:, 2] != want[:, 2]))
    assert bad.size == 0, f"(u,v) not the cell centre (h = {h}): {got[bad[:3], 1:3]} vs {want[bad[:3], 1:]}"
    rel = np.abs(got[:, 0] - want[:, 0]) / want[:, 0]
    assert rel.max() < 2e-5, f"t off the plane by {rel.max():.2e} relative"


@pytest.mark.parametrize("threshold", [2.0 ** -4, 2.0 ** -3, 0.1, 2.0 ** -6])
def test_affine_patch_cell_centres(oracle, threshold):
    check_affine_cells(oracle, threshold)


def second_derivative_bound(P):
    """max |S_uu|, |S_uv|, |S_vv| over the patch (FP64, dense sampling)."""
    from test_oracle_spline import bern
    g = np.linspace(0, 1, 33)
    B, dB = bern(g), None
    # second derivatives of the Bernstein basis by differences of the control net
    C = P.astype(np.float64).reshape(4, 4, 4)[:, :, :3]     # [j][i]
    d2u = 6 * (C[:, 2:] - 2 * C[:, 1:-1] + C[:, :-2])       # [j][i'] quadratic in v? cubic in v, linear in u
    d2v = 6 * (C[2:] - 2 * C[1:-1] + C[:-2])
    duv = 9 * (C[1:, 1:] - C[1:, :-1] - C[:-1, 1:] + C[:-1, :-1])
    return max(np.abs(d2u).sum(axis=-1).max(), np.abs(d2v).sum(axis=-1).max(), np.abs(duv).sum(axis=-1).max())


def check_curved_against_newton(oracle, threshold=2.0 ** -4, seed=41, n=150):
    """Bumpy patches, rays from above: t within the triangles' deviation of
    the FP64 root; the reported centre in the root's cell or a neighbour,
    and in the root's own cell for nearly every ray."""
    h = cell_size(threshold)
    r = rng(seed)
    rays, pats, roots, bound = [], [], [], []
    while len(rays) < n:
        P = curved_patch(r)
        u, v = r.uniform(0.05, 0.95, 2)
        X = surface(P, u, v)[:3]
        o = X + np.array([r.normal() * 0.2, r.normal() * 0.2, 3.0])
        d = (X - o) / np.linalg.norm(X - o)
        ok, uvt = oracle.newton(P, o, d, [u, v, np.linalg.norm(X - o)])
        if not ok or not (0 < uvt[0] < 1 and 0 < uvt[1] < 1):
            continue
        ray = ray_at(o.astype(np.float32), X.astype(np.float32))
        o32 = ray[0:3].astype(np.float64)
        d32 = ray[4:7].astype(np.float64)
        d32 /= np.linalg.norm(d32)
        ok, uvt = oracle.newton(P, o32, d32, uvt)
        if not ok:
            continue
        Su = (surface(P, uvt[0] + 1e-6, uvt[1]) - surface(P, uvt[0] - 1e-6, uvt[1]))[:3]
        Sv = (surface(P, uvt[0], uvt[1] + 1e-6) - surface(P, uvt[0], uvt[1] - 1e-6))[:3]
        nrm = np.cross(Su, Sv)
        cos = abs(nrm @ d32) / np.linalg.norm(nrm)
        if cos < 0.3:
            continue
        rays.append(ray)
        pats.append(P)
        roots.append(uvt)
        bound.append(h * h * second_derivative_bound(P) / cos)
    rays, pats, roots, bound = np.array(rays), np.array(pats), np.array(roots), np.array(bound)
    got, _ = oracle.clip_batch(rays, pats, method=1, threshold=threshold)
    found = got[:, 3] == 1
    assert found.mean() > 0.97
    g, rt, b = got[found], roots[found], bound[found]
    assert (np.abs(g[:, 0] - rt[:, 2]) <= b + 1e-5 * rt[:, 2]).all(), "t beyond the triangles' deviation"
    duv = np.abs(g[:, 1:3] - rt[:, 0:2]).max(axis=1)
    assert duv.max() <= 1.5 * h + 1e-6, f"reported centre {duv.max() / h:.2f} cells from the root"
    own = duv <= 0.5 * h + 1e-6
    assert own.mean() > 0.9, f"only {own.mean():.2f} of hits reported in the root's own cell"


def test_curved_patch_against_newton_root(oracle):
    check_curved_against_newton(oracle)


def test_curved_patch_finer_threshold(oracle):
    check_curved_against_newton(oracle, threshold=2.0 ** -6, seed=42, n=100)


@pytest.mark.parametrize("plant", [1, 2])
def test_planted_subdivision_mistakes_fail_the_pins(plant):
    """A corner reported instead of the centre (1), or sides compared with
    <= (2, cells twice as large at a power-of-two threshold), fail the
    planar closed form."""
    planted.check_fails(plant, "test_oracle_subdivision", "check_affine_cells")


def test_planted_corner_fails_the_newton_pin():
    planted.check_fails(1, "test_oracle_subdivision", "check_curved_against_newton")
