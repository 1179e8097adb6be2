"""Pins of the oracle's spline math against closed forms (CPU only).

The reference ships no tests or golden vectors for this path, and it cannot
be compiled here, so the oracle is pinned by mathematics: Bernstein closed
forms in float64, identities of the blossom (slice) and libm's own fmin/fmax.
"""
from __future__ import annotations

import ctypes
import ctypes.util
import math

import numpy as np
import pytest

from conftest import rng

BIN = np.array([1.0, 3.0, 3.0, 1.0])


def bern(t):
    t = np.asarray(t, np.float64)
    return BIN * t[..., None] ** np.arange(4) * (1 - t[..., None]) ** (3 - np.arange(4))


def dbern(t):
    # d/dt of the cubic Bernstein basis
    t = float(t)
    return np.array([-3 * (1 - t) ** 2, 3 * (1 - t) ** 2 - 6 * t * (1 - t), 6 * t * (1 - t) - 3 * t * t, 3 * t * t])


def surface(P, u, v):
    """S(u,v) with point(i,j) = P[4j+i] (i along u)."""
    P = np.asarray(P, np.float64).reshape(4, 4, -1)   # [j][i]
    return np.einsum("j,i,jid->d", bern(v), bern(u), P)


def rand_patch(r, w=False):
    P = r.normal(size=(16, 4)).astype(np.float32)
    P[:, 3] = r.uniform(0.5, 2.0, 16) if w else 1.0
    return P


# --- scalar helpers -------------------------------------------------------------

def test_fmin_fmax_match_libm(oracle):
    libm = ctypes.CDLL(ctypes.util.find_library("m"))
    libm.fminf.argtypes = libm.fmaxf.argtypes = [ctypes.c_float, ctypes.c_float]
    libm.fminf.restype = libm.fmaxf.restype = ctypes.c_float
    vals = [0.0, -0.0, 1.0, -1.0, float("nan"), float("inf"), -float("inf"), 1e-45, 3.5]
    for a in vals:
        for b in vals:
            for f, g in ((oracle.fminf, libm.fminf), (oracle.fmaxf, libm.fmaxf)):
                x, y = f(a, b), g(a, b)
                assert (math.isnan(x) and math.isnan(y)) or (x == y and math.copysign(1, x) == math.copysign(1, y)), (a, b)


# --- evaluation -----------------------------------------------------------------

@pytest.mark.parametrize("kind", ["interp_dc", "interp_mat"])
def test_interpolate_equals_bernstein(oracle, kind):
    r = rng(1)
    for _ in range(50):
        P = rand_patch(r, w=True)
        u, v = r.random(2)
        exact = surface(P, u, v)
        got64 = oracle.evaluate(P, u, v, kind, fp64=True)
        assert np.allclose(got64, exact, rtol=1e-13, atol=1e-13)
        got32 = oracle.evaluate(P, np.float32(u), np.float32(v), kind)
        assert np.allclose(got32, surface(P, np.float32(u), np.float32(v)), rtol=2e-5, atol=2e-5)


def test_interpolate_corners_are_control_points(oracle):
    P = rand_patch(rng(2))
    for (u, v), k in (((0, 0), 0), ((1, 0), 3), ((0, 1), 12), ((1, 1), 15)):
        for kind in ("interp_dc", "interp_mat"):
            assert np.array_equal(oracle.evaluate(P, u, v, kind).astype(np.float32), P[k]), (kind, u, v)


@pytest.mark.parametrize("form", ["dc", "mat"])
def test_derivatives_equal_bernstein(oracle, form):
    r = rng(3)
    for _ in range(30):
        P = rand_patch(r)
        u, v = r.random(2)
        Pg = P.astype(np.float64).reshape(4, 4, 4)
        du = np.einsum("j,i,jid->d", bern(v), dbern(u), Pg)
        dv = np.einsum("j,i,jid->d", dbern(v), bern(u), Pg)
        # de Casteljau derivative omits the factor 3 of the hodograph (Spline.h:183-195)
        scale = 1.0 / 3.0 if form == "dc" else 1.0
        assert np.allclose(oracle.evaluate(P, u, v, f"du_{form}", fp64=True)[:3], scale * du[:3], atol=1e-12)
        assert np.allclose(oracle.evaluate(P, u, v, f"dv_{form}", fp64=True)[:3], scale * dv[:3], atol=1e-12)


def test_normal_degenerate_fix(oracle):
    """Spline.cu:10-27: a zero derivative (collapsed row) is re-evaluated at
    |v - 1e-4|; the matrix form has no fix and returns a zero normal."""
    P = rand_patch(rng(4))
    P[0:4, :3] = P[0, :3]            # row v = 0 collapsed to a point: dS/du = 0 at v = 0
    n_dc = oracle.evaluate(P, 0.3, 0.0, "normal_dc", fp64=True)[:3]
    n_mat = oracle.evaluate(P, 0.3, 0.0, "normal_mat", fp64=True)[:3]
    assert np.linalg.norm(n_dc) > 0
    assert np.allclose(n_mat, 0.0)


# --- slicing ----------------------------------------------------------------------

def curve_at(c, t):
    return np.einsum("i,id->d", bern(t), np.asarray(c, np.float64))


@pytest.mark.parametrize("side", [0, 1])
def test_subpatch_is_reparametrization(oracle, side):
    r = rng(5 + side)
    for _ in range(30):
        p2 = r.normal(size=(16, 2)).astype(np.float32)
        a, b = np.sort(r.uniform(0, 0.98, 2)).astype(np.float32)
        b = np.float32(max(b, a + 0.01))
        q = oracle.subpatch2d(p2, side, a, b)
        for k in range(4):
            idx = [4 * k + i for i in range(4)] if side == 0 else [k + 4 * i for i in range(4)]
            for s in (0.0, 0.3, 1.0):
                want = curve_at(p2[idx], float(a) + s * (float(b) - float(a)))
                assert np.allclose(curve_at(q[idx], s), want, atol=3e-5)


def test_subpatch_full_range_is_identity(oracle):
    p2 = rng(7).normal(size=(16, 2)).astype(np.float32)
    for side in (0, 1):
        assert np.array_equal(oracle.subpatch2d(p2, side, 0.0, 1.0), p2)


def test_slice_matrix_equals_de_casteljau_slice(oracle):
    """SplineMat.h:64-82 closed form vs Spline.h:324-364 (float32 within ulps)."""
    r = rng(8)
    for _ in range(30):
        a, b = np.sort(r.uniform(0, 1, 2)).astype(np.float32)
        M = oracle.slice_matrix(a, b).astype(np.float64)       # columns
        c = r.normal(size=(4, 2)).astype(np.float32)
        p2 = np.zeros((16, 2), np.float32)
        p2[0:4] = c
        q = oracle.subpatch2d(p2, 0, a, b)[0:4]
        # mat4(P0..P3) * slice_matrix: new point k = sum_i column_k[i] * P_i
        via_matrix = np.einsum("ki,id->kd", M, c.astype(np.float64))
        assert np.allclose(q, via_matrix, atol=5e-6)


# --- hull helpers ---------------------------------------------------------------------

def test_x_axis_intersection_closed_form(oracle):
    r = rng(9)
    for _ in range(200):
        A = r.normal(size=2).astype(np.float32)
        B = r.normal(size=2).astype(np.float32)
        ok, s = oracle.x_axis_intersection(A, B)
        crosses = (A[1] <= 0 <= B[1]) or (B[1] <= 0 <= A[1])
        assert ok == crosses
        if ok:
            x = A[0] + (B[0] - A[0]) * (A[1] / (A[1] - B[1]))
            assert s[0] == s[1] and abs(s[0] - x) < 1e-5 * (1 + abs(x))
    ok, s = oracle.x_axis_intersection([2.0, 0.0], [-1.0, 0.0])      # on the axis: whole segment
    assert ok and tuple(s) == (-1.0, 2.0)
    ok, _ = oracle.x_axis_intersection([0.0, 1.0], [1.0, 1.0])       # parallel above
    assert not ok


def test_triangle_origin_barycentrics(oracle):
    r = rng(10)
    hits = 0
    for _ in range(300):
        P, A, B = (r.normal(size=2).astype(np.float32) for _ in range(3))
        ok, c = oracle.triangle_origin(P, A, B)
        M = np.array([A - P, B - P], np.float64).T
        uv = np.linalg.solve(M, -P.astype(np.float64))
        inside = uv[0] >= 0 and uv[1] >= 0 and uv.sum() <= 1
        if abs(uv[0]) > 1e-5 and abs(uv[1]) > 1e-5 and abs(uv.sum() - 1) > 1e-5:
            assert ok == inside
        if ok:
            hits += 1
            assert np.allclose(c, uv, atol=1e-5)
    assert hits > 10


def test_subpatch_is_a_reparametrisation(oracle):
    """spline::subpatch (Spline.h:384-396): the sub-patch's surface at (s, t)
    is the patch's at (umin + s (umax - umin), vmin + t (vmax - vmin)), in
    homogeneous coordinates (FP64 Bernstein evaluation; FP32 slicing error
    only); the full range leaves the patch as it is."""
    r = rng(12)
    for k in range(40):
        P = rand_patch(r, w=(k % 2 == 1))
        a = np.sort(r.random(2)).astype(np.float32)
        b = np.sort(r.random(2)).astype(np.float32)
        Q = oracle.subpatch(P, a[0], a[1], b[0], b[1])
        for s, t in r.random((6, 2)):
            want = surface(P, a[0] + s * (a[1] - a[0]), b[0] + t * (b[1] - b[0]))
            got = surface(Q, s, t)
            assert np.allclose(got, want, rtol=2e-5, atol=2e-5), (k, s, t)
        assert np.array_equal(oracle.subpatch(P, 0.0, 1.0, 0.0, 1.0), P)
    # a planted transposition (u and v ranges swapped) fails the identity
    P = rand_patch(r)
    Q = oracle.subpatch(P, 0.1, 0.4, 0.5, 0.9)
    assert not np.allclose(surface(Q, 0.3, 0.7), surface(P, 0.5 + 0.3 * 0.4, 0.1 + 0.7 * 0.3), rtol=1e-3)
