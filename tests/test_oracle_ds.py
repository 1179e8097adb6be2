"""Pins of the oracle's restatements of the absent Ds library's boundary
functions (CPU; the GPU path equals the oracle bit for bit through the clip,
trace and render parity tests):

* Vector::versor — unit length and direction in FP64; unchanged at or below
  length FLT_EPSILON (the boundary itself included);
* triangle::intersect as Möller–Trumbore (Spline_Subdivision.cu:225-226) —
  hit/miss, t and barycentrics against an FP64 linear solve;
* pack_sRGB (RayTracer.cu:262-266) — the IEC 61966-2-1 encode in FP64,
  red in the low byte, alpha 255;
* Bounds3::intersect as the slab test — interior points of the returned
  interval lie in the box, points just outside it do not.
"""
from __future__ import annotations

import math

import numpy as np
import pytest

FLT_EPS = float(np.finfo(np.float32).eps)


def rng(seed):
    return np.random.default_rng(seed)


def test_versor(oracle):
    r = rng(1)
    for _ in range(400):
        v = r.normal(size=3) * math.exp(r.uniform(math.log(1e-3), math.log(1e3)))
        got = oracle.ds("versor", *v)[:3].astype(np.float64)
        v32 = np.float32(v).astype(np.float64)
        assert abs(np.linalg.norm(got) - 1) < 4 * 2.0 ** -23
        assert np.allclose(got, v32 / np.linalg.norm(v32), atol=4 * 2.0 ** -23)
    for tiny in ([2.0 ** -23, 0, 0], [0, 0, 0], [1e-8, -3e-8, 2e-8], [0, 2.0 ** -24, 2.0 ** -24]):
        assert np.array_equal(oracle.ds("versor", *tiny)[:3], np.float32(tiny))     # length <= FLT_EPSILON
    assert np.array_equal(oracle.ds("versor", 2.0 ** -22, 0, 0)[:3], np.float32([1, 0, 0]))


def test_moller_trumbore_against_linear_solve(oracle):
    r = rng(2)
    agree = hits = 0
    for _ in range(3000):
        p = r.normal(size=(3, 3))
        o = r.normal(size=3) * 2
        target = p.mean(axis=0) + r.normal(size=3) * 0.8
        d = target - o
        a = np.float32(np.concatenate([o, d, p.ravel()])).astype(np.float64)
        o, d, p0, p1, p2 = a[0:3], a[3:6], a[6:9], a[9:12], a[12:15]
        M = np.stack([p1 - p0, p2 - p0, -d], axis=1)
        if abs(np.linalg.det(M)) < 1e-6:
            continue
        b1, b2, t = np.linalg.solve(M, o - p0)
        margin = min(b1, b2, 1 - b1 - b2)
        if abs(margin) < 1e-4 or abs(t) < 1e-4:
            continue
        out = oracle.ds("triangle", *a)
        inside = margin > 0 and t > 0
        assert bool(out[0]) == inside, (b1, b2, t)
        agree += 1
        if inside:
            hits += 1
            assert abs(out[1] - t) <= 1e-4 * (1 + abs(t))
            assert abs(out[2] - b1) <= 1e-4 and abs(out[3] - b2) <= 1e-4
            assert abs(out[4] - (1 - b1 - b2)) <= 2e-4
    assert agree > 2000 and hits > 300
    # a ray in the triangle's plane (det = 0) misses
    assert oracle.ds("triangle", -1, 0.2, 0, 1, 0, 0, 0, 0, 0, 1, 0, 0, 0, 1, 0)[0] == 0


def srgb64(c):
    c = min(max(c, 0.0), 1.0) if c == c else 0.0
    e = 12.92 * c if c <= 0.0031308 else 1.055 * c ** (1 / 2.4) - 0.055
    return int(math.floor(e * 255 + 0.5))


def test_pack_srgb(oracle):
    cs = np.concatenate([np.linspace(0, 1, 2049), [0.0031308, 0.0031309, 0.001, 0.5, -0.5, 1.5]]).astype(np.float32)
    for c in cs:
        q = int(oracle.ds("srgb", c, 0, 1)[:1].view(np.uint32)[0])
        assert q >> 24 == 255 and (q >> 8) & 255 == 0 and (q >> 16) & 255 == 255
        assert abs((q & 255) - srgb64(float(c))) <= 1, c
    known = {0.0: 0, 1.0: 255, 0.5: 188, 0.001: 3, 0.25: 137, -1.0: 0, 2.0: 255}
    for c, want in known.items():
        assert int(oracle.ds("srgb", c, c, c)[:1].view(np.uint32)[0]) & 255 == want, c


def test_box_slab(oracle):
    r = rng(4)
    checked = 0
    for _ in range(2000):
        lo = r.normal(size=3)
        hi = lo + r.uniform(0.05, 2, size=3)
        o = r.normal(size=3) * 3
        # most rays aimed near the box, the rest anywhere
        d = (0.5 * (lo + hi) + r.normal(size=3) * 0.6 - o) if r.random() < 0.7 else r.normal(size=3)
        a = np.float32(np.concatenate([lo, hi, o, d])).astype(np.float64)
        lo, hi, o, d = a[0:3], a[3:6], a[6:9], a[9:12]
        ok, tmin, tmax = (float(x) for x in oracle.ds("box", *a)[:3])
        inside = lambda t: np.all(o + t * d >= lo - 1e-9) and np.all(o + t * d <= hi + 1e-9)
        # FP64 slab interval
        t0 = (lo - o) / d
        t1 = (hi - o) / d
        n = max(0.0, np.minimum(t0, t1).max())
        f = np.maximum(t0, t1).min()
        if abs(f - n) < 1e-4 * (1 + abs(f)):
            continue
        assert bool(ok) == (n <= f)
        if ok:
            checked += 1
            assert inside(0.5 * (tmin + tmax))
            if tmin > 1e-3:
                assert not inside(tmin - 1e-3 * (1 + tmin))
            assert not inside(tmax + 1e-3 * (1 + tmax))
            assert abs(tmin - n) <= 1e-5 * (1 + abs(n)) and abs(tmax - f) <= 1e-5 * (1 + abs(f))
    assert checked > 200
