"""Pins of the oracle's PBR functions (rt/PBR.h, rt/PBR.cu; CPU only) by the
closed forms of the published models (Karis 2013, "Real Shading in Unreal
Engine 4"): Schlick's Fresnel, Smith–Schlick G with k = (r+1)²/8, the GGX
distribution with α = r² and its normalisation ∫ D(h)(n·h) dω = 1, and the
light falloff.  The GPU renderer uses the same functions bit for bit (the
render parity tests)."""
from __future__ import annotations

import math

import numpy as np
import pytest


def test_schlick_endpoints_and_closed_form(oracle):
    f0 = (0.04, 0.5, 0.95)
    assert np.array_equal(oracle.pbr("schlick", *f0, 1.0)[:3], np.float32(f0))    # normal incidence: f0
    assert np.array_equal(oracle.pbr("schlick", *f0, 0.0)[:3], np.float32([1, 1, 1]))  # grazing: 1
    for x in np.linspace(0, 1, 37):
        want = np.array(f0) + (1 - np.array(f0)) * (1 - x) ** 5
        assert np.allclose(oracle.pbr("schlick", *f0, x)[:3], want, atol=2e-7)


@pytest.mark.parametrize("r", [0.05, 0.3, 0.7, 1.0])
def test_smith_g_closed_form(oracle, r):
    k = (r + 1) ** 2 / 8
    g1 = lambda x: x / (x * (1 - k) + k)
    assert oracle.pbr("G", 1.0, 1.0, r)[0] == np.float32(1.0)
    for nl, nv in [(0.2, 0.9), (0.5, 0.5), (0.99, 0.1), (0.01, 0.7)]:
        assert math.isclose(oracle.pbr("G", nl, nv, r)[0], g1(nl) * g1(nv), rel_tol=2e-6)


@pytest.mark.parametrize("r", [0.2, 0.4, 0.7, 1.0])
def test_ggx_is_normalised(oracle, r):
    """∫ over the hemisphere of D(n·h)(n·h) dω = 1 (θ-integral, FP64 trapezoid
    over the oracle's FP32 values)."""
    th = np.linspace(0, math.pi / 2, 20001)
    D = np.array([oracle.pbr("D", math.cos(t), r)[0] for t in th[::20]], np.float64)
    t = th[::20]
    integral = np.trapezoid(D * np.cos(t) * np.sin(t) * 2 * math.pi, t)
    assert abs(integral - 1) < 3e-3


def test_ggx_peak_and_roughness_clamp(oracle):
    for r in (0.3, 0.8):
        a2 = r ** 4
        # n·h = 1: D = 1/(π α²); FP32 d = (α²-1)+1 cancels to ~ulp(1)/α² relative
        assert math.isclose(oracle.pbr("D", 1.0, r)[0], 1 / (math.pi * a2), rel_tol=4 * 2 ** -23 / a2)
    # roughness below 1e-3 is clamped (PBR.h:87-96)
    assert oracle.pbr("D", 0.999, 0.0)[0] == oracle.pbr("D", 0.999, 1e-3)[0]


def test_microfacet_is_g_d_over_4_nl_nv(oracle):
    for nv, nl, nh, r in [(0.7, 0.4, 0.9, 0.5), (0.2, 0.9, 0.6, 0.1)]:
        g = oracle.pbr("G", nl, nv, r)[0]
        d = oracle.pbr("D", nh, r)[0]
        assert math.isclose(oracle.pbr("microfacet", nv, nl, nh, r)[0], g * d / (4 * nl * nv), rel_tol=3e-6)


def light(color=(1, 2, 3), pos=(0, 0, 5), range_=0.0, direction=(0, 0, -1), angle=math.pi):
    return [*color, 1.0, *pos, range_, *direction, angle]


def test_light_falloff(oracle):
    for d in (0.5, 2.0, 7.0):
        assert np.allclose(oracle.pbr("light_color", d, *light())[:3], np.array([1, 2, 3]) / d ** 2, rtol=2e-7)
        R = 10.0
        assert np.allclose(oracle.pbr("light_color", d, *light(range_=R))[:3],
                           np.array([1, 2, 3]) * (1 - d / R) ** 2, rtol=2e-6)
    # directional (angle < 0.01): no falloff
    assert np.array_equal(oracle.pbr("light_color", 3.0, *light(angle=0.0))[:3], np.float32([1, 2, 3]))


def test_light_vector(oracle):
    ok, d, *L = oracle.pbr("light_vector", 0, 0, 1, *light())[:5]
    assert ok == 1 and math.isclose(d, 4.0) and np.allclose(L, [0, 0, 1])
    ok = oracle.pbr("light_vector", 0, 0, 1, *light(range_=3.0))[0]          # beyond range
    assert ok == 0
    # A spot light (0.01 <= angle <= 1.57, Scene.h:73-76) lights nothing in
    # the reference: PBR.cu:41 asks DL < 0 and angle > 2 acos(DL), i.e.
    # angle > pi.  Kept for parity; checked on and off the cone axis.
    for ang in (0.02, math.radians(60), 1.57):
        for dirn in ((0, 0, -1), (0, 0, 1)):
            spot = light(pos=(0, 0, 5), direction=dirn, angle=ang)
            for P in ((0, 0, 0), (0.5, 0, 0), (0, 0, 10), (5, 0, 5)):
                assert oracle.pbr("light_vector", *P, *spot)[0] == 0
    # just above 1.57 it is a point light and lights everything in range
    assert oracle.pbr("light_vector", 5, 0, 5, *light(angle=1.5701))[0] == 1
    # directional: L = -direction, infinite distance
    ok, d, *L = oracle.pbr("light_vector", 1, 2, 3, *light(direction=(0, -1, 0), angle=0.0))[:5]
    assert ok == 1 and math.isinf(d) and np.allclose(L, [0, 1, 0])
