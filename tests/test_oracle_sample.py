"""Pins of the oracle's sampled (Monte Carlo) render (CPU only): the Philox
generator by the Random123 known-answer vectors, the draws' float mapping,
the Van der Corput positions (PBR.h:145-157), the angle functions against
libm in double precision, the sampled bounce (RayTracer.cu:473-522) against
the GGX sampling of Walter et al. (EGSR 2007: tan θm = α sqrt(x / (1 - x)),
φm = 2π y, weight F G |v·m| / (|n·v| |n·m|)), and whole sampled frames of a
flat square against an FP64 closed form over every sample and draw.  Each
render pin also runs against oracle copies with a planted mistake
(tests/planted.py), which must fail it.

The reference's own draws come from cuRAND and are not reproduced: its
frames agree with these in distribution only."""
from __future__ import annotations

import math

import numpy as np
import pytest

import planted
from conftest import rng
from test_oracle_render import BG, LIGHT_COL, LIGHT_POS, MAT, expected_colour, flat_square_scene


# Random123 kat_vectors, philox4x32 with 10 rounds: counter, key, result.
PHILOX_KAT = [
    ((0, 0, 0, 0), (0, 0), (0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8)),
    ((0xFFFFFFFF,) * 4, (0xFFFFFFFF,) * 2, (0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD)),
    ((0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344), (0xA4093822, 0x299F31D0),
     (0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1)),
]


def test_philox_known_answers(oracle):
    ctr = np.array([k[0] for k in PHILOX_KAT], np.uint32)
    key = np.array([k[1] for k in PHILOX_KAT], np.uint32)
    out = oracle.philox(ctr, key)
    assert np.array_equal(out, np.array([k[2] for k in PHILOX_KAT], np.uint32))


def test_draws_are_philox_blocks_and_open_unit_floats(oracle):
    """Draw m of stream s under seed k is word m mod 4 of the block with
    counter (m / 4, 0, s, 0) and key (k, 0); as a float it is an odd multiple
    of 2^-24 in (0, 1), uniform (Kolmogorov-Smirnov) and different per
    stream."""
    words, units = oracle.sample_words(42, 1234, 0, 4096)
    m = np.arange(4096)
    ctr = np.zeros((1024, 4), np.uint32)
    ctr[:, 0], ctr[:, 2] = np.arange(1024), 1234
    key = np.tile(np.array([42, 0], np.uint32), (1024, 1))
    assert np.array_equal(words, oracle.philox(ctr, key).reshape(-1)[m])
    q = units.astype(np.float64) * 2 ** 24
    assert np.all(q == np.round(q)) and np.all(q % 2 == 1)
    assert np.array_equal(q, 2 * (words >> 9).astype(np.float64) + 1)
    assert units.min() > 0 and units.max() < 1
    from scipy import stats
    assert stats.kstest(units, "uniform").pvalue > 1e-3
    other, _ = oracle.sample_words(42, 1235, 0, 4096)
    reseeded, _ = oracle.sample_words(43, 1234, 0, 4096)
    assert (words == other).mean() < 0.01 and (words == reseeded).mean() < 0.01
    tail, _ = oracle.sample_words(42, 1234, 4000, 96)
    assert np.array_equal(tail, words[4000:])


def test_vdc_radical_inverse(oracle):
    i = np.concatenate([np.arange(1 << 12), rng(5).integers(0, 1 << 32, 4096, dtype=np.uint64)]).astype(np.uint32)
    want = np.array([int(format(int(k), "032b")[::-1], 2) for k in i], np.float64) * 2.0 ** -32
    assert np.array_equal(oracle.vdc(i), want.astype(np.float32))
    assert list(oracle.vdc(np.arange(8))) == [0, 0.5, 0.25, 0.75, 0.125, 0.625, 0.375, 0.875]


def _f32_ulps(a, b):
    a = np.asarray(a, np.float32).view(np.int32).astype(np.int64)
    b = np.asarray(b, np.float32).view(np.int32).astype(np.int64)
    return np.abs(a - b)


def test_angle_functions_against_libm(oracle):
    """sin/cos on [-8, 8] and atan2 over the first quadrant equal libm's double
    results rounded to float, to within an ulp (nearly always exactly)."""
    r = rng(6)
    a = np.concatenate([r.uniform(-8, 8, 200000), [0.0, math.pi / 2, math.pi, 2 * math.pi, 1e-30, -1e-30],
                        np.arange(-16, 17) * (math.pi / 4)]).astype(np.float32)
    s, c = oracle.sincos(a)
    ws = np.sin(a.astype(np.float64)).astype(np.float32)
    wc = np.cos(a.astype(np.float64)).astype(np.float32)
    assert _f32_ulps(s, ws).max() <= 1 and _f32_ulps(c, wc).max() <= 1
    assert (s != ws).mean() < 1e-4 and (c != wc).mean() < 1e-4
    y = np.concatenate([r.uniform(0, 3, 200000), r.uniform(0, 1e-6, 1000), [0.0, 1.0, 2.0]]).astype(np.float32)
    x = np.concatenate([r.uniform(0, 3, 200000), r.uniform(0.5, 1, 1000), [1.0, 0.0, 2.0]]).astype(np.float32)
    t = oracle.atan2q(y, x)
    wt = np.arctan2(y.astype(np.float64), x.astype(np.float64)).astype(np.float32)
    assert _f32_ulps(t, wt).max() <= 1 and (t != wt).mean() < 1e-4
    assert t[-3] == 0 and t[-2] == np.float32(math.pi / 2) and t[-1] == np.float32(math.pi / 4)


def ggx_bounce64(N, V, mat, x, y):
    """FP64 GGX importance sampling (Walter et al. 2007) as the reference
    applies it: M at θm = atan(α sqrt(x/(1-x))), α = roughness², φm = 2π y
    about N in the frame (N×V, N×(N×V)); R = 2(M·V)M - V; the path ends where
    N·R < 1e-14; weight = (1-metalness) albedo/π + metalness F(M·V) G(N·R, N·V)
    (M·V)/((N·V)(N·M)), Schlick F and Smith-Schlick G with k = (r+1)²/8."""
    N, V = np.asarray(N, np.float64), np.asarray(V, np.float64)
    diffuse, f0, metal, r = mat[0:3].astype(np.float64), mat[3:6].astype(np.float64), float(mat[6]), float(mat[7])
    alpha = r * r
    th = math.atan(alpha * math.sqrt(x / (1 - x)))
    ph = 2 * math.pi * y
    ts = np.cross(N, V)
    ts = ts / np.linalg.norm(ts)
    tp = np.cross(N, ts)
    tp = tp / np.linalg.norm(tp)
    M = math.sin(th) * math.cos(ph) * ts + math.sin(th) * math.sin(ph) * tp + math.cos(th) * N
    M = M / np.linalg.norm(M)
    mv = M @ V
    R = 2 * mv * M - V
    nv, nm, nr = N @ V, N @ M, N @ R
    if nr < 1e-14:
        return R, None
    k = (r + 1) ** 2 / 8
    G = (nr / (nr * (1 - k) + k)) * (nv / (nv * (1 - k) + k))
    F = f0 + (1 - f0) * (1 - mv) ** 5
    return R, (1 - metal) * diffuse / math.pi + metal * F * G * mv / (nv * nm)


def test_sampled_bounce_is_ggx_importance_sampling(oracle):
    r = rng(7)
    n = 3000
    N = r.normal(size=(n, 3))
    N /= np.linalg.norm(N, axis=1, keepdims=True)
    V = r.normal(size=(n, 3))
    V /= np.linalg.norm(V, axis=1, keepdims=True)
    V *= np.sign(np.sum(N * V, axis=1, keepdims=True))
    mats = np.column_stack([r.uniform(0, 1, (n, 6)), r.uniform(0, 1, n), r.uniform(0.05, 1, n)]).astype(np.float32)
    xy = r.uniform(0.001, 0.999, (n, 2)).astype(np.float32)
    N32, V32 = N.astype(np.float32), V.astype(np.float32)
    R, brdf, ok = oracle.sampled_bounce(N32, V32, mats, xy)
    ended = 0
    for k in range(n):
        R64, w64 = ggx_bounce64(N32[k], V32[k], mats[k], float(xy[k, 0]), float(xy[k, 1]))
        assert np.abs(R[k] - R64).max() < 2e-5, k
        if (w64 is None) != (not ok[k]):
            # the two may disagree only on a reflection grazing the surface
            assert abs(float(N32[k].astype(np.float64) @ R64)) < 1e-4, k
        if w64 is None or not ok[k]:
            ended += 1
            continue
        assert (np.abs(brdf[k] - w64) / np.maximum(np.abs(w64), 1e-3)).max() < 1e-3, k
    assert ok.sum() > n // 2 and (~ok).sum() > 10 and ended == (~ok).sum()


def test_sampled_bounce_distribution(oracle):
    """Over uniform draws, cos θm follows GGX's D(m) cos θm density: P(θm <= θ)
    = tan²θ / (α² + tan²θ)."""
    n = 20000
    r = rng(8)
    xy = r.uniform(0, 1, (n, 2)).astype(np.float32)
    N = np.tile(np.array([0, 0, 1], np.float32), (n, 1))
    V = np.tile(np.array([0.6, 0, 0.8], np.float32), (n, 1))
    mat = np.tile(np.array([0.5, 0.5, 0.5, 0.9, 0.9, 0.9, 1.0, 0.6], np.float32), (n, 1))
    R, _, _ = oracle.sampled_bounce(N, V, mat, xy)
    H = R.astype(np.float64) + V
    H /= np.linalg.norm(H, axis=1, keepdims=True)
    th = np.arccos(np.clip(np.abs(H[:, 2]), 0, 1))       # R + V = 2 (M·V) M: M up to sign
    alpha = 0.6 ** 2
    from scipy import stats
    cdf = lambda t: np.tan(t) ** 2 / (alpha ** 2 + np.tan(t) ** 2)
    assert stats.kstest(th, cdf).pvalue > 1e-3


def check_sampled_frames(oracle, depth=1, samples=6, seed=42):
    """Every pixel of a sampled view of the flat square equals the FP64
    closed form: the mean over its samples, through (i + s/n, j + vdc(s)), of
    the hit's direct light without ambient (times π) and, while depth allows,
    the background along the GGX-sampled reflection, weighted as Walter et
    al.; draws 2k, 2k+1 of the pixel's stream go to its k-th hitting sample."""
    from cgspline_amd import scenes
    sc = flat_square_scene()
    cam = scenes.make_camera((0.4, -1.8, 2.2), (0.0, 0.1, 0.0), (0.0, 0.0, 1.0), 50.0, 20, 16)
    color, frame, counts = oracle.OracleScene(sc).render_sampled(cam, samples=samples, seed=seed, eps=1e-4, depth=depth)
    n = samples
    nInv = np.float32(1) / np.float32(n)
    s0 = np.arange(n, dtype=np.float32) * nInv
    s1 = oracle.vdc(np.arange(n))
    amb = np.array((0.05, 0.1, 0.15))
    dif = (1 - float(MAT[6])) * MAT[0:3].astype(np.float64) / math.pi
    hit_rays = 0
    for j in range(cam.height):
        for i in range(cam.width):
            xy = np.column_stack([np.float32(i) + s0, np.float32(j) + s1])
            d = oracle.OracleScene.pixel_rays(cam, xy)
            stream = (i & 255) + 256 * (j & 255)
            _, u = oracle.sample_words(seed, stream, 0, 2 * n)
            total = np.zeros(3)
            k = 0
            for s in range(n):
                o = cam.position.astype(np.float64)
                direct = expected_colour(o, d[s], 0)
                if np.array_equal(direct, np.array(BG)):
                    total += direct
                    continue
                hit_rays += 1
                total += direct - math.pi * amb * dif
                if depth <= 0:
                    continue
                V = -d[s].astype(np.float64)
                N = np.array([0.0, 0.0, 1.0]) * (1 if V[2] >= 0 else -1)
                _, w = ggx_bounce64(N, V, MAT, float(u[2 * k]), float(u[2 * k + 1]))
                k += 1
                if w is None or w.max() <= 2 ** -8:
                    continue
                total += w * np.array(BG)
            want = total / n
            got = color[j, i].astype(np.float64)
            rel = (np.abs(got - want) / np.maximum(np.abs(want), 1e-3)).max()
            assert rel < 2e-4, f"pixel ({i}, {j}) off the closed form by {rel:.2e}"
    assert 0.2 < hit_rays / (cam.width * cam.height * n) < 0.95
    assert counts["primary"] == cam.width * cam.height * n and counts["shadow"] > 0
    assert (counts["secondary"] > 0) == (depth > 0)


def test_sampled_frames_closed_form(oracle):
    check_sampled_frames(oracle)


def test_sampled_frames_closed_form_without_bounce(oracle):
    check_sampled_frames(oracle, depth=0, samples=3)


def test_sampled_render_deterministic_and_seeded(oracle):
    from cgspline_amd import scenes
    sc = flat_square_scene()
    cam = scenes.make_camera((0.4, -1.8, 2.2), (0.0, 0.1, 0.0), (0.0, 0.0, 1.0), 50.0, 24, 16)
    osc = oracle.OracleScene(sc)
    a, fa, _ = osc.render_sampled(cam, samples=4, seed=7, depth=1)
    b, fb, _ = osc.render_sampled(cam, samples=4, seed=7, depth=1, threads=3)
    c, _, _ = osc.render_sampled(cam, samples=4, seed=8, depth=1)
    assert np.array_equal(a.view(np.uint32), b.view(np.uint32)) and np.array_equal(fa, fb)
    assert not np.array_equal(a, c)
    with pytest.raises(ValueError):
        osc.render_sampled(cam, samples=0)


@pytest.mark.parametrize("plant", [10, 11, 12])
def test_planted_sampling_mistakes_fail_the_pin(plant):
    planted.check_fails(plant, "test_oracle_sample", "check_sampled_frames")
