"""Pins of the oracle's BVH build, traversal and perfect hash (CPU only).

* BVH (rt/BVH.cpp:9-124): structural invariants of a pre-order median split
  with heap keys; the Y-only split that follows from the never-inflated box.
* Traversal (rt/BVH.h:178-350): the hashed stackless closest hit equals the
  explicit-stack traversal (rt/BVH.cpp:126-189) and brute force over every
  patch; the any-hit query agrees with the closest hit.
* Perfect hash (rt/PerfectHashFunction.{h,cu}): injectivity and table
  sizing, and the (A,B) draws against an independent restatement of
  std::mt19937 + libstdc++'s uniform_int_distribution.
"""
from __future__ import annotations

import math

import numpy as np
import pytest

from conftest import rng

EMPTY = 0xFFFFFFFF


def node_box(nodes, i, side):
    w = nodes[i].view(np.float32)
    return (w[0:3], w[3:6]) if side == 0 else (w[6:9], w[9:12])


def elements(n, seed, ties=True):
    r = rng(seed)
    c = r.normal(size=(n, 3)).astype(np.float32)
    if ties:
        c[:, 1] = np.round(c[:, 1] * 4) / 4
    e = np.zeros((n, 10), np.float32)
    e[:, 0:3] = c - 0.1
    e[:, 3:6] = c + 0.1
    e[:, 6:9] = c
    e[:, 9] = np.arange(n, dtype=np.uint32).view(np.float32)
    return e


@pytest.mark.parametrize("n", [1, 2, 3, 7, 100, 1000])
def test_bvh_structure(oracle, n):
    e = elements(n, n)
    nodes, keys, idx = oracle.bvh_build_elements(e)
    assert nodes.shape[0] == 2 * n - 1 and keys.shape[0] == 2 * n - 1 and idx.shape[0] == 2 * n
    assert keys[0] == 1
    seen = []
    for i in range(nodes.shape[0]):
        left, right = nodes[i, 12], nodes[i, 13]
        if left == EMPTY:
            assert idx[right + 1] == EMPTY
            seen.append(idx[right])
            lo, hi = node_box(nodes, i, 0)
            el = e[idx[right]]
            assert np.array_equal(lo, el[0:3]) and np.array_equal(hi, el[3:6])
            continue
        assert left == i + 1                                   # pre-order
        assert keys[left] == 2 * keys[i] and keys[right] == 2 * keys[i] + 1
        for side, c in ((0, left), (1, right)):                # parent box = union of the child's boxes
            lo, hi = node_box(nodes, i, side)
            clo, chi = node_box(nodes, c, 0)
            if nodes[c, 12] != EMPTY:
                rlo, rhi = node_box(nodes, c, 1)
                clo, chi = np.minimum(clo, rlo), np.maximum(chi, rhi)
            assert np.array_equal(lo, clo) and np.array_equal(hi, chi)
    assert sorted(seen) == list(range(n))


def test_bvh_splits_on_y(oracle):
    """dim comes from an empty box (BVH.cpp:46-52): always Y.  Every left
    subtree's centroids are <= every right subtree's in y."""
    e = elements(500, 9, ties=False)
    nodes, keys, idx = oracle.bvh_build_elements(e)
    cy = e[:, 7]

    def leaves(i):
        if nodes[i, 12] == EMPTY:
            return [idx[nodes[i, 13]]]
        return leaves(nodes[i, 12]) + leaves(nodes[i, 13])

    for i in range(nodes.shape[0]):
        if nodes[i, 12] != EMPTY:
            assert cy[leaves(nodes[i, 12])].max() <= cy[leaves(nodes[i, 13])].min()


def test_hashed_traversal_equals_stack_and_brute_force(oracle, c1):
    from cgspline_amd import scenes
    sc = c1.scene
    osc = oracle.OracleScene(sc)
    rays = scenes.camera_rays(c1.camera)[::37]
    h, p, o, _ = osc.trace(rays)
    hs, ps, os_, _ = osc.trace(rays, stack=True)
    assert np.array_equal(o, os_) and np.array_equal(p, ps)
    assert np.array_equal(h.view(np.uint32), hs.view(np.uint32))
    # brute force: every ray against every patch, keep the smallest t
    P = sc.surface_patches(0)
    R = np.repeat(rays, P.shape[0], axis=0)
    res, _ = oracle.clip_batch(R, np.tile(P, (rays.shape[0], 1, 1)))
    res = res.reshape(rays.shape[0], P.shape[0], 4)
    t = np.where(res[:, :, 3] == 1, res[:, :, 0], np.inf)
    best = t.min(axis=1)
    hit = o >= 0
    assert np.array_equal(hit, np.isfinite(best))
    # the scene path clips in object space and scales back (t_world = t_local / |M d|):
    # a few ulps against clipping the world ray directly
    assert np.allclose(h[hit, 0], best[hit], rtol=1e-6, atol=0)
    assert hit.sum() > 100


def test_grazing_rays_postpone_many_siblings(oracle, grazing):
    """The grazing height field, where traversals postpone more than 8
    siblings at once: the hashed traversal still equals the stack traversal,
    and the count of siblings postponed at once stays within the deepest
    leaf's depth (one postponement per interior node on a root-to-leaf
    path)."""
    from cgspline_amd import scenes
    osc = oracle.OracleScene(grazing.scene)
    rays = scenes.camera_rays(grazing.camera)
    h, p, o, w = osc.trace(rays)
    hs, ps, os_, _ = osc.trace(rays, stack=True)
    assert np.array_equal(o, os_) and np.array_equal(p, ps)
    assert np.array_equal(h.view(np.uint32), hs.view(np.uint32))
    depth = max(int(k).bit_length() - 1 for s in (0, -1) for k in osc.bvh(s)["keys"])
    assert 8 < w["maxPostponed"] <= depth, (w["maxPostponed"], depth)
    assert (o >= 0).sum() > rays.shape[0] // 4


def test_stacked_objects_postpone_past_the_lds_levels(oracle, stacked):
    """The stacked-arcs scene: rays along the stack find both children hit
    at every level of the top-level BVH, 17 siblings postponed at once.  A
    traversal that keeps the 8 most recent in a ring and the heap keys below
    512 (9 levels) in LDS must then find a sibling at depth 9 or deeper
    through the perfect hash.  Every ray ends on the top patch."""
    sc, rays = stacked
    osc = oracle.OracleScene(sc)
    h, p, o, w = osc.trace(rays)
    assert w["maxPostponed"] >= 17, w
    assert (o == sc.nobj - 1).all()
    occ, wa = osc.trace(rays, any_hit=True)
    assert occ.all() and wa["maxPostponed"] >= 17


def test_any_hit_agrees_with_closest(oracle, c2_small):
    from cgspline_amd import scenes
    osc = oracle.OracleScene(c2_small.scene)
    rays = scenes.camera_rays(c2_small.camera)[::53].copy()
    h, p, o, _ = osc.trace(rays)
    occ, _ = osc.trace(rays, any_hit=True)
    assert np.array_equal(occ, o >= 0)
    # limited to just before the hit: occluded only if another object is closer
    rays2 = rays[o >= 0].copy()
    rays2[:, 3] = h[o >= 0, 0] * 0.999
    occ2, _ = osc.trace(rays2, any_hit=True)
    assert occ2.mean() < 0.05


# --- perfect hash -------------------------------------------------------------

class MT19937:
    """std::mt19937 (Matsumoto & Nishimura 1998), 32-bit."""

    def __init__(self, seed):
        self.mt = [0] * 624
        self.mt[0] = seed & 0xFFFFFFFF
        for i in range(1, 624):
            self.mt[i] = (1812433253 * (self.mt[i - 1] ^ (self.mt[i - 1] >> 30)) + i) & 0xFFFFFFFF
        self.i = 624

    def __call__(self):
        if self.i >= 624:
            for k in range(624):
                y = (self.mt[k] & 0x80000000) | (self.mt[(k + 1) % 624] & 0x7FFFFFFF)
                self.mt[k] = self.mt[(k + 397) % 624] ^ (y >> 1) ^ (0x9908B0DF if y & 1 else 0)
            self.i = 0
        y = self.mt[self.i]
        self.i += 1
        y ^= y >> 11
        y ^= (y << 7) & 0x9D2C5680
        y ^= (y << 15) & 0xEFC60000
        y ^= y >> 18
        return y & 0xFFFFFFFF


def uniform_u32(gen, lo, hi):
    """libstdc++ (GCC >= 11) uniform_int_distribution<uint32_t> on a 32-bit
    engine: Lemire's nearly-divisionless method on the range hi - lo + 1."""
    rng_ = (hi - lo + 1) & 0xFFFFFFFF
    if rng_ == 0:                      # full range
        return lo + gen()
    prod = gen() * rng_
    low = prod & 0xFFFFFFFF
    if low < rng_:
        thresh = ((1 << 32) - rng_) % rng_
        while low < thresh:
            prod = gen() * rng_
            low = prod & 0xFFFFFFFF
    return lo + (prod >> 32)


def reference_draws(n, seed=42, count=16):
    """(A,B) pairs of PerfectHashFunction::build (.cu:452-473)."""
    g = MT19937(seed)
    out = []
    for _ in range(count):
        while True:
            A = uniform_u32(g, 0, n)
            if A == 0:
                A = 31
            B = uniform_u32(g, 0, n)
            if math.gcd(n, A) == 1:
                break
        out.append((A, B))
    return out


def heap_keys(npatch):
    keys, stack = [], [(1, npatch)]
    while stack:
        k, c = stack.pop()
        keys.append(k)
        if c > 1:
            m = c // 2
            stack.append(((k << 1) ^ 1, c - m))
            stack.append((k << 1, m))
    return np.array(keys, np.uint32)


def test_mt19937_restatement_is_standard():
    # the C++ standard fixes the 10000th output of default-seeded mt19937 (seed 5489)
    g = MT19937(5489)
    for _ in range(9999):
        g()
    assert g() == 4123659995


@pytest.mark.parametrize("npatch", [2, 256, 4096, 40000])
def test_phf_draws_and_invariants(oracle, npatch):
    keys = heap_keys(npatch)
    n = keys.shape[0]
    st, p, D = oracle.phf_build(keys)
    draws = reference_draws(n)
    assert (p["A"], p["B"]) == draws[p["attempts"] - 1]
    assert p["mask"] + 1 == 1 << (max(2, (n - 1) >> 1).bit_length() - 1)
    assert p["table_size"] >= n
    step = max(4, int(np.float32(0.05) * np.float32(n)))
    assert (p["table_size"] - n) % step == 0
    if st == 0:
        h = (np.uint64(p["A"]) * keys.astype(np.uint64) + np.uint64(p["B"]) +
             D[keys & np.uint32(p["mask"])].astype(np.uint64)) & np.uint64(0xFFFFFFFF)
        h = (h % np.uint64(p["table_size"])).astype(np.uint32)
        assert np.unique(h).size == n and h.max() < p["table_size"]
        # buckets hold at most 4 median-split heap keys
        assert np.bincount(keys & np.uint32(p["mask"])).max() <= 4


def test_phf_bucket_overflow_reported(oracle):
    keys = np.arange(0, 64 * 10, 64, dtype=np.uint32) + 1     # all in one bucket
    st, _, _ = oracle.phf_build(keys)
    assert st == 2


def test_phf_fails_at_two_million_heap_keys(oracle):
    """The reference's scheme cannot hash one BVH over 2^20 patches: all 16
    functions fail and the final function has duplicates (build() == 1)."""
    keys = heap_keys(1 << 20)
    st, p, _ = oracle.phf_build(keys)
    assert st == 1 and p["attempts"] == 16


def with_geometryless_objects(sc, where=(0, 2, 5)):
    """The scene with objects that have no surface (the reference's mesh and
    sphere primitives, whose bounds are empty) inserted at `where`."""
    import dataclasses
    n = sc.nobj + len(where)
    keep = np.ones(n, bool)
    keep[list(where)] = False
    def spread(a, fill):
        out = np.empty((n,) + a.shape[1:], a.dtype)
        out[keep] = a
        out[~keep] = fill
        return out
    eye = np.eye(4, dtype=np.float32).reshape(16)
    return dataclasses.replace(
        sc, obj_surface=spread(sc.obj_surface, 0xFFFFFFFF),
        obj_local2world=spread(sc.obj_local2world, eye), obj_world2local=spread(sc.obj_world2local, eye),
        obj_material=spread(sc.obj_material, 0.5)), keep


def test_geometryless_objects_are_never_hit(oracle, c2_small):
    """Objects without geometry take part in the top-level build with an
    empty box (Scene.cu:39-53 with Mesh::bounds() = {}): the top-level BVH
    gains their leaves, and every ray hits what it hit without them."""
    from cgspline_amd import scenes
    sc = c2_small.scene
    rays = scenes.camera_rays(c2_small.camera)[::7]
    base = oracle.OracleScene(sc)
    tuv0, p0, o0, _ = base.trace(rays)
    sc2, keep = with_geometryless_objects(sc)
    ext = oracle.OracleScene(sc2)
    top = ext.bvh(-1)
    assert top["nodes"].shape[0] == 2 * sc2.nobj - 1
    tuv1, p1, o1, _ = ext.trace(rays)
    ids = np.flatnonzero(keep)                  # new id of each original object
    assert not np.isin(o1, np.flatnonzero(~keep)).any()
    assert np.array_equal(np.where(o0 >= 0, ids[np.maximum(o0, 0)], -1), o1)
    assert np.array_equal(p0, p1) and np.array_equal(tuv0.view(np.uint32), tuv1.view(np.uint32))
    occ0, _ = base.trace(rays, any_hit=True)
    occ1, _ = ext.trace(rays, any_hit=True)
    assert np.array_equal(occ0, occ1)


def test_mesh_elements_and_bvh_match_oracle(oracle):
    """Mesh::buildBVH (rt/Primitive.cu:90-108): cgs_mesh_elements equals the
    element data restated here in float32 (vertex box; centroid
    (1/3)((p0+p1)+p2)); the 16-triangles-per-leaf BVH built from it equals the
    oracle's, and each leaf holds at most 16 triangles."""
    import cgspline_amd as cgs
    r = rng(21)
    V = r.normal(size=(400, 3)).astype(np.float32) * np.float32(3)
    T = r.integers(0, 400, size=(700, 3)).astype(np.uint32)
    el = cgs.Context.mesh_elements(V, T)
    p0, p1, p2 = V[T[:, 0]], V[T[:, 1]], V[T[:, 2]]
    assert np.array_equal(el[:, 0:3], np.minimum(np.minimum(p0, p1), p2))
    assert np.array_equal(el[:, 3:6], np.maximum(np.maximum(p0, p1), p2))
    assert np.array_equal(el[:, 6:9], np.float32(1.0 / 3.0) * ((p0 + p1) + p2))
    assert np.array_equal(el[:, 9].view(np.uint32), np.arange(700, dtype=np.uint32))
    n1, k1, i1 = cgs.Context.bvh_build(el, 16)
    n2, k2, i2 = oracle.bvh_build_elements(el, 16)
    assert np.array_equal(n1, n2) and np.array_equal(k1, k2) and np.array_equal(i1, i2)
    leaves = n1.view(np.uint32).reshape(-1, 16)
    assert (leaves[:, 12] == 0xFFFFFFFF).sum() >= 700 // 16
    runs = np.diff(np.flatnonzero(np.append(i1 == 0xFFFFFFFF, True)))
    assert runs.max() <= 17
    with pytest.raises(cgs.CgsError, match="out of range"):
        cgs.Context.mesh_elements(V[:3], np.array([[0, 1, 3]], np.uint32))


# ---------------------------------------------------------------------------
# The wide family: the reference's function in exact arithmetic, 64-bit keys
# ---------------------------------------------------------------------------

def wide_hash_bigint(params, D, keys):
    """h(k) = (A·k + B + D[k & mask]) mod T with Python integers (no wrap)."""
    A, B, mask, T = params["A"], params["B"], params["mask"], params["table_size"]
    return np.array([(A * int(k) + B + int(D[int(k) & mask])) % T for k in keys], np.uint64)


def check_wide(oracle, keys, seed=42):
    st, p, D = oracle.phf_build_wide(keys, seed)
    assert st == 0, st
    n = keys.shape[0]
    assert p["mask"] + 1 == oracle.lib().orc_phf_displacement_size(n)
    step = max(4, int(np.float32(0.05) * np.float32(n)))
    assert p["table_size"] >= n and (p["table_size"] - n) % step == 0
    assert p["table_size"] == n + (p["attempts"] - 1) * step
    h = oracle.phf_hash_wide(p, D, keys)
    assert h.max() < p["table_size"]
    assert np.unique(h).shape[0] == n, "not injective"
    # the function is the reference's family evaluated without wrapping
    sub = keys[:: max(1, n // 20000)]
    assert np.array_equal(oracle.phf_hash_wide(p, D, sub).astype(np.uint64), wide_hash_bigint(p, D, sub))
    return st, p, D


def test_wide_family_hashes_two_million_heap_keys(oracle):
    """One BVH over 2^20 patches (2,097,151 heap keys), where every reference
    attempt fails, hashes with the wide family; so does 2^21 patches."""
    for npatch in (1 << 20, 1 << 21):
        _, p, _ = check_wide(oracle, heap_keys(npatch))
        assert p["attempts"] <= 2


def test_wide_family_equals_the_reference_where_nothing_wraps(oracle):
    """For key sets small enough that A·k + B + d stays below 2^32, the wide
    family's build is the reference's build: the same (A,B), table size and
    displacements."""
    for npatch in (5, 100, 3000, 1 << 14):
        keys = heap_keys(npatch)
        st, p, D = oracle.phf_build(keys)
        stw, pw, Dw = check_wide(oracle, keys.astype(np.uint64))
        assert st == stw == 0
        assert p == pw and np.array_equal(D, Dw)


def test_wide_family_64_bit_keys(oracle):
    """Heap keys lifted above 2^32 (the same low bits, so the same buckets):
    the build succeeds and equals the big-integer formula."""
    keys = heap_keys(1 << 16).astype(np.uint64) + (np.uint64(0x5A5) << np.uint64(40))
    check_wide(oracle, keys)
    # keys differing only above the bucket bits share buckets: five of them overflow one
    bad = np.array([7 + (k << 33) for k in range(5)], np.uint64)
    st, _, _ = oracle.phf_build_wide(bad)
    assert st == 2


def test_wide_family_pin_catches_a_wrap():
    """The big-integer pin fails for a function that wraps A·k at 32 bits,
    the reference's arithmetic, on keys past 2^32 / A."""
    p = dict(A=2_000_003, B=17, mask=7, table_size=1_000_003)
    D = np.arange(8, dtype=np.uint32)
    keys = np.arange(3_000_000, 3_000_100, dtype=np.uint64)
    wrapped = np.array([((p["A"] * int(k)) % (1 << 32) + p["B"] + int(D[int(k) & 7])) % p["table_size"]
                        for k in keys], np.uint64)
    assert not np.array_equal(wrapped, wide_hash_bigint(p, D, keys))
