"""cgs_phf_build_static, the hash build that allocates nothing, against
cgs_phf_build (the entry point it mirrors) and the oracle: the same
parameters, displacement table and status on calls in a row that share its
static scratch, the second smaller than the first; sizes past its scratch
are refused."""
from __future__ import annotations

import ctypes as C

import numpy as np
import pytest

from conftest import rng, say

pytestmark = pytest.mark.gpu


def heap_keys(n, seed):
    """The node keys of the reference's median-split BVH over n random boxes."""
    import cgspline_amd as cgs
    r = rng(seed)
    lo = r.uniform(0, 100, (n, 3)).astype(np.float32)
    hi = lo + r.uniform(0, 1, (n, 3)).astype(np.float32)
    el = np.zeros((n, 10), np.float32)
    el[:, 0:3], el[:, 3:6], el[:, 6:9] = lo, hi, 0.5 * (lo + hi)
    el.view(np.uint32)[:, 9] = np.arange(n, dtype=np.uint32)
    _, keys, _ = cgs.Context.bvh_build(el)
    return keys


def build(ctx, fn, dkeys, n, seed=42):
    import cgspline_amd as cgs
    m = cgs.lib().cgs_phf_displacement_size(n)
    dd = cgs.DeviceArray(ctx, 4 * m)
    cgs.check(cgs.lib().cgs_memset(ctx.h, dd.dptr, 0xAB, 4 * m, None))
    h = cgs._abi.PHF()
    st = fn(ctx.h, dkeys.dptr, n, seed, dd.dptr, C.byref(h), None)
    ctx.sync()
    return st, (h.A, h.B, h.mask, h.table_size, h.attempts, h.displacement_size), dd.to_host(np.uint32, (m,))


def test_static_hash_build_equals_allocating_build(ctx, oracle):
    import cgspline_amd as cgs
    lib = cgs.lib()
    cases = [("large", 150_000, 1), ("smaller", 3_000, 2), ("largest", 262_144, 3), ("smallest", 2, 4)]
    for name, n, seed in cases:
        keys = heap_keys(n, seed)
        say("static hash build", name, keys.shape[0], "keys")
        dk = cgs.DeviceArray.from_host(ctx, keys)
        a = build(ctx, lib.cgs_phf_build_static, dk, keys.shape[0])
        b = build(ctx, lib.cgs_phf_build, dk, keys.shape[0])
        assert a[0] == b[0] and a[1] == b[1], (name, a[:2], b[:2])
        assert np.array_equal(a[2], b[2]), name
        ost, op, oD = oracle.phf_build(keys)
        assert a[0] == (0 if ost == 0 else a[0]) and (ost == 0) == (a[0] == 0), (name, a[0], ost)
        assert (a[1][0], a[1][1], a[1][2], a[1][3]) == (op["A"], op["B"], op["mask"], op["table_size"]), name
        assert np.array_equal(a[2], oD), name
        if a[0] == 0:
            h = cgs._abi.PHF(*a[1])
            out = cgs.DeviceArray(ctx, 4 * keys.shape[0])
            dd = cgs.DeviceArray.from_host(ctx, a[2])
            cgs.check(lib.cgs_phf_hash(ctx.h, C.byref(h), dd.dptr, dk.dptr, keys.shape[0], out.dptr, None))
            ctx.sync()
            hv = out.to_host(np.uint32, keys.shape)
            assert np.unique(hv).shape[0] == keys.shape[0] and hv.max() < a[1][3], name


def test_static_hash_build_refuses_past_its_scratch(ctx):
    import cgspline_amd as cgs
    lib = cgs.lib()
    n = (1 << 19) + 1
    say("static hash build over", n, "keys (refused)")
    dk = cgs.DeviceArray(ctx, 4 * n)
    dd = cgs.DeviceArray(ctx, 4 * lib.cgs_phf_displacement_size(n))
    h = cgs._abi.PHF()
    assert lib.cgs_phf_build_static(ctx.h, dk.dptr, n, 42, dd.dptr, C.byref(h), None) == 1
    assert b"static scratch" in lib.cgs_last_error()
    assert lib.cgs_phf_build_static(ctx.h, dk.dptr, 0, 42, dd.dptr, C.byref(h), None) == 1
