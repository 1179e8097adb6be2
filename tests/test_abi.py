"""The C ABI: the library loads, exports every symbol include/cgs.h declares,
and refuses to run without a device (no fallback).  No GPU needed."""
from __future__ import annotations

import os
import re

import numpy as np
import pytest

import conftest

HEADER = os.path.join(conftest.ROOT, "include", "cgs.h")


def declared_symbols():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(cgs_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_the_binding_list():
    from cgspline_amd import _abi
    assert declared_symbols() == sorted(_abi.EXPORTS)


def test_library_exports_every_declared_symbol():
    import cgspline_amd
    lib = cgspline_amd.lib()
    missing = [s for s in declared_symbols() if not hasattr(lib, s)]
    assert not missing, missing


def test_abi_version_and_status_strings():
    import cgspline_amd
    lib = cgspline_amd.lib()
    assert lib.cgs_abi_version() == 2
    assert lib.cgs_status_string(3) == b"hash table build failed"
    assert lib.cgs_status_string(7) == b"invalid"        # BVH::wrap's logic_error text


def test_struct_layouts_match_header():
    import ctypes as C
    from cgspline_amd import _abi
    assert C.sizeof(_abi.Ray) == 32
    assert C.sizeof(_abi.Hit) == 16
    assert C.sizeof(_abi.BVHNode) == 64          # rt::BVH::Node is 64 B (BVH.h:23-39)
    assert C.sizeof(_abi.BVHElement) == 40
    assert C.sizeof(_abi.RayCounts) == 32


@pytest.mark.skipif(os.path.exists("/dev/kfd") and os.access("/dev/kfd", os.R_OK), reason="a device is present")
def test_open_without_device_fails_loudly():
    import cgspline_amd
    with pytest.raises(cgspline_amd.CgsError) as e:
        cgspline_amd.Context(0)
    assert e.value.status == 9   # CGS_ERR_NO_DEVICE


def test_host_bvh_build_rejects_empty_and_matches_oracle(oracle, c1):
    """The host BVH build through the ABI equals the oracle's restatement."""
    import cgspline_amd
    from cgspline_amd import _abi
    el = cgspline_amd.Context.patch_elements(c1.scene.vertices, c1.scene.indices)
    n1, k1, i1 = cgspline_amd.Context.bvh_build(el)
    n2, k2, i2 = oracle.bvh_build_patches(c1.scene.vertices, c1.scene.indices)
    assert np.array_equal(n1, n2) and np.array_equal(k1, k2) and np.array_equal(i1, i2)
    with pytest.raises(_abi.CgsError) as e:
        cgspline_amd.Context.bvh_build(np.zeros((0, 10), np.float32))
    assert e.value.status == 7
