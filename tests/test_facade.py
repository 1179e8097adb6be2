"""The C++ facade (include/cgspline/cgspline.hpp), driven by
tests/cpp/facade_check the way the reference's callers use rt::BVH,
rt::PerfectHashFunction, rt::Scene and rt::RayTracer.  The BVH build and the
error behaviour run on the CPU; the hash and the render need the GPU and are
checked against the oracle."""
from __future__ import annotations

import os
import subprocess

import numpy as np
import pytest

import conftest
from conftest import say

EXE = os.path.join(conftest.ROOT, "tests", "cpp", "facade_check")


def run(*args, timeout=120):
    if not os.path.exists(EXE):
        pytest.fail(f"{EXE} is not built (run __graft_entry__.build())")
    r = subprocess.run([EXE, *map(str, args)], stdout=subprocess.PIPE, stderr=subprocess.STDOUT, timeout=timeout)
    return r.returncode, r.stdout.decode(errors="replace")


def write_elements(path, el, epn=1):
    with open(path, "wb") as f:
        f.write(np.array([el.shape[0], epn], np.uint32).tobytes())
        f.write(np.ascontiguousarray(el, np.float32).tobytes())


def elements_of(scene):
    import cgspline_amd
    return cgspline_amd.Context.patch_elements(scene.vertices, scene.indices)


def test_facade_bvh_build_equals_oracle(tmp_path, oracle, c1):
    el = elements_of(c1.scene)
    write_elements(tmp_path / "in", el)
    rc, out = run("bvh", tmp_path / "in", tmp_path / "out")
    assert rc == 0, out
    raw = np.fromfile(tmp_path / "out", np.uint32)
    nn, ni = int(raw[0]), int(raw[1])
    nodes = raw[2:2 + 16 * nn].reshape(nn, 16)
    keys = raw[2 + 16 * nn:2 + 17 * nn]
    idx = raw[2 + 17 * nn:2 + 17 * nn + ni]
    n2, k2, i2 = oracle.bvh_build_patches(c1.scene.vertices, c1.scene.indices)
    assert np.array_equal(nodes, n2) and np.array_equal(keys, k2) and np.array_equal(idx, i2)


@pytest.mark.parametrize("name", ["teapot", "teaspoon"])
def test_facade_load_be_equals_oracle(tmp_path, oracle, name):
    """Surface::load_be through the facade on the reference's models."""
    src = os.path.join(conftest.ROOT, "tests", "golden", "assets", "bezier", name)
    rc, out = run("load_be", src, tmp_path / "s.bin")
    assert rc == 0, out
    raw = (tmp_path / "s.bin").read_bytes()
    nv, npch = np.frombuffer(raw[:8], np.uint32)
    v = np.frombuffer(raw[8:8 + 16 * nv], np.float32).reshape(-1, 4)
    ix = np.frombuffer(raw[8 + 16 * nv:], np.uint32).reshape(-1, 16)
    with open(src, "rb") as f:
        ov, oi = oracle.load_be(f.read())
    assert ix.shape[0] == npch
    assert np.array_equal(v.view(np.uint32), ov.view(np.uint32)) and np.array_equal(ix, oi)


@pytest.mark.parametrize("name", ["tube", "ball0", "dice16"])
def test_facade_load_se_equals_oracle(tmp_path, oracle, name):
    """Surface::load_se through the facade: points, the leaf test's 16 indices
    per patch (16 zeros past the table's end) and PatchBVH's element boxes,
    against the oracle's restatement."""
    src = os.path.join(conftest.ROOT, "tests", "golden", "assets", "surfaces", name + ".se")
    rc, out = run("load_se", src, tmp_path / "s.bin")
    assert rc == 0, out
    raw = (tmp_path / "s.bin").read_bytes()
    nv, npch = np.frombuffer(raw[:8], np.uint32)
    o = 8
    v = np.frombuffer(raw[o:o + 16 * nv], np.float32).reshape(-1, 4)
    o += 16 * nv
    ix = np.frombuffer(raw[o:o + 64 * npch], np.uint32).reshape(-1, 16)
    o += 64 * npch
    el = np.frombuffer(raw[o:], np.float32).reshape(-1, 10)
    with open(src, "rb") as f:
        pts, idx, mat, refs = oracle.load_se(f.read())
    want = np.zeros((refs.shape[0], 16), np.uint32)
    for k, (off, _, _, _) in enumerate(refs):
        if off + 16 <= idx.shape[0]:
            want[k] = idx[off:off + 16]
    assert np.array_equal(v.view(np.uint32), pts.view(np.uint32)) and np.array_equal(ix, want)
    assert np.array_equal(el.view(np.uint32), oracle.se_elements(pts, idx, mat, refs).view(np.uint32))


@pytest.mark.skipif(os.path.exists("/dev/kfd") and os.access("/dev/kfd", os.R_OK), reason="a device is present")
def test_facade_error_behaviour_without_device():
    rc, out = run("errors")
    assert rc == 0, out


@pytest.mark.gpu
def test_facade_hash_table(tmp_path, oracle, c2_small):
    say("facade hash", c2_small.scene.nobj, "objects")
    P = c2_small.scene.surface_patches(0)
    el = elements_of(c2_small.scene)[: P.shape[0]]
    write_elements(tmp_path / "in", el)
    rc, out = run("hash", tmp_path / "in", tmp_path / "out")
    assert rc == 0, out
    raw = np.fromfile(tmp_path / "out", np.uint32)
    A, B, mask, T, attempts, dsize = (int(x) for x in raw[:6])
    _, keys, _ = oracle.bvh_build_elements(el)
    hashes = raw[6:6 + keys.shape[0]]
    table = raw[6 + keys.shape[0]:]
    st, p, D = oracle.phf_build(keys)
    assert st == 0 and (A, B, mask, T) == (p["A"], p["B"], p["mask"], p["table_size"])
    want = ((np.uint64(A) * keys.astype(np.uint64) + np.uint64(B) + D[keys & np.uint32(mask)].astype(np.uint64))
            & np.uint64(0xFFFFFFFF)) % np.uint64(T)
    assert np.array_equal(hashes, want.astype(np.uint32))
    assert table.shape[0] == T and np.array_equal(table[hashes], np.arange(keys.shape[0], dtype=np.uint32))


@pytest.mark.gpu
def test_facade_render_matches_oracle(tmp_path, oracle, c1):
    wl, sc, cam = c1, c1.scene, c1.camera
    say("facade render", cam.width, cam.height)
    with open(tmp_path / "in", "wb") as f:
        u32 = lambda *v: f.write(np.array(v, np.uint32).tobytes())
        f32 = lambda v: f.write(np.ascontiguousarray(v, np.float32).tobytes())
        u32(sc.nsurf, sc.nobj, sc.lights.shape[0])
        for s in range(sc.nsurf):
            v0, v1 = int(sc.surf_vbase[s]), int(sc.surf_vbase[s + 1])
            p0, p1 = int(sc.surf_pbase[s]), int(sc.surf_pbase[s + 1])
            u32(v1 - v0, p1 - p0)
            f32(sc.vertices[v0:v1])
            f.write(np.ascontiguousarray(sc.indices[p0:p1], np.uint32).tobytes())
        for o in range(sc.nobj):
            u32(sc.obj_surface[o])
            f32(sc.obj_local2world[o])
            f32(sc.obj_world2local[o])
            f32(sc.obj_material[o])
        f32(sc.lights)
        f32(sc.background)
        f32(sc.ambient)
        u32(sc.hash_seed)
        f32(cam.to_world)
        f32(cam.position)
        f32([cam.fovy])
        u32(cam.width, cam.height)
        f32([wl.eps])
        f.write(np.array([wl.depth], np.int32).tobytes())
    rc, out = run("render", tmp_path / "in", tmp_path / "out")
    assert rc == 0, out
    # the tracer's own frame, then the frames of two RayTracers rendering one
    # independently built scene at the same time on two streams, then a Monte
    # Carlo tracer's frame (2 samples, seed 42)
    frames = np.fromfile(tmp_path / "out", np.uint32).reshape(4, cam.height, cam.width)
    _, oframe, _, _ = oracle.OracleScene(sc).render(cam, eps=wl.eps, depth=wl.depth)
    ch = lambda fr: np.stack([(fr >> s) & 0xFF for s in (0, 8, 16, 24)], -1).astype(np.int32)
    assert np.abs(ch(frames[0]) - ch(oframe)).max() <= 1
    assert np.array_equal(frames[1], frames[0]) and np.array_equal(frames[2], frames[0])
    _, sframe, _ = oracle.OracleScene(sc).render_sampled(cam, samples=2, seed=42, eps=wl.eps, depth=wl.depth)
    assert np.abs(ch(frames[3]) - ch(sframe)).max() <= 1 and not np.array_equal(frames[3], frames[0])


@pytest.mark.gpu
def test_facade_pick_and_per_call_functions(tmp_path, oracle):
    """cg::PatchBVH::intersect (closest and any), the per-call
    spline::doBezierClipping / doSubdivision and spline::evaluate, driven from
    C++ on the teapot, against the Python binding's PatchBVH and the oracle."""
    import cgspline_amd as cgs
    from test_gpu_patchbvh import ASSETS, _pick_rays
    model = os.path.join(ASSETS, "teapot")
    v, i = cgs.load_be(model)
    rays = _pick_rays(v, 300, 75)
    rays[::3, 3] = 5.0
    rays.tofile(tmp_path / "rays")
    say("facade pick", rays.shape[0])
    rc, out = run("pick", model, tmp_path / "rays", tmp_path / "out")
    assert rc == 0, out
    rec = np.fromfile(tmp_path / "out", np.uint32).reshape(rays.shape[0], 18)
    f32 = rec.view(np.float32)
    ctx = cgs.Context(0)
    pb = ctx.patch_bvh(v, i)
    tuv, patch = pb.intersect(rays)
    occ = pb.intersect_any(rays)
    found = rec[:, 0] == 1
    assert np.array_equal(found, patch != 0xFFFFFFFF)
    assert np.array_equal(rec[found, 1:4], tuv[found].view(np.uint32)) and np.array_equal(rec[found, 4], patch[found])
    assert np.array_equal(rec[:, 5], occ)
    assert found.sum() > 50
    P = v[i][np.arange(rays.shape[0]) % i.shape[0]]
    for col, fn in ((6, lambda: oracle.clip_batch(rays, P)[0]),
                    (10, lambda: oracle.clip_batch(rays, P, method=1)[0])):
        o = fn()
        of = o[:, 3] == 1
        assert np.array_equal(rec[:, col] == 1, of)
        assert np.array_equal(f32[of, col + 1:col + 4], o[of, :3])
    nrm = np.stack([oracle.evaluate(P[k], np.float32(0.3), np.float32(0.6), "normal_dc") for k in range(P.shape[0])])
    nrm = nrm.astype(np.float32)
    nrm[:, 3] = 0.0
    assert np.array_equal(rec[:, 14:18], nrm.view(np.uint32))
    pb.destroy()
    ctx.close()
