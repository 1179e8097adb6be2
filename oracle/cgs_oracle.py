"""ctypes wrapper of the oracle library — TEST INFRASTRUCTURE ONLY.

The oracle is a CPU restatement of the reference's ray–patch path
(cgs_oracle.hpp).  Only tests/, ``__graft_entry__.smoke()`` and bench.py's
cpu_baseline leg use it, as the reference results compare against; the product never does.
"""
from __future__ import annotations

import ctypes as C
import os
import re
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
# CGS_ORACLE_LIB loads another build of the oracle (the tests' planted-mistake
# copies, tests/planted.py).
LIB_PATH = os.environ.get("CGS_ORACLE_LIB") or os.path.join(_HERE, "liboracle_cgs.so")

_P = C.c_void_p
_U32, _U64, _I32, _F32, _F64 = C.c_uint32, C.c_uint64, C.c_int32, C.c_float, C.c_double


class Work(C.Structure):
    _fields_ = [(n, C.c_uint64) for n in (
        "nodes", "leaves", "hashProbes", "patchTests", "clipIters", "subdivs", "prunes",
        "converged", "onHits", "maxStack", "objects", "maxPostponed", "overflows", "topNodes")]

    def asdict(self):
        return {n: int(getattr(self, n)) for n, _ in self._fields_}


class SceneDesc(C.Structure):
    _fields_ = [
        ("vertices", _P), ("indices", _P), ("surf_vbase", _P), ("surf_pbase", _P), ("nsurf", _U32),
        ("obj_surface", _P), ("obj_local2world", _P), ("obj_world2local", _P), ("obj_material", _P),
        ("nobj", _U32), ("lights", _P), ("nlight", _U32),
        ("background", _F32 * 3), ("ambient", _F32 * 3), ("seed", _U32)]


class TraceOpts(C.Structure):
    _fields_ = [("method", C.c_int), ("tol", _F32), ("threshold", _F32), ("maxIters", _U32), ("polish", C.c_int)]


class CameraDesc(C.Structure):
    _fields_ = [("toWorld", _F32 * 16), ("position", _F32 * 3), ("fovy", _F32), ("flipY", C.c_int)]


def build(force: bool = False) -> None:
    """Compiles the oracle library with its Makefile."""
    if force or not os.path.exists(LIB_PATH):
        subprocess.run(["make", "-s", "-C", _HERE] + (["-B"] if force else []), check=True)


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        L.orc_clip_batch.argtypes = [C.c_int, _P, _P, _U64, _F32, _P, _P, _P, C.c_int, _F32]
        L.orc_clip2d_candidates.argtypes = [_P, _F32, _P, C.c_int, _P]
        L.orc_eval.argtypes = [C.c_int, _P, _F64, _F64, C.c_int, _P]
        L.orc_subpatch2d.argtypes = [_P, C.c_int, _F32, _F32]
        L.orc_subpatch.argtypes = [_P, _F32, _F32, _F32, _F32]
        L.orc_slice_matrix.argtypes = [_P, _P]
        L.orc_x_axis_intersection.argtypes = [_P, _P, _P]
        L.orc_triangle_origin.argtypes = [_P, _P, _P, _P]
        L.orc_pbr.argtypes = [C.c_int, _P, _P]
        L.orc_ds.argtypes = [C.c_int, _P, _P]
        L.orc_fminf.argtypes = [_F32, _F32]
        L.orc_fminf.restype = _F32
        L.orc_fmaxf.argtypes = [_F32, _F32]
        L.orc_fmaxf.restype = _F32
        L.orc_newton.argtypes = [_P, _P, _P, _P, C.c_int]
        L.orc_clip_batch64.argtypes = [_P, _P, _U64, _F64, _P]
        L.orc_scene_record_tests.argtypes = [_P, _P, _U64, _P, _P, _U64]
        L.orc_scene_record_tests.restype = _U64
        L.orc_newton_batch.argtypes = [_P, _P, _U64, _P, _P, _P, C.c_int]
        L.orc_bvh_build_patches.argtypes = [_P, _P, _U32]
        L.orc_bvh_build_patches.restype = _P
        L.orc_bvh_build_elements.argtypes = [_P, _U32, _U32]
        L.orc_bvh_build_elements.restype = _P
        L.orc_bvh_sizes.argtypes = [_P, _P, _P]
        L.orc_bvh_copy.argtypes = [_P, _P, _P, _P]
        L.orc_bvh_free.argtypes = [_P]
        L.orc_phf_displacement_size.argtypes = [_U64]
        L.orc_phf_displacement_size.restype = _U32
        L.orc_phf_build.argtypes = [_P, _U64, _U32, _P, _P]
        L.orc_phf_build_wide.argtypes = [_P, _U64, _U32, _P, _P]
        L.orc_phf_hash_wide.argtypes = [_P, _P, _P, _U64, _P]
        L.orc_scene_bvh_wide.argtypes = [_P, C.c_int]
        L.orc_scene_create_elements.argtypes = [C.POINTER(SceneDesc), _P, C.POINTER(C.c_int)]
        L.orc_scene_create_elements.restype = _P
        L.orc_scene_pick.argtypes = [_P, _P, _U64, C.c_int, _P, _P, _P, C.c_int]
        L.orc_scene_create.argtypes = [C.POINTER(SceneDesc), C.POINTER(C.c_int)]
        L.orc_scene_create.restype = _P
        L.orc_scene_destroy.argtypes = [_P]
        L.orc_scene_bvh_sizes.argtypes = [_P, C.c_int, _P, _P, _P, _P]
        L.orc_scene_bvh_copy.argtypes = [_P, C.c_int, _P, _P, _P, _P, _P, _P]
        L.orc_scene_trace.argtypes = [_P, _P, _U64, C.c_int, _P, _P, _P, _P, C.c_int]
        L.orc_scene_render.argtypes = [_P, C.POINTER(CameraDesc), _F32, C.c_int, _U32, _U32, _P, _P, _P, _P,
                                       C.c_int, _U32, _U32]
        L.orc_cpu_render.argtypes = [_P, C.POINTER(CameraDesc), _F32, C.c_int, _U32, _U32, _P, _P, C.c_int,
                                     _U32, _U32, _U32, _P]
        L.orc_cpu_render.restype = _F64
        L.orc_set_clip_stack_cap.argtypes = [C.c_int]
        L.orc_scene_trace_ex.argtypes = [_P, _P, _U64, C.c_int, _P, _P, _P, _P, C.c_int, C.POINTER(TraceOpts)]
        L.orc_pixel_rays.argtypes = [C.POINTER(CameraDesc), _U32, _U32, _P, _U64, _P]
        L.orc_scene_shade.argtypes = [_P, _P, _U64, _F32, C.c_int, _P, _P, _P, _P, _P, _P, C.c_int]
        L.orc_scene_render_sampled.argtypes = [_P, C.POINTER(CameraDesc), _F32, C.c_int, _U32, _U32, _U32, _U32,
                                               _P, _P, _P, C.c_int, _U32, _U32]
        L.orc_philox.argtypes = [_P, _P, _U64, _P]
        L.orc_sample_words.argtypes = [_U32, _U32, _U32, _U64, _P, _P]
        L.orc_vdc.argtypes = [_P, _U64, _P]
        L.orc_sincos.argtypes = [_P, _U64, _P, _P]
        L.orc_atan2q.argtypes = [_P, _P, _U64, _P]
        L.orc_sampled_bounce.argtypes = [_P, _P, _P, _P, _U64, _P]
        _lib = L
    return _lib


def _ptr(a):
    return None if a is None else a.ctypes.data


# ---------------------------------------------------------------------------
# per-patch
# ---------------------------------------------------------------------------

def clip_batch(rays, patches, tol=2.0 ** -12, fp64=False, method=0, threshold=2.0 ** -4):
    """rays (n,8) f32, patches (n,16,4) f32 → (n,4) results (t,u,v,found), Work."""
    rays = np.ascontiguousarray(rays, np.float32)
    patches = np.ascontiguousarray(patches, np.float32)
    n = rays.shape[0]
    w = Work()
    if fp64:
        out = np.zeros((n, 4), np.float64)
        lib().orc_clip_batch(1, _ptr(rays), _ptr(patches), n, tol, None, _ptr(out), C.byref(w), method, threshold)
    else:
        out = np.zeros((n, 4), np.float32)
        lib().orc_clip_batch(0, _ptr(rays), _ptr(patches), n, tol, _ptr(out), None, C.byref(w), method, threshold)
    return out, w.asdict()


def set_clip_stack_cap(cap: int) -> None:
    """Caps the clipper's region stack on this thread like the GPU clipper
    (stop and count an overflow when a split would store more than `cap`
    regions); 0 restores the reference's unbounded stack."""
    lib().orc_set_clip_stack_cap(int(cap))


def clip2d_candidates(p2d, tol=2.0 ** -12, cap=256):
    p2d = np.ascontiguousarray(p2d, np.float32).reshape(16, 2)
    out = np.zeros((cap, 2), np.float32)
    w = Work()
    k = lib().orc_clip2d_candidates(_ptr(p2d), tol, _ptr(out), cap, C.byref(w))
    return out[:min(k, cap)], w.asdict()


EVAL_KINDS = {"interp_dc": 0, "interp_mat": 1, "du_dc": 2, "dv_dc": 3, "du_mat": 4, "dv_mat": 5,
              "normal_dc": 6, "normal_mat": 7}


def evaluate(patch, u, v, kind, fp64=False):
    p = np.ascontiguousarray(patch, np.float32).reshape(16, 4)
    out = np.zeros(4, np.float64)
    rc = lib().orc_eval(int(fp64), _ptr(p), float(u), float(v), EVAL_KINDS[kind], _ptr(out))
    assert rc == 0
    return out


def subpatch(patch, umin, umax, vmin, vmax):
    """Spline.h:384-396 on a (16, 4) float32 patch (subpatchV, then subpatchU)."""
    p = np.array(patch, np.float32).reshape(16, 4).copy()
    lib().orc_subpatch(_ptr(p), umin, umax, vmin, vmax)
    return p


def subpatch2d(p2d, side, a, b):
    p = np.array(p2d, np.float32).reshape(16, 2).copy()
    lib().orc_subpatch2d(_ptr(p), side, a, b)
    return p


def slice_matrix(a, b):
    ab = np.array([a, b], np.float32)
    m = np.zeros(16, np.float32)
    lib().orc_slice_matrix(_ptr(ab), _ptr(m))
    return m.reshape(4, 4)   # rows are columns of the reference's mat4


def x_axis_intersection(A, B):
    A = np.asarray(A, np.float32)
    B = np.asarray(B, np.float32)
    out = np.zeros(2, np.float32)
    ok = lib().orc_x_axis_intersection(_ptr(A), _ptr(B), _ptr(out))
    return bool(ok), out


def triangle_origin(P, A, B):
    P, A, B = (np.asarray(x, np.float32) for x in (P, A, B))
    out = np.zeros(2, np.float32)
    ok = lib().orc_triangle_origin(_ptr(P), _ptr(A), _ptr(B), _ptr(out))
    return bool(ok), out


DS_KINDS = {"versor": 0, "triangle": 1, "srgb": 2, "box": 3}


def ds(kind, *args):
    """The restated Ds boundary functions (orc_ds); returns 8 floats."""
    a = np.array(args, np.float32).ravel()
    out = np.zeros(8, np.float32)
    assert lib().orc_ds(DS_KINDS[kind], _ptr(a), _ptr(out)) == 0
    return out


PBR_KINDS = {"schlick": 0, "G": 1, "D": 2, "microfacet": 3, "light_color": 4, "light_vector": 5}


def pbr(kind, *args):
    a = np.array(args, np.float32).ravel()
    out = np.zeros(8, np.float32)
    assert lib().orc_pbr(PBR_KINDS[kind], _ptr(a), _ptr(out)) == 0
    return out


def fminf(a, b):
    return lib().orc_fminf(a, b)


def fmaxf(a, b):
    return lib().orc_fmaxf(a, b)


def newton(patch, o, d, uvt, iters=30):
    p = np.ascontiguousarray(patch, np.float32).reshape(16, 4)
    o = np.asarray(o, np.float64)
    d = np.asarray(d, np.float64)
    x = np.array(uvt, np.float64)
    ok = lib().orc_newton(_ptr(p), _ptr(o), _ptr(d), _ptr(x), iters)
    return bool(ok), x


def clip_batch64(rays64, patches, tol=2.0 ** -12):
    """FP64 clipping of FP64 rays (n,8) against patches (n,16,4) → (n,4) float64 (t,u,v,found)."""
    rays64 = np.ascontiguousarray(rays64, np.float64).reshape(-1, 8)
    patches = np.ascontiguousarray(patches, np.float32)
    n = rays64.shape[0]
    out = np.zeros((n, 4), np.float64)
    if n:
        lib().orc_clip_batch64(_ptr(rays64), _ptr(patches), n, tol, _ptr(out))
    return out


def newton_batch(patches, rays64, uvt, iters=30):
    """Polishes (u,v,t) seeds (n,3) to the FP64 root; returns (uvt, |N·d| cosine, converged)."""
    patches = np.ascontiguousarray(patches, np.float32)
    rays64 = np.ascontiguousarray(rays64, np.float64).reshape(-1, 8)
    x = np.array(uvt, np.float64).reshape(-1, 3).copy()
    n = x.shape[0]
    cos = np.zeros(n, np.float64)
    ok = np.zeros(n, np.int32)
    if n:
        lib().orc_newton_batch(_ptr(patches), _ptr(rays64), n, _ptr(x), _ptr(cos), _ptr(ok), iters)
    return x, cos, ok.astype(bool)


# ---------------------------------------------------------------------------
# BVH and hash
# ---------------------------------------------------------------------------

def _bvh_fetch(h):
    nn, ni = C.c_uint64(), C.c_uint64()
    lib().orc_bvh_sizes(h, C.byref(nn), C.byref(ni))
    nodes = np.zeros((nn.value, 16), np.uint32)
    keys = np.zeros(nn.value, np.uint32)
    idx = np.zeros(ni.value, np.uint32)
    lib().orc_bvh_copy(h, _ptr(nodes), _ptr(keys), _ptr(idx))
    lib().orc_bvh_free(h)
    return nodes, keys, idx


def bvh_build_patches(vertices, indices):
    v = np.ascontiguousarray(vertices, np.float32)
    i = np.ascontiguousarray(indices, np.uint32)
    h = lib().orc_bvh_build_patches(_ptr(v), _ptr(i), i.shape[0])
    assert h
    return _bvh_fetch(h)


def bvh_build_elements(elements10, elements_per_node=1):
    e = np.ascontiguousarray(elements10, np.float32)
    h = lib().orc_bvh_build_elements(_ptr(e), e.shape[0], elements_per_node)
    if not h:
        raise ValueError("invalid")
    return _bvh_fetch(h)


def phf_build(keys, seed=42):
    keys = np.ascontiguousarray(keys, np.uint32)
    p = np.zeros(5, np.uint32)
    D = np.zeros(lib().orc_phf_displacement_size(keys.shape[0]), np.uint32)
    rc = lib().orc_phf_build(_ptr(keys), keys.shape[0], seed, _ptr(p), _ptr(D))
    return rc, dict(A=int(p[0]), B=int(p[1]), mask=int(p[2]), table_size=int(p[3]), attempts=int(p[4])), D


# ---------------------------------------------------------------------------
# `.se` surfaces (Surface::load_se, Surface.cpp:214-348), the patch table
# (PatchTable::set, :36-148) and cg::PatchBVH's boxes (PatchBVH.cpp:40-80):
# an independent Python restatement, token by token
# ---------------------------------------------------------------------------

class SeError(ValueError):
    pass


def _se_tokens(text):
    """Whitespace-separated tokens, '#' comments removed (load_se's discard()
    runs before every item that may follow a comment; the reference's files
    only hold comments there)."""
    out = []
    for line in text.splitlines():
        k = line.find("#")
        out.extend((line if k < 0 else line[:k]).split())
    return out


def load_se(text):
    """Returns (points (n,4) f32, indices u32, matrices f32, patches (m,4):
    offset, matrix offset or -1, node count, type 1 B-spline / 2 Gregory)."""
    tok = _se_tokens(text.decode() if isinstance(text, (bytes, bytearray)) else text)
    pos = 0

    def take(kind, msg):
        nonlocal pos
        if pos >= len(tok):
            raise SeError(msg)
        t = tok[pos]
        pos += 1
        try:
            return int(t) if kind == "i" else np.float32(float(t))
        except ValueError:
            raise SeError(msg)

    n = take("i", "could not read no. points / nodes")
    pts = np.zeros((n, 4), np.float32)
    for i in range(n):
        if take("i", "unexpected point id") != i:
            raise SeError("unexpected point id")
        for c in range(4):
            pts[i, c] = take("f", "unexpected point id")
    m = take("i", "could not read no. elements")
    elements = []
    for _ in range(m):
        _id, typ, nodes, degree, boundary = (take("i", "read failed during element header") for _ in range(5))
        if degree != 3:
            raise SeError("degree value unsupported")
        if typ == 1:
            lines = 16 if boundary else 0
        elif typ == 2:
            lines = 20
        else:
            raise SeError("unknown patch type")
        idx = [take("i", "Invalid/corrupted file. Indices out of range") for _ in range(nodes)]
        if any(x >= n for x in idx):
            raise SeError("Invalid/corrupted file. Indices out of range")
        for _ in range(nodes):
            take("i", "read failed during element data")
        mat = np.zeros(lines * nodes, np.float32)
        for j in range(lines):
            for k in range(nodes):
                mat[j + k * lines] = take("f", "read failed during element data")
        elements.append((typ, idx, mat))
    groups = []            # (type, nodes, has matrix) in order of first appearance
    members = {}
    for e, (typ, idx, mat) in enumerate(elements):
        key = (typ, len(idx), mat.size > 0)
        if key not in members:
            groups.append(key)
            members[key] = []
        if mat.size and mat.size != (20 if typ == 2 else 16) * len(idx):
            raise SeError("invalid element matrix size")
        members[key].append(e)
    indices, matrices, refs = [], [], []
    for key in groups:
        for e in members[key]:
            typ, idx, mat = elements[e]
            refs.append((len(indices), len(matrices) if mat.size else -1, len(idx), typ))
            indices.extend(idx)
            matrices.extend(mat.tolist())
    return (pts, np.array(indices, np.uint32), np.array(matrices, np.float32),
            np.array(refs, np.int64).reshape(-1, 4))


def se_elements(points, indices, matrices, patches):
    """cg::PatchBVH's element boxes (PatchBVH.cpp:40-80), FP32 in the
    reference's order; the empty box for a patch whose 16-index leaf read runs
    past the table.  Returns (m, 10) float32 (min, max, centroid, index bits)."""
    f32 = np.float32
    out = np.zeros((patches.shape[0], 10), np.float32)
    for i, (off, moff, size, typ) in enumerate(patches):
        n = 20 if typ == 2 else 16
        if moff >= 0:
            pts = []
            for r in range(n):
                s = None
                for col in range(size):
                    w = matrices[moff + r + col * n]
                    term = w * points[indices[off + col]]
                    s = term if s is None else (s + term).astype(f32)
                pts.append(s[:3] / s[3])
            pts = np.array(pts, np.float32)
        else:
            pts = points[indices[off:off + n], :3]
        mn, mx = pts.min(axis=0), pts.max(axis=0)
        if off + 16 > indices.shape[0]:
            mn = np.full(3, np.finfo(np.float32).max, np.float32)
            mx = -mn
        out[i, 0:3], out[i, 3:6] = mn, mx
        out[i, 6:9] = f32(0.5) * (mn + mx)
        out[i, 9] = np.array([i], np.uint32).view(np.float32)[0]
    return out


def philox(ctr, key):
    """Philox4x32-10 blocks: ctr (n,4) u32, key (n,2) u32 -> (n,4) u32."""
    ctr = np.ascontiguousarray(ctr, np.uint32).reshape(-1, 4)
    key = np.ascontiguousarray(key, np.uint32).reshape(-1, 2)
    out = np.zeros_like(ctr)
    lib().orc_philox(_ptr(ctr), _ptr(key), ctr.shape[0], _ptr(out))
    return out


def sample_words(seed, stream, first, n):
    """Draws first .. first+n-1 of a stream: (words u32, floats in (0,1))."""
    words = np.zeros(n, np.uint32)
    units = np.zeros(n, np.float32)
    lib().orc_sample_words(seed, stream, first, n, _ptr(words), _ptr(units))
    return words, units


def vdc(i):
    i = np.ascontiguousarray(i, np.uint32)
    out = np.zeros(i.shape, np.float32)
    lib().orc_vdc(_ptr(i), i.size, _ptr(out))
    return out


def sincos(a):
    a = np.ascontiguousarray(a, np.float32)
    s, c = np.zeros_like(a), np.zeros_like(a)
    lib().orc_sincos(_ptr(a), a.size, _ptr(s), _ptr(c))
    return s, c


def atan2q(y, x):
    y = np.ascontiguousarray(y, np.float32)
    x = np.ascontiguousarray(x, np.float32)
    out = np.zeros_like(y)
    lib().orc_atan2q(_ptr(y), _ptr(x), y.size, _ptr(out))
    return out


def sampled_bounce(N, V, material, xy):
    """One sampled reflection per row: N, V (n,3), material (n,8) diffuse3,
    specular3, metalness, roughness; xy (n,2) draws -> (R (n,3), brdf (n,3),
    continued (n,) bool)."""
    N = np.ascontiguousarray(N, np.float32).reshape(-1, 3)
    V = np.ascontiguousarray(V, np.float32).reshape(-1, 3)
    material = np.ascontiguousarray(material, np.float32).reshape(-1, 8)
    xy = np.ascontiguousarray(xy, np.float32).reshape(-1, 2)
    out = np.zeros((N.shape[0], 7), np.float32)
    lib().orc_sampled_bounce(_ptr(N), _ptr(V), _ptr(material), _ptr(xy), N.shape[0], _ptr(out))
    return out[:, 0:3], out[:, 3:6], out[:, 6] == 1


def phf_build_wide(keys, seed=42):
    """The wide family (exact arithmetic, 64-bit keys): (status, params, D)."""
    keys = np.ascontiguousarray(keys, np.uint64)
    p = np.zeros(5, np.uint32)
    D = np.zeros(lib().orc_phf_displacement_size(keys.shape[0]), np.uint32)
    rc = lib().orc_phf_build_wide(_ptr(keys), keys.shape[0], seed, _ptr(p), _ptr(D))
    return rc, dict(A=int(p[0]), B=int(p[1]), mask=int(p[2]), table_size=int(p[3]), attempts=int(p[4])), D


def phf_hash_wide(params, D, keys):
    keys = np.ascontiguousarray(keys, np.uint64)
    p = np.array([params["A"], params["B"], params["mask"], params["table_size"]], np.uint32)
    D = np.ascontiguousarray(D, np.uint32)
    out = np.zeros(keys.shape[0], np.uint32)
    lib().orc_phf_hash_wide(_ptr(p), _ptr(D), _ptr(keys), keys.shape[0], _ptr(out))
    return out


# ---------------------------------------------------------------------------
# scenes
# ---------------------------------------------------------------------------

class OracleScene:
    """The oracle's copy of a scene: BVHs, hashes and tables built on the CPU."""

    def __init__(self, scene, surface0_elements=None):
        """surface0_elements: (npatch, 10) float32 BVH elements of surface 0
        (cg::PatchBVH's boxes of a patch table) instead of its patch boxes."""
        self._keep = [np.ascontiguousarray(a) for a in (
            scene.vertices, scene.indices, scene.surf_vbase, scene.surf_pbase, scene.obj_surface,
            scene.obj_local2world, scene.obj_world2local, scene.obj_material, scene.lights)]
        k = self._keep
        d = SceneDesc(_ptr(k[0]), _ptr(k[1]), _ptr(k[2]), _ptr(k[3]), scene.nsurf,
                      _ptr(k[4]), _ptr(k[5]), _ptr(k[6]), _ptr(k[7]), scene.nobj,
                      _ptr(k[8]), k[8].shape[0], (_F32 * 3)(*scene.background), (_F32 * 3)(*scene.ambient),
                      scene.hash_seed)
        st = C.c_int(0)
        if surface0_elements is None:
            self.h = lib().orc_scene_create(C.byref(d), C.byref(st))
        else:
            e = np.ascontiguousarray(surface0_elements, np.float32).reshape(-1, 10)
            self._keep.append(e)
            self.h = lib().orc_scene_create_elements(C.byref(d), _ptr(e), C.byref(st))
        self.status = st.value
        if not self.h:
            raise RuntimeError(f"oracle scene build failed (status {st.value})")

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_scene_destroy(self.h)
            self.h = None

    def bvh(self, surface):
        nn, ni, nd, nt = C.c_uint64(), C.c_uint64(), C.c_uint64(), C.c_uint64()
        lib().orc_scene_bvh_sizes(self.h, surface, C.byref(nn), C.byref(ni), C.byref(nd), C.byref(nt))
        nodes = np.zeros((nn.value, 16), np.uint32)
        keys = np.zeros(nn.value, np.uint32)
        idx = np.zeros(ni.value, np.uint32)
        p = np.zeros(5, np.uint32)
        D = np.zeros(nd.value, np.uint32)
        T = np.zeros(nt.value, np.uint32)
        lib().orc_scene_bvh_copy(self.h, surface, _ptr(nodes), _ptr(keys), _ptr(idx), _ptr(p), _ptr(D), _ptr(T))
        return dict(nodes=nodes, keys=keys, indices=idx, A=int(p[0]), B=int(p[1]), mask=int(p[2]),
                    table_size=int(p[3]), attempts=int(p[4]), displacement=D, table=T,
                    wide=bool(lib().orc_scene_bvh_wide(self.h, surface)))

    def trace(self, rays, any_hit=False, threads=None, stack=False, method=0, clip_tol=0.0, subdiv_depth=0,
              max_clip_iters=0, polish=False):
        """Closest (or any) hits of world rays, with cgs_trace_options' meaning:
        method 0 clips (clip_tol, 0: the reference's 2^-12), 1 subdivides
        (subdiv_depth D: threshold 2^(1-D), 0: the reference's 2^-4);
        max_clip_iters caps each clip; polish refines closest hits."""
        tol = clip_tol if clip_tol > 0 else 2.0 ** -12
        rays = np.ascontiguousarray(rays, np.float32)
        n = rays.shape[0]
        hits = np.zeros((n, 3), np.float32)
        patch = np.zeros(n, np.uint32)
        obj = np.zeros(n, np.int32)
        w = Work()
        th = threads or min(16, os.cpu_count() or 1)
        mode = 1 if any_hit else (2 if stack else 0)
        thr = 2.0 ** (1 - subdiv_depth) if subdiv_depth else 2.0 ** -4
        op = TraceOpts(int(method), float(tol), float(thr), int(max_clip_iters), int(bool(polish)))
        lib().orc_scene_trace_ex(self.h, _ptr(rays), n, mode, _ptr(hits), _ptr(patch), _ptr(obj), C.byref(w), th,
                                 C.byref(op))
        if any_hit:
            return obj.astype(bool), w.asdict()
        return hits, patch, obj, w.asdict()

    def pick(self, rays, any_hit=False, threads=None):
        """cg::PatchBVH's queries on surface 0 (orc_scene_pick): closest
        (t,u,v (n,3), patch (n,)), or occlusion (n,) bool."""
        rays = np.ascontiguousarray(rays, np.float32)
        n = rays.shape[0]
        hits = np.zeros((n, 3), np.float32)
        patch = np.zeros(n, np.uint32)
        occ = np.zeros(n, np.uint32)
        th = threads or min(16, os.cpu_count() or 1)
        lib().orc_scene_pick(self.h, _ptr(rays), n, 1 if any_hit else 0, _ptr(hits), _ptr(patch), _ptr(occ), th)
        return occ.astype(bool) if any_hit else (hits, patch)

    def record_patch_tests(self, rays, cap=1 << 22):
        """Every patch test of the closest-hit traces of `rays`, in call order:
        (clipper rays (k,8): local o, initial best t, local d; patches (k,16,4))."""
        rays = np.ascontiguousarray(rays, np.float32)
        ro = np.zeros((cap, 8), np.float32)
        po = np.zeros((cap, 16, 4), np.float32)
        k = lib().orc_scene_record_tests(self.h, _ptr(rays), rays.shape[0], _ptr(ro), _ptr(po), cap)
        k = min(k, cap)
        return ro[:k], po[:k]

    @staticmethod
    def pixel_rays(cam, xy):
        """Unit directions of the camera's rays through fractional pixel
        coordinates xy (n,2) of its width x height frame (orc_pixel_rays)."""
        xy = np.ascontiguousarray(xy, np.float32).reshape(-1, 2)
        out = np.zeros((xy.shape[0], 3), np.float32)
        c = OracleScene._cam(cam)
        lib().orc_pixel_rays(C.byref(c), cam.width, cam.height, _ptr(xy), xy.shape[0], _ptr(out))
        return out

    @staticmethod
    def _cam(cam):
        return CameraDesc((_F32 * 16)(*cam.to_world), (_F32 * 3)(*cam.position), cam.fovy, int(cam.flip_y))

    def render(self, cam, eps=1e-4, depth=0, rows=None, threads=None):
        """GPU-renderer semantics; returns (color (h,w,3), frame (h,w) u32, Work, ray counts)."""
        w, h = cam.width, cam.height
        r0, r1 = (0, h) if rows is None else rows
        color = np.zeros((h, w, 3), np.float32)
        frame = np.zeros((h, w), np.uint32)
        wk = Work()
        rc = np.zeros(3, np.uint64)
        th = threads or min(16, os.cpu_count() or 1)
        c = self._cam(cam)
        lib().orc_scene_render(self.h, C.byref(c), eps, depth, w, h, _ptr(color), _ptr(frame), C.byref(wk),
                               _ptr(rc), th, r0, r1)
        return color, frame, wk.asdict(), dict(primary=int(rc[0]), shadow=int(rc[1]), secondary=int(rc[2]))

    def render_sampled(self, cam, samples=8, seed=42, eps=1e-4, depth=0, rows=None, threads=None):
        """The sampled (Monte Carlo) render: returns (color (h,w,3), frame
        (h,w) u32, ray counts); rows outside `rows` stay zero."""
        w, h = cam.width, cam.height
        r0, r1 = (0, h) if rows is None else rows
        color = np.zeros((h, w, 3), np.float32)
        frame = np.zeros((h, w), np.uint32)
        rc = np.zeros(3, np.uint64)
        th = threads or min(16, os.cpu_count() or 1)
        c = self._cam(cam)
        st = lib().orc_scene_render_sampled(self.h, C.byref(c), eps, depth, samples, seed, w, h, _ptr(color),
                                            _ptr(frame), _ptr(rc), th, r0, r1)
        if st:
            raise ValueError("orc_scene_render_sampled: samples must be positive")
        return color, frame, dict(primary=int(rc[0]), shadow=int(rc[1]), secondary=int(rc[2]))

    def shade(self, rays, eps=1e-4, depth=0, threads=None):
        """The render's trace loop from given world rays (n,8): returns
        (color (n,3), packed pixels (n,), first-segment hits (n,3) t,u,v,
        patch, object, ray counts)."""
        rays = np.ascontiguousarray(rays, np.float32)
        n = rays.shape[0]
        color = np.zeros((n, 3), np.float32)
        px = np.zeros(n, np.uint32)
        hits = np.zeros((n, 3), np.float32)
        patch = np.zeros(n, np.uint32)
        obj = np.zeros(n, np.int32)
        rc = np.zeros(3, np.uint64)
        th = threads or min(16, os.cpu_count() or 1)
        lib().orc_scene_shade(self.h, _ptr(rays), n, eps, depth, _ptr(color), _ptr(px), _ptr(hits), _ptr(patch),
                              _ptr(obj), _ptr(rc), th)
        return color, px, hits, patch, obj, dict(primary=int(rc[0]), shadow=int(rc[1]), secondary=int(rc[2]))

    def cpu_render(self, cam, eps=1e-4, depth=0, rows=None, threads=None, stride=1):
        """CPURayTracer restatement over rows r0, r0+stride, ... < r1; returns
        (seconds, color, frame, rays, shadow rays)."""
        w, h = cam.width, cam.height
        r0, r1 = (0, h) if rows is None else rows
        color = np.zeros((h, w, 3), np.float32)
        frame = np.zeros((h, w), np.uint32)
        rc = np.zeros(2, np.uint64)
        th = threads or min(16, os.cpu_count() or 1)
        c = self._cam(cam)
        sec = lib().orc_cpu_render(self.h, C.byref(c), eps, depth, w, h, _ptr(color), _ptr(frame), th, r0, r1, stride,
                                   _ptr(rc))
        return sec, color, frame, int(rc[0]), int(rc[1])


# ---------------------------------------------------------------------------
# model files
# ---------------------------------------------------------------------------

_BE_INT = re.compile(rb"\s*\+?([0-9]+)")
_BE_FLOAT = re.compile(rb"\s*([+-]?(?:[0-9]+\.?[0-9]*|\.[0-9]+)(?:[eE][+-]?[0-9]+)?)")


class BeError(ValueError):
    pass


def load_be(text: bytes):
    """Surface::load_be (Surface.cpp:149-213) restated in Python: each
    `input >> x` skips whitespace and reads one number, each `input.ignore()`
    drops the next character.  Returns (points (n,4) f32 with w = 1, indices
    (p,16) u32 zero-based); raises BeError with the reference's messages,
    and on an index equal to the point count (which the reference's
    `points.size() < i` lets through).  Coordinates go decimal → float64 →
    float32, which equals the direct rounding unless the decimal lies within
    half a float64 ulp of a float32 midpoint."""
    pos = 0

    def num(rx, what):
        nonlocal pos
        m = rx.match(text, pos)
        if not m:
            raise BeError(what)
        pos = m.end()
        return m.group(1)

    npatch = int(num(_BE_INT, "Could not read the no. patches"))
    idx = np.zeros(16 * npatch, np.uint32)
    for k in range(16 * npatch):
        idx[k] = (int(num(_BE_INT, "Parsing patch failed")) - 1) & 0xFFFFFFFF
        pos += 1
    npts = int(num(_BE_INT, "Could not read the no. vertices"))
    pts = np.ones((npts, 4), np.float32)
    for i in range(npts):
        for c in range(3):
            pts[i, c] = np.float32(float(num(_BE_FLOAT, "Parsing point failed")))
            pos += 1
    if idx.size and int(idx.max()) >= npts:
        raise BeError("Invalid/corrupted file. Indices out of range")
    return pts, idx.reshape(-1, 16)
