"""cgspline_amd — MI355X-native ray–bicubic-Bézier-patch path of cgspline.

The product is ``libcgs.so`` (hand-written gfx950 HIP kernels behind the C ABI
in ``include/cgs.h``); this package is its Python binding plus the synthetic
scene generators.  There is no CPU execution path: on a machine without a HIP
device :class:`Context` raises.
"""
from __future__ import annotations

import ctypes as C
import dataclasses

import numpy as np

from . import _abi, scenes
from ._abi import CgsError, EMPTY, PATCH_OVERFLOW, check, lib, ptr

__all__ = ["Batch", "CLIPPING", "Context", "DeviceArray", "DeviceScene", "CgsError", "DlScene", "EMPTY", "PATCH_OVERFLOW",
           "PatchBVH", "SUBDIVISION", "load_be", "read_dl", "scenes", "lib", "trace_options"]


def load_be(source) -> tuple[np.ndarray, np.ndarray]:
    """Surface::load_be (Surface.cpp:149-213) through cgs_load_be: `source`
    is the file's text as bytes, or a path (str or os.PathLike).  Returns the
    control points (npoints, 4) float32 with w = 1 and the zero-based patch
    indices (npatches, 16) uint32, u fastest.  Malformed text raises CgsError
    with the reference's message."""
    if isinstance(source, (bytes, bytearray, memoryview)):
        text = bytes(source)
    else:
        with open(source, "rb") as f:
            text = f.read()
    vp, ip = C.c_void_p(), C.c_void_p()
    nv, npch = C.c_uint32(), C.c_uint32()
    check(lib().cgs_load_be(text, len(text), C.byref(vp), C.byref(nv), C.byref(ip), C.byref(npch)), "cgs_load_be")
    try:
        v = np.ctypeslib.as_array(C.cast(vp, C.POINTER(C.c_float)), (max(nv.value, 1) * 4,))[: nv.value * 4].copy()
        i = np.ctypeslib.as_array(C.cast(ip, C.POINTER(C.c_uint32)), (max(npch.value, 1) * 16,))[: npch.value * 16].copy()
    finally:
        lib().cgs_free_host(vp)
        lib().cgs_free_host(ip)
    return v.reshape(-1, 4), i.reshape(-1, 16)


@dataclasses.dataclass
class PatchTable:
    """A surface's patch table as Surface::load_se fills it (cgs_load_se):
    points (n, 4); node indices; basis matrices; per patch, in the table's
    group order, (offset, matrix offset or -1, node count, type)."""

    points: np.ndarray
    indices: np.ndarray
    matrices: np.ndarray
    patches: np.ndarray       # (npatches, 4) int64: offset, matrix_offset, size, type

    def _c(self):
        keep = (np.ascontiguousarray(self.points, np.float32), np.ascontiguousarray(self.indices, np.uint32),
                np.ascontiguousarray(self.matrices, np.float32),
                np.ascontiguousarray(self.patches.astype(np.int64).astype(np.uint32)))
        t = _abi.PatchTable(keep[0].ctypes.data_as(C.POINTER(C.c_float)), keep[0].shape[0],
                            keep[1].ctypes.data_as(C.POINTER(C.c_uint32)), keep[1].shape[0],
                            keep[2].ctypes.data_as(C.POINTER(C.c_float)), keep[2].shape[0],
                            keep[3].ctypes.data_as(C.POINTER(_abi.PatchRef)), self.patches.shape[0])
        return t, keep

    def elements(self) -> np.ndarray:
        """cgs_patch_table_elements: (npatches, 10) float32 view of the PatchBVH
        element boxes (bmin, bmax, centroid, index bits)."""
        t, keep = self._c()
        out = np.zeros((self.patches.shape[0], 10), np.float32)
        check(lib().cgs_patch_table_elements(C.byref(t), out.ctypes.data), "cgs_patch_table_elements")
        return out

    def patch_indices(self) -> np.ndarray:
        """cgs_patch_table_patches: (npatches, 16) the indices the leaf test clips."""
        t, keep = self._c()
        out = np.zeros((self.patches.shape[0], 16), np.uint32)
        check(lib().cgs_patch_table_patches(C.byref(t), out.ctypes.data), "cgs_patch_table_patches")
        return out


def load_se(source) -> PatchTable:
    """Surface::load_se (Surface.cpp:214-348) through cgs_load_se: `source`
    is the file's text as bytes, or a path."""
    if isinstance(source, (bytes, bytearray, memoryview)):
        text = bytes(source)
    else:
        with open(source, "rb") as f:
            text = f.read()
    tp = C.POINTER(_abi.PatchTable)()
    check(lib().cgs_load_se(text, len(text), C.byref(tp)), "cgs_load_se")
    try:
        t = tp.contents
        arr = lambda ptr, n, dt: np.ctypeslib.as_array(ptr, (max(n, 1),))[:n].astype(dt).copy()
        pts = arr(t.points, 4 * t.npoints, np.float32).reshape(-1, 4)
        idx = arr(t.indices, t.nindices, np.uint32)
        mat = arr(t.matrices, t.nmatrices, np.float32)
        refs = np.array([(t.patches[k].offset, t.patches[k].matrix_offset, t.patches[k].size, t.patches[k].type)
                         for k in range(t.npatches)], np.int64).reshape(-1, 4)
    finally:
        lib().cgs_patch_table_free(tp)
    return PatchTable(pts, idx, mat, refs)


@dataclasses.dataclass
class DlScene:
    """A scene read from a `.dl` file (cgs_dl_read): the scene, the file's
    view (camera→world, column-major, and position) and the names of its
    objects and lights."""
    scene: "scenes.Scene"
    view_to_world: np.ndarray
    view_position: np.ndarray
    object_names: list
    light_names: list


def _copy(p, ctype, n, shape):
    if n == 0:
        return np.zeros(tuple(0 if k < 0 else k for k in shape), np.dtype(ctype))
    return np.ctypeslib.as_array(C.cast(p, C.POINTER(ctype)), (n,)).copy().reshape(shape)


def read_dl(source, name: str = "<text>", working_dir=None) -> DlScene:
    """The reference's scene description language (reader/SceneReader.cpp)
    through cgs_dl_read: `source` is a path (str or os.PathLike; model paths
    resolve against its directory) or the text as bytes (model paths resolve
    against `working_dir`).  Errors raise CgsError with "file:line:col:
    message"."""
    out = C.POINTER(_abi.DlScene)()
    if isinstance(source, (bytes, bytearray, memoryview)):
        text = bytes(source)
        wd = None if working_dir is None else str(working_dir).encode()
        check(lib().cgs_dl_parse(text, len(text), name.encode(), wd, C.byref(out)), "cgs_dl_parse")
    else:
        import os
        check(lib().cgs_dl_read(os.fsencode(source), C.byref(out)), "cgs_dl_read")
    try:
        r = out.contents
        d = r.scene
        ns = d.nsurf
        vb = _copy(d.surf_vbase, C.c_uint32, ns + 1, (ns + 1,))
        pb = _copy(d.surf_pbase, C.c_uint32, ns + 1, (ns + 1,))
        sc = scenes.Scene(
            vertices=_copy(d.vertices, C.c_float, int(vb[-1]) * 4, (-1, 4)),
            indices=_copy(d.indices, C.c_uint32, int(pb[-1]) * 16, (-1, 16)),
            surf_vbase=vb, surf_pbase=pb,
            obj_surface=_copy(d.obj_surface, C.c_uint32, d.nobj, (d.nobj,)),
            obj_local2world=_copy(d.obj_local2world, C.c_float, d.nobj * 16, (-1, 16)),
            obj_world2local=_copy(d.obj_world2local, C.c_float, d.nobj * 16, (-1, 16)),
            obj_material=_copy(d.obj_material, C.c_float, d.nobj * 8, (-1, 8)),
            lights=_copy(d.lights, C.c_float, d.nlight * 12, (-1, 12)),
            background=tuple(float(x) for x in d.background), ambient=tuple(float(x) for x in d.ambient),
            hash_seed=int(d.hash_seed), name=str(name if isinstance(source, (bytes, bytearray, memoryview)) else source))
        res = DlScene(sc, np.array(r.view_to_world[:], np.float32), np.array(r.view_position[:], np.float32),
                      [r.object_names[k].decode() for k in range(d.nobj)],
                      [r.light_names[k].decode() for k in range(d.nlight)])
    finally:
        lib().cgs_dl_free(out)
    return res


CLIPPING = 0
SUBDIVISION = 1


def trace_options(method=CLIPPING, clip_tol=0.0, subdiv_depth=0, max_clip_iters=0, polish=False):
    """A cgs_trace_options record; every default (zero) is the reference's
    leaf call: clipping at 2^-12, no cap, no polish."""
    if method not in (CLIPPING, SUBDIVISION):
        raise ValueError(f"unknown leaf method {method!r}")
    return _abi.TraceOptions(int(method), float(clip_tol), int(subdiv_depth), int(max_clip_iters), int(bool(polish)))


class DeviceArray:
    """A device allocation owned by a :class:`Context` (cgs_malloc)."""

    def __init__(self, ctx: "Context", nbytes: int):
        self.ctx = ctx
        self.nbytes = int(nbytes)
        p = C.c_void_p()
        check(lib().cgs_malloc(ctx.h, max(1, self.nbytes), C.byref(p)), "cgs_malloc")
        self.dptr = p.value

    @classmethod
    def from_host(cls, ctx: "Context", a: np.ndarray) -> "DeviceArray":
        a = np.ascontiguousarray(a)
        d = cls(ctx, a.nbytes)
        if a.nbytes:
            check(lib().cgs_memcpy_h2d(ctx.h, d.dptr, a.ctypes.data, a.nbytes, None), "cgs_memcpy_h2d")
        return d

    def to_host(self, dtype, shape) -> np.ndarray:
        out = np.empty(shape, dtype)
        if out.nbytes:
            assert out.nbytes <= self.nbytes
            check(lib().cgs_memcpy_d2h(self.ctx.h, out.ctypes.data, self.dptr, out.nbytes, None), "cgs_memcpy_d2h")
        return out

    def free(self):
        if self.dptr is not None and self.ctx.h is not None:
            lib().cgs_free(self.ctx.h, self.dptr)
        self.dptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class Context:
    """One HIP device (cgs_open)."""

    def __init__(self, device: int = 0):
        p = C.c_void_p()
        check(lib().cgs_open(device, C.byref(p)), "cgs_open")
        self.h = p.value
        self.device = device

    def close(self):
        if self.h is not None:
            lib().cgs_close(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # --- per-patch entry points -------------------------------------------------
    def bezier_clip(self, rays: np.ndarray, patches: np.ndarray, tol: float = 2.0 ** -12) -> np.ndarray:
        """doBezierClipping of ray r against patch r; returns (n, 4) records
        (t, u, v as float32; patch word as uint32 view in column 3)."""
        return self._per_patch(lib().cgs_bezier_clip, rays, patches, tol)

    def subdivision(self, rays: np.ndarray, patches: np.ndarray, threshold: float = 2.0 ** -4) -> np.ndarray:
        return self._per_patch(lib().cgs_subdivision, rays, patches, threshold)

    def _per_patch(self, fn, rays, patches, param):
        rays = np.ascontiguousarray(rays, np.float32).reshape(-1, 8)
        patches = np.ascontiguousarray(patches, np.float32).reshape(-1, 64)
        n = rays.shape[0]
        assert patches.shape[0] == n
        dr, dp = DeviceArray.from_host(self, rays), DeviceArray.from_host(self, patches)
        dh = DeviceArray(self, n * 16)
        check(fn(self.h, dr.dptr, dp.dptr, n, param, dh.dptr, None))
        check(lib().cgs_stream_sync(self.h, None))
        return dh.to_host(np.float32, (n, 4))

    def patch_eval(self, kind: int, patches: np.ndarray, uv: np.ndarray) -> np.ndarray:
        """cgs_patch_eval: patch r (16 × 4 floats) at uv[r] by `kind`
        (CGS_EVAL_*: interpolate, derivatives, normal; de Casteljau or matrix
        form); returns (n, 4) float32."""
        patches = np.ascontiguousarray(patches, np.float32).reshape(-1, 64)
        uv = np.ascontiguousarray(uv, np.float32).reshape(-1, 2)
        n = patches.shape[0]
        assert uv.shape[0] == n
        dp, du = DeviceArray.from_host(self, patches), DeviceArray.from_host(self, uv)
        do = DeviceArray(self, n * 16)
        check(lib().cgs_patch_eval(self.h, int(kind), dp.dptr, du.dptr, n, do.dptr, None), "cgs_patch_eval")
        self.sync()
        return do.to_host(np.float32, (n, 4))

    def subpatch(self, patches: np.ndarray, ranges: np.ndarray) -> np.ndarray:
        """cgs_patch_subpatch: patch r cut to ranges[r] = (umin, umax, vmin,
        vmax); returns (n, 16, 4) float32."""
        patches = np.ascontiguousarray(patches, np.float32).reshape(-1, 64)
        ranges = np.ascontiguousarray(ranges, np.float32).reshape(-1, 4)
        n = patches.shape[0]
        assert ranges.shape[0] == n
        dp, dr = DeviceArray.from_host(self, patches), DeviceArray.from_host(self, ranges)
        do = DeviceArray(self, n * 256)
        check(lib().cgs_patch_subpatch(self.h, dp.dptr, dr.dptr, n, do.dptr, None), "cgs_patch_subpatch")
        self.sync()
        return do.to_host(np.float32, (n, 16, 4))

    def patch_bvh(self, vertices: np.ndarray, indices: np.ndarray) -> "PatchBVH":
        return PatchBVH(self, vertices, indices)

    # --- BVH / hash -------------------------------------------------------------
    @staticmethod
    def patch_elements(vertices: np.ndarray, indices: np.ndarray) -> np.ndarray:
        v = np.ascontiguousarray(vertices, np.float32)
        i = np.ascontiguousarray(indices, np.uint32)
        out = np.zeros((i.shape[0], 10), np.float32)
        check(lib().cgs_patch_elements(v.ctypes.data, i.ctypes.data, i.shape[0], out.ctypes.data))
        return out

    @staticmethod
    def mesh_elements(vertices3: np.ndarray, triangles: np.ndarray) -> np.ndarray:
        """Mesh::buildBVH element data (cgs_mesh_elements): (ntri, 10) float32
        rows of bounds min, max, centroid and the triangle index."""
        v = np.ascontiguousarray(vertices3, np.float32).reshape(-1, 3)
        t = np.ascontiguousarray(triangles, np.uint32).reshape(-1, 3)
        out = np.zeros((t.shape[0], 10), np.float32)
        check(lib().cgs_mesh_elements(v.ctypes.data, v.shape[0], t.ctypes.data, t.shape[0], out.ctypes.data),
              "cgs_mesh_elements")
        return out

    @staticmethod
    def bvh_build(elements10: np.ndarray, elements_per_node: int = 1):
        """Host BVH build; returns (nodes (N,16) u32 words, keys, indices)."""
        e = np.ascontiguousarray(elements10, np.float32)
        n_, k_, i_ = C.c_void_p(), C.c_void_p(), C.c_void_p()
        nc, ic = C.c_uint32(), C.c_uint32()
        check(lib().cgs_bvh_build(e.ctypes.data, e.shape[0], elements_per_node, C.byref(n_), C.byref(k_),
                                  C.byref(nc), C.byref(i_), C.byref(ic)), "cgs_bvh_build")
        try:
            nodes = np.ctypeslib.as_array((C.c_uint32 * (16 * nc.value)).from_address(n_.value)).reshape(-1, 16).copy()
            keys = np.ctypeslib.as_array((C.c_uint32 * nc.value).from_address(k_.value)).copy()
            idx = np.ctypeslib.as_array((C.c_uint32 * ic.value).from_address(i_.value)).copy()
        finally:
            for p in (n_, k_, i_):
                lib().cgs_free_host(p)
        return nodes, keys, idx

    def phf_build(self, keys: np.ndarray, seed: int = 42):
        """GPU perfect-hash build; returns (status, params dict, displacement)."""
        keys = np.ascontiguousarray(keys, np.uint32)
        n = keys.shape[0]
        dk = DeviceArray.from_host(self, keys)
        m = lib().cgs_phf_displacement_size(n)
        dd = DeviceArray(self, 4 * m)
        h = _abi.PHF()
        st = lib().cgs_phf_build(self.h, dk.dptr, n, seed, dd.dptr, C.byref(h), None)
        check(lib().cgs_stream_sync(self.h, None))
        params = dict(A=h.A, B=h.B, mask=h.mask, table_size=h.table_size, attempts=h.attempts)
        return st, params, dd.to_host(np.uint32, (m,))

    def phf_hash(self, params: dict, displacement: np.ndarray, keys: np.ndarray) -> np.ndarray:
        keys = np.ascontiguousarray(keys, np.uint32)
        h = _abi.PHF(params["A"], params["B"], params["mask"], params["table_size"], 0, displacement.shape[0])
        dk, dd = DeviceArray.from_host(self, keys), DeviceArray.from_host(self, displacement.astype(np.uint32))
        out = DeviceArray(self, 4 * keys.shape[0])
        check(lib().cgs_phf_hash(self.h, C.byref(h), dd.dptr, dk.dptr, keys.shape[0], out.dptr, None))
        check(lib().cgs_stream_sync(self.h, None))
        return out.to_host(np.uint32, keys.shape)

    def phf_build_wide(self, keys: np.ndarray, seed: int = 42):
        """GPU build of the wide family over 64-bit keys (cgs_phf_build_wide);
        returns (status, params dict, displacement)."""
        keys = np.ascontiguousarray(keys, np.uint64)
        n = keys.shape[0]
        dk = DeviceArray.from_host(self, keys)
        m = lib().cgs_phf_displacement_size(n)
        dd = DeviceArray(self, 4 * m)
        h = _abi.PHF()
        st = lib().cgs_phf_build_wide(self.h, dk.dptr, n, seed, dd.dptr, C.byref(h), None)
        check(lib().cgs_stream_sync(self.h, None))
        params = dict(A=h.A, B=h.B, mask=h.mask, table_size=h.table_size, attempts=h.attempts)
        return st, params, dd.to_host(np.uint32, (m,))

    def phf_hash_wide(self, params: dict, displacement: np.ndarray, keys: np.ndarray) -> np.ndarray:
        keys = np.ascontiguousarray(keys, np.uint64)
        h = _abi.PHF(params["A"], params["B"], params["mask"], params["table_size"], 0, displacement.shape[0])
        dk, dd = DeviceArray.from_host(self, keys), DeviceArray.from_host(self, displacement.astype(np.uint32))
        out = DeviceArray(self, 4 * keys.shape[0])
        check(lib().cgs_phf_hash_wide(self.h, C.byref(h), dd.dptr, dk.dptr, keys.shape[0], out.dptr, None))
        check(lib().cgs_stream_sync(self.h, None))
        return out.to_host(np.uint32, keys.shape)

    def phf_fill_table_wide(self, params: dict, displacement: np.ndarray, keys: np.ndarray) -> np.ndarray:
        keys = np.ascontiguousarray(keys, np.uint64)
        h = _abi.PHF(params["A"], params["B"], params["mask"], params["table_size"], 0, displacement.shape[0])
        dk, dd = DeviceArray.from_host(self, keys), DeviceArray.from_host(self, displacement.astype(np.uint32))
        out = DeviceArray(self, 4 * params["table_size"])
        check(lib().cgs_phf_fill_table_wide(self.h, C.byref(h), dd.dptr, dk.dptr, keys.shape[0], out.dptr, None))
        check(lib().cgs_stream_sync(self.h, None))
        return out.to_host(np.uint32, (params["table_size"],))

    # --- scenes -------------------------------------------------------------------
    def scene(self, scene: "scenes.Scene", hash_retries: int = 0) -> "DeviceScene":
        """Makes `scene` resident.  Like the reference's BVH::buildHashTable, a
        perfect hash that cannot be built raises; with hash_retries > 0 the
        build is retried with the next hash seeds (PerfectHashFunction(seed))."""
        for k in range(hash_retries + 1):
            try:
                return DeviceScene(self, scene if k == 0 else dataclasses.replace(scene, hash_seed=scene.hash_seed + k))
            except CgsError as e:
                if e.status != 3 or k == hash_retries:
                    raise

    def sync(self, stream=None):
        check(lib().cgs_stream_sync(self.h, stream), "cgs_stream_sync")

    def create_stream(self) -> int:
        """A non-blocking HIP stream on this context's device (cgs_stream_create)."""
        s = C.c_void_p()
        check(lib().cgs_stream_create(self.h, C.byref(s)), "cgs_stream_create")
        return s.value

    def destroy_stream(self, stream: int):
        check(lib().cgs_stream_destroy(self.h, stream), "cgs_stream_destroy")

    # --- batched queries over resident scenes -----------------------------------
    @staticmethod
    def batch_jobs(jobs) -> "C.Array":
        """ctypes array of cgs_batch_job from dicts: scene (DeviceScene), rays
        (device), n, and hits/objects (closest) or occluded (any=True)."""
        arr = (_abi.BatchJob * max(1, len(jobs)))()
        for k, j in enumerate(jobs):
            arr[k] = _abi.BatchJob(j["scene"].h, ptr(j["rays"]), j["n"], 1 if j.get("any") else 0,
                                   ptr(j.get("hits")), ptr(j.get("objects")), ptr(j.get("occluded")))
        return arr

    def trace_batch(self, jobs, stream=None):
        """One persistent launch over every job (cgs_trace_batch)."""
        arr = self.batch_jobs(jobs)
        check(lib().cgs_trace_batch(self.h, arr, len(jobs), stream), "cgs_trace_batch")

    def batch(self, jobs) -> "Batch":
        """A fixed batch captured in a HIP graph (cgs_batch_create)."""
        return Batch(self, jobs)

    def batch_config(self) -> tuple[int, int]:
        c = _abi.BatchConfig()
        check(lib().cgs_batch_get_config(self.h, C.byref(c)), "cgs_batch_get_config")
        return c.chunk, c.blocks_per_cu

    def set_batch_config(self, chunk: int, blocks_per_cu: int = 0):
        c = _abi.BatchConfig(chunk, blocks_per_cu)
        check(lib().cgs_batch_set_config(self.h, C.byref(c)), "cgs_batch_set_config")

    def batch_autotune(self, jobs) -> tuple[int, int]:
        arr = self.batch_jobs(jobs)
        c = _abi.BatchConfig()
        check(lib().cgs_batch_autotune(self.h, arr, len(jobs), C.byref(c)), "cgs_batch_autotune")
        return c.chunk, c.blocks_per_cu


    def set_batch_trace_options(self, **options):
        """Leaf method of later batched launches (trace_options' keywords;
        none: the reference's clipping)."""
        op = trace_options(**options) if options else None
        check(lib().cgs_batch_set_trace_options(self.h, C.byref(op) if op is not None else None),
              "cgs_batch_set_trace_options")

    def batch_trace_options(self) -> dict:
        op = _abi.TraceOptions()
        check(lib().cgs_batch_get_trace_options(self.h, C.byref(op)), "cgs_batch_get_trace_options")
        return _options_dict(op)

    def batch_autotune_accuracy(self, jobs, max_rel_t=1e-4, max_uv=1e-4, max_disagree=0.0, allow_polish=True):
        """Accuracy-gated tuning of the batched leaf method
        (cgs_batch_autotune_accuracy): returns (the kept options, a list of
        every candidate's options, ms, max_rel_t, max_uv, disagree,
        admissible)."""
        arr = self.batch_jobs(jobs)
        lim = _abi.AccuracyLimits(max_rel_t, max_uv, max_disagree, int(bool(allow_polish)))
        best = _abi.TraceOptions()
        rep = (_abi.AutotuneEntry * 64)()
        cnt = C.c_uint32()
        check(lib().cgs_batch_autotune_accuracy(self.h, arr, len(jobs), C.byref(lim), C.byref(best), rep, 64,
                                                C.byref(cnt)), "cgs_batch_autotune_accuracy")
        report = [dict(options=_options_dict(e.options), ms=e.ms, max_rel_t=e.max_rel_t, max_uv=e.max_uv,
                       disagree=e.disagree, admissible=bool(e.admissible)) for e in rep[:cnt.value]]
        return _options_dict(best), report


def _options_dict(op) -> dict:
    return dict(method=int(op.method), clip_tol=float(op.clip_tol), subdiv_depth=int(op.subdiv_depth),
                max_clip_iters=int(op.max_clip_iters), polish=bool(op.polish))


class Batch:
    """A batched launch captured once in a HIP graph and replayed."""

    def __init__(self, ctx: Context, jobs):
        self.ctx = ctx
        arr = ctx.batch_jobs(jobs)
        p = C.c_void_p()
        check(lib().cgs_batch_create(ctx.h, arr, len(jobs), C.byref(p)), "cgs_batch_create")
        self.h = p.value

    def launch(self, stream=None):
        check(lib().cgs_batch_launch(self.h, stream), "cgs_batch_launch")

    def destroy(self):
        if self.h is not None and self.ctx.h is not None:
            lib().cgs_batch_destroy(self.h)
        self.h = None

    def __del__(self):
        try:
            self.destroy()
        except Exception:
            pass


class PatchBVH:
    """cg::PatchBVH over one surface's Bézier patches (cgs_patchbvh_create):
    closest and any hits of rays in the surface's frame."""

    def __init__(self, ctx: Context, vertices: np.ndarray, indices: np.ndarray, elements: np.ndarray = None):
        v = np.ascontiguousarray(vertices, np.float32).reshape(-1, 4)
        i = np.ascontiguousarray(indices, np.uint32).reshape(-1, 16)
        p = C.c_void_p()
        if elements is None:
            check(lib().cgs_patchbvh_create(ctx.h, v.ctypes.data, v.shape[0], i.ctypes.data, i.shape[0], C.byref(p)),
                  "cgs_patchbvh_create")
        else:
            e = np.ascontiguousarray(elements, np.float32).reshape(-1, 10)
            check(lib().cgs_patchbvh_create_elements(ctx.h, v.ctypes.data, v.shape[0], i.ctypes.data, i.shape[0],
                                                     e.ctypes.data, C.byref(p)), "cgs_patchbvh_create_elements")
        self.ctx, self.h = ctx, p.value

    @classmethod
    def from_table(cls, ctx: Context, table: PatchTable) -> "PatchBVH":
        """cg::PatchBVH(const Surface*) of a `.se` surface: boxes from the
        table's basis matrices, leaf tests on each patch's 16 indices."""
        return cls(ctx, table.points, table.patch_indices(), table.elements())

    def intersect(self, rays: np.ndarray):
        """Closest hits of host rays (n, 8): (t,u,v (n,3), patch (n,))."""
        rays = np.ascontiguousarray(rays, np.float32).reshape(-1, 8)
        n = rays.shape[0]
        dr, dh = DeviceArray.from_host(self.ctx, rays), DeviceArray(self.ctx, 16 * n)
        check(lib().cgs_patchbvh_intersect(self.ctx.h, self.h, dr.dptr, n, dh.dptr, None), "cgs_patchbvh_intersect")
        self.ctx.sync()
        h = dh.to_host(np.float32, (n, 4))
        return h[:, :3].copy(), h.view(np.uint32)[:, 3].copy()

    def intersect_any(self, rays: np.ndarray) -> np.ndarray:
        rays = np.ascontiguousarray(rays, np.float32).reshape(-1, 8)
        n = rays.shape[0]
        dr, do = DeviceArray.from_host(self.ctx, rays), DeviceArray(self.ctx, 4 * n)
        check(lib().cgs_patchbvh_intersect_any(self.ctx.h, self.h, dr.dptr, n, do.dptr, None),
              "cgs_patchbvh_intersect_any")
        self.ctx.sync()
        return do.to_host(np.uint32, (n,))

    def destroy(self):
        if self.h is not None:
            lib().cgs_patchbvh_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.destroy()
        except Exception:
            pass


class DeviceScene:
    """A scene resident in HBM (cgs_scene_create).  Any context on the same
    device may trace or render it: `via(ctx)` gives a view through another
    context."""

    def via(self, ctx: Context) -> "DeviceScene":
        """This resident scene used through another context (no copy)."""
        v = object.__new__(DeviceScene)
        v.ctx, v.scene, v.h, v._owner = ctx, self.scene, self.h, self
        return v

    def __init__(self, ctx: Context, scene: "scenes.Scene"):
        self.ctx = ctx
        self.scene = scene
        keep = [np.ascontiguousarray(a) for a in (
            scene.vertices, scene.indices, scene.surf_vbase, scene.surf_pbase, scene.obj_surface,
            scene.obj_local2world, scene.obj_world2local, scene.obj_material, scene.lights)]
        d = _abi.SceneDesc(keep[0].ctypes.data, keep[1].ctypes.data, keep[2].ctypes.data, keep[3].ctypes.data,
                           scene.nsurf, keep[4].ctypes.data, keep[5].ctypes.data, keep[6].ctypes.data,
                           keep[7].ctypes.data, scene.nobj, keep[8].ctypes.data, keep[8].shape[0],
                           (C.c_float * 3)(*scene.background), (C.c_float * 3)(*scene.ambient), scene.hash_seed)
        p = C.c_void_p()
        check(lib().cgs_scene_create(ctx.h, C.byref(d), C.byref(p)), "cgs_scene_create")
        self.h = p.value

    def bvh_culled(self, surface: int) -> bool:
        """Whether that BVH (-1: the top level) has a cull tree."""
        r = lib().cgs_scene_bvh_culled(self.h, surface)
        if r < 0:
            raise ValueError(f"no BVH {surface}")
        return r == 1

    def bvh_wide(self, surface: int) -> bool:
        """Whether that BVH (-1: the top level) is hashed with the wide family."""
        r = lib().cgs_scene_bvh_wide(self.h, surface)
        if r < 0:
            raise ValueError(f"no BVH {surface}")
        return r == 1

    @property
    def culled(self) -> bool:
        """Whether the queries find objects through the cull tree (cgs_scene_culled)."""
        return lib().cgs_scene_culled(self.h) == 1

    def destroy(self):
        if getattr(self, "_owner", None) is not None:   # a view: the owner destroys
            self.h = None
            return
        if self.h is not None and self.ctx.h is not None:
            lib().cgs_scene_destroy(self.h)
        self.h = None

    def __del__(self):
        try:
            self.destroy()
        except Exception:
            pass

    def bvh_info(self, surface: int) -> tuple[int, int, dict]:
        """(node count, index count, hash parameters) of that BVH (-1: the top level)."""
        nc, ic, h = C.c_uint32(), C.c_uint32(), _abi.PHF()
        check(lib().cgs_scene_bvh_info(self.h, surface, C.byref(nc), C.byref(ic), C.byref(h)))
        return nc.value, ic.value, dict(A=h.A, B=h.B, mask=h.mask, table_size=h.table_size, attempts=h.attempts)

    def bvh(self, surface: int) -> dict:
        nc, ic, h = C.c_uint32(), C.c_uint32(), _abi.PHF()
        check(lib().cgs_scene_bvh_info(self.h, surface, C.byref(nc), C.byref(ic), C.byref(h)))
        nodes = np.zeros((nc.value, 16), np.uint32)
        keys = np.zeros(nc.value, np.uint32)
        idx = np.zeros(ic.value, np.uint32)
        D = np.zeros(h.displacement_size, np.uint32)
        T = np.zeros(h.table_size, np.uint32)
        check(lib().cgs_scene_bvh_copy(self.h, surface, nodes.ctypes.data, keys.ctypes.data, idx.ctypes.data,
                                       D.ctypes.data, T.ctypes.data))
        return dict(nodes=nodes, keys=keys, indices=idx, A=h.A, B=h.B, mask=h.mask, table_size=h.table_size,
                    attempts=h.attempts, displacement=D, table=T, wide=self.bvh_wide(surface))

    # device-pointer entry points (for resident buffers)
    def trace_closest_dev(self, rays, n, hits, objects=None, stream=None, **options):
        """cgs_trace_closest, or cgs_trace_closest_ex with trace_options(**options)."""
        if options:
            check(lib().cgs_trace_closest_ex(self.ctx.h, self.h, ptr(rays), n, ptr(hits), ptr(objects),
                                             C.byref(trace_options(**options)), stream), "cgs_trace_closest_ex")
        else:
            check(lib().cgs_trace_closest(self.ctx.h, self.h, ptr(rays), n, ptr(hits), ptr(objects), stream),
                  "cgs_trace_closest")

    def trace_any_dev(self, rays, n, occluded, stream=None, **options):
        """cgs_trace_any, or cgs_trace_any_ex with trace_options(**options)."""
        if options:
            check(lib().cgs_trace_any_ex(self.ctx.h, self.h, ptr(rays), n, ptr(occluded),
                                         C.byref(trace_options(**options)), stream), "cgs_trace_any_ex")
        else:
            check(lib().cgs_trace_any(self.ctx.h, self.h, ptr(rays), n, ptr(occluded), stream), "cgs_trace_any")

    def render_dev(self, camera: "scenes.Camera", frame, color=None, counts=None, eps=1e-4, depth=0, stream=None,
                   stride=0):
        """cgs_render into device buffers; `stride` is the frame's row pitch in
        pixels (0: the width), pixel (x, y) at frame[stride*y + x]."""
        cam = _abi.Camera((C.c_float * 16)(*camera.to_world), (C.c_float * 3)(*camera.position), camera.fovy,
                          int(camera.flip_y))
        op = _abi.RenderOptions(eps, depth, stride)
        return lib().cgs_render(self.ctx.h, self.h, C.byref(cam), C.byref(op), camera.width, camera.height,
                                ptr(frame), ptr(color), ptr(counts), stream)

    def render_sampled_dev(self, camera: "scenes.Camera", frame, color=None, counts=None, samples=8, seed=42, eps=1e-4,
                           depth=0, stream=None, stride=0):
        """cgs_render_sampled into device buffers: the Monte Carlo render with
        `samples` camera rays per pixel, draws from `seed`."""
        cam = _abi.Camera((C.c_float * 16)(*camera.to_world), (C.c_float * 3)(*camera.position), camera.fovy,
                          int(camera.flip_y))
        op = _abi.RenderOptions(eps, depth, stride)
        return lib().cgs_render_sampled(self.ctx.h, self.h, C.byref(cam), C.byref(op), samples, seed, camera.width,
                                        camera.height, ptr(frame), ptr(color), ptr(counts), stream)

    # host convenience wrappers
    def trace_closest(self, rays: np.ndarray, **options):
        """Closest hits of host rays (n, 8): returns (t,u,v (n,3), patch (n,), object (n,)).
        options: trace_options' keywords (method, clip_tol, subdiv_depth,
        max_clip_iters, polish)."""
        rays = np.ascontiguousarray(rays, np.float32).reshape(-1, 8)
        n = rays.shape[0]
        dr = DeviceArray.from_host(self.ctx, rays)
        dh = DeviceArray(self.ctx, 16 * n)
        do = DeviceArray(self.ctx, 4 * n)
        self.trace_closest_dev(dr, n, dh, do, **options)
        self.ctx.sync()
        h = dh.to_host(np.float32, (n, 4))
        return h[:, :3].copy(), h.view(np.uint32)[:, 3].copy(), do.to_host(np.int32, (n,))

    def trace_any(self, rays: np.ndarray, **options) -> np.ndarray:
        """Occlusion of host rays (n, 8): 1 / 0 (2: clip-stack overflow)."""
        rays = np.ascontiguousarray(rays, np.float32).reshape(-1, 8)
        n = rays.shape[0]
        dr = DeviceArray.from_host(self.ctx, rays)
        do = DeviceArray(self.ctx, 4 * n)
        self.trace_any_dev(dr, n, do, **options)
        self.ctx.sync()
        return do.to_host(np.uint32, (n,))

    def shade_dev(self, rays, n, color, pixels=None, hits=None, objects=None, counts=None, eps=1e-4, depth=0,
                  stream=None):
        """cgs_shade into device buffers."""
        op = _abi.RenderOptions(eps, depth, 0)
        check(lib().cgs_shade(self.ctx.h, self.h, ptr(rays), n, C.byref(op), ptr(color), ptr(pixels), ptr(hits),
                              ptr(objects), ptr(counts), stream), "cgs_shade")

    def shade(self, rays: np.ndarray, eps=1e-4, depth=0):
        """Host rays (n, 8) traced and shaded as the render shades camera rays:
        returns (color (n,3), packed pixels (n,), first hits t,u,v (n,3),
        patch (n,), object (n,), counts dict)."""
        rays = np.ascontiguousarray(rays, np.float32).reshape(-1, 8)
        n = rays.shape[0]
        dr = DeviceArray.from_host(self.ctx, rays)
        dc = DeviceArray(self.ctx, 12 * n)
        dp = DeviceArray(self.ctx, 4 * n)
        dh = DeviceArray(self.ctx, 16 * n)
        do = DeviceArray(self.ctx, 4 * n)
        dn = DeviceArray.from_host(self.ctx, np.zeros(4, np.uint64))
        self.shade_dev(dr, n, dc, dp, dh, do, dn, eps, depth)
        self.ctx.sync()
        h = dh.to_host(np.float32, (n, 4))
        cnt = dn.to_host(np.uint64, (4,))
        return (dc.to_host(np.float32, (n, 3)), dp.to_host(np.uint32, (n,)), h[:, :3].copy(),
                h.view(np.uint32)[:, 3].copy(), do.to_host(np.int32, (n,)),
                dict(primary=int(cnt[0]), shadow=int(cnt[1]), secondary=int(cnt[2]), overflow=int(cnt[3])))

    def render(self, camera: "scenes.Camera", eps=1e-4, depth=0, samples=None, seed=42):
        """Returns (color (h,w,3) f32, frame (h,w) u32, counts dict).
        samples: None for the plain render, n for the sampled (Monte Carlo)
        render with n camera rays per pixel and draws from `seed`."""
        w, h = camera.width, camera.height
        df = DeviceArray(self.ctx, 4 * w * h)
        dc = DeviceArray(self.ctx, 12 * w * h)
        dn = DeviceArray.from_host(self.ctx, np.zeros(4, np.uint64))
        if samples is None:
            check(self.render_dev(camera, df, dc, dn, eps, depth), "cgs_render")
        else:
            check(self.render_sampled_dev(camera, df, dc, dn, samples, seed, eps, depth), "cgs_render_sampled")
        self.ctx.sync()
        cnt = dn.to_host(np.uint64, (4,))
        return (dc.to_host(np.float32, (h, w, 3)), df.to_host(np.uint32, (h, w)),
                dict(primary=int(cnt[0]), shadow=int(cnt[1]), secondary=int(cnt[2]), overflow=int(cnt[3])))
