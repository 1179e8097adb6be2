"""FP64 accuracy of the closest-hit query (test infrastructure).

Three measurements per workload and ray sample, for the FP32 path (the GPU
path equals the FP32 oracle bit for bit, which the GPU parity tests check):

(ii) against the same algorithm in FP64: every patch whose control-point box
     the FP64 object-local ray crosses is clipped in FP64 (Bézier clipping,
     tolerance 2^-12, rt/Primitive.cu:222-248 + Spline_BezierClipping.cu) and
     the nearest world t kept (RayTracer.cu:296-350 scaling t by 1/|M d|);
(iii) against the true root: the FP64 hit is polished with Newton on
     S(u,v) - O - t·d = 0 (the clipper's method error).

Flags, from the FP64 run: grazing rays (|N̂·d̂| < 1e-3 at the hit) and seam
rays (a second patch hit within 1e-5 relative in t).  Only flagged rays may
disagree on hit/miss or patch id.

Run as a script to print the table of every workload:
    python tests/accuracy.py
"""
from __future__ import annotations

import dataclasses
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
for p in (ROOT, os.path.join(ROOT, "oracle")):
    if p not in sys.path:
        sys.path.insert(0, p)

GRAZING_COS = 1e-3
SEAM_REL = 1e-5


@dataclasses.dataclass
class Fp64Hits:
    t: np.ndarray        # world t (inf = miss)
    u: np.ndarray
    v: np.ndarray
    obj: np.ndarray      # -1 = miss
    patch: np.ndarray    # patch id within the object's surface
    t2: np.ndarray       # nearest world t of any other (object, patch)
    local: np.ndarray    # (n, 8) object-local FP64 ray of the hit object (zeros on a miss)
    ell: np.ndarray      # |M d| of the hit object (local t = world t · ell)


def _local_rays(scene, o, rays64):
    M = scene.obj_world2local[o].astype(np.float64).reshape(4, 4).T      # column-major storage
    O4 = rays64[:, 0:3] @ M[:3, :3].T + M[:3, 3]
    w = rays64[:, 0:3] @ M[3, :3] + M[3, 3]
    org = O4 / w[:, None]
    d = rays64[:, 4:7] @ M[:3, :3].T
    ell = np.linalg.norm(d, axis=1)
    out = np.zeros_like(rays64)
    out[:, 0:3] = org
    out[:, 3] = rays64[:, 3] * ell
    out[:, 4:7] = d / ell[:, None]
    return out, ell


def fp64_closest(scene, rays, chunk=2048):
    """Closest hit of world rays (n,8 float32) by FP64 clipping of every
    candidate patch of every object."""
    import cgs_oracle
    rays64 = np.asarray(rays, np.float64).copy()
    rays64[:, 4:7] /= np.linalg.norm(rays64[:, 4:7], axis=1, keepdims=True)
    n = rays64.shape[0]
    best = np.full(n, np.inf)
    second = np.full(n, np.inf)
    bu = np.zeros(n)
    bv = np.zeros(n)
    bobj = np.full(n, -1, np.int64)
    bpatch = np.full(n, -1, np.int64)
    blocal = np.zeros((n, 8))
    bell = np.ones(n)
    boxes = {}
    for o in range(scene.nobj):
        s = int(scene.obj_surface[o])
        if s not in boxes:
            P = scene.surface_patches(s)
            xyz = P[:, :, :3].astype(np.float64) / P[:, :, 3:4].astype(np.float64)
            lo, hi = xyz.min(axis=1), xyz.max(axis=1)
            pad = 1e-9 * (np.abs(lo) + np.abs(hi)) + 1e-12
            boxes[s] = (P, lo - pad, hi + pad)
        P, lo, hi = boxes[s]
        lr, ell = _local_rays(scene, o, rays64)
        for c0 in range(0, n, chunk):
            L = lr[c0:c0 + chunk]
            with np.errstate(divide="ignore", invalid="ignore"):
                inv = 1.0 / L[:, None, 4:7]
                t0 = (lo[None] - L[:, None, 0:3]) * inv
                t1 = (hi[None] - L[:, None, 0:3]) * inv
            tn = np.nanmax(np.minimum(t0, t1), axis=2)
            tf = np.nanmin(np.maximum(t0, t1), axis=2)
            ri, pj = np.nonzero((tn <= tf) & (tf >= 0) & (tn <= L[:, 3:4]))
            if ri.size == 0:
                continue
            res = cgs_oracle.clip_batch64(L[ri], P[pj])
            found = res[:, 3] == 1
            ri, pj, res = ri[found], pj[found], res[found]
            gi = ri + c0
            tw = res[:, 0] / ell[gi]
            order = np.argsort(tw, kind="stable")
            for k in order:
                g = gi[k]
                if tw[k] < best[g]:
                    if bobj[g] != o or bpatch[g] != pj[k]:
                        second[g] = best[g]
                    best[g], bu[g], bv[g], bobj[g], bpatch[g] = tw[k], res[k, 1], res[k, 2], o, pj[k]
                    blocal[g] = L[ri[k]]
                    bell[g] = ell[g]
                elif (bobj[g] != o or bpatch[g] != pj[k]) and tw[k] < second[g]:
                    second[g] = tw[k]
    return Fp64Hits(best, bu, bv, bobj, bpatch, second, blocal, bell)


@dataclasses.dataclass
class Report:
    name: str
    rays: int
    hits32: int
    hits64: int
    grazing: int
    seam: int
    hitmiss_unflagged: int
    patch_unflagged: int
    hitmiss_flagged: int
    patch_flagged: int
    max_rel_t_fp64: float
    max_uv_fp64: float
    p99_rel_t_fp64: float
    max_rel_t_root: float
    max_uv_root: float
    max_rel_t_root32: float
    newton_failed: int
    max_uv_root32: float = 0.0

    def row(self) -> str:
        return (f"| {self.name} | {self.rays} | {self.hits32} | {self.grazing} / {self.seam} | "
                f"{self.hitmiss_unflagged} / {self.patch_unflagged} | {self.hitmiss_flagged} / {self.patch_flagged} | "
                f"{self.max_rel_t_fp64:.2e} (p99 {self.p99_rel_t_fp64:.1e}) | {self.max_uv_fp64:.2e} | "
                f"{self.max_rel_t_root:.2e} / {self.max_uv_root:.2e} | {self.max_rel_t_root32:.2e} |")


HEADER = ("| workload | rays | hits | grazing / seam | hit-miss / patch disagreements (unflagged) | "
          "(flagged) | max rel t, FP32 vs FP64 | max uv, FP32 vs FP64 | FP64 vs root: rel t / uv | "
          "FP32 vs root: rel t |\n|---|---|---|---|---|---|---|---|---|---|")


def fp32_closest(scene, rays, **options):
    """(t,u,v), patch, object of the FP32 path (the oracle's restatement),
    with the trace options of cgs_trace_options (e.g. polish=True)."""
    import cgs_oracle
    h32, p32, o32, _ = cgs_oracle.OracleScene(scene).trace(rays, **options)
    return h32, p32, o32


def measure(name, scene, rays, fp32=None, fp64=None):
    """Report of the FP32 results (default: the oracle's) against FP64."""
    with np.errstate(invalid="ignore", divide="ignore"):
        return _measure(name, scene, rays, fp32, fp64)


def _measure(name, scene, rays, fp32, fp64):
    import cgs_oracle
    h32, p32, o32 = fp32 if fp32 is not None else fp32_closest(scene, rays)
    f = fp64 if fp64 is not None else fp64_closest(scene, rays)
    hit64 = f.obj >= 0
    hit32 = o32 >= 0
    # Newton polish of the FP64 hits in the hit object's local frame
    idx = np.nonzero(hit64)[0]
    cos = np.ones(rays.shape[0])
    root = np.zeros((rays.shape[0], 3))
    ok = np.ones(rays.shape[0], bool)
    if idx.size:
        P = np.stack([scene.surface_patches(int(scene.obj_surface[f.obj[i]]))[f.patch[i]] for i in idx])
        L = f.local[idx]
        ell = f.ell[idx]
        seed = np.stack([f.u[idx], f.v[idx], f.t[idx] * ell], 1)
        x, c, k = cgs_oracle.newton_batch(P, L, seed)
        root[idx] = np.stack([x[:, 0], x[:, 1], x[:, 2] / ell], 1)
        cos[idx] = c
        ok[idx] = k
    grazing = hit64 & (cos < GRAZING_COS)
    seam = hit64 & (np.abs(f.t2 - f.t) <= SEAM_REL * np.abs(f.t))
    flagged = grazing | seam
    hm = hit32 != hit64
    both = hit32 & hit64
    same = both & (o32 == f.obj) & (p32 == f.patch)
    pd = both & ~same
    rel = np.abs(h32[:, 0].astype(np.float64) - f.t) / f.t
    duv = np.maximum(np.abs(h32[:, 1] - f.u), np.abs(h32[:, 2] - f.v))
    good = same & ~flagged
    polished = hit64 & ok & ~flagged
    rel_root = np.abs(f.t - root[:, 2]) / root[:, 2]
    uv_root = np.maximum(np.abs(f.u - root[:, 0]), np.abs(f.v - root[:, 1]))
    pr32 = polished & same
    rel_root32 = np.abs(h32[:, 0].astype(np.float64) - root[:, 2]) / root[:, 2]
    uv_root32 = np.maximum(np.abs(h32[:, 1] - root[:, 0]), np.abs(h32[:, 2] - root[:, 1]))
    mx = lambda a, m: float(a[m].max()) if m.any() else 0.0
    return Report(name, int(rays.shape[0]), int(hit32.sum()), int(hit64.sum()), int(grazing.sum()), int(seam.sum()),
                  int((hm & ~flagged).sum()), int((pd & ~flagged).sum()), int((hm & flagged).sum()),
                  int((pd & flagged).sum()), mx(rel, good), mx(duv, good),
                  float(np.percentile(rel[good], 99)) if good.any() else 0.0,
                  mx(rel_root, polished), mx(uv_root, polished), mx(rel_root32, pr32),
                  int((hit64 & ~ok).sum()), mx(uv_root32, pr32))


def workloads():
    from cgspline_amd import scenes
    c1 = scenes.heightfield_c1()
    c2 = scenes.closed_surfaces(width=320, height=180)
    c3 = scenes.displaced_grids(grids=96, width=256, height=256, box=6.0)
    return [("heightfield 256 patches, 256x256", c1, 1), ("64 closed surfaces, 320x180", c2, 2),
            ("96 displaced grids (24,576 patches), 256x256", c3, 2)]


if __name__ == "__main__":
    from cgspline_amd import scenes
    print(HEADER)
    for name, wl, step in workloads():
        rays = scenes.camera_rays(wl.camera)[::step]
        print(measure(name, wl.scene, rays).row(), flush=True)
