"""Writes tests/golden/c1_heightfield_hits.npz: the closest hits of the plain
CPU workload, as the oracle computes them, kept as a fixture so that a change
that moved the oracle and the kernels together would still show.

Workload: the 256-patch C1-continuous height field of scenes.heightfield_c1
(49×49 control grid, four seeded sinusoids) and its 256×256 camera rays from
(0.5, -1, 1).  Stored per ray: t, u, v (float32) and the patch id (uint32,
0xFFFFFFFF on a miss) of the Bézier-clipping closest hit
(Context::intersect, RayTracer.cu:296-350), and SHA-256 digests of the
scene's control points and indices and of the rays, so that a change of the
generator is reported as such.

Run from the repository root after building the oracle:
    python tests/golden/make_c1_golden.py
"""
from __future__ import annotations

import hashlib
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
OUT = os.path.join(HERE, "c1_heightfield_hits.npz")


def digest(*arrays) -> str:
    h = hashlib.sha256()
    for a in arrays:
        h.update(np.ascontiguousarray(a).tobytes())
    return h.hexdigest()


def compute():
    sys.path[:0] = [ROOT, os.path.join(ROOT, "oracle")]
    import cgs_oracle
    from cgspline_amd import scenes
    wl = scenes.heightfield_c1()
    rays = scenes.camera_rays(wl.camera)
    hits, patch, obj, _ = cgs_oracle.OracleScene(wl.scene).trace(rays)
    patch = np.where(obj >= 0, patch, np.uint32(0xFFFFFFFF)).astype(np.uint32)
    return dict(t=hits[:, 0].astype(np.float32), u=hits[:, 1].astype(np.float32), v=hits[:, 2].astype(np.float32),
                patch=patch, scene_sha256=np.array(digest(wl.scene.vertices, wl.scene.indices)),
                rays_sha256=np.array(digest(rays)), width=np.array(wl.camera.width), height=np.array(wl.camera.height))


if __name__ == "__main__":
    d = compute()
    np.savez_compressed(OUT, **d)
    print(f"{OUT}: {int((d['patch'] != 0xFFFFFFFF).sum())} hits of {d['patch'].shape[0]} rays, "
          f"{os.path.getsize(OUT)} bytes")
