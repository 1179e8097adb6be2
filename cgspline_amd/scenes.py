"""Synthetic, seeded scene generators for the ray–patch workloads.

Every scene is a set of bicubic Bézier surfaces (16 control points per patch,
``point(i, j)`` at ``[4*j + i]`` as in the reference's ``PatchRef``,
Spline.h:107-163), placed as objects with a local→world transform and a
material, plus point lights and a pinhole camera (the reference's
``rt::Scene``/``rt::Camera``, rt/Scene.h:29-159).  All arrays are float32 /
uint32 in the layouts the C ABI takes (``include/cgs.h``).

Workloads:

* ``heightfield_c1``  — 256 patches tiling a C1-continuous 16×16 height field
  over [0,1]², 256×256 primary rays from (0.5,-1,1).
* ``closed_surfaces`` — 64 jittered bicubic spheres/tori (16,384 patches) in a
  10³ box, 1920×1080 primary rays + one shadow ray per hit, Lambert shading.
* ``displaced_grids`` — 4,096 displaced 16×16-patch grids (1,048,576 patches)
  with log-uniform scales 1e-3–1 in a 25³ box, 10% nearly-degenerate patches,
  2048² primary rays + shadow + one reflection bounce, GGX shading.
* ``teaset``          — the reference's own `.be` models (teapot, teacup,
  teaspoon on a `longwave` floor) arranged as in assets/scenes/Teaset01.dl.
"""
from __future__ import annotations

import dataclasses
import math

import numpy as np

BASE_SEED = 20261202

_GRID_N = 16               # patches per side of a height-field grid


@dataclasses.dataclass
class Scene:
    """Host-side scene description (the C ABI's ``cgs_scene_desc``)."""

    vertices: np.ndarray          # (nv, 4) float32, all surfaces concatenated
    indices: np.ndarray           # (npatch, 16) uint32, relative to the surface's first vertex
    surf_vbase: np.ndarray        # (nsurf + 1,) uint32 prefix over vertices
    surf_pbase: np.ndarray        # (nsurf + 1,) uint32 prefix over patches
    obj_surface: np.ndarray       # (nobj,) uint32
    obj_local2world: np.ndarray   # (nobj, 16) float32, column-major
    obj_world2local: np.ndarray   # (nobj, 16) float32, column-major
    obj_material: np.ndarray      # (nobj, 8) float32: diffuse rgb, specular rgb, metalness, roughness
    lights: np.ndarray            # (nlight, 12) float32: color rgb, strength, position xyz, range, direction xyz, angle
    background: tuple = (0.1, 0.1, 0.1)
    ambient: tuple = (0.1, 0.1, 0.1)
    hash_seed: int = 42
    name: str = ""

    @property
    def npatch(self) -> int:
        return int(self.indices.shape[0])

    @property
    def nsurf(self) -> int:
        return int(self.surf_vbase.shape[0] - 1)

    @property
    def nobj(self) -> int:
        return int(self.obj_surface.shape[0])

    def surface_patches(self, s: int) -> np.ndarray:
        """(n, 16, 4) float32 control points of surface ``s``'s patches."""
        v0, v1 = int(self.surf_vbase[s]), int(self.surf_vbase[s + 1])
        p0, p1 = int(self.surf_pbase[s]), int(self.surf_pbase[s + 1])
        return self.vertices[v0:v1][self.indices[p0:p1]]


@dataclasses.dataclass
class Camera:
    """Pinhole camera: looks along its local -Z, up +Y (RayTracer.cu:59-71,141-163)."""

    to_world: np.ndarray   # (16,) float32, column-major camera→world
    position: np.ndarray   # (3,) float32
    fovy: float            # radians
    width: int
    height: int
    flip_y: bool = False


@dataclasses.dataclass
class Workload:
    scene: Scene
    camera: Camera
    eps: float = 1e-4
    depth: int = 0
    desc: str = ""


# ---------------------------------------------------------------------------
# transforms
# ---------------------------------------------------------------------------

def look_at(position, target, up) -> np.ndarray:
    """Camera→world matrix (column-major, 16 floats): X right, Y up, Z back."""
    p = np.asarray(position, np.float64)
    f = np.asarray(target, np.float64) - p
    f /= np.linalg.norm(f)
    r = np.cross(f, np.asarray(up, np.float64))
    r /= np.linalg.norm(r)
    u = np.cross(r, f)
    m = np.eye(4)
    m[:3, 0], m[:3, 1], m[:3, 2], m[:3, 3] = r, u, -f, p
    return m.T.reshape(16).astype(np.float32)  # column-major


def make_camera(position, target, up, fovy_deg, width, height) -> Camera:
    return Camera(look_at(position, target, up), np.asarray(position, np.float32),
                  math.radians(fovy_deg), int(width), int(height))


def random_rotation(rng) -> np.ndarray:
    q = rng.normal(size=4)
    q /= np.linalg.norm(q)
    x, y, z, w = q
    return np.array([
        [1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
        [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
        [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)]])


def trs(t, r3, s) -> tuple[np.ndarray, np.ndarray]:
    """local→world and world→local (column-major float32) of T·R·S(s)."""
    m = np.eye(4)
    m[:3, :3] = r3 * s
    m[:3, 3] = t
    inv = np.linalg.inv(m)
    return (m.T.reshape(16).astype(np.float32), inv.T.reshape(16).astype(np.float32))


IDENTITY16 = np.eye(4, dtype=np.float32).reshape(16)


def point_light(position, color) -> np.ndarray:
    """Point light row (angle > 1.57 means point, rt/Scene.h:72-76)."""
    c = np.asarray(color, np.float32)
    return np.array([c[0], c[1], c[2], 1.0, *position, 0.0, 0.0, 0.0, -1.0, 3.1415],
                    np.float32)


# ---------------------------------------------------------------------------
# surfaces
# ---------------------------------------------------------------------------

def heightfield_grid(rng, amplitude=0.1, n=_GRID_N):
    """(3n+1)² control grid over [0,1]² (n patches per side, 49×49 for the
    default 16), z = amp·Σ_k sin(2π(a_k x + b_k y) + φ_k), made C1 across
    patch seams (z_i = (z_{i-1} + z_{i+1})/2 at i ≡ 0 mod 3, along x then y).
    Returns (vertices ((3n+1)²,4) f32, indices (n²,16) u32)."""
    a = rng.uniform(0.5, 2.0, 4)
    b = rng.uniform(0.5, 2.0, 4)
    ph = rng.uniform(0.0, 2 * np.pi, 4)
    ctrl = 3 * n + 1
    g = np.arange(ctrl) / (ctrl - 1)
    x, y = np.meshgrid(g, g, indexing="xy")   # x varies along axis 1 (ix), y along axis 0 (iy)
    z = amplitude * np.sum(np.sin(2 * np.pi * (a[:, None, None] * x + b[:, None, None] * y)
                                  + ph[:, None, None]), axis=0)
    for i in range(3, ctrl - 1, 3):
        z[:, i] = 0.5 * (z[:, i - 1] + z[:, i + 1])
    for i in range(3, ctrl - 1, 3):
        z[i, :] = 0.5 * (z[i - 1, :] + z[i + 1, :])
    verts = np.stack([x, y, z, np.ones_like(z)], axis=-1).reshape(-1, 4).astype(np.float32)
    I, J = np.meshgrid(np.arange(n), np.arange(n), indexing="xy")
    aa, bb = np.meshgrid(np.arange(4), np.arange(4), indexing="xy")   # aa = u (fast), bb = v
    idx = (3 * I.reshape(-1, 1) + aa.reshape(1, -1)) + ctrl * (3 * J.reshape(-1, 1) + bb.reshape(1, -1))
    return verts, idx.astype(np.uint32)


def make_rows_collinear(rng, verts, idx, fraction, jitter):
    """Gives `fraction` of the patches one control row (fixed v index) of
    collinear points (own vertex copies, `jitter`-sized noise)."""
    n = idx.shape[0]
    k = int(round(fraction * n))
    if k == 0:
        return verts, idx
    sel = rng.choice(n, size=k, replace=False)
    rows = rng.integers(0, 4, size=k)
    idx = idx.copy()
    base = verts.shape[0]
    new = np.empty((k, 4, 4), np.float64)
    for m, (p, r) in enumerate(zip(sel, rows)):
        ids = idx[p, 4 * r:4 * r + 4]
        p0 = verts[ids[0], :3].astype(np.float64)
        p3 = verts[ids[3], :3].astype(np.float64)
        for a in range(4):
            q = p0 + (a / 3.0) * (p3 - p0) + rng.uniform(-jitter, jitter, 3)
            new[m, a] = (*q, 1.0)
        idx[p, 4 * r:4 * r + 4] = base + 4 * m + np.arange(4)
    verts = np.concatenate([verts, new.reshape(-1, 4).astype(np.float32)])
    return verts, idx


def _hermite_patches(S, dSu, dSv, dSuv, u_knots, v_knots):
    """Bicubic Hermite→Bézier fit of a parametric surface over a knot grid.

    S, dSu, dSv, dSuv: callables of (u, v) arrays returning (..., 3)."""
    pts = []
    for j in range(len(v_knots) - 1):
        for i in range(len(u_knots) - 1):
            u0, u1 = u_knots[i], u_knots[i + 1]
            v0, v1 = v_knots[j], v_knots[j + 1]
            du, dv = (u1 - u0) / 3.0, (v1 - v0) / 3.0
            P = np.zeros((4, 4, 3))
            for (ci, uu, su) in ((0, u0, 1), (3, u1, -1)):
                for (cj, vv, sv) in ((0, v0, 1), (3, v1, -1)):
                    p = S(uu, vv)
                    fu, fv, fuv = dSu(uu, vv), dSv(uu, vv), dSuv(uu, vv)
                    P[cj, ci] = p
                    P[cj, ci + su] = p + su * du * fu
                    P[cj + sv, ci] = p + sv * dv * fv
                    P[cj + sv, ci + su] = p + su * du * fu + sv * dv * fv + su * sv * du * dv * fuv
            pts.append(P.reshape(16, 3))
    return np.array(pts)   # (n, 16, 3), [4j+i]


def sphere_patches(n=16):
    def S(u, v):
        return np.array([np.cos(u) * np.sin(v), np.sin(u) * np.sin(v), np.cos(v)])

    def Su(u, v):
        return np.array([-np.sin(u) * np.sin(v), np.cos(u) * np.sin(v), 0.0])

    def Sv(u, v):
        return np.array([np.cos(u) * np.cos(v), np.sin(u) * np.cos(v), -np.sin(v)])

    def Suv(u, v):
        return np.array([-np.sin(u) * np.cos(v), np.cos(u) * np.cos(v), 0.0])

    return _hermite_patches(S, Su, Sv, Suv, np.linspace(0, 2 * np.pi, n + 1), np.linspace(0, np.pi, n + 1))


def torus_patches(n=16, r=0.5):
    def S(u, v):
        return np.array([(1 + r * np.cos(v)) * np.cos(u), (1 + r * np.cos(v)) * np.sin(u), r * np.sin(v)])

    def Su(u, v):
        return np.array([-(1 + r * np.cos(v)) * np.sin(u), (1 + r * np.cos(v)) * np.cos(u), 0.0])

    def Sv(u, v):
        return np.array([-r * np.sin(v) * np.cos(u), -r * np.sin(v) * np.sin(u), r * np.cos(v)])

    def Suv(u, v):
        return np.array([r * np.sin(v) * np.sin(u), -r * np.sin(v) * np.cos(u), 0.0])

    return _hermite_patches(S, Su, Sv, Suv, np.linspace(0, 2 * np.pi, n + 1), np.linspace(0, 2 * np.pi, n + 1))


# ---------------------------------------------------------------------------
# scene assembly
# ---------------------------------------------------------------------------

def assemble(surfaces, objects, lights, name="", background=(0.1, 0.1, 0.1), ambient=(0.1, 0.1, 0.1)) -> Scene:
    """surfaces: list of (verts (nv,4), idx (np,16)); objects: list of
    (surface, l2w16, w2l16, material8)."""
    vb = np.zeros(len(surfaces) + 1, np.uint32)
    pb = np.zeros(len(surfaces) + 1, np.uint32)
    for s, (v, i) in enumerate(surfaces):
        vb[s + 1] = vb[s] + v.shape[0]
        pb[s + 1] = pb[s] + i.shape[0]
    verts = np.concatenate([v for v, _ in surfaces]).astype(np.float32)
    idx = np.concatenate([i for _, i in surfaces]).astype(np.uint32)
    return Scene(
        vertices=np.ascontiguousarray(verts), indices=np.ascontiguousarray(idx),
        surf_vbase=vb, surf_pbase=pb,
        obj_surface=np.array([o[0] for o in objects], np.uint32),
        obj_local2world=np.array([o[1] for o in objects], np.float32).reshape(-1, 16),
        obj_world2local=np.array([o[2] for o in objects], np.float32).reshape(-1, 16),
        obj_material=np.array([o[3] for o in objects], np.float32).reshape(-1, 8),
        lights=np.array(lights, np.float32).reshape(-1, 12),
        background=tuple(background), ambient=tuple(ambient), name=name)


def heightfield_c1(seed=BASE_SEED + 1, width=256, height=256) -> Workload:
    """256 C1 height-field patches; camera at (0.5,-1,1) looking at (0.5,0.5,0)."""
    rng = np.random.default_rng(seed)
    v, i = heightfield_grid(rng)
    mat = np.array([0.8, 0.7, 0.6, 0.04, 0.04, 0.04, 0.0, 0.5], np.float32)
    sc = assemble([(v, i)], [(0, IDENTITY16, IDENTITY16, mat)],
                  [point_light((0.5, -0.5, 2.0), (4.0, 4.0, 4.0))], name="heightfield_c1")
    cam = make_camera((0.5, -1.0, 1.0), (0.5, 0.5, 0.0), (0.0, 0.0, 1.0), 45.0, width, height)
    return Workload(sc, cam, depth=0, desc="256 C1 height-field patches, 256x256 primary rays")


def closed_surfaces(seed=BASE_SEED + 2, count=64, width=1920, height=1080, jitter=0.02) -> Workload:
    """64 jittered bicubic spheres/tori (256 patches each) in a 10³ box."""
    rng = np.random.default_rng(seed)
    sph, tor = sphere_patches(), torus_patches()
    surfaces, objects = [], []
    for k in range(count):
        base = sph if rng.random() < 0.5 else tor
        pts = base + rng.uniform(-jitter, jitter, base.shape)
        verts = np.concatenate([pts.reshape(-1, 3), np.ones((pts.shape[0] * 16, 1))], axis=1)
        idx = np.arange(pts.shape[0] * 16, dtype=np.uint32).reshape(-1, 16)
        surfaces.append((verts.astype(np.float32), idx))
        l2w, w2l = trs(rng.uniform(0, 10, 3), random_rotation(rng), rng.uniform(0.3, 1.5))
        mat = np.array([*rng.uniform(0.2, 0.9, 3), 0.04, 0.04, 0.04, 0.0, 1.0], np.float32)
        objects.append((k, l2w, w2l, mat))
    sc = assemble(surfaces, objects, [point_light((5.0, -5.0, 15.0), (300.0, 300.0, 300.0))],
                  name="closed_surfaces")
    cam = make_camera((5.0, -14.0, 8.0), (5.0, 5.0, 5.0), (0.0, 0.0, 1.0), 50.0, width, height)
    return Workload(sc, cam, depth=0, desc="64 spheres/tori (16,384 patches), Lambert, primary + shadow")


def displaced_grids(seed=BASE_SEED + 3, grids=4096, width=2048, height=2048,
                    degenerate=0.10, box=25.0) -> Workload:
    """`grids` displaced 16×16-patch grids with log-uniform scale in [1e-3, 1],
    random rotation and position in [0, box]³; 10% of the patches get a
    collinear control row.  One object (own BVH and hash) per grid: a single
    BVH over all 1M patches cannot be hashed by the reference's perfect-hash
    scheme (every (A,B) attempt fails at 2M keys)."""
    rng = np.random.default_rng(seed)
    surfaces, objects = [], []
    for k in range(grids):
        v, i = heightfield_grid(rng)
        v, i = make_rows_collinear(rng, v, i, degenerate, 1e-6)
        s = math.exp(rng.uniform(math.log(1e-3), math.log(1.0)))
        l2w, w2l = trs(rng.uniform(0, box, 3), random_rotation(rng), s)
        mat = np.array([*rng.uniform(0.2, 0.9, 3), *np.full(3, rng.uniform(0.02, 0.9)),
                        rng.uniform(0.0, 1.0), rng.uniform(0.05, 1.0)], np.float32)
        surfaces.append((v, i))
        objects.append((k, l2w, w2l, mat))
    c = box / 2
    sc = assemble(surfaces, objects,
                  [point_light((c, -0.5 * box, 1.5 * box), (3e4, 3e4, 3e4))], name="displaced_grids")
    cam = make_camera((c, -0.85 * box, c), (c, c, c), (0.0, 0.0, 1.0), 60.0, width, height)
    return Workload(sc, cam, depth=1, desc=f"{grids} displaced grids ({grids * 256} patches), GGX, "
                                            "primary + shadow + 1 reflection bounce")


def displaced_grids_single(seed=BASE_SEED + 3, grids=4096, width=2048, height=2048,
                           degenerate=0.10, box=25.0) -> Workload:
    """The displaced grids of displaced_grids (the same seeded generator) as
    one surface in world space: one BVH over all `grids`×256 patches under
    one identity object, with the first grid's material.  At 4096 grids its
    BVH has 2,097,151 nodes, which the reference's hash family cannot index;
    the scene hashes it with the wide family."""
    wl = displaced_grids(seed, grids, width, height, degenerate, box)
    sc = wl.scene
    parts, idx, nv = [], [], 0
    for k in range(sc.nobj):
        v0, v1 = int(sc.surf_vbase[k]), int(sc.surf_vbase[k + 1])
        p0, p1 = int(sc.surf_pbase[k]), int(sc.surf_pbase[k + 1])
        M = sc.obj_local2world[k].reshape(4, 4).T.astype(np.float64)     # column-major
        w = sc.vertices[v0:v1].astype(np.float64) @ M.T
        parts.append(np.column_stack([w[:, :3] / w[:, 3:4], np.ones(v1 - v0)]).astype(np.float32))
        idx.append(sc.indices[p0:p1] + nv)
        nv += v1 - v0
    one = assemble([(np.concatenate(parts), np.concatenate(idx))],
                   [(0, IDENTITY16, IDENTITY16, sc.obj_material[0])], sc.lights, name="displaced_grids_single")
    return Workload(one, wl.camera, depth=wl.depth,
                    desc=f"{grids} displaced grids ({grids * 256} patches) in one BVH, GGX, "
                         "primary + shadow + 1 reflection bounce")


def grazing_heightfield(seed=BASE_SEED + 5, n=128, width=128, height=128) -> Workload:
    """One n×n-patch height field (16,384 patches at n=128: a 14-level BVH)
    seen at a grazing angle, so that rays find both children of a node hit at
    many levels in a row and postpone up to 12 siblings at once.  Shading
    adds shadow rays and one reflection bounce."""
    rng = np.random.default_rng(seed)
    v, i = heightfield_grid(rng, amplitude=0.01, n=n)
    mat = np.array([0.8, 0.7, 0.6, 0.04, 0.04, 0.04, 0.0, 0.5], np.float32)
    sc = assemble([(v, i)], [(0, IDENTITY16, IDENTITY16, mat)],
                  [point_light((0.5, 2.0, 0.3), (4.0, 4.0, 4.0))], name="grazing_heightfield")
    cam = make_camera((0.5, -0.1, 0.04), (0.5, 0.5, -0.02), (0.0, 0.0, 1.0), 40.0, width, height)
    return Workload(sc, cam, depth=1, desc=f"{n * n} height-field patches at a grazing angle, "
                                            "primary + shadow + 1 reflection bounce")


def stacked_arcs(n=180_000):
    """n objects stacked along +y, each a quarter-cylinder patch (axis y,
    radius 1, height 1) whose box [0,1]×[k,k+1]×[0,1] a ray along +y at
    x, z ≈ 0.3 crosses without touching the surface, and on top one bumped
    square patch that the ray hits.  The top-level BVH over the objects is
    about 18 levels deep and such a ray finds both children hit at every
    level.  (n = 180,000: the top level's 359,999 heap keys hash; 2^17 of
    them do not.)  Returns (scene, rays (64, 8))."""
    k = 0.5523
    arc = [(1.0, 0.0), (1.0, k), (k, 1.0), (0.0, 1.0)]
    cyl = np.array([[x, j / 3.0, z, 1.0] for j in range(4) for (x, z) in arc], np.float32)
    bump = np.array([[i / 3.0, 0.3 if (0 < i < 3 and 0 < j < 3) else 0.0, j / 3.0, 1.0]
                     for j in range(4) for i in range(4)], np.float32)
    idx = np.arange(16, dtype=np.uint32).reshape(1, 16)
    mat = np.array([0.8, 0.8, 0.8, 0.04, 0.04, 0.04, 0.0, 0.5], np.float32)
    objects = []
    for i in range(n):
        l2w, w2l = trs(np.array([0.0, float(i), 0.0]), np.eye(3), 1.0)
        objects.append((1 if i == n - 1 else 0, l2w, w2l, mat))
    sc = assemble([(cyl, idx), (bump, idx)], objects, [point_light((0.5, -2.0, 0.5), (1.0, 1.0, 1.0))],
                  name="stacked_arcs")
    rng = np.random.default_rng(BASE_SEED + 6)
    rays = np.zeros((64, 8), np.float32)
    rays[:, 0] = rng.uniform(0.2, 0.4, 64)
    rays[:, 1] = -1.0
    rays[:, 2] = rng.uniform(0.2, 0.4, 64)
    rays[:, 3] = np.inf
    rays[:, 5] = 1.0
    return sc, rays


@dataclasses.dataclass
class BatchItem:
    scene: Scene
    rays: np.ndarray     # (n, 8) float32
    any: bool = False


def batched_scenes(seed=BASE_SEED + 4, count=64, patches=(1_000, 200_000), rays=(4_096, 1 << 21),
                   any_fraction=0.0) -> list[BatchItem]:
    """Many resident scenes with uneven ray batches: each scene is the
    displaced-grid generator with a log-uniform patch count (whole 16×16-patch
    grids, density of the 1M-patch scene) and a log-uniform ray count, half
    camera rays through the scene centre and half random origins in the scene
    box with uniform directions.  `any_fraction` of the jobs are occlusion
    queries."""
    rng = np.random.default_rng(seed)
    out = []
    for k in range(count):
        npatch = math.exp(rng.uniform(math.log(patches[0]), math.log(patches[1])))
        grids = max(1, int(round(npatch / 256)))
        nray = int(round(math.exp(rng.uniform(math.log(rays[0]), math.log(rays[1])))))
        box = 25.0 * (grids / 4096.0) ** (1.0 / 3.0)
        side = max(8, int(math.ceil(math.sqrt(nray / 2))))
        wl = displaced_grids(seed=int(rng.integers(1 << 31)), grids=grids, width=side, height=side, box=box)
        cam = camera_rays(wl.camera)[: nray // 2]
        m = nray - cam.shape[0]
        rnd = np.zeros((m, 8), np.float32)
        rnd[:, 0:3] = rng.uniform(0.0, box, (m, 3))
        d = rng.normal(size=(m, 3))
        rnd[:, 4:7] = d / np.linalg.norm(d, axis=1, keepdims=True)
        rnd[:, 3] = np.inf
        out.append(BatchItem(wl.scene, np.concatenate([cam, rnd]).astype(np.float32),
                             bool(rng.random() < any_fraction)))
    return out


def rotation_xyz(degrees) -> np.ndarray:
    """R_y(y)·R_x(x)·R_z(z) for Euler angles in degrees (yaw, then pitch,
    then roll, the usual camera order).  Single-axis rotations, the only kind
    the teaset objects use, are the same under every order."""
    x, y, z = (math.radians(a) for a in degrees)
    cx, sx, cy, sy, cz, sz = math.cos(x), math.sin(x), math.cos(y), math.sin(y), math.cos(z), math.sin(z)
    Rx = np.array([[1, 0, 0], [0, cx, -sx], [0, sx, cx]])
    Ry = np.array([[cy, 0, sy], [0, 1, 0], [-sy, 0, cy]])
    Rz = np.array([[cz, -sz, 0], [sz, cz, 0], [0, 0, 1]])
    return Ry @ Rx @ Rz


def euler_camera(position, rotation_deg, fovy_deg, width, height) -> Camera:
    """Camera at `position` looking down its −Z axis turned by
    :func:`rotation_xyz` (Y up)."""
    m = np.eye(4)
    m[:3, :3] = rotation_xyz(rotation_deg)
    m[:3, 3] = position
    return Camera(m.T.reshape(16).astype(np.float32), np.asarray(position, np.float32),
                  math.radians(fovy_deg), int(width), int(height))


def teaset(assets_dir, width=320, height=240, depth=2, fovy_deg=60.0) -> Workload:
    """assets/scenes/Teaset01.dl restated: a `longwave` floor scaled (10, 0.1,
    10), the teapot turned −90° about X and halved, a teacup at (2.4, 0, 0),
    a teaspoon at (3.7, 0.1, 0) turned −85° about X, one point light of
    colour 90·(1, 0.95, 0.8) at (−10, 5, 5), the file's materials,
    background and ambient, and its view (position (3, 2.9, 3.5), rotation
    (−34°, 20°, −0.0012°)).  `assets_dir` is laid out like the reference's
    assets/ folder (bezier/teapot, ...); the surfaces are read with
    cgs_load_be.  The view's Euler order
    and the field of view are this module's choice (the scene file does not
    fix them)."""
    import os

    from . import load_be
    surf = {n: load_be(os.path.join(assets_dir, "bezier", n)) for n in ("longwave", "teapot", "teacup", "teaspoon")}
    names = list(surf)
    f0, silver = (0.04, 0.04, 0.04), (0.90, 0.91, 0.98)
    ceramic = (0.82, 0.83, 0.84)
    mat = {
        "CeramicWhite": [*ceramic, *ceramic, 0.1, 0.9],
        "Silver": [*f0, *silver, 0.9, 0.1],
        "Floor": [*silver, *silver, 0.05, 0.9],
    }
    placed = [  # (model, position, rotation degrees, scale, material)
        ("longwave", (0, 0, 0), (0, 0, 0), (10, 0.1, 10), "Floor"),
        ("teapot", (0, 0, 0), (-90, 0, 0), (0.5, 0.5, 0.5), "CeramicWhite"),
        ("teacup", (2.4, 0, 0), (0, 0, 0), (1, 1, 1), "CeramicWhite"),
        ("teaspoon", (3.7, 0.1, 0), (-85, 0, 0), (1, 1, 1), "Silver"),
    ]
    objects = []
    for model, pos, rot, scale, m in placed:
        l2w, w2l = trs(pos, rotation_xyz(rot), np.asarray(scale, np.float64))
        objects.append((names.index(model), l2w, w2l, np.array(mat[m], np.float32)))
    light = point_light((-10.0, 5.0, 5.0), 90.0 * np.array([1.0, 0.95, 0.8]))
    sc = assemble([surf[n] for n in names], objects, [light], name="teaset",
                  background=(0.5, 0.5, 0.5), ambient=(0.025, 0.025, 0.025))
    cam = euler_camera((3.0, 2.9, 3.5), (-34.0, 20.0, -0.0012), fovy_deg, width, height)
    return Workload(sc, cam, depth=depth, desc="reference teaset models (78 patches), GGX, "
                                               f"{depth} reflection bounces")


def from_dl(path, width=320, height=240, depth=2, fovy_deg=60.0) -> Workload:
    """A workload from one of the reference's `.dl` scene files (read with
    cgs_dl_read) seen from the file's view.  The field of view is not part of
    the file; 60° is this module's choice."""
    from . import read_dl
    r = read_dl(path)
    cam = Camera(r.view_to_world, r.view_position, math.radians(fovy_deg), int(width), int(height))
    return Workload(r.scene, cam, depth=depth, desc=f"{r.scene.name}: {r.scene.nobj} objects, "
                                                    f"{r.scene.npatch} patches, {r.scene.lights.shape[0]} lights")


_libm = None


def _tanf(x) -> np.float32:
    """The C library's single-precision tangent, which the render's camera
    setup calls on the host (std::tan of a float, RayTracer.cu:141-163):
    it is not always the correctly rounded value (tan(pi/6 in float32)
    rounds up, where the double tangent rounded to float32 rounds down)."""
    global _libm
    if _libm is None:
        import ctypes
        import ctypes.util
        _libm = ctypes.CDLL(ctypes.util.find_library("m") or "libm.so.6")
        _libm.tanf.restype = ctypes.c_float
        _libm.tanf.argtypes = [ctypes.c_float]
    return np.float32(_libm.tanf(float(np.float32(x))))


def camera_rays(cam: Camera, rows=None) -> np.ndarray:
    """World rays (n, 8) float32 (o xyz, tmax, d xyz, 0) through pixel centres,
    in float32 with the reference's camera formula (RayTracer.cu:59-71,
    141-163)."""
    f32 = np.float32
    w, h = cam.width, cam.height
    half_h = _tanf(f32(0.5) * f32(cam.fovy))
    half_dy = f32(half_h / f32(h))
    half_dx = half_dy
    tl_x = f32(-half_dx * f32(w - 1))
    tl_y = f32(half_h - half_dy)
    if cam.flip_y:
        tl_y, half_dy = -tl_y, -half_dy
    js = np.arange(h) if rows is None else np.asarray(rows)
    ii, jj = np.meshgrid(np.arange(w, dtype=f32), js.astype(f32), indexing="xy")
    px = tl_x + half_dx * (f32(2) * ii)
    py = tl_y - half_dy * (f32(2) * jj)
    M = cam.to_world.reshape(4, 4).astype(f32)   # rows are columns
    P = (px[..., None] * M[0] + py[..., None] * M[1]) + f32(-1) * M[2] + f32(1) * M[3]
    P3 = P[..., :3] / P[..., 3:4]
    d = P3 - cam.position.astype(f32)
    n = np.sqrt((d[..., 0] * d[..., 0] + d[..., 1] * d[..., 1]) + d[..., 2] * d[..., 2])
    d = d * (f32(1) / n)[..., None]
    rays = np.zeros(d.shape[:-1] + (8,), f32)
    rays[..., 0:3] = cam.position
    rays[..., 3] = np.inf
    rays[..., 4:7] = d
    return rays.reshape(-1, 8)


WORKLOADS = {
    "heightfield_c1": heightfield_c1,
    "closed_surfaces": closed_surfaces,
    "displaced_grids": displaced_grids,
    "displaced_grids_single": displaced_grids_single,
}
