"""Serving over resident scenes: the continuous-batching server of libcgs
(cgs_server_*) and a Poisson load generator that reports latency
percentiles and sustained throughput."""
from __future__ import annotations

import ctypes as C
import math
import time

import numpy as np

from . import _abi
from ._abi import check, lib


class Server:
    """cgs_server over the given DeviceScenes.  Requests are host arrays; the
    output arrays are owned by the returned request record until it is done."""

    def __init__(self, ctx, scenes, hbm_budget: int = 1 << 30, slots: int = 3, max_batch_rays: int = 0):
        self.ctx = ctx
        self.scenes = list(scenes)
        arr = (C.c_void_p * max(1, len(self.scenes)))(*[s.h for s in self.scenes])
        opt = _abi.ServerOptions(hbm_budget, slots, max_batch_rays)
        p = C.c_void_p()
        check(lib().cgs_server_create(ctx.h, arr, len(self.scenes), C.byref(opt), C.byref(p)), "cgs_server_create")
        self.h = p.value
        self.live = {}

    def submit(self, scene: int, rays: np.ndarray, any_hit: bool = False) -> int:
        rays = np.ascontiguousarray(rays, np.float32).reshape(-1, 8)
        n = rays.shape[0]
        rec = {"rays": rays, "any": any_hit}
        if any_hit:
            rec["occluded"] = np.empty(n, np.uint32)
            outs = (None, None, rec["occluded"].ctypes.data)
        else:
            rec["hits"] = np.empty((n, 4), np.float32)
            rec["objects"] = np.empty(n, np.int32)
            outs = (rec["hits"].ctypes.data, rec["objects"].ctypes.data, None)
        rid = C.c_uint64()
        check(lib().cgs_server_submit(self.h, scene, rays.ctypes.data, n, 1 if any_hit else 0, *outs, C.byref(rid)),
              "cgs_server_submit")
        self.live[rid.value] = rec
        return rid.value

    def submit_shaded(self, scene: int, rays: np.ndarray, eps: float = 1e-4, depth: int = 0) -> int:
        """A shaded request (cgs_server_submit_shaded): the record gets
        "color" (n,3), "pixels" (n,), "hits" (n,4) and "objects" (n,)."""
        rays = np.ascontiguousarray(rays, np.float32).reshape(-1, 8)
        n = rays.shape[0]
        rec = {"rays": rays, "any": False, "shaded": True, "color": np.empty((n, 3), np.float32),
               "pixels": np.empty(n, np.uint32), "hits": np.empty((n, 4), np.float32),
               "objects": np.empty(n, np.int32)}
        op = _abi.RenderOptions(eps, depth, 0)
        rid = C.c_uint64()
        check(lib().cgs_server_submit_shaded(self.h, scene, rays.ctypes.data, n, C.byref(op), rec["color"].ctypes.data,
                                             rec["pixels"].ctypes.data, rec["hits"].ctypes.data,
                                             rec["objects"].ctypes.data, C.byref(rid)), "cgs_server_submit_shaded")
        self.live[rid.value] = rec
        return rid.value

    def poll(self, rid: int):
        """None while pending, else (record, latency in seconds)."""
        lat = C.c_uint64()
        r = lib().cgs_server_poll(self.h, rid, C.byref(lat))
        if r == 0:
            return None
        if r < 0:
            raise _abi.CgsError(2, f"request {rid} failed")
        return self.live.pop(rid), lat.value * 1e-9

    def wait(self, rid: int):
        lat = C.c_uint64()
        check(lib().cgs_server_wait(self.h, rid, C.byref(lat)), "cgs_server_wait")
        return self.live.pop(rid), lat.value * 1e-9

    def close(self):
        if self.h is not None:
            lib().cgs_server_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def request_sizes(rng, count, lo=1_000, hi=4_000_000):
    """Log-uniform request sizes in [lo, hi] rays."""
    return np.exp(rng.uniform(math.log(lo), math.log(hi), count)).round().astype(np.int64)


def run_load(server: Server, requests, rate_hz: float, seed: int = 0, depth: int = 1):
    """Submits `requests` [(scene, rays, kind)] at Poisson arrival times of
    the given rate (in-process, open loop) and waits for all of them; kind is
    False (closest hits), True (occlusion) or "shade" (shaded at `depth`).
    Returns the latencies (s), the requests' rays (a shaded request's shadow
    and reflection rays not counted) and the wall time from the first arrival
    to the last completion."""
    rng = np.random.default_rng(seed)
    gaps = rng.exponential(1.0 / rate_hz, len(requests)) if rate_hz > 0 else np.zeros(len(requests))
    start = time.perf_counter()
    due = start
    ids = []
    for (scene, rays, kind), gap in zip(requests, gaps):
        due += gap
        while True:
            now = time.perf_counter()
            if now >= due:
                break
            time.sleep(min(due - now, 5e-4))
        if kind == "shade":
            ids.append(server.submit_shaded(scene, rays, depth=depth))
        else:
            ids.append(server.submit(scene, rays, bool(kind)))
    lats = [server.wait(i)[1] for i in ids]
    wall = time.perf_counter() - start
    return np.array(lats), int(sum(r[1].shape[0] for r in requests)), wall
