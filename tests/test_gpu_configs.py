"""The batched and serving configurations at their full size.

Batched: 64 resident scenes with patch counts log-uniform in [1k, 200k] and
ray batches log-uniform in [4k, 2M] (half camera, half random rays): one
persistent batched launch, its graph replay and the autotuned launch shape
each give every job the results of a separate cgs_trace_closest call, bit
for bit, and a strided sample of every job's rays equals the oracle.

Serving: 8 resident scenes (7 of 10k-1M patches and the 1,048,576-patch
scene) and 24 requests of log-uniform 1k-4M rays, 40 % of them occlusion
queries: every request's results are the same served alone and in the mix
(bit for bit), and a strided sample of every request equals the oracle."""
from __future__ import annotations

import numpy as np
import pytest

from conftest import rng, say

pytestmark = pytest.mark.gpu

SAMPLE = 1000   # oracle rays per job or request


def bits(a):
    return np.ascontiguousarray(a, np.float32).view(np.uint32)


def check_against_oracle(osc, rays, any_hit, got, what):
    k = max(1, rays.shape[0] // SAMPLE)
    sub = rays[::k]
    if any_hit:
        occ, _ = osc.trace(sub, any_hit=True)
        g = got[::k]
        assert not np.any(g == 2), what
        assert np.array_equal(g == 1, occ), f"{what}: occlusion differs from the oracle"
        return
    hits, objects = got
    oh, op, oo, _ = osc.trace(sub)
    h = hits[::k]
    assert np.array_equal(objects[::k], oo), f"{what}: object ids differ from the oracle"
    assert np.array_equal(h[:, 3], op), f"{what}: patch ids differ from the oracle"
    assert np.array_equal(h[:, :3], bits(oh)), f"{what}: t, u, v differ from the oracle"


@pytest.fixture(scope="module")
def batched_config(ctx):
    import cgspline_amd as cgs
    from cgspline_amd import scenes
    say("batched config: generating 64 scenes")
    items = scenes.batched_scenes()
    jobs = []
    for it in items:
        ds = ctx.scene(it.scene, hash_retries=8)
        n = it.rays.shape[0]
        jobs.append(dict(scene=ds, rays=cgs.DeviceArray.from_host(ctx, it.rays), n=n,
                         hits=cgs.DeviceArray(ctx, 16 * n), objects=cgs.DeviceArray(ctx, 4 * n)))
    yield items, jobs
    for j in jobs:
        j["scene"].destroy()


def read_jobs(ctx, jobs):
    ctx.sync()
    return [(j["hits"].to_host(np.uint32, (j["n"], 4)), j["objects"].to_host(np.int32, (j["n"],))) for j in jobs]


def scrub(ctx, jobs):
    import cgspline_amd as cgs
    for j in jobs:
        for key in ("hits", "objects"):
            cgs.check(cgs.lib().cgs_memset(ctx.h, j[key].dptr, 0x5A, j[key].nbytes, None))
    ctx.sync()


def test_batched_config_equals_per_job_and_oracle(ctx, oracle, batched_config):
    items, jobs = batched_config
    npatch = sum(i.scene.npatch for i in items)
    nray = sum(i.rays.shape[0] for i in items)
    assert len(items) == 64 and 1_000_000 < npatch and 10_000_000 < nray
    assert min(i.scene.npatch for i in items) >= 256 and max(i.rays.shape[0] for i in items) <= 1 << 21
    say("batched config: per-job", len(jobs), "scenes", npatch, "patches", nray, "rays")
    for j in jobs:
        j["scene"].trace_closest_dev(j["rays"], j["n"], j["hits"], j["objects"])
    ref = read_jobs(ctx, jobs)
    forms = {}
    scrub(ctx, jobs)
    say("batched config: one batched launch")
    ctx.trace_batch(jobs)
    forms["batched"] = read_jobs(ctx, jobs)
    scrub(ctx, jobs)
    say("batched config: graph replay x3")
    g = ctx.batch(jobs)
    for _ in range(3):
        g.launch()
        ctx.sync()
    g.destroy()
    forms["graph"] = read_jobs(ctx, jobs)
    scrub(ctx, jobs)
    say("batched config: autotuned")
    shape = ctx.batch_autotune(jobs)
    scrub(ctx, jobs)
    ctx.trace_batch(jobs)
    forms["autotuned"] = read_jobs(ctx, jobs)
    for name, res in forms.items():
        for k, (a, b) in enumerate(zip(res, ref)):
            assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1]), f"{name}: job {k} differs from per-job"
    say("batched config: oracle on", SAMPLE, "rays per job; autotuned shape", shape)
    hits_total = 0
    for k, (it, r) in enumerate(zip(items, ref)):
        check_against_oracle(oracle.OracleScene(it.scene), it.rays, False, r, f"job {k}")
        hits_total += int((r[1] >= 0).sum())
    assert hits_total > nray // 100      # about 3 % of these rays hit (random rays mostly leave the box)


@pytest.fixture(scope="module")
def serving_config(ctx):
    from cgspline_amd import scenes
    say("serving config: generating 8 scenes")
    items = scenes.batched_scenes(seed=20261202 + 5, count=7, patches=(10_000, 1_000_000), rays=(1 << 20, 1 << 21))
    big = scenes.displaced_grids()
    items.append(scenes.BatchItem(big.scene, scenes.camera_rays(big.camera)))
    ds = [ctx.scene(it.scene, hash_retries=8) for it in items]
    yield items, ds
    for d in ds:
        d.destroy()


def test_serving_config_alone_equals_mix_and_oracle(ctx, oracle, serving_config):
    from cgspline_amd.serving import Server, request_sizes
    items, ds = serving_config
    assert len(items) == 8 and max(i.scene.npatch for i in items) == 1 << 20
    r = rng(23)
    sizes = request_sizes(r, 24)
    sizes[0] = 1000                       # the smallest and the largest request of the configuration
    sizes[1] = 4_000_000
    reqs = []
    for n in sizes:
        s = int(r.integers(0, len(items))) if n < 4_000_000 else len(items) - 1
        pool = items[s].rays
        n = int(min(n, pool.shape[0]))
        start = int(r.integers(0, pool.shape[0] - n + 1))
        reqs.append((s, pool[start:start + n], bool(r.random() < 0.4)))
    nray = sum(x[1].shape[0] for x in reqs)
    say("serving config:", len(reqs), "requests,", nray, "rays, served alone")
    sv = Server(ctx, ds, hbm_budget=1 << 30, slots=3, max_batch_rays=1 << 22)
    alone = []
    for s, rays, a in reqs:
        rec, _ = sv.wait(sv.submit(s, rays, a))
        alone.append(rec)
    say("serving config: the same requests all at once")
    ids = [sv.submit(s, rays, a) for s, rays, a in reqs]
    mix = [sv.wait(i)[0] for i in ids]
    sv.close()
    for k, ((s, rays, a), x, y) in enumerate(zip(reqs, alone, mix)):
        if a:
            assert np.array_equal(x["occluded"], y["occluded"]), f"request {k}"
        else:
            assert np.array_equal(x["hits"].view(np.uint32), y["hits"].view(np.uint32)), f"request {k}"
            assert np.array_equal(x["objects"], y["objects"]), f"request {k}"
    say("serving config: oracle on", SAMPLE, "rays per request")
    oscs = {}
    for k, ((s, rays, a), x) in enumerate(zip(reqs, alone)):
        if s not in oscs:
            oscs[s] = oracle.OracleScene(ds[s].scene)    # with the hash seed it was built with
        got = x["occluded"] if a else (x["hits"].view(np.uint32), x["objects"])
        check_against_oracle(oscs[s], rays, a, got, f"request {k} (scene {s}, {rays.shape[0]} rays)")
