"""Batched queries over many resident scenes (cgs_trace_batch, graph-captured
cgs_batch, the launch-shape autotuner) give every job the results of a
separate cgs_trace_closest / cgs_trace_any call, bit for bit."""
from __future__ import annotations

import numpy as np
import pytest

from conftest import say

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def small_batch():
    from cgspline_amd import scenes
    items = scenes.batched_scenes(count=6, patches=(256, 6_000), rays=(500, 20_000), any_fraction=0.35)
    items.append(scenes.BatchItem(items[0].scene, items[0].rays[:0]))   # an empty job
    return items


def upload(ctx, items):
    import cgspline_amd as cgs
    jobs, outs = [], []
    for it in items:
        ds = ctx.scene(it.scene)
        n = it.rays.shape[0]
        dr = cgs.DeviceArray.from_host(ctx, it.rays)
        if it.any:
            occ = cgs.DeviceArray(ctx, 4 * max(1, n))
            jobs.append(dict(scene=ds, rays=dr, n=n, any=True, occluded=occ))
        else:
            dh, do = cgs.DeviceArray(ctx, 16 * max(1, n)), cgs.DeviceArray(ctx, 4 * max(1, n))
            jobs.append(dict(scene=ds, rays=dr, n=n, hits=dh, objects=do))
        outs.append(it)
    return jobs


def results(ctx, jobs):
    ctx.sync()
    out = []
    for j in jobs:
        n = j["n"]
        if j.get("any"):
            out.append(j["occluded"].to_host(np.uint32, (n,)))
        else:
            out.append((j["hits"].to_host(np.uint32, (n, 4)), j["objects"].to_host(np.int32, (n,))))
    return out


def reference(items, jobs):
    ref = []
    for it, j in zip(items, jobs):
        if it.any:
            ref.append(j["scene"].trace_any(it.rays))
        else:
            h, p, o = j["scene"].trace_closest(it.rays)
            hb = np.zeros((it.rays.shape[0], 4), np.uint32)
            hb[:, :3] = h.view(np.uint32)
            hb[:, 3] = p
            ref.append((hb, o))
    return ref


def same(a, b):
    if isinstance(a, tuple):
        return np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])
    return np.array_equal(a, b)


def scrub(ctx, jobs):
    import cgspline_amd as cgs
    for j in jobs:
        for key in ("hits", "objects", "occluded"):
            if j.get(key) is not None:
                cgs.check(cgs.lib().cgs_memset(ctx.h, j[key].dptr, 0x5A, j[key].nbytes, None))
    ctx.sync()


def test_batch_equals_separate_queries(ctx, small_batch):
    say("batch", len(small_batch), "jobs", sum(i.rays.shape[0] for i in small_batch), "rays")
    jobs = upload(ctx, small_batch)
    ref = reference(small_batch, jobs)
    for chunk, bpc in ((64, 0), (256, 1)):
        scrub(ctx, jobs)
        ctx.set_batch_config(chunk, bpc)
        ctx.trace_batch(jobs)
        got = results(ctx, jobs)
        assert all(same(g, r) for g, r in zip(got, ref)), (chunk, bpc)
    ctx.set_batch_config(64, 0)


def test_batch_of_deep_traversals_equals_oracle(ctx, oracle, grazing, stacked):
    """Jobs whose traversals overflow the sibling ring, in the batched
    kernel, which stages no BVH levels in LDS (every sibling the ring has
    overwritten comes from the perfect hash): grazing rays on the 128×128-patch
    height field, and rays through 180,000 stacked objects, closest and any
    hit.  Every job equals the oracle."""
    from cgspline_amd import scenes
    sc_s, rays_s = stacked
    rays_g = scenes.camera_rays(grazing.camera)[::3]
    items = [scenes.BatchItem(grazing.scene, rays_g), scenes.BatchItem(sc_s, rays_s),
             scenes.BatchItem(sc_s, rays_s, any=True)]
    say("batch of deep traversals", sum(i.rays.shape[0] for i in items), "rays")
    jobs = upload(ctx, items)
    scrub(ctx, jobs)
    ctx.trace_batch(jobs)
    got = results(ctx, jobs)
    for it, g in zip(items, got):
        osc = oracle.OracleScene(it.scene)
        if it.any:
            o, w = osc.trace(it.rays, any_hit=True)
            assert np.array_equal(g == 1, o) and w["maxPostponed"] > 8
        else:
            h, p, o, w = osc.trace(it.rays)
            assert np.array_equal(g[1], o) and np.array_equal(g[0][:, 3], p)
            assert np.array_equal(g[0][:, :3], h.view(np.uint32)) and w["maxPostponed"] > 8


def test_graph_replays_match_eager(ctx, small_batch):
    jobs = upload(ctx, small_batch)
    ref = reference(small_batch, jobs)
    b = ctx.batch(jobs)
    for rep in range(3):
        scrub(ctx, jobs)
        say("graph replay", rep)
        b.launch()
        got = results(ctx, jobs)
        assert all(same(g, r) for g, r in zip(got, ref)), rep
        # the program's own eager work between replays
        reference(small_batch[:1], jobs[:1])
    b.destroy()


def test_autotune_keeps_results(ctx, small_batch):
    jobs = upload(ctx, small_batch)
    ref = reference(small_batch, jobs)
    chunk, bpc = ctx.batch_autotune(jobs)
    assert chunk in (64, 128, 256) and bpc >= 0
    assert ctx.batch_config() == (chunk, bpc)
    scrub(ctx, jobs)
    ctx.trace_batch(jobs)
    assert all(same(g, r) for g, r in zip(results(ctx, jobs), ref))
    ctx.set_batch_config(64, 0)


def test_batch_rejects_bad_jobs(ctx, small_batch):
    import cgspline_amd as cgs
    jobs = upload(ctx, small_batch[:1])
    jobs[0]["hits"] = None
    jobs[0]["any"] = False
    with pytest.raises(cgs.CgsError) as e:
        ctx.trace_batch(jobs)
    assert e.value.status == 1
    with pytest.raises(cgs.CgsError):
        ctx.set_batch_config(100, 0)
