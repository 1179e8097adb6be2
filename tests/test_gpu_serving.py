"""The continuous-batching server: every request's results are the same
alone and in the mix (bit for bit), requests join launches while others run
under the byte budget, and oversized requests are refused."""
from __future__ import annotations

import numpy as np
import pytest

from conftest import rng, say

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def served(ctx):
    from cgspline_amd import scenes
    items = scenes.batched_scenes(seed=7, count=3, patches=(512, 8_000), rays=(1_000, 30_000))
    ds = [ctx.scene(it.scene) for it in items]
    return items, ds


def requests_for(items, count, seed):
    r = rng(seed)
    reqs = []
    for k in range(count):
        s = int(r.integers(0, len(items)))
        pool = items[s].rays
        n = int(r.integers(1, pool.shape[0]))
        start = int(r.integers(0, pool.shape[0] - n + 1))
        reqs.append((s, pool[start:start + n].copy(), bool(r.random() < 0.4)))
    return reqs


def alone(ds, s, rays, any_hit):
    if any_hit:
        return ds[s].trace_any(rays)
    h, p, o = ds[s].trace_closest(rays)
    return h, p, o


def same(rec, ref, any_hit):
    if any_hit:
        return np.array_equal(rec["occluded"], ref)
    h, p, o = ref
    hb = rec["hits"].view(np.uint32)
    return (np.array_equal(hb[:, :3], h.view(np.uint32)) and np.array_equal(hb[:, 3], p)
            and np.array_equal(rec["objects"], o))


def test_results_alone_equal_results_in_the_mix(ctx, served):
    from cgspline_amd.serving import Server
    items, ds = served
    reqs = requests_for(items, 24, seed=5)
    refs = [alone(ds, s, r, a) for s, r, a in reqs]
    # a budget of a few requests forces several launches, each admitting what fits
    budget = 60 * max(r.shape[0] for _, r, _ in reqs) * 52
    sv = Server(ctx, ds, hbm_budget=budget, slots=2)
    say("serving", len(reqs), "requests, budget", budget)
    ids = [sv.submit(s, r, a) for s, r, a in reqs]
    for rid, (s, r, a), ref in zip(ids, reqs, refs):
        rec, lat = sv.wait(rid)
        assert lat > 0
        assert same(rec, ref, a), rid
    sv.close()


def test_requests_join_while_others_run(ctx, served):
    from cgspline_amd.serving import Server, run_load
    items, ds = served
    reqs = requests_for(items, 40, seed=6)
    sv = Server(ctx, ds, slots=2)
    lats, nrays, wall = run_load(sv, reqs, rate_hz=2000.0, seed=1)
    sv.close()
    assert lats.shape[0] == 40 and (lats > 0).all() and nrays > 0 and wall > 0


def test_oversized_request_is_refused(ctx, served):
    import cgspline_amd as cgs
    from cgspline_amd.serving import Server
    items, ds = served
    sv = Server(ctx, ds, hbm_budget=52 * 100)
    with pytest.raises(cgs.CgsError) as e:
        sv.submit(0, items[0].rays[:1000])
    assert e.value.status == 1
    with pytest.raises(cgs.CgsError):
        sv.submit(7, items[0].rays[:10])
    sv.close()
