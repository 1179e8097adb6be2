"""Leaf methods of batched launches (cgs_batch_set_trace_options) and the
accuracy-gated tuner over them (cgs_batch_autotune_accuracy): a batched
launch with a method gives every job the results of cgs_trace_closest_ex /
cgs_trace_any_ex with that method, bit for bit; the tuner's reference
candidate agrees exactly, every reported error is what the kept method's
results show against the reference's, and it keeps the fastest candidate
within the limits."""
from __future__ import annotations

import numpy as np
import pytest

from conftest import say
from test_gpu_batch import results, same, upload

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def tune_batch():
    from cgspline_amd import scenes
    return scenes.batched_scenes(seed=11, count=5, patches=(256, 6_000), rays=(2_000, 20_000), any_fraction=0.3)


@pytest.fixture
def restore(ctx):
    yield
    ctx.set_batch_trace_options()


def per_job(items, jobs, **options):
    ref = []
    for it, j in zip(items, jobs):
        if it.any:
            ref.append(j["scene"].trace_any(it.rays, **options))
        else:
            h, p, o = j["scene"].trace_closest(it.rays, **options)
            hb = np.zeros((it.rays.shape[0], 4), np.uint32)
            hb[:, :3] = h.view(np.uint32)
            hb[:, 3] = p
            ref.append((hb, o))
    return ref


@pytest.mark.parametrize("options", [dict(max_clip_iters=8), dict(method=1, subdiv_depth=6),
                                     dict(polish=True), dict(method=1, subdiv_depth=4, polish=True)])
def test_batched_leaf_methods_equal_per_job(ctx, tune_batch, restore, options):
    jobs = upload(ctx, tune_batch)
    ctx.set_batch_trace_options(**options)
    got_opts = ctx.batch_trace_options()
    for k, v in options.items():
        assert got_opts[k] == v
    say("cgs_trace_batch with", options, sum(j["n"] for j in jobs), "rays")
    ctx.trace_batch(jobs)
    got = results(ctx, jobs)
    ref = per_job(tune_batch, jobs, **options)
    for k, (a, b) in enumerate(zip(got, ref)):
        assert same(a, b), (options, k)
    # the graph-captured batch takes the method set when it is created
    b = ctx.batch(jobs)
    b.launch()
    got2 = results(ctx, jobs)
    b.destroy()
    for a, c in zip(got2, ref):
        assert same(a, c)


def errors_against(ref, got, items, max_rel_t, max_uv):
    """max rel t, max uv over same-patch hits, fraction of rays outside the
    limits (as the library defines them)."""
    relT = uv = 0.0
    dis = tot = 0
    for it, a, b in zip(items, ref, got):
        tot += it.rays.shape[0]
        if it.any:
            dis += int(np.sum(a != b))
            continue
        (ha, oa), (hb, ob) = a, b
        ta, tb = ha[:, :3].view(np.float32), hb[:, :3].view(np.float32)
        pa, pb = ha[:, 3], hb[:, 3]
        hita, hitb = pa < 0xFFFFFFFE, pb < 0xFFFFFFFE
        bad = (hita != hitb) | ((hita & hitb) & (oa != ob))
        both = hita & hitb & ~bad
        # misses keep t = inf: compare distances only where both hit
        t64a = np.where(both, ta[:, 0].astype(np.float64), 1.0)
        t64b = np.where(both, tb[:, 0].astype(np.float64), 1.0)
        dt = np.abs(t64b - t64a) / np.maximum(np.abs(t64a), 1e-30)
        seam = both & (pa != pb)
        bad |= seam & (dt > max_rel_t)
        same_p = both & (pa == pb)
        duv = np.abs(tb[:, 1:3].astype(np.float64) - ta[:, 1:3]).max(axis=1)
        bad |= same_p & ((dt > max_rel_t) | (duv > max_uv))
        if same_p.any():
            relT = max(relT, float(dt[same_p].max()))
            uv = max(uv, float(duv[same_p].max()))
        dis += int(bad.sum())
    return relT, uv, dis / max(tot, 1)


@pytest.mark.parametrize("allow_polish", [True, False])
def test_accuracy_gated_autotune(ctx, tune_batch, restore, allow_polish):
    jobs = upload(ctx, tune_batch)
    lim = dict(max_rel_t=1e-4, max_uv=1e-4, max_disagree=1e-3)
    say("cgs_batch_autotune_accuracy", sum(j["n"] for j in jobs), "rays, polish", allow_polish)
    best, report = ctx.batch_autotune_accuracy(jobs, allow_polish=allow_polish, **lim)
    assert len(report) == (24 if allow_polish else 12)
    reference = dict(method=0, clip_tol=0.0, subdiv_depth=0, max_clip_iters=0, polish=allow_polish)
    refent = [e for e in report if e["options"] == reference]
    assert len(refent) == 1 and refent[0]["admissible"]
    assert refent[0]["max_rel_t"] == 0 and refent[0]["max_uv"] == 0 and refent[0]["disagree"] == 0
    adm = [e for e in report if e["admissible"]]
    assert best == min(adm, key=lambda e: e["ms"])["options"]
    for e in report:
        assert e["admissible"] == (e["disagree"] <= lim["max_disagree"])
    assert ctx.batch_trace_options() == best
    # the kept method's results, measured here against the reference's
    ref = per_job(tune_batch, jobs, polish=allow_polish)
    ctx.trace_batch(jobs)
    got = results(ctx, jobs)
    relT, uv, dis = errors_against(ref, got, tune_batch, lim["max_rel_t"], lim["max_uv"])
    ent = next(e for e in report if e["options"] == best)
    assert dis <= lim["max_disagree"] and abs(dis - ent["disagree"]) <= 1e-6
    assert abs(relT - ent["max_rel_t"]) <= 1e-6 * max(1.0, relT) and abs(uv - ent["max_uv"]) <= 1e-6
    # a capped or subdivided candidate's reported error is what its results show
    other = next(e for e in report if e["options"]["max_clip_iters"] == 8 and e["options"]["polish"] == allow_polish)
    ctx.set_batch_trace_options(**other["options"])
    ctx.trace_batch(jobs)
    relT8, uv8, dis8 = errors_against(ref, results(ctx, jobs), tune_batch, lim["max_rel_t"], lim["max_uv"])
    assert abs(relT8 - other["max_rel_t"]) <= 1e-6 * max(1.0, relT8)
    assert abs(uv8 - other["max_uv"]) <= 1e-6 and abs(dis8 - other["disagree"]) <= 1e-6
    say("autotune picked", best, "report", [(e["options"], round(e["ms"], 3), e["admissible"]) for e in report])
