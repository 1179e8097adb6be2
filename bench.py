"""Benchmark: Mrays/s of the ray–patch renderer on the 1M-patch scene.

One step = one render launch (RayTracer::render semantics) of the
displaced-grid workload: 4,096 displaced 16×16-patch grids (1,048,576 bicubic
patches), 2048×2048 primary rays, one shadow ray per lit hit, one reflection
bounce (+ its shadow ray), GGX shading, fp32.  The scene (BVHs and perfect
hashes) is built and resident in HBM before the timed region; `value` counts
every ray the launch traces (primary + shadow + secondary).

Prints one JSON line.  Besides the contract's fields:

* `roofline`: the render kernel against its bound, FP32 VALU: FP32
  operations per launch (the library's own event counts, from its
  CGS_WORK_COUNTS build rendering the same frame in a child process, × the
  per-event operation counts of cgspline_amd/workmodel.py) ÷ the kernel's
  time measured with HIP events on the launch stream; beside it the
  reference algorithm's operations on the same rays (the oracle's counts).  `traffic` is the measured HBM bytes
  per launch (profiles/kRender_displaced_grids_hbm_traffic.md) and `hbm` that
  traffic over the same kernel time; `issue` gives the SQ counters of
  profiles/kRender_displaced_grids_valu.md.
* `cpu_baseline`: the CPU restatement of the reference's CPURayTracer, built
  with -march=native on this host, on the same frame.
* `first_render`, `pcie_inclusive`: the scene's first render (tiles row by
  row) and the render with the frame downloaded after each step.
* `moving_camera`: an orbit of the same view, a small step per frame, each
  render taking its tiles in the order the previous frame measured.
* `batched`: 64 resident scenes (1k–200k patches, 4k–2M rays), one
  persistent batched launch against one trace call per scene on the same rays;
  `accuracy_tuned`: the leaf method cgs_batch_autotune_accuracy keeps within
  1e-4 of the FP64-polished reference, and its pass time.
* `single_bvh`: the same 1M patches as one BVH (hashed with the wide family,
  patches found through a cull tree over its leaf boxes), same frame.
* `sampled`: the Monte Carlo render of the same frame, 4 camera rays per
  pixel with GGX-sampled reflections, and the oracle's sampled render on 8
  rows beside it.
* `serving`: 8 resident scenes (one of them this 1M-patch scene) under
  Poisson arrivals of 1k–4M-ray requests (40 % occlusion, 20 % shaded, the
  rest closest hits) at 30/60/90 % of capacity: p50/p99 latency and sustained
  Mrays/s.

Each render records its tiles' times and the next render of the same frame
size takes the tiles longest first; the warmup renders set that order up, and
every timed render still traces every ray.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


def measured_traffic(workload):
    """HBM bytes per kRender launch from the committed PMC summary of this
    workload (FETCH_SIZE and WRITE_SIZE passes, each scaled by the same
    counter's reading on a streaming copy of known size), or None."""
    path = os.path.join(ROOT, "profiles", f"kRender_{workload}_hbm_traffic.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        return int(json.load(f)["traffic_bytes_per_launch"])


def measured_issue(workload):
    """The SQ counter summary of kRender on this workload
    (profiles/kRender_<workload>_valu.json), or None."""
    path = os.path.join(ROOT, "profiles", f"kRender_{workload}_valu.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        return json.load(f)


_COUNT_CHILD = r"""
import json, os, sys
sys.path.insert(0, os.environ["CGS_ROOT"])
import ctypes as C
import cgspline_amd as cgs
from cgspline_amd import scenes, workmodel
a = json.loads(sys.argv[1])
if a["scene_file"]:
    wl = scenes.from_dl(a["scene_file"], width=a["width"], height=a["height"], depth=a["depth"])
else:
    wl = scenes.WORKLOADS[a["workload"]]()
ctx = cgs.Context(0)
ds = ctx.scene(wl.scene)
out = (C.c_uint64 * len(workmodel.GPU_EVENTS))()
cgs.check(cgs.lib().cgs_work_counts(ctx.h, out, len(out), 1), "cgs_work_counts")
color, frame, counts = ds.render(wl.camera, eps=wl.eps, depth=wl.depth)
cgs.check(cgs.lib().cgs_work_counts(ctx.h, out, len(out), 1), "cgs_work_counts")
print(json.dumps({"events": dict(zip(workmodel.GPU_EVENTS, [int(x) for x in out])), "rays": counts}))
"""


def gpu_work_counts(args):
    """Work events of one render of the workload, counted by the library's
    CGS_WORK_COUNTS=1 build (cgspline_amd/libcgs_counts.so) in a child
    process (the counting build is never the one timed), or None."""
    import subprocess
    from cgspline_amd import _build
    if not os.path.exists(_build.COUNTING_LIB):
        return None
    env = dict(os.environ, CGS_LIBRARY=_build.COUNTING_LIB, CGS_ROOT=ROOT)
    a = {"workload": args.workload, "scene_file": args.scene_file, "width": args.width, "height": args.height,
         "depth": args.depth}
    r = subprocess.run([sys.executable, "-c", _COUNT_CHILD, json.dumps(a)], env=env, stdout=subprocess.PIPE,
                       stderr=subprocess.PIPE, text=True, timeout=600)
    if r.returncode != 0:
        log(f"work count child failed ({r.returncode}): {r.stderr[-400:]}")
        return None
    return json.loads(r.stdout.strip().splitlines()[-1])


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def native_oracle():
    """Builds the oracle with -march=native for this host's CPU baseline (the
    reference's CXX flags, CMakeLists.txt:83) into a temporary directory;
    returns (path, flags), or (None, reason) when the build fails and the
    prebuilt x86-64-v3 library is used instead."""
    import subprocess
    import tempfile
    out = os.path.join(tempfile.mkdtemp(prefix="cgs_bench_"), "liboracle_native.so")
    flags = ["-O3", "-march=native", "-ffp-contract=off", "-fno-fast-math", "-std=c++20", "-fPIC"]
    r = subprocess.run(["g++", *flags, "-shared", "-o", out, os.path.join(ROOT, "oracle", "cgs_oracle.cpp"),
                        "-lpthread"], stdout=subprocess.PIPE, stderr=subprocess.STDOUT)
    if r.returncode != 0:
        return None, "g++ -march=native failed; the prebuilt -march=x86-64-v3 oracle was used"
    return out, " ".join(flags)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="displaced_grids")
    ap.add_argument("--cpu-seconds", type=float, default=12.0, help="target CPU-baseline sample duration")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-work-count", action="store_true", help="skip the oracle's work count (no roofline)")
    ap.add_argument("--mode", choices=["plain", "batched", "serving"], default="plain",
                    help="plain: the 1M-patch render (the headline line); batched: 64 resident scenes traced "
                         "per job vs in one persistent launch; serving: continuous batching under Poisson load")
    ap.add_argument("--scenes", type=int, default=None,
                    help="batched/serving: number of resident scenes (default 64 batched, 8 serving)")
    ap.add_argument("--slots", type=int, default=3, help="serving: launch slots (streams) of the server")
    ap.add_argument("--max-batch-rays", type=int, default=1 << 22, help="serving: rays per launch")
    ap.add_argument("--scene-file", default=None,
                    help="plain mode on one of the reference's .dl scene files instead of the 1M-patch scene "
                         "(e.g. tests/golden/assets/scenes/Test152.dl), seen from the file's view")
    ap.add_argument("--width", type=int, default=1920, help="--scene-file: frame width")
    ap.add_argument("--height", type=int, default=1080, help="--scene-file: frame height")
    ap.add_argument("--depth", type=int, default=6, help="--scene-file: recursion depth (reference default 6)")
    ap.add_argument("--samples", type=int, default=0,
                    help="time the sampled (Monte Carlo) render with this many camera rays per pixel instead of the "
                         "plain render (no roofline, CPU baseline or extra legs)")
    ap.add_argument("--no-extras", dest="extras", action="store_false",
                    help="plain mode: skip the moving-camera, batched and serving sub-objects")
    ap.add_argument("--calibration-copy-mib", type=int, default=0,
                    help="also run one streaming copy kernel (bitwise_not over int32) of this many MiB after the "
                         "warmup: the reference kernel that scales the HBM byte counters (FETCH_SIZE, WRITE_SIZE) "
                         "of a profile")
    args = ap.parse_args()

    if args.gpus != 1:
        print(f"bench.py: the library runs on one GPU; --gpus {args.gpus} is not supported "
              "(rays are independent: the multi-GPU form replicates the scene and splits the frame).",
              file=sys.stderr)
        sys.exit(2)
    if args.mode == "batched":
        args.scenes = args.scenes or 64
        return run_batched(args)
    if args.mode == "serving":
        args.scenes = args.scenes or 8
        return run_serving(args)

    import torch
    import cgspline_amd as cgs
    from cgspline_amd import scenes, workmodel

    t0 = time.time()
    if args.scene_file:
        wl = scenes.from_dl(args.scene_file, width=args.width, height=args.height, depth=args.depth)
        args.workload = os.path.splitext(os.path.basename(args.scene_file))[0]
        metric = f"Mrays/s on the reference scene {args.workload} (plain launch)"
        data = f"the reference's scene file {os.path.basename(args.scene_file)} and its .be models"
    else:
        wl = scenes.WORKLOADS[args.workload]()
        metric = "Mrays/s at 1M-patch scene (plain launch)"
        data = "synthetic (seeded displaced bicubic grids)"
    if args.samples > 0:
        metric = metric.replace("(plain launch)", f"(sampled render, {args.samples} samples per pixel)")
        args.no_work_count = args.no_cpu_baseline = True
        args.extras = False
    cam = wl.camera
    log(f"workload {args.workload}: {wl.scene.npatch} patches, {wl.scene.nobj} objects, "
        f"{cam.width}x{cam.height}, depth {wl.depth} (generated in {time.time() - t0:.1f}s)")

    ctx = cgs.Context(0)
    t0 = time.time()
    ds = ctx.scene(wl.scene)
    build_s = time.time() - t0
    log(f"scene resident: BVH + perfect-hash build {build_s * 1e3:.0f} ms")

    dev = torch.device("cuda:0")
    stream = torch.cuda.current_stream(dev)
    sh = stream.cuda_stream
    frame = torch.empty(cam.width * cam.height, dtype=torch.int32, device=dev)
    counts = torch.zeros(4, dtype=torch.int64, device=dev)

    def step():
        if args.samples > 0:
            st = ds.render_sampled_dev(cam, frame, None, counts, samples=args.samples, seed=42, eps=wl.eps,
                                       depth=wl.depth, stream=sh)
        else:
            st = ds.render_dev(cam, frame, None, counts, eps=wl.eps, depth=wl.depth, stream=sh)
        cgs.check(st, "cgs_render")

    # The scene's first render at this size takes the tiles row by row; every
    # later one takes them longest first by the previous render's tile times.
    # The first warmup render is timed on its own to report that case.
    first_ms = None
    for w in range(args.warmup):
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        step()
        e1.record(stream)
        torch.cuda.synchronize()   # render refuses a second launch while one is in flight
        if w == 0:
            first_ms = e0.elapsed_time(e1)
    counts.zero_()
    torch.cuda.synchronize()
    if args.calibration_copy_mib:
        nb = args.calibration_copy_mib << 20
        src = torch.ones(nb // 4, dtype=torch.int32, device=dev)
        dst = torch.empty_like(src)
        torch.bitwise_not(src, out=dst)
        torch.cuda.synchronize()
        del src, dst
        log(f"calibration copy: {nb} bytes read, {nb} bytes written")

    # Timed region: K launches, each bracketed by HIP events on the launch stream.
    ev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in range(args.steps)]
    torch.cuda.synchronize()
    wall0 = time.perf_counter()
    for i in range(args.steps):
        ev[i][0].record(stream)
        step()
        ev[i][1].record(stream)
        ev[i][1].synchronize()
    torch.cuda.synchronize()
    wall = time.perf_counter() - wall0
    kernel_ms = [a.elapsed_time(b) for a, b in ev]
    c = counts.cpu().numpy().astype(np.int64)
    rays_total = int(c[0] + c[1] + c[2])
    rays_per_step = rays_total / args.steps
    assert int(c[3]) == 0, "clip stack overflow in the benchmark workload"
    ms_step = wall * 1e3 / args.steps
    value = rays_total / wall / 1e6
    avg_kernel_s = float(np.mean(kernel_ms)) / 1e3
    log(f"{args.steps} steps: {ms_step:.2f} ms/step wall, kernel {np.mean(kernel_ms):.2f} ms "
        f"(min {np.min(kernel_ms):.2f}), {rays_per_step:.0f} rays/step -> {value:.1f} Mrays/s")

    # Outside the timed region: the same frames with each one downloaded to
    # pinned host memory after its render (what a caller that shows or saves
    # every frame pays over PCIe).
    host_frame = torch.empty(frame.numel(), dtype=torch.int32, pin_memory=True)
    dl_steps = max(1, min(args.steps, 5))
    torch.cuda.synchronize()
    t_dl = time.perf_counter()
    for _ in range(dl_steps):
        step()
        host_frame.copy_(frame, non_blocking=True)
        torch.cuda.synchronize()
    dl_ms = (time.perf_counter() - t_dl) * 1e3 / dl_steps
    log(f"render + frame download ({frame.numel() * 4} B): {dl_ms:.2f} ms/frame")

    out = {
        "metric": metric,
        "value": round(value, 3),
        "unit": "Mrays/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": data,
        "config": {
            "workload": args.workload,
            "patches": wl.scene.npatch,
            "objects": wl.scene.nobj,
            "width": cam.width,
            "height": cam.height,
            "recursion_depth": wl.depth,
            **({"samples_per_pixel": args.samples, "seed": 42} if args.samples > 0 else {}),
            "rays_per_step": rays_per_step,
            "primary_shadow_secondary_per_step": [int(c[0]) // args.steps, int(c[1]) // args.steps,
                                                  int(c[2]) // args.steps],
            "scene_build_ms": round(build_s * 1e3, 1),
        },
        "first_render": None if first_ms is None else {
            "ms": round(first_ms, 4),
            "value": round(rays_per_step / (first_ms / 1e3) / 1e6, 3),
            "unit": "Mrays/s",
            "note": "the scene's first render at this size (tiles row by row, one-time setup included); "
                    "later renders take the tiles longest first by the previous render's tile times",
        },
        "pcie_inclusive": {
            "value": round(rays_per_step / (dl_ms / 1e3) / 1e6, 3),
            "unit": "Mrays/s",
            "ms_per_step": round(dl_ms, 4),
            "frame_bytes": frame.numel() * 4,
            "note": "render then frame download to pinned host memory, outside the timed region",
        },
    }

    osc = None
    threads = max(1, min(16, os.cpu_count() or 1))
    if not args.no_work_count or not args.no_cpu_baseline:
        build = "prebuilt -O3 -march=x86-64-v3 -ffp-contract=off"
        if not args.no_cpu_baseline:
            t0 = time.time()
            lib, build = native_oracle()
            if lib:
                os.environ["CGS_ORACLE_LIB"] = lib
            log(f"oracle for the CPU baseline: {build} ({time.time() - t0:.1f}s)")
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import cgs_oracle
        osc = cgs_oracle.OracleScene(wl.scene)

    if not args.no_work_count:
        t0 = time.time()
        _, _, work, rc = osc.render(cam, eps=wl.eps, depth=wl.depth, threads=threads)
        log(f"oracle work count ({time.time() - t0:.1f}s): {work}")
        oracle_rays = rc["primary"] + rc["shadow"] + rc["secondary"]
        if oracle_rays != int(round(rays_per_step)):
            log(f"WARNING: oracle ray count {oracle_rays} != GPU {rays_per_step}")
        ref_flops = workmodel.flops_of(work)
        t0 = time.time()
        gc = gpu_work_counts(args)
        log(f"library work count ({time.time() - t0:.1f}s): {gc}")
        if gc is not None:
            events = gc["events"]
            nbytes = workmodel.bytes_of_counts(events, cam.width * cam.height)
            nflops = workmodel.flops_of_counts(events)
            source = "the library's own event counts (its CGS_WORK_COUNTS build, one render of this frame)"
        else:
            events = None
            nbytes = workmodel.bytes_of(work, cam.width * cam.height, work["objects"])
            nflops = ref_flops
            source = "the oracle's event counts of the reference algorithm"
        tflops = nflops / avg_kernel_s / 1e12
        traffic = measured_traffic(args.workload)
        roof = {
            "bound": "valu",
            "achieved": round(tflops, 3),
            "peak": workmodel.FP32_PEAK_TFLOPS,
            "unit": "TFLOP/s",
            "frac": round(tflops / workmodel.FP32_PEAK_TFLOPS, 4),
            "traffic": traffic,
            "kernel": "kRender",
            "kernel_ms": round(avg_kernel_s * 1e3, 4),
            "algorithmic_flops_per_launch": int(nflops),
            "peak_nonpacked": workmodel.FP32_PEAK_NONPACKED_TFLOPS,
            "frac_nonpacked": round(tflops / workmodel.FP32_PEAK_NONPACKED_TFLOPS, 4),
            "reference_algorithm_flops_per_launch": int(ref_flops),
            "reference_equivalent_tflops": round(ref_flops / avg_kernel_s / 1e12, 3),
            "note": "FP32 VALU bound: FP32 operations of the launch (" + source + " x the per-event costs of "
                    "workmodel.py) over the kernel time measured with HIP events; 157.3 TFLOP/s assumes packed "
                    "FP32, which this build disables, so 78.6 is the attainable peak.  The reference algorithm "
                    "on the same rays (its top-level BVH walked node by node, which the cull tree replaces) "
                    "would take reference_algorithm_flops_per_launch.",
        }
        if events is not None:
            roof["events"] = events
        if traffic is not None:
            roof["hbm"] = {"achieved": round(traffic / avg_kernel_s / 1e9, 2), "peak": workmodel.HBM_PEAK_GBS,
                           "unit": "GB/s", "frac": round(traffic / avg_kernel_s / 1e9 / workmodel.HBM_PEAK_GBS, 4),
                           "note": "measured HBM bytes per launch (FETCH_SIZE / 0.5 + WRITE_SIZE, "
                                   "profiles/kRender_displaced_grids_hbm_traffic.md) over the kernel time; "
                                   f"the algorithmic node, hash and patch bytes are {nbytes / 1e9:.1f} GB, "
                                   "served from L2 and the MALL"}
        issue = measured_issue(args.workload)
        if issue is not None:
            roof["issue"] = issue
        roof["work_counts"] = work
        out["roofline"] = roof

    if not args.no_cpu_baseline:
        # bounded sample: every `stride`-th row of the frame (uniform over the
        # image), the stride shrunk until one render takes about --cpu-seconds;
        # when the whole frame takes less, it is rendered repeatedly (at most 9
        # times, the reference harness's run count) and the median run kept
        h = cam.height
        stride = 256
        while True:
            sec, _, _, nr, ns = osc.cpu_render(cam, eps=wl.eps, depth=wl.depth, rows=(0, h), threads=threads,
                                               stride=stride)
            if sec * 2 > args.cpu_seconds or stride == 1:
                break
            stride = max(1, int(stride / max(2.0, min(8.0, args.cpu_seconds / max(sec, 1e-3) * 0.8))))
        runs = [sec]
        while sum(runs) < args.cpu_seconds and len(runs) < 9:
            runs.append(osc.cpu_render(cam, eps=wl.eps, depth=wl.depth, rows=(0, h), threads=threads,
                                       stride=stride)[0])
        sec = float(np.median(runs))
        nrows = (h + stride - 1) // stride
        cpu_value = (nr + ns) / sec / 1e6
        log(f"cpu baseline: every {stride}th row ({nrows} rows): {nr + ns} rays, median of {len(runs)} runs "
            f"{sec:.2f}s on {threads} threads -> {cpu_value:.3f} Mrays/s")
        what = "the whole" if stride == 1 else f"every {stride}th row of the"
        out["cpu_baseline"] = {
            "value": round(cpu_value, 4),
            "unit": "Mrays/s",
            "cores": threads,
            "kind": "port",
            "sample": f"{what} {cam.width}x{h} frame ({nrows * cam.width} pixels, {nr + ns} rays per run), "
                      f"median of {len(runs)} runs ({sum(runs):.1f}s CPU wall in total): CPURayTracer "
                      f"restatement (32x32 tiles, {threads} worker threads), same scene",
            "host_cpus": os.cpu_count(),
            "cpu_model": cpu_model(),
            "build": build,
            "threads_note": f"{threads} threads: the host CPU share one GPU is allotted on this machine "
                            f"({os.cpu_count()} logical CPUs are visible)",
            "gpu_over_cpu": round(value / cpu_value, 2),
        }

    if args.extras and not args.scene_file:
        out["moving_camera"] = moving_camera(ds, wl, frame, counts, stream, args)
        out["batched"] = batched_leg(ctx, args)
        out["serving"] = serving_leg(ctx, args, ds, wl)
        out["single_bvh"] = single_bvh_leg(ctx, args, frame, counts, stream)
        out["sampled"] = sampled_leg(ds, wl, frame, counts, stream, args, osc, threads)
        if "batched" in out and out.get("value"):
            out["batched"]["batched_over_plain_render"] = round(out["batched"]["value"] / value, 3)
    print(json.dumps(out), flush=True)


def sampled_leg(ds, wl, frame, counts, stream, args, osc, threads, samples=4):
    """The Monte Carlo render (the reference's USE_MONTECARLO_SAMPLING build)
    of the headline frame: `samples` camera rays per pixel, GGX-sampled
    reflections, same scene, camera and depth; rays counted as the plain
    render counts them.  Beside it, the oracle's sampled render on a few rows
    of the same frame (the CPU restatement of the GPU path)."""
    import torch
    import cgspline_amd as cgs
    sh = stream.cuda_stream

    def render():
        cgs.check(ds.render_sampled_dev(wl.camera, frame, None, counts, samples=samples, seed=42, eps=wl.eps,
                                        depth=wl.depth, stream=sh), "cgs_render_sampled")

    for _ in range(2):
        render()
        torch.cuda.synchronize()
    counts.zero_()
    torch.cuda.synchronize()
    steps = max(3, min(args.steps, 6))
    total = 0.0
    for _ in range(steps):
        a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        a.record(stream)
        render()
        b.record(stream)
        b.synchronize()
        total += a.elapsed_time(b)
    ms = total / steps
    cnt = counts.cpu().numpy().astype(np.int64)
    rays = int(cnt[0] + cnt[1] + cnt[2]) // steps
    res = {"value": round(rays / ms / 1e3, 3), "unit": "Mrays/s", "samples_per_pixel": samples, "seed": 42,
           "ms_per_step": round(ms, 4), "rays_per_step": rays,
           "primary_shadow_secondary_per_step": [int(c) // steps for c in cnt[:3]], "overflow": int(cnt[3])}
    if osc is not None:
        h = wl.camera.height
        rows = (h // 2 - 4, h // 2 + 4)
        t0 = time.perf_counter()
        _, _, orc = osc.render_sampled(wl.camera, samples=samples, seed=42, eps=wl.eps, depth=wl.depth, rows=rows,
                                       threads=threads)
        sec = time.perf_counter() - t0
        orays = orc["primary"] + orc["shadow"] + orc["secondary"]
        res["cpu_oracle"] = {"value": round(orays / sec / 1e6, 3), "unit": "Mrays/s", "cores": threads,
                             "sample": f"rows {rows[0]}-{rows[1] - 1} of the frame, {orays} rays, {sec:.2f}s",
                             "kind": "port"}
    log(f"sampled render: {res}")
    return res


def single_bvh_leg(ctx, args, frame, counts, stream):
    """The same 1,048,576 displaced-grid patches as one BVH under one object
    (the configuration as the survey states it): the reference's hash family
    cannot index its 2,097,151 nodes, so the scene hashes it with the wide
    family, and a cull tree over its leaf boxes finds the patches a ray
    reaches.  Same frame, camera and shading as the headline."""
    import torch
    import cgspline_amd as cgs
    from cgspline_amd import scenes
    t0 = time.time()
    wl = scenes.displaced_grids_single()
    t1 = time.time()
    ds = ctx.scene(wl.scene)
    torch.cuda.synchronize()
    build_ms = (time.time() - t1) * 1e3
    sh = stream.cuda_stream

    def render():
        cgs.check(ds.render_dev(wl.camera, frame, None, counts, eps=wl.eps, depth=wl.depth, stream=sh), "cgs_render")

    for _ in range(max(2, args.warmup)):
        render()
        torch.cuda.synchronize()   # render refuses a second launch while one is in flight
    counts.zero_()
    torch.cuda.synchronize()
    steps = max(3, min(args.steps, 10))
    total = 0.0
    for _ in range(steps):
        a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        a.record(stream)
        render()
        b.record(stream)
        b.synchronize()
        total += a.elapsed_time(b)
    ms = total / steps
    cnt = counts.cpu().numpy().astype(np.int64)
    rays = int(cnt[0] + cnt[1] + cnt[2]) // steps
    res = {"value": round(rays / ms / 1e3, 3), "unit": "Mrays/s", "ms_per_step": round(ms, 4), "rays_per_step": rays,
           "patches": wl.scene.npatch, "objects": wl.scene.nobj, "bvh_nodes": int(ds.bvh_info(0)[0]),
           "hash_family": "wide" if ds.bvh_wide(0) else "reference", "surface_cull_tree": ds.bvh_culled(0),
           "scene_build_ms": round(build_ms, 1), "generated_s": round(t1 - t0, 1),
           "note": "one BVH over all patches in world space (one identity object, the first grid's material); "
                   "results equal to the reference traversal (the oracle's) on sampled rays in the GPU tests"}
    log(f"single BVH: {res}")
    ds.destroy()
    return res


def moving_camera(ds, wl, frame, counts, stream, args):
    """An orbit about the vertical through the point of the view axis nearest
    the objects' mean position, 0.25 degrees per frame: each render takes its
    tiles in the order the previous frame measured (the first in the static
    view's order)."""
    import math as _m
    import torch
    import cgspline_amd as cgs
    from cgspline_amd import scenes
    cam0 = wl.camera
    pos = np.asarray(cam0.position, np.float64)
    M = cam0.to_world.reshape(4, 4).astype(np.float64)
    fwd = -M[2, :3] / np.linalg.norm(M[2, :3])
    mean = wl.scene.obj_local2world[:, 12:15].astype(np.float64).mean(axis=0)
    centre = pos + max(1.0, float((mean - pos) @ fwd)) * fwd
    rel = pos - centre
    frames = max(4, args.steps)
    cams = []
    for k in range(1, frames + 1):
        a = _m.radians(0.25 * k)
        R = np.array([[_m.cos(a), -_m.sin(a), 0.0], [_m.sin(a), _m.cos(a), 0.0], [0.0, 0.0, 1.0]])
        p = centre + R @ rel
        cams.append(scenes.make_camera(p, centre, (0.0, 0.0, 1.0), _m.degrees(cam0.fovy), cam0.width, cam0.height))
    sh = stream.cuda_stream
    counts.zero_()
    torch.cuda.synchronize()
    ms = []
    t0 = time.perf_counter()
    for c in cams:
        a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        a.record(stream)
        cgs.check(ds.render_dev(c, frame, None, counts, eps=wl.eps, depth=wl.depth, stream=sh), "cgs_render")
        b.record(stream)
        b.synchronize()
        ms.append(a.elapsed_time(b))
    wall = time.perf_counter() - t0
    cnt = counts.cpu().numpy().astype(np.int64)
    rays = int(cnt[0] + cnt[1] + cnt[2])
    res = {"value": round(rays / wall / 1e6, 3), "unit": "Mrays/s", "frames": frames,
           "ms_per_frame": round(wall * 1e3 / frames, 4), "kernel_ms_mean": round(float(np.mean(ms)), 4),
           "kernel_ms_max": round(float(np.max(ms)), 4), "degrees_per_frame": 0.25,
           "note": "orbit about a point on the view axis; each render takes its tiles in the order measured "
                   "by the previous frame"}
    log(f"moving camera: {res}")
    return res


def batched_leg(ctx, args):
    """The 64-scene batched workload: one persistent batched launch (and its
    autotuned shape and graph replay) against one cgs_trace_closest per scene
    on the same rays, results compared."""
    import torch
    import cgspline_amd as cgs
    from cgspline_amd import scenes
    t0 = time.time()
    items = scenes.batched_scenes()
    npatch = sum(i.scene.npatch for i in items)
    nray = sum(i.rays.shape[0] for i in items)
    jobs = []
    for it in items:
        d = ctx.scene(it.scene, hash_retries=8)
        n = it.rays.shape[0]
        jobs.append(dict(scene=d, rays=cgs.DeviceArray.from_host(ctx, it.rays), n=n,
                         hits=cgs.DeviceArray(ctx, 16 * n), objects=cgs.DeviceArray(ctx, 4 * n)))
    log(f"batched: {len(items)} scenes, {npatch} patches, {nray} rays resident ({time.time() - t0:.1f}s)")
    stream = torch.cuda.current_stream(torch.device("cuda:0"))
    sh = stream.cuda_stream
    reps = max(3, min(args.steps, 10))

    def timed(fn):
        fn()
        torch.cuda.synchronize()
        a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        a.record(stream)
        for _ in range(reps):
            fn()
        b.record(stream)
        b.synchronize()
        return a.elapsed_time(b) / reps

    def per_job():
        for j in jobs:
            j["scene"].trace_closest_dev(j["rays"], j["n"], j["hits"], j["objects"], stream=sh)

    ms_per_job = timed(per_job)
    ref = [j["hits"].to_host(np.uint32, (j["n"], 4)) for j in jobs]
    tuned = ctx.batch_autotune(jobs)
    ms_batch = timed(lambda: ctx.trace_batch(jobs, stream=sh))
    same = all(np.array_equal(j["hits"].to_host(np.uint32, (j["n"], 4)), r) for j, r in zip(jobs, ref))
    graph = ctx.batch(jobs)
    ms_graph = timed(lambda: graph.launch(sh))
    graph.destroy()
    best = min(ms_batch, ms_graph)
    # the leaf method tuned under the accuracy limits (t and (u,v) within
    # 1e-4 of the FP64-polished reference, at most 1e-4 of rays disagreeing)
    t0 = time.time()
    limits = dict(max_rel_t=1e-4, max_uv=1e-4, max_disagree=1e-4)
    picked, report = ctx.batch_autotune_accuracy(jobs, allow_polish=True, **limits)
    ms_acc = timed(lambda: ctx.trace_batch(jobs, stream=sh))
    ctx.set_batch_trace_options()
    ent = next(e for e in report if e["options"] == picked)
    ref_ent = next(e for e in report if e["options"]["method"] == 0 and e["options"]["max_clip_iters"] == 0
                   and e["options"]["polish"])
    acc = {"limits": limits, "picked": picked, "ms_per_pass": round(ms_acc, 4),
           "value": round(nray / ms_acc / 1e3, 3), "max_rel_t": ent["max_rel_t"], "max_uv": ent["max_uv"],
           "disagree": ent["disagree"], "reference_polished_ms": round(ref_ent["ms"], 4),
           "candidates": [{"options": {k: v for k, v in e["options"].items() if v}, "ms": round(e["ms"], 3),
                           "max_rel_t": float(f"{e['max_rel_t']:.3g}"), "max_uv": float(f"{e['max_uv']:.3g}"),
                           "disagree": float(f"{e['disagree']:.3g}"), "admissible": e["admissible"]}
                          for e in report],
           "tune_s": round(time.time() - t0, 1),
           "note": "cgs_batch_autotune_accuracy over clip caps 8-32, subdivision depths 4-12 and the reference "
                   "clipping, each with and without the FP64 polish; errors against the reference clipping with "
                   "the polish (north_star's 1e-4 on t and (u,v) is met only by polished candidates)"}
    log(f"batched, accuracy-tuned: {picked} {ms_acc:.3f} ms ({acc['tune_s']} s to tune)")
    res = {"value": round(nray / best / 1e3, 3), "unit": "Mrays/s", "scenes": len(items), "patches": npatch,
           "rays": nray, "ms_per_pass": round(best, 4), "per_job_value": round(nray / ms_per_job / 1e3, 3),
           "batched_over_per_job": round(ms_per_job / best, 3),
           "forms_ms": {"per_job": round(ms_per_job, 3), "batched": round(ms_batch, 3), "graph": round(ms_graph, 3)},
           "autotuned_shape": {"chunk": tuned[0], "blocks_per_cu": tuned[1]},
           "results_equal_per_job": bool(same),
           "accuracy_tuned": acc,
           "note": "closest hits of half camera, half random rays; the batched results equal the per-job "
                   "results bit for bit (same per-ray algorithm), so the accuracy is equal"}
    log(f"batched: {res}")
    for j in jobs:
        j["scene"].destroy()
    return res


def serving_leg(ctx, args, big_ds, big_wl):
    """8 resident scenes (this 1M-patch scene and 7 of 10k-1M patches):
    capacity, then Poisson arrivals at 30/60/90 % of it."""
    from cgspline_amd import scenes
    from cgspline_amd.serving import Server, request_sizes, run_load
    t0 = time.time()
    items = scenes.batched_scenes(seed=20261202 + 5, count=7, patches=(10_000, 1_000_000), rays=(1 << 20, 1 << 21))
    ds = [ctx.scene(it.scene, hash_retries=8) for it in items] + [big_ds]
    pools = [it.rays for it in items] + [scenes.camera_rays(big_wl.camera)]
    log(f"serving: {len(ds)} resident scenes ({time.time() - t0:.1f}s)")

    def make_requests(count, seed):
        r = np.random.default_rng(seed)
        sizes = request_sizes(r, count)
        reqs = []
        for n in sizes:
            k = int(r.integers(0, len(ds)))
            pool = pools[k]
            n = int(min(n, pool.shape[0]))
            start = int(r.integers(0, pool.shape[0] - n + 1))
            u = r.random()
            reqs.append((k, pool[start:start + n], True if u < 0.4 else ("shade" if u < 0.6 else False)))
        return reqs

    budget = 8 << 30
    sv = Server(ctx, ds, hbm_budget=budget, slots=args.slots, max_batch_rays=args.max_batch_rays)
    run_load(sv, make_requests(8, 1), rate_hz=0.0)
    nreq = max(24, 2 * args.steps)
    cap = make_requests(nreq, 2)
    _, nrays, wall = run_load(sv, cap, rate_hz=0.0)
    capacity = nrays / wall
    mean_rays = nrays / len(cap)
    loads = {}
    for frac in (0.3, 0.6, 0.9):
        reqs = make_requests(nreq, int(frac * 100))
        rate = frac * capacity / mean_rays
        lats, nr, w = run_load(sv, reqs, rate_hz=rate, seed=int(frac * 10))
        loads[f"{int(frac * 100)}%"] = {"offered_req_per_s": round(rate, 2), "requests": len(reqs),
                                        "p50_ms": round(float(np.percentile(lats, 50)) * 1e3, 2),
                                        "p99_ms": round(float(np.percentile(lats, 99)) * 1e3, 2),
                                        "sustained_mrays_s": round(nr / w / 1e6, 2)}
    sv.close()
    res = {"capacity": round(capacity / 1e6, 3), "unit": "Mrays/s", "scenes": len(ds),
           "patches": sum(int(d.scene.npatch) for d in ds), "hbm_budget_bytes": budget, "slots": args.slots,
           "max_batch_rays": args.max_batch_rays, "request_rays": "log-uniform [1e3, 4e6]",
           "occlusion_fraction": 0.4, "shaded_fraction": 0.2, "loads": loads,
           "note": "latency is submit to results in the caller's host arrays (host copies and PCIe included); "
                   "shaded requests (closest hit, one shadow ray per light, reflection depth 1) count their "
                   "request rays only"}
    log(f"serving: {res}")
    for d in ds[:-1]:
        d.destroy()
    return res


def run_batched(args):
    """64 resident scenes (patch counts log-uniform in [1k, 200k], ray batches
    log-uniform in [4k, 2M], half camera and half random rays): the per-job
    form (one cgs_trace_closest per scene, back to back on one stream) against
    one persistent batched launch, its graph replay and the autotuned shape."""
    import torch
    import cgspline_amd as cgs
    from cgspline_amd import scenes

    t0 = time.time()
    items = scenes.batched_scenes(count=args.scenes)
    npatch = sum(i.scene.npatch for i in items)
    nray = sum(i.rays.shape[0] for i in items)
    log(f"{len(items)} scenes, {npatch} patches, {nray} rays (generated in {time.time() - t0:.1f}s)")
    ctx = cgs.Context(0)
    t0 = time.time()
    jobs = []
    for it in items:
        ds = ctx.scene(it.scene, hash_retries=8)
        n = it.rays.shape[0]
        jobs.append(dict(scene=ds, rays=cgs.DeviceArray.from_host(ctx, it.rays), n=n,
                         hits=cgs.DeviceArray(ctx, 16 * n), objects=cgs.DeviceArray(ctx, 4 * n)))
    log(f"scenes resident in {time.time() - t0:.1f}s")
    stream = torch.cuda.current_stream(torch.device("cuda:0"))
    sh = stream.cuda_stream

    def timed(fn):
        for _ in range(args.warmup):
            fn()
        torch.cuda.synchronize()
        a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        a.record(stream)
        for _ in range(args.steps):
            fn()
        b.record(stream)
        b.synchronize()
        return a.elapsed_time(b) / args.steps

    def per_job():
        for j in jobs:
            j["scene"].trace_closest_dev(j["rays"], j["n"], j["hits"], j["objects"], stream=sh)

    ms_plain = timed(per_job)
    ms_batch = timed(lambda: ctx.trace_batch(jobs, stream=sh))
    graph = ctx.batch(jobs)
    ms_graph = timed(lambda: graph.launch(sh))
    graph.destroy()
    tuned = ctx.batch_autotune(jobs)
    ms_tuned = timed(lambda: ctx.trace_batch(jobs, stream=sh))
    rate = lambda ms: round(nray / ms / 1e3, 3)
    log(f"per-job {ms_plain:.2f} ms, batched {ms_batch:.2f} ms, graph {ms_graph:.2f} ms, autotuned {tuned} "
        f"{ms_tuned:.2f} ms")
    best = min(ms_batch, ms_graph, ms_tuned)
    out = {
        "metric": f"Mrays/s over {len(items)} resident scenes (batched)",
        "value": rate(best),
        "unit": "Mrays/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(best, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded displaced-grid scenes, camera and random rays)",
        "config": {"workload": "batched_scenes", "scenes": len(items), "patches": npatch, "rays": nray},
        "per_job_value": rate(ms_plain),
        "batched_over_per_job": round(ms_plain / best, 3),
        "forms_ms": {"per_job": round(ms_plain, 3), "batched": round(ms_batch, 3), "graph": round(ms_graph, 3),
                     "autotuned": round(ms_tuned, 3)},
        "autotuned_shape": {"chunk": tuned[0], "blocks_per_cu": tuned[1]},
    }
    print(json.dumps(out), flush=True)


def run_serving(args):
    """8 resident scenes (10k-1M patches, one the 1,048,576-patch scene),
    requests of log-uniform 1k-4M rays (half camera, half random rays; 40 %
    occlusion queries) arriving as a Poisson process at 30/60/90 % of the
    server's measured capacity; p50/p99 latency (submit to results in host
    memory) and sustained Mrays/s per load."""
    import cgspline_amd as cgs
    from cgspline_amd import scenes
    from cgspline_amd.serving import Server, request_sizes, run_load

    t0 = time.time()
    rng = np.random.default_rng(20261202 + 5)
    items = scenes.batched_scenes(seed=20261202 + 5, count=args.scenes - 1, patches=(10_000, 1_000_000),
                                  rays=(1 << 20, 1 << 21))
    big = scenes.displaced_grids()
    items.append(scenes.BatchItem(big.scene, scenes.camera_rays(big.camera)))
    ctx = cgs.Context(0)
    ds = [ctx.scene(it.scene, hash_retries=8) for it in items]
    log(f"{len(items)} resident scenes, {sum(i.scene.npatch for i in items)} patches ({time.time() - t0:.1f}s)")

    def make_requests(count, seed):
        r = np.random.default_rng(seed)
        sizes = request_sizes(r, count)
        reqs = []
        for n in sizes:
            s = int(r.integers(0, len(items)))
            pool = items[s].rays
            n = int(min(n, pool.shape[0]))
            start = int(r.integers(0, pool.shape[0] - n + 1))
            reqs.append((s, pool[start:start + n], bool(r.random() < 0.4)))
        return reqs

    budget = 8 << 30
    max_batch = args.max_batch_rays   # rays per launch: several launches in flight so copies and tracing overlap
    sv = Server(ctx, ds, hbm_budget=budget, slots=args.slots, max_batch_rays=max_batch)
    warm = make_requests(8, 1)
    run_load(sv, warm, rate_hz=0.0)
    cap_reqs = make_requests(args.steps * 4, 2)
    lats, nrays, wall = run_load(sv, cap_reqs, rate_hz=0.0)
    capacity = nrays / wall
    mean_rays = nrays / len(cap_reqs)
    log(f"capacity (all requests at once): {capacity / 1e6:.1f} Mrays/s, mean request {mean_rays:.0f} rays")
    loads = {}
    for frac in (0.3, 0.6, 0.9):
        reqs = make_requests(args.steps * 4, int(frac * 100))
        rate = frac * capacity / mean_rays
        lats, nrays, wall = run_load(sv, reqs, rate_hz=rate, seed=int(frac * 10))
        loads[f"{int(frac * 100)}%"] = {"offered_req_per_s": round(rate, 2), "requests": len(reqs),
                                        "p50_ms": round(float(np.percentile(lats, 50)) * 1e3, 2),
                                        "p99_ms": round(float(np.percentile(lats, 99)) * 1e3, 2),
                                        "sustained_mrays_s": round(nrays / wall / 1e6, 2)}
        log(f"load {frac:.0%}: {loads[f'{int(frac * 100)}%']}")
    sv.close()
    out = {
        "metric": "Mrays/s serving capacity; p50/p99 request latency at 30/60/90% load",
        "value": round(capacity / 1e6, 3),
        "unit": "Mrays/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": 1,
        "ms_per_step": None,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded displaced-grid scenes; camera and random rays; Poisson arrivals)",
        "config": {"workload": "serving", "scenes": len(items), "hbm_budget_bytes": budget, "slots": args.slots,
                   "max_batch_rays": max_batch,
                   "request_rays": "log-uniform [1e3, 4e6]", "occlusion_fraction": 0.4},
        "loads": loads,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
