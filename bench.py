"""Benchmark: Mrays/s of the ray–patch renderer on the 1M-patch scene.

One step = one render launch (RayTracer::render semantics) of the
displaced-grid workload: 4,096 displaced 16×16-patch grids (1,048,576 bicubic
patches), 2048×2048 primary rays, one shadow ray per lit hit, one reflection
bounce (+ its shadow ray), GGX shading, fp32.  The scene (BVHs and perfect
hashes) is built and resident in HBM before the timed region; `value` counts
every ray the launch traces (primary + shadow + secondary).

Prints one JSON line.  Extra objects: `roofline` (algorithmic bytes of the
render kernel from the oracle's event counts ÷ the kernel time measured with
HIP events on the launch stream) and `cpu_baseline` (the CPU restatement of
the reference's CPURayTracer on a bounded row sample, all host threads).
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


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="displaced_grids")
    ap.add_argument("--cpu-seconds", type=float, default=12.0, help="target CPU-baseline sample duration")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-work-count", action="store_true", help="skip the oracle's work count (no roofline)")
    ap.add_argument("--calibration-copy-mib", type=int, default=0,
                    help="also run one streaming copy kernel (bitwise_not over int32) of this many MiB after the "
                         "warmup: the reference kernel that scales the HBM byte counters (FETCH_SIZE, WRITE_SIZE) "
                         "of a profile")
    args = ap.parse_args()

    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {args.gpus} is not supported "
              "(rays are independent: the multi-GPU form replicates the scene and splits the frame).",
              file=sys.stderr)
        sys.exit(2)

    import torch
    import cgspline_amd as cgs
    from cgspline_amd import scenes, workmodel

    t0 = time.time()
    wl = scenes.WORKLOADS[args.workload]()
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
        st = ds.render_dev(cam, frame, None, counts, eps=wl.eps, depth=wl.depth, stream=sh)
        cgs.check(st, "cgs_render")

    for _ in range(args.warmup):
        step()
        torch.cuda.synchronize()   # render refuses a second launch while one is in flight
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

    out = {
        "metric": "Mrays/s at 1M-patch scene (plain launch)",
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
        "data": "synthetic (seeded displaced bicubic grids)",
        "config": {
            "workload": args.workload,
            "patches": wl.scene.npatch,
            "objects": wl.scene.nobj,
            "width": cam.width,
            "height": cam.height,
            "recursion_depth": wl.depth,
            "rays_per_step": rays_per_step,
            "primary_shadow_secondary_per_step": [int(c[0]) // args.steps, int(c[1]) // args.steps,
                                                  int(c[2]) // args.steps],
            "scene_build_ms": round(build_s * 1e3, 1),
        },
    }

    osc = None
    if not args.no_work_count or not args.no_cpu_baseline:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import cgs_oracle
        osc = cgs_oracle.OracleScene(wl.scene)
    threads = max(1, min(16, os.cpu_count() or 1))

    if not args.no_work_count:
        t0 = time.time()
        _, _, work, rc = osc.render(cam, eps=wl.eps, depth=wl.depth, threads=threads)
        log(f"oracle work count ({time.time() - t0:.1f}s): {work}")
        oracle_rays = rc["primary"] + rc["shadow"] + rc["secondary"]
        if oracle_rays != int(round(rays_per_step)):
            log(f"WARNING: oracle ray count {oracle_rays} != GPU {rays_per_step}")
        nbytes = workmodel.bytes_of(work, cam.width * cam.height, work["objects"])
        nflops = workmodel.flops_of(work)
        achieved = nbytes / avg_kernel_s / 1e9
        out["roofline"] = {
            "bound": "hbm",
            "achieved": round(achieved, 2),
            "peak": workmodel.HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / workmodel.HBM_PEAK_GBS, 4),
            "traffic": measured_traffic(args.workload),
            "kernel": "kRender",
            "algorithmic_bytes_per_launch": int(nbytes),
            "kernel_ms": round(avg_kernel_s * 1e3, 4),
            "valu": {
                "achieved": round(nflops / avg_kernel_s / 1e12, 3),
                "peak": workmodel.FP32_PEAK_TFLOPS,
                "unit": "TFLOP/s",
                "frac": round(nflops / avg_kernel_s / 1e12 / workmodel.FP32_PEAK_TFLOPS, 4),
                "algorithmic_flops_per_launch": int(nflops),
            },
            "work_counts": work,
        }

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
        }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
