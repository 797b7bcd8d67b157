"""Benchmark: Mrays/s of barney's single-GPU ray-tracing hot path on MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload c1]

A step renders one frame of the workload from a cleared frame buffer
(raygen -> trace/shade generations -> accumulate), with the scene and its
BVHs resident in HBM.  Prints one JSON line with whole-job Mrays/s, the
frame latency percentiles, the trace kernel's roofline figures and the CPU
oracle's throughput on the same scene (timed on this host).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "Mrays/s + p50/p99 frame latency at fixed RMSE, served batched vs plain GPU"
HBM_PEAK_GBS = 8000.0


def algorithmic_bytes(rays, nodes, tris, spheres):
    # per traced ray: 32 B ray in + 16 B hit out; 64 B per node visit,
    # 48 B per triangle record, 16 B per sphere record
    return 48.0 * rays + 64.0 * nodes + 48.0 * tris + 16.0 * spheres


def workload_c1(args):
    from barney_amd import scenes
    sc = scenes.config1(width=args.width or 512, height=args.height or 512, spp=1)
    desc = {"workload": f"config 1: plain CPU-reference scene, {sc.width}x{sc.height}, 1 spp, icosphere(5) "
                        f"20,480 triangles + 1,000 spheres, 1 dir light + ambient, Lambertian",
            "width": sc.width, "height": sc.height, "spp": 1, "triangles": 20480, "spheres": 1000,
            "form": "plain single request"}
    return sc, desc


def run_gpu(sc, steps, warmup):
    import numpy as np

    import barney_amd as bx
    from barney_amd import abi
    dev = bx.Device(0)
    w = bx.SurfaceWorld(dev, sc)
    cam = bx.camera_perspective(sc.camera["position"], sc.camera["direction"], sc.camera["up"], sc.camera["fovy"])
    rend = abi.renderer(bg=sc.bg_color, ambient=sc.ambient)
    fr = bx.Frame(dev, sc.width, sc.height)
    for _ in range(warmup):
        fr.reset()
        bx.render(w.world, cam, rend, fr, sc.spp)
    dev.sync()
    lat, tot = [], dict(rays=0, nodes=0, tris=0, spheres=0, trace_ms=0.0, trace_launches=0)
    t0 = time.perf_counter()
    for _ in range(steps):
        ts = time.perf_counter()
        fr.reset()
        st = bx.render(w.world, cam, rend, fr, sc.spp)
        lat.append((time.perf_counter() - ts) * 1e3)
        for k in tot:
            tot[k] += st[k]
    dev.sync()
    elapsed = time.perf_counter() - t0
    fr.close()
    dev.close()
    return elapsed, np.array(lat), tot


def run_cpu(sc, budget_s):
    import oracle
    from barney_amd import abi
    if not os.path.exists(oracle.LIB_PATH):
        oracle.build()
    threads = oracle.default_threads()
    o = oracle.surface_scene(sc)
    cam = oracle.camera_perspective(sc.camera["position"], sc.camera["direction"], sc.camera["up"],
                                    sc.camera["fovy"])
    rend = abi.renderer(bg=sc.bg_color, ambient=sc.ambient)
    fr = oracle.Frame(sc.width, sc.height)
    rays, frames = 0, 0
    t0 = time.perf_counter()
    while frames == 0 or time.perf_counter() - t0 < budget_s:
        fr.reset()
        _, st = oracle.render(o, cam, rend, fr, sc.spp, threads=threads)
        rays += st["rays"]
        frames += 1
    el = time.perf_counter() - t0
    return {"value": rays / el / 1e6, "unit": "Mrays/s", "cores": threads, "kind": "port",
            "sample": f"{frames} full {sc.width}x{sc.height} {sc.spp}-spp frames of the same scene "
                      f"({el:.1f} s wall on {threads} threads; shadow rays traced as closest hit as in "
                      f"the reference CPU path)"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="c1", choices=["c1"])
    ap.add_argument("--width", type=int, default=0)
    ap.add_argument("--height", type=int, default=0)
    ap.add_argument("--cpu-seconds", type=float, default=8.0)
    ap.add_argument("--no-cpu", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU only (the multi-GPU path is out of scope); "
              f"--gpus {args.gpus} is not supported", file=sys.stderr)
        sys.exit(2)
    sc, desc = workload_c1(args)
    elapsed, lat, tot = run_gpu(sc, args.steps, args.warmup)
    mrays = tot["rays"] / elapsed / 1e6
    byts = algorithmic_bytes(tot["rays"], tot["nodes"], tot["tris"], tot["spheres"])
    achieved = byts / (tot["trace_ms"] * 1e-3) / 1e9 if tot["trace_ms"] > 0 else 0.0
    line = {
        "metric": METRIC, "value": round(mrays, 3), "unit": "Mrays/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(elapsed * 1e3 / args.steps, 4), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic (seeded PCG32 scene)",
        "config": desc,
        "frame_latency_ms": {"p50": round(float(__import__("numpy").percentile(lat, 50)), 4),
                             "p99": round(float(__import__("numpy").percentile(lat, 99)), 4)},
        "rays_per_step": tot["rays"] // args.steps,
        "roofline": {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 5), "traffic": None,
                     "kernel": "k_trace_queue",
                     "bytes_per_launch": round(byts / max(1, tot["trace_launches"])),
                     "avg_launch_ms": round(tot["trace_ms"] / max(1, tot["trace_launches"]), 5)},
    }
    if not args.no_cpu:
        line["cpu_baseline"] = run_cpu(sc, args.cpu_seconds)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
