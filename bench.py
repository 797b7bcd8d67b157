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


def workload_c2(args):
    from barney_amd import scenes
    sc = scenes.config2(width=args.width or 1024, height=args.height or 1024, spp=args.spp or 16)
    desc = {"workload": f"config 2: structured volume, 256^3 fp32 Gaussian-blob field, 256-entry TF (opacity "
                        f"0->0.8), density 50, 16^3 majorant grid, delta tracking + single-scatter shadow rays, "
                        f"{sc.width}x{sc.height}, {sc.spp} spp",
            "width": sc.width, "height": sc.height, "spp": sc.spp, "grid": 256, "majorant_grid": 16,
            "form": "plain single request"}
    return sc, desc


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
    from barney_amd import abi, scenes
    dev = bx.Device(0)
    w = bx.VolumeWorld(dev, sc) if isinstance(sc, scenes.StructuredVolumeScene) else bx.SurfaceWorld(dev, sc)
    cam = bx.camera_perspective(sc.camera["position"], sc.camera["direction"], sc.camera["up"], sc.camera["fovy"])
    rend = abi.renderer(bg=sc.bg_color, ambient=sc.ambient)
    fr = bx.Frame(dev, sc.width, sc.height)
    for _ in range(warmup):
        fr.reset()
        bx.render(w.world, cam, rend, fr, sc.spp)
    dev.sync()
    lat, tot = [], dict(rays=0, nodes=0, tris=0, spheres=0, trace_ms=0.0, trace_launches=0, volume_samples=0,
                        march_ms=0.0, march_launches=0, mc_cells=0, march_rays=0)
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
    from barney_amd import abi, scenes
    if not os.path.exists(oracle.LIB_PATH):
        oracle.build()
    threads = oracle.default_threads()
    vol = isinstance(sc, scenes.StructuredVolumeScene)
    o = oracle.volume_scene(sc) if vol else oracle.surface_scene(sc)
    spp = 1 if vol else sc.spp
    cam = oracle.camera_perspective(sc.camera["position"], sc.camera["direction"], sc.camera["up"],
                                    sc.camera["fovy"])
    rend = abi.renderer(bg=sc.bg_color, ambient=sc.ambient)
    fr = oracle.Frame(sc.width, sc.height)
    rays, frames, vsamp = 0, 0, 0
    t0 = time.perf_counter()
    while frames == 0 or time.perf_counter() - t0 < budget_s:
        fr.reset()
        _, st = oracle.render(o, cam, rend, fr, spp, threads=threads)
        rays += st["rays"]
        vsamp += st["volume_samples"]
        frames += 1
    el = time.perf_counter() - t0
    out = {"value": rays / el / 1e6, "unit": "Mrays/s", "cores": threads, "kind": "port",
           "sample": f"{frames} full {sc.width}x{sc.height} frames at {spp} spp of the same scene "
                     f"({el:.1f} s wall on {threads} threads; shadow rays traced as closest hit as in "
                     f"the reference CPU path)"}
    if vol:
        out["volume_samples_per_s"] = vsamp / el
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="c2", choices=["c1", "c2"])
    ap.add_argument("--spp", type=int, default=0)
    ap.add_argument("--width", type=int, default=0)
    ap.add_argument("--height", type=int, default=0)
    ap.add_argument("--cpu-seconds", type=float, default=8.0)
    ap.add_argument("--no-cpu", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU only (the multi-GPU path is out of scope); "
              f"--gpus {args.gpus} is not supported", file=sys.stderr)
        sys.exit(2)
    sc, desc = (workload_c2 if args.workload == "c2" else workload_c1)(args)
    elapsed, lat, tot = run_gpu(sc, args.steps, args.warmup)
    mrays = tot["rays"] / elapsed / 1e6
    if tot["march_launches"] > 0:
        # volume march kernel: 8 texels (32 B) per sample, 4 B per macrocell
        # majorant, 32 B ray in + 16 B hit out per marched ray
        kernel, kms, launches = "k_volume_march", tot["march_ms"], tot["march_launches"]
        byts = 32.0 * tot["volume_samples"] + 4.0 * tot["mc_cells"] + 48.0 * tot["march_rays"]
    else:
        kernel, kms, launches = "k_trace_queue", tot["trace_ms"], tot["trace_launches"]
        byts = algorithmic_bytes(tot["rays"], tot["nodes"], tot["tris"], tot["spheres"])
    achieved = byts / (kms * 1e-3) / 1e9 if kms > 0 else 0.0
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
                     "kernel": kernel, "bytes_per_launch": round(byts / max(1, launches)),
                     "avg_launch_ms": round(kms / max(1, launches), 5), "launches_per_step": launches // args.steps},
    }
    if tot["volume_samples"]:
        line["volume_samples_per_s"] = round(tot["volume_samples"] / elapsed, 1)
    if not args.no_cpu:
        line["cpu_baseline"] = run_cpu(sc, args.cpu_seconds)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
