"""Benchmark: Mrays/s of barney's single-GPU ray-tracing hot path on MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload c5|c1|c2|c3|c4]

c5 (default): the serving stream.  Requests over resident worlds (config 1's
surface scene, config 2's structured volume and config 3's 33.5M-triangle
terrain with 4M spheres and an HDRI light) arrive with seeded Poisson
arrivals, each with a resolution in {256^2, 512^2, 1024^2, 2048^2}, spp in
{1, 4, 16, 64} and a camera on a seeded orbit.  The timed workload is a
fixed sample: the first 200 requests of the seeded stream (--requests), after
8 warm-up requests; --steps does not resize it (a step is one request of the
sample, and the line reports the sample's size as `steps`).  The same
requests are rendered (a) plain: one bnx_render per request, in arrival
order, and (b) served: through the continuous-batching server.  Reported:
served Mrays/s at saturation (all requests queued at once) as `value`,
plain Mrays/s and their ratio, p50/p99 frame latency (arrival -> frame
readable) of both forms under the same Poisson stream, and the plain
single-request latency.  Served and plain frames are identical (checked on
a sample at the end).

c1 / c2 / c3 / c4: one request per step, from a cleared frame, of config 1
(512^2, 1 spp surfaces), config 2 (1024^2, 16 spp structured volume),
config 3 (1920x1080, 64 spp, 33.5M-triangle terrain + 4M spheres, mixed
matte / glossy, HDRI light) or config 4 (1024^2, 32 spp, 20.5M-tet volume
+ isosurface).

Every line carries the dominant kernel's roofline figures and the CPU
oracle's throughput on a bounded sample of the same workload.  The roofline's
algorithmic bytes follow the per-ray model 32 (ray in) + 16 (hit out) +
64 N_node + 48 N_tri + 16 N_sph, with N_node counting binary-BVH node visits,
and the counts taken by the CPU oracle's instrumented traversal of its own
binary binned-SAH BVH on a stated ray subset of the same workload (shadow
rays as occlusion queries, as the GPU traces them); per volume sample
32 B (eight texels) + 4 B per macrocell step.  Beside it: the measured HBM
traffic of the kernel (rocprofv3 FETCH_SIZE / WRITE_SIZE, in profiles/) and
its fraction of the HBM peak.
"""
from __future__ import annotations

import argparse
import copy
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "Mrays/s + p50/p99 frame latency at fixed RMSE, served batched vs plain GPU"
C5_REQUESTS = 200           # the fixed sample of the serving stream
C5_WARMUP = 8
BYTE_MODEL_REQUESTS = 3     # per world: the oracle counts the rays of its first requests in the sample
BYTE_MODEL_RES = 256        # ... rendered at 256^2, 1 spp
HBM_PEAK_GBS = 8000.0
CPU_TET_RES = 64            # config 4 CPU sample: the same scene and camera at 64x64
CPU_BIG_RES = (160, 90)     # config 3 CPU sample: the same scene and camera at 160x90, 1 spp


def ray_bytes(rays, nodes, tris, spheres):
    """per traced ray: 32 B ray in + 16 B hit out; 64 B per binary-BVH node
    visit (two child boxes + refs), 48 B per triangle test (12 B index + 3 x
    12 B vertices), 16 B per sphere test (centre + radius)"""
    return 48.0 * rays + 64.0 * nodes + 48.0 * tris + 16.0 * spheres


def march_bytes(samples, cells):
    """structured volume: 32 B per sample (eight texels) + 4 B per
    macrocell majorant read"""
    return 32.0 * samples + 4.0 * cells


def oracle_ray_model(o, cam, rend, w, h, spp=1):
    """the oracle's traversal counts on the rays of one render (shadow rays
    as occlusion queries): rays, nodes, tris, spheres"""
    import oracle
    fr = oracle.Frame(w, h)
    _, st = oracle.render(o, cam, rend, fr, spp, shadow_any_hit=True)
    return {k: int(st[k]) for k in ("rays", "nodes", "tris", "spheres")}


def per_ray_bytes(counts):
    return ray_bytes(counts["rays"], counts["nodes"], counts["tris"], counts["spheres"]) / max(1, counts["rays"])


# rocprofv3 summaries of each workload's dominant kernel; each holds a line
# "traffic_bytes_per_launch <kernel> <bytes>" measured with FETCH_SIZE /
# WRITE_SIZE counter runs scaled by a plain copy of known size
PROFILES = {"c5": "k_trace_config5_serving.md", "c2": "k_volume_march_config2_1024x1024x16spp.md",
            "c1": "k_trace_static_config1_512x512x1spp.md", "c3": "k_trace_config3_1920x1080x64spp.md",
            "c4": "k_volume_march_config4_1024x1024x32spp.md"}


def profile_traffic(workload, kernel):
    path = os.path.join(ROOT, "profiles", PROFILES.get(workload, ""))
    if not os.path.isfile(path):
        return None
    with open(path) as f:
        for line in f:
            # "traffic_bytes_per_launch <kernel> <bytes>"; kernel names may hold spaces
            key, _, rest = line.strip().partition(" ")
            name, _, value = rest.rpartition(" ")
            if key == "traffic_bytes_per_launch" and name == kernel:
                return float(value)
    return None


def umesh_bytes(nodes, tets, cells, rays):
    # tet point queries: 64 B per element-BVH node, 80 B per tet test (4 x 16 B
    # vertex+scalar, 16 B index); 4 B per macrocell; 48 B per marched ray
    return 64.0 * nodes + 80.0 * tets + 4.0 * cells + 48.0 * rays


def trace_kernel_name(tuning, max_world_prims):
    """the queue-trace kernels that ran: the LDS-stack and walk variant of
    the tuning, for worlds without instance transforms (the bench worlds);
    the pooled kernel too when a world is large enough to use it"""
    sh = tuning["trace_stack_lds"]
    walk = "true" if tuning.get("trace_walk", 0) == 1 else "false"
    pooled = tuning["trace_refill"] < 64 and max_world_prims >= tuning["trace_refill_prims"]
    static = f"k_trace_static<{sh}, {walk}, false>"
    return f"{static}+k_trace_pool<{sh}, false>" if pooled else static


def roofline(tot, workload=None, trace_kernel="k_trace_static<16, false, false>", trace_bytes=None, model=None):
    """dominant kernel (by summed HIP-event time) and its GB/s: algorithmic
    (the byte model) and measured (HBM traffic per launch from profiles/).
    trace_bytes: the trace kernels' algorithmic bytes over the timed region
    from the oracle's counts (model: how they were counted); without it the
    GPU's own counters are priced at the same per-item sizes"""
    if tot.get("march_ms", 0.0) >= tot.get("trace_ms", 0.0) and tot.get("march_launches", 0) > 0:
        # the full variant (MARCH_FULL = 1) runs for tet fields and
        # isosurfaces, the lean one (0) otherwise
        full = tot.get("umesh_nodes", 0) > 0 or workload == "c4"
        kernel = "k_volume_march<1>" if full else "k_volume_march<0>"
        kms, launches = tot["march_ms"], tot["march_launches"]
        if tot.get("umesh_nodes", 0) > 0:
            byts = umesh_bytes(tot["umesh_nodes"], tot["umesh_tets"], tot["mc_cells"], tot["march_rays"])
            model = ("tet point queries: 64 B per element-BVH node + 80 B per tet test (the GPU walk's counts; "
                     "the oracle's depth-first query visits more) + 4 B per macrocell + 48 B per marched ray")
        else:
            byts = march_bytes(tot["volume_samples"], tot["mc_cells"])
            model = ("32 B per volume sample + 4 B per macrocell step; sample and macrocell counts equal the "
                     "oracle's (the march is bit-exact with it)")
    else:
        # queue traversal: one ray per thread, or pooled for the secondary rays
        # of large scenes when the tuning says so; the HIP events bracket
        # whichever ran
        kernel = trace_kernel
        kms, launches = tot["trace_ms"], tot["trace_launches"]
        if trace_bytes is not None:
            byts = trace_bytes
        else:
            byts = ray_bytes(tot["rays"], tot["nodes"], tot["tris"], tot["spheres"])
            model = ("48 B per ray + 64 B per node fetch + 48 B per triangle + 16 B per sphere, priced on the "
                     "GPU's own counts of its 4-wide walk (no oracle counts in this run)")
    achieved = byts / (kms * 1e-3) / 1e9 if kms > 0 else 0.0
    traffic = profile_traffic(workload, kernel)
    avg_ms = kms / max(1, launches)
    out = {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
           "frac": round(achieved / HBM_PEAK_GBS, 5), "traffic": traffic, "kernel": kernel,
           "bytes_per_launch": round(byts / max(1, launches)), "avg_launch_ms": round(avg_ms, 5),
           "launches": launches, "kernel_ms": round(kms, 3), "byte_model": model}
    if traffic:
        dram = traffic / (avg_ms * 1e-3) / 1e9
        out["dram_gbs"] = round(dram, 2)
        out["dram_frac"] = round(dram / HBM_PEAK_GBS, 5)
    return out


def pct(a, q):
    import numpy as np
    return round(float(np.percentile(np.asarray(a, dtype=np.float64), q)), 4) if len(a) else None


def workload_c2(args):
    from barney_amd import scenes
    sc = scenes.config2(width=args.width or 1024, height=args.height or 1024, spp=args.spp or 16)
    desc = {"workload": f"config 2: structured volume, 256^3 fp32 Gaussian-blob field, 256-entry TF (opacity "
                        f"0->0.8), density 50, 16^3 majorant grid, delta tracking + single-scatter shadow rays, "
                        f"{sc.width}x{sc.height}, {sc.spp} spp",
            "width": sc.width, "height": sc.height, "spp": sc.spp, "grid": 256, "majorant_grid": 16,
            "form": "plain single request"}
    return sc, desc


def workload_c4(args):
    from barney_amd import scenes
    n = args.umesh_n
    sc = scenes.config4(n=n, width=args.width or 1024, height=args.height or 1024, spp=args.spp or 32)
    desc = {"workload": f"config 4: unstructured volume, jittered {n}^3 hex lattice split into 5 tets per hex "
                        f"({len(sc.tets):,} tets, {len(sc.vertices):,} vertices), per-vertex sin(4x)cos(4y)sin(4z)"
                        f" + N(0,0.05^2), element-BVH point location, 32^3 majorant grid, delta tracking + "
                        f"isosurface at 0.3, {sc.width}x{sc.height}, {sc.spp} spp",
            "width": sc.width, "height": sc.height, "spp": sc.spp, "tets": int(len(sc.tets)),
            "majorant_grid": 32, "iso": 0.3, "form": "plain single request"}
    return sc, desc


def workload_c3(args):
    from barney_amd import scenes
    n = args.terrain_n
    sc = scenes.config3(n=n, num_spheres=args.spheres, width=args.width or 1920, height=args.height or 1080,
                        spp=args.spp or 64)
    ntri = sum(len(m.indices) for m in sc.meshes)
    nsph = sum(len(x.radii) for x in sc.spheres)
    desc = {"workload": f"config 3: large triangle scene, {n}x{n} fBm heightfield ({ntri:,} triangles) + "
                        f"{nsph:,} spheres, 50/50 matte / glossy (AnariPBR) per primitive, 1024x512 4-lobe HDRI "
                        f"light, {sc.width}x{sc.height}, {sc.spp} spp, primary + 4 bounce generations",
            "width": sc.width, "height": sc.height, "spp": sc.spp, "triangles": int(ntri), "spheres": int(nsph),
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


def autotune(dev, world, cam, rend, width, height, spp):
    """bnx_autotune on probe renders of this request shape, at up to 16M
    paths per probe (untimed set-up, like the BVH build)"""
    import barney_amd as bx
    pspp = min(spp, max(1, (16 << 20) // (width * height)))
    t0 = time.perf_counter()
    chosen, rep = bx.autotune(world, cam, rend, width, height, pspp)
    el = time.perf_counter() - t0
    log(f"autotune: {el:.1f} s, {rep['probes']} probes at {width}x{height}x{pspp}: "
        f"{rep['ms_before']:.2f} -> {rep['ms_after']:.2f} ms; {chosen}")
    return {"tuning": chosen, "autotune": {"seconds": round(el, 2), "probe": f"{width}x{height}, {pspp} spp",
                                           "probes": rep["probes"], "probe_ms_before": round(rep["ms_before"], 3),
                                           "probe_ms_after": round(rep["ms_after"], 3)}}


def parse_tuning(text):
    """--tuning "k=v,k=v": launch-tuning fields to set; "default": the
    device's defaults; "": run the autotuner"""
    if text in ("", None):
        return None
    if text == "default":
        return {}
    out = {}
    for item in text.split(","):
        k, v = item.split("=")
        out[k.strip()] = int(v)
    return out


def set_tuning(dev, tuning):
    """apply a given tuning (no autotuner); None when the autotuner should run"""
    if tuning is None:
        return None
    if tuning:
        dev.set_tuning(**tuning)
    return {"tuning": dev.tuning(), "autotune": "not run (tuning given)"}


def run_gpu(sc, steps, warmup, tuning=None, captured_form=True):
    import numpy as np

    import barney_amd as bx
    from barney_amd import abi, scenes
    dev = bx.Device(0)
    t0 = time.perf_counter()
    if isinstance(sc, scenes.StructuredVolumeScene):
        w = bx.VolumeWorld(dev, sc)
    elif isinstance(sc, scenes.UMeshVolumeScene):
        w = bx.UMeshWorld(dev, sc)
    else:
        w = bx.SurfaceWorld(dev, sc)
    dev.sync()
    log(f"world built in {time.perf_counter() - t0:.2f} s")
    cam = bx.camera_perspective(sc.camera["position"], sc.camera["direction"], sc.camera["up"], sc.camera["fovy"])
    rend = abi.renderer(bg=sc.bg_color, ambient=sc.ambient, max_generations=getattr(sc, "max_generations", 0))
    tune = set_tuning(dev, tuning) or autotune(dev, w.world, cam, rend, sc.width, sc.height, sc.spp)
    fr = bx.Frame(dev, sc.width, sc.height)
    for _ in range(warmup):
        fr.reset()
        bx.render(w.world, cam, rend, fr, sc.spp)
    dev.sync()
    log("warm-up done")
    lat, tot = [], dict(rays=0, nodes=0, tris=0, spheres=0, trace_ms=0.0, trace_launches=0, volume_samples=0,
                        march_ms=0.0, march_launches=0, mc_cells=0, march_rays=0, umesh_nodes=0, umesh_tets=0,
                        captured_batches=0, uncaptured_batches=0)
    t0 = time.perf_counter()
    for _ in range(steps):
        ts = time.perf_counter()
        fr.reset()
        st = bx.render(w.world, cam, rend, fr, sc.spp)
        lat.append((time.perf_counter() - ts) * 1e3)
        if lat[-1] > 1000.0:
            log(f"frame {len(lat)}: {lat[-1]:.0f} ms")
        for k in tot:
            tot[k] += st[k]
    dev.sync()
    elapsed = time.perf_counter() - t0
    # the same frames with every batch in a captured bounce loop (graph_paths
    # raised to the batch size), when the tuned form ran its batches on the
    # host loop (none captured nor device driven): the batching form of the
    # design, timed beside the tuned one
    cap = None
    if captured_form and tot["uncaptured_batches"] == 0 and tot["captured_batches"] == 0:
        t = dict(tune["tuning"])
        t["graph_paths"] = max(t["graph_paths"], t["render_paths"])
        dev.set_tuning(**t)
        fr.reset()
        bx.render(w.world, cam, rend, fr, sc.spp)        # learns the generations, captures the loop
        dev.sync()
        rays, nb = 0, 0
        t1 = time.perf_counter()
        for _ in range(steps):
            fr.reset()
            st = bx.render(w.world, cam, rend, fr, sc.spp)
            rays += st["rays"]
            nb += st["captured_batches"]
        dev.sync()
        el = time.perf_counter() - t1
        cap = {"mrays_s": round(rays / el / 1e6, 3), "ms_per_step": round(el * 1e3 / steps, 4),
               "captured_batches_per_step": nb // steps, "graph_paths": t["graph_paths"]}
        dev.set_tuning(**tune["tuning"])
    fr.close()
    dev.close()
    return elapsed, np.array(lat), tot, tune, cap


def run_cpu(sc, budget_s):
    import oracle
    from barney_amd import abi, scenes
    march = oracle.use_host_native()
    threads = oracle.default_threads()
    vol = isinstance(sc, (scenes.StructuredVolumeScene, scenes.UMeshVolumeScene))
    if isinstance(sc, scenes.UMeshVolumeScene):
        t0 = time.perf_counter()
        o = oracle.umesh_scene(sc)
        log(f"oracle tet scene built in {time.perf_counter() - t0:.1f} s")
        # the same scene and camera at a reduced resolution: a bounded sample
        sc = copy.copy(sc)
        sc.width, sc.height = CPU_TET_RES, CPU_TET_RES
    elif isinstance(sc, scenes.SurfaceScene) and sc.num_prims > 1_000_000:
        t0 = time.perf_counter()
        o = oracle.surface_scene(sc)
        log(f"oracle surface scene built in {time.perf_counter() - t0:.1f} s")
        sc = copy.copy(sc)
        sc.width, sc.height, sc.spp = CPU_BIG_RES[0], CPU_BIG_RES[1], 1
    else:
        o = oracle.volume_scene(sc) if vol else oracle.surface_scene(sc)
    spp = 1 if vol else sc.spp
    cam = oracle.camera_perspective(sc.camera["position"], sc.camera["direction"], sc.camera["up"],
                                    sc.camera["fovy"])
    rend = abi.renderer(bg=sc.bg_color, ambient=sc.ambient, max_generations=getattr(sc, "max_generations", 0))
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
           "sample": f"{frames} {sc.width}x{sc.height} frames at {spp} spp of the same scene and camera "
                     f"({el:.1f} s wall on {threads} threads, built -march={march}; shadow rays traced as "
                     f"closest hit as in the reference CPU path)"}
    if vol:
        out["volume_samples_per_s"] = vsamp / el
    counts = None if vol else oracle_ray_model(o, cam, rend, sc.width, sc.height, spp)
    return out, counts


# ---------------------------------------------------------------- serving
_C5_SCENES = {}


def serve_scenes(args):
    """the resident scenes of the serving stream: config 1's surface scene,
    config 2's structured volume and config 3's large triangle scene (built
    once per process; the CPU sample reuses them)"""
    from barney_amd import scenes
    key = (args.terrain_n, args.spheres)
    if key not in _C5_SCENES:
        _C5_SCENES.clear()
        _C5_SCENES[key] = [scenes.config1(), scenes.config2(),
                           scenes.config3(n=args.terrain_n, num_spheres=args.spheres)]
    return _C5_SCENES[key]


def serve_renderer(sc):
    from barney_amd import abi
    return abi.renderer(bg=sc.bg_color, ambient=sc.ambient, max_generations=getattr(sc, "max_generations", 0))


def serve_setup(dev, args, n, seed=None):
    import barney_amd as bx
    from barney_amd import scenes
    c1, c2, c3 = serve_scenes(args)
    worlds = [(bx.SurfaceWorld(dev, c1), c1), (bx.VolumeWorld(dev, c2), c2), (bx.SurfaceWorld(dev, c3), c3)]
    stream = scenes.serving_stream(n, 1.0, [c1, c2, c3], **({} if seed is None else {"seed": seed}))
    reqs = []
    for r in stream:
        gw, sc = worlds[r.scene]
        cam = bx.camera_perspective(r.camera["position"], r.camera["direction"], r.camera["up"], r.camera["fovy"])
        rend = serve_renderer(sc)
        reqs.append(dict(world=gw.world, cam=cam, rend=rend, w=r.width, h=r.height, spp=r.spp, a1=r.arrival_s,
                         scene=r.scene, camera=r.camera))
    return worlds, reqs


def serve_plain(dev, reqs, frames, per_sample=False):
    """one bnx_render per request; returns service times (s) and totals.
    per_sample: each request renders one pixel wave per sample, each with its
    own generation loop, as the reference's Context::renderTiles does
    (native/Context.cpp:154-178) -- bnx_render with one sample per batch and
    no captured bounce loop; otherwise bnx_render merges samples into
    batches of up to render_paths paths"""
    import barney_amd as bx
    saved = dev.tuning()
    svc, tot, rays = [], {}, []
    try:
        for r in reqs:
            if per_sample:
                dev.set_tuning(render_paths=max(1 << 16, r["w"] * r["h"]), graph_paths=0)
            fr = frames[(r["w"], r["h"])]
            t = time.perf_counter()
            fr.reset()
            st = bx.render(r["world"], r["cam"], r["rend"], fr, r["spp"])
            svc.append(time.perf_counter() - t)
            rays.append(st["rays"])
            for k, v in st.items():
                tot[k] = tot.get(k, 0) + v
    finally:
        dev.set_tuning(**saved)
    return svc, tot, rays


def serve_served(dev, reqs, arrivals=None, cfg=None, gens=0):
    """submit at the arrival times (all at once when arrivals is None) and
    step the server until every request is done; returns per-request
    completion times (s after the start), the makespan and server stats.
    gens: generations per server step (0: bnx_server_step, one host
    read-back per generation; n: bnx_server_step_n)"""
    import barney_amd as bx
    srv = bx.Server(dev, **(cfg or {}))
    done_at = {}
    tix = {}
    t0 = time.perf_counter()
    i, n = 0, len(reqs)
    while i < n or srv.pending:
        now = time.perf_counter() - t0
        while i < n and (arrivals is None or arrivals[i] <= now):
            r = reqs[i]
            tix[srv.submit(r["world"], r["cam"], r["rend"], r["w"], r["h"], r["spp"])] = i
            i += 1
        if srv.pending:
            for t in srv.step(gens):
                done_at[tix[t]] = time.perf_counter() - t0
                srv.release(t)
        elif i < n:
            time.sleep(max(0.0, min(arrivals[i] - now, 0.01)))
    makespan = time.perf_counter() - t0
    st = srv.stats()
    srv.close()
    return [done_at[k] for k in range(n)], makespan, st


def serve_check(dev, reqs, frames, k=6):
    """served frames equal plain frames on the first k small requests"""
    import numpy as np

    import barney_amd as bx
    from barney_amd import abi
    small = [r for r in reqs if r["w"] * r["h"] * r["spp"] <= 1024 * 1024 * 4][:k]
    srv = bx.Server(dev)
    tix = {srv.submit(r["world"], r["cam"], r["rend"], r["w"], r["h"], r["spp"]): j for j, r in enumerate(small)}
    srv.run_until_idle()
    same = 0
    for t, j in tix.items():
        r = small[j]
        fr = frames[(r["w"], r["h"])]
        fr.reset()
        bx.render(r["world"], r["cam"], r["rend"], fr, r["spp"])
        same += int(np.array_equal(fr.read(abi.BNX_FB_COLOR), srv.read(t, abi.BNX_FB_COLOR)) and
                    np.array_equal(fr.read(abi.BNX_FB_PRIMID), srv.read(t, abi.BNX_FB_PRIMID)))
        srv.release(t)
    srv.close()
    return same, len(small)


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def run_c5(args, n, warmup):
    """forms: "all" runs the plain pass, the served pass at saturation (the
    timed region behind `value` and `roofline`), the served pass under
    Poisson arrivals and the parity sample; "served" runs only the served
    pass at saturation, so that a profiler sees exactly the launches behind
    `roofline`"""
    import numpy as np

    import barney_amd as bx
    dev = bx.Device(0)
    # the stream's first C5_WARMUP requests are the warm-up pool, the next n
    # the timed sample, whatever the warm-up count: a run with fewer warm-up
    # requests (a profiler's) times the same sample
    warmup = min(warmup, C5_WARMUP)
    worlds, reqs = serve_setup(dev, args, C5_WARMUP + n)
    log("worlds resident")
    max_prims = max(int(getattr(sc, "num_prims", 0)) for _, sc in worlds)
    # one launch tuning for both forms: bnx_autotune on the stream's heaviest
    # world (config 3) at a 1024^2 request with 16 spp
    tune = set_tuning(dev, parse_tuning(args.tuning))
    if tune is None:
        gw3, c3 = worlds[2]
        cam3 = bx.camera_perspective(c3.camera["position"], c3.camera["direction"], c3.camera["up"],
                                     c3.camera["fovy"])
        tune = autotune(dev, gw3.world, cam3, serve_renderer(c3), 1024, 1024, 16)
    warm, reqs = reqs[C5_WARMUP - warmup:C5_WARMUP], reqs[C5_WARMUP:]
    a1 = np.array([r["a1"] for r in reqs]) - reqs[0]["a1"]
    frames = {}
    for r in warm + reqs:
        if (r["w"], r["h"]) not in frames:
            frames[(r["w"], r["h"])] = bx.Frame(dev, r["w"], r["h"])
    only_served = args.forms == "served"
    # warm-up: both forms
    if warm:
        serve_plain(dev, warm, frames)
        serve_served(dev, warm, gens=args.server_generations)
    dev.sync()
    # plain: service times alone, Mrays/s over the busy time; samples merged
    # into batches (bnx_render), and one pixel wave per sample (the
    # reference's frame loop)
    if not only_served:
        svc, ptot, req_rays = serve_plain(dev, reqs, frames)
        svc = np.array(svc)
        plain_mrays = ptot["rays"] / svc.sum() / 1e6
        log(f"plain (merged samples): {svc.sum():.3f} s busy, {plain_mrays:.1f} Mrays/s")
        svc_ps, ptot_ps, _ = serve_plain(dev, reqs, frames, per_sample=True)
        svc_ps = np.array(svc_ps)
        plain_ps_mrays = ptot_ps["rays"] / svc_ps.sum() / 1e6
        log(f"plain (one pixel wave per sample): {svc_ps.sum():.3f} s busy, {plain_ps_mrays:.1f} Mrays/s")
    # served at saturation
    _, makespan, sst = serve_served(dev, reqs, gens=args.server_generations)
    served_mrays = sst["rays_traced"] / makespan / 1e6
    log(f"served at saturation: {makespan:.3f} s, {served_mrays:.1f} Mrays/s")
    tot = dict(rays=sst["rays_traced"], nodes=sst["node_visits"], tris=sst["tri_tests"],
               spheres=sst["sphere_tests"], trace_ms=sst["trace_ms"], trace_launches=sst["trace_launches"],
               volume_samples=sst["volume_samples"], march_ms=sst["march_ms"],
               march_launches=sst["march_launches"], mc_cells=sst["mc_cells"], march_rays=sst["march_rays"])
    if only_served:
        for f in frames.values():
            f.close()
        dev.close()
        desc = {"workload": f"config 5 serving stream, served pass at saturation only ({n} requests)",
                "requests": n, "warmup_requests": warmup, "form": "served (continuous batching)",
                "server_generations_per_step": args.server_generations, **tune,
                "max_world_prims": max_prims}
        return served_mrays, makespan, tot, desc, {"frame_latency_ms": {"p50": None, "p99": None}}
    # the same stream, Poisson arrivals at a fixed rate: load args.load of
    # the plain form's capacity
    qps = args.load / float(svc.mean())
    arr = a1 / qps
    done, _, _ = serve_served(dev, reqs, arrivals=arr, gens=args.server_generations)
    lat_served = (np.array(done) - arr) * 1e3
    log(f"served at {qps:.1f} QPS: p50 {np.percentile(lat_served, 50):.2f} ms, p99 {np.percentile(lat_served, 99):.2f} ms")
    # plain form under the same arrivals: one request at a time, in order
    fin, lat_plain = 0.0, []
    for a, s in zip(arr, svc):
        fin = max(fin, a) + s
        lat_plain.append((fin - a) * 1e3)
    same, checked = serve_check(dev, reqs, frames)
    for f in frames.values():
        f.close()
    res = sorted({r["w"] for r in reqs})
    c3 = worlds[2][1]
    ntri = sum(len(m.indices) for m in c3.meshes)
    nsph = sum(len(x.radii) for x in c3.spheres)
    per_scene = [sum(1 for r in reqs if r["scene"] == k) for k in range(3)]
    desc = {"workload": f"config 5: serving stream, a fixed sample of {n} requests (requests {C5_WARMUP}.."
                        f"{C5_WARMUP + n - 1} of the seeded 10,000-request stream; the {warmup} before them run "
                        f"untimed as warm-up; a step is one request of the sample) "
                        f"over resident worlds config 1 (20,480 triangles + 1,000 spheres), config 2 "
                        f"(256^3 structured volume) and config 3 ({ntri:,} triangles + {nsph:,} spheres, "
                        f"matte / glossy, HDRI light, 5 generations); resolution in {{256..2048}}^2, spp in "
                        f"{{1,4,16,64}}, seeded orbit cameras; Poisson arrivals at {qps:.2f} QPS (load "
                        f"{args.load} of the plain form's capacity)",
            "requests": n, "warmup_requests": warmup, "qps": round(qps, 3), "resolutions": res,
            "server_generations_per_step": args.server_generations,
            "form": "served (continuous batching) vs plain (one bnx_render per request)",
            "target_judged_against": "plain: bnx_render per request in arrival order, samples merged into "
                                     "batches of up to render_paths paths (the stronger of the two plain forms)",
            "scenes": ["config 1", "config 2", "config 3"], "requests_per_scene": per_scene, **tune,
            "max_world_prims": max_prims}
    extra = {
        "frame_latency_ms": {"p50": pct(lat_served, 50), "p99": pct(lat_served, 99)},
        "served": {"mrays_s_saturated": round(served_mrays, 3), "makespan_s": round(makespan, 4),
                   "latency_ms": {"p50": pct(lat_served, 50), "p99": pct(lat_served, 99)},
                   "generations": sst["generations"], "peak_queue": sst["peak_queue"]},
        "plain": {"mrays_s": round(plain_mrays, 3), "busy_s": round(float(svc.sum()), 4),
                  "latency_alone_ms": {"p50": pct(svc * 1e3, 50), "p99": pct(svc * 1e3, 99)},
                  "latency_queued_ms": {"p50": pct(lat_plain, 50), "p99": pct(lat_plain, 99)}},
        "served_vs_plain_mrays": round(served_mrays / plain_mrays, 3),
        "served_p99_vs_plain_alone_p99": round(pct(lat_served, 99) / pct(svc * 1e3, 99), 3),
        "plain_per_sample": {"mrays_s": round(plain_ps_mrays, 3), "busy_s": round(float(svc_ps.sum()), 4),
                             "latency_alone_ms": {"p50": pct(svc_ps * 1e3, 50), "p99": pct(svc_ps * 1e3, 99)},
                             "form": "one pixel wave per sample with its own generation loop, as "
                                     "Context::renderTiles (native/Context.cpp:154-178)"},
        "served_vs_plain_per_sample_mrays": round(served_mrays / plain_ps_mrays, 3),
        "served_p99_vs_plain_per_sample_alone_p99": round(pct(lat_served, 99) / pct(svc_ps * 1e3, 99), 3),
        "parity_sample": f"{same}/{checked} served frames identical to plain",
        "volume_samples_per_s": round(sst["volume_samples"] / makespan, 1),
        "_rays_per_world": [int(sum(x for x, r in zip(req_rays, reqs) if r["scene"] == k)) for k in range(3)],
        "_sample_reqs": reqs,
    }
    dev.close()
    return served_mrays, makespan, tot, desc, extra


def run_cpu_c5(args, budget_s, sample_reqs=None):
    """the oracle on the smallest requests of the same stream; with the timed
    sample's requests, also the oracle's traversal counts per world (the
    byte model's ray subset: each world's first BYTE_MODEL_REQUESTS requests
    of the sample, rendered at BYTE_MODEL_RES^2, 1 spp)"""
    import oracle
    from barney_amd import scenes
    march = oracle.use_host_native()
    threads = oracle.default_threads()
    c1, c2, c3 = serve_scenes(args)
    built = {}

    def oscene(k):
        # the oracle's scene (and BVH) of world k, built on first use, outside
        # the timed sample
        if k not in built:
            built[k] = (oracle.volume_scene if k == 1 else oracle.surface_scene)((c1, c2, c3)[k])
        return built[k]
    stream = scenes.serving_stream(2000, 1.0, [c1, c2, c3])
    small = [r for r in stream if r.width == 256 and r.spp == 1]
    for k in sorted({r.scene for r in small}):
        oscene(k)
    # the small requests in stream order, from the start again once all are
    # done, until the budget is spent
    rays, frames, t0 = 0, 0, time.perf_counter()
    while time.perf_counter() - t0 <= budget_s:
        r = small[frames % len(small)]
        sc = (c1, c2, c3)[r.scene]
        os_ = oscene(r.scene)
        cam = oracle.camera_perspective(r.camera["position"], r.camera["direction"], r.camera["up"], r.camera["fovy"])
        fr = oracle.Frame(r.width, r.height)
        _, st = oracle.render(os_, cam, serve_renderer(sc), fr, r.spp, threads=threads)
        rays += st["rays"]
        frames += 1
    el = time.perf_counter() - t0
    what = (f"the first {frames}" if frames <= len(small) else
            f"{frames} renders cycling over the {len(small)}")
    out = {"value": round(rays / el / 1e6, 4), "unit": "Mrays/s", "cores": threads, "kind": "port",
           "sample": f"{what} 256^2 1-spp requests among the stream's first 2000 (all three worlds), {el:.1f} s "
                     f"wall on {threads} threads, built -march={march}"}
    counts = None
    if sample_reqs:
        counts = []
        for k in range(3):
            tot = dict(rays=0, nodes=0, tris=0, spheres=0)
            for r in [r for r in sample_reqs if r["scene"] == k][:BYTE_MODEL_REQUESTS]:
                c = r["camera"]
                cam = oracle.camera_perspective(c["position"], c["direction"], c["up"], c["fovy"])
                m = oracle_ray_model(oscene(k), cam, serve_renderer((c1, c2, c3)[k]), BYTE_MODEL_RES,
                                     BYTE_MODEL_RES)
                for key in tot:
                    tot[key] += m[key]
            counts.append(tot)
    return out, counts


def copy_probe(mib):
    """a plain device-to-device copy of `mib` MiB, five times: the reference
    point that scales the FETCH_SIZE / WRITE_SIZE counters to bytes"""
    import torch
    n = mib << 20
    a = torch.ones(n, dtype=torch.uint8, device="cuda")
    b = torch.empty_like(a)
    for _ in range(5):
        b.copy_(a)
    torch.cuda.synchronize()
    print(json.dumps({"copy_probe_bytes": n, "copies": 5}), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=None,
                    help="timed steps (c1-c4: renders); c5 times its fixed sample (--requests) whatever the value")
    ap.add_argument("--requests", type=int, default=C5_REQUESTS, help="c5: size of the fixed sample")
    ap.add_argument("--no-captured-form", action="store_true",
                    help="c1-c4: skip the second timing with every batch in a captured bounce loop")
    ap.add_argument("--server-generations", type=int, default=0,
                    help="c5: generations per server step (0: one host read-back per generation, "
                         "bnx_server_step; 1..16: bnx_server_step_n)")
    ap.add_argument("--warmup", type=int, default=None)
    ap.add_argument("--workload", default="c5", choices=["c1", "c2", "c3", "c4", "c5"])
    ap.add_argument("--terrain-n", type=int, default=4096, help="c3: heightfield size (4096: 33.5M triangles)")
    ap.add_argument("--spheres", type=int, default=4_000_000, help="c3: sphere count")
    ap.add_argument("--umesh-n", type=int, default=160, help="c4: hex lattice size (160: 20.48M tets)")
    ap.add_argument("--spp", type=int, default=0)
    ap.add_argument("--width", type=int, default=0)
    ap.add_argument("--height", type=int, default=0)
    ap.add_argument("--load", type=float, default=0.7)
    ap.add_argument("--forms", default="all", choices=["all", "served"],
                    help="c5 only: run every form, or only the served pass at saturation (for profiling)")
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--tuning", default="",
                    help='launch tuning instead of running bnx_autotune at set-up: "default" (the device\'s '
                         'defaults) or "field=value,..." (bnx_tuning fields)')
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MIB",
                    help="only run a plain device copy of MIB MiB (counter calibration)")
    args = ap.parse_args()
    if args.copy_probe:
        copy_probe(args.copy_probe)
        return
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU only (the multi-GPU path is out of scope); "
              f"--gpus {args.gpus} is not supported", file=sys.stderr)
        sys.exit(2)
    import numpy as np
    if args.workload == "c5":
        # the fixed sample: --steps does not resize it
        steps = args.requests
        warmup = args.warmup if args.warmup is not None else C5_WARMUP
        mrays, elapsed, tot, desc, extra = run_c5(args, steps, warmup)
        lat_line = extra.pop("frame_latency_ms")
        if args.steps is not None:
            desc["steps_requested"] = args.steps
    else:
        heavy = args.workload in ("c3", "c4")
        steps = args.steps if args.steps is not None else (3 if heavy else 20)
        warmup = args.warmup if args.warmup is not None else (1 if heavy else 3)
        sc, desc = {"c1": workload_c1, "c2": workload_c2, "c3": workload_c3,
                    "c4": workload_c4}[args.workload](args)
        elapsed, lat, tot, tune, cap = run_gpu(sc, steps, warmup, parse_tuning(args.tuning),
                                               captured_form=not args.no_captured_form)
        desc.update(tune)
        desc["max_world_prims"] = int(getattr(sc, "num_prims", 0))
        mrays = tot["rays"] / elapsed / 1e6
        lat_line = {"p50": pct(lat, 50), "p99": pct(lat, 99)}
        # batches per frame that ran as a captured bounce loop, or device
        # driven without a graph; neither: the host loop (one read-back per
        # generation)
        extra = {"rays_per_step": tot["rays"] // steps,
                 "batches_per_step": {"captured": tot["captured_batches"] // steps,
                                      "device_driven_uncaptured": tot["uncaptured_batches"] // steps}}
        if cap is not None:
            extra["captured_form"] = cap
        if tot["volume_samples"]:
            extra["volume_samples_per_s"] = round(tot["volume_samples"] / elapsed, 1)
    rays_per_world = extra.pop("_rays_per_world", None)
    sample_reqs = extra.pop("_sample_reqs", None)
    cpu, trace_bytes, model, byte_detail = None, None, None, None
    if not args.no_cpu:
        if args.workload == "c5":
            cpu, counts = run_cpu_c5(args, args.cpu_seconds, sample_reqs)
            if counts and rays_per_world:
                per = [per_ray_bytes(c) for c in counts]
                trace_bytes = sum(r * b for r, b in zip(rays_per_world, per))
                model = (f"oracle binary-BVH counts on each world's first {BYTE_MODEL_REQUESTS} requests of the "
                         f"sample at {BYTE_MODEL_RES}^2, 1 spp (shadow rays as occlusion queries), per-ray bytes "
                         f"times that world's rays in the timed region")
                byte_detail = {"worlds": ["config 1", "config 2", "config 3"], "oracle_counts": counts,
                               "bytes_per_ray": [round(b, 2) for b in per], "rays": rays_per_world,
                               "note": "config 2 rays (volume world, no surfaces) cost 48 B in the trace launch"}
        else:
            cpu, counts = run_cpu(sc, args.cpu_seconds)
            if counts:
                trace_bytes = tot["rays"] * per_ray_bytes(counts)
                model = ("oracle binary-BVH counts on the CPU sample's rays (shadow rays as occlusion queries): "
                         "per-ray bytes times the GPU's rays")
                byte_detail = {"oracle_counts": counts, "bytes_per_ray": round(per_ray_bytes(counts), 2)}
    line = {
        "metric": METRIC, "value": round(mrays, 3), "unit": "Mrays/s", "n_gpus": 1, "steps": steps,
        "warmup": warmup, "ms_per_step": round(elapsed * 1e3 / steps, 4), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic (seeded PCG32 scenes and stream)",
        "config": desc, "frame_latency_ms": lat_line,
        "roofline": roofline(tot, args.workload, trace_kernel_name(desc["tuning"], desc.get("max_world_prims", 0)),
                             trace_bytes, model),
    }
    if byte_detail and line["roofline"]["kernel"].startswith("k_trace"):
        line["roofline"]["byte_model_detail"] = byte_detail
    line.update(extra)
    if cpu is not None:
        line["cpu_baseline"] = cpu
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
