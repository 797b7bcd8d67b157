"""Launch tuning (bnx_device_set_tuning) and the autotuner (bnx_autotune).

No tuning field may change a result: the raygen tile only reorders the
primary rays in the queue (the pixel of a path, its random numbers and the
fixed-point accumulation do not depend on the order), the LDS share of the
traversal stack only moves entries between LDS and private memory, and the
refill threshold, the pooled launch size and the paths per batch only change
which lane or launch traces a ray.  So every channel of every frame must be
bit-identical to the one rendered with the default tuning -- for odd frame
sizes that the tiles do not divide (which would expose a tile order that
misses or repeats a pixel), for several batches per frame, and for a
surface world with PBR and an HDRI light, a matte surface world and a
structured volume."""
import numpy as np
import pytest

import barney_amd as bx
from barney_amd import abi, scenes

pytestmark = pytest.mark.gpu

CHANNELS = (abi.BNX_FB_COLOR, abi.BNX_FB_DEPTH, abi.BNX_FB_PRIMID, abi.BNX_FB_INSTID, abi.BNX_FB_OBJID,
            abi.BNX_FB_NORMAL)


@pytest.fixture
def keep_tuning(dev):
    saved = dev.tuning()
    yield saved
    dev.set_tuning(**saved)


@pytest.fixture(scope="module")
def worlds(dev):
    c1 = scenes.config1(width=64, height=64)
    c2 = scenes.config2(n=48, width=64, height=64, spp=1, mc_grid=8)
    c3 = scenes.config3(n=96, num_spheres=2000, width=64, height=64, spp=1, env_size=(64, 32))
    return {"c1": (bx.SurfaceWorld(dev, c1), c1), "c2": (bx.VolumeWorld(dev, c2), c2),
            "c3": (bx.SurfaceWorld(dev, c3), c3)}


def _render(dev, gw, sc, w, h, spp, stats=None):
    cam = bx.camera_perspective(sc.camera["position"], sc.camera["direction"], sc.camera["up"], sc.camera["fovy"])
    rend = abi.renderer(bg=sc.bg_color, ambient=sc.ambient, max_generations=getattr(sc, "max_generations", 0))
    fr = bx.Frame(dev, w, h)
    st = bx.render(gw.world, cam, rend, fr, spp)
    if stats is not None:
        stats.append(st)
    out = {ch: fr.read(ch) for ch in CHANNELS}
    fr.close()
    return out


def _same(got, ref, tag):
    for ch in CHANNELS:
        assert np.array_equal(got[ch].view(np.uint8), ref[ch].view(np.uint8)), (tag, ch)


TUNINGS = [
    dict(raygen_tile_w=8, raygen_tile_h=8),
    dict(raygen_tile_w=5, raygen_tile_h=3),
    dict(raygen_tile_w=256, raygen_tile_h=1),
    dict(raygen_tile_w=1, raygen_tile_h=256),
    dict(trace_stack_lds=8),
    dict(trace_stack_lds=32),
    dict(trace_walk=1),
    dict(sort_rays=1, graph_paths=0),
    dict(shade_block=256),
    dict(shade_block=256, graph_paths=1 << 26),
    dict(sort_rays=1, graph_paths=0, trace_walk=1, render_paths=1 << 16),
    dict(trace_walk=1, trace_stack_lds=8, raygen_tile_w=0, raygen_tile_h=0),
    dict(trace_refill=4, trace_refill_prims=0, trace_blocks_per_cu=1),
    dict(trace_refill=64, trace_refill_prims=0),
    dict(render_paths=1 << 16, raygen_tile_w=16, raygen_tile_h=4),
    dict(graph_paths=0),
    dict(graph_paths=1 << 26, render_paths=1 << 16),
]


@pytest.mark.parametrize("name,w,h,spp", [("c1", 67, 45, 3), ("c3", 61, 37, 2), ("c2", 43, 29, 2),
                                          ("c1", 128, 96, 9), ("c3", 160, 120, 4)])
def test_results_independent_of_tuning(dev, worlds, keep_tuning, name, w, h, spp):
    gw, sc = worlds[name]
    ref = _render(dev, gw, sc, w, h, spp)
    for t in TUNINGS:
        dev.set_tuning(**keep_tuning)
        dev.set_tuning(**t)
        _same(_render(dev, gw, sc, w, h, spp), ref, (name, w, h, t))


def test_tuning_roundtrip_and_rejects(dev, keep_tuning):
    t0 = dev.tuning()
    assert t0["trace_stack_lds"] in (8, 16, 32) and t0["render_paths"] >= 1 << 16
    dev.set_tuning(raygen_tile_w=8, raygen_tile_h=8, trace_stack_lds=32, render_paths=1 << 22)
    t1 = dev.tuning()
    assert (t1["raygen_tile_w"], t1["raygen_tile_h"], t1["trace_stack_lds"], t1["render_paths"]) == (8, 8, 32, 1 << 22)
    for bad in (dict(raygen_tile_w=8, raygen_tile_h=0), dict(raygen_tile_w=300, raygen_tile_h=1),
                dict(trace_stack_lds=12), dict(trace_blocks_per_cu=0), dict(trace_refill=65), dict(trace_walk=2), dict(sort_rays=3), dict(shade_block=512),
                dict(trace_refill_prims=-1), dict(render_paths=1000), dict(render_paths=1 << 27)):
        with pytest.raises(bx.BnxError):
            dev.set_tuning(**bad)
        assert dev.tuning() == t1, bad


def test_served_under_tuning_equals_alone(dev, worlds, keep_tuning):
    dev.set_tuning(raygen_tile_w=8, raygen_tile_h=8, trace_stack_lds=8, trace_refill=8, trace_refill_prims=0)
    srv = bx.Server(dev, max_paths_in_flight=5000)
    jobs = []
    for k, name in enumerate(("c3", "c1", "c2", "c3")):
        gw, sc = worlds[name]
        cam = bx.camera_perspective(sc.camera["position"], sc.camera["direction"], sc.camera["up"],
                                    sc.camera["fovy"])
        rend = abi.renderer(bg=sc.bg_color, ambient=sc.ambient, max_generations=getattr(sc, "max_generations", 0))
        w, h, spp = (40 + 3 * k, 31 + k, 1 + k)
        jobs.append((srv.submit(gw.world, cam, rend, w, h, spp), gw, sc, w, h, spp))
    srv.run_until_idle()
    dev.set_tuning(**keep_tuning)
    for t, gw, sc, w, h, spp in jobs:
        ref = _render(dev, gw, sc, w, h, spp)
        got = {ch: srv.read(t, ch) for ch in CHANNELS}
        _same(got, ref, ("served", w, h, spp))
        srv.release(t)
    srv.close()


def test_autotune_sets_valid_tuning_and_keeps_results(dev, worlds, keep_tuning):
    gw, sc = worlds["c3"]
    ref = _render(dev, gw, sc, 64, 64, 4)
    cam = bx.camera_perspective(sc.camera["position"], sc.camera["direction"], sc.camera["up"], sc.camera["fovy"])
    rend = abi.renderer(bg=sc.bg_color, ambient=sc.ambient, max_generations=sc.max_generations)
    chosen, rep = bx.autotune(gw.world, cam, rend, 64, 64, probe_spp=2)
    assert chosen == dev.tuning()
    assert rep["probes"] >= 2 * (1 + 4 + 2 + 4) and rep["ms_after"] <= rep["ms_before"] > 0.0
    _same(_render(dev, gw, sc, 64, 64, 4), ref, ("autotuned", chosen))
    with pytest.raises(bx.BnxError):
        bx.autotune(gw.world, cam, rend, 0, 64, probe_spp=1)
    assert dev.tuning() == chosen


def _render_mg(dev, gw, sc, w, h, spp, mg):
    cam = bx.camera_perspective(sc.camera["position"], sc.camera["direction"], sc.camera["up"], sc.camera["fovy"])
    rend = abi.renderer(bg=sc.bg_color, ambient=sc.ambient, max_generations=mg)
    fr = bx.Frame(dev, w, h)
    bx.render(gw.world, cam, rend, fr, spp)
    out = {ch: fr.read(ch) for ch in CHANNELS}
    fr.close()
    return out


@pytest.mark.parametrize("name", ["c1", "c3", "c2"])
def test_captured_bounce_loop_replays(dev, worlds, keep_tuning, name):
    """A batch shape's first render runs the host loop (one read-back per
    generation) and records its generation count; later renders of that
    shape run the generations as a captured graph (captured on the second
    render, replayed after).  Every replay, with eager renders of other
    shapes in between, must equal the host loop's frame bit for bit."""
    gw, sc = worlds[name]
    w, h, spp = 53, 41, 3
    dev.set_tuning(graph_paths=0)
    ref = _render(dev, gw, sc, w, h, spp)
    other = _render(dev, gw, sc, 29, 31, 2)
    dev.set_tuning(graph_paths=1 << 26)
    stats = []
    for k in range(4):
        _same(_render(dev, gw, sc, w, h, spp, stats), ref, (name, "replay", k))
        _same(_render(dev, gw, sc, 29, 31, 2), other, (name, "other", k))
    # the captured loop ran (a failed capture is reported, not silent)
    assert all(st["capture_error"] == 0 for st in stats), stats[-1]["capture_error"]
    assert sum(st["captured_batches"] for st in stats) >= 3 and stats[-1]["uncaptured_batches"] == 0


def test_captured_loop_with_too_few_generations(dev, worlds, keep_tuning):
    """The generation count recorded for a batch shape is only a hint: a
    request whose paths live longer than the captured loop's generations
    finishes them through the host loop, and one whose paths end sooner
    runs empty generations."""
    gw, sc = worlds["c3"]
    w, h, spp = 47, 33, 2
    dev.set_tuning(graph_paths=0)
    ref_full = _render_mg(dev, gw, sc, w, h, spp, 0)
    ref_two = _render_mg(dev, gw, sc, w, h, spp, 2)
    dev.set_tuning(graph_paths=1 << 26)
    _same(_render_mg(dev, gw, sc, w, h, spp, 2), ref_two, "records 2 generations")
    _same(_render_mg(dev, gw, sc, w, h, spp, 0), ref_full, "graph of 2, rest on the host loop")
    _same(_render_mg(dev, gw, sc, w, h, spp, 0), ref_full, "graph of the full count")
    _same(_render_mg(dev, gw, sc, w, h, spp, 2), ref_two, "graph longer than the paths")


def test_captured_loop_follows_tuning(dev, worlds, keep_tuning):
    """A captured bounce loop replays the kernels it was captured with, so a
    change of tuning must not reuse it.  The plain walk's node-visit count
    is a sum of per-ray counts, the same for the host loop, the capture, a
    replay and after switching back; every render with the speculative walk
    steps other nodes (it steps some that the plain walk culls; its count
    also depends on which rays share a wave, and the order in which shade
    appends secondary rays varies from run to run)."""
    gw, sc = worlds["c3"]
    cam = bx.camera_perspective(sc.camera["position"], sc.camera["direction"], sc.camera["up"], sc.camera["fovy"])
    rend = abi.renderer(bg=sc.bg_color, ambient=sc.ambient, max_generations=sc.max_generations)
    fr = bx.Frame(dev, 57, 43)

    def visits():
        fr.reset()
        return bx.render(gw.world, cam, rend, fr, 2)["nodes"]

    dev.set_tuning(graph_paths=1 << 26, trace_walk=0)
    n0 = [visits() for _ in range(3)]
    dev.set_tuning(trace_walk=1)
    n1 = [visits() for _ in range(3)]
    dev.set_tuning(trace_walk=0)
    n0b = visits()
    fr.close()
    assert n0[0] == n0[1] == n0[2] == n0b, (n0, n0b)
    assert all(n != n0[0] for n in n1), (n0, n1)
