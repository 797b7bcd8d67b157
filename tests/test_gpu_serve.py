"""Request server (continuous batching) on the GPU.

A served request must give the same frame as rendering it alone with
bnx_render into a cleared frame: the per-pixel RNG streams depend only on
(pixel, sample), the fixed-point accumulation does not depend on the order
in which rays of different requests arrive, and the primary-hit channels are
written by sample 0's primary ray only.  So the comparison is bit-exact, for
every channel, while requests of different worlds (matte surfaces, a
structured volume, PBR surfaces under an HDRI map), resolutions and spp share
the queue under either order of the primary top-up, arrive while others are
in flight, wait for request slots and frame memory, and are split into
partial-sample runs of primary rays."""
import ctypes as C

import numpy as np
import pytest

import barney_amd as bx
from barney_amd import abi, scenes

pytestmark = pytest.mark.gpu

CHANNELS = (abi.BNX_FB_COLOR, abi.BNX_FB_DEPTH, abi.BNX_FB_PRIMID, abi.BNX_FB_INSTID, abi.BNX_FB_OBJID,
            abi.BNX_FB_NORMAL)


@pytest.fixture(scope="module")
def worlds(dev):
    c1 = scenes.config1(width=64, height=64)
    c2 = scenes.config2(n=48, width=64, height=64, spp=1, mc_grid=8)
    c3 = scenes.config3(n=64, num_spheres=400, width=64, height=64, spp=1, env_size=(64, 32))
    sw = bx.SurfaceWorld(dev, c1)
    vw = bx.VolumeWorld(dev, c2)
    tw = bx.SurfaceWorld(dev, c3)
    return {"c1": (sw, c1), "c2": (vw, c2), "c3": (tw, c3)}


def _orbit_camera(sc, k):
    rng = np.random.default_rng(1000 + k)
    c = scenes.orbit_camera(sc, rng.uniform(), rng.uniform())
    return bx.camera_perspective(c["position"], c["direction"], c["up"], c["fovy"])


def _plain(dev, world, cam, rend, w, h, spp):
    fr = bx.Frame(dev, w, h)
    bx.render(world, cam, rend, fr, spp)
    out = {ch: fr.read(ch) for ch in CHANNELS}
    fr.close()
    return out


def _mix(worlds):
    reqs = []
    sizes = [(64, 64), (48, 32), (96, 80), (17, 23)]
    spps = [1, 4, 2, 7]
    for k in range(12):
        name = ("c2", "c1", "c3")[k % 3]
        gw, sc = worlds[name]
        cam = _orbit_camera(sc, k) if k % 4 else bx.camera_perspective(
            sc.camera["position"], sc.camera["direction"], sc.camera["up"], sc.camera["fovy"])
        mg = 2 if k == 5 else getattr(sc, "max_generations", 0)
        rend = abi.renderer(bg=sc.bg_color, ambient=sc.ambient, max_generations=mg)
        w, h = sizes[k % len(sizes)]
        reqs.append((gw.world, cam, rend, w, h, spps[k % len(spps)]))
    return reqs


def _check_same(got, ref, tag):
    for ch in CHANNELS:
        assert np.array_equal(got[ch].view(np.uint8), ref[ch].view(np.uint8)), (tag, ch)


def _gens(mode, step):
    """generations per server step: 0 = bnx_server_step (host read-back after
    every generation), n = bnx_server_step_n(n), "mix" = the two alternating
    with 2..5 device-driven generations (the request state hands over)"""
    if mode == "mix":
        return 0 if step % 2 == 0 else 2 + step % 4
    return mode


@pytest.mark.parametrize("cfg", [dict(), dict(max_paths_in_flight=3000, max_active=3),
                                 dict(max_paths_in_flight=3000, deadline_slack=-1.0),
                                 dict(max_paths_in_flight=20000, frame_memory_budget=80 * 96 * 80 + 80 * 64 * 64)])
@pytest.mark.parametrize("mode", [0, 3, 16, "mix"])
def test_served_equals_alone(dev, worlds, cfg, mode):
    print(f"served equals alone cfg={cfg} generations={mode}", flush=True)
    reqs = _mix(worlds)
    refs = [_plain(dev, *r) for r in reqs]
    srv = bx.Server(dev, **cfg)
    tickets = {}
    done = set()

    def finish(ts):
        # read each result as it completes and release its frame, so that
        # frame memory frees up for the requests still waiting
        for t in ts:
            k = tickets[t]
            s = srv.poll(t)
            assert s["state"] == abi.BNX_REQ_DONE and s["paths_started"] == s["paths_total"]
            assert s["submit_ms"] <= s["admit_ms"] <= s["done_ms"]
            got = {ch: srv.read(t, ch) for ch in CHANNELS}
            _check_same(got, refs[k], k)
            srv.release(t)
            done.add(k)

    # requests arrive while others are in flight: a few up front, then one
    # more after every second step
    for k in range(3):
        tickets[srv.submit(*reqs[k])] = k
    nxt, steps = 3, 0
    while srv.pending or nxt < len(reqs):
        finish(srv.step(_gens(mode, steps)))
        steps += 1
        if nxt < len(reqs) and steps % 2 == 0:
            tickets[srv.submit(*reqs[nxt])] = nxt
            nxt += 1
        assert steps < 20000
    st = srv.stats()
    print("steps", steps, "stats", st, flush=True)
    assert st["done"] == st["released"] == len(reqs) and st["active"] == 0 and st["queued"] == 0
    assert done == set(range(len(reqs)))
    if mode == 16:
        # the same rays as one generation per step, in fewer steps
        assert st["generations"] >= 5
    srv.close()


def test_step_n_counts_and_rejects(dev, worlds):
    """bnx_server_step_n: the stats (rays, generations, shadow rays, node
    visits) of device-driven steps equal those of host steps over the same
    requests; generations outside 1..16 are rejected"""
    reqs = _mix(worlds)[:6]
    out = {}
    for mode in (0, 4):
        srv = bx.Server(dev, max_paths_in_flight=4000)
        for r in reqs:
            srv.submit(*r)
        srv.run_until_idle(generations=mode)
        out[mode] = srv.stats()
        n = C.c_int()
        for bad in (0, 17, -3):
            assert srv.L.bnx_server_step_n(srv.h, bad, srv._done, len(srv._done), C.byref(n)) != 0, bad
        srv.close()
    for k in ("rays_traced", "shadow_rays", "volume_samples", "node_visits", "tri_tests", "sphere_tests", "done"):
        assert out[0][k] == out[4][k], k
    assert out[4]["generations"] == out[0]["generations"]


def test_served_oracle_parity(dev, orc, worlds):
    """a served 1-spp config-1 request against the CPU oracle"""
    sc = scenes.config1(width=64, height=64)
    gw, _ = worlds["c1"]
    cam = bx.camera_perspective(sc.camera["position"], sc.camera["direction"], sc.camera["up"], sc.camera["fovy"])
    rend = abi.renderer(bg=sc.bg_color, ambient=sc.ambient)
    srv = bx.Server(dev, max_paths_in_flight=1000)
    t = srv.submit(gw.world, cam, rend, 64, 64, 1)
    srv.run_until_idle()
    o = orc.surface_scene(sc)
    ocam = orc.camera_perspective(sc.camera["position"], sc.camera["direction"], sc.camera["up"], sc.camera["fovy"])
    of = orc.Frame(64, 64)
    orc.render(o, ocam, rend, of, 1)
    assert np.array_equal(srv.read(t, abi.BNX_FB_PRIMID), of.read(abi.BNX_FB_PRIMID))
    assert np.array_equal(srv.read(t, abi.BNX_FB_INSTID), of.read(abi.BNX_FB_INSTID))
    d = np.abs(srv.read(t, abi.BNX_FB_COLOR) - of.read(abi.BNX_FB_COLOR))
    assert d.max() <= 1e-3
    srv.close()


def test_server_rejects_bad_requests(dev, worlds):
    gw, sc = worlds["c1"]
    cam = bx.camera_perspective(sc.camera["position"], sc.camera["direction"], sc.camera["up"], sc.camera["fovy"])
    rend = abi.renderer()
    srv = bx.Server(dev, frame_memory_budget=80 * 32 * 32)
    for w, h, spp in [(0, 8, 1), (8, 8, 0), (8, 8, 5000), (64, 64, 1)]:
        with pytest.raises(bx.BnxError):
            srv.submit(gw.world, cam, rend, w, h, spp)
    t = srv.submit(gw.world, cam, rend, 32, 32, 1)
    with pytest.raises(bx.BnxError):
        srv.release(t)                      # not done yet
    srv.run_until_idle()
    assert srv.poll(t)["state"] == abi.BNX_REQ_DONE
    srv.release(t)
    with pytest.raises(bx.BnxError):
        srv.poll(t)
    assert srv.step() == []                 # idle
    srv.close()
