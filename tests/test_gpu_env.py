"""HDRI environment light on the GPU against the CPU oracle: the
config-1 scene lit by a seeded 1024x512 map of Gaussian lobes instead of
the constant environment.  Gates: primary IDs bit-exact, 1-spp pixels within
1e-3 absolute, converged RMSE within 1e-3; a served request with the map
equals the same request rendered alone, bit for bit."""
import ctypes as C

import numpy as np
import pytest

import barney_amd as bx
from barney_amd import abi, scenes
from tests import helpers

pytestmark = pytest.mark.gpu


def _env():
    env = scenes.lobe_env_map(1024, 512)
    env.direction, env.up, env.scale = (0.3, 1.0, -0.2), (0.1, 0.2, 1.0), 0.2
    return env


def _render_both(dev, orc, sc, env, spp):
    rend = abi.renderer(bg=sc.bg_color, ambient=sc.ambient)
    spec = helpers.c1_spec(sc)
    world = bx.World(dev, spec.gpu(dev), sc.dir_lights)
    world.set_env_map(env.radiance, env.direction, env.up, env.scale)
    cam = bx.camera_perspective(sc.camera["position"], sc.camera["direction"], sc.camera["up"], sc.camera["fovy"])
    fr = bx.Frame(dev, sc.width, sc.height)
    st = bx.render(world, cam, rend, fr, spp)
    o = spec.oracle(orc)
    o.set_env_map(env.radiance, env.direction, env.up, env.scale)
    ocam = orc.camera_perspective(sc.camera["position"], sc.camera["direction"], sc.camera["up"],
                                  sc.camera["fovy"])
    of = orc.Frame(sc.width, sc.height)
    _, ost = orc.render(o, ocam, rend, of, spp)
    out = {ch: (fr.read(ch), of.read(ch)) for ch in (abi.BNX_FB_COLOR, abi.BNX_FB_PRIMID, abi.BNX_FB_INSTID,
                                                       abi.BNX_FB_OBJID, abi.BNX_FB_DEPTH)}
    fr.close()
    return out, st, ost


def test_env_one_spp_parity(dev, orc):
    sc = scenes.config1(width=160, height=120)
    print("case: env 1spp 160x120", flush=True)
    out, st, ost = _render_both(dev, orc, sc, _env(), 1)
    for ch in (abi.BNX_FB_PRIMID, abi.BNX_FB_INSTID, abi.BNX_FB_OBJID):
        g, o = out[ch]
        assert np.array_equal(g, o), ch
    gd, od = out[abi.BNX_FB_DEPTH]
    assert np.allclose(gd, od, rtol=1e-5, atol=0)
    gc, oc = out[abi.BNX_FB_COLOR]
    diff = np.abs(gc - oc)
    print("env 1spp max abs diff", float(diff.max()), "exact pixels", float((diff == 0).mean()), flush=True)
    assert diff.max() <= 1e-3
    assert st["rays"] == ost["rays"] and st["generations"] == ost["generations"]
    # the map shows where primary rays miss
    miss = out[abi.BNX_FB_PRIMID][0] < 0
    assert miss.any() and np.ptp(gc[miss][:, 0]) > 0.05


def test_env_converged_rmse(dev, orc):
    sc = scenes.config1(width=64, height=64)
    print("case: env 64spp 64x64", flush=True)
    out, _, _ = _render_both(dev, orc, sc, _env(), 64)
    gc, oc = out[abi.BNX_FB_COLOR]
    rmse = float(np.sqrt(np.mean((gc[..., :3] - oc[..., :3]) ** 2)))
    print("env 64spp rmse", rmse, flush=True)
    assert rmse <= 1e-3


def test_env_served_equals_alone(dev):
    sc = scenes.config1(width=64, height=64)
    env = _env()
    plain = bx.SurfaceWorld(dev, sc)
    lit = bx.SurfaceWorld(dev, sc)
    lit.world.set_env_map(env.radiance, env.direction, env.up, env.scale)
    rend = abi.renderer(bg=sc.bg_color, ambient=sc.ambient)
    cam = bx.camera_perspective(sc.camera["position"], sc.camera["direction"], sc.camera["up"], sc.camera["fovy"])
    reqs = [(lit.world, 64, 64, 4), (plain.world, 48, 40, 2), (lit.world, 33, 17, 3)]
    alone = []
    for w, W, H, spp in reqs:
        fr = bx.Frame(dev, W, H)
        bx.render(w, cam, rend, fr, spp)
        alone.append(fr.read(abi.BNX_FB_COLOR))
        fr.close()
    srv = bx.Server(dev)
    tickets = [srv.submit(w, cam, rend, W, H, spp) for w, W, H, spp in reqs]
    srv.run_until_idle()
    for t, (w, W, H, spp), ref in zip(tickets, reqs, alone):
        got = srv.read(t, abi.BNX_FB_COLOR)
        assert np.array_equal(got.view(np.uint32), ref.view(np.uint32)), t
        srv.release(t)
    srv.close()
    # with and without the map differ; removing it restores the plain frame
    fr = bx.Frame(dev, 48, 40)
    bx.render(lit.world, cam, rend, fr, 2)
    assert not np.array_equal(fr.read(abi.BNX_FB_COLOR), alone[1])
    fr.close()
    lit.world.set_env_map(None)
    fr = bx.Frame(dev, 48, 40)
    bx.render(lit.world, cam, rend, fr, 2)
    assert np.array_equal(fr.read(abi.BNX_FB_COLOR), alone[1])
    fr.close()


def test_env_rejects_bad_maps(dev):
    sc = scenes.config1(width=16, height=16)
    w = bx.SurfaceWorld(dev, sc)
    d, keep = abi.env_map_desc(np.ones((4, 4, 3), np.float32))
    d.width = 0
    assert dev.L.bnx_world_set_env_map(w.world.h, C.byref(d)) != 0
    with pytest.raises(ValueError):
        w.world.set_env_map(np.zeros((4, 4), np.float32))
