"""Wavefront frame loop on the GPU against the CPU oracle (config 1: plain
CPU-reference scene).  Gates: primary primID/instID/objID bit-exact, depth
within 1e-5 relative, 1-spp pixels within 1e-3 absolute, converged RMSE
within 1e-3."""
import numpy as np
import pytest

import barney_amd as bx
from barney_amd import abi, scenes
from tests import helpers

pytestmark = pytest.mark.gpu


def _render_both(dev, orc, sc, spp, rend=None):
    rend = rend or abi.renderer(bg=sc.bg_color, ambient=sc.ambient)
    spec = helpers.c1_spec(sc)
    gi = spec.gpu(dev)
    world = bx.World(dev, gi, sc.dir_lights)
    cam = bx.camera_perspective(sc.camera["position"], sc.camera["direction"], sc.camera["up"], sc.camera["fovy"])
    fr = bx.Frame(dev, sc.width, sc.height)
    st = bx.render(world, cam, rend, fr, spp)
    o = spec.oracle(orc)
    ocam = orc.camera_perspective(sc.camera["position"], sc.camera["direction"], sc.camera["up"],
                                  sc.camera["fovy"])
    assert bytes(cam) == bytes(ocam)
    of = orc.Frame(sc.width, sc.height)
    _, ost = orc.render(o, ocam, rend, of, spp)
    out = {}
    for ch in (abi.BNX_FB_COLOR, abi.BNX_FB_DEPTH, abi.BNX_FB_PRIMID, abi.BNX_FB_INSTID, abi.BNX_FB_OBJID,
               abi.BNX_FB_NORMAL):
        out[ch] = (fr.read(ch), of.read(ch))
    fr.close()
    return out, st, ost


def test_c1_one_spp_parity(dev, orc):
    sc = scenes.config1()
    out, st, ost = _render_both(dev, orc, sc, 1)
    for ch in (abi.BNX_FB_PRIMID, abi.BNX_FB_INSTID, abi.BNX_FB_OBJID):
        g, o = out[ch]
        assert np.array_equal(g, o), (ch, int((g != o).sum()))
    gd, od = out[abi.BNX_FB_DEPTH]
    assert np.allclose(gd, od, rtol=1e-5, atol=0)
    gc, oc = out[abi.BNX_FB_COLOR]
    diff = np.abs(gc - oc)
    print("1spp max abs diff", float(diff.max()), "exact pixels", float((diff == 0).mean()), flush=True)
    assert diff.max() <= 1e-3
    assert st["rays"] == ost["rays"] and st["generations"] == ost["generations"]


def test_c1_converged_rmse(dev, orc):
    sc = scenes.config1(width=128, height=128)
    out, _, _ = _render_both(dev, orc, sc, 32)
    gc, oc = out[abi.BNX_FB_COLOR]
    rmse = float(np.sqrt(np.mean((gc[..., :3] - oc[..., :3]) ** 2)))
    print("32spp rmse", rmse, flush=True)
    assert rmse <= 1e-3


def test_accumulation_split_across_calls(dev):
    sc = scenes.config1(width=64, height=64)
    spec = helpers.c1_spec(sc)
    gi = spec.gpu(dev)
    world = bx.World(dev, gi, sc.dir_lights)
    cam = bx.camera_perspective(sc.camera["position"], sc.camera["direction"], sc.camera["up"], sc.camera["fovy"])
    rend = abi.renderer(bg=sc.bg_color, ambient=sc.ambient)
    a = bx.Frame(dev, 64, 64)
    bx.render(world, cam, rend, a, 4)
    b = bx.Frame(dev, 64, 64)
    for _ in range(4):
        bx.render(world, cam, rend, b, 1)
    assert a.accum_id == b.accum_id == 4
    assert np.array_equal(a.read(abi.BNX_FB_COLOR), b.read(abi.BNX_FB_COLOR))
    b.reset()
    assert b.accum_id == 0 and (b.read(abi.BNX_FB_PRIMID) == -1).all()
