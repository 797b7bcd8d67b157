"""Structured volume (config 2) on the GPU against the CPU oracle: macrocell
majorants bit-exact, 1-spp pixels within 1e-3, converged RMSE within 1e-3,
and the same number of volume samples."""
import numpy as np
import pytest

import barney_amd as bx
from barney_amd import abi, scenes

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("mc", [16, 32])
def test_majorants_match(dev, orc, mc):
    vs = scenes.config2(mc_grid=mc)
    g = bx.Volume(dev, vs)
    gd, gm = g.majorants()
    o = orc.volume_scene(vs)
    od, om = o.volume_majorants()
    assert tuple(gd) == tuple(od) == (mc, mc, mc)
    assert np.array_equal(gm.view(np.uint32), om.view(np.uint32))


def _render_both(dev, orc, vs, spp, flags=0):
    rend = abi.renderer(bg=vs.bg_color, ambient=vs.ambient, flags=flags)
    w = bx.VolumeWorld(dev, vs)
    cam = bx.camera_perspective(vs.camera["position"], vs.camera["direction"], vs.camera["up"], vs.camera["fovy"])
    fr = bx.Frame(dev, vs.width, vs.height)
    st = bx.render(w.world, cam, rend, fr, spp)
    o = orc.volume_scene(vs)
    ocam = orc.camera_perspective(vs.camera["position"], vs.camera["direction"], vs.camera["up"], vs.camera["fovy"])
    of = orc.Frame(vs.width, vs.height)
    _, ost = orc.render(o, ocam, rend, of, spp)
    out = {ch: (fr.read(ch), of.read(ch)) for ch in (abi.BNX_FB_COLOR, abi.BNX_FB_DEPTH, abi.BNX_FB_PRIMID,
                                                       abi.BNX_FB_INSTID, abi.BNX_FB_OBJID)}
    fr.close()
    return out, st, ost


def test_volume_one_spp_parity(dev, orc):
    vs = scenes.config2(width=160, height=160, spp=1)
    out, st, ost = _render_both(dev, orc, vs, 1)
    for ch in (abi.BNX_FB_PRIMID, abi.BNX_FB_INSTID, abi.BNX_FB_OBJID):
        assert np.array_equal(*out[ch]), ch
    gd, od = out[abi.BNX_FB_DEPTH]
    assert np.allclose(gd, od, rtol=1e-5, atol=0)
    gc, oc = out[abi.BNX_FB_COLOR]
    d = np.abs(gc - oc)
    print("volume 1spp max abs diff", float(d.max()), "exact", float((d == 0).mean()), flush=True)
    assert d.max() <= 1e-3
    assert st["rays"] == ost["rays"]
    assert st["volume_samples"] == ost["volume_samples"]


def test_volume_converged_rmse(dev, orc):
    vs = scenes.config2(width=64, height=64, spp=16)
    out, _, _ = _render_both(dev, orc, vs, 16)
    gc, oc = out[abi.BNX_FB_COLOR]
    rmse = float(np.sqrt(np.mean((gc[..., :3] - oc[..., :3]) ** 2)))
    print("volume 16spp rmse", rmse, flush=True)
    assert rmse <= 1e-3


def test_transfer_function_update_rebuilds_majorants(dev, orc):
    vs = scenes.config2(n=65, mc_grid=8)
    g = bx.Volume(dev, vs)
    tf2 = scenes.ramp_tf(64, 0.3)
    g.set_transfer_function(tf2, (0.0, 1.0), 20.0)
    vs2 = scenes.StructuredVolumeScene(scalars=vs.scalars, grid_origin=vs.grid_origin, grid_spacing=vs.grid_spacing,
                                       tf=tf2, tf_domain=(0.0, 1.0), base_density=20.0,
                                       cells_per_mc=vs.cells_per_mc)
    _, om = orc.volume_scene(vs2).volume_majorants()
    _, gm = g.majorants()
    assert np.array_equal(gm, om)


def test_ratio_tracking_parity(dev, orc):
    """BNX_RENDER_RATIO_TRACKING (shadow rays carry a ratio-tracking
    transmittance estimate): GPU == oracle channel for channel at 1 and 4
    spp, the frame differs from delta tracking's, and the two agree where
    both have converged (same expectation)"""
    print("ratio tracking parity", flush=True)
    vs = scenes.config2(width=96, height=96, spp=4)
    for spp in (1, 4):
        out, st, ost = _render_both(dev, orc, vs, spp, flags=abi.BNX_RENDER_RATIO_TRACKING)
        for ch in (abi.BNX_FB_PRIMID, abi.BNX_FB_INSTID, abi.BNX_FB_OBJID):
            assert np.array_equal(*out[ch]), ch
        gc, oc = out[abi.BNX_FB_COLOR]
        assert np.abs(gc - oc).max() <= 1e-3, float(np.abs(gc - oc).max())
        assert st["volume_samples"] == ost["volume_samples"]
    delta, _, _ = _render_both(dev, orc, vs, 4)
    assert not np.array_equal(delta[abi.BNX_FB_COLOR][0], out[abi.BNX_FB_COLOR][0])
    vc = scenes.config2(width=32, height=32, spp=256)
    r, _, _ = _render_both(dev, orc, vc, 256, flags=abi.BNX_RENDER_RATIO_TRACKING)
    dl, _, _ = _render_both(dev, orc, vc, 256)
    diff = float(np.abs(r[abi.BNX_FB_COLOR][0][..., :3].mean() - dl[abi.BNX_FB_COLOR][0][..., :3].mean()))
    print("ratio vs delta mean diff at 256 spp", diff, flush=True)
    assert diff < 0.01 * float(dl[abi.BNX_FB_COLOR][0][..., :3].mean()) + 1e-4


@pytest.mark.parametrize("shift", [2, 3])
def test_bricked_layout_parity(dev, orc, shift, monkeypatch):
    """BNX_VOLUME_BRICK: the sampler reads a bricked copy of the field
    (2^shift cells a side plus an apron texel), a layout only; frames equal
    the oracle's as with the linear array, with the same sample count, on a
    field whose size is not a multiple of the brick"""
    print(f"bricked layout shift={shift}", flush=True)
    monkeypatch.setenv("BNX_VOLUME_BRICK", str(shift))
    vs = scenes.config2(n=45, width=96, height=96, spp=2)
    out, st, ost = _render_both(dev, orc, vs, 2)
    for ch in (abi.BNX_FB_PRIMID, abi.BNX_FB_INSTID, abi.BNX_FB_OBJID):
        assert np.array_equal(*out[ch]), ch
    gc, oc = out[abi.BNX_FB_COLOR]
    assert np.abs(gc - oc).max() <= 1e-3, float(np.abs(gc - oc).max())
    assert st["volume_samples"] == ost["volume_samples"] > 0
