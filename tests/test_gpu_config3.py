"""Config 3 (fBm terrain + spheres, 50/50 matte / glossy, HDRI light,
primary + 4 bounce generations) on the GPU against the CPU oracle at a
reduced size: a 160^2 heightfield over [-2, 2]^2 and 20,000 spheres.  Gates: primary IDs
bit-exact, 1-spp pixels within 1e-3 absolute, converged RMSE within 1e-3."""
import numpy as np
import pytest

import barney_amd as bx
from barney_amd import abi, scenes
from tests import helpers

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def small_c3():
    return scenes.config3(n=160, num_spheres=20000, width=192, height=108, spp=1, env_size=(256, 128), extent=2.0)


def _render_both(dev, orc, sc, spp):
    rend = abi.renderer(bg=sc.bg_color, ambient=sc.ambient, max_generations=sc.max_generations)
    world = bx.SurfaceWorld(dev, sc)
    o = orc.surface_scene(sc)
    cam = bx.camera_perspective(sc.camera["position"], sc.camera["direction"], sc.camera["up"], sc.camera["fovy"])
    fr = bx.Frame(dev, sc.width, sc.height)
    st = bx.render(world.world, cam, rend, fr, spp)
    ocam = orc.camera_perspective(sc.camera["position"], sc.camera["direction"], sc.camera["up"],
                                  sc.camera["fovy"])
    of = orc.Frame(sc.width, sc.height)
    _, ost = orc.render(o, ocam, rend, of, spp)
    out = {ch: (fr.read(ch), of.read(ch)) for ch in (abi.BNX_FB_COLOR, abi.BNX_FB_PRIMID, abi.BNX_FB_INSTID,
                                                       abi.BNX_FB_OBJID, abi.BNX_FB_DEPTH)}
    fr.close()
    return out, st, ost


def test_c3_small_one_spp_parity(dev, orc, small_c3):
    print("case: small config 3, 1 spp", flush=True)
    out, st, ost = _render_both(dev, orc, small_c3, 1)
    for ch in (abi.BNX_FB_PRIMID, abi.BNX_FB_INSTID, abi.BNX_FB_OBJID):
        g, o = out[ch]
        assert np.array_equal(g, o), ch
    g, o = out[abi.BNX_FB_OBJID]
    assert len(np.unique(g)) >= 4                  # both materials on both primitive kinds are in view
    gd, od = out[abi.BNX_FB_DEPTH]
    assert np.allclose(gd, od, rtol=1e-5, atol=0)
    gc, oc = out[abi.BNX_FB_COLOR]
    diff = np.abs(gc - oc)
    print("c3 1spp max abs diff", float(diff.max()), "exact pixels", float((diff == 0).mean()), flush=True)
    assert diff.max() <= 1e-3
    assert st["rays"] == ost["rays"] and st["generations"] == ost["generations"] <= 5


def test_c3_small_converged_rmse(dev, orc, small_c3):
    print("case: small config 3, 16 spp", flush=True)
    out, _, _ = _render_both(dev, orc, small_c3, 16)
    gc, oc = out[abi.BNX_FB_COLOR]
    rmse = float(np.sqrt(np.mean((gc[..., :3] - oc[..., :3]) ** 2)))
    print("c3 16spp rmse", rmse, flush=True)
    assert rmse <= 1e-3
