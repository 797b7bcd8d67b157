"""The renderer's cut plane on the GPU against the oracle: primary rays are
clipped at raygen (generateRays.cu:173-197) and every trace, shadow and
bounce rays included, rejects surface hits on the plane's negative side
(OptixGlobals::hitOnInvisibleSide, native/render/OptixGlobals.h:36-45) for
triangles, spheres and cylinders.  1-spp frames: IDs bit-exact, pixels within
1e-3; the same request served equals its render alone bit for bit."""
import numpy as np
import pytest

import barney_amd as bx
from barney_amd import abi, scenes
from tests import helpers

pytestmark = pytest.mark.gpu

PLANES = [(0.3, -0.5, 0.81, 0.2), (0.0, 0.0, -1.0, 0.25)]


def _scene():
    sc = scenes.config1(width=160, height=160)
    sc.cylinders.append(scenes.sticks(500, extent=2.0))
    return sc


@pytest.mark.parametrize("plane", PLANES)
def test_render_cut_plane_one_spp(dev, orc, plane):
    print(f"cut plane render {plane}", flush=True)
    sc = _scene()
    spec = helpers.c1_spec(sc)
    gi = spec.gpu(dev)
    world = bx.World(dev, gi, sc.dir_lights)
    cam = bx.camera_perspective(sc.camera["position"], sc.camera["direction"], sc.camera["up"], sc.camera["fovy"])
    rend = abi.renderer(bg=sc.bg_color, ambient=sc.ambient, cut_plane=plane)
    fr = bx.Frame(dev, sc.width, sc.height)
    st_g = bx.render(world, cam, rend, fr, 1)
    o = spec.oracle(orc)
    ocam = orc.camera_perspective(sc.camera["position"], sc.camera["direction"], sc.camera["up"], sc.camera["fovy"])
    of = orc.Frame(sc.width, sc.height)
    _, st_o = orc.render(o, ocam, rend, of, 1)
    for ch in (abi.BNX_FB_PRIMID, abi.BNX_FB_INSTID, abi.BNX_FB_OBJID):
        gch, och = fr.read(ch), of.read(ch)
        assert np.array_equal(gch, och), (ch, int((gch != och).sum()))
    gc, oc = fr.read(abi.BNX_FB_COLOR), of.read(abi.BNX_FB_COLOR)
    d = float(np.abs(gc - oc).max())
    print("cut plane 1spp max abs diff", d, flush=True)
    assert d <= 1e-3
    assert st_g["rays"] == st_o["rays"] and st_g["shadow"] == st_o["shadow"]
    # the plane changes the image against an uncut render (it is applied)
    of2 = orc.Frame(sc.width, sc.height)
    orc.render(o, ocam, abi.renderer(bg=sc.bg_color, ambient=sc.ambient), of2, 1)
    assert (of2.read(abi.BNX_FB_PRIMID) != of.read(abi.BNX_FB_PRIMID)).sum() > 500
    fr.close()


def test_served_cut_plane_equals_alone(dev):
    sc = _scene()
    w = bx.SurfaceWorld(dev, sc)
    cam = bx.camera_perspective(sc.camera["position"], sc.camera["direction"], sc.camera["up"], sc.camera["fovy"])
    plain = abi.renderer(bg=sc.bg_color, ambient=sc.ambient)
    cut = abi.renderer(bg=sc.bg_color, ambient=sc.ambient, cut_plane=PLANES[0])
    reqs = [(cut, 64, 48, 2), (plain, 56, 64, 1), (cut, 40, 40, 3)]
    srv = bx.Server(dev, max_paths_in_flight=6000)
    tickets = [srv.submit(w.world, cam, r, wd, ht, spp) for r, wd, ht, spp in reqs]
    srv.run_until_idle()
    for t, (r, wd, ht, spp) in zip(tickets, reqs):
        fr = bx.Frame(dev, wd, ht)
        bx.render(w.world, cam, r, fr, spp)
        for ch in (abi.BNX_FB_COLOR, abi.BNX_FB_DEPTH, abi.BNX_FB_PRIMID, abi.BNX_FB_OBJID):
            assert np.array_equal(srv.read(t, ch).view(np.uint8), fr.read(ch).view(np.uint8)), (wd, ht, ch)
        fr.close()
    srv.close()
