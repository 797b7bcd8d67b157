"""Orthographic, omnidirectional and depth-of-field cameras and 8-bit colour
output on the GPU against the oracle, and the reference's own fixture
(unitTests/reference_unitTest_omnicamera.png, tests/golden/) rendered on the
GPU: IDs bit-exact, float colour within 1e-3 per pixel, 8-bit sRGB within one
step (the device's powf against the host's), and the fixture's hit mask,
silhouette bleed and per-sphere mean colour."""
import numpy as np
import pytest

import barney_amd as bx
from barney_amd import abi, scenes
from tests import helpers

pytestmark = pytest.mark.gpu


def _compare(fr, of, spp_ids=True):
    if spp_ids:
        for ch in (abi.BNX_FB_PRIMID, abi.BNX_FB_INSTID, abi.BNX_FB_OBJID):
            g, o = fr.read(ch), of.read(ch)
            assert np.array_equal(g, o), (ch, int((g != o).sum()))
    d = float(np.abs(fr.read(abi.BNX_FB_COLOR) - of.read(abi.BNX_FB_COLOR)).max())
    assert d <= 1e-3, d
    for ch in (abi.BNX_FB_COLOR_RGBA8, abi.BNX_FB_COLOR_SRGB8):
        g, o = fr.read(ch).astype(int), of.read(ch).astype(int)
        assert np.abs(g - o).max() <= 1, ch
        assert (g != o).mean() < 1e-3, (ch, float((g != o).mean()))
    return d


def test_omni_fixture_on_gpu(dev, orc):
    print("omni fixture 2048x1024 1 spp", flush=True)
    spec = helpers.omni_spec()
    world = bx.World(dev, spec.gpu(dev), [])
    pos, d, u = helpers.omni_camera_args()
    cam = bx.camera_omni(pos, d, u)
    W, H = helpers.OMNI_SIZE
    rend = abi.renderer(bg=(0.0, 0.0, 0.0, 1.0), ambient=0.8)
    fr = bx.Frame(dev, W, H)
    bx.render(world, cam, rend, fr, 1)
    of = orc.Frame(W, H)
    orc.render(spec.oracle(orc), orc.camera_omni(pos, d, u), rend, of, 1)
    assert bytes(cam) == bytes(orc.camera_omni(pos, d, u))
    _compare(fr, of)
    helpers.check_against_omni_fixture(fr.read(abi.BNX_FB_COLOR_SRGB8), fr.read(abi.BNX_FB_OBJID))
    fr.close()


def _c1(dev, orc, w=128, h=96):
    sc = scenes.config1(width=w, height=h)
    spec = helpers.c1_spec(sc)
    return sc, bx.World(dev, spec.gpu(dev), sc.dir_lights), spec.oracle(orc)


@pytest.mark.parametrize("spp", [1, 3])
def test_orthographic_camera(dev, orc, spp):
    sc, world, o = _c1(dev, orc)
    args = ((0.3, -0.2, 5.0), (0.0, 0.1, -1.0), (0.0, 1.0, 0.0), 3.5, 128 / 96)
    cam, ocam = bx.camera_orthographic(*args), orc.camera_orthographic(*args)
    assert bytes(cam) == bytes(ocam)
    rend = abi.renderer(bg=sc.bg_color, ambient=sc.ambient)
    fr = bx.Frame(dev, 128, 96)
    st = bx.render(world, cam, rend, fr, spp)
    of = orc.Frame(128, 96)
    _, ost = orc.render(o, ocam, rend, of, spp)
    _compare(fr, of)
    assert st["rays"] == ost["rays"]
    assert (of.read(abi.BNX_FB_PRIMID) >= 0).mean() > 0.2
    fr.close()


def test_depth_of_field(dev, orc):
    sc, world, o = _c1(dev, orc)
    cam = bx.camera_perspective(sc.camera["position"], sc.camera["direction"], sc.camera["up"], sc.camera["fovy"],
                                aperture_radius=0.15, focus_distance=4.0)
    ocam = orc.camera_perspective(sc.camera["position"], sc.camera["direction"], sc.camera["up"], sc.camera["fovy"])
    ocam.aperture_radius, ocam.focus_distance = 0.15, 4.0
    assert bytes(cam) == bytes(ocam)
    rend = abi.renderer(bg=sc.bg_color, ambient=sc.ambient)
    fr = bx.Frame(dev, 128, 96)
    bx.render(world, cam, rend, fr, 4)
    of = orc.Frame(128, 96)
    orc.render(o, ocam, rend, of, 4)
    _compare(fr, of)
    fr.close()


def test_served_cameras_equal_alone(dev):
    sc = scenes.config1(width=64, height=64)
    w = bx.SurfaceWorld(dev, sc)
    rend = abi.renderer(bg=sc.bg_color, ambient=sc.ambient)
    cams = [bx.camera_omni((0.0, 0.0, 3.0), (0.0, 0.0, -1.0), (0.0, 1.0, 0.0)),
            bx.camera_orthographic((0.0, 0.0, 5.0), (0.0, 0.0, -1.0), (0.0, 1.0, 0.0), 3.0, 1.0),
            bx.camera_perspective((0.0, 0.0, 5.0), (0.0, 0.0, -1.0), (0.0, 1.0, 0.0), 40.0, 0.2, 4.5)]
    reqs = [(cams[0], 96, 48, 2), (cams[1], 64, 64, 3), (cams[2], 56, 40, 5)]
    srv = bx.Server(dev, max_paths_in_flight=5000)
    tickets = [srv.submit(w.world, c, rend, wd, ht, spp) for c, wd, ht, spp in reqs]
    srv.run_until_idle()
    for t, (c, wd, ht, spp) in zip(tickets, reqs):
        fr = bx.Frame(dev, wd, ht)
        bx.render(w.world, c, rend, fr, spp)
        for ch in (abi.BNX_FB_COLOR, abi.BNX_FB_COLOR_SRGB8, abi.BNX_FB_DEPTH, abi.BNX_FB_PRIMID):
            assert np.array_equal(srv.read(t, ch).view(np.uint8), fr.read(ch).view(np.uint8)), (wd, ht, ch)
        fr.close()
    srv.close()
