"""Tetrahedral volume and isosurface (config 4) on the GPU against the CPU
oracle: macrocell value ranges and majorants bit-exact, primary IDs
bit-exact, depth within 1e-5 relative, 1-spp pixels within 1e-3,
converged RMSE within 1e-3, and served == alone for a tet request."""
import numpy as np
import pytest

import barney_amd as bx
from barney_amd import abi, scenes

pytestmark = pytest.mark.gpu


def small(n=10, w=96, h=96, volume=True, mc=(16, 16, 16)):
    return scenes.config4(n=n, width=w, height=h, spp=1, mc_dims=mc, volume=volume)


@pytest.mark.parametrize("mc", [(16, 16, 16), (7, 9, 5), (0, 0, 0)])
def test_macrocells_match(dev, orc, mc):
    us = small(n=8, mc=mc)
    g = bx.Volume(dev, us)
    gd, glo, ghi = g.mc_ranges()
    o = orc.umesh_scene(us)
    od, olo, ohi = o.mc_ranges(0)
    assert tuple(gd) == tuple(od)
    assert np.array_equal(glo.view(np.uint32), olo.view(np.uint32))
    assert np.array_equal(ghi.view(np.uint32), ohi.view(np.uint32))
    _, gm = g.majorants()
    _, om = o.volume_majorants(0)
    assert np.array_equal(gm.view(np.uint32), om.view(np.uint32))


def _render_both(dev, orc, us, spp):
    rend = abi.renderer(bg=us.bg_color, ambient=us.ambient)
    w = bx.UMeshWorld(dev, us)
    cam = bx.camera_perspective(us.camera["position"], us.camera["direction"], us.camera["up"], us.camera["fovy"])
    fr = bx.Frame(dev, us.width, us.height)
    st = bx.render(w.world, cam, rend, fr, spp)
    o = orc.umesh_scene(us)
    ocam = orc.camera_perspective(us.camera["position"], us.camera["direction"], us.camera["up"], us.camera["fovy"])
    of = orc.Frame(us.width, us.height)
    _, ost = orc.render(o, ocam, rend, of, spp)
    out = {ch: (fr.read(ch), of.read(ch)) for ch in (abi.BNX_FB_COLOR, abi.BNX_FB_DEPTH, abi.BNX_FB_PRIMID,
                                                       abi.BNX_FB_INSTID, abi.BNX_FB_OBJID, abi.BNX_FB_NORMAL)}
    fr.close()
    return out, st, ost


@pytest.mark.parametrize("volume", [True, False])
def test_one_spp_parity(dev, orc, volume):
    us = small(volume=volume)
    out, st, ost = _render_both(dev, orc, us, 1)
    for ch in (abi.BNX_FB_PRIMID, abi.BNX_FB_INSTID, abi.BNX_FB_OBJID):
        g, o = out[ch]
        assert np.array_equal(g, o), (ch, int((g != o).sum()))
    gobj = out[abi.BNX_FB_OBJID][0]
    assert (gobj == us.iso_user_id).sum() > 50          # the isosurface is in view
    if volume:
        assert (gobj == us.user_id).sum() > 50          # and so is the volume
    gd, od = out[abi.BNX_FB_DEPTH]
    assert np.allclose(gd, od, rtol=1e-5, atol=0)
    gn, on = out[abi.BNX_FB_NORMAL]
    assert np.abs(gn - on).max() <= 1e-6
    gc, oc = out[abi.BNX_FB_COLOR]
    d = np.abs(gc - oc)
    print("umesh 1spp max abs diff", float(d.max()), "exact", float((d == 0).mean()), flush=True)
    assert d.max() <= 1e-3
    assert st["rays"] == ost["rays"]


def test_converged_rmse(dev, orc):
    us = small(w=64, h=64)
    out, _, _ = _render_both(dev, orc, us, 16)
    gc, oc = out[abi.BNX_FB_COLOR]
    rmse = float(np.sqrt(np.mean((gc[..., :3] - oc[..., :3]) ** 2)))
    print("umesh 16spp rmse", rmse, flush=True)
    assert rmse <= 1e-3


def test_served_equals_alone(dev):
    us = small(w=48, h=40)
    w = bx.UMeshWorld(dev, us)
    rend = abi.renderer(bg=us.bg_color, ambient=us.ambient)
    cams = []
    for k in range(3):
        c = scenes.orbit_camera(us, 0.1 + 0.3 * k, 0.5)
        cams.append(bx.camera_perspective(c["position"], c["direction"], c["up"], c["fovy"]))
    refs = []
    for cam in cams:
        fr = bx.Frame(dev, 48, 40)
        bx.render(w.world, cam, rend, fr, 2)
        refs.append(fr.read(abi.BNX_FB_COLOR))
        fr.close()
    srv = bx.Server(dev, max_paths_in_flight=2000)
    t = [srv.submit(w.world, cam, rend, 48, 40, 2) for cam in cams]
    srv.run_until_idle()
    for k in range(3):
        assert np.array_equal(srv.read(t[k], abi.BNX_FB_COLOR), refs[k])
    srv.close()


def test_rejects_other_cell_types(dev):
    """polyhedra (_VTK_POLYHEDRON = 42) and unknown codes: BNX_ERR_UNSUPPORTED;
    a hexahedron whose 8 indices run past the index array: BNX_ERR_INVALID"""
    us = small(n=2)
    d, t, keep = abi.umesh_desc(us)
    h = abi.C.c_void_p()
    for code, want in ((42, -4), (7, -4), (12, -1)):
        types = np.zeros(len(us.tets), np.uint8)
        types[-1] = code                                  # the last cell
        d.cell_types = types.ctypes.data_as(abi.C.POINTER(abi.C.c_uint8))
        rc = dev.L.bnx_volume_create_umesh(dev.h, abi.C.byref(d), abi.C.byref(t), 0, abi.C.byref(h))
        assert rc == want, (code, rc)
