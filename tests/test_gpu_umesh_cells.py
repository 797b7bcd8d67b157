"""Unstructured volumes with pyramids, prisms and hexahedra besides tets
(UMeshField::DD::eltScalar, umesh/common/UMeshField.h:372-495, with the
Newton inversions of ElementIntersection.h) on the GPU against the CPU
oracle: macrocell value ranges and majorants bit-exact (tets rasterized
split, other cells by their boxes), then 1-spp frames of the volume and an
isosurface over it -- primary IDs bit-exact, depth within 1e-5 relative,
pixels within 1e-3, and the same rays and volume samples -- on a jittered
mesh (non-planar faces: neighbouring cells overlap or leave gaps, where the
first cell of the depth-first query decides) and on a planar-faced one."""
import numpy as np
import pytest

import barney_amd as bx
from barney_amd import abi, scenes

pytestmark = pytest.mark.gpu

MESHES = {
    "jittered": dict(n=8, jitter=0.15),
    "planar": dict(n=7, jitter=0.0),
    "hex_pyr": dict(n=6, jitter=0.1, kinds=(scenes.CELL_HEX, scenes.CELL_PYR)),
}


def _scene(name, w=96, h=96, volume=True):
    return scenes.mixed_volume(width=w, height=h, volume=volume, **MESHES[name])


@pytest.mark.parametrize("name", sorted(MESHES))
def test_cell_macrocells_match(dev, orc, name):
    print(f"mixed cells macrocells {name}", flush=True)
    us = _scene(name)
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


@pytest.mark.parametrize("name,volume", [("jittered", True), ("jittered", False), ("planar", True),
                                         ("hex_pyr", True)])
def test_cell_one_spp_parity(dev, orc, name, volume):
    print(f"mixed cells 1spp {name} volume={volume}", flush=True)
    us = _scene(name, volume=volume)
    rend = abi.renderer(bg=us.bg_color, ambient=us.ambient)
    w = bx.UMeshWorld(dev, us)
    cam = bx.camera_perspective(us.camera["position"], us.camera["direction"], us.camera["up"], us.camera["fovy"])
    fr = bx.Frame(dev, us.width, us.height)
    st = bx.render(w.world, cam, rend, fr, 1)
    o = orc.umesh_scene(us)
    ocam = orc.camera_perspective(us.camera["position"], us.camera["direction"], us.camera["up"], us.camera["fovy"])
    of = orc.Frame(us.width, us.height)
    _, ost = orc.render(o, ocam, rend, of, 1)
    for ch in (abi.BNX_FB_PRIMID, abi.BNX_FB_INSTID, abi.BNX_FB_OBJID):
        g, r = fr.read(ch), of.read(ch)
        assert np.array_equal(g, r), (ch, int((g != r).sum()))
    obj = of.read(abi.BNX_FB_OBJID)
    assert (obj == us.iso_user_id).sum() > 50
    if volume:
        assert (obj == us.user_id).sum() > 50
    assert np.allclose(fr.read(abi.BNX_FB_DEPTH), of.read(abi.BNX_FB_DEPTH), rtol=1e-5, atol=0)
    d = np.abs(fr.read(abi.BNX_FB_COLOR) - of.read(abi.BNX_FB_COLOR))
    print("mixed cells 1spp max abs diff", float(d.max()), flush=True)
    assert d.max() <= 1e-3
    assert st["rays"] == ost["rays"]
    assert st["volume_samples"] == ost["volume_samples"]
    fr.close()
