"""Capped cylinders (the 'Cylinders' user geometry, native/geometry/
Cylinders.dev.cu:107-266) on the GPU against the CPU oracle, through the
C-ABI: hit IDs and distances bit for bit, next to spheres and triangles and
under a rotated instance; 1-spp frames within 1e-3; and the argument checks
of bnx_geom_create_cylinders and bnx_group_create_user."""
import numpy as np
import pytest

import barney_amd as bx
from barney_amd import abi, scenes
from tests import helpers

pytestmark = pytest.mark.gpu


def _same_hits(g, o):
    bad = helpers.hits_equal(g, o)
    assert len(bad) == 0, (len(bad), g[bad[:3]], o[bad[:3]])


@pytest.mark.parametrize("rotated", [False, True])
def test_trace_cylinders(dev, orc, rotated):
    print(f"cylinders trace rotated={rotated}", flush=True)
    spec = helpers.sticks_spec(rotated)
    g, o = spec.gpu(dev), spec.oracle(orc)
    rays = scenes.random_rays(21, 200000, radius=4.0, spread=2.0)
    gh = g.trace(rays)
    oh = o.trace(rays)[0]
    _same_hits(gh, oh)
    cyl = np.isin(oh["obj_id"], (3, 4))
    assert cyl.sum() > 5000                 # cylinders are hit, not only spheres and triangles
    if rotated:
        assert (oh["user_inst_id"][cyl] == 7).sum() > 500
    # any hit: the same rays are occluded
    assert np.array_equal(g.trace(rays, any_hit=True)["prim_id"] >= 0, oh["prim_id"] >= 0)


def _warm_plain_graphs(dev, sc):
    """render the scene without its cylinders until its bounce loop of this
    batch shape is captured: the world with cylinders must not replay it"""
    spec = helpers.c1_spec(scenes.config1(width=sc.width, height=sc.height))
    gi = spec.gpu(dev)
    world = bx.World(dev, gi, sc.dir_lights)
    cam = bx.camera_perspective(sc.camera["position"], sc.camera["direction"], sc.camera["up"], sc.camera["fovy"])
    fr = bx.Frame(dev, sc.width, sc.height)
    for _ in range(3):
        fr.reset()
        bx.render(world, cam, abi.renderer(bg=sc.bg_color, ambient=sc.ambient), fr, 1)
    fr.close()


def test_render_with_cylinders_one_spp(dev, orc):
    sc = scenes.config1(width=256, height=256)
    _warm_plain_graphs(dev, sc)
    st = scenes.sticks(800, extent=2.0)
    glossy = abi.pbr(base_color=(0.7, 0.5, 0.3, 1.0), from_attribute=False, metallic=0.6, roughness=0.3)
    st2 = scenes.sticks(200, seed=scenes.SEEDS["c1"] + 1, extent=1.5)
    st2.user_id, st2.material = 5, glossy
    sc.cylinders += [st, st2]
    spec = helpers.c1_spec(sc)
    gi = spec.gpu(dev)
    world = bx.World(dev, gi, sc.dir_lights)
    cam = bx.camera_perspective(sc.camera["position"], sc.camera["direction"], sc.camera["up"], sc.camera["fovy"])
    fr = bx.Frame(dev, sc.width, sc.height)
    rend = abi.renderer(bg=sc.bg_color, ambient=sc.ambient)
    st_g = bx.render(world, cam, rend, fr, 1)
    o = spec.oracle(orc)
    of = orc.Frame(sc.width, sc.height)
    _, st_o = orc.render(o, orc.camera_perspective(sc.camera["position"], sc.camera["direction"], sc.camera["up"],
                                                   sc.camera["fovy"]), rend, of, 1)
    for ch in (abi.BNX_FB_PRIMID, abi.BNX_FB_INSTID, abi.BNX_FB_OBJID):
        gch, och = fr.read(ch), of.read(ch)
        assert np.array_equal(gch, och), (ch, int((gch != och).sum()))
    objs = of.read(abi.BNX_FB_OBJID)
    assert (objs == 4).sum() > 500 and (objs == 5).sum() > 50
    gd, od = fr.read(abi.BNX_FB_DEPTH), of.read(abi.BNX_FB_DEPTH)
    assert np.allclose(gd, od, rtol=1e-5, atol=0)
    gc, oc = fr.read(abi.BNX_FB_COLOR), of.read(abi.BNX_FB_COLOR)
    diff = np.abs(gc - oc)
    print("cylinders 1spp max abs diff", float(diff.max()), flush=True)
    assert diff.max() <= 1e-3
    assert st_g["rays"] == st_o["rays"] and st_g["generations"] == st_o["generations"]
    fr.close()


def test_cylinder_argument_checks(dev):
    st = scenes.sticks(10)
    bad = st.indices.copy()
    bad[3, 1] = len(st.vertices)
    with pytest.raises(bx.BnxError):
        bx.Geom.cylinders(dev, st.vertices, bad, st.radii)
    neg = st.indices.copy()
    neg[0, 0] = -1
    with pytest.raises(bx.BnxError):
        bx.Geom.cylinders(dev, st.vertices, neg, st.radii)
    # one user group holds one user type
    cyl = bx.Geom.cylinders(dev, st.vertices, st.indices, st.radii)
    sph = bx.Geom.spheres(dev, np.zeros((2, 3), np.float32), np.full(2, 0.1, np.float32))
    with pytest.raises(bx.BnxError):
        bx.Group.geoms(dev, bx.USER, [sph, cyl])
    # an empty set builds and is never hit
    empty = bx.Geom.cylinders(dev, st.vertices, np.zeros((0, 2), np.int32), None)
    g = bx.Group.geoms(dev, bx.USER, [empty]).build()
    w = bx.Group.instances(dev, [g], [0]).build()
    h = w.trace(scenes.random_rays(3, 1000))
    assert np.all(h["prim_id"] == -1)


def test_served_cylinder_world(dev, orc):
    """requests on a world with cylinders (bx.SurfaceWorld: spheres, then
    cylinders, then triangles) served next to one without: each equals its
    render alone bit for bit, and the 1-spp one matches the oracle's
    surface_scene (the same flattening)"""
    sc = scenes.config1(width=64, height=64)
    sc.cylinders.append(scenes.sticks(400, extent=1.8))
    gw = bx.SurfaceWorld(dev, sc)
    plain_sc = scenes.config1(width=64, height=64)
    pw = bx.SurfaceWorld(dev, plain_sc)
    cam = bx.camera_perspective(sc.camera["position"], sc.camera["direction"], sc.camera["up"], sc.camera["fovy"])
    rend = abi.renderer(bg=sc.bg_color, ambient=sc.ambient)
    reqs = [(gw.world, 64, 64, 1), (pw.world, 48, 40, 2), (gw.world, 40, 56, 3)]
    srv = bx.Server(dev, max_paths_in_flight=4000)
    tickets = [srv.submit(w, cam, rend, wd, ht, spp) for w, wd, ht, spp in reqs]
    srv.run_until_idle()
    chans = (abi.BNX_FB_COLOR, abi.BNX_FB_DEPTH, abi.BNX_FB_PRIMID, abi.BNX_FB_INSTID, abi.BNX_FB_OBJID)
    for t, (w, wd, ht, spp) in zip(tickets, reqs):
        fr = bx.Frame(dev, wd, ht)
        bx.render(w, cam, rend, fr, spp)
        for ch in chans:
            assert np.array_equal(srv.read(t, ch).view(np.uint8), fr.read(ch).view(np.uint8)), (wd, ht, ch)
        fr.close()
    o = orc.surface_scene(sc)
    of = orc.Frame(64, 64)
    orc.render(o, orc.camera_perspective(sc.camera["position"], sc.camera["direction"], sc.camera["up"],
                                         sc.camera["fovy"]), rend, of, 1)
    for ch in (abi.BNX_FB_PRIMID, abi.BNX_FB_INSTID, abi.BNX_FB_OBJID):
        assert np.array_equal(srv.read(tickets[0], ch), of.read(ch)), ch
    assert (of.read(abi.BNX_FB_OBJID) == 4).sum() > 200
    assert np.abs(srv.read(tickets[0], abi.BNX_FB_COLOR) - of.read(abi.BNX_FB_COLOR)).max() <= 1e-3
    srv.close()
