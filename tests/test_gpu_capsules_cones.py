"""Capsules (the 'Capsules' user geometry, native/geometry/Capsules.dev.cu:
19-306) and cones ('Cones', native/geometry/Cones.dev.cu:21-270) on the GPU
against the CPU oracle, through the C-ABI: hit IDs and distances bit for bit
beside spheres and triangles and under a rotated instance; 1-spp frames of
the plain CPU-reference scene with both added (per-prim colours, a
per-vertex colour attribute, a PBR material) within 1e-3 and ID channels
equal; capsules with stochastic opacity; the argument checks."""
import numpy as np
import pytest

import barney_amd as bx
from barney_amd import abi, scenes
from tests import helpers

pytestmark = pytest.mark.gpu


def _same_hits(g, o):
    bad = helpers.hits_equal(g, o)
    assert len(bad) == 0, (len(bad), g[bad[:3]], o[bad[:3]])


def _spec(rotated):
    verts, idx, col, cs, rs = helpers.soup_data()
    caps = scenes.capsule_chains(400, extent=1.8)
    cones = scenes.cone_set(300, extent=1.8)
    spec = helpers.SceneSpec()
    gs = spec.group(1, [spec.sph(cs, rs, None, user_id=2)])
    gc = spec.group(1, [spec.cap(caps.vertices, caps.indices[:250], caps.colors[:250], user_id=5),
                        spec.cap(caps.vertices, caps.indices[250:], None, user_id=7)])
    gk = spec.group(1, [spec.cone(cones.vertices, cones.radii, cones.indices, cones.colors, user_id=6)])
    gt = spec.group(0, [spec.tri(verts, idx[:1000], col[:1000], user_id=0)])
    for g in (gs, gc, gk, gt):
        spec.inst(g, 0)
    if rotated:
        o2w, w2o = helpers.affine_pair((0, 1, 0, 0, 0, 1, 1, 0, 0), (0.5, -0.25, 1.0))
        spec.inst(gc, 8, o2w, w2o)
        spec.inst(gk, 9, o2w, w2o)
    return spec


@pytest.mark.parametrize("rotated", [False, True])
def test_trace_capsules_cones(dev, orc, rotated):
    print(f"capsules/cones trace rotated={rotated}", flush=True)
    spec = _spec(rotated)
    g, o = spec.gpu(dev), spec.oracle(orc)
    rays = scenes.random_rays(22, 200000, radius=4.0, spread=2.0)
    gh = g.trace(rays)
    oh = o.trace(rays)[0]
    _same_hits(gh, oh)
    assert np.isin(oh["obj_id"], (5, 7)).sum() > 5000
    assert (oh["obj_id"] == 6).sum() > 2000
    if rotated:
        assert (oh["user_inst_id"] == 8).sum() > 500 and (oh["user_inst_id"] == 9).sum() > 200
    assert np.array_equal(g.trace(rays, any_hit=True)["prim_id"] >= 0, oh["prim_id"] >= 0)


def _render_both(dev, orc, sc, spp=1):
    spec = helpers.c1_spec(sc)
    gi = spec.gpu(dev)
    world = bx.World(dev, gi, sc.dir_lights)
    cam = bx.camera_perspective(sc.camera["position"], sc.camera["direction"], sc.camera["up"], sc.camera["fovy"])
    fr = bx.Frame(dev, sc.width, sc.height)
    rend = abi.renderer(bg=sc.bg_color, ambient=sc.ambient)
    st_g = bx.render(world, cam, rend, fr, spp)
    o = spec.oracle(orc)
    of = orc.Frame(sc.width, sc.height)
    _, st_o = orc.render(o, orc.camera_perspective(sc.camera["position"], sc.camera["direction"], sc.camera["up"],
                                                   sc.camera["fovy"]), rend, of, spp)
    return fr, of, st_g, st_o


def test_render_with_capsules_cones_one_spp(dev, orc):
    print("capsules/cones 1spp frame", flush=True)
    sc = scenes.config1(width=256, height=256)
    caps = scenes.capsule_chains(500, extent=2.0)
    # per-vertex colour attribute, interpolated between the ends
    vcol = (0.2 + 0.6 * scenes.uniform_table(7, 46, len(caps.vertices), 3)).astype(np.float32)
    caps.attributes = {abi.ATTRIB_SLOT_COLOR: (abi.SCOPE_PER_VERTEX, vcol)}
    caps.material = abi.matte(from_attribute=True)
    cones = scenes.cone_set(300, extent=2.0)
    cones.material = abi.pbr(base_color=(0.7, 0.5, 0.3, 1.0), from_attribute=False, metallic=0.6, roughness=0.3)
    sc.capsules.append(caps)
    sc.cones.append(cones)
    fr, of, st_g, st_o = _render_both(dev, orc, sc)
    for ch in (abi.BNX_FB_PRIMID, abi.BNX_FB_INSTID, abi.BNX_FB_OBJID):
        gch, och = fr.read(ch), of.read(ch)
        assert np.array_equal(gch, och), (ch, int((gch != och).sum()))
    objs = of.read(abi.BNX_FB_OBJID)
    assert (objs == 5).sum() > 500 and (objs == 6).sum() > 200
    assert np.allclose(fr.read(abi.BNX_FB_DEPTH), of.read(abi.BNX_FB_DEPTH), rtol=1e-5, atol=0)
    diff = np.abs(fr.read(abi.BNX_FB_COLOR) - of.read(abi.BNX_FB_COLOR))
    print("capsules/cones 1spp max abs diff", float(diff.max()), flush=True)
    assert diff.max() <= 1e-3
    assert st_g["rays"] == st_o["rays"] and st_g["generations"] == st_o["generations"]
    fr.close()


def test_translucent_capsules(dev, orc):
    # opacity 0.5: each candidate is kept by its own draw (Capsules.dev.cu:
    # 266-280), the same on both sides; 4 spp converge on both
    print("translucent capsules", flush=True)
    sc = scenes.config1(width=128, height=128)
    caps = scenes.capsule_chains(600, extent=1.5)
    m = abi.matte()
    m.opacity = 0.5
    caps.material = m
    sc.capsules.append(caps)
    fr, of, st_g, st_o = _render_both(dev, orc, sc, spp=4)
    for ch in (abi.BNX_FB_PRIMID, abi.BNX_FB_OBJID):
        assert np.array_equal(fr.read(ch), of.read(ch)), ch
    diff = np.abs(fr.read(abi.BNX_FB_COLOR) - of.read(abi.BNX_FB_COLOR))
    assert diff.max() <= 1e-3, float(diff.max())
    assert st_g["rays"] == st_o["rays"]
    fr.close()


def test_capsule_cone_argument_checks(dev):
    caps = scenes.capsule_chains(10)
    bad = caps.indices.copy()
    bad[3, 1] = len(caps.vertices)
    with pytest.raises(bx.BnxError):
        bx.Geom.capsules(dev, caps.vertices, bad)
    cones = scenes.cone_set(10)
    with pytest.raises(bx.BnxError):
        bx.Geom.cones(dev, cones.vertices, cones.radii, cones.indices + len(cones.vertices))
    cap = bx.Geom.capsules(dev, caps.vertices, caps.indices)
    cone = bx.Geom.cones(dev, cones.vertices, cones.radii)
    with pytest.raises(bx.BnxError):
        bx.Group.geoms(dev, bx.USER, [cap, cone])
    # without indices cone i joins vertices 2i, 2i+1: the same hits as explicit pairs
    g1 = bx.Group.geoms(dev, bx.USER, [cone]).build()
    g2 = bx.Group.geoms(dev, bx.USER, [bx.Geom.cones(dev, cones.vertices, cones.radii, cones.indices)]).build()
    w1 = bx.Group.instances(dev, [g1], [0]).build()
    w2 = bx.Group.instances(dev, [g2], [0]).build()
    rays = scenes.random_rays(5, 5000, spread=2.0)
    h1, h2 = w1.trace(rays), w2.trace(rays)
    assert (h1["prim_id"] >= 0).sum() > 50
    _same_hits(h1, h2)
