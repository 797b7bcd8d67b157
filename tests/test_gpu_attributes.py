"""Geometry attributes, vertex normals, instance attributes and stochastic
opacity on the GPU against the CPU oracle, through the C-ABI: every frame
channel bit for bit at 1 and 4 spp (the oracle's attribute evaluation and
opacity test are pinned in tests/test_attributes_cpu.py)."""
import numpy as np
import pytest

import barney_amd as bx
from barney_amd import abi, scenes

pytestmark = pytest.mark.gpu

CHANNELS = (abi.BNX_FB_COLOR, abi.BNX_FB_DEPTH, abi.BNX_FB_PRIMID, abi.BNX_FB_INSTID, abi.BNX_FB_OBJID,
            abi.BNX_FB_NORMAL)
W, H = 96, 80


def _quad(z, lo=-2.0, hi=2.0):
    v = np.array([[lo, lo, z], [hi, lo, z], [hi, hi, z], [lo, hi, z]], np.float32)
    return v, np.array([[0, 1, 2], [0, 2, 3]], np.int32)


def _light():
    return [((-0.3, -0.2, -1.0), 2.0, (1.0, 0.9, 0.8))]


def _attr_scene():
    """triangles, spheres and cylinders carrying every attribute scope"""
    rng = np.random.default_rng(11)
    v, f = scenes.icosphere(2)
    v = (v * 0.8 + np.array([0.9, 0.4, 0.0])).astype(np.float32)
    nv, nt = len(v), len(f)
    ico = scenes.TriMesh(v, f, None, user_id=3, normals=v - np.array([0.9, 0.4, 0.0], np.float32),
                         attributes={abi.ATTRIB_SLOT_COLOR: (abi.SCOPE_PER_VERTEX,
                                                             rng.uniform(0.1, 1, (nv, 3)).astype(np.float32))})
    qv, qi = _quad(-0.6)
    fv = scenes.TriMesh(qv, qi, None, user_id=4, material=abi.matte(from_attribute=abi.ATTRIB_0),
                        attributes={abi.ATTRIB_SLOT_0: (abi.SCOPE_FACE_VARYING,
                                                        rng.uniform(0.1, 1, (6, 2)).astype(np.float32)),
                                    abi.ATTRIB_SLOT_NORMAL: (abi.SCOPE_PER_VERTEX,
                                                             np.array([[0, 0, 1], [0.2, 0, 1], [0, 0.2, 1],
                                                                       [-0.1, 0.1, 1]], np.float32))})
    tv, ti = _quad(0.3, -1.6, -0.4)
    pos = scenes.TriMesh(tv, ti, None, user_id=5, material=abi.pbr(from_attribute=abi.ATTRIB_OBJECT_POSITION,
                                                                     metallic=0.3, roughness=0.4))
    sph = scenes.SphereSet(rng.uniform(-1.5, 1.5, (60, 3)).astype(np.float32),
                           rng.uniform(0.05, 0.2, 60).astype(np.float32),
                           rng.uniform(0.2, 0.9, (60, 3)).astype(np.float32), user_id=6,
                           material=abi.matte(from_attribute=abi.ATTRIB_OBJECT_POSITION))
    sph2 = scenes.SphereSet(rng.uniform(-1.5, 1.5, (30, 3)).astype(np.float32),
                            rng.uniform(0.05, 0.2, 30).astype(np.float32), None, user_id=7,
                            attributes={abi.ATTRIB_SLOT_COLOR: (abi.SCOPE_PER_VERTEX,
                                                                rng.uniform(0.2, 0.9, (30, 4)).astype(np.float32))})
    st = scenes.sticks(40, extent=1.5)
    st.colors = None
    st.attributes = {abi.ATTRIB_SLOT_COLOR: (abi.SCOPE_PER_VERTEX,
                                             rng.uniform(0.1, 1, (len(st.vertices), 3)).astype(np.float32))}
    return scenes.SurfaceScene(meshes=[ico, fv, pos], spheres=[sph, sph2], cylinders=[st], dir_lights=_light(),
                               ambient=0.4, width=W, height=H,
                               camera=dict(position=(0.3, -0.5, 6.0), direction=(-0.05, 0.08, -1.0),
                                           up=(0.0, 1.0, 0.0), fovy=40.0))


def _inst_attr_scene():
    """a geometry without a color attribute takes its instance's (one per
    rtc instance: spheres group, then triangles group)"""
    qv, qi = _quad(0.0)
    mesh = scenes.TriMesh(qv, qi, None, user_id=2)
    sph = scenes.SphereSet(np.array([[0.5, 0.5, 0.6], [-0.7, 0.2, 0.4]], np.float32),
                           np.array([0.4, 0.3], np.float32), None, user_id=8)
    sc = scenes.SurfaceScene(meshes=[mesh], spheres=[sph], dir_lights=_light(), ambient=0.5, width=W, height=H,
                             camera=dict(position=(0.0, 0.0, 6.0), direction=(0.0, 0.0, -1.0), up=(0.0, 1.0, 0.0),
                                         fovy=40.0))
    sc.inst_attr = {abi.ATTRIB_SLOT_COLOR: np.array([[0.9, 0.3, 0.2, 1.0], [0.2, 0.5, 0.9, 1.0]], np.float32)}
    return sc


def _translucent_scene(pbr=False):
    """layers of translucent quads (matte opacity, per-vertex color alpha, PBR
    transmission) over an opaque floor and spheres"""
    rng = np.random.default_rng(5)
    meshes = []
    fv, fi = _quad(-1.2, -3.0, 3.0)
    meshes.append(scenes.TriMesh(fv, fi, None, user_id=1,
                                 material=abi.matte((0.6, 0.6, 0.6, 1.0), from_attribute=False)))
    v0, i0 = _quad(0.2, -1.5, 0.5)
    meshes.append(scenes.TriMesh(v0, i0, None, user_id=2,
                                 material=abi.matte((0.9, 0.2, 0.2, 1.0), from_attribute=False, opacity=0.45)))
    v1, i1 = _quad(-0.3, -0.5, 1.6)
    meshes.append(scenes.TriMesh(v1, i1, None, user_id=3, material=abi.matte(from_attribute=True),
                                 attributes={abi.ATTRIB_SLOT_COLOR: (abi.SCOPE_PER_VERTEX, np.array(
                                     [[0.2, 0.9, 0.2, 0.0], [0.2, 0.9, 0.2, 1.0], [0.2, 0.9, 0.2, 0.7],
                                      [0.2, 0.9, 0.2, 0.2]], np.float32))}))
    if pbr:
        v2, i2 = _quad(0.6, -1.0, 1.0)
        meshes.append(scenes.TriMesh(v2, i2, None, user_id=4,
                                     material=abi.pbr((0.3, 0.3, 0.9, 1.0), from_attribute=False, metallic=0.1,
                                                      roughness=0.5, transmission=0.4, ior=1.0, opacity=0.9)))
    sph = scenes.SphereSet(rng.uniform(-1.5, 1.5, (20, 3)).astype(np.float32) * np.array([1, 1, 0.3], np.float32)
                           - np.array([0, 0, 0.8], np.float32),
                           rng.uniform(0.1, 0.3, 20).astype(np.float32),
                           rng.uniform(0.2, 0.9, (20, 3)).astype(np.float32), user_id=9)
    return scenes.SurfaceScene(meshes=meshes, spheres=[sph], dir_lights=_light(), ambient=0.6, width=W, height=H,
                               camera=dict(position=(0.2, 0.1, 6.0), direction=(0.0, 0.0, -1.0),
                                           up=(0.0, 1.0, 0.0), fovy=45.0))


SCENES = {"attributes": _attr_scene, "instance_attributes": _inst_attr_scene,
          "translucent": _translucent_scene, "translucent_pbr": lambda: _translucent_scene(True)}


def _render_both(dev, orc, sc, spp):
    cam_args = (sc.camera["position"], sc.camera["direction"], sc.camera["up"], sc.camera["fovy"])
    rend = abi.renderer(bg=(0.1, 0.1, 0.1, 1.0), ambient=sc.ambient)
    gw = bx.SurfaceWorld(dev, sc)
    fr = bx.Frame(dev, sc.width, sc.height)
    st = bx.render(gw.world, bx.camera_perspective(*cam_args), rend, fr, spp)
    g = {ch: fr.read(ch) for ch in CHANNELS}
    fr.close()
    o = orc.surface_scene(sc)
    of = orc.Frame(sc.width, sc.height)
    orc.render(o, orc.camera_perspective(*cam_args), rend, of, spp)
    return g, {ch: of.read(ch) for ch in CHANNELS}, st


@pytest.mark.parametrize("spp", [1, 4])
@pytest.mark.parametrize("name", list(SCENES))
def test_attribute_frames_equal_oracle(dev, orc, name, spp):
    print(f"attributes {name} {spp} spp", flush=True)
    sc = SCENES[name]()
    g, o, st = _render_both(dev, orc, sc, spp)
    for ch in CHANNELS:
        same = g[ch].view(np.uint8) == o[ch].view(np.uint8)
        assert same.all(), (name, spp, ch, int((~same).sum()))
    objs = o[abi.BNX_FB_OBJID]
    present = set(np.unique(objs[objs >= 0]).tolist())
    want = {m.user_id for m in sc.meshes} | {s.user_id for s in sc.spheres}
    assert want <= present, (want, present)


def test_translucent_ids_report_rejected_candidates(dev, orc):
    """an opacity-0 quad in front of an opaque one: the color is the opaque
    quad's alone, the ID channels the transparent quad's (Triangles.dev.cu:
    45-60), on the GPU as in the oracle"""
    qv, qi = _quad(0.5, -1.0, 1.0)
    bv, bi = _quad(-0.5)
    sc = scenes.SurfaceScene(
        meshes=[scenes.TriMesh(bv, bi, None, user_id=1, material=abi.matte((0.5, 0.7, 0.3, 1.0), False)),
                scenes.TriMesh(qv, qi, None, user_id=2, material=abi.matte((1, 0, 0, 1), False, opacity=0.0))],
        dir_lights=_light(), ambient=0.3, width=W, height=H,
        camera=dict(position=(0.0, 0.0, 6.0), direction=(0.0, 0.0, -1.0), up=(0.0, 1.0, 0.0), fovy=40.0))
    g, o, _ = _render_both(dev, orc, sc, 2)
    for ch in CHANNELS:
        assert np.array_equal(g[ch].view(np.uint8), o[ch].view(np.uint8)), ch
    assert (g[abi.BNX_FB_OBJID] == 2).sum() > 500 and (g[abi.BNX_FB_OBJID] == 1).sum() > 500
    alone = scenes.SurfaceScene(meshes=sc.meshes[:1], dir_lights=sc.dir_lights, ambient=sc.ambient, width=W,
                                height=H, camera=sc.camera)
    ga, _, _ = _render_both(dev, orc, alone, 2)
    assert np.array_equal(ga[abi.BNX_FB_COLOR], g[abi.BNX_FB_COLOR])


def test_api_trace_ignores_opacity(dev):
    """bnx_trace carries no RNG state: every surface is opaque there"""
    qv, qi = _quad(0.0)
    gm = bx.Geom.triangles(dev, qv, qi, None, user_id=1, material=abi.matte((1, 1, 1, 1), False, opacity=0.0))
    grp = bx.Group.geoms(dev, bx.TRIANGLES, [gm]).build()
    inst = bx.Group.instances(dev, [grp], [0]).build()
    rays = np.zeros((256, 8), np.float32)
    rays[:, 0] = np.linspace(-1.5, 1.5, 256)
    rays[:, 2] = 3.0
    rays[:, 6] = -1.0
    rays[:, 7] = 1e30
    hits = inst.trace(rays)
    assert (hits["prim_id"] >= 0).all() and np.allclose(hits["t"], 3.0)


def test_attribute_argument_checks(dev):
    qv, qi = _quad(0.0)
    gm = bx.Geom.triangles(dev, qv, qi, None)
    L = dev.L
    with pytest.raises(RuntimeError):
        gm.set_attribute(9, abi.SCOPE_CONSTANT, value=(1, 1, 1, 1))
    with pytest.raises(RuntimeError):
        gm.set_attribute(abi.ATTRIB_SLOT_COLOR, abi.SCOPE_PER_VERTEX, data=np.ones((3, 3), np.float32))  # 4 needed
    with pytest.raises(RuntimeError):
        gm.set_attribute(abi.ATTRIB_SLOT_COLOR, abi.SCOPE_FACE_VARYING, data=np.ones((5, 3), np.float32))
    with pytest.raises(RuntimeError):
        gm.set_attribute(abi.ATTRIB_SLOT_COLOR, abi.SCOPE_PER_PRIM, data=np.ones((2, 5), np.float32))
    with pytest.raises(RuntimeError):
        gm.set_vertex_normals(np.ones((3, 3), np.float32))
    gm.set_attribute(abi.ATTRIB_SLOT_COLOR, abi.SCOPE_FACE_VARYING, data=np.ones((6, 3), np.float32))
    gm.set_attribute(abi.ATTRIB_SLOT_COLOR, abi.SCOPE_NONE)
    assert L is dev.L
