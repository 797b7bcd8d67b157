"""Geometry attributes, vertex normals and stochastic opacity in the CPU
oracle, pinned by closed forms and statistics (CPU only).

Closed forms use an orthographic camera looking down -z onto triangles in
the z = 0 plane, one directional light along -z and no ambient light: a
matte pixel is then the shadow ray's a * cos^2 / pi^2 * radiance with
a = 0.85 * color / pi (AnariMatte.h:58-63, Lambertian.h:57-65,
shadeRays.cu:680-700), so each pixel gives the color the hit's attributes
produced, to the precision of the fp16 path throughput the shadow ray
carries (2^-11 relative).  Barycentrics are solved in float64 from the
pixel's ray."""
import math

import numpy as np
import pytest

from barney_amd import abi, scenes
from tests import helpers

W = 48
RAD = 2.0


def _ortho(orc):
    return orc.camera_orthographic((0.0, 0.0, 5.0), (0.0, 0.0, -1.0), (0.0, 1.0, 0.0), 5.0, 1.0)


def _render(orc, sc, spp=1):
    s = orc.surface_scene(sc)
    fr = orc.Frame(W, W)
    orc.render(s, _ortho(orc), abi.renderer(bg=(0.0, 0.0, 0.0, 1.0), ambient=0.0), fr, spp, threads=4)
    return fr


def _bary(orc, tri):
    """float64 barycentrics (u, v) of every pixel's hit on the plane of tri"""
    rays = orc.primary_rays(_ortho(orc), W, W).astype(np.float64)
    o, d = rays[:, 0:3], rays[:, 4:7]
    t = -o[:, 2] / d[:, 2]
    P = o + t[:, None] * d
    v0, v1, v2 = (np.asarray(x, np.float64) for x in tri)
    e1, e2 = v1 - v0, v2 - v0
    M = np.array([[e1[0], e2[0]], [e1[1], e2[1]]])
    uv = np.linalg.solve(M, (P[:, :2] - v0[:2]).T).T
    return uv[:, 0].reshape(W, W), uv[:, 1].reshape(W, W)


def _expected(color):
    """pixel of a matte surface facing the light (cos = 1) of this color"""
    return (0.85 * np.asarray(color) / math.pi) / math.pi ** 2 * RAD


TRI = np.array([[-2.0, -2.0, 0.0], [2.0, -2.0, 0.0], [0.0, 2.0, 0.0]], np.float32)
C3 = np.array([[1.0, 0.0, 0.0], [0.0, 1.0, 0.0], [0.0, 0.0, 1.0]], np.float32)


def _scene(mesh, inst_attr=None):
    sc = scenes.SurfaceScene(meshes=[mesh], dir_lights=[((0.0, 0.0, -1.0), RAD, (1.0, 1.0, 1.0))], ambient=0.0,
                             width=W, height=W)
    if inst_attr:
        sc.inst_attr = inst_attr
    return sc


def _inside(u, v, m=0.03):
    return (u > m) & (v > m) & (1 - u - v > m)


def test_per_vertex_color_interpolates(orc):
    mesh = scenes.TriMesh(TRI, np.array([[0, 1, 2]], np.int32), None, user_id=3,
                          attributes={abi.ATTRIB_SLOT_COLOR: (abi.SCOPE_PER_VERTEX, C3)})
    img = _render(orc, _scene(mesh)).read(abi.BNX_FB_COLOR)
    u, v = _bary(orc, TRI)
    ok = _inside(u, v)
    assert ok.sum() > 300
    col = (1 - u - v)[..., None] * C3[0] + u[..., None] * C3[1] + v[..., None] * C3[2]
    assert np.allclose(img[ok][:, :3], _expected(col[ok]), rtol=3e-3, atol=1e-6)


def test_face_varying_differs_from_per_vertex(orc):
    """face-varying elements are taken per corner (3 prim + k), not per vertex"""
    fv = C3[[2, 0, 1]]
    mesh = scenes.TriMesh(TRI, np.array([[0, 1, 2]], np.int32), None,
                          attributes={abi.ATTRIB_SLOT_COLOR: (abi.SCOPE_FACE_VARYING, fv)})
    img = _render(orc, _scene(mesh)).read(abi.BNX_FB_COLOR)
    u, v = _bary(orc, TRI)
    ok = _inside(u, v)
    col = (1 - u - v)[..., None] * fv[0] + u[..., None] * fv[1] + v[..., None] * fv[2]
    assert np.allclose(img[ok][:, :3], _expected(col[ok]), rtol=3e-3, atol=1e-6)


def test_per_prim_and_constant_scopes(orc):
    quad_v = np.array([[-2, -2, 0], [2, -2, 0], [2, 2, 0], [-2, 2, 0]], np.float32)
    quad_i = np.array([[0, 1, 2], [0, 2, 3]], np.int32)
    pp = np.array([[0.9, 0.1, 0.2, 1.0], [0.1, 0.7, 0.3, 1.0]], np.float32)
    img = _render(orc, _scene(scenes.TriMesh(quad_v, quad_i, None, attributes={
        abi.ATTRIB_SLOT_COLOR: (abi.SCOPE_PER_PRIM, pp)}))).read(abi.BNX_FB_COLOR)
    cy, cx = np.mgrid[0:W, 0:W]
    x, y = (cx + 0.5) / W * 5 - 2.5, (cy + 0.5) / W * 5 - 2.5
    inside = (np.abs(x) < 1.9) & (np.abs(y) < 1.9)
    lower = inside & (y < x - 0.1)              # triangle 0 (0, 1, 2) lies below the diagonal
    upper = inside & (y > x + 0.1)
    assert np.allclose(img[lower][:, :3], _expected(pp[0, :3]), rtol=1e-3)
    assert np.allclose(img[upper][:, :3], _expected(pp[1, :3]), rtol=1e-3)
    k = (0.25, 0.5, 0.75, 1.0)
    img = _render(orc, _scene(scenes.TriMesh(quad_v, quad_i, pp[:, :3], attributes={
        abi.ATTRIB_SLOT_COLOR: (abi.SCOPE_CONSTANT, k)}))).read(abi.BNX_FB_COLOR)
    assert np.allclose(img[inside][:, :3], _expected(k[:3]), rtol=1e-3)


def test_instance_attribute_fallback(orc):
    """a geometry without the attribute reads its instance's value
    (Geometry.h:103-110); one set on the geometry wins"""
    ia = np.array([[0.3, 0.6, 0.9, 1.0]], np.float32)
    mesh = scenes.TriMesh(TRI, np.array([[0, 1, 2]], np.int32), None)
    img = _render(orc, _scene(mesh, {abi.ATTRIB_SLOT_COLOR: ia})).read(abi.BNX_FB_COLOR)
    u, v = _bary(orc, TRI)
    ok = _inside(u, v)
    assert np.allclose(img[ok][:, :3], _expected(ia[0, :3]), rtol=1e-3)
    mesh.attributes = {abi.ATTRIB_SLOT_COLOR: (abi.SCOPE_CONSTANT, (0.5, 0.5, 0.5, 1.0))}
    img = _render(orc, _scene(mesh, {abi.ATTRIB_SLOT_COLOR: ia})).read(abi.BNX_FB_COLOR)
    assert np.allclose(img[ok][:, :3], _expected((0.5, 0.5, 0.5)), rtol=1e-3)


def test_attribute0_one_component_and_positions(orc):
    """a material reading attribute0 of one float per vertex sees (x, 0, 0, 1)
    (AttributeArray::DD::valueAt); one reading the world position sees P"""
    a0 = np.array([0.2, 0.5, 0.8], np.float32)
    mesh = scenes.TriMesh(TRI, np.array([[0, 1, 2]], np.int32), None,
                          material=abi.matte(from_attribute=abi.ATTRIB_0),
                          attributes={abi.ATTRIB_SLOT_0: (abi.SCOPE_PER_VERTEX, a0)})
    img = _render(orc, _scene(mesh)).read(abi.BNX_FB_COLOR)
    u, v = _bary(orc, TRI)
    ok = _inside(u, v)
    x = (1 - u - v) * a0[0] + u * a0[1] + v * a0[2]
    assert np.allclose(img[ok][:, 0], _expected(x[ok]), rtol=3e-3)
    assert (img[ok][:, 1:3] == 0).all()
    tri = TRI + np.array([2.5, 2.5, 0.0], np.float32)
    mesh = scenes.TriMesh(tri, np.array([[0, 1, 2]], np.int32), None,
                          material=abi.matte(from_attribute=abi.ATTRIB_WORLD_POSITION))
    sc = _scene(mesh)
    s = orc.surface_scene(sc)
    fr = orc.Frame(W, W)
    cam = orc.camera_orthographic((2.5, 2.5, 5.0), (0.0, 0.0, -1.0), (0.0, 1.0, 0.0), 5.0, 1.0)
    orc.render(s, cam, abi.renderer(bg=(0.0, 0.0, 0.0, 1.0), ambient=0.0), fr, 1, threads=4)
    img = fr.read(abi.BNX_FB_COLOR)
    rays = orc.primary_rays(cam, W, W)
    P = rays[:, 0:2].reshape(W, W, 2)
    hit = fr.read(abi.BNX_FB_PRIMID) >= 0
    assert hit.sum() > 300
    assert np.allclose(img[hit][:, :2], _expected(P[hit]), rtol=1e-3)
    assert (img[hit][:, 2] == 0).all()


def test_vertex_normals_interpolate(orc):
    """Triangles::DD::normals: the shading normal is the normalized
    barycentric blend of the vertex normals (Triangles.dev.cu:68-75), so the
    pixel follows cos^2 of it against the light"""
    n = np.array([[-0.4, -0.3, 1.0], [0.5, -0.2, 1.0], [0.0, 0.45, 1.0]], np.float32)
    n /= np.linalg.norm(n, axis=1, keepdims=True)
    mesh = scenes.TriMesh(TRI, np.array([[0, 1, 2]], np.int32), None, normals=n,
                          material=abi.matte((0.7, 0.7, 0.7, 1.0), from_attribute=False))
    fr = _render(orc, _scene(mesh))
    img = fr.read(abi.BNX_FB_COLOR)
    u, v = _bary(orc, TRI)
    ok = _inside(u, v)
    ns = (1 - u - v)[..., None] * n[0].astype(np.float64) + u[..., None] * n[1] + v[..., None] * n[2]
    ns /= np.linalg.norm(ns, axis=-1, keepdims=True)
    cos = ns[..., 2]
    exp = (0.85 * 0.7 / math.pi) * cos ** 2 / math.pi ** 2 * RAD
    assert np.allclose(img[ok][:, 0], exp[ok], rtol=3e-3)
    # the normal channel holds the interpolated normal (fp16), not the face normal
    assert np.allclose(fr.read(abi.BNX_FB_NORMAL)[ok], ns[ok], atol=2e-3)
    # without them: the face normal, cos = 1
    mesh.normals = None
    img2 = _render(orc, _scene(mesh)).read(abi.BNX_FB_COLOR)
    assert np.allclose(img2[ok][:, 0], (0.85 * 0.7 / math.pi) / math.pi ** 2 * RAD, rtol=1e-3)


# ------------------------------------------------------- stochastic opacity
def _layers(orc, alphas, pbr_transmission=None):
    """square layers (two triangles each) at z = 0, -0.5, ... of the given
    opacities, one geometry each, in one triangles group"""
    s = orc.Scene()
    gs = []
    for k, a in enumerate(alphas):
        z = -0.5 * k
        v = np.array([[-2, -2, z], [2, -2, z], [2, 2, z], [-2, 2, z]], np.float32)
        i = np.array([[0, 1, 2], [0, 2, 3]], np.int32)
        if pbr_transmission is None:
            m = abi.matte((0.5, 0.5, 0.5, 1.0), from_attribute=False, opacity=a)
        else:
            m = abi.pbr((0.5, 0.5, 0.5, 1.0), from_attribute=False, opacity=a, transmission=pbr_transmission,
                        ior=1.0)
        gs.append(s.add_triangles(v, i, None, user_id=10 + k, material=m))
    s.add_instance(s.add_group(0, gs), 0)
    s.build()
    return s


def _down_rays(n, seed=1):
    rng = np.random.default_rng(seed)
    r = np.zeros((n, 8), np.float32)
    r[:, 0:2] = rng.uniform(-1.5, 1.5, (n, 2))
    r[:, 2] = 5.0
    r[:, 6] = -1.0
    r[:, 7] = 1e30
    seeds = rng.integers(0, 2 ** 63, n, dtype=np.uint64)
    return r, seeds


def test_opacity_kept_set_is_order_independent(orc):
    """each candidate's draw depends on the ray and the candidate alone, so
    the BVH walk and brute force (different test orders) keep the same
    closest candidate and reject the same one in front of it (every
    candidate in front of the kept one is tested in any order; which ones
    behind it are tested depends on the order, and nothing uses them)"""
    s = _layers(orc, [0.5, 0.35, 0.6, 0.25, 0.5])
    rays, seeds = _down_rays(4000)
    hb, rb = s.trace_opacity(rays, seeds, mode=0)
    hf, rf = s.trace_opacity(rays, seeds, mode=1)
    assert len(helpers.hits_equal(hb, hf)) == 0
    front_b, front_f = rb["t"] < hb["t"], rf["t"] < hf["t"]
    assert np.array_equal(front_b, front_f)
    assert len(helpers.hits_equal(rb[front_b], rf[front_f])) == 0
    assert 0 < (hb["prim_id"] >= 0).sum() < len(hb) and front_b.sum() > 0


def test_opacity_layer_statistics(orc):
    """k layers of alpha a: the first kept layer is j with probability
    (1 - a)^j a, and a ray passes all with (1 - a)^k; the closest rejected
    candidate is layer 0 whenever layer 0 is not kept"""
    a, k, n = 0.4, 4, 40000
    s = _layers(orc, [a] * k)
    rays, seeds = _down_rays(n, seed=7)
    hits, rej = s.trace_opacity(rays, seeds)
    layer = np.where(hits["prim_id"] >= 0, hits["geom_id"], k)
    for j in range(k + 1):
        p = (1 - a) ** j * (a if j < k else 1.0)
        got = (layer == j).mean()
        assert abs(got - p) < 5 * math.sqrt(p * (1 - p) / n), (j, got, p)
    miss0 = layer != 0
    assert (rej["geom_id"][miss0] == 0).all() and (rej["prim_id"][~miss0] == -1).all()


def test_opacity_extremes_and_pbr_alpha(orc):
    """alpha 0: never kept; alpha 1: always, with no draw; PBR alpha =
    (1 - transmission) * baseColor.w * opacity as half (AnariPBR.h:87-90)"""
    rays, seeds = _down_rays(5000, seed=3)
    h0, r0 = _layers(orc, [0.0]).trace_opacity(rays, seeds)
    assert (h0["prim_id"] == -1).sum() >= 4999 and (r0["prim_id"] >= 0).sum() >= 4999
    h1, r1 = _layers(orc, [1.0]).trace_opacity(rays, seeds)
    assert (h1["prim_id"] >= 0).all() and (r1["prim_id"] == -1).all()
    # transmission 0.25, opacity 0.8: alpha 0.6
    hp, _ = _layers(orc, [0.8], pbr_transmission=0.25).trace_opacity(rays, seeds)
    p = (hp["prim_id"] >= 0).mean()
    assert abs(p - 0.6) < 5 * math.sqrt(0.24 / 5000), p


def test_opacity_from_per_vertex_alpha(orc):
    """alpha from the w of an interpolated per-vertex color: the kept
    fraction follows it across the square"""
    s = orc.Scene()
    v = np.array([[-2, -2, 0], [2, -2, 0], [2, 2, 0], [-2, 2, 0]], np.float32)
    i = np.array([[0, 1, 2], [0, 2, 3]], np.int32)
    g = s.add_triangles(v, i, None, material=abi.matte(from_attribute=True))
    # w = 0 on the left edge (x = -2), 1 on the right (x = 2)
    s.set_attribute(g, abi.ATTRIB_SLOT_COLOR, abi.SCOPE_PER_VERTEX,
                    np.array([[1, 1, 1, 0], [1, 1, 1, 1], [1, 1, 1, 1], [1, 1, 1, 0]], np.float32))
    s.add_instance(s.add_group(0, [g]), 0)
    s.build()
    rays, seeds = _down_rays(40000, seed=5)
    hits, _ = s.trace_opacity(rays, seeds)
    x = rays[:, 0]
    for lo in np.arange(-1.5, 1.5, 0.5):
        sel = (x >= lo) & (x < lo + 0.5)
        p = ((lo + 0.25) + 2.0) / 4.0          # mean alpha over the bin
        got = (hits["prim_id"][sel] >= 0).mean()
        assert abs(got - p) < 5 * math.sqrt(p * (1 - p) / sel.sum()) + 0.01, (lo, got, p)


def _floor_and_sheet(alpha, sheet=True):
    floor = scenes.TriMesh(np.array([[-3, -3, -1], [3, -3, -1], [3, 3, -1], [-3, 3, -1]], np.float32),
                           np.array([[0, 1, 2], [0, 2, 3]], np.int32), None, user_id=1,
                           material=abi.matte((0.6, 0.6, 0.6, 1.0), from_attribute=False))
    meshes = [floor]
    if sheet:
        # a sheet over the left half of the floor, seen edge-on from the light
        meshes.append(scenes.TriMesh(np.array([[-3, -3, 1], [0, -3, 1], [0, 3, 1], [-3, 3, 1]], np.float32),
                                     np.array([[0, 1, 2], [0, 2, 3]], np.int32), None, user_id=2,
                                     material=abi.matte((0.9, 0.9, 0.9, 1.0), from_attribute=False,
                                                        opacity=alpha)))
    return scenes.SurfaceScene(meshes=meshes, dir_lights=[((0.0, 0.0, -1.0), RAD, (1.0, 1.0, 1.0))], ambient=0.0,
                               width=W, height=W)


def test_translucent_sheet_shadow_and_ids(orc):
    """An opacity-0 sheet: the frame's color equals the scene without it
    (every candidate rejected, no RNG state changed), while the ID channels
    report the sheet where it covers the floor (IDs ignore opacity,
    Triangles.dev.cu:45-60).  An alpha-0.3 sheet lets 0.7 of the light
    through to the floor under it, on average over samples."""
    s0 = _render(orc, _floor_and_sheet(0.0))
    sn = _render(orc, _floor_and_sheet(0.0, sheet=False))
    assert np.array_equal(s0.read(abi.BNX_FB_COLOR), sn.read(abi.BNX_FB_COLOR))
    cy, cx = np.mgrid[0:W, 0:W]
    x = (cx + 0.5) / W * 5 - 2.5
    left, right = x < -0.1, x > 0.1
    obj = s0.read(abi.BNX_FB_OBJID)
    assert (obj[left] == 2).all() and (obj[right] == 1).all()
    assert (sn.read(abi.BNX_FB_OBJID)[left] == 1).all()
    fr = _render(orc, _floor_and_sheet(0.3), spp=64)
    img = fr.read(abi.BNX_FB_COLOR)
    floor_lit = _expected((0.6, 0.6, 0.6))[0]
    sheet_px = _expected((0.9, 0.9, 0.9))[0]
    # left: the sheet kept with 0.3 (its own shading), else the floor, lit
    # through the sheet with 0.7
    exp_left = 0.3 * sheet_px + 0.7 * (0.7 * floor_lit)
    got = img[left][:, 0].mean()
    assert abs(got - exp_left) < 0.03 * exp_left, (got, exp_left)
    assert np.allclose(img[right][:, 0], floor_lit, rtol=1e-3)
