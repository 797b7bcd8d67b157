"""Shared helpers for the parity tests (scene construction on both sides,
canonical BVH comparison)."""
import math
import os

import numpy as np

from barney_amd import abi, scenes


def canonical_bvh(nodes, order, max_leaf):
    """DFS serialisation of a BVH independent of node numbering: a list of
    (child box, leaf slot range or subtree marker) in pre-order."""
    out = []
    if len(nodes) == 0:
        return out
    stack = [0]
    while stack:
        i = stack.pop()
        nd = nodes[i]
        rec = []
        for s in (0, 1):
            lo = nd["lo0"] if s == 0 else nd["lo1"]
            hi = nd["hi0"] if s == 0 else nd["hi1"]
            cnt = int(nd["cnt"][s])
            ref = int(nd["ref"][s])
            if cnt < 0:
                rec.append(("empty",))
            elif cnt > 0:
                assert cnt <= max_leaf
                rec.append(("leaf", tuple(lo.tolist()), tuple(hi.tolist()), ref, cnt))
            else:
                rec.append(("node", tuple(lo.tolist()), tuple(hi.tolist())))
        out.append(tuple(rec))
        for s in (1, 0):
            if int(nd["cnt"][s]) == 0:
                stack.append(int(nd["ref"][s]))
    return out


def random_boxes(seed, n, degenerate=False):
    rng = np.random.default_rng(seed)
    if degenerate:
        c = np.zeros((n, 3), np.float32) + np.float32(0.25)
    else:
        c = rng.random((n, 3)).astype(np.float32) * 10
    e = (rng.random((n, 3)).astype(np.float32) * 0.3 + 0.01).astype(np.float32)
    return np.concatenate([c - e, c + e], 1).astype(np.float32)


def triangle_soup(seed, n, extent=4.0, size=0.4):
    rng = np.random.default_rng(seed)
    c = (rng.random((n, 3)) - 0.5) * extent
    v = c[:, None, :] + (rng.random((n, 3, 3)) - 0.5) * size
    verts = v.reshape(-1, 3).astype(np.float32)
    idx = np.arange(3 * n, dtype=np.int32).reshape(n, 3)
    col = rng.random((n, 3)).astype(np.float32)
    return verts, idx, col


def tie_scene():
    """Coplanar, shared-edge and exactly duplicated triangles plus two
    identical spheres: rays through shared edges and duplicates produce equal
    t for several prims, so the lower-ID rule decides."""
    g = np.linspace(-1.0, 1.0, 9, dtype=np.float32)
    verts, idx = [], []
    for j in range(8):
        for i in range(8):
            a = len(verts)
            verts += [(g[i], g[j], 0.0), (g[i + 1], g[j], 0.0), (g[i + 1], g[j + 1], 0.0), (g[i], g[j + 1], 0.0)]
            idx += [(a, a + 1, a + 2), (a, a + 2, a + 3)]
    verts = np.array(verts, np.float32)
    idx = np.array(idx, np.int32)
    idx = np.concatenate([idx, idx[:10]])            # exact duplicates of prims 0..9
    centers = np.array([[0.0, 0.0, 2.0], [0.0, 0.0, 2.0], [0.5, 0.5, 2.0]], np.float32)
    radii = np.array([0.25, 0.25, 0.125], np.float32)
    return verts, idx, centers, radii


def grid_rays(z0=5.0, n=64, span=1.2, target_z=0.0):
    """Rays from a plane above the scene straight down through a grid that
    includes exact edge and vertex coordinates of tie_scene()."""
    xs = np.linspace(-span, span, n, dtype=np.float32)
    X, Y = np.meshgrid(xs, xs)
    rays = np.zeros((n * n, 8), np.float32)
    rays[:, 0], rays[:, 1], rays[:, 2] = X.ravel(), Y.ravel(), z0
    rays[:, 4:7] = (0.0, 0.0, -1.0)
    rays[:, 7] = 1e30
    return rays


def camera_rays(sc):
    """Primary rays of config scene `sc` at pixel centres (generateRays.cu:
    accumID 0), in bnx_ray layout, computed in float32 like the kernels."""
    import oracle
    cam = oracle.camera_perspective(sc.camera["position"], sc.camera["direction"], sc.camera["up"],
                                    sc.camera["fovy"])
    W, H = sc.width, sc.height
    ix, iy = np.meshgrid(np.arange(W, dtype=np.float32), np.arange(H, dtype=np.float32))
    u = (ix + np.float32(0.5)) / np.float32(W)
    v = (iy + np.float32(0.5)) / np.float32(H)
    aspect = np.float32(W) / np.float32(H)
    d00, du, dv = (np.array(list(x), np.float32) for x in (cam.dir_00, cam.dir_du, cam.dir_dv))
    su = (np.float32(1.0) * aspect) * (u - np.float32(0.5))
    sv = np.float32(1.0) * (v - np.float32(0.5))
    d = (d00[None, None, :] + su[..., None] * du[None, None, :]) + sv[..., None] * dv[None, None, :]
    d = d.astype(np.float32)
    nrm = np.float32(1.0) / np.sqrt(((d[..., 0] * d[..., 0] + d[..., 1] * d[..., 1]) + d[..., 2] * d[..., 2]))
    d = d * nrm[..., None]
    rays = np.zeros((H * W, 8), np.float32)
    rays[:, 0:3] = np.array(sc.camera["position"], np.float32)
    rays[:, 4:7] = d.reshape(-1, 3)
    rays[:, 7] = 1e30
    return rays


def hits_equal(a, b):
    """Bit-exact comparison of two bnx_hit arrays; returns mismatch indices."""
    bad = np.zeros(len(a), bool)
    for f in ("prim_id", "geom_id", "inst", "user_inst_id", "obj_id"):
        bad |= a[f] != b[f]
    for f in ("t", "u", "v"):
        bad |= a[f].view(np.uint32) != b[f].view(np.uint32)
    return np.nonzero(bad)[0]


class SceneSpec:
    """A scene described once and built on both sides: groups of geometries
    (kind 0 triangles, 1 spheres, 2 cylinders, 3 capsules, 4 cones; a group
    of kind 1 holds one user type) and instances (group, user id, o2w, w2o)."""

    def __init__(self):
        self.geoms = []      # (kind, kwargs)
        self.groups = []     # (kind, [geom indices])
        self.insts = []      # (group, user_id, o2w or None, w2o or None)
        self.lights = []

    def tri(self, vertices, indices, colors=None, user_id=0, material=None):
        self.geoms.append((0, dict(vertices=vertices, indices=indices, colors=colors, user_id=user_id,
                                   material=material)))
        return len(self.geoms) - 1

    def sph(self, centers, radii, colors=None, user_id=0, material=None):
        self.geoms.append((1, dict(centers=centers, radii=radii, colors=colors, user_id=user_id,
                                   material=material)))
        return len(self.geoms) - 1

    def cyl(self, vertices, indices, radii=None, colors=None, user_id=0, material=None, default_radius=0.1):
        self.geoms.append((2, dict(vertices=vertices, indices=indices, radii=radii, colors=colors, user_id=user_id,
                                   material=material, default_radius=default_radius)))
        return len(self.geoms) - 1

    def cap(self, vertices, indices, colors=None, user_id=0, material=None):
        self.geoms.append((3, dict(vertices=vertices, indices=indices, colors=colors, user_id=user_id,
                                   material=material)))
        return len(self.geoms) - 1

    def cone(self, vertices, radii, indices=None, colors=None, user_id=0, material=None):
        self.geoms.append((4, dict(vertices=vertices, radii=radii, indices=indices, colors=colors, user_id=user_id,
                                   material=material)))
        return len(self.geoms) - 1

    def group(self, kind, geoms):
        self.groups.append((kind, list(geoms)))
        return len(self.groups) - 1

    def inst(self, group, user_id=0, o2w=None, w2o=None):
        self.insts.append((group, user_id, o2w, w2o))

    def oracle(self, orc):
        s = orc.Scene()
        ids = []
        for kind, kw in self.geoms:
            if kind == 0:
                ids.append(s.add_triangles(kw["vertices"], kw["indices"], kw["colors"], user_id=kw["user_id"],
                                           material=kw["material"]))
            elif kind == 1:
                ids.append(s.add_spheres(kw["centers"], kw["radii"], kw["colors"], user_id=kw["user_id"],
                                         material=kw["material"]))
            elif kind == 2:
                ids.append(s.add_cylinders(**kw))
            elif kind == 3:
                ids.append(s.add_capsules(**kw))
            else:
                ids.append(s.add_cones(**kw))
        gids = [s.add_group(k, [ids[g] for g in gl]) for k, gl in self.groups]
        for g, uid, o2w, w2o in self.insts:
            s.add_instance(gids[g], uid, o2w, w2o)
        s.set_dir_lights(self.lights)
        s.build()
        return s

    def gpu(self, dev):
        import barney_amd as bx
        geoms = []
        for kind, kw in self.geoms:
            if kind == 0:
                geoms.append(bx.Geom.triangles(dev, kw["vertices"], kw["indices"], kw["colors"],
                                               user_id=kw["user_id"], material=kw["material"]))
            elif kind == 1:
                geoms.append(bx.Geom.spheres(dev, kw["centers"], kw["radii"], kw["colors"],
                                             user_id=kw["user_id"], material=kw["material"]))
            elif kind == 2:
                geoms.append(bx.Geom.cylinders(dev, **kw))
            elif kind == 3:
                geoms.append(bx.Geom.capsules(dev, **kw))
            else:
                geoms.append(bx.Geom.cones(dev, **kw))
        groups = [bx.Group.geoms(dev, k, [geoms[g] for g in gl]).build() for k, gl in self.groups]
        xf = None
        if any(o is not None for _, _, o, _ in self.insts):
            ident = np.array([1, 0, 0, 0, 1, 0, 0, 0, 1, 0, 0, 0], np.float32)
            xf = np.stack([o if o is not None else ident for _, _, o, _ in self.insts])
        insts = bx.Group.instances(dev, [groups[g] for g, _, _, _ in self.insts],
                                   [u for _, u, _, _ in self.insts], xf).build()
        insts._keep = (geoms, groups)
        return insts


def c1_spec(sc):
    spec = SceneSpec()
    sph = [spec.sph(s.centers, s.radii, s.colors, user_id=s.user_id, material=s.material) for s in sc.spheres]
    cyl = [spec.cyl(c.vertices, c.indices, c.radii, c.colors, user_id=c.user_id, material=c.material)
           for c in sc.cylinders]
    cap = [spec.cap(c.vertices, c.indices, c.colors, user_id=c.user_id, material=c.material)
           for c in getattr(sc, "capsules", [])]
    cone = [spec.cone(c.vertices, c.radii, c.indices, c.colors, user_id=c.user_id, material=c.material)
            for c in getattr(sc, "cones", [])]
    tri = [spec.tri(m.vertices, m.indices, m.colors, user_id=m.user_id, material=m.material) for m in sc.meshes]
    if sph:
        spec.inst(spec.group(1, sph), 0)
    if cyl:
        spec.inst(spec.group(1, cyl), 0)
    if cap:
        spec.inst(spec.group(1, cap), 0)
    if cone:
        spec.inst(spec.group(1, cone), 0)
    if tri:
        spec.inst(spec.group(0, tri), 0)
    spec.lights = list(sc.dir_lights)
    return spec


def soup_data():
    """the triangle soup (3000 triangles) and 500 spheres of soup_spec"""
    verts, idx, col = triangle_soup(9, 3000)
    rng = np.random.default_rng(1)
    cs = ((rng.random((500, 3)) - 0.5) * 4).astype(np.float32)
    rs = (rng.random(500) * 0.1 + 0.01).astype(np.float32)
    return verts, idx, col, cs, rs


def soup_spec(move=None, sticks=None, vertex_move=None):
    """instance 0: the spheres' group; instance 1: a group of two triangle
    geometries (the soup's halves), translated by `move` if given; with
    `sticks` (a CylinderSet), instance 2: its cylinders (user ID 3);
    `vertex_move` adds an offset to every triangle vertex (float32)"""
    verts, idx, col, cs, rs = soup_data()
    if vertex_move is not None:
        verts = (verts + np.asarray(vertex_move, np.float32)).astype(np.float32)
    spec = SceneSpec()
    g0 = spec.tri(verts, idx[:1500], col[:1500], user_id=0)
    g1 = spec.tri(verts, idx[1500:], col[1500:], user_id=1)
    g2 = spec.sph(cs, rs, None, user_id=2)
    spec.inst(spec.group(1, [g2]), 0)
    if move is None:
        spec.inst(spec.group(0, [g0, g1]), 0)
    else:
        o2w, w2o = affine_pair((1, 0, 0, 0, 1, 0, 0, 0, 1), move)
        spec.inst(spec.group(0, [g0, g1]), 0, o2w, w2o)
    if sticks is not None:
        spec.inst(spec.group(1, [spec.cyl(sticks.vertices, sticks.indices, sticks.radii, None, user_id=3)]), 0)
    return spec


def tie_spec():
    v, i, c, r = tie_scene()
    spec = SceneSpec()
    ga = spec.tri(v, i, None, user_id=5)
    gb = spec.tri(v, i[:20], None, user_id=6)
    gs = spec.sph(c, r, None, user_id=7)
    spec.inst(spec.group(1, [gs]), 0)
    spec.inst(spec.group(0, [ga, gb]), 0)
    return spec


INSTANCE_OFFSETS = [(0, 0, 0), (4, 0, 0), (-4, 2, 0), (0, -4, 1)]


def affine_pair(lin, off):
    """object-to-world (12 floats: 3x3 column-major, then translation) and
    its inverse, for a linear part whose inverse is its transpose with
    entries in {-1, 0, 1} (so both are exact in float32)"""
    L = np.array(lin, np.float64).reshape(3, 3).T          # columns are the images of x, y, z
    t = np.array(off, np.float64)
    Li = L.T
    ti = -Li @ t
    o2w = np.concatenate([L.T.reshape(-1), t]).astype(np.float32)
    w2o = np.concatenate([Li.T.reshape(-1), ti]).astype(np.float32)
    return o2w, w2o


def sticks_spec(rotated=False):
    """cylinders (two geometries: sticks with and without per-cylinder radii)
    beside the soup's spheres and triangles; with `rotated`, the cylinders'
    group is also instanced a second time, rotated and moved"""
    verts, idx, col, cs, rs = soup_data()
    st = scenes.sticks(600, extent=1.8)
    spec = SceneSpec()
    c0 = spec.cyl(st.vertices, st.indices[:400], st.radii[:400], st.colors[:400], user_id=3)
    c1 = spec.cyl(st.vertices, st.indices[400:], None, None, user_id=4, default_radius=0.03)
    gs = spec.group(1, [spec.sph(cs, rs, None, user_id=2)])
    gc = spec.group(1, [c0, c1])
    gt = spec.group(0, [spec.tri(verts, idx[:1000], col[:1000], user_id=0)])
    spec.inst(gs, 0)
    spec.inst(gc, 0)
    spec.inst(gt, 0)
    if rotated:
        o2w, w2o = affine_pair((0, 1, 0, 0, 0, 1, 1, 0, 0), (0.5, -0.25, 1.0))
        spec.inst(gc, 7, o2w, w2o)
    return spec


def instanced_spec(xforms=None):
    """four instances of one triangle group; xforms: (linear 3x3 column-major,
    translation) per instance, translations only by default"""
    verts, idx, col = triangle_soup(2, 400, extent=2.0)
    spec = SceneSpec()
    grp = spec.group(0, [spec.tri(verts, idx, col, user_id=0)])
    ident = (1, 0, 0, 0, 1, 0, 0, 0, 1)
    xforms = xforms or [(ident, off) for off in INSTANCE_OFFSETS]
    for k, (lin, off) in enumerate(xforms):
        o2w, w2o = affine_pair(lin, off)
        spec.inst(grp, 10 + k, o2w, w2o)
    return spec


# ---------------------------------------------------------------------------
# the scene of the reference's unitTests/omnicamera.py (its fixture is
# tests/golden/reference_unitTest_omnicamera.png): three matte spheres of
# radius 0.2 at camera_from + right / front / top, an omnidirectional camera
# at camera_from, ambient radiance 0.8, black background, 2048 x 1024
OMNI_FROM = (-3.0, 1.5, -2.0)
OMNI_DIR = (1.0, 0.1, 0.3)
OMNI_UP = (0.0, 1.0, 0.0)
OMNI_SIZE = (2048, 1024)
OMNI_PNG = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "reference_unitTest_omnicamera.png")


def _nrm64(a):
    a = np.asarray(a, np.float64)
    return a / math.sqrt(float(a @ a))


def omni_spec():
    """the spheres as omnicamera.py places them (positions in float64, then
    float32 as its numpy arrays are); one user group, one instance"""
    front = _nrm64(OMNI_DIR)
    right = _nrm64(np.cross(front, OMNI_UP))
    top = _nrm64(np.cross(right, front))
    spec = SceneSpec()
    gs = []
    for k, (off, col) in enumerate(((right, (0.7, 0.1, 0.1)), (front, (0.1, 0.7, 0.1)), (top, (0.1, 0.1, 0.7)))):
        pos = np.array([np.add(OMNI_FROM, off)], np.float32)
        gs.append(spec.sph(pos, np.array([0.2], np.float32), None, user_id=k,
                           material=abi.matte(col + (1.0,), from_attribute=False)))
    spec.inst(spec.group(1, gs), 0)
    return spec


def omni_camera_args():
    """position, direction, up as the ANARI camera passes them on
    (anari/Camera.cpp:37-46 normalizes direction and up, and bnSet3f of
    'direction' normalizes again, native/Camera.cu:226-229)"""
    d = _nrm64(OMNI_DIR).astype(np.float32)
    d = (d / np.float32(math.sqrt(float(d.astype(np.float64) @ d.astype(np.float64))))).astype(np.float32)
    u = _nrm64(OMNI_UP).astype(np.float32)
    return np.array(OMNI_FROM, np.float32), d, u


def srgb_to_linear(c8):
    c = np.asarray(c8, np.float64) / 255.0
    return np.where(c <= 0.04045, c / 12.92, ((c + 0.055) / 1.055) ** 2.4)


def check_against_omni_fixture(srgb8, objid):
    """srgb8: (H, W, 4) uint8 frame in framebuffer row order (row 0 at the
    bottom), objid: (H, W) object IDs.  Compares with the reference PNG:
    every pixel hit is lit in the fixture; every lit fixture pixel is hit or
    next to a hit pixel (the denoiser's bleed); per-sphere mean colour over
    interior pixels, in linear space, within 1.5%."""
    from PIL import Image
    ref = np.array(Image.open(OMNI_PNG).convert("RGB")).astype(np.int64)[::-1]   # back to framebuffer rows
    img = np.asarray(srgb8)[..., :3].astype(np.int64)
    assert ref.shape == img.shape, (ref.shape, img.shape)
    hit = objid >= 0
    lit = (ref != 0).any(-1)
    assert not (hit & ~lit).any(), int((hit & ~lit).sum())
    grown = hit.copy()
    for dy in (-1, 0, 1):
        for dx in (-1, 0, 1):
            grown |= np.roll(np.roll(hit, dy, 0), dx, 1)
    assert not (lit & ~grown).any(), int((lit & ~grown).sum())
    assert hit.sum() > 150000 and (lit & ~hit).sum() < 0.02 * hit.sum()
    interior = hit.copy()
    for dy in (-1, 0, 1):
        for dx in (-1, 0, 1):
            interior &= np.roll(np.roll(objid, dy, 0), dx, 1) == objid
    for k, col in enumerate(((0.7, 0.1, 0.1), (0.1, 0.7, 0.1), (0.1, 0.1, 0.7))):
        m = interior & (objid == k)
        assert m.sum() > 10000, (k, int(m.sum()))
        ours = srgb_to_linear(img[m]).mean(0)
        theirs = srgb_to_linear(ref[m]).mean(0)
        rel = ours / theirs - 1.0
        print(f"omni fixture sphere {k}: ours {ours}, reference {theirs}, rel {rel}", flush=True)
        assert np.all(np.abs(rel) < 0.015), (k, rel)
        # the colour ratio of the material survives: 0.7 / 0.1 per channel
        big = int(np.argmax(col))
        assert 6.0 < ours[big] / np.delete(ours, big).mean() < 8.0
