"""CPU oracle of the unstructured (tet) field, its macrocells, element BVH
and the isosurface program, pinned by closed forms and brute force."""
import numpy as np
import pytest

from barney_amd import abi, scenes


def affine_scene(n=5, mc=(8, 8, 8), coef=(0.7, -0.4, 0.25), bias=0.1, volume=True):
    us = scenes.config4(n=n, width=32, height=32, spp=1, mc_dims=mc, volume=volume)
    v = us.vertices.astype(np.float64)
    us.scalars = (v @ np.array(coef) + bias).astype(np.float32)
    us.tf_domain = (float(us.scalars.min()), float(us.scalars.max()))
    return us


def test_lattice_mesh_is_conforming():
    us = scenes.config4(n=4, width=8, height=8, spp=1)
    v = us.vertices.astype(np.float64)
    a, b, c, d = (v[us.tets[:, q]] for q in range(4))
    vol = np.einsum("ij,ij->i", np.cross(b - a, c - a), d - a) / 6.0
    assert (vol > 0).all()
    assert abs(vol.sum() - 1.0) < 1e-9
    # every face is shared by two tets, or lies on the domain boundary
    faces = np.sort(np.concatenate([us.tets[:, [0, 1, 2]], us.tets[:, [0, 1, 3]], us.tets[:, [0, 2, 3]],
                                    us.tets[:, [1, 2, 3]]]), axis=1)
    uf, cnt = np.unique(faces, axis=0, return_counts=True)
    assert cnt.max() == 2
    fc = v[uf].mean(axis=1)
    on_boundary = ((fc < 1e-9) | (fc > 1 - 1e-9)).any(axis=1)
    assert (on_boundary == (cnt == 1)).all()
    assert len(us.tets) == 5 * 4 ** 3 and len(us.vertices) == 5 ** 3


def test_affine_field_is_reproduced(orc):
    """tet interpolation is exact for an affine field: every point query
    inside the domain returns the affine value (BVH walk + containment +
    barycentric weights), and NaN outside"""
    us = affine_scene()
    o = orc.umesh_scene(us)
    rng = np.random.default_rng(5)
    P = rng.uniform(0.0, 1.0, (400, 3)).astype(np.float32)
    for p in P:
        f, _ = o.umesh_sample(p)
        ref = float(p.astype(np.float64) @ np.array([0.7, -0.4, 0.25]) + 0.1)
        assert abs(f - ref) < 2e-5, (p, f, ref)
    for p in [(-0.01, 0.5, 0.5), (0.5, 1.01, 0.5), (2, 2, 2)]:
        assert np.isnan(o.umesh_sample(np.array(p, np.float32))[0])


def test_sample_agrees_with_brute_force(orc):
    us = scenes.config4(n=4, width=8, height=8, spp=1)
    o = orc.umesh_scene(us)
    v = us.vertices.astype(np.float64)
    T = v[us.tets]                                # (nt, 4, 3)
    S = us.scalars.astype(np.float64)[us.tets]
    M = np.transpose(T[:, 1:] - T[:, :1], (0, 2, 1))
    Minv = np.linalg.inv(M)
    rng = np.random.default_rng(9)
    for p in rng.uniform(-0.05, 1.05, (300, 3)):
        lam = np.einsum("tij,tj->ti", Minv, p - T[:, 0])
        bary = np.concatenate([1 - lam.sum(1, keepdims=True), lam], 1)
        inside = (bary >= -1e-9).all(1)
        f, _ = o.umesh_sample(p.astype(np.float32))
        if not inside.any():
            assert np.isnan(f)
            continue
        k = np.nonzero(inside)[0][0]
        ref = (bary[k] * S[k]).sum()
        assert abs(f - ref) < 1e-5


def test_macrocells_bound_the_field(orc):
    us = scenes.config4(n=6, width=8, height=8, spp=1, mc_dims=(8, 8, 8))
    o = orc.umesh_scene(us)
    dims, lo, hi = o.mc_ranges(0)
    assert tuple(dims) == (8, 8, 8)
    rng = np.random.default_rng(3)
    P = rng.uniform(0.0, 1.0, (2000, 3)).astype(np.float32)
    for p in P:
        f, _ = o.umesh_sample(p)
        c = np.minimum((p * 8).astype(int), 7)
        i = c[0] + 8 * (c[1] + 8 * c[2])
        assert lo[i] <= f <= hi[i]
    # every cell is covered by some tet, and its range lies in the field's range
    assert np.isfinite(lo).all() and (lo <= hi).all()
    assert lo.min() >= us.scalars.min() and hi.max() <= us.scalars.max()
    # majorants = TF majorant over the ranges, zero only where alpha is zero
    mdims, maj = o.volume_majorants(0)
    assert (maj >= 0).all() and tuple(mdims) == (8, 8, 8)


def test_reference_grid_size_formula(orc):
    us = scenes.config4(n=3, width=8, height=8, spp=1, mc_dims=(0, 0, 0))
    o = orc.umesh_scene(us)
    dims, _, _ = o.mc_ranges(0)
    # G = 200 + int(sqrt(135 / 1000)) = 200 on the unit cube: 1 + 1 * 199 / 1
    assert tuple(dims) == (200, 200, 200)


def test_element_bvh_is_the_spec_build(orc):
    us = scenes.config4(n=4, width=8, height=8, spp=1)
    o = orc.umesh_scene(us)
    nodes, order, depth = o.element_bvh(0)
    v = us.vertices
    T = v[us.tets]
    boxes = np.concatenate([T.min(1), T.max(1)], 1).astype(np.float32)
    rn, ro, _ = orc.bvh_build_boxes(boxes, 4)
    assert np.array_equal(order, ro)
    assert nodes.tobytes() == rn.tobytes()
    assert sorted(order.tolist()) == list(range(len(us.tets)))


def test_isosurface_of_a_linear_field(orc):
    """f = x: the surface f = 0.3 is the plane x = 0.3 with normal +x"""
    us = affine_scene(n=5, coef=(1.0, 0.0, 0.0), bias=0.0, volume=False)
    us.iso_value = 0.3
    o = orc.umesh_scene(us)
    # along +x through the middle: hits x = 0.3 at t = 1.3
    t, N = o.iso_march([-1.0, 0.5, 0.45], [1.0, 0.0, 0.0], 1e30)
    assert abs(t - 1.3) < 2e-5 and np.allclose(N, [1, 0, 0], atol=1e-4)
    # from the other side: t = 0.7 from x = 1, normal still along +x (gradient)
    t, N = o.iso_march([1.0, 0.37, 0.61], [-1.0, 0.0, 0.0], 1e30)
    assert abs(t - 0.7) < 2e-5 and np.allclose(N, [1, 0, 0], atol=1e-4)
    # oblique ray: analytic crossing
    org = np.array([-0.5, 0.2, 0.1])
    d = np.array([1.0, 0.4, 0.5])
    d /= np.linalg.norm(d)
    t, N = o.iso_march(org, d, 1e30)
    assert abs(t - (0.3 - org[0]) / d[0]) < 5e-5
    # clipped by tmax, parallel to the plane, or missing the box: no hit
    assert o.iso_march([-1.0, 0.5, 0.5], [1.0, 0.0, 0.0], 1.2)[0] == -1.0
    assert o.iso_march([0.2, -1.0, 0.5], [0.0, 1.0, 0.0], 1e30)[0] == -1.0
    assert o.iso_march([-1.0, 2.0, 0.5], [1.0, 0.0, 0.0], 1e30)[0] == -1.0


def test_isosurface_render_ids(orc):
    """an iso-only render of f = x: pixels whose primary ray meets the plane
    x = 0.3 inside the unit cube show the isosurface's IDs and depth"""
    us = affine_scene(n=4, coef=(1.0, 0.0, 0.0), bias=0.0, volume=False)
    us.iso_value = 0.3
    us.camera = dict(position=(-2.0, 0.5, 0.5), direction=(1.0, 0.0, 0.0), up=(0.0, 1.0, 0.0), fovy=30.0,
                     target=(0.5, 0.5, 0.5))
    o = orc.umesh_scene(us)
    cam = orc.camera_perspective(us.camera["position"], us.camera["direction"], us.camera["up"], us.camera["fovy"])
    fr = orc.Frame(32, 32)
    orc.render(o, cam, abi.renderer(), fr, 1)
    obj = fr.read(abi.BNX_FB_OBJID)
    inst = fr.read(abi.BNX_FB_INSTID)
    depth = fr.read(abi.BNX_FB_DEPTH)
    hit = obj == us.iso_user_id
    assert hit.sum() > 100 and (inst[hit] == 1).all()
    # depth along the ray = 2.3 / cos(angle to +x) >= 2.3
    assert (depth[hit] >= 2.3 - 1e-4).all() and (depth[hit] < 2.4).all()
    assert (obj[~hit] == -1).all()
