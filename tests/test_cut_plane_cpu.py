"""The renderer's cut plane in the oracle's traversal (CPU only).

Every surface program of the reference rejects a hit whose world position
lies on the plane's negative side (OptixGlobals::hitOnInvisibleSide,
native/render/OptixGlobals.h:36-45; Triangles.dev.cu:39, Spheres.dev.cu:170,
Cylinders.dev.cu:258), in every trace: primary, shadow and bounce rays.
Pinned here by closed forms (a sphere whose near side is cut shows no far
side, a quad behind it shows through), by brute force, and by the defining
property on random scenes: every accepted hit is on the visible side, and a
ray whose uncut hit is visible keeps it."""
import numpy as np

from barney_amd import scenes
from tests import helpers


def _sphere_and_quad(orc):
    s = orc.Scene()
    sp = s.add_spheres(np.zeros((1, 3), np.float32), np.ones(1, np.float32), None, user_id=1)
    v = np.array([[-3, -3, -2], [3, -3, -2], [3, 3, -2], [-3, 3, -2]], np.float32)
    q = s.add_triangles(v, np.array([[0, 1, 2], [0, 2, 3]], np.int32), None, user_id=2)
    s.add_instance(s.add_group(1, [sp]), 0)
    s.add_instance(s.add_group(0, [q]), 0)
    s.build()
    return s


def _down_ray(x=0.0, y=0.0):
    r = np.zeros((1, 8), np.float32)
    r[0, 0:3] = (x, y, 5.0)
    r[0, 4:7] = (0.0, 0.0, -1.0)
    r[0, 7] = 1e30
    return r


def test_cut_plane_closed_form(orc):
    s = _sphere_and_quad(orc)
    ray = _down_ray()
    # no plane / plane keeping z >= 0.5: the sphere's near side at t = 4
    for cp in (None, (0.0, 0.0, 1.0, -0.5), (0.0, 0.0, 0.0, -1e30)):
        h = s.trace(ray, cut_plane=cp)[0][0]
        assert h["obj_id"] == 1 and h["t"] == np.float32(4.0), cp
    # keeping z <= 0.5: the near side (z = 1) is cut; the sphere program
    # reports only its first root past tmin, so the ray goes on to the quad
    h = s.trace(ray, cut_plane=(0.0, 0.0, -1.0, 0.5))[0][0]
    assert h["obj_id"] == 2 and h["t"] == np.float32(7.0)
    # keeping z <= -3: everything is cut
    h = s.trace(ray, cut_plane=(0.0, 0.0, -1.0, -3.0))[0][0]
    assert h["prim_id"] == -1
    # a plane x = 0 keeping x <= 0: the left half of the scene stays
    # visible, the right half (sphere and quad alike) is gone
    for x, obj in ((-0.5, 1), (0.5, -1)):
        h = s.trace(_down_ray(x), cut_plane=(-1.0, 0.0, 0.0, 0.0))[0][0]
        assert h["obj_id"] == obj, x
    # any-hit queries see the same visibility
    occ = s.trace(ray, mode=2, cut_plane=(0.0, 0.0, -1.0, -3.0))[0][0]
    assert occ["prim_id"] == -1


def _world_hit_points(rays, hits):
    d = rays[:, 4:7].astype(np.float32)
    d = np.where(d == 0, np.float32(1e-6), d)
    d = np.where(np.abs(d) < 1e-6, np.float32(1e-6), d)
    return rays[:, 0:3] + hits["t"][:, None] * d


def test_cut_plane_property_random_scenes(orc):
    """sticks (cylinders), spheres and triangles, with a rotated instance:
    BVH == brute force under the plane; accepted hits lie on the visible
    side; rays whose uncut hit is visible keep it bit for bit"""
    s = helpers.sticks_spec(rotated=True).oracle(orc)
    rays = scenes.random_rays(5, 40000, radius=4.0, spread=2.0)
    plane = np.array([0.3, -0.5, 0.81, 0.2], np.float32)
    cut, _ = s.trace(rays, cut_plane=plane)
    brute, _ = s.trace(rays, mode=1, cut_plane=plane)
    assert len(helpers.hits_equal(cut, brute)) == 0
    free, _ = s.trace(rays)
    hit = cut["prim_id"] >= 0
    P = _world_hit_points(rays, cut)
    side = (P.astype(np.float32) @ plane[:3]).astype(np.float64) + plane[3]
    assert (side[hit] >= -1e-4).all()
    Pf = _world_hit_points(rays, free)
    fside = (Pf @ plane[:3]).astype(np.float64) + plane[3]
    keep = (free["prim_id"] >= 0) & (fside > 1e-4)
    assert keep.sum() > 2000
    assert len(helpers.hits_equal(cut[keep], free[keep])) == 0
    cutaway = (free["prim_id"] >= 0) & (fside < -1e-4)
    assert cutaway.sum() > 2000
    # every kind of surface is cut somewhere
    for obj in (0, 2, 3):
        assert ((free["obj_id"] == obj) & cutaway).sum() > 20, obj
    moved = cutaway & ((cut["prim_id"] < 0) | (cut["t"] > free["t"]))
    assert moved.sum() == cutaway.sum()
