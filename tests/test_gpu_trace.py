"""GPU BVH builder and traversal against the CPU oracle, through the C-ABI.
Integer work (tree structure, prim order, hit IDs) and hit distances are
compared bit for bit."""
import numpy as np
import pytest

import barney_amd as bx
from barney_amd import scenes
from tests import helpers

pytestmark = pytest.mark.gpu


def _same_hits(g, o):
    bad = helpers.hits_equal(g, o)
    assert len(bad) == 0, (len(bad), g[bad[:3]], o[bad[:3]])


@pytest.mark.parametrize("n,leaf,deg", [(0, 4, False), (1, 4, False), (4, 4, False), (5, 4, False),
                                        (33, 4, False), (257, 4, False), (5000, 4, False),
                                        (100000, 4, False), (3000, 4, True), (1000, 1, False)])
def test_builder_matches_spec(dev, orc, n, leaf, deg):
    print(f"builder n={n} leaf={leaf} degenerate={deg}", flush=True)
    boxes = helpers.random_boxes(n + 11, n, deg)
    gn, go, ginfo = bx.bvh_build_boxes(dev, boxes, leaf)
    on, oo, od = orc.bvh_build_boxes(boxes, leaf)
    assert np.array_equal(go, oo)
    assert helpers.canonical_bvh(gn, go, leaf) == helpers.canonical_bvh(on, oo, leaf)
    assert ginfo.max_depth == od and ginfo.num_nodes == len(on)


def test_c1_blas_matches_spec(dev, orc):
    sc = scenes.config1()
    spec = helpers.c1_spec(sc)
    g = spec.gpu(dev)
    o = spec.oracle(orc)
    for gi, grp in enumerate(g._keep[1]):
        gn, go, gd = grp.bvh()
        on, oo, od = o.bvh(gi)
        assert np.array_equal(go, oo)
        assert helpers.canonical_bvh(gn, go, 4) == helpers.canonical_bvh(on, oo, 4)
        assert gd == od


def test_trace_c1_primary_and_random(dev, orc):
    sc = scenes.config1()
    spec = helpers.c1_spec(sc)
    g, o = spec.gpu(dev), spec.oracle(orc)
    prim = helpers.camera_rays(sc)
    _same_hits(g.trace(prim), o.trace(prim)[0])
    rays = scenes.random_rays(3, 200000)
    _same_hits(g.trace(rays), o.trace(rays)[0])


@pytest.mark.parametrize("name", ["soup", "ties", "instanced"])
def test_trace_scenes(dev, orc, name):
    spec = {"soup": helpers.soup_spec, "ties": helpers.tie_spec, "instanced": helpers.instanced_spec}[name]()
    g, o = spec.gpu(dev), spec.oracle(orc)
    if name == "ties":
        rays = helpers.grid_rays()
    elif name == "instanced":
        rays = scenes.random_rays(8, 50000, radius=9.0, spread=5.0)
    else:
        rays = scenes.random_rays(4, 50000, radius=5.0, spread=2.0)
    _same_hits(g.trace(rays), o.trace(rays)[0])


def test_any_hit_occlusion(dev, orc):
    spec = helpers.c1_spec(scenes.config1())
    g, o = spec.gpu(dev), spec.oracle(orc)
    rays = scenes.random_rays(5, 100000)
    ga = g.trace(rays, any_hit=True)
    oc = o.trace(rays, 0)[0]
    assert np.array_equal(ga["prim_id"] >= 0, oc["prim_id"] >= 0)


def test_edge_cases(dev, orc):
    # empty group next to a single triangle; zero direction components; tiny tmax
    spec = helpers.SceneSpec()
    e = spec.tri(np.zeros((0, 3), np.float32), np.zeros((0, 3), np.int32), None, user_id=1)
    t = spec.tri(np.array([[-1, -1, 0], [1, -1, 0], [0, 1, 0]], np.float32), np.array([[0, 1, 2]], np.int32),
                 None, user_id=2)
    spec.inst(spec.group(0, [e]), 0)
    spec.inst(spec.group(0, [t]), 1)
    g, o = spec.gpu(dev), spec.oracle(orc)
    rays = np.zeros((6, 8), np.float32)
    rays[:, 2] = 3.0
    rays[:, 4:7] = (0.0, 0.0, -1.0)           # x, y components exactly zero
    rays[:, 7] = [1e30, 3.0, 2.9999, 0.5, 1e30, 1e30]
    rays[4, 0] = 2.0                          # miss
    rays[5, 3] = 3.5                          # tmin beyond the hit
    _same_hits(g.trace(rays), o.trace(rays)[0])
    # t == tmax is not a hit (rtcore/hiprt/TraceInterface.h:51 rejects t >= tMax)
    assert g.trace(rays)["prim_id"].tolist() == [0, -1, -1, -1, -1, -1]


def test_instance_transform_update(dev, orc):
    """bnx_group_set_transforms (bnUpdateInstanceTransforms: new transforms,
    then the instance BVH built again): hits after the update equal those of
    a group built from scratch with the new transforms and the oracle's,
    bit for bit; a count mismatch or a geometry group is rejected."""
    rays = scenes.random_rays(8, 50000, radius=9.0, spread=5.0)
    g = helpers.instanced_spec().gpu(dev)
    before = g.trace(rays)
    ident = (1, 0, 0, 0, 1, 0, 0, 0, 1)
    rot_z = (0, 1, 0, -1, 0, 0, 0, 0, 1)              # x -> y, y -> -x
    moved = [(ident, (1, 0, 0)), (rot_z, (4, 1, 0)), (ident, (-3, 2, -1)), (rot_z, (0, -4, 2))]
    g.set_transforms(np.stack([helpers.affine_pair(lin, off)[0] for lin, off in moved]))
    after = g.trace(rays)
    spec2 = helpers.instanced_spec(moved)
    _same_hits(after, spec2.gpu(dev).trace(rays))
    _same_hits(after, spec2.oracle(orc).trace(rays)[0])
    assert not np.array_equal(after["prim_id"], before["prim_id"])
    # back to the creation transforms: the first hits again
    g.set_transforms(np.stack([helpers.affine_pair(ident, off)[0] for off in helpers.INSTANCE_OFFSETS]))
    _same_hits(g.trace(rays), before)
    with pytest.raises(bx.BnxError):
        g.set_transforms(np.zeros((3, 12), np.float32))
    with pytest.raises(bx.BnxError):
        g._keep[1][0].set_transforms(np.zeros((1, 12), np.float32))
