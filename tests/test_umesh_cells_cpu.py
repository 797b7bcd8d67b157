"""Pyramids, prisms and hexahedra in the oracle (CPU only).

The oracle inverts each cell's isoparametric map by Newton's method as
ElementIntersection.h:77-476 does, with the shape functions in VTK vertex
order.  Pinned by what the mathematics fixes:

* Each cell's interpolation reproduces every field in the span of its shape
  functions.  On the unit cube the trilinear hexahedron gives f = x y z
  exactly; the unit right prism gives f = x z + y; the pyramid over the unit
  square with its apex above the origin maps (r, s, t) to ((1-t) r,
  (1-t) s, t), so a value of 1 at vertex 2 alone gives x y / (1 - z).  A
  vertex order other than VTK's fails these.
* Under any affine map, linear fields are reproduced (to the Newton
  tolerance) and containment equals containment of the preimage in the
  reference cell (away from its boundary).
* A flat hexahedron (det J = 0) contains nothing.
* A mesh of cubes split into hexahedra, prisms, pyramids and tets under an
  affine map (planar faces: the cells tile the domain) with a linear field
  gives f(P) at every interior point and NaN outside; the macrocell ranges
  of a hexahedra-only mesh equal the boxes' rasterization computed
  independently, and cover every cell's value range.
"""
import numpy as np
import pytest

import oracle as orc
from barney_amd import scenes

HEX, PRISM, PYR = 1, 2, 3
CUBE = np.array([(0, 0, 0), (1, 0, 0), (1, 1, 0), (0, 1, 0), (0, 0, 1), (1, 0, 1), (1, 1, 1), (0, 1, 1)],
                np.float64)
PRISM_V = np.array([(0, 0, 0), (1, 0, 0), (0, 1, 0), (0, 0, 1), (1, 0, 1), (0, 1, 1)], np.float64)
PYR_V = np.array([(0, 0, 0), (1, 0, 0), (1, 1, 0), (0, 1, 0), (0, 0, 1)], np.float64)


def _cell(V, f):
    return np.concatenate([V, np.asarray(f, np.float64)[:, None]], 1).astype(np.float32)


def _points(n, lo=-0.25, hi=1.25, seed=3):
    return np.random.default_rng(seed).uniform(lo, hi, (n, 3)).astype(np.float32)


def _away(d, eps=1e-4):
    """points whose signed distances to the boundary all exceed eps"""
    return np.all(np.abs(d) > eps, 1)


def test_hex_trilinear_exact():
    P = _points(20000)
    inside, val = orc.cell_invert(HEX, _cell(CUBE, CUBE[:, 0] * CUBE[:, 1] * CUBE[:, 2]), P)
    d = np.concatenate([P, 1 - P], 1)
    keep = _away(d)
    assert np.array_equal(inside[keep], np.all(d[keep] > 0, 1))
    x = P[inside].astype(np.float64)
    assert np.abs(val[inside] - x[:, 0] * x[:, 1] * x[:, 2]).max() < 2e-6
    assert inside.sum() > 5000


def test_prism_exact():
    P = _points(20000)
    f = PRISM_V[:, 0] * PRISM_V[:, 2] + PRISM_V[:, 1]
    inside, val = orc.cell_invert(PRISM, _cell(PRISM_V, f), P)
    d = np.stack([P[:, 0], P[:, 1], 1 - P[:, 0] - P[:, 1], P[:, 2], 1 - P[:, 2]], 1)
    keep = _away(d)
    assert np.array_equal(inside[keep], np.all(d[keep] > 0, 1))
    x = P[inside].astype(np.float64)
    assert np.abs(val[inside] - (x[:, 0] * x[:, 2] + x[:, 1])).max() < 2e-6
    assert inside.sum() > 2000


def test_pyramid_exact():
    # below z = 0.9: toward the apex det J = (1 - t)^2 falls under the 1e-6
    # tolerance and the reference's inversion gives up (checked below)
    P = _points(40000)
    P = P[P[:, 2] < 0.9]
    f = np.array([0, 0, 1, 0, 0], np.float64)
    inside, val = orc.cell_invert(PYR, _cell(PYR_V, f), P)
    z = P[:, 2].astype(np.float64)
    d = np.stack([P[:, 0], P[:, 1], (1 - z) - P[:, 0], (1 - z) - P[:, 1], z], 1)
    keep = _away(d)
    assert np.array_equal(inside[keep], np.all(d[keep] > 0, 1))
    x = P[inside].astype(np.float64)
    assert np.abs(val[inside] - x[:, 0] * x[:, 1] / (1 - x[:, 2])).max() < 1e-5
    assert inside.sum() > 2000
    # at the apex the Jacobian is singular: no containment
    ins, _ = orc.cell_invert(PYR, _cell(PYR_V, f), np.array([[0.0, 0.0, 0.9999995]], np.float32))
    assert not ins[0]


@pytest.mark.parametrize("cell,V", [(HEX, CUBE), (PRISM, PRISM_V), (PYR, PYR_V)])
def test_affine_linear_reproduction(cell, V):
    rng = np.random.default_rng(11 + cell)
    A = np.eye(3) + 0.4 * rng.uniform(-1, 1, (3, 3))
    b = rng.uniform(-2, 2, 3)
    X = V @ A.T + b
    g, c0 = rng.uniform(-1, 1, 3), 0.3
    P = (_points(30000, -0.25, 1.25) @ A.T + b).astype(np.float32)
    inside, val = orc.cell_invert(cell, _cell(X, X @ g + c0), P)
    # preimage in the reference cell
    R = np.linalg.solve(A, (P.astype(np.float64) - b).T).T
    if cell == HEX:
        d = np.concatenate([R, 1 - R], 1)
    elif cell == PRISM:
        d = np.stack([R[:, 0], R[:, 1], 1 - R[:, 0] - R[:, 1], R[:, 2], 1 - R[:, 2]], 1)
    else:
        d = np.stack([R[:, 0], R[:, 1], (1 - R[:, 2]) - R[:, 0], (1 - R[:, 2]) - R[:, 1], R[:, 2]], 1)
    # away from the boundary, and for pyramids from the apex (singular J)
    keep = _away(d, 1e-3) & ((R[:, 2] < 0.9) if cell == PYR else True)
    assert np.array_equal(inside[keep], np.all(d[keep] > 0, 1)), cell
    ref = P[inside].astype(np.float64) @ g + c0
    assert np.abs(val[inside] - ref).max() < 1e-4 * (1 + np.abs(g).sum() * 3), cell


def test_flat_hex_contains_nothing():
    V = CUBE.copy()
    V[:, 2] = 0.0                                        # all eight corners in z = 0
    inside, _ = orc.cell_invert(HEX, _cell(V, np.ones(8)), _points(2000, 0.0, 1.0) * np.array([1, 1, 0], np.float32))
    assert not inside.any()


def _mesh(n, shear, kinds, g=(0.7, -0.4, 0.25), c0=0.1):
    V, idx, offs, types = scenes.mixed_cells(n, 0xC311, kinds, jitter=0.0, shear=shear)
    f = (V.astype(np.float64) @ np.array(g) + c0).astype(np.float32)
    us = scenes.UMeshVolumeScene(vertices=V, scalars=f, tets=np.zeros((0, 4), np.int32), tf=scenes.ramp_tf(16, 0.8),
                                 tf_domain=(float(f.min()), float(f.max())), base_density=5.0, mc_dims=(5, 6, 7),
                                 cells=(idx, offs, types))
    return us, np.array(g), c0


def test_mesh_point_queries_linear_field():
    """every interior point of a sheared, planar-faced mixed mesh finds a
    cell, with the value f(P); points outside find none"""
    shear = 0.3
    us, g, c0 = _mesh(4, shear, (0, 1, 2, 3))
    types = us.cells[2]
    assert all((types == k).any() for k in range(4))
    s = orc.umesh_scene(us)
    rng = np.random.default_rng(5)
    R = rng.uniform(-0.1, 1.1, (3000, 3))
    # the lattice's preimage coordinates R map to P by the shear
    P = np.stack([R[:, 0] + shear * R[:, 1], R[:, 1] + shear * R[:, 2], R[:, 2] + shear * R[:, 0]], 1)
    inside_dom = np.all((R > 1e-3) & (R < 1 - 1e-3), 1)
    outside_dom = np.any((R < -1e-3) | (R > 1 + 1e-3), 1)
    vals = np.array([s.umesh_sample(p)[0] for p in P.astype(np.float32)])
    assert not np.isnan(vals[inside_dom]).any()
    assert np.isnan(vals[outside_dom]).all()
    ref = P @ g + c0
    assert np.abs(vals[inside_dom] - ref[inside_dom]).max() < 1e-4


def test_hex_mesh_macrocell_ranges():
    """hexahedra rasterize their boxes (rasterBox, MCGrid.cuh:66-93): the
    oracle's ranges equal an independent rasterization of the cell boxes,
    and every cell's value range lies inside the range of each macrocell its
    box reaches"""
    us, _, _ = _mesh(3, 0.2, (1,))
    s = orc.umesh_scene(us)
    dims, lo, hi = s.mc_ranges(0)
    idx, offs, types = us.cells
    V = us.vertices
    cells = np.stack([idx[offs + q] for q in range(8)], 1)
    P = V[cells]                                          # (nc, 8, 3)
    F = us.scalars[cells]
    blo, bhi = P.min(1), P.max(1)
    wlo, whi = blo.min(0), bhi.max(0)
    size = whi - wlo
    spacing = size * (np.float32(1.0) / dims.astype(np.float32))
    rsp = np.float32(1.0) / spacing
    elo = np.full(dims[::-1], np.inf, np.float32)
    ehi = np.full(dims[::-1], -np.inf, np.float32)
    for c in range(len(cells)):
        a = np.clip(((blo[c] - wlo) * rsp).astype(np.int64), 0, dims - 1)
        b = np.clip(((bhi[c] - wlo) * rsp).astype(np.int64), 0, dims - 1)
        sl = (slice(a[2], b[2] + 1), slice(a[1], b[1] + 1), slice(a[0], b[0] + 1))
        elo[sl] = np.minimum(elo[sl], F[c].min())
        ehi[sl] = np.maximum(ehi[sl], F[c].max())
    assert np.array_equal(lo, elo.ravel())
    assert np.array_equal(hi, ehi.ravel())


def test_unknown_cell_type_rejected():
    us, _, _ = _mesh(2, 0.0, (1,))
    idx, offs, types = us.cells
    t = types.copy()
    t[0] = 42                                             # _VTK_POLYHEDRON
    us.cells = (idx, offs, t)
    with pytest.raises(ValueError):
        orc.umesh_scene(us)
