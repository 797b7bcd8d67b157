"""The oracle's capped cylinders (CylindersPrograms::intersect,
native/geometry/Cylinders.dev.cu:107-236) pinned on the CPU: textbook closed
forms, an independent float64 solution in the cylinder's own frame, the
reference's own quirks (a ray parallel to the axis misses, :133-134; a back
hit through a cap takes the tube formula's normal, :195), and the BVH trace
of a scene of cylinders against a brute-force loop over every cylinder."""
import numpy as np
import pytest

import oracle
from barney_amd import scenes

V0, V1, R = (0.0, 0.0, -1.0), (0.0, 0.0, 1.0), 0.5


def one(org, d, tmin=0.0, tmax=1e30, v0=V0, v1=V1, r=R):
    t, n = oracle.cylinder_intersect([org], [d], v0, v1, r, tmin, tmax)
    return float(t[0]), n[0]


def unit(v):
    v = np.asarray(v, np.float64)
    return v / np.linalg.norm(v)


def test_perpendicular_tube_hit():
    # distance to the axis 3, radius 0.5: t = 2.5, normal +x
    t, n = one((3, 0, 0), (-1, 0, 0))
    assert t == pytest.approx(2.5, rel=1e-6)
    assert np.allclose(unit(n), (1, 0, 0), atol=1e-6)


def test_tube_hit_off_centre():
    # offset 0.3 from the axis: t = 3 - sqrt(0.25 - 0.09) = 2.6
    t, n = one((3, 0.3, 0.2), (-1, 0, 0))
    assert t == pytest.approx(2.6, rel=1e-6)
    assert np.allclose(unit(n), unit((0.4, 0.3, 0)), atol=1e-5)


def test_cap_hit_oblique():
    # enters through the v1 cap (z = 1) at x = 0.2: t = 2 |d| / |d.z|
    d = unit((0.05, 0, -1))
    t, n = one((0.1, 0, 3), d)
    assert t == pytest.approx(2.0 * np.sqrt(1.0025), rel=1e-5)
    assert np.allclose(unit(n), (0, 0, 1), atol=1e-6)
    # the same from below enters through the v0 cap: normal -z
    t, n = one((0.1, 0, -3), unit((0.05, 0, 1)))
    assert t == pytest.approx(2.0 * np.sqrt(1.0025), rel=1e-5)
    assert np.allclose(unit(n), (0, 0, -1), atol=1e-6)


def test_misses():
    assert one((3, 0.6, 0), (-1, 0, 0))[0] < 0          # passes 0.6 from the axis
    assert one((3, 0, 1.5), (-1, 0, 0))[0] < 0          # above the v1 cap
    assert one((3, 0, -1.5), (-1, 0, 0))[0] < 0         # below the v0 cap
    assert one((3, 0, 0), (1, 0, 0))[0] < 0             # points away
    assert one((3, 0, 0), (-1, 0, 0), tmax=2.0)[0] < 0  # beyond tmax


def test_parallel_to_axis_misses():
    # (s x d)^2 == 0 returns before anything else (Cylinders.dev.cu:133-134),
    # even for a ray straight down the axis through both caps
    assert one((0.1, 0, 3), (0, 0, -1))[0] < 0
    assert one((0, 0, 0), (0, 0, 1))[0] < 0


def test_inside_gives_back_hit():
    # from the axis the exit is one radius away, normal outward
    t, n = one((0, 0, 0), (1, 0, 0))
    assert t == pytest.approx(0.5, rel=1e-6)
    assert np.allclose(unit(n), (1, 0, 0), atol=1e-6)
    # tmin past the entry: the exit counts (Cylinders.dev.cu:189-199)
    t, _ = one((3, 0, 0), (-1, 0, 0), tmin=3.0)
    assert t == pytest.approx(3.5, rel=1e-6)


def back_cap_normal(o, d, v0, v1, t_hit, tube_t1):
    """The normal a back hit through a cap gets: the reference tests
    t0 == cap_t1 (Cylinders.dev.cu:195), which fails there, and so takes the
    tube formula td*d - f' - u*s, whose first two terms are the tube's exit
    point (relative to v0) while u is the cap hit's"""
    o, d, v0, v1 = (np.asarray(x, np.float64) for x in (o, d, v0, v1))
    s = v1 - v0
    u = (o + t_hit * d - v0) @ s / (s @ s)
    return (o + tube_t1 * d - v0) - u * s


def test_back_cap_hit_keeps_tube_formula():
    # from inside, leaving through the v1 cap at z = 1 (t = |d| / d.z); the
    # tube would be left at x = 0.5, z = 3.5 (t = 4.5 |d| / d.z)
    d = unit((0.1, 0, 1))
    t, n = one((0.05, 0, 0), d)
    L = np.sqrt(1.01)
    assert t == pytest.approx(L, rel=1e-5)
    want = back_cap_normal((0.05, 0, 0), d, V0, V1, L, 4.5 * L)
    assert np.allclose(want, (0.5, 0, 3.5), atol=1e-9)
    assert unit(n) @ unit(want) > 0.99999


def ref64(o, d, v0, v1, r, tmin):
    """float64 closest hit of the solid capped cylinder in its own frame:
    (t, outward normal, kind) with kind 0 tube, 1 cap, or None."""
    o, d, v0, v1 = (np.asarray(x, np.float64) for x in (o, d, v0, v1))
    h = np.linalg.norm(v1 - v0)
    a = (v1 - v0) / h
    p = o - v0
    z, dz = p @ a, d @ a
    pp, dp = p - z * a, d - dz * a
    cands = []
    A, B, Cc = dp @ dp, 2 * (pp @ dp), pp @ pp - r * r
    disc = B * B - 4 * A * Cc
    tube_t1 = None
    if A > 0 and disc >= 0:
        tube_t1 = (-B + np.sqrt(disc)) / (2 * A)
        for s in (-1, 1):
            t = (-B + s * np.sqrt(disc)) / (2 * A)
            if 0 <= z + t * dz <= h:
                q = pp + t * dp
                cands.append((t, q / np.linalg.norm(q), 0, abs(z + t * dz) / h, abs(h - z - t * dz) / h))
    if dz != 0:
        for zc, nrm in ((0.0, -a), (h, a)):
            t = (zc - z) / dz
            q = pp + t * dp
            if q @ q <= r * r:
                cands.append((t, nrm, 1, abs(np.sqrt(q @ q) - r) / r, 1.0))
    if not cands:
        return None
    cands.sort(key=lambda c: c[0])
    entry, exit_ = cands[0], cands[-1]
    if entry[0] >= tmin:
        return entry, True, tube_t1
    if exit_[0] >= tmin:
        return exit_, False, tube_t1
    return None


def test_random_rays_against_float64():
    rng = np.random.default_rng(7)
    checked = 0
    for _ in range(3000):
        v0 = rng.uniform(-1, 1, 3)
        v1 = v0 + rng.uniform(-1, 1, 3)
        r = rng.uniform(0.05, 0.5)
        inside = rng.random() < 0.2
        if inside:
            o = v0 + rng.uniform(0.2, 0.8) * (v1 - v0)
        else:
            o = rng.uniform(-4, 4, 3)
        tgt = v0 + rng.uniform(-0.1, 1.1) * (v1 - v0) + rng.uniform(-r, r, 3) * 1.3
        d = unit(tgt - o)
        ref = ref64(o, d, v0, v1, r, 0.0)
        t, n = one(o, d, 0.0, 1e30, v0, v1, r)
        if ref is None:
            # grazing misses may be found by float32: only check clear misses
            continue
        (tr, nr, kind, edge0, edge1), front, tube_t1 = ref
        if min(edge0, edge1) < 1e-3:
            continue                                  # too close to a rim to be decided
        assert t >= 0, (o, d, v0, v1, r)
        scale = np.linalg.norm(o - v0) + np.linalg.norm(v1 - v0) + 1.0
        assert abs(t - tr) <= 2e-5 * scale, (t, tr)
        if front or kind == 0:
            assert unit(n) @ nr > 0.9999, (n, nr, kind, front)
        else:
            assert unit(n) @ unit(back_cap_normal(o, d, v0, v1, tr, tube_t1)) > 0.999
        checked += 1
    assert checked > 1500


def test_trace_matches_brute_force():
    cs = scenes.sticks(300, extent=1.5)
    s = oracle.Scene()
    g = s.add_cylinders(cs.vertices, cs.indices, cs.radii, cs.colors, user_id=4)
    s.add_instance(s.add_group(1, [g]), 0)
    s.build()
    rays = scenes.random_rays(11, 3000, radius=4.0, spread=1.5)
    hits, _ = s.trace(rays)
    orgs, dirs = rays[:, 0:3], rays[:, 4:7]
    best_t = np.full(len(rays), np.inf)
    best_p = np.full(len(rays), -1)
    for k, (i0, i1) in enumerate(cs.indices):
        t, _ = oracle.cylinder_intersect(orgs, dirs, cs.vertices[i0], cs.vertices[i1], cs.radii[k], 0.0, 1e30)
        better = (t >= 0) & (t < best_t)
        best_t[better], best_p[better] = t[better], k
    hit = best_p >= 0
    assert hit.sum() > 200
    assert np.array_equal(hits["prim_id"] >= 0, hit)
    assert np.array_equal(hits["prim_id"][hit], best_p[hit])
    assert np.allclose(hits["t"][hit], best_t[hit], rtol=1e-6)
    assert np.all(hits["obj_id"][hit] == 4)


def test_shared_end_points_and_default_radius():
    cs = scenes.sticks(50)
    chained = np.nonzero(cs.indices[1:, 0] == cs.indices[:-1, 1])[0]
    assert len(chained) > 3
    s = oracle.Scene()
    g = s.add_cylinders(cs.vertices, cs.indices, None, None, user_id=4, default_radius=0.07)
    s.add_instance(s.add_group(1, [g]), 0)
    s.build()
    # a ray at the centre of cylinder k, perpendicular to it, hits it at
    # distance - default radius unless another one is closer
    k = 10
    a, b = cs.vertices[cs.indices[k, 0]], cs.vertices[cs.indices[k, 1]]
    mid = 0.5 * (a + b)
    ax = unit(b - a)
    perp = unit(np.cross(ax, (0.3, 0.7, 0.2)))
    org = mid + 0.2 * perp
    ray = np.zeros((1, 8), np.float32)
    ray[0, 0:3], ray[0, 4:7], ray[0, 7] = org, -perp, 1e30
    hits, _ = s.trace(ray)
    assert hits["prim_id"][0] == k
    assert hits["t"][0] == pytest.approx(0.2 - 0.07, rel=1e-4)
