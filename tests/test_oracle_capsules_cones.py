"""The oracle's capsules (CapsulesPrograms::intersect with intersectRoundedCone,
native/geometry/Capsules.dev.cu:19-300) and cones (ConesPrograms::intersect,
native/geometry/Cones.dev.cu:21-266) pinned on the CPU: closed forms of
textbook cases, an independent float64 solution of the rounded cone (two
spheres and the cone tangent to both) and of the capped cone, the
reference's own rules (a ray along a capsule's axis misses, a cone's cap on
the origin's side only), the cone's box holding the cone, and the BVH trace
of a scene of each against a brute-force loop over the primitives."""
import numpy as np
import pytest

import oracle
from barney_amd import scenes


def unit(v):
    v = np.asarray(v, np.float64)
    return v / np.linalg.norm(v)


def cap1(org, d, a, ra, b, rb, tmax=1e30):
    t, n, p = oracle.capsule_intersect([org], [d], a, ra, b, rb, tmax)
    return float(t[0]), n[0], p[0]


def cone1(org, d, a, ra, b, rb, tmin=0.0, tmax=1e30):
    t, n, w = oracle.cone_intersect([org], [d], a, ra, b, rb, tmin, tmax)
    return float(t[0]), n[0], float(w[0])


# ------------------------------------------------------------------ capsules
A, B = (-1.0, 0.0, 0.0), (1.0, 0.0, 0.0)


def test_capsule_body_end_spheres_closed_form():
    # equal radii: a cylinder with hemispherical ends
    t, n, p = cap1((0, 0, 5), (0, 0, -1), A, 0.5, B, 0.5)
    assert t == pytest.approx(4.5, rel=1e-5)
    assert np.allclose(n, (0, 0, 1), atol=1e-5)
    assert np.allclose(p, (0, 0, 0.5), atol=1e-5)
    # obliquely into the first end sphere (centre (-1,0,0), r .5)
    o, d = np.array((-5, 0, 0.3)), unit((4, 0, -0.3))
    oc = o - np.array(A)
    tt = -(oc @ d) - np.sqrt((oc @ d) ** 2 - (oc @ oc - 0.25))
    t, n, _ = cap1(o, d, A, 0.5, B, 0.5)
    assert t == pytest.approx(tt, rel=1e-5)
    assert np.allclose(n, unit(o + tt * d - np.array(A)), atol=1e-5)
    # past the second centre, from above: the end sphere at x = 1.2
    t, n, _ = cap1((1.2, 0, 5), (0, 0, -1), A, 0.5, B, 0.5)
    z = np.sqrt(0.25 - 0.04)
    assert t == pytest.approx(5 - z, rel=1e-5)
    assert np.allclose(n, unit((0.2, 0, z)), atol=1e-5)
    # misses: beside the body, beyond the ends, away, beyond tmax
    assert cap1((0, 0.6, 5), (0, 0, -1), A, 0.5, B, 0.5)[0] < 0
    assert cap1((1.6, 0, 5), (0, 0, -1), A, 0.5, B, 0.5)[0] < 0
    assert cap1((0, 0, 5), (0, 0, 1), A, 0.5, B, 0.5)[0] < 0
    assert cap1((0, 0, 5), (0, 0, -1), A, 0.5, B, 0.5, tmax=4.0)[0] < 0


def test_tapered_capsule_body_is_the_tangent_cone():
    # spheres (0,0,0) r 1 and (4,0,0) r .5: the body is the cone tangent to
    # both, s x + c z = ra with s = (ra - rb) / L, between the tangent points
    ra, rb, L = 1.0, 0.5, 4.0
    s = (ra - rb) / L
    c = np.sqrt(1 - s * s)
    for x0 in (0.5, 1.5, 2.5, 3.5):
        assert ra * s < x0 < L + rb * s
        t, n, p = cap1((x0, 0, 10), (0, 0, -1), (0, 0, 0), ra, (L, 0, 0), rb)
        z = (ra - s * x0) / c
        assert t == pytest.approx(10 - z, rel=2e-5), x0
        assert np.allclose(n, (s, 0, c), atol=2e-5), (x0, n)
    # outside the tangent points the end spheres take over
    t, n, _ = cap1((-0.05, 0, 10), (0, 0, -1), (0, 0, 0), ra, (L, 0, 0), rb)
    assert t == pytest.approx(10 - np.sqrt(1 - 0.05 ** 2), rel=1e-5)


def test_capsule_ray_parallel_to_the_axis_misses():
    # k2 = d2 - m2^2 == 0 returns before the end spheres are tested
    # (Capsules.dev.cu:56-57): a ray exactly along the axis misses
    assert cap1((-5, 0, 0), (1, 0, 0), A, 0.5, B, 0.5)[0] < 0
    assert cap1((-5, 0.2, 0), (1, 0, 0), A, 0.5, B, 0.5)[0] < 0


def test_capsule_one_sphere_inside_the_other_is_empty():
    # |ab| + rb <= ra: intersectRoundedCone returns no hit (Capsules.dev.cu:38-43)
    assert cap1((0, 0, 5), (0, 0, -1), (0, 0, 0), 1.0, (0.2, 0, 0), 0.5)[0] < 0
    assert cap1((0, 0, 5), (0, 0, -1), (0, 0, 0), 0.5, (0.2, 0, 0), 1.0)[0] < 0


def rounded_cone64(o, d, a, ra, b, rb):
    """nearest t > 0 of the union of two spheres and the cone tangent to
    both between its tangent circles, in float64 (None: miss)"""
    o, d, a, b = (np.asarray(x, np.float64) for x in (o, d, a, b))
    best = None

    def sph(c, r):
        oc = o - c
        bq = oc @ d
        disc = bq * bq - (oc @ oc - r * r)
        if disc < 0:
            return None
        for tt in (-bq - np.sqrt(disc), -bq + np.sqrt(disc)):
            if tt > 1e-9:
                return tt
        return None

    for c, r in ((a, ra), (b, rb)):
        tt = sph(c, r)
        if tt is not None and (best is None or tt < best):
            best = tt
    L = np.linalg.norm(b - a)
    u = (b - a) / L
    s = (ra - rb) / L
    cth = np.sqrt(1 - s * s)
    # tangent cone: points P with axial x = (P-a).u and radial rho satisfy
    # s x + cth rho = ra, for x in [ra s, L + rb s]
    oa = o - a
    ox, dx = oa @ u, d @ u
    orad, drad = oa - ox * u, d - dx * u
    # (ra - s x)^2 = cth^2 rho^2, rho^2 = |orad + t drad|^2, x = ox + t dx
    k = ra - s * ox
    qa = cth ** 2 * (drad @ drad) - (s * dx) ** 2
    qb = 2 * (cth ** 2 * (orad @ drad) + s * dx * k)
    qc = cth ** 2 * (orad @ orad) - k * k
    disc = qb * qb - 4 * qa * qc
    if abs(qa) > 1e-12 and disc >= 0:
        for tt in sorted(((-qb - np.sqrt(disc)) / (2 * qa), (-qb + np.sqrt(disc)) / (2 * qa))):
            x = ox + tt * dx
            if tt > 1e-9 and ra * s <= x <= L + rb * s and ra - s * x >= 0:
                if best is None or tt < best:
                    best = tt
                break
    return best


def test_capsules_against_float64():
    rng = np.random.default_rng(11)
    checked = 0
    for _ in range(3000):
        a = rng.uniform(-1, 1, 3)
        b = a + rng.uniform(-1, 1, 3)
        ra, rb = rng.uniform(0.05, 0.4, 2)
        if np.linalg.norm(b - a) < abs(ra - rb) + 0.05:
            continue
        o = rng.uniform(-4, 4, 3)
        # origins outside the capsule (from inside the reference takes the
        # entry root, behind the origin, and reports no hit)
        w = np.clip((o - a) @ (b - a) / ((b - a) @ (b - a)), 0, 1)
        if np.linalg.norm(o - a - w * (b - a)) < max(ra, rb) + 0.05:
            continue
        tgt = a + rng.uniform(-0.2, 1.2) * (b - a) + rng.uniform(-0.5, 0.5, 3)
        d = unit(tgt - o)
        ref = rounded_cone64(o, d, a, ra, b, rb)
        t, n, _ = cap1(o, d, a, ra, b, rb)
        # decide only rays clear of the silhouette: the same ray moved by
        # +-1e-3 across its direction agrees on hit / miss
        near = [rounded_cone64(o + e, d, a, ra, b, rb) for e in 1e-3 * np.eye(3)] + \
               [rounded_cone64(o - e, d, a, ra, b, rb) for e in 1e-3 * np.eye(3)]
        if any((x is None) != (ref is None) for x in near):
            continue
        if ref is None:
            assert t < 0, (o, d, a, b, ra, rb)
        else:
            assert t > 0, (o, d, a, b, ra, rb)
            assert abs(t - ref) <= 2e-4 * (1 + ref), (t, ref)
        checked += 1
    assert checked > 2000


def test_capsule_trace_matches_brute_force():
    cs = scenes.capsule_chains(240, extent=1.5)
    s = oracle.Scene()
    g = s.add_capsules(cs.vertices, cs.indices, cs.colors, user_id=5)
    s.add_instance(s.add_group(1, [g]), 0)
    s.build()
    rays = scenes.random_rays(12, 3000, radius=4.0, spread=1.5)
    hits, _ = s.trace(rays)
    orgs, dirs = rays[:, 0:3], rays[:, 4:7]
    best_t = np.full(len(rays), np.inf)
    best_p = np.full(len(rays), -1)
    for k, (i0, i1) in enumerate(cs.indices):
        va, vb = cs.vertices[i0], cs.vertices[i1]
        t, _, _ = oracle.capsule_intersect(orgs, dirs, va[:3], va[3], vb[:3], vb[3], 1e30)
        better = (t >= 0) & (t < best_t)
        best_t[better], best_p[better] = t[better], k
    hit = best_p >= 0
    assert hit.sum() > 200
    assert np.array_equal(hits["prim_id"] >= 0, hit)
    assert np.array_equal(hits["prim_id"][hit], best_p[hit])
    assert np.array_equal(hits["t"][hit], best_t[hit].astype(np.float32))
    assert np.all(hits["obj_id"][hit] == 5)


# --------------------------------------------------------------------- cones
P0, P1 = (0.0, 0.0, 0.0), (0.0, 0.0, 2.0)


def test_cone_caps_and_body_closed_form():
    # from below through the first cap (m1 < 0): t 5, normal -axis, weight 0
    t, n, w = cone1((0, 0, -5), (0, 0, 1), P0, 1.0, P1, 0.5)
    assert t == pytest.approx(5.0, rel=1e-6) and np.allclose(unit(n), (0, 0, -1)) and w == 0.0
    # from above through the second cap (m2 > 0): t 3, normal +axis, weight 1
    t, n, w = cone1((0, 0, 5), (0, 0, -1), P0, 1.0, P1, 0.5)
    assert t == pytest.approx(3.0, rel=1e-6) and np.allclose(unit(n), (0, 0, 1)) and w == 1.0
    # the side at z = 1, radius 0.75: t 4.25, normal (1, 0, 0.25) / |.|, weight 0.5
    t, n, w = cone1((5, 0, 1), (-1, 0, 0), P0, 1.0, P1, 0.5)
    assert t == pytest.approx(4.25, rel=1e-6)
    assert np.allclose(unit(n), unit((1, 0, 0.25)), atol=1e-6)
    assert w == pytest.approx(0.5, abs=1e-6)
    # misses: beside the cone, past the apex end, away from it, beyond tmax
    assert cone1((5, 0.8, 1), (-1, 0, 0), P0, 1.0, P1, 0.5)[0] < 0
    assert cone1((5, 0, 2.5), (-1, 0, 0), P0, 1.0, P1, 0.5)[0] < 0
    assert cone1((5, 0, 1), (1, 0, 0), P0, 1.0, P1, 0.5)[0] < 0
    assert cone1((5, 0, 1), (-1, 0, 0), P0, 1.0, P1, 0.5, tmax=4.0)[0] < 0


def test_cone_only_the_cap_on_the_origins_side():
    # an origin below the first end's plane (m1 < 0) that misses its cap
    # goes to the body; the second cap is never tested from there
    # (Cones.dev.cu:117-205).  Ray from below, outside the base disc, rising
    # into the side:
    o, d = np.array((2.5, 0, -1.0)), unit((-1.0, 0, 1.0))
    t, n, w = cone1(o, d, P0, 1.0, P1, 0.5)
    # x = 1.5 > 1 where the ray crosses z = 0; then the side r(z) = 1 - z / 4
    # meets x(z) = 1.5 - z at z = 2/3
    z = 2.0 / 3.0
    assert t == pytest.approx((z + 1.0) * np.sqrt(2.0), rel=1e-5)
    assert w == pytest.approx(z / 2, abs=1e-5)
    assert np.allclose(unit(n), unit((1, 0, 0.25)), atol=1e-5)


def capped_cone64(o, d, a, ra, b, rb):
    """nearest t > 0 of the solid capped cone (two discs + lateral surface), float64"""
    o, d, a, b = (np.asarray(x, np.float64) for x in (o, d, a, b))
    L = np.linalg.norm(b - a)
    u = (b - a) / L
    k = (rb - ra) / L
    best = None
    oa = o - a
    ox, dx = oa @ u, d @ u
    orad, drad = oa - ox * u, d - dx * u
    # |orad + t drad|^2 = (ra + k (ox + t dx))^2
    c0 = ra + k * ox
    qa = drad @ drad - (k * dx) ** 2
    qb = 2 * (orad @ drad - k * dx * c0)
    qc = orad @ orad - c0 * c0
    disc = qb * qb - 4 * qa * qc
    if abs(qa) > 1e-12 and disc >= 0:
        for tt in ((-qb - np.sqrt(disc)) / (2 * qa), (-qb + np.sqrt(disc)) / (2 * qa)):
            x = ox + tt * dx
            if tt > 1e-9 and 0 <= x <= L and ra + k * x >= 0 and (best is None or tt < best):
                best = tt
    for x0, r in ((0.0, ra), (L, rb)):
        if abs(dx) > 1e-12:
            tt = (x0 - ox) / dx
            if tt > 1e-9 and np.linalg.norm(orad + tt * drad) <= r and (best is None or tt < best):
                best = tt
    return best


def test_cones_against_float64():
    # rays from outside the cone's bounding sphere, not steeper toward the
    # axis than the side (where the body's root formula stays the entry)
    rng = np.random.default_rng(13)
    checked = 0
    for _ in range(4000):
        a = rng.uniform(-1, 1, 3)
        b = a + unit(rng.normal(size=3)) * rng.uniform(0.5, 1.5)
        ra = rng.uniform(0.1, 0.5)
        rb = ra * rng.uniform(0.0, 1.0)
        ctr, rad = 0.5 * (a + b), 0.5 * np.linalg.norm(b - a) + ra
        o = ctr + unit(rng.normal(size=3)) * rng.uniform(1.5, 3.0) * rad
        tgt = a + rng.uniform(-0.2, 1.2) * (b - a) + rng.uniform(-0.6, 0.6, 3)
        d = unit(tgt - o)
        axis = unit(b - a)
        if abs(d @ axis) > 0.7:
            continue
        ref = capped_cone64(o, d, a, ra, b, rb)
        near = [capped_cone64(o + e, d, a, ra, b, rb) for e in 1e-3 * np.eye(3)] + \
               [capped_cone64(o - e, d, a, ra, b, rb) for e in 1e-3 * np.eye(3)]
        if any((x is None) != (ref is None) for x in near):
            continue
        t, _, _ = cone1(o, d, a, ra, b, rb)
        if ref is None:
            assert t < 0, (o, d, a, b, ra, rb)
        else:
            assert t > 0, (o, d, a, b, ra, rb)
            assert abs(t - ref) <= 2e-4 * (1 + ref), (t, ref)
        checked += 1
    assert checked > 1500


def test_cone_box_holds_the_cone():
    # ConesPrograms::bounds builds the second end's box as (pb - ra, pb + rb)
    # (Cones.dev.cu:38), which need not hold the cone: a cone from (0,0,0)
    # r .1 to (2,0,0) r .6 reaches y = -.6 at its second cap, below that
    # box's y = -.1.  A hit outside its box is then found or not depending on
    # the order the traversal meets boxes in; the oracle (like the GPU) boxes
    # a cone with its end spheres, so the trace finds both of these hits.
    verts = np.array([[0, 0, 0], [2, 0, 0]], np.float32)
    radii = np.array([0.1, 0.6], np.float32)
    s = oracle.Scene()
    g = s.add_cones(verts, radii, None, None, user_id=6)
    s.add_instance(s.add_group(1, [g]), 0)
    s.build()
    ray = np.zeros((2, 8), np.float32)
    ray[0, 0:3], ray[0, 4:7], ray[0, 7] = (5, -0.3, 0), (-1, 0, 0), 1e30
    ray[1, 0:3], ray[1, 4:7], ray[1, 7] = (5, 0.3, 0), (-1, 0, 0), 1e30
    hits, _ = s.trace(ray)
    for k in range(2):
        t, _, _ = cone1(tuple(ray[k, 0:3]), tuple(ray[k, 4:7]), verts[0], 0.1, verts[1], 0.6)
        assert t == pytest.approx(3.0)                # the second cap
        assert hits["prim_id"][k] == 0 and hits["t"][k] == np.float32(t)


def test_cone_trace_matches_brute_force():
    cs = scenes.cone_set(300, extent=1.5)
    s = oracle.Scene()
    g = s.add_cones(cs.vertices, cs.radii, cs.indices, None, user_id=6)
    s.add_instance(s.add_group(1, [g]), 0)
    s.build()
    rays = scenes.random_rays(13, 3000, radius=4.0, spread=1.5)
    hits, _ = s.trace(rays)
    orgs, dirs = rays[:, 0:3], rays[:, 4:7]
    best_t = np.full(len(rays), np.inf)
    best_p = np.full(len(rays), -1)
    for k, (i0, i1) in enumerate(cs.indices):
        t, _, _ = oracle.cone_intersect(orgs, dirs, cs.vertices[i0], cs.radii[i0], cs.vertices[i1], cs.radii[i1])
        better = (t >= 0) & (t < best_t)
        best_t[better], best_p[better] = t[better], k
    hit = best_p >= 0
    assert hit.sum() > 200
    assert np.array_equal(hits["prim_id"] >= 0, hit)
    assert np.array_equal(hits["prim_id"][hit], best_p[hit])
    assert np.array_equal(hits["t"][hit], best_t[hit].astype(np.float32))
    assert np.all(hits["obj_id"][hit] == 6)


def test_implicit_cone_indices():
    cs = scenes.cone_set(40, indexed=False)
    s1, s2 = oracle.Scene(), oracle.Scene()
    g1 = s1.add_cones(cs.vertices, cs.radii, None, cs.colors)
    g2 = s2.add_cones(cs.vertices, cs.radii, np.stack([2 * np.arange(40), 2 * np.arange(40) + 1], 1), cs.colors)
    for s, g in ((s1, g1), (s2, g2)):
        s.add_instance(s.add_group(1, [g]), 0)
        s.build()
    rays = scenes.random_rays(14, 500, radius=4.0, spread=1.5)
    h1, _ = s1.trace(rays)
    h2, _ = s2.trace(rays)
    assert (h1["prim_id"] >= 0).sum() > 20
    assert np.array_equal(h1["prim_id"], h2["prim_id"]) and np.array_equal(h1["t"], h2["t"])
