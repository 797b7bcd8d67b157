"""CPU oracle of the physically-based material (AnariPBR -> NVisii Disney
BSDF, material/AnariPBR.h:62-96, packedBSDF/NVisii.h), pinned by a float64
numpy restatement of the BSDF at random directions, closed forms at normal
incidence, the distributions of the lobe samplers, and a closed-form render
of a PBR sphere under one directional light."""
import math

import numpy as np
import pytest

from barney_amd import abi, scenes


def f16(x):
    return np.asarray(x, np.float32).astype(np.float16).astype(np.float64)


def unit(v):
    v = np.asarray(v, np.float64)
    return v / np.linalg.norm(v, axis=-1, keepdims=True)


def sw(c):
    return np.clip(1.0 - c, 0.0, 1.0) ** 5


def np_nvisii(base, metallic, roughness, ng, wo, wi):
    """float64 Disney BSDF of NVisii.h (isotropic, no sheen / clearcoat /
    transmission / flatness, specular 0.5, specular tint 0): value and pdf"""
    b = f16(base)
    m = float(f16(metallic))
    r = float(f16(min(max(roughness, 0.001), 0.9)))
    n, wo, wi = (np.atleast_2d(np.asarray(x, np.float64)) for x in (ng, wo, wi))
    wh = unit(wi + wo)
    ndo = np.abs((wo * n).sum(1))
    ndi = np.abs((wi * n).sum(1))
    ih = (wi * wh).sum(1)
    fi, fo = sw(ndi), sw(ndo)
    fd90 = 0.5 + r * ih * ih
    diffuse = (1 + (fd90 - 1) * fi) * (1 + (fd90 - 1) * fo)
    spec = (1 - m) * 0.04 + m * b
    a = max(0.002, r * r)
    nh = (n * wh).sum(1)
    d = (1 / math.pi) / np.maximum(1 + (a * a - 1) * nh * nh, 2e-10)
    fw = sw(np.abs(ih)) * ((1 - max(m, a)) * 0.5 + max(m, a))
    F = spec[None, :] + (1 - spec[None, :]) * fw[:, None]

    def smith(x):
        return 1 / (x + np.sqrt(a * a + x * x - a * a * x * x))

    G = smith(ndi) * smith(ndo)
    value = diffuse[:, None] * b[None, :] * (1 - m) + (d * G)[:, None] * F
    dw, gw = 1 - m, 1 + m + 0.5 + r
    dn = (wi * n).sum(1)
    dpdf = np.where(dn > 0, dn / math.pi, 0.0)
    same = (wo * n).sum(1) * dn > 0
    mf = np.where(same, d * np.abs(nh) / (4 * np.abs((wo * wh).sum(1))), 0.0)
    return value, dw * dpdf + gw * mf


def rand_dirs(rng, n):
    return unit(rng.standard_normal((n, 3)))


MATS = [((0.7, 0.5, 0.3), 0.0, 0.5), ((0.9, 0.8, 0.2), 1.0, 0.2), ((0.2, 0.6, 0.9), 0.35, 0.05),
        ((0.5, 0.5, 0.5), 0.6, 0.95), ((0.3, 0.1, 0.8), 0.1, 0.0)]


@pytest.mark.parametrize("mat", MATS)
def test_eval_matches_float64_restatement(orc, mat):
    rng = np.random.default_rng(hash(mat) % 2 ** 32)
    n = rand_dirs(rng, 3000)
    wo = rand_dirs(rng, 3000)
    wi = rand_dirs(rng, 3000)
    # keep away from the wi = -wo singularity of the half vector
    ok = (wo * wi).sum(1) > -0.95
    n, wo, wi = n[ok], wo[ok], wi[ok]
    val, pdf = orc.pbr_eval(mat[0], mat[1], mat[2], n, wo, wi)
    rv, rp = np_nvisii(mat[0], mat[1], mat[2], n, wo, wi)
    assert np.allclose(val, rv, rtol=2e-4, atol=1e-6), np.abs(val - rv).max()
    assert np.allclose(pdf, rp, rtol=2e-4, atol=1e-6)
    assert np.array_equal(orc.pbr_pdf(mat[0], mat[1], mat[2], n, wo, wi), pdf)


def test_normal_incidence_closed_form(orc):
    """wo = wi = n: half vector n, Schlick weights 0, Smith terms 1/2 each,
    D = 1 / (pi alpha^2): value = b (1 - m) + spec / (4 pi alpha^2),
    pdf = (1 - m) / pi + (1.5 + m + r) / (4 pi alpha^2)"""
    for base, m, r in [((0.5, 0.25, 0.75), 0.25, 0.5), ((1.0, 1.0, 1.0), 0.0, 0.9), ((0.1, 0.2, 0.3), 1.0, 0.3)]:
        n = np.array([[0.0, 0.0, 1.0]])
        val, pdf = orc.pbr_eval(base, m, r, n, n, n)
        b, mm, rr = f16(base), float(f16(m)), float(f16(r))
        a = max(0.002, rr * rr)
        spec = (1 - mm) * 0.04 + mm * b
        assert np.allclose(val[0], b * (1 - mm) + spec / (4 * math.pi * a * a), rtol=1e-5)
        assert np.isclose(pdf[0], (1 - mm) / math.pi + (1.5 + mm + rr) / (4 * math.pi * a * a), rtol=1e-5)


def test_roughness_clamp_and_half_packing(orc):
    n = np.array([[0.0, 0.0, 1.0]])
    # roughness 0 is clamped to 0.001 and 1 to 0.9 (AnariPBR.h:86), both stored as halves
    for r, rc in [(0.0, 0.001), (1.0, 0.9), (0.3333, 0.3333)]:
        _, p = orc.pbr_eval((0.5, 0.5, 0.5), 0.0, r, n, n, n)
        a = np.float32(max(0.002, float(np.float32(f16(rc)) * np.float32(f16(rc)))))
        # 1 + (alpha^2 - 1) cancels in float32 for tiny alpha: evaluate it as the BSDF does
        den = np.float32(1) + (a * a - np.float32(1))
        exp = 1 / math.pi + (1.5 + float(f16(rc))) * (1 / math.pi) / float(den) / 4
        assert np.isclose(p[0], exp, rtol=1e-5), (r, p[0], exp)


def test_lobe_selection_and_diffuse_distribution(orc):
    """diffuse lobe chosen with probability (1 - m) / (2.5 + r); its
    directions are cosine distributed about Ng (cos^2 uniform on [0, 1])"""
    base, m, r = (0.6, 0.6, 0.6), 0.25, 0.5
    ng = np.array([0.0, 0.0, 1.0])
    wo = unit([0.3, -0.2, 0.9])
    d, pdf, fr, diff = orc.pbr_scatter(base, m, r, ng, wo, 40000, seed=(5, 6))
    mm, rr = float(f16(m)), float(f16(r))
    pd = (1 - mm) / ((1 - mm) + (1 + mm + 0.5 + rr))
    assert abs(diff.mean() - pd) < 4 * math.sqrt(pd * (1 - pd) / len(diff))
    c = d[diff == 1] @ ng
    assert (c > 0).all()
    c2 = np.sort(c.astype(np.float64) ** 2)
    ks = np.abs(c2 - (np.arange(1, len(c2) + 1) / len(c2))).max()
    assert ks < 1.63 / math.sqrt(len(c2))                 # 1% KS level
    # every valid sample's pdf and f_r are the eval at its direction
    ok = pdf > 0
    v, p = orc.pbr_eval(base, m, r, np.broadcast_to(ng, (ok.sum(), 3)), np.broadcast_to(wo, (ok.sum(), 3)), d[ok])
    assert np.allclose(p, pdf[ok], rtol=1e-5) and np.allclose(v, fr[ok], rtol=1e-5, atol=1e-7)


def test_glossy_half_vector_distribution(orc):
    """metallic 1: only the glossy lobe.  The sampler draws cos^2(theta_h) =
    (1 - u) / (1 + (alpha^2 - 1) u) and builds the half vector from
    (s cos phi, s sin phi, c) with s = 1 - c^2 before normalising
    (NVisii.h:180-188); the reflected direction's half vector follows that law"""
    base, m, r = (0.9, 0.9, 0.9), 1.0, 0.6
    ng = np.array([0.0, 0.0, 1.0])
    wo = unit([0.0, 1e-4, 1.0])          # (almost) along Ng; exactly along it has no tangent frame
    d, pdf, _, diff = orc.pbr_scatter(base, m, r, ng, wo, 40000, seed=(7, 8))
    assert (diff == 0).all()
    ok = pdf > 0
    h = unit(d[ok].astype(np.float64) + wo[None, :])
    ch = h @ ng
    a = max(0.002, float(f16(r)) ** 2)
    # invert the normalisation: ch = c / sqrt(s^2 + c^2), s = 1 - c^2
    cgrid = np.linspace(0, 1, 200001)
    chgrid = cgrid / np.sqrt((1 - cgrid ** 2) ** 2 + cgrid ** 2)

    def law(c):                           # P(cos^2 theta_h <= c^2)
        x = c * c
        return 1 - (1 - x) / (1 + x * (a * a - 1))

    # with wo along Ng the reflection stays above the surface iff the half
    # vector is within 45 degrees of Ng: the kept samples follow the law
    # truncated there
    c0 = np.interp(1 / math.sqrt(2), chgrid, cgrid)
    f0 = law(c0)
    assert abs((~ok).mean() - f0) < 4 * math.sqrt(f0 * (1 - f0) / len(ok))
    c = np.sort(np.interp(ch, chgrid, cgrid))
    emp = np.arange(1, len(c) + 1) / len(c)
    ks = np.abs((law(c) - f0) / (1 - f0) - emp).max()
    assert ks < 2.0 / math.sqrt(len(c)), ks


def test_pbr_sphere_under_directional_light(orc):
    """one directional light, no ambient: a sphere pixel is
    min(f(wo, l), 13) * radiance * |N.l| / pi (the shadow ray), with f the
    NVisii value; bounces escape into a black environment.  When the glossy
    sample reflects below the surface the scatter is invalid and the ray is
    traced again unchanged (shadeRays.cu:771-775), so it reaches the same
    point and adds the shadow ray once more: pixels hold whole multiples."""
    base, m, r = (0.7, 0.5, 0.3), 0.3, 0.4
    L = unit([0.3, 0.5, 0.8])
    s = orc.Scene()
    g = s.add_spheres(np.zeros((1, 3), np.float32), np.ones(1, np.float32), None, user_id=0,
                      material=abi.pbr(base_color=base + (1.0,), from_attribute=False, metallic=m, roughness=r))
    s.add_instance(s.add_group(1, [g]), 0)
    s.set_dir_lights([(-L.astype(np.float32), 2.0, (1.0, 1.0, 1.0))])
    s.build()
    cam = orc.camera_perspective((0, 0, 5), (0, 0, -1), (0, 1, 0), 30.0)
    W = 48
    fr = orc.Frame(W, W)
    orc.render(s, cam, abi.renderer(bg=(0.0, 0.0, 0.0, 1.0), ambient=0.0), fr, 1, threads=4)
    img = fr.read(abi.BNX_FB_COLOR)[..., :3].astype(np.float64)
    nrm = fr.read(abi.BNX_FB_NORMAL).astype(np.float64)
    hit = fr.read(abi.BNX_FB_PRIMID) >= 0
    iy, ix = np.mgrid[0:W, 0:W]
    dirs = (np.array(cam.dir_00)[None, None, :] + (((ix + 0.5) / W) - 0.5)[..., None] * np.array(cam.dir_du)
            + (((iy + 0.5) / W) - 0.5)[..., None] * np.array(cam.dir_dv))
    N = unit(nrm[hit])
    wo = -unit(dirs[hit])
    cosl = N @ L
    lit = cosl > 0.05
    f, _ = np_nvisii(base, m, r, N[lit], wo[lit], np.broadcast_to(L, (lit.sum(), 3)))
    exp = np.minimum(f, 13.0) * 2.0 * cosl[lit][:, None] / math.pi
    got = img[hit][lit]
    assert lit.sum() > 300
    k = np.round(got[:, 0] / exp[:, 0])
    assert (k >= 1).all() and (k == 1).mean() > 0.5
    assert np.allclose(got, k[:, None] * exp, rtol=3e-3, atol=1e-5)
    assert np.allclose(img[hit][cosl < -0.05], 0.0)


@pytest.mark.parametrize("mat", MATS)
def test_eval_reciprocity(orc, mat):
    """Helmholtz reciprocity, independent of any transcription of the
    formulas: the Disney diffuse term, the GGX distribution, the Smith
    masking product and the Schlick term on the half vector are each
    symmetric in (wo, wi), so f(wo, wi) = f(wi, wo) for directions on the
    same side of the surface; a term that takes the wrong angle breaks it"""
    rng = np.random.default_rng(31)
    n = np.tile([[0.0, 0.0, 1.0]], (4000, 1))
    wo = rand_dirs(rng, 4000)
    wi = rand_dirs(rng, 4000)
    wo[:, 2], wi[:, 2] = np.abs(wo[:, 2]) + 0.05, np.abs(wi[:, 2]) + 0.05
    wo, wi = unit(wo), unit(wi)
    a, _ = orc.pbr_eval(mat[0], mat[1], mat[2], n, wo, wi)
    b, _ = orc.pbr_eval(mat[0], mat[1], mat[2], n, wi, wo)
    assert np.allclose(a, b, rtol=3e-5, atol=1e-7), np.abs(a - b).max()
