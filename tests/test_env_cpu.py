"""CPU oracle of the HDRI environment light (light/EnvMap.h, EnvMap.cu,
rtcore/cpu/Texture.cpp, kernels/shadeRays.cu:265-291,435-470,728-741,
866-878), pinned by closed forms, numpy restatements in float64 and
integrals over the sphere, and the shared atan2/acos against libm."""
import ctypes as C
import math

import numpy as np
import pytest

from barney_amd import abi, scenes


# ------------------------------------------------------------------ numpy side
def np_tex(img, u, v):
    """bilinear lookup in float64: u wraps (texel centres at (i + .5) / W),
    v clamps to the first and last rows"""
    H, W = img.shape[:2]
    u = np.asarray(u, np.float64)
    v = np.asarray(v, np.float64)
    tc = np.mod(u * W - 0.5, W)
    x0 = np.floor(tc).astype(np.int64) % W
    fx = tc - np.floor(tc)
    x1 = (x0 + 1) % W
    t = v * H - 0.5
    y0 = np.clip(np.floor(t), 0, H - 1).astype(np.int64)
    fy = np.where(t <= 0, 0.0, np.where(t >= H - 1, 0.0, t - np.floor(t)))
    y1 = np.minimum(y0 + 1, H - 1)
    a = img[y0, x0] * (1 - fx)[:, None] + img[y0, x1] * fx[:, None]
    b = img[y1, x0] * (1 - fx)[:, None] + img[y1, x1] * fx[:, None]
    return a * (1 - fy)[:, None] + b * fy[:, None]


def np_frame(direction, up):
    vz = -np.asarray(up, np.float64) / np.linalg.norm(up)
    vy = np.cross(vz, direction)
    vy = -vy / np.linalg.norm(vy)
    vx = np.cross(vy, vz)
    vx /= np.linalg.norm(vx)
    return np.stack([vx, vy, vz], 1)              # columns


def np_uv(dirs, to_local):
    d = np.asarray(dirs, np.float64)
    d = d / np.linalg.norm(d, axis=1, keepdims=True)
    loc = d @ to_local.T
    th = np.arccos(np.clip(loc[:, 2], -1, 1))
    ph = np.mod(np.arctan2(loc[:, 1], loc[:, 0]), 2 * np.pi)
    return ph / (2 * np.pi), th / np.pi


def sphere_grid(nt, nph):
    th = np.pi * (np.arange(nt) + 0.5) / nt
    ph = 2 * np.pi * (np.arange(nph) + 0.5) / nph
    T, P = np.meshgrid(th, ph, indexing="ij")
    d = np.stack([np.sin(T) * np.cos(P), np.sin(T) * np.sin(P), np.cos(T)], -1).reshape(-1, 3)
    w = (np.sin(T) * (np.pi / nt) * (2 * np.pi / nph)).ravel()
    return d, w


def env_scene(orc, env):
    s = orc.Scene()
    s.set_env_map(env.radiance, env.direction, env.up, env.scale)
    return s


def broad_map(W=128, H=64):
    """smooth map: dim floor plus one broad lobe, so Monte Carlo converges fast"""
    th = np.pi * (np.arange(H) + 0.5) / H
    ph = 2 * np.pi * (np.arange(W) + 0.5) / W
    d = np.stack([np.sin(th)[:, None] * np.cos(ph)[None, :], np.sin(th)[:, None] * np.sin(ph)[None, :],
                  np.broadcast_to(np.cos(th)[:, None], (H, W))], 2)
    ax = np.array([0.3, -0.5, 0.8])
    ax /= np.linalg.norm(ax)
    lobe = 3.0 * np.exp((d @ ax - 1.0) / 0.5 ** 2)
    img = 0.3 + lobe[:, :, None] * np.array([1.0, 0.8, 0.6])[None, None, :]
    return img.astype(np.float32)


# ------------------------------------------------------------------- numerics
def _ulp(ref):
    r32 = np.abs(ref).astype(np.float32)
    return np.maximum(np.spacing(r32).astype(np.float64), np.spacing(np.float32(1e-30)))


def test_atan2_within_three_ulp_of_libm(orc):
    L = orc.lib()
    rng = np.random.default_rng(3)
    ang = rng.uniform(-np.pi, np.pi, 100000)
    rad = np.exp(rng.uniform(-20, 20, 100000))
    y = (rad * np.sin(ang)).astype(np.float32)
    x = (rad * np.cos(ang)).astype(np.float32)
    # axes, diagonals and the tan(pi/12) split of the reduction
    t12 = np.float32(math.tan(math.pi / 12))
    extra = np.array([[0, 1], [0, -1], [1, 0], [-1, 0], [1, 1], [-1, 1], [1, -1], [-1, -1], [t12, 1],
                      [np.nextafter(t12, 2), 1], [np.nextafter(t12, 0), 1], [1, t12], [1e-30, 1], [1, 1e-30],
                      [0, 0], [3, 1e-7], [-3, 1e-7], [1e-7, -3], [-1e-7, -3]], np.float32)
    y = np.concatenate([y, extra[:, 0]])
    x = np.concatenate([x, extra[:, 1]])
    out = np.zeros(len(x), np.float32)
    L.orc_atan2_n(abi.fptr(y), abi.fptr(x), len(x), abi.fptr(out))
    ref = np.arctan2(y.astype(np.float64), x.astype(np.float64))
    err = np.abs(out - ref) / _ulp(ref)
    # y / x rounds once before the reduction: 3 ulp at most, 1 ulp for 95%
    assert err.max() <= 3.0, (err.max(), y[err.argmax()], x[err.argmax()])
    assert np.mean(err <= 1.0) > 0.95
    assert out[len(x) - len(extra) + 14] == 0.0          # atan2(0, 0)


def test_acos_within_three_ulp_of_libm(orc):
    L = orc.lib()
    x = np.concatenate([np.linspace(-1, 1, 200001, dtype=np.float32),
                        np.array([1.0, -1.0, 0.0, -0.0, 1 - 2 ** -24, -1 + 2 ** -24, 0.5, -0.5], np.float32)])
    out = np.zeros(len(x), np.float32)
    L.orc_acos_n(abi.fptr(x), len(x), abi.fptr(out))
    ref = np.arccos(x.astype(np.float64))
    err = np.abs(out - ref) / _ulp(ref)
    # acos(x) = atan2(sqrt((1 - x)(1 + x)), x): the product rounds once, so
    # near |x| = 1 the error is bounded in absolute terms instead
    assert err[np.abs(x) < 0.999].max() <= 3.0
    assert np.abs(out - ref).max() < 1e-6
    assert out[-8] == 0.0 and abs(out[-7] - np.float32(np.pi)) == 0


def test_env_frame_default_and_inverse(orc):
    L = orc.lib()
    tw = np.zeros(9, np.float32)
    tl = np.zeros(9, np.float32)
    L.orc_env_frame(abi.fptr(np.array([1, 0, 0], np.float32)), abi.fptr(np.array([0, 0, 1], np.float32)),
                    abi.fptr(tw), abi.fptr(tl))
    # default HDRI: vz = -up, vy = -normalize(cross(vz, dir)), vx = cross(vy, vz)
    assert np.array_equal(tw.reshape(3, 3).T, [[-1, 0, 0], [0, 1, 0], [0, 0, -1]])
    rng = np.random.default_rng(7)
    for _ in range(20):
        d = rng.standard_normal(3).astype(np.float32)
        u = rng.standard_normal(3).astype(np.float32)
        L.orc_env_frame(abi.fptr(d), abi.fptr(u), abi.fptr(tw), abi.fptr(tl))
        Mw = tw.reshape(3, 3).T.astype(np.float64)
        Ml = tl.reshape(3, 3).T.astype(np.float64)
        assert np.allclose(Mw, np_frame(d.astype(np.float64), u.astype(np.float64)), atol=2e-6)
        assert np.allclose(Ml @ Mw, np.eye(3), atol=2e-6)


# -------------------------------------------------------------------- texture
def test_texture_lookup_matches_bilinear(orc):
    env = scenes.lobe_env_map(64, 32)
    s = env_scene(orc, env)
    img = env.radiance.astype(np.float64)
    H, W = img.shape[:2]
    # texel centres return the texel exactly
    i, j = np.meshgrid(np.arange(W), np.arange(H))
    uv = np.stack([(i.ravel() + 0.5) / W, (j.ravel() + 0.5) / H], 1)
    got = s.env_tex(uv)
    assert np.array_equal(got[:, :3], env.radiance[j.ravel(), i.ravel()])
    assert (got[:, 3] == 0).all()
    # the seam blends the last and first columns; the poles clamp
    got = s.env_tex([[0.0, 0.5 / H], [1.0, 0.5 / H]])
    assert np.allclose(got[0, :3], 0.5 * (img[0, W - 1] + img[0, 0]), rtol=1e-6)
    assert np.allclose(got[1, :3], got[0, :3], rtol=1e-6)
    got = s.env_tex([[0.5 / W, 0.0], [0.5 / W, 1.0]])
    assert np.array_equal(got[:, :3], env.radiance[[0, H - 1], 0])
    # anywhere: float64 bilinear
    rng = np.random.default_rng(2)
    uv = rng.uniform(-0.2, 1.2, (5000, 2)).astype(np.float32)
    got = s.env_tex(uv)[:, :3]
    ref = np_tex(img, uv[:, 0].astype(np.float64), uv[:, 1].astype(np.float64))
    assert np.allclose(got, ref, rtol=1e-4, atol=1e-5)


# ----------------------------------------------------------------------- CDFs
def test_cdfs_rows_only_closed_form(orc):
    """img[j, i] = g(j): each row's CDF is (i + 1) / W and the row weight is
    W * w_j * sin(theta_j) with w_j the largest of g over rows j - 1/2, j,
    j + 1/2 (clamped at the poles)"""
    H, W = 32, 48
    g = 0.5 + np.arange(H) % 5 * 0.7
    img = np.repeat(np.repeat(g[:, None, None], W, 1), 3, 2).astype(np.float32)
    s = env_scene(orc, scenes.EnvMap(img))
    cx, cy, _, _ = s.env_cdfs()
    assert np.allclose(cx, ((np.arange(W) + 1) / W)[None, :], atol=2e-6)
    assert (cx[:, -1] == 1).all() and cy[-1] == 1
    gm = np.concatenate([[g[0]], g, [g[-1]]])
    w = np.maximum(np.maximum(g, 0.5 * (gm[:-2] + g)), 0.5 * (g + gm[2:]))
    rw = W * w * np.sin(np.pi * (np.arange(H) + 0.5) / H)
    assert np.allclose(cy, np.cumsum(rw) / rw.sum(), atol=2e-6)


def test_cdfs_columns_only_closed_form(orc):
    """img[j, i] = h(i): every row has the CDF of w_i = max(h(i), midpoints
    with its wrapped neighbours); the row CDF follows sin(theta)"""
    H, W = 16, 40
    h = 1.0 + (np.arange(W) * 7 % 11).astype(np.float64)
    img = np.repeat(np.repeat(h[None, :, None], H, 0), 3, 2).astype(np.float32)
    s = env_scene(orc, scenes.EnvMap(img))
    cx, cy, _, _ = s.env_cdfs()
    w = np.maximum(np.maximum(h, 0.5 * (np.roll(h, 1) + h)), 0.5 * (h + np.roll(h, -1)))
    assert np.allclose(cx, (np.cumsum(w) / w.sum())[None, :], atol=2e-6)
    st = np.sin(np.pi * (np.arange(H) + 0.5) / H)
    assert np.allclose(cy, np.cumsum(st) / st.sum(), atol=2e-6)


def test_cdfs_are_monotone(orc):
    s = env_scene(orc, scenes.lobe_env_map(256, 128))
    cx, cy, _, _ = s.env_cdfs()
    assert (np.diff(cx, axis=1) >= 0).all() and (np.diff(cy) >= 0).all()
    assert (cx[:, -1] == 1).all() and cy[-1] == 1 and cx.min() > 0


# ------------------------------------------------------------- sample and pdf
def test_samples_land_in_their_texel(orc):
    """the sampled direction maps back (float64 acos/atan2 through the
    inverse frame) to the texel that the CDFs picked, the pdf is
    pdf_x * pdf_y / (2 pi^2 sin theta) there, and the radiance is the
    map's value along that direction"""
    env = scenes.lobe_env_map(128, 64)
    env.direction, env.up, env.scale = (0.3, 1.0, -0.2), (0.1, 0.2, 1.0), 0.5
    s = env_scene(orc, env)
    cx, cy, tw, tl = s.env_cdfs()
    H, W = cx.shape
    d, rad, pdf = s.env_sample(20000, seed=(3, 9))
    assert np.allclose(np.linalg.norm(d, axis=1), 1, atol=1e-6)
    u, v = np_uv(d, np.linalg.inv(tw.astype(np.float64)))
    ix = np.minimum((u * W).astype(int), W - 1)
    iy = np.minimum((v * H).astype(int), H - 1)
    pdf_y = H * (cy[iy] - np.where(iy > 0, cy[np.maximum(iy - 1, 0)], 0))
    pdf_x = W * (cx[iy, ix] - np.where(ix > 0, cx[iy, np.maximum(ix - 1, 0)], 0))
    ref = pdf_x * pdf_y / (2 * np.pi ** 2 * np.sin(np.pi * v))
    ok = np.abs(pdf / ref - 1) < 1e-4
    assert ok.mean() > 0.995, ok.mean()                  # texel-boundary samples aside
    rref = 0.5 * np_tex(env.radiance.astype(np.float64), u, v)
    assert np.allclose(rad, rref, rtol=2e-3, atol=1e-5)


def test_pdf_integrates_to_one(orc):
    s = env_scene(orc, scenes.lobe_env_map(128, 64))
    d, w = sphere_grid(512, 1024)
    p = s.env_pdf(d.astype(np.float32)).astype(np.float64)
    assert abs((p * w).sum() - 1.0) < 2e-3


def test_pdf_without_map_is_uniform(orc):
    s = orc.Scene()
    p = s.env_pdf(np.array([[0, 0, 1], [1, 0, 0], [0.3, -0.2, 0.9]], np.float32))
    assert np.array_equal(p, np.full(3, np.float32(1.0 / (4.0 * np.float32(np.pi))), np.float32))


def test_sampling_is_unbiased(orc):
    """E[L / pdf] over light samples = the integral of L over the sphere"""
    env = scenes.lobe_env_map(128, 64)
    s = env_scene(orc, env)
    d, rad, pdf = s.env_sample(200000, seed=(11, 12))
    est = (rad.astype(np.float64) / pdf[:, None]).mean(0)
    g, w = sphere_grid(512, 1024)
    ref = (s.env_radiance(g.astype(np.float32)).astype(np.float64) * w[:, None]).sum(0)
    assert np.allclose(est, ref, rtol=1e-2), (est, ref)


def test_radiance_along_directions(orc):
    env = scenes.lobe_env_map(128, 64)
    env.direction, env.up, env.scale = (0.0, 1.0, 0.0), (0.0, 0.0, 1.0), 2.0
    s = env_scene(orc, env)
    rng = np.random.default_rng(4)
    d = rng.standard_normal((5000, 3)).astype(np.float32) * 3
    got = s.env_radiance(d)
    _, _, tw, _ = s.env_cdfs()
    u, v = np_uv(d, np.linalg.inv(tw.astype(np.float64)))
    ref = 2.0 * np_tex(env.radiance.astype(np.float64), u, v)
    assert np.allclose(got, ref, rtol=2e-3, atol=1e-5)


# --------------------------------------------------------------------- render
def _sphere_env_scene(orc, env, color=(0.8, 0.8, 0.8)):
    s = orc.Scene()
    g = s.add_spheres(np.zeros((1, 3), np.float32), np.ones(1, np.float32), np.array([color], np.float32),
                      user_id=0)
    s.add_instance(s.add_group(1, [g]), 0)
    s.set_env_map(env.radiance, env.direction, env.up, env.scale)
    s.build()
    return s


def test_primary_miss_shows_the_map(orc):
    env = scenes.lobe_env_map(128, 64)
    env.direction, env.up, env.scale = (0.2, 1.0, 0.0), (0.0, 0.3, 1.0), 0.25
    s = _sphere_env_scene(orc, env)
    cam = orc.camera_perspective((0, 0, 5), (0, 0, -1), (0, 1, 0), 60.0)
    W, H = 48, 32
    fr = orc.Frame(W, H)
    orc.render(s, cam, abi.renderer(bg=(0.1, 0.2, 0.3, 0.5)), fr, 1, threads=4)
    img = fr.read(abi.BNX_FB_COLOR)
    miss = fr.read(abi.BNX_FB_PRIMID) < 0
    assert miss.sum() > 500 and (~miss).sum() > 50
    # accumID 0: rays through pixel centres (generateRays.cu:180-200)
    iy, ix = np.mgrid[0:H, 0:W]
    u = (ix + 0.5) / W
    v = (iy + 0.5) / H
    d = (np.array(cam.dir_00)[None, None, :] + ((W / H) * (u - 0.5))[..., None] * np.array(cam.dir_du)
         + (v - 0.5)[..., None] * np.array(cam.dir_dv))
    _, _, tw, _ = s.env_cdfs()
    uu, vv = np_uv(d.reshape(-1, 3), np.linalg.inv(tw.astype(np.float64)))
    ref = np.minimum(0.25 * np_tex(env.radiance.astype(np.float64), uu, vv), 10.0).reshape(H, W, 3)
    assert np.allclose(img[miss][:, :3], ref[miss], rtol=2e-3, atol=1e-5)
    assert (img[miss][:, 3] == 0.5).all()                # alpha stays the background's


def test_sphere_under_env_map_expectation(orc):
    """Convex sphere (albedo a = 0.85 * 0.8 / pi) lit by the map only: the
    pixel mean with normal N is
      int_{H(N)} L cos^2 * a / pi^2 * p_l / (p_l + cos/pi + 1e-10) dw
    + int_{H(N)} L * a * pi * cos/pi * (cos/pi) / (cos/pi + p_l) dw
    (light sample: f = a/pi cos, geometry term cos/pi, MIS weight from the
    map's pdf p_l; bounce: throughput a*pi, escaping ray sees L with the
    complementary weight), shadeRays.cu:680-744,848-879."""
    env = scenes.EnvMap(broad_map())
    s = _sphere_env_scene(orc, env)
    cam = orc.camera_perspective((0, 0, 3.5), (0, 0, -1), (0, 1, 0), 40.0)
    W = 12
    fr = orc.Frame(W, W)
    orc.render(s, cam, abi.renderer(ambient=0.0), fr, 2048, threads=8)
    img = fr.read(abi.BNX_FB_COLOR)[..., :3].astype(np.float64)
    nrm = fr.read(abi.BNX_FB_NORMAL).astype(np.float64)
    hit = fr.read(abi.BNX_FB_PRIMID) >= 0
    g, w = sphere_grid(256, 512)
    Lg = s.env_radiance(g.astype(np.float32)).astype(np.float64)
    pl = s.env_pdf(g.astype(np.float32)).astype(np.float64)
    a = 0.85 * 0.8 / math.pi
    N = nrm[hit] / np.linalg.norm(nrm[hit], axis=1, keepdims=True)
    cos = np.clip(N @ g.T, 0, None)                      # (pixels, dirs)
    ps = cos / math.pi
    t1 = (cos ** 2 * a / math.pi ** 2 * pl / (pl + ps + 1e-10) * w) @ Lg
    t2 = (a * math.pi * ps * ps / (ps + pl) * w) @ Lg
    exp = t1 + t2
    got = img[hit]
    # silhouette pixels mix hit and miss over the pixel footprint
    cy, cx = np.mgrid[0:W, 0:W]
    inner = ((cx - W / 2 + 0.5) ** 2 + (cy - W / 2 + 0.5) ** 2 < (0.6 * hit.sum() / math.pi))[hit]
    ratio = got[inner].mean(0) / exp[inner].mean(0)
    assert np.allclose(ratio, 1.0, atol=0.01), ratio
    assert np.allclose(got[inner], exp[inner], rtol=0.06)
    # the uniform-pdf weights (no map) would give a visibly different mean
    pu = 1.0 / (4 * math.pi)
    alt = ((cos ** 2 * a / math.pi ** 2 * pu / (pu + ps + 1e-10) * w) @ Lg + (a * math.pi * ps * ps / (ps + pu) * w) @ Lg)
    assert np.abs(alt[inner].mean(0) / exp[inner].mean(0) - 1).max() > 0.03


def test_removing_the_map_restores_ambient(orc):
    env = scenes.lobe_env_map(64, 32)
    s = _sphere_env_scene(orc, env)
    cam = orc.camera_perspective((0, 0, 5), (0, 0, -1), (0, 1, 0), 30.0)
    rend = abi.renderer(bg=(0.1, 0.2, 0.3, 1.0), ambient=1.0)
    fr = orc.Frame(16, 16)
    orc.render(s, cam, rend, fr, 1, threads=2)
    a = fr.read(abi.BNX_FB_COLOR)
    s.set_env_map(None)
    fr2 = orc.Frame(16, 16)
    orc.render(s, cam, rend, fr2, 1, threads=2)
    b = fr2.read(abi.BNX_FB_COLOR)
    miss = fr2.read(abi.BNX_FB_PRIMID) < 0
    assert np.allclose(b[miss][:, :3], [0.1, 0.2, 0.3]) and not np.allclose(a[miss][:, :3], [0.1, 0.2, 0.3])
