"""Cameras and 8-bit colour output of the oracle (CPU only).

* omnidirectional camera (native/Camera.cu:230-242, generateRays.cu:24-38,
  140-146): pixel (u, v) looks along toWorld (cos 2 pi u sin pi v,
  sin 2 pi u sin pi v, cos pi v); the centre pixel along the camera
  direction, the bottom row down, the top row up.  The reference's own
  fixture (unitTests/reference_unitTest_omnicamera.png) is reproduced: hit
  mask, silhouettes, per-sphere mean colour.
* orthographic camera (Camera.cu:168-181, generateRays.cu:130-139):
  parallel rays from the image plane, closed form.
* depth of field (generateRays.cu:101-126): sample 0 is the pinhole ray
  (renormalized);
  later samples start on the lens disk and pass through the pinhole ray's
  point on the focal plane.
* BN_UFIXED8_RGBA(_SRGB) (fb/FrameBuffer.cu:111-130, common/math.h:61-125):
  closed-form values of the sRGB curve and of min(255, int(256 x))."""
import math

import numpy as np

from barney_amd import abi, scenes
from tests import helpers


def test_omni_camera_closed_form(orc):
    cam = orc.camera_omni((1.0, 2.0, 3.0), (0.0, 0.0, 1.0), (0.0, 1.0, 0.0))
    W, H = 64, 32
    rays = orc.primary_rays(cam, W, H, 0)
    M = np.array(cam.dir_00[:] + cam.dir_du[:] + cam.dir_dv[:], np.float64).reshape(3, 3).T   # to_world columns
    assert np.allclose(M.T @ M, np.eye(3), atol=1e-6)
    iy, ix = np.divmod(np.arange(W * H), W)
    u, v = (ix + 0.5) / W, (iy + 0.5) / H
    loc = np.stack([np.cos(2 * math.pi * u) * np.sin(math.pi * v), np.sin(2 * math.pi * u) * np.sin(math.pi * v),
                    np.cos(math.pi * v)], 1)
    assert np.allclose(rays[:, 4:7], loc @ M.T, atol=2e-6)
    assert np.allclose(rays[:, 0:3], (1.0, 2.0, 3.0))
    d = rays[:, 4:7].reshape(H, W, 3)
    # centre column / middle row looks along the direction; v -> 0 looks down
    c = d[H // 2, W // 2]
    assert np.dot(c, (0.0, 0.0, 1.0)) > 0.99
    assert np.dot(d[0, 5], (0.0, -1.0, 0.0)) > 0.99 and np.dot(d[H - 1, 5], (0.0, 1.0, 0.0)) > 0.99


def test_orthographic_camera_closed_form(orc):
    pos, dirn, up = (0.5, -1.0, 4.0), (0.0, 0.0, -2.0), (0.0, 1.0, 0.0)
    cam = orc.camera_orthographic(pos, dirn, up, 3.0, 1.5)
    W, H = 40, 30
    rays = orc.primary_rays(cam, W, H, 0)
    assert np.allclose(rays[:, 4:7], (0.0, 0.0, -1.0), atol=0)
    iy, ix = np.divmod(np.arange(W * H), W)
    u, v = (ix + 0.5) / W, (iy + 0.5) / H
    # du = normalize(cross(dir, up)) = +x, dv = normalize(cross(du, dir)) = +y
    exp = np.stack([0.5 + (u - 0.5) * 1.5 * 3.0, -1.0 + (v - 0.5) * 3.0, np.full_like(u, 4.0)], 1)
    assert np.allclose(rays[:, 0:3], exp, atol=1e-6)


def test_depth_of_field_focus(orc):
    pos, look, up = (0.0, 0.0, 5.0), (0.0, 0.0, -1.0), (0.0, 1.0, 0.0)
    pin = orc.camera_perspective(pos, look, up, 40.0)
    dof = orc.camera_perspective(pos, look, up, 40.0)
    dof.aperture_radius, dof.focus_distance = 0.25, 4.0
    W, H = 48, 40
    # sample 0: no lens offset; the direction is the pinhole one normalized again
    r0p, r0d = orc.primary_rays(pin, W, H, 0), orc.primary_rays(dof, W, H, 0)
    assert np.array_equal(r0p[:, 0:3], r0d[:, 0:3]) and np.allclose(r0p[:, 4:7], r0d[:, 4:7], atol=2e-7)
    offs = []
    for accum in (1, 2, 3, 7):
        p, d = orc.primary_rays(pin, W, H, accum), orc.primary_rays(dof, W, H, accum)
        off = d[:, 0:3].astype(np.float64) - np.asarray(pos)
        assert np.allclose(off[:, 2], 0.0, atol=1e-7)                   # on the lens plane
        assert (np.linalg.norm(off, axis=1) <= 0.25 * (1 + 1e-6)).all()  # on the lens disk
        offs.append(off)
        # the pinhole ray's point on the focal plane z = 5 - 4
        D = p[:, 4:7].astype(np.float64)
        F = np.asarray(pos) + D * (4.0 / np.abs(D[:, 2:3]))
        t = (F[:, 2] - d[:, 2]) / d[:, 6]
        Q = d[:, 0:3] + t[:, None] * d[:, 4:7]
        assert np.allclose(Q, F, atol=2e-5)
    offs = np.concatenate(offs)
    assert np.abs(offs[:, :2].mean(0)).max() < 0.01                     # uniform on the disk
    assert abs((offs[:, :2] ** 2).sum(1).mean() - 0.25 ** 2 / 2) < 0.002


def test_srgb8_closed_form(orc):
    """one pixel per value: a frame whose colour is set through a render of
    a constant background, read as RGBA8 and as sRGB8"""
    s = orc.Scene()
    s.build()
    for val in (0.0, 0.001, 0.0031308, 0.02, 0.214, 0.5, 0.999, 1.0, 2.0):
        fr = orc.Frame(4, 2)
        orc.render(s, orc.camera_perspective((0, 0, 5), (0, 0, -1), (0, 1, 0), 40.0),
                   abi.renderer(bg=(val, val * 0.5, 0.25, 1.0)), fr, 1)
        c = fr.read(abi.BNX_FB_COLOR)[0, 0]
        rgba = fr.read(abi.BNX_FB_COLOR_RGBA8)[0, 0].astype(int)
        srgb = fr.read(abi.BNX_FB_COLOR_SRGB8)[0, 0].astype(int)

        def enc(x, curve):
            x = min(max(float(x), 0.0), 1.0)
            if curve:
                x = 12.92 * x if x <= 0.0031308 else 1.055 * x ** (1 / 2.4) - 0.055
            return min(255, max(0, int(math.floor(x * 256.0))))
        exp_lin = [enc(c[k], False) for k in range(4)]
        exp_srgb = [enc(c[k], k < 3) for k in range(3)] + [enc(c[3], False)]
        assert list(rgba) == exp_lin, (val, rgba, exp_lin)
        assert np.all(np.abs(srgb - np.array(exp_srgb)) <= (0 if val in (0.0, 1.0, 2.0) else 1)), (val, srgb)
    # a few exact values of the curve
    assert enc(0.0, True) == 0 and enc(1.0, True) == 255 and enc(0.5, True) == 188 and enc(0.2158605, True) == 128


def test_omni_fixture_reproduced(orc):
    spec = helpers.omni_spec()
    s = spec.oracle(orc)
    pos, d, u = helpers.omni_camera_args()
    cam = orc.camera_omni(pos, d, u)
    W, H = helpers.OMNI_SIZE
    fr = orc.Frame(W, H)
    orc.render(s, cam, abi.renderer(bg=(0.0, 0.0, 0.0, 1.0), ambient=0.8), fr, 1)
    helpers.check_against_omni_fixture(fr.read(abi.BNX_FB_COLOR_SRGB8), fr.read(abi.BNX_FB_OBJID))
