"""The denoiser hook (bnx_frame_set_denoiser; rtc::Denoiser, rtcore/cpu/
Denoiser.h:16-30, as fb/FrameBuffer.cu:196-290 drives it) on the GPU.

A Python hook reads its device images with bnx_copy and writes a 3x3 box
filter of the colour.  Checked: the hook receives the frame's linear colour
(bit-identical to the colour channel without a hook) and its normals
(bit-identical to the normal channel), width, height and the blend factor
(0, or (accumID - 1) / (accumID + 100) with fade-out); the colour read
returns exactly what the hook wrote; a hook that writes nothing returns the
input (alpha and all); the 8-bit reads convert the hook's output (min(255,
int(256 x)) of the saturated value, the sRGB curve within one step); a
failing hook fails the read; removing the hook restores the plain read."""
import ctypes as C

import numpy as np
import pytest

import barney_amd as bx
from barney_amd import abi, scenes

pytestmark = pytest.mark.gpu


def _box3(img):
    p = np.pad(img, ((1, 1), (1, 1), (0, 0)), mode="edge").astype(np.float32)
    acc = np.zeros_like(img, dtype=np.float32)
    for dy in range(3):
        for dx in range(3):
            acc += p[dy:dy + img.shape[0], dx:dx + img.shape[1]]
    return acc / np.float32(9.0)


@pytest.fixture(scope="module")
def frame(dev):
    sc = scenes.config1(width=64, height=48)
    w = bx.SurfaceWorld(dev, sc)
    cam = bx.camera_perspective(sc.camera["position"], sc.camera["direction"], sc.camera["up"], sc.camera["fovy"])
    fr = bx.Frame(dev, sc.width, sc.height)
    bx.render(w.world, cam, abi.renderer(bg=sc.bg_color, ambient=sc.ambient), fr, 3)
    yield fr
    fr.close()


def test_denoiser_hook(dev, frame):
    print("denoiser hook", flush=True)
    fr = frame
    H, W = 48, 64
    plain = fr.read(abi.BNX_FB_COLOR)
    normals = fr.read(abi.BNX_FB_NORMAL)
    plain8 = fr.read(abi.BNX_FB_COLOR_RGBA8)
    seen = {}

    def hook(user, d_in, d_nrm, d_out, w, h, blend, stream):
        inp = np.empty((h, w, 4), np.float32)
        nrm = np.empty((h, w, 3), np.float32)
        cur = np.empty((h, w, 4), np.float32)
        bx.check(dev.L.bnx_copy(dev.h, inp.ctypes.data, C.c_void_p(d_in), inp.nbytes))
        bx.check(dev.L.bnx_copy(dev.h, nrm.ctypes.data, C.c_void_p(d_nrm), nrm.nbytes))
        bx.check(dev.L.bnx_copy(dev.h, cur.ctypes.data, C.c_void_p(d_out), cur.nbytes))
        out = _box3(inp)
        out[..., 3] = inp[..., 3]
        bx.check(dev.L.bnx_copy(dev.h, C.c_void_p(d_out), out.ctypes.data, out.nbytes))
        seen.update(inp=inp, nrm=nrm, cur=cur, out=out, w=w, h=h, blend=blend)
        return 0

    fr.set_denoiser(hook, fade_out=False)
    got = fr.read(abi.BNX_FB_COLOR)
    assert (seen["w"], seen["h"]) == (W, H) and seen["blend"] == 0.0
    assert np.array_equal(seen["inp"].view(np.uint32), plain.view(np.uint32))
    assert np.array_equal(seen["cur"].view(np.uint32), plain.view(np.uint32))     # out starts as a copy of in
    assert np.array_equal(seen["nrm"].view(np.uint32), normals.view(np.uint32))
    assert np.array_equal(got.view(np.uint32), seen["out"].view(np.uint32))
    assert not np.array_equal(got, plain)
    # 8-bit reads convert the hook's output
    g8 = fr.read(abi.BNX_FB_COLOR_RGBA8).view(np.uint32)
    c = np.clip(seen["out"], 0.0, 1.0)
    want = np.minimum(255, (c * np.float32(256.0)).astype(np.int64)).astype(np.uint32)
    want8 = want[..., 0] | (want[..., 1] << 8) | (want[..., 2] << 16) | (want[..., 3] << 24)
    assert np.array_equal(g8.reshape(H, W), want8)
    s8 = fr.read(abi.BNX_FB_COLOR_SRGB8).view(np.uint32).reshape(H, W)
    lin = c[..., :3].astype(np.float64)
    srgb = np.where(lin <= 0.0031308, 12.92 * lin, 1.055 * lin ** (1 / 2.4) - 0.055)
    ws = np.minimum(255, (srgb * 256.0).astype(np.int64))
    for k in range(3):
        assert np.abs(((s8 >> (8 * k)) & 255).astype(np.int64) - ws[..., k]).max() <= 1
    # fade-out: blend = (accumID - 1) / (accumID + 100), accumID = 3
    fr.set_denoiser(hook, fade_out=True)
    fr.read(abi.BNX_FB_COLOR)
    assert seen["blend"] == np.float32(2.0) / np.float32(103.0)

    # a hook that writes nothing: the output is the input, alpha included
    def noop(user, d_in, d_nrm, d_out, w, h, blend, stream):
        return 0

    fr.set_denoiser(noop)
    assert np.array_equal(fr.read(abi.BNX_FB_COLOR).view(np.uint32), plain.view(np.uint32))

    def fail(user, d_in, d_nrm, d_out, w, h, blend, stream):
        return 7

    fr.set_denoiser(fail)
    with pytest.raises(bx.BnxError):
        fr.read(abi.BNX_FB_COLOR)
    fr.set_denoiser(None)
    assert np.array_equal(fr.read(abi.BNX_FB_COLOR).view(np.uint32), plain.view(np.uint32))
    assert np.array_equal(fr.read(abi.BNX_FB_COLOR_RGBA8), plain8)
