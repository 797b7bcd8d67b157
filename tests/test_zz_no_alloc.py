"""bnx_render_noalloc through the C-ABI: the frame loop of bnx_render with its
queues in the library's static scratch array (no allocation).  Two calls in a
row, the second on a smaller frame, equal bnx_render of the same requests
bit for bit in every channel; a frame above the array's limit is refused
with an error, and the device keeps working."""
import ctypes as C

import numpy as np
import pytest

import barney_amd as bx
from barney_amd import abi, scenes
from tests import helpers

pytestmark = pytest.mark.gpu

CHANNELS = (abi.BNX_FB_COLOR, abi.BNX_FB_DEPTH, abi.BNX_FB_PRIMID, abi.BNX_FB_INSTID, abi.BNX_FB_OBJID,
            abi.BNX_FB_NORMAL)


def _noalloc(world, cam, rend, fr, spp):
    st = abi.RenderStats()
    rc = world.dev.L.bnx_render_noalloc(world.h, C.byref(cam), C.byref(rend), fr.h_, spp, C.byref(st))
    return rc, st


def _same_as_render(dev, world, cam, rend, w, h, spp):
    """bnx_render_noalloc against bnx_render of the same request: every
    channel bit for bit; with bnx_render's batches as large as the scratch
    array's (2^16 paths) and no captured loop, every statistic too"""
    print(f"noalloc {w}x{h} {spp} spp", flush=True)
    fa = bx.Frame(dev, w, h)
    rc, st = _noalloc(world, cam, rend, fa, spp)
    assert rc == abi.BNX_OK, world.dev.L.bnx_last_error().decode()
    fb = bx.Frame(dev, w, h)
    saved = dev.tuning()
    dev.set_tuning(render_paths=1 << 16, graph_paths=0)
    try:
        ref = bx.render(world, cam, rend, fb, spp)
    finally:
        dev.set_tuning(**saved)
    for ch in CHANNELS:
        assert np.array_equal(fa.read(ch).view(np.uint8), fb.read(ch).view(np.uint8)), (w, h, ch)
    assert st.rays_traced == ref["rays"] and st.generations == ref["generations"]
    assert st.node_visits == ref["nodes"] and st.volume_samples == ref["volume_samples"]
    fa.close()
    fb.close()


def test_noalloc_equals_render_twice(dev):
    sc = scenes.config1(width=200, height=160)
    sw = bx.SurfaceWorld(dev, sc)
    cam = bx.camera_perspective(sc.camera["position"], sc.camera["direction"], sc.camera["up"], sc.camera["fovy"])
    rend = abi.renderer(bg=sc.bg_color, ambient=sc.ambient)
    _same_as_render(dev, sw.world, cam, rend, 200, 160, 5)        # 32000 px: batches of 2 samples
    _same_as_render(dev, sw.world, cam, rend, 72, 40, 3)          # smaller, after the larger one
    # a volume world, after the surfaces
    vs = scenes.config2(n=33, width=64, height=48, spp=2, mc_grid=4)
    vw = bx.VolumeWorld(dev, vs)
    vcam = bx.camera_perspective(vs.camera["position"], vs.camera["direction"], vs.camera["up"], vs.camera["fovy"])
    _same_as_render(dev, vw.world, vcam, abi.renderer(bg=vs.bg_color, ambient=vs.ambient), 64, 48, 2)


def test_noalloc_refuses_large_frames(dev):
    sc = scenes.config1(width=300, height=300)
    sw = bx.SurfaceWorld(dev, sc)
    cam = bx.camera_perspective(sc.camera["position"], sc.camera["direction"], sc.camera["up"], sc.camera["fovy"])
    rend = abi.renderer(bg=sc.bg_color, ambient=sc.ambient)
    fr = bx.Frame(dev, 300, 300)                                  # 90000 px > 65536
    rc, _ = _noalloc(sw.world, cam, rend, fr, 1)
    assert rc == abi.BNX_ERR_UNSUPPORTED
    assert "65536" in dev.L.bnx_last_error().decode()
    # the device still renders, both ways
    small = bx.Frame(dev, 64, 64)
    rc, st = _noalloc(sw.world, cam, rend, small, 1)
    assert rc == abi.BNX_OK and st.rays_traced > 0
    fr.close()
    small.close()
