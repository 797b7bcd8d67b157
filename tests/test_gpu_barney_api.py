"""barney's public bn* API (include/bnx_barney.h, after native/include/
barney.h) on the GPU: tests/cpp/bn_render (built by build()) renders a frame
through the call sequence barney's ANARI device uses -- bnContextCreate,
bnDataCreate, bnGeometryCreate + bnSetData / bnSetString / bnSetObject /
bnCommit, bnMaterialCreate, bnLightCreate, bnGroupCreate + "lights" +
bnGroupBuild, bnModelCreate + bnSetInstances + bnBuild, bnCameraCreate,
bnRendererCreate, bnFrameBufferResize, bnRender, bnFrameBufferRead.

Config 1 (surfaces) and a structured volume: the frames equal, bit for bit,
the same scene rendered through the C-ABI (the Python binding), and match the
oracle (IDs bit-exact, pixels within 1e-3, 8-bit sRGB within one step)."""
import os
import subprocess

import numpy as np
import pytest

import barney_amd as bx
from barney_amd import abi, scenes
from tests import helpers

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "tests", "cpp", "bn_render")


def _set3f_direction(d):
    """the camera direction as set3f("direction") stores it: normalized in
    float32 (native/Camera.cu:68-71) before commit normalizes it again"""
    d = np.asarray(d, np.float32)
    dd = np.float32(np.float32(d[0] * d[0] + d[1] * d[1]) + d[2] * d[2])
    return tuple(float(x) for x in (d * (np.float32(1.0) / np.sqrt(dd))).astype(np.float32))


def _header(f, mode, W, H, spp, cam, bg, ambient, light):
    d, rad, col = light
    for a in (np.array([mode, W, H, spp], np.int32), np.array(cam["position"], np.float32),
              np.array(cam["direction"], np.float32), np.array(cam["up"], np.float32),
              np.array([cam["fovy"]], np.float32), np.array(bg, np.float32), np.array([ambient], np.float32),
              np.array(d, np.float32), np.array([rad], np.float32), np.array(col, np.float32)):
        f.write(np.ascontiguousarray(a).tobytes())


def _run(tmp_path, write, W, H):
    inp, out = tmp_path / "in.bin", tmp_path / "out.bin"
    with open(inp, "wb") as f:
        write(f)
    assert os.access(EXE, os.X_OK), f"{EXE} not built (run __graft_entry__.build())"
    r = subprocess.run([EXE, str(inp), str(out)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0 and "bn_render ok" in r.stdout, (r.returncode, r.stdout, r.stderr)
    raw = np.fromfile(out, dtype=np.uint8)
    n = W * H
    color = raw[:16 * n].view(np.float32).reshape(H, W, 4)
    o = 16 * n
    prim, inst, obj = (raw[o + 4 * n * k:o + 4 * n * (k + 1)].view(np.int32).reshape(H, W) for k in range(3))
    o += 12 * n
    depth = raw[o:o + 4 * n].view(np.float32).reshape(H, W)
    srgb = raw[o + 4 * n:o + 8 * n].reshape(H, W, 4)
    o += 8 * n
    denoised = raw[o:o + 16 * n].view(np.float32).reshape(H, W, 4)
    again = raw[o + 16 * n:o + 32 * n].view(np.float32).reshape(H, W, 4)
    return dict(color=color, prim=prim, inst=inst, obj=obj, depth=depth, srgb=srgb, denoised=denoised,
                again=again)


def _check(bn, fr, of):
    """bn frame == C-ABI frame bit for bit; vs the oracle within tolerance"""
    pairs = [("color", abi.BNX_FB_COLOR), ("prim", abi.BNX_FB_PRIMID), ("inst", abi.BNX_FB_INSTID),
             ("obj", abi.BNX_FB_OBJID), ("depth", abi.BNX_FB_DEPTH), ("srgb", abi.BNX_FB_COLOR_SRGB8)]
    for k, ch in pairs:
        assert np.array_equal(bn[k].view(np.uint8), fr.read(ch).view(np.uint8)), k
    # "enableDenoising" with the context's denoiser (1 - rgb, alpha kept),
    # then off again
    want = bn["color"].copy()
    want[..., :3] = np.float32(1.0) - want[..., :3]
    assert np.array_equal(bn["denoised"].view(np.uint32), want.view(np.uint32))
    assert np.array_equal(bn["again"].view(np.uint32), bn["color"].view(np.uint32))
    for k, ch in pairs[1:4]:
        assert np.array_equal(bn[k], of.read(ch)), k
    d = float(np.abs(bn["color"] - of.read(abi.BNX_FB_COLOR)).max())
    assert d <= 1e-3, d
    assert np.abs(bn["srgb"].astype(int) - of.read(abi.BNX_FB_COLOR_SRGB8).astype(int)).max() <= 1


@pytest.mark.parametrize("spp", [1, 3])
def test_bn_api_config1(dev, orc, tmp_path, spp):
    W, H = 96, 80
    sc = scenes.config1(width=W, height=H)
    mesh, sph = sc.meshes[0], sc.spheres[0]
    light = sc.dir_lights[0]

    def write(f):
        _header(f, 0, W, H, spp, sc.camera, sc.bg_color, sc.ambient, light)
        for a in (np.int32(len(mesh.vertices)), mesh.vertices, np.int32(len(mesh.indices)),
                  mesh.indices.astype(np.int32), mesh.colors, np.int32(len(sph.centers)), sph.centers, sph.radii,
                  sph.colors):
            f.write(np.ascontiguousarray(a).tobytes())
    print(f"bn_render config 1 {W}x{H} {spp} spp", flush=True)
    bn = _run(tmp_path, write, W, H)
    # the same scene through the C-ABI: spheres (instance 0), triangles, both user instance 0
    spec = helpers.c1_spec(sc)
    world = bx.World(dev, spec.gpu(dev), sc.dir_lights)
    cam = bx.camera_perspective(sc.camera["position"], sc.camera["direction"], sc.camera["up"], sc.camera["fovy"])
    rend = abi.renderer(bg=sc.bg_color, ambient=sc.ambient)
    fr = bx.Frame(dev, W, H)
    bx.render(world, cam, rend, fr, spp)
    of = orc.Frame(W, H)
    orc.render(spec.oracle(orc), orc.camera_perspective(sc.camera["position"], sc.camera["direction"],
                                                        sc.camera["up"], sc.camera["fovy"]), rend, of, spp)
    _check(bn, fr, of)
    assert (bn["obj"] == 1).sum() > 50 and (bn["obj"] == 0).sum() > 1000
    fr.close()


def test_bn_api_structured_volume(dev, orc, tmp_path):
    W, H, spp = 64, 64, 2
    vs = scenes.config2(n=33, width=W, height=H, spp=spp, mc_grid=4)
    assert vs.cells_per_mc == 8            # StructuredData's fixed macrocell size (StructuredData.cu:26)
    light = vs.dir_lights[0]

    def write(f):
        _header(f, 1, W, H, spp, vs.camera, vs.bg_color, vs.ambient, light)
        nz, ny, nx = vs.scalars.shape
        for a in (np.array([nx, ny, nz], np.int32), np.array(vs.grid_origin, np.float32),
                  np.array(vs.grid_spacing, np.float32), vs.scalars.astype(np.float32),
                  np.int32(len(vs.tf)), vs.tf.astype(np.float32), np.array(vs.tf_domain, np.float32),
                  np.float32(vs.base_density)):
            f.write(np.ascontiguousarray(a).tobytes())
    print(f"bn_render structured volume {W}x{H} {spp} spp", flush=True)
    bn = _run(tmp_path, write, W, H)
    w = bx.VolumeWorld(dev, vs)
    d = _set3f_direction(vs.camera["direction"])
    cam = bx.camera_perspective(vs.camera["position"], d, vs.camera["up"], vs.camera["fovy"])
    rend = abi.renderer(bg=vs.bg_color, ambient=vs.ambient)
    fr = bx.Frame(dev, W, H)
    bx.render(w.world, cam, rend, fr, spp)
    of = orc.Frame(W, H)
    orc.render(orc.volume_scene(vs), orc.camera_perspective(vs.camera["position"], d, vs.camera["up"],
                                                            vs.camera["fovy"]), rend, of, spp)
    _check(bn, fr, of)
    assert (bn["prim"] == 0).sum() > 200            # volume events on primary rays
