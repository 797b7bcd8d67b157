"""Configs 3 and 4 at their full size on the GPU against the CPU oracle.

Config 3: the 33.5M-triangle terrain with 4M spheres (37.5M primitives, the
GPU builder's largest input, and BVHs deep enough that walks spill the LDS
stack ring into the private stack).  200,000 rays over the whole terrain:
closest hits (t, u, v, primitive and object IDs) bit-exact against the
oracle's traversal of its own build, any-hit occlusion equal; the same with
the smallest LDS ring (8 entries, the most private-stack spills) and the
speculative walk; a 1-spp frame at 320x180 (primary + 4 bounces, HDRI
light): primary IDs bit-exact, pixels within 1e-3, the same rays.

Config 4: the 20.48M-tet volume with its isosurface (element BVH, face
neighbours and seed grid at scale): macrocell ranges and majorants
bit-exact, a 1-spp frame at 48x48: IDs bit-exact, pixels within 1e-3, the
same rays and volume samples.

The oracle builds these scenes in about 30 s (config 3) and 80 s (config 4)
on the host."""
import numpy as np
import pytest

import barney_amd as bx
from barney_amd import abi, scenes
from tests import helpers

pytestmark = pytest.mark.gpu


@pytest.fixture
def keep_tuning(dev):
    saved = dev.tuning()
    yield saved
    dev.set_tuning(**saved)


@pytest.fixture(scope="module")
def c3(dev, orc):
    print("case: full config 3, building", flush=True)
    sc = scenes.config3(width=320, height=180, spp=1)
    world = bx.SurfaceWorld(dev, sc)
    o = orc.surface_scene(sc)
    print("case: full config 3 built", flush=True)
    return sc, world, o


def _terrain_rays(n, seed=31, extent=16.0):
    """rays from above the terrain toward ground points within a few units,
    and a fifth aimed up and sideways (mostly misses, some grazing hits)"""
    rng = np.random.default_rng(seed)
    org = np.stack([rng.uniform(-extent, extent, n), rng.uniform(1.0, 8.0, n), rng.uniform(-extent, extent, n)], 1)
    tgt = org + np.stack([rng.uniform(-4, 4, n), -org[:, 1] - rng.uniform(0, 2, n), rng.uniform(-4, 4, n)], 1)
    up = rng.uniform(size=n) < 0.2
    tgt[up] = org[up] + np.stack([rng.uniform(-1, 1, up.sum()), rng.uniform(0, 0.3, up.sum()),
                                  rng.uniform(-1, 1, up.sum())], 1)
    d = tgt - org
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    rays = np.zeros((n, 8), np.float32)
    rays[:, 0:3], rays[:, 4:7], rays[:, 7] = org, d, 1e30
    return rays


def _same(g, o):
    bad = np.zeros(len(g), bool)
    for f in ("prim_id", "obj_id"):
        bad |= g[f] != o[f]
    for f in ("t", "u", "v"):
        bad |= g[f].view(np.uint32) != o[f].view(np.uint32)
    return np.nonzero(bad)[0]


def test_c3_full_trace(dev, orc, c3, keep_tuning):
    sc, world, o = c3
    rays = _terrain_rays(200000)
    oh = o.trace(rays)[0]
    assert 0.5 < (oh["prim_id"] >= 0).mean() < 0.95
    assert (oh["obj_id"] >= 2).sum() > 1000                 # spheres are hit too
    for t in (dict(), dict(trace_stack_lds=8, trace_walk=1), dict(trace_stack_lds=32)):
        print(f"case: full config 3 trace, tuning {t}", flush=True)
        dev.set_tuning(**keep_tuning)
        dev.set_tuning(**t)
        gh = world.insts.trace(rays)
        bad = _same(gh, oh)
        assert len(bad) == 0, (t, len(bad), gh[bad[:3]], oh[bad[:3]])
        ga = world.insts.trace(rays, any_hit=True)
        assert np.array_equal(ga["prim_id"] >= 0, oh["prim_id"] >= 0), t


def test_c3_full_frame(dev, orc, c3):
    print("case: full config 3 frame, 320x180 1 spp", flush=True)
    sc, world, o = c3
    rend = abi.renderer(bg=sc.bg_color, ambient=sc.ambient, max_generations=sc.max_generations)
    cam = bx.camera_perspective(sc.camera["position"], sc.camera["direction"], sc.camera["up"], sc.camera["fovy"])
    fr = bx.Frame(dev, sc.width, sc.height)
    st = bx.render(world.world, cam, rend, fr, 1)
    ocam = orc.camera_perspective(sc.camera["position"], sc.camera["direction"], sc.camera["up"], sc.camera["fovy"])
    of = orc.Frame(sc.width, sc.height)
    _, ost = orc.render(o, ocam, rend, of, 1)
    for ch in (abi.BNX_FB_PRIMID, abi.BNX_FB_OBJID):
        g, r = fr.read(ch), of.read(ch)
        assert np.array_equal(g, r), (ch, int((g != r).sum()))
    d = np.abs(fr.read(abi.BNX_FB_COLOR) - of.read(abi.BNX_FB_COLOR))
    print("full config 3 1spp max abs diff", float(d.max()), flush=True)
    assert d.max() <= 1e-3
    assert st["rays"] == ost["rays"]
    fr.close()


def test_c4_full(dev, orc):
    print("case: full config 4, building", flush=True)
    us = scenes.config4(width=48, height=48, spp=1)
    w = bx.UMeshWorld(dev, us)
    o = orc.umesh_scene(us)
    print("case: full config 4 built", flush=True)
    gd, glo, ghi = w.volume.mc_ranges()
    od, olo, ohi = o.mc_ranges(0)
    assert tuple(gd) == tuple(od)
    assert np.array_equal(glo.view(np.uint32), olo.view(np.uint32))
    assert np.array_equal(ghi.view(np.uint32), ohi.view(np.uint32))
    _, gm = w.volume.majorants()
    _, om = o.volume_majorants(0)
    assert np.array_equal(gm.view(np.uint32), om.view(np.uint32))
    rend = abi.renderer(bg=us.bg_color, ambient=us.ambient)
    cam = bx.camera_perspective(us.camera["position"], us.camera["direction"], us.camera["up"], us.camera["fovy"])
    fr = bx.Frame(dev, us.width, us.height)
    st = bx.render(w.world, cam, rend, fr, 1)
    ocam = orc.camera_perspective(us.camera["position"], us.camera["direction"], us.camera["up"], us.camera["fovy"])
    of = orc.Frame(us.width, us.height)
    _, ost = orc.render(o, ocam, rend, of, 1)
    for ch in (abi.BNX_FB_PRIMID, abi.BNX_FB_INSTID, abi.BNX_FB_OBJID):
        g, r = fr.read(ch), of.read(ch)
        assert np.array_equal(g, r), (ch, int((g != r).sum()))
    obj = of.read(abi.BNX_FB_OBJID)
    assert (obj == us.iso_user_id).sum() > 50 and (obj == us.user_id).sum() > 50
    d = np.abs(fr.read(abi.BNX_FB_COLOR) - of.read(abi.BNX_FB_COLOR))
    print("full config 4 1spp max abs diff", float(d.max()), flush=True)
    assert d.max() <= 1e-3
    assert st["rays"] == ost["rays"]
    assert st["volume_samples"] == ost["volume_samples"]
    fr.close()
