"""Physically-based (AnariPBR -> NVisii) surfaces on the GPU against the CPU
oracle: the config-1 scene with the icosphere and every other sphere made
glossy.  Gates: primary IDs bit-exact, 1-spp pixels within 1e-3 absolute,
converged RMSE within 1e-3, with the constant environment and with an HDRI
map."""
import ctypes as C

import numpy as np
import pytest

import barney_amd as bx
from barney_amd import abi, scenes
from tests import helpers

pytestmark = pytest.mark.gpu


def glossy_c1(width, height):
    sc = scenes.config1(width=width, height=height)
    sp = sc.spheres[0]
    odd = np.arange(len(sp.radii)) % 2 == 1
    sc.spheres = [scenes.SphereSet(sp.centers[~odd], sp.radii[~odd], sp.colors[~odd], user_id=1),
                  scenes.SphereSet(sp.centers[odd], sp.radii[odd], sp.colors[odd], user_id=2,
                                   material=abi.pbr(metallic=0.6, roughness=0.25))]
    sc.meshes[0].material = abi.pbr(metallic=0.0, roughness=0.4)
    return sc


def _render_both(dev, orc, sc, spp, env=None):
    rend = abi.renderer(bg=sc.bg_color, ambient=sc.ambient)
    spec = helpers.c1_spec(sc)
    world = bx.World(dev, spec.gpu(dev), sc.dir_lights)
    o = spec.oracle(orc)
    if env is not None:
        world.set_env_map(env.radiance, env.direction, env.up, env.scale)
        o.set_env_map(env.radiance, env.direction, env.up, env.scale)
    cam = bx.camera_perspective(sc.camera["position"], sc.camera["direction"], sc.camera["up"], sc.camera["fovy"])
    fr = bx.Frame(dev, sc.width, sc.height)
    st = bx.render(world, cam, rend, fr, spp)
    ocam = orc.camera_perspective(sc.camera["position"], sc.camera["direction"], sc.camera["up"],
                                  sc.camera["fovy"])
    of = orc.Frame(sc.width, sc.height)
    _, ost = orc.render(o, ocam, rend, of, spp)
    out = {ch: (fr.read(ch), of.read(ch)) for ch in (abi.BNX_FB_COLOR, abi.BNX_FB_PRIMID, abi.BNX_FB_INSTID,
                                                       abi.BNX_FB_OBJID)}
    fr.close()
    return out, st, ost


@pytest.mark.parametrize("with_env", [False, True])
def test_pbr_one_spp_parity(dev, orc, with_env):
    sc = glossy_c1(160, 120)
    env = None
    if with_env:
        env = scenes.lobe_env_map(512, 256)
        env.scale = 0.2
    print("case: pbr 1spp 160x120 env", with_env, flush=True)
    out, st, ost = _render_both(dev, orc, sc, 1, env)
    for ch in (abi.BNX_FB_PRIMID, abi.BNX_FB_INSTID, abi.BNX_FB_OBJID):
        g, o = out[ch]
        assert np.array_equal(g, o), ch
    gc, oc = out[abi.BNX_FB_COLOR]
    diff = np.abs(gc - oc)
    print("pbr 1spp max abs diff", float(diff.max()), "exact pixels", float((diff == 0).mean()), flush=True)
    assert diff.max() <= 1e-3
    assert st["rays"] == ost["rays"] and st["generations"] == ost["generations"]


def test_pbr_converged_rmse(dev, orc):
    sc = glossy_c1(64, 64)
    print("case: pbr 32spp 64x64", flush=True)
    out, _, _ = _render_both(dev, orc, sc, 32)
    gc, oc = out[abi.BNX_FB_COLOR]
    rmse = float(np.sqrt(np.mean((gc[..., :3] - oc[..., :3]) ** 2)))
    print("pbr 32spp rmse", rmse, flush=True)
    assert rmse <= 1e-3


def test_pbr_rejects_what_is_not_provided(dev):
    v = np.zeros((3, 3), np.float32)
    v[1, 0] = v[2, 1] = 1
    i = np.array([[0, 1, 2]], np.int32)
    with pytest.raises(bx.BnxError):            # the Glass BSDF
        bx.Geom.triangles(dev, v, i, None, material=abi.pbr(transmission=0.5, ior=1.5))
    # translucent triangles are provided (stochastic opacity, tests/test_gpu_attributes.py)
    bx.Geom.triangles(dev, v, i, None, material=abi.pbr(base_color=(1, 1, 1, 0.5), from_attribute=False))
    # spheres ignore opacity, as the reference's sphere programs do
    bx.Geom.spheres(dev, np.zeros((1, 3), np.float32), np.ones(1, np.float32), None,
                    material=abi.pbr(opacity=0.5))
