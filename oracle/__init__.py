"""ctypes binding of the CPU oracle (oracle/_build/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, always as the checker and never as the path
being measured.  The product package (barney_amd) never imports it.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

from barney_amd import abi

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "_build", "liboracle.so")
_lib = None


def build(force: bool = False) -> str:
    """Compile the oracle with its Makefile (g++)."""
    if force and os.path.exists(LIB_PATH):
        os.remove(LIB_PATH)
    subprocess.run(["make", "-C", _HERE, "-s"], check=True)
    return LIB_PATH


def use_host_native() -> str:
    """Compile the oracle for the instruction set of the machine this runs on
    (g++ -march=native, with the same -ffp-contract=off arithmetic, so the
    same results) into a fresh temporary folder and load that library instead
    of the portable x86-64-v2 one.  For timing the CPU baseline on the host
    that measures; call before anything loads the library.  Returns the
    -march used: "native", or "x86-64-v2" when the native build fails (the
    portable library is then kept)."""
    global LIB_PATH
    if _lib is not None:
        raise RuntimeError("oracle library already loaded")
    import tempfile
    out = os.path.join(tempfile.mkdtemp(prefix="oracle_native_"), "liboracle.so")
    flags = "-O3 -march=native -std=c++17 -fPIC -ffp-contract=off -fno-fast-math -Wall -Wno-unused-function"
    r = subprocess.run(["make", "-C", _HERE, "-s", f"OUT={out}", f"CXXFLAGS={flags}"], capture_output=True)
    if r.returncode == 0 and os.path.exists(out):
        LIB_PATH = out
        return "native"
    if not os.path.exists(LIB_PATH):
        build()
    return "x86-64-v2"


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"oracle library missing: {LIB_PATH} (run `make -C oracle`)")
        L = C.CDLL(LIB_PATH)
        vp = C.c_void_p
        L.orc_scene_create.restype = vp
        L.orc_scene_destroy.argtypes = [vp]
        L.orc_scene_add_triangles.argtypes = [vp, C.POINTER(abi.TrianglesDesc)]
        L.orc_scene_add_spheres.argtypes = [vp, C.POINTER(abi.SpheresDesc)]
        L.orc_scene_add_cylinders.argtypes = [vp, C.POINTER(abi.CylindersDesc)]
        L.orc_cylinder_intersect_n.argtypes = [abi.fp, C.c_int, abi.fp, C.c_float, C.c_float, abi.fp]
        L.orc_volume_ratio_n.argtypes = [vp, C.c_int, vp, C.c_int, C.c_uint32, C.c_uint32, abi.fp, vp, C.c_int]
        L.orc_volume_ratio_n.restype = None
        L.orc_scene_add_capsules.argtypes = [vp, C.POINTER(abi.CapsulesDesc)]
        L.orc_scene_add_cones.argtypes = [vp, C.POINTER(abi.ConesDesc)]
        L.orc_capsule_intersect_n.argtypes = [abi.fp, C.c_int, abi.fp, C.c_float, abi.fp]
        L.orc_cone_intersect_n.argtypes = [abi.fp, C.c_int, abi.fp, C.c_float, C.c_float, abi.fp]
        L.orc_scene_add_group.argtypes = [vp, C.c_int, C.POINTER(C.c_int), C.c_int]
        L.orc_geom_set_attribute.argtypes = [vp, C.c_int, C.c_int, C.c_int, C.c_int, abi.fp, C.c_int64, abi.fp]
        L.orc_geom_set_vertex_normals.argtypes = [vp, C.c_int, abi.fp, C.c_int]
        L.orc_scene_set_instance_attribute.argtypes = [vp, C.c_int, abi.fp, C.c_int]
        L.orc_trace_opacity.argtypes = [vp, vp, C.POINTER(C.c_uint64), C.c_int, vp, vp, C.c_int]
        L.orc_scene_add_instance.argtypes = [vp, C.c_int, C.c_int, abi.fp, abi.fp]
        L.orc_scene_set_dir_lights.argtypes = [vp, C.POINTER(abi.DirLight), C.c_int]
        L.orc_scene_set_env_map.argtypes = [vp, C.POINTER(abi.EnvMapDesc)]
        L.orc_env_cdfs.argtypes = [vp, abi.fp, abi.fp, abi.fp, abi.fp]
        L.orc_env_tex.argtypes = [vp, abi.fp, C.c_int, abi.fp]
        L.orc_env_pdf.argtypes = [vp, abi.fp, C.c_int, abi.fp]
        L.orc_env_radiance.argtypes = [vp, abi.fp, C.c_int, abi.fp]
        L.orc_env_sample.argtypes = [vp, C.c_uint32, C.c_uint32, C.c_int, abi.fp, abi.fp, abi.fp]
        L.orc_env_uv_to_world.argtypes = [vp, abi.fp, C.c_int, abi.fp]
        L.orc_atan2_n.argtypes = [abi.fp, abi.fp, C.c_int, abi.fp]
        L.orc_acos_n.argtypes = [abi.fp, C.c_int, abi.fp]
        L.orc_env_frame.argtypes = [abi.fp, abi.fp, abi.fp, abi.fp]
        L.orc_pbr_eval.argtypes = [abi.fp, abi.fp, C.c_int, abi.fp, abi.fp]
        L.orc_pbr_pdf.argtypes = [abi.fp, abi.fp, C.c_int, abi.fp]
        L.orc_pbr_scatter.argtypes = [abi.fp, abi.fp, abi.fp, C.c_uint32, C.c_uint32, C.c_int, abi.fp, abi.fp,
                                      abi.fp, abi.ip]
        L.orc_scene_add_volume.argtypes = [vp, C.POINTER(abi.StructuredDesc), C.POINTER(abi.TransferFunction),
                                           C.c_int, C.c_int]
        L.orc_volume_majorants.argtypes = [vp, C.c_int, abi.ip, abi.fp]
        L.orc_scene_add_umesh.argtypes = [vp, C.POINTER(abi.UMeshDesc), C.POINTER(abi.TransferFunction), C.c_int,
                                          C.c_int, C.c_int]
        L.orc_scene_add_isosurface.argtypes = [vp, C.c_int, C.c_float, abi.fp, C.c_int, C.c_int]
        L.orc_volume_mc_ranges.argtypes = [vp, C.c_int, abi.ip, abi.fp, abi.fp]
        L.orc_volume_ebvh.argtypes = [vp, C.c_int, vp, abi.ip, abi.ip]
        L.orc_cell_invert_n.argtypes = [C.c_int, abi.fp, C.c_int, abi.fp, abi.ip, abi.fp]
        L.orc_umesh_sample.argtypes = [vp, C.c_int, abi.fp, C.POINTER(C.c_uint64)]
        L.orc_umesh_sample.restype = C.c_float
        L.orc_iso_march.argtypes = [vp, C.c_int, abi.fp, abi.fp, C.c_float, C.c_uint32, C.c_uint32, abi.fp]
        L.orc_iso_march.restype = C.c_float
        L.orc_volume_sample.argtypes = [vp, C.c_int, abi.fp]
        L.orc_volume_sample.restype = C.c_float
        L.orc_volume_tf_map.argtypes = [vp, C.c_int, C.c_float, abi.fp]
        L.orc_dda_cells.argtypes = [abi.fp, abi.fp, C.c_float, abi.ip, C.c_int, abi.ip, abi.fp]
        L.orc_volume_march_n.argtypes = [vp, C.c_int, vp, C.c_int, C.c_uint32, C.c_uint32, abi.fp, C.c_int]
        L.orc_camera_orthographic.argtypes = [abi.fp, abi.fp, abi.fp, C.c_float, C.c_float, vp]
        L.orc_camera_omni.argtypes = [abi.fp, abi.fp, abi.fp, vp]
        L.orc_primary_rays.argtypes = [vp, C.c_int, C.c_int, C.c_int, vp]
        L.orc_fast_log.argtypes = [C.c_float]
        L.orc_fast_log.restype = C.c_float
        L.orc_volume_march.argtypes = [vp, C.c_int, abi.fp, abi.fp, C.c_float, C.c_uint32, C.c_uint32, abi.fp,
                                       C.POINTER(C.c_uint64)]
        L.orc_volume_march.restype = C.c_float
        L.orc_scene_build.argtypes = [vp]
        L.orc_scene_build.restype = C.c_double
        L.orc_trace.argtypes = [vp, vp, C.c_int, vp, C.c_int, C.POINTER(C.c_uint64), C.c_int, abi.fp]
        L.orc_bvh_build_boxes.argtypes = [abi.fp, C.c_int, C.c_int, vp, abi.ip, abi.ip]
        L.orc_scene_bvh.argtypes = [vp, C.c_int, vp, abi.ip, abi.ip]
        L.orc_camera_perspective.argtypes = [abi.fp, abi.fp, abi.fp, C.c_float, C.POINTER(abi.Camera)]
        L.orc_frame_create.restype = vp
        L.orc_frame_create.argtypes = [C.c_int, C.c_int]
        L.orc_frame_destroy.argtypes = [vp]
        L.orc_frame_reset.argtypes = [vp]
        L.orc_render.argtypes = [vp, C.POINTER(abi.Camera), C.POINTER(abi.Renderer), vp, C.c_int, C.c_int,
                                 C.POINTER(C.c_uint64), C.c_int]
        L.orc_render.restype = C.c_double
        L.orc_frame_read.argtypes = [vp, C.c_int, vp]
        L.orc_sincos.argtypes = [C.c_float, abi.fp, abi.fp]
        L.orc_f32_to_f16.argtypes = [C.c_float]
        L.orc_f32_to_f16.restype = C.c_uint16
        L.orc_f16_to_f32.argtypes = [C.c_uint16]
        L.orc_f16_to_f32.restype = C.c_float
        L.orc_rng_seq.argtypes = [C.c_uint32, C.c_uint32, C.c_int, C.POINTER(C.c_uint32)]
        L.orc_random2_seq.argtypes = [C.c_uint32, C.c_uint32, C.c_uint32, C.c_int, abi.fp, C.POINTER(C.c_uint64)]
        _lib = L
    return _lib


def _mat6(base, metallic, roughness, ior):
    return np.array([base[0], base[1], base[2], metallic, roughness, ior], np.float32)


def pbr_eval(base, metallic, roughness, ng, wo, wi, ior=1.45):
    """NVisii eval of an AnariPBR material for rows of (Ng, wo, wi)"""
    m = _mat6(base, metallic, roughness, ior)
    v = np.ascontiguousarray(np.concatenate([np.atleast_2d(ng), np.atleast_2d(wo), np.atleast_2d(wi)], 1),
                             np.float32)
    val = np.zeros((len(v), 3), np.float32)
    pdf = np.zeros(len(v), np.float32)
    lib().orc_pbr_eval(abi.fptr(m), abi.fptr(v), len(v), abi.fptr(val), abi.fptr(pdf))
    return val, pdf


def pbr_pdf(base, metallic, roughness, ng, wo, wi, ior=1.45):
    m = _mat6(base, metallic, roughness, ior)
    v = np.ascontiguousarray(np.concatenate([np.atleast_2d(ng), np.atleast_2d(wo), np.atleast_2d(wi)], 1),
                             np.float32)
    pdf = np.zeros(len(v), np.float32)
    lib().orc_pbr_pdf(abi.fptr(m), abi.fptr(v), len(v), abi.fptr(pdf))
    return pdf


def pbr_scatter(base, metallic, roughness, ng, wo, n, seed=(1, 2), ior=1.45):
    """n scatter samples: (dirs, pdf, f_r, diffuse flag)"""
    m = _mat6(base, metallic, roughness, ior)
    d = np.zeros((n, 3), np.float32)
    p = np.zeros(n, np.float32)
    f = np.zeros((n, 3), np.float32)
    fl = np.zeros(n, np.int32)
    lib().orc_pbr_scatter(abi.fptr(m), abi.fptr(np.asarray(ng, np.float32)), abi.fptr(np.asarray(wo, np.float32)),
                          seed[0], seed[1], n, abi.fptr(d), abi.fptr(p), abi.fptr(f), abi.iptr(fl))
    return d, p, f, fl


def default_threads() -> int:
    return max(1, min(16, os.cpu_count() or 1))


class Scene:
    """Oracle scene: geometries -> groups -> instances, as in rtc."""

    def __init__(self):
        self.L = lib()
        self.h = self.L.orc_scene_create()
        self._keep = []

    def __del__(self):
        try:
            self.L.orc_scene_destroy(self.h)
        except Exception:
            pass

    def add_triangles(self, vertices, indices, colors=None, user_id=0, material=None):
        v = np.ascontiguousarray(vertices, dtype=np.float32)
        i = np.ascontiguousarray(indices, dtype=np.int32)
        c = None if colors is None else np.ascontiguousarray(colors, dtype=np.float32)
        self._keep += [v, i, c]
        d = abi.TrianglesDesc(abi.fptr(v), len(v), len(i), abi.iptr(i), abi.fptr(c), user_id, 0,
                              material or abi.matte())
        return self.L.orc_scene_add_triangles(self.h, C.byref(d))

    def add_spheres(self, centers, radii, colors=None, user_id=1, material=None, default_radius=0.1):
        ce = np.ascontiguousarray(centers, dtype=np.float32)
        r = None if radii is None else np.ascontiguousarray(radii, dtype=np.float32)
        c = None if colors is None else np.ascontiguousarray(colors, dtype=np.float32)
        self._keep += [ce, r, c]
        d = abi.SpheresDesc(abi.fptr(ce), abi.fptr(r), abi.fptr(c), len(ce), default_radius, user_id, 0,
                            material or abi.matte())
        return self.L.orc_scene_add_spheres(self.h, C.byref(d))

    def add_cylinders(self, vertices, indices, radii=None, colors=None, user_id=1, material=None,
                      default_radius=0.1):
        v = np.ascontiguousarray(vertices, dtype=np.float32)
        i = np.ascontiguousarray(indices, dtype=np.int32)
        r = None if radii is None else np.ascontiguousarray(radii, dtype=np.float32)
        c = None if colors is None else np.ascontiguousarray(colors, dtype=np.float32)
        self._keep += [v, i, r, c]
        d = abi.CylindersDesc(abi.fptr(v), abi.iptr(i), abi.fptr(r), abi.fptr(c), len(v), len(i), default_radius,
                              user_id, material or abi.matte())
        return self.L.orc_scene_add_cylinders(self.h, C.byref(d))

    def add_capsules(self, vertices, indices, colors=None, user_id=1, material=None):
        v = np.ascontiguousarray(vertices, dtype=np.float32)
        i = np.ascontiguousarray(indices, dtype=np.int32)
        c = None if colors is None else np.ascontiguousarray(colors, dtype=np.float32)
        self._keep += [v, i, c]
        d = abi.CapsulesDesc(abi.fptr(v), abi.iptr(i), abi.fptr(c), len(v), len(i), user_id, 0,
                             material or abi.matte())
        return self.L.orc_scene_add_capsules(self.h, C.byref(d))

    def add_cones(self, vertices, radii, indices=None, colors=None, user_id=1, material=None, num_cones=None):
        v = np.ascontiguousarray(vertices, dtype=np.float32)
        r = np.ascontiguousarray(radii, dtype=np.float32)
        i = None if indices is None else np.ascontiguousarray(indices, dtype=np.int32)
        c = None if colors is None else np.ascontiguousarray(colors, dtype=np.float32)
        self._keep += [v, r, i, c]
        n = len(i) if i is not None else (len(v) // 2 if num_cones is None else num_cones)
        d = abi.ConesDesc(abi.fptr(v), abi.fptr(r), abi.iptr(i), abi.fptr(c), len(v), n, user_id, 0,
                          material or abi.matte())
        return self.L.orc_scene_add_cones(self.h, C.byref(d))

    def set_attribute(self, geom: int, slot: int, scope: int, data=None, value=None):
        """a geometry attribute (bnx_geom_set_attribute): data (n, comps) floats, or a constant rgba"""
        d = None if data is None else np.ascontiguousarray(data, dtype=np.float32)
        comps = 0 if d is None else (1 if d.ndim == 1 else d.shape[1])
        v = None if value is None else np.ascontiguousarray(value, dtype=np.float32)
        self._keep += [d, v]
        assert self.L.orc_geom_set_attribute(self.h, geom, slot, scope, comps, abi.fptr(d),
                                             0 if d is None else len(d), abi.fptr(v)) == 0

    def set_vertex_normals(self, geom: int, normals):
        n = None if normals is None else np.ascontiguousarray(normals, dtype=np.float32)
        self._keep += [n]
        assert self.L.orc_geom_set_vertex_normals(self.h, geom, abi.fptr(n), 0 if n is None else len(n)) == 0

    def set_instance_attribute(self, slot: int, values):
        v = None if values is None else np.ascontiguousarray(values, dtype=np.float32)
        self._keep += [v]
        assert self.L.orc_scene_set_instance_attribute(self.h, slot, abi.fptr(v), 0 if v is None else len(v)) == 0

    def trace_opacity(self, rays, seeds, mode=0):
        """closest hits with stochastic opacity (ray i's RNG state seeds[i]) and the
        closest candidates the opacity test rejected; mode 0 BVH walk, 1 brute force"""
        r = np.ascontiguousarray(rays)
        s = np.ascontiguousarray(seeds, dtype=np.uint64)
        hits = np.zeros(len(r), abi.HIT_DTYPE)
        rej = np.zeros(len(r), abi.HIT_DTYPE)
        self.L.orc_trace_opacity(self.h, r.ctypes.data, s.ctypes.data_as(C.POINTER(C.c_uint64)), len(r),
                                 hits.ctypes.data, rej.ctypes.data, mode)
        return hits, rej

    def add_group(self, kind: int, geoms):
        g = (C.c_int * len(geoms))(*geoms)
        return self.L.orc_scene_add_group(self.h, kind, g, len(geoms))

    def add_instance(self, group: int, user_id: int = 0, o2w=None, w2o=None):
        a = None if o2w is None else np.ascontiguousarray(o2w, dtype=np.float32)
        b = None if w2o is None else np.ascontiguousarray(w2o, dtype=np.float32)
        self._keep += [a, b]
        return self.L.orc_scene_add_instance(self.h, group, user_id, abi.fptr(a), abi.fptr(b))

    def add_volume(self, vs, user_id=0, user_inst=0):
        d, t, keep = abi.structured_desc(vs)
        k = self.L.orc_scene_add_volume(self.h, C.byref(d), C.byref(t), user_id, user_inst)
        return k

    def add_umesh(self, us, user_id=0, user_inst=0, rendered=True):
        d, t, keep = abi.umesh_desc(us)
        k = self.L.orc_scene_add_umesh(self.h, C.byref(d), C.byref(t), user_id, user_inst, int(rendered))
        if k < 0:
            raise ValueError("oracle: unsupported cell type")
        return k

    def add_isosurface(self, field: int, iso: float, color=(0.8, 0.8, 0.8), user_id=0, user_inst=0):
        c = np.ascontiguousarray(color[:3], np.float32)
        return self.L.orc_scene_add_isosurface(self.h, field, float(iso), abi.fptr(c), user_id, user_inst)

    def mc_ranges(self, k=0):
        dims = np.zeros(3, np.int32)
        self.L.orc_volume_mc_ranges(self.h, k, abi.iptr(dims), None, None)
        lo = np.zeros(int(np.prod(dims)), np.float32)
        hi = np.zeros_like(lo)
        self.L.orc_volume_mc_ranges(self.h, k, abi.iptr(dims), abi.fptr(lo), abi.fptr(hi))
        return dims, lo, hi

    def element_bvh(self, k=0):
        info = np.zeros(3, np.int32)
        self.L.orc_volume_ebvh(self.h, k, None, None, abi.iptr(info))
        nodes = np.zeros(info[0], abi.NODE_DTYPE)
        order = np.zeros(info[1], np.int32)
        self.L.orc_volume_ebvh(self.h, k, nodes.ctypes.data, abi.iptr(order), abi.iptr(info))
        return nodes, order, int(info[2])

    def umesh_sample(self, P, k=0):
        p = np.ascontiguousarray(P, np.float32)
        n = C.c_uint64()
        f = self.L.orc_umesh_sample(self.h, k, abi.fptr(p), C.byref(n))
        return f, n.value

    def iso_march(self, org, dir, tmax, seed=(1, 2), k=0):
        o = np.ascontiguousarray(org, np.float32)
        d = np.ascontiguousarray(dir, np.float32)
        N = np.zeros(3, np.float32)
        t = self.L.orc_iso_march(self.h, k, abi.fptr(o), abi.fptr(d), tmax, seed[0], seed[1], abi.fptr(N))
        return t, N

    def volume_majorants(self, k=0):
        dims = np.zeros(3, np.int32)
        self.L.orc_volume_majorants(self.h, k, abi.iptr(dims), None)
        maj = np.zeros(int(np.prod(dims)), np.float32)
        self.L.orc_volume_majorants(self.h, k, abi.iptr(dims), abi.fptr(maj))
        return dims, maj

    def volume_sample(self, P, k=0):
        p = np.ascontiguousarray(P, np.float32)
        return self.L.orc_volume_sample(self.h, k, abi.fptr(p))

    def tf_map(self, s, k=0):
        out = np.zeros(4, np.float32)
        self.L.orc_volume_tf_map(self.h, k, float(s), abi.fptr(out))
        return out

    def volume_march(self, org, dir, tmax, seed=(1, 2), k=0):
        o = np.ascontiguousarray(org, np.float32)
        d = np.ascontiguousarray(dir, np.float32)
        rgb = np.zeros(3, np.float32)
        ns = C.c_uint64()
        t = self.L.orc_volume_march(self.h, k, abi.fptr(o), abi.fptr(d), tmax, seed[0], seed[1], abi.fptr(rgb),
                                    C.byref(ns))
        return t, rgb, ns.value

    def volume_march_n(self, rays, seed=(0, 7), k=0, threads=None):
        """event t (or -1) of each ray through volume k; ray i marches with
        the parent seed (seed[0] + i, seed[1])"""
        rays = np.ascontiguousarray(rays, np.float32).reshape(-1, 8)
        out = np.zeros(len(rays), np.float32)
        self.L.orc_volume_march_n(self.h, k, rays.ctypes.data, len(rays), seed[0], seed[1], abi.fptr(out),
                                  threads or default_threads())
        return out

    def volume_ratio_n(self, rays, seed=(0, 7), k=0, threads=None):
        """ratio-tracking transmittance estimate of each ray through volume
        k (ray i with the parent seed (seed[0] + i, seed[1])), and the
        tentative collisions each took"""
        rays = np.ascontiguousarray(rays, np.float32).reshape(-1, 8)
        out = np.zeros(len(rays), np.float32)
        ns = np.zeros(len(rays), np.uint32)
        self.L.orc_volume_ratio_n(self.h, k, rays.ctypes.data, len(rays), seed[0], seed[1], abi.fptr(out),
                                  ns.ctypes.data, threads or default_threads())
        return out, ns

    def set_dir_lights(self, lights):
        arr = abi.dir_lights(lights)
        self._keep.append(arr)
        self.L.orc_scene_set_dir_lights(self.h, arr, len(lights))

    def set_env_map(self, radiance, direction=(1.0, 0.0, 0.0), up=(0.0, 0.0, 1.0), scale=1.0):
        """HDRI light from an (height, width, 3) radiance image; None removes it"""
        if radiance is None:
            self.L.orc_scene_set_env_map(self.h, None)
            self.env_shape = None
            return
        d, keep = abi.env_map_desc(radiance, direction, up, scale)
        assert self.L.orc_scene_set_env_map(self.h, C.byref(d)) == 0
        self.env_shape = keep.shape[:2]

    def env_cdfs(self):
        """(cdf_x (H, W), cdf_y (H,), to_world (3x3, columns), to_local)"""
        H, W = self.env_shape
        cx = np.zeros((H, W), np.float32)
        cy = np.zeros(H, np.float32)
        tw = np.zeros(9, np.float32)
        tl = np.zeros(9, np.float32)
        self.L.orc_env_cdfs(self.h, abi.fptr(cx), abi.fptr(cy), abi.fptr(tw), abi.fptr(tl))
        return cx, cy, tw.reshape(3, 3).T.copy(), tl.reshape(3, 3).T.copy()

    def env_tex(self, uv):
        uv = np.ascontiguousarray(uv, np.float32).reshape(-1, 2)
        out = np.zeros((len(uv), 4), np.float32)
        self.L.orc_env_tex(self.h, abi.fptr(uv), len(uv), abi.fptr(out))
        return out

    def env_pdf(self, dirs):
        d = np.ascontiguousarray(dirs, np.float32).reshape(-1, 3)
        out = np.zeros(len(d), np.float32)
        self.L.orc_env_pdf(self.h, abi.fptr(d), len(d), abi.fptr(out))
        return out

    def env_radiance(self, dirs):
        d = np.ascontiguousarray(dirs, np.float32).reshape(-1, 3)
        out = np.zeros((len(d), 3), np.float32)
        self.L.orc_env_radiance(self.h, abi.fptr(d), len(d), abi.fptr(out))
        return out

    def env_sample(self, n, seed=(1, 2)):
        """n samples in sequence from Random2(seed): (dirs, radiance, pdf)"""
        d = np.zeros((n, 3), np.float32)
        r = np.zeros((n, 3), np.float32)
        p = np.zeros(n, np.float32)
        self.L.orc_env_sample(self.h, seed[0], seed[1], n, abi.fptr(d), abi.fptr(r), abi.fptr(p))
        return d, r, p

    def env_uv_to_world(self, uv):
        uv = np.ascontiguousarray(uv, np.float32).reshape(-1, 2)
        out = np.zeros((len(uv), 3), np.float32)
        self.L.orc_env_uv_to_world(self.h, abi.fptr(uv), len(uv), abi.fptr(out))
        return out

    def build(self) -> float:
        return self.L.orc_scene_build(self.h)

    def trace(self, rays: np.ndarray, mode: int = 0, threads: int | None = None, cut_plane=None):
        """mode 0: BVH closest hit, 1: brute force, 2: BVH any hit.  cut_plane
        (nx, ny, nz, w): surface hits on its negative side are rejected."""
        rays = np.ascontiguousarray(rays, dtype=np.float32)
        n = len(rays)
        hits = np.zeros(n, dtype=abi.HIT_DTYPE)
        ctr = (C.c_uint64 * 4)()
        cp = None if cut_plane is None else np.ascontiguousarray(cut_plane, np.float32)
        self.L.orc_trace(self.h, rays.ctypes.data, n, hits.ctypes.data, mode, ctr,
                         threads or default_threads(), None if cp is None else abi.fptr(cp))
        return hits, dict(nodes=ctr[0], tris=ctr[1], spheres=ctr[2], rays=ctr[3])

    def bvh(self, group: int):
        info = (C.c_int32 * 3)()
        self.L.orc_scene_bvh(self.h, group, None, None, info)
        nodes = np.zeros(info[0], dtype=abi.NODE_DTYPE)
        order = np.zeros(info[2], dtype=np.int32)
        self.L.orc_scene_bvh(self.h, group, nodes.ctypes.data, abi.iptr(order), info)
        return nodes, order, int(info[1])


def dda_cells(dorg, ddir, tmax, dims, max_cells=4096):
    """dda3 (volume/DDA.h:60-201) of one grid-space ray: (cells n x 3, ranges n x 2)"""
    o = np.ascontiguousarray(dorg, np.float32)
    d = np.ascontiguousarray(ddir, np.float32)
    dm = np.ascontiguousarray(dims, np.int32)
    cells = np.zeros((max_cells, 3), np.int32)
    ts = np.zeros((max_cells, 2), np.float32)
    n = lib().orc_dda_cells(abi.fptr(o), abi.fptr(d), float(tmax), abi.iptr(dm), max_cells, abi.iptr(cells),
                            abi.fptr(ts))
    assert n <= max_cells
    return cells[:n], ts[:n]


def bvh_build_boxes(boxes: np.ndarray, max_leaf: int = 4):
    boxes = np.ascontiguousarray(boxes, dtype=np.float32).reshape(-1, 6)
    n = len(boxes)
    nodes = np.zeros(max(1, n), dtype=abi.NODE_DTYPE)
    order = np.zeros(max(1, n), dtype=np.int32)
    info = (C.c_int32 * 2)()
    lib().orc_bvh_build_boxes(abi.fptr(boxes), n, max_leaf, nodes.ctypes.data, abi.iptr(order), info)
    return nodes[: info[0]], order[:n], int(info[1])


def camera_orthographic(position, direction, up, height, aspect) -> abi.Camera:
    cam = abi.Camera()
    p, d, u = (np.ascontiguousarray(x, np.float32) for x in (position, direction, up))
    lib().orc_camera_orthographic(abi.fptr(p), abi.fptr(d), abi.fptr(u), float(height), float(aspect),
                                  C.byref(cam))
    return cam


def camera_omni(position, direction, up) -> abi.Camera:
    cam = abi.Camera()
    p, d, u = (np.ascontiguousarray(x, np.float32) for x in (position, direction, up))
    lib().orc_camera_omni(abi.fptr(p), abi.fptr(d), abi.fptr(u), C.byref(cam))
    return cam


def primary_rays(cam: abi.Camera, w: int, h: int, accum_id: int = 0) -> np.ndarray:
    """primary rays (bnx_ray layout, pixel ix + w iy) of sample accum_id"""
    out = np.zeros((w * h, 8), np.float32)
    lib().orc_primary_rays(C.byref(cam), w, h, accum_id, out.ctypes.data)
    return out


def camera_perspective(position, direction, up, fovy_deg) -> abi.Camera:
    cam = abi.Camera()
    p = np.asarray(position, np.float32)
    d = np.asarray(direction, np.float32)
    u = np.asarray(up, np.float32)
    lib().orc_camera_perspective(abi.fptr(p), abi.fptr(d), abi.fptr(u), fovy_deg, C.byref(cam))
    return cam


class Frame:
    def __init__(self, w: int, h: int):
        self.L = lib()
        self.w, self.h = w, h
        self.h_ = self.L.orc_frame_create(w, h)

    def __del__(self):
        try:
            self.L.orc_frame_destroy(self.h_)
        except Exception:
            pass

    def reset(self):
        self.L.orc_frame_reset(self.h_)

    def read(self, channel: int) -> np.ndarray:
        n = self.w * self.h
        if channel == abi.BNX_FB_COLOR:
            out = np.zeros((self.h, self.w, 4), np.float32)
        elif channel in (abi.BNX_FB_COLOR_RGBA8, abi.BNX_FB_COLOR_SRGB8):
            out = np.zeros((self.h, self.w, 4), np.uint8)
        elif channel == abi.BNX_FB_NORMAL:
            out = np.zeros((self.h, self.w, 3), np.float32)
        elif channel == abi.BNX_FB_DEPTH:
            out = np.zeros((self.h, self.w), np.float32)
        else:
            out = np.zeros((self.h, self.w), np.int32)
        assert out.size >= n
        self.L.orc_frame_read(self.h_, channel, out.ctypes.data)
        return out


def render(scene: Scene, cam: abi.Camera, rend: abi.Renderer, frame: Frame, spp: int,
           threads: int | None = None, shadow_any_hit: bool = False):
    """shadow_any_hit: count shadow rays as occlusion queries (the image is
    the same; only the traversal counters change)"""
    st = (C.c_uint64 * 7)()
    ms = scene.L.orc_render(scene.h, C.byref(cam), C.byref(rend), frame.h_, spp, threads or default_threads(), st,
                            1 if shadow_any_hit else 0)
    return ms, dict(rays=st[0], shadow=st[1], generations=st[2], volume_samples=st[3], nodes=st[4],
                    tris=st[5], spheres=st[6])


def cylinder_intersect(orgs, dirs, v0, v1, radius, tmin=0.0, tmax=1e30):
    """The oracle's capped-cylinder test (Cylinders.dev.cu:107-203) for n
    rays against one cylinder: (t, object normal); t = -1 where missed."""
    r = np.ascontiguousarray(np.concatenate([np.asarray(orgs, np.float32).reshape(-1, 3),
                                             np.asarray(dirs, np.float32).reshape(-1, 3)], axis=1))
    cyl = np.ascontiguousarray(np.concatenate([np.asarray(v0, np.float32), np.asarray(v1, np.float32),
                                               [np.float32(radius)]]).astype(np.float32))
    out = np.zeros((len(r), 4), np.float32)
    lib().orc_cylinder_intersect_n(abi.fptr(r), len(r), abi.fptr(cyl), tmin, tmax, abi.fptr(out))
    return out[:, 0], out[:, 1:]


def _rays(orgs, dirs):
    return np.ascontiguousarray(np.concatenate([np.asarray(orgs, np.float32).reshape(-1, 3),
                                                np.asarray(dirs, np.float32).reshape(-1, 3)], axis=1))


def capsule_intersect(orgs, dirs, v0, r0, v1, r1, tmax=1e30):
    """The oracle's capsule test (Capsules.dev.cu:154-212 with
    intersectRoundedCone) for n rays against one capsule: (t, normalized
    object normal, object point); t = -1 where missed."""
    r = _rays(orgs, dirs)
    cap = np.array([*np.asarray(v0, np.float32), r0, *np.asarray(v1, np.float32), r1], np.float32)
    out = np.zeros((len(r), 7), np.float32)
    lib().orc_capsule_intersect_n(abi.fptr(r), len(r), abi.fptr(cap), tmax, abi.fptr(out))
    return out[:, 0], out[:, 1:4], out[:, 4:7]


def cone_intersect(orgs, dirs, p0, ra, p1, rb, tmin=0.0, tmax=1e30):
    """The oracle's cone test (Cones.dev.cu:60-266) for n rays against one
    cone: (t, unnormalized object normal, attribute weight); t = -1 where
    missed."""
    r = _rays(orgs, dirs)
    cone = np.array([*np.asarray(p0, np.float32), ra, *np.asarray(p1, np.float32), rb], np.float32)
    out = np.zeros((len(r), 5), np.float32)
    lib().orc_cone_intersect_n(abi.fptr(r), len(r), abi.fptr(cone), tmin, tmax, abi.fptr(out))
    return out[:, 0], out[:, 1:4], out[:, 4]


def surface_scene(sc) -> Scene:
    """Oracle scene for a barney_amd.scenes.SurfaceScene: one user instance
    (ID 0) flattened into user-geometry groups (spheres, cylinders, capsules,
    cones) and a triangles group, in the order of native/ModelSlot.cpp:
    128-155."""
    s = Scene()
    sph = [s.add_spheres(x.centers, x.radii, x.colors, user_id=x.user_id, material=x.material) for x in sc.spheres]
    cyl = [s.add_cylinders(x.vertices, x.indices, x.radii, x.colors, user_id=x.user_id, material=x.material)
           for x in getattr(sc, "cylinders", [])]
    cap = [s.add_capsules(x.vertices, x.indices, x.colors, user_id=x.user_id, material=x.material)
           for x in getattr(sc, "capsules", [])]
    cone = [s.add_cones(x.vertices, x.radii, x.indices, x.colors, user_id=x.user_id, material=x.material)
            for x in getattr(sc, "cones", [])]
    tri = [s.add_triangles(m.vertices, m.indices, m.colors, user_id=m.user_id, material=m.material)
           for m in sc.meshes]
    objs = (list(sc.spheres) + list(getattr(sc, "cylinders", [])) + list(getattr(sc, "capsules", [])) +
            list(getattr(sc, "cones", [])) + list(sc.meshes))
    for gi, obj in zip(sph + cyl + cap + cone + tri, objs):
        for slot, (scope, arr) in getattr(obj, "attributes", {}).items():
            if scope == abi.SCOPE_CONSTANT:
                s.set_attribute(gi, slot, scope, value=arr)
            else:
                s.set_attribute(gi, slot, scope, data=arr)
        if getattr(obj, "normals", None) is not None:
            s.set_vertex_normals(gi, obj.normals)
    for slot, vals in getattr(sc, "inst_attr", {}).items():
        s.set_instance_attribute(slot, vals)
    if sph:
        s.add_instance(s.add_group(1, sph), 0)
    if cyl:
        s.add_instance(s.add_group(1, cyl), 0)
    if cap:
        s.add_instance(s.add_group(1, cap), 0)
    if cone:
        s.add_instance(s.add_group(1, cone), 0)
    if tri:
        s.add_instance(s.add_group(0, tri), 0)
    s.set_dir_lights(sc.dir_lights)
    env = getattr(sc, "env", None)
    if env is not None:
        s.set_env_map(env.radiance, env.direction, env.up, env.scale)
    s.build()
    return s


def volume_scene(vs) -> Scene:
    """Oracle scene for a barney_amd.scenes.StructuredVolumeScene: one volume
    as user instance 0, no surfaces."""
    s = Scene()
    s.add_volume(vs, user_id=vs.user_id, user_inst=0)
    s.set_dir_lights(vs.dir_lights)
    env = getattr(vs, "env", None)
    if env is not None:
        s.set_env_map(env.radiance, env.direction, env.up, env.scale)
    s.build()
    return s


def cell_invert(cell_type: int, verts4, P):
    """the Newton inversion of one pyramid / prism / hexahedron (type 3 / 2 /
    1; verts4: (n, 4) x, y, z, scalar in VTK order) at points P (m, 3):
    (inside bool[m], value float32[m], NaN outside)"""
    L = lib()
    v = np.ascontiguousarray(verts4, np.float32)
    p = np.ascontiguousarray(P, np.float32).reshape(-1, 3)
    inside = np.zeros(len(p), np.int32)
    val = np.zeros(len(p), np.float32)
    L.orc_cell_invert_n(int(cell_type), abi.fptr(v), len(p), abi.fptr(p), abi.iptr(inside), abi.fptr(val))
    return inside.astype(bool), val


def umesh_scene(us) -> Scene:
    """Oracle scene for a barney_amd.scenes.UMeshVolumeScene: the tet field as
    a volume (user instance 0, unless us.volume is false) and the isosurface
    over it (user instance 1), no surfaces."""
    s = Scene()
    k = s.add_umesh(us, user_id=us.user_id, user_inst=0, rendered=us.volume)
    s.add_isosurface(k, us.iso_value, us.iso_color, user_id=us.iso_user_id, user_inst=1)
    s.set_dir_lights(us.dir_lights)
    env = getattr(us, "env", None)
    if env is not None:
        s.set_env_map(env.radiance, env.direction, env.up, env.scale)
    s.build()
    return s
