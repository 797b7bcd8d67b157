"""barney_amd -- MI355X (gfx950) ray-tracing and volume backend for barney.

Python mirror of the C-ABI in include/bnx_rt.h (loaded with ctypes from the
in-tree barney_amd/libbnx.so).  The objects follow barney's rtcore backend
interface: Device, Buffer, Geom, Group (triangles / user geometry /
instances), plus the World, Frame and render() of the native layer.

There is no fallback: if the HIP library is missing or no GPU is present,
every call raises.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from . import abi

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libbnx.so")
_lib = None

vp = C.c_void_p


class BnxError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"bnx error {code}: {msg}")
        self.code = code


# bnx_denoise_fn: (user, in_rgba, in_normal, out_rgba, width, height, blend, stream) -> int
DENOISE_FN = C.CFUNCTYPE(C.c_int, vp, vp, vp, vp, C.c_int32, C.c_int32, C.c_float, vp)

# (name, argtypes, restype)
_PROTOS = [
    ("bnx_last_error", [], C.c_char_p),
    ("bnx_abi_version", [], C.c_int),
    ("bnx_device_create", [C.c_int, C.POINTER(vp)], C.c_int),
    ("bnx_device_destroy", [vp], C.c_int),
    ("bnx_device_sync", [vp], C.c_int),
    ("bnx_device_mem_info", [vp, C.POINTER(C.c_size_t), C.POINTER(C.c_size_t)], C.c_int),
    ("bnx_mem_alloc", [vp, C.c_size_t, C.POINTER(vp)], C.c_int),
    ("bnx_mem_free", [vp, vp], C.c_int),
    ("bnx_host_alloc", [vp, C.c_size_t, C.POINTER(vp)], C.c_int),
    ("bnx_host_free", [vp, vp], C.c_int),
    ("bnx_mem_copy_async", [vp, vp, vp, C.c_size_t], C.c_int),
    ("bnx_mem_set_async", [vp, vp, C.c_int, C.c_size_t], C.c_int),
    ("bnx_device_set_active", [vp, C.POINTER(C.c_int)], C.c_int),
    ("bnx_device_restore_active", [C.c_int], C.c_int),
    ("bnx_buffer_create", [vp, C.c_size_t, vp, C.POINTER(vp)], C.c_int),
    ("bnx_buffer_upload", [vp, vp, C.c_size_t, C.c_size_t], C.c_int),
    ("bnx_buffer_download", [vp, vp, C.c_size_t, C.c_size_t], C.c_int),
    ("bnx_buffer_destroy", [vp], C.c_int),
    ("bnx_buffer_dptr", [vp], vp),
    ("bnx_buffer_size", [vp], C.c_size_t),
    ("bnx_geom_create_triangles", [vp, C.POINTER(abi.TrianglesDesc), C.POINTER(vp)], C.c_int),
    ("bnx_geom_create_spheres", [vp, C.POINTER(abi.SpheresDesc), C.POINTER(vp)], C.c_int),
    ("bnx_geom_create_cylinders", [vp, C.POINTER(abi.CylindersDesc), C.POINTER(vp)], C.c_int),
    ("bnx_geom_create_capsules", [vp, C.POINTER(abi.CapsulesDesc), C.POINTER(vp)], C.c_int),
    ("bnx_geom_create_cones", [vp, C.POINTER(abi.ConesDesc), C.POINTER(vp)], C.c_int),
    ("bnx_geom_destroy", [vp], C.c_int),
    ("bnx_group_create_triangles", [vp, C.POINTER(vp), C.c_int, C.POINTER(vp)], C.c_int),
    ("bnx_group_create_user", [vp, C.POINTER(vp), C.c_int, C.POINTER(vp)], C.c_int),
    ("bnx_group_create_instances", [vp, C.POINTER(vp), C.POINTER(C.c_int32), C.POINTER(C.c_float), C.c_int,
                                    C.POINTER(vp)], C.c_int),
    ("bnx_group_build", [vp], C.c_int),
    ("bnx_group_set_transforms", [vp, C.POINTER(C.c_float), C.c_int], C.c_int),
    ("bnx_group_set_instance_attribute", [vp, C.c_int, C.POINTER(C.c_float), C.c_int], C.c_int),
    ("bnx_geom_set_attribute", [vp, C.c_int, C.c_int, C.c_int, C.POINTER(C.c_float), C.c_int64,
                                C.POINTER(C.c_float)], C.c_int),
    ("bnx_geom_set_vertex_normals", [vp, C.POINTER(C.c_float), C.c_int], C.c_int),
    ("bnx_group_bvh_info", [vp, C.POINTER(abi.BvhInfo)], C.c_int),
    ("bnx_group_download_bvh", [vp, vp, C.POINTER(C.c_int32)], C.c_int),
    ("bnx_group_destroy", [vp], C.c_int),
    ("bnx_bvh_build_boxes", [vp, vp, C.c_int, C.c_int, vp, C.POINTER(C.c_int32), C.POINTER(abi.BvhInfo)], C.c_int),
    ("bnx_trace", [vp, vp, C.c_int, vp, C.c_uint32], C.c_int),
    ("bnx_volume_create_structured", [vp, C.POINTER(abi.StructuredDesc), C.POINTER(abi.TransferFunction), C.c_int32,
                                      C.POINTER(vp)], C.c_int),
    ("bnx_volume_set_transfer_function", [vp, C.POINTER(abi.TransferFunction)], C.c_int),
    ("bnx_volume_majorants", [vp, C.POINTER(C.c_int32), C.POINTER(C.c_float)], C.c_int),
    ("bnx_volume_destroy", [vp], C.c_int),
    ("bnx_volume_create_umesh", [vp, C.POINTER(abi.UMeshDesc), C.POINTER(abi.TransferFunction), C.c_int32,
                                 C.POINTER(vp)], C.c_int),
    ("bnx_volume_mc_ranges", [vp, C.POINTER(C.c_int32), C.POINTER(C.c_float), C.POINTER(C.c_float)], C.c_int),
    ("bnx_isosurface_create", [vp, C.c_float, C.POINTER(abi.Material), C.c_int32, C.POINTER(vp)], C.c_int),
    ("bnx_isosurface_destroy", [vp], C.c_int),
    ("bnx_world_add_volume", [vp, vp, C.c_int32], C.c_int),
    ("bnx_world_add_isosurface", [vp, vp, C.c_int32], C.c_int),
    ("bnx_world_create", [vp, vp, C.POINTER(abi.DirLight), C.c_int, C.POINTER(vp)], C.c_int),
    ("bnx_world_set_env_map", [vp, C.POINTER(abi.EnvMapDesc)], C.c_int),
    ("bnx_world_destroy", [vp], C.c_int),
    ("bnx_camera_perspective", [C.POINTER(C.c_float), C.POINTER(C.c_float), C.POINTER(C.c_float), C.c_float,
                                C.POINTER(abi.Camera)], C.c_int),
    ("bnx_camera_orthographic", [C.POINTER(C.c_float), C.POINTER(C.c_float), C.POINTER(C.c_float), C.c_float,
                                 C.c_float, C.POINTER(abi.Camera)], C.c_int),
    ("bnx_camera_omni", [C.POINTER(C.c_float), C.POINTER(C.c_float), C.POINTER(C.c_float),
                         C.POINTER(abi.Camera)], C.c_int),
    ("bnx_frame_create", [vp, C.c_int, C.c_int, C.POINTER(vp)], C.c_int),
    ("bnx_frame_reset", [vp], C.c_int),
    ("bnx_frame_read", [vp, C.c_int, vp], C.c_int),
    ("bnx_frame_set_denoiser", [vp, DENOISE_FN, vp, C.c_int32], C.c_int),
    ("bnx_copy", [vp, vp, vp, C.c_size_t], C.c_int),
    ("bnx_frame_accum_id", [vp], C.c_int),
    ("bnx_frame_destroy", [vp], C.c_int),
    ("bnx_render", [vp, C.POINTER(abi.Camera), C.POINTER(abi.Renderer), vp, C.c_int, C.POINTER(abi.RenderStats)],
     C.c_int),
    ("bnx_render_noalloc", [vp, C.POINTER(abi.Camera), C.POINTER(abi.Renderer), vp, C.c_int,
                            C.POINTER(abi.RenderStats)], C.c_int),
    ("bnx_device_count", [C.POINTER(C.c_int)], C.c_int),
    ("bnx_device_get_tuning", [vp, C.POINTER(abi.Tuning)], C.c_int),
    ("bnx_device_set_tuning", [vp, C.POINTER(abi.Tuning)], C.c_int),
    ("bnx_autotune", [vp, C.POINTER(abi.Camera), C.POINTER(abi.Renderer), C.c_int, C.c_int, C.c_int,
                      C.POINTER(abi.Tuning), C.POINTER(abi.AutotuneReport)], C.c_int),
    ("bnx_server_create", [vp, C.POINTER(abi.ServerConfig), C.POINTER(vp)], C.c_int),
    ("bnx_server_destroy", [vp], C.c_int),
    ("bnx_server_submit", [vp, C.POINTER(abi.Request), C.POINTER(C.c_int64)], C.c_int),
    ("bnx_server_step", [vp, C.POINTER(C.c_int64), C.c_int, C.POINTER(C.c_int)], C.c_int),
    ("bnx_server_step_n", [vp, C.c_int, C.POINTER(C.c_int64), C.c_int, C.POINTER(C.c_int)], C.c_int),
    ("bnx_server_poll", [vp, C.c_int64, C.POINTER(abi.RequestStatus)], C.c_int),
    ("bnx_server_read", [vp, C.c_int64, C.c_int, vp], C.c_int),
    ("bnx_server_release", [vp, C.c_int64], C.c_int),
    ("bnx_server_get_stats", [vp, C.POINTER(abi.ServerStats)], C.c_int),
]
EXPORTED = [p[0] for p in _PROTOS]


def lib():
    """The loaded HIP library; raises when it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"HIP library missing: {LIB_PATH} (run `python -m barney_amd.build`)")
        L = C.CDLL(LIB_PATH)
        for name, args, res in _PROTOS:
            f = getattr(L, name)
            f.argtypes = args
            f.restype = res
        _lib = L
    return _lib


def check(rc: int):
    if rc != abi.BNX_OK:
        raise BnxError(rc, lib().bnx_last_error().decode())


def _handle_array(objs):
    arr = (vp * max(1, len(objs)))()
    for i, o in enumerate(objs):
        arr[i] = o.h
    return arr


class Device:
    """rtc::Device: one GPU and its stream."""

    def __init__(self, gpu: int = 0):
        self.L = lib()
        h = vp()
        check(self.L.bnx_device_create(gpu, C.byref(h)))
        self.h = h

    def sync(self):
        check(self.L.bnx_device_sync(self.h))

    def mem_info(self):
        f, t = C.c_size_t(), C.c_size_t()
        check(self.L.bnx_device_mem_info(self.h, C.byref(f), C.byref(t)))
        return f.value, t.value

    def tuning(self) -> dict:
        """the launch tuning in place (bnx_device_get_tuning)"""
        t = abi.Tuning()
        check(self.L.bnx_device_get_tuning(self.h, C.byref(t)))
        return {k: getattr(t, k) for k in abi.TUNING_FIELDS}

    def set_tuning(self, **fields):
        """change some fields of the launch tuning (bnx_device_set_tuning)"""
        cur = self.tuning()
        unknown = set(fields) - set(cur)
        if unknown:
            raise ValueError(f"unknown tuning fields {sorted(unknown)}")
        cur.update(fields)
        t = abi.Tuning(**cur)
        check(self.L.bnx_device_set_tuning(self.h, C.byref(t)))

    def close(self):
        if self.h:
            self.L.bnx_device_destroy(self.h)
            self.h = None


class Buffer:
    """rtc::Buffer: device memory."""

    def __init__(self, dev: Device, nbytes: int = 0, data: np.ndarray | None = None):
        self.dev, self.L = dev, dev.L
        if data is not None:
            data = np.ascontiguousarray(data)
            nbytes = data.nbytes
        h = vp()
        check(self.L.bnx_buffer_create(dev.h, nbytes, data.ctypes.data if data is not None else None, C.byref(h)))
        self.h = h
        self.nbytes = nbytes

    @property
    def dptr(self) -> int:
        return self.L.bnx_buffer_dptr(self.h)

    def upload(self, data: np.ndarray, offset: int = 0):
        data = np.ascontiguousarray(data)
        check(self.L.bnx_buffer_upload(self.h, data.ctypes.data, data.nbytes, offset))

    def download(self, out: np.ndarray, offset: int = 0) -> np.ndarray:
        check(self.L.bnx_buffer_download(self.h, out.ctypes.data, out.nbytes, offset))
        return out

    def close(self):
        if self.h:
            self.L.bnx_buffer_destroy(self.h)
            self.h = None


class Geom:
    def __init__(self, dev, h, keep):
        self.dev, self.h, self._keep = dev, h, keep

    @staticmethod
    def triangles(dev: Device, vertices, indices, colors=None, user_id=0, material=None) -> "Geom":
        v = np.ascontiguousarray(vertices, dtype=np.float32)
        i = np.ascontiguousarray(indices, dtype=np.int32)
        c = None if colors is None else np.ascontiguousarray(colors, dtype=np.float32)
        d = abi.TrianglesDesc(abi.fptr(v), len(v), len(i), abi.iptr(i), abi.fptr(c), user_id, 0,
                              material or abi.matte())
        h = vp()
        check(dev.L.bnx_geom_create_triangles(dev.h, C.byref(d), C.byref(h)))
        return Geom(dev, h, None)

    @staticmethod
    def spheres(dev: Device, centers, radii=None, colors=None, user_id=1, material=None,
                default_radius=0.1) -> "Geom":
        ce = np.ascontiguousarray(centers, dtype=np.float32)
        r = None if radii is None else np.ascontiguousarray(radii, dtype=np.float32)
        c = None if colors is None else np.ascontiguousarray(colors, dtype=np.float32)
        d = abi.SpheresDesc(abi.fptr(ce), abi.fptr(r), abi.fptr(c), len(ce), default_radius, user_id, 0,
                            material or abi.matte())
        h = vp()
        check(dev.L.bnx_geom_create_spheres(dev.h, C.byref(d), C.byref(h)))
        return Geom(dev, h, None)

    @staticmethod
    def cylinders(dev: Device, vertices, indices, radii=None, colors=None, user_id=1, material=None,
                  default_radius=0.1) -> "Geom":
        v = np.ascontiguousarray(vertices, dtype=np.float32)
        i = np.ascontiguousarray(indices, dtype=np.int32)
        r = None if radii is None else np.ascontiguousarray(radii, dtype=np.float32)
        c = None if colors is None else np.ascontiguousarray(colors, dtype=np.float32)
        d = abi.CylindersDesc(abi.fptr(v), abi.iptr(i), abi.fptr(r), abi.fptr(c), len(v), len(i), default_radius,
                              user_id, material or abi.matte())
        h = vp()
        check(dev.L.bnx_geom_create_cylinders(dev.h, C.byref(d), C.byref(h)))
        return Geom(dev, h, None)

    @staticmethod
    def capsules(dev: Device, vertices, indices, colors=None, user_id=1, material=None) -> "Geom":
        """vertices (nv, 4): position and radius per vertex; indices (n, 2)"""
        v = np.ascontiguousarray(vertices, dtype=np.float32)
        i = np.ascontiguousarray(indices, dtype=np.int32)
        c = None if colors is None else np.ascontiguousarray(colors, dtype=np.float32)
        d = abi.CapsulesDesc(abi.fptr(v), abi.iptr(i), abi.fptr(c), len(v), len(i), user_id, 0,
                             material or abi.matte())
        h = vp()
        check(dev.L.bnx_geom_create_capsules(dev.h, C.byref(d), C.byref(h)))
        return Geom(dev, h, None)

    @staticmethod
    def cones(dev: Device, vertices, radii, indices=None, colors=None, user_id=1, material=None,
              num_cones=None) -> "Geom":
        """vertices (nv, 3), radii (nv,); indices (n, 2) or None (cone i
        joins vertices 2i and 2i+1)"""
        v = np.ascontiguousarray(vertices, dtype=np.float32)
        r = np.ascontiguousarray(radii, dtype=np.float32)
        i = None if indices is None else np.ascontiguousarray(indices, dtype=np.int32)
        c = None if colors is None else np.ascontiguousarray(colors, dtype=np.float32)
        n = len(i) if i is not None else (len(v) // 2 if num_cones is None else num_cones)
        d = abi.ConesDesc(abi.fptr(v), abi.fptr(r), abi.iptr(i), abi.fptr(c), len(v), n, user_id, 0,
                          material or abi.matte())
        h = vp()
        check(dev.L.bnx_geom_create_cones(dev.h, C.byref(d), C.byref(h)))
        return Geom(dev, h, None)

    def set_attribute(self, slot: int, scope: int, data=None, value=None) -> "Geom":
        """bnx_geom_set_attribute: data (n, comps) floats, or an rgba constant"""
        d = None if data is None else np.ascontiguousarray(data, dtype=np.float32)
        comps = 0 if d is None else (1 if d.ndim == 1 else d.shape[1])
        v = None if value is None else np.ascontiguousarray(value, dtype=np.float32)
        check(self.dev.L.bnx_geom_set_attribute(self.h, slot, scope, comps, abi.fptr(d), 0 if d is None else len(d),
                                                abi.fptr(v)))
        return self

    def set_vertex_normals(self, normals) -> "Geom":
        n = None if normals is None else np.ascontiguousarray(normals, dtype=np.float32)
        check(self.dev.L.bnx_geom_set_vertex_normals(self.h, abi.fptr(n), 0 if n is None else len(n)))
        return self

    def set_attributes(self, obj) -> "Geom":
        """the attributes (and triangles' vertex normals) of a scene container"""
        for slot, (scope, arr) in getattr(obj, "attributes", {}).items():
            if scope == abi.SCOPE_CONSTANT:
                self.set_attribute(slot, scope, value=arr)
            else:
                self.set_attribute(slot, scope, data=arr)
        if getattr(obj, "normals", None) is not None:
            self.set_vertex_normals(obj.normals)
        return self


TRIANGLES, USER = 0, 1


class Group:
    """rtc::Group: a triangles group, a user-geometry group, or an instance group."""

    def __init__(self, dev: Device, h, members):
        self.dev, self.h, self.members = dev, h, members

    @staticmethod
    def geoms(dev: Device, kind: int, geoms) -> "Group":
        arr = _handle_array(geoms)
        h = vp()
        fn = dev.L.bnx_group_create_triangles if kind == TRIANGLES else dev.L.bnx_group_create_user
        check(fn(dev.h, arr, len(geoms), C.byref(h)))
        return Group(dev, h, list(geoms))

    @staticmethod
    def instances(dev: Device, groups, inst_ids=None, xfms=None) -> "Group":
        arr = _handle_array(groups)
        ids = (C.c_int32 * max(1, len(groups)))(*(inst_ids if inst_ids is not None else range(len(groups))))
        x = None
        if xfms is not None:
            xa = np.ascontiguousarray(xfms, dtype=np.float32).reshape(-1)
            x = (C.c_float * len(xa))(*xa)
        h = vp()
        check(dev.L.bnx_group_create_instances(dev.h, arr, ids, x, len(groups), C.byref(h)))
        return Group(dev, h, list(groups))

    def build(self):
        check(self.dev.L.bnx_group_build(self.h))
        return self

    def set_transforms(self, xfms):
        """new object-to-world transforms (n x 12, as at creation) for every
        instance, and the instance BVH rebuilt (bnx_group_set_transforms)"""
        xa = np.ascontiguousarray(xfms, dtype=np.float32).reshape(-1)
        check(self.dev.L.bnx_group_set_transforms(self.h, (C.c_float * max(1, len(xa)))(*xa), len(xa) // 12))
        return self

    def info(self) -> abi.BvhInfo:
        inf = abi.BvhInfo()
        check(self.dev.L.bnx_group_bvh_info(self.h, C.byref(inf)))
        return inf

    def bvh(self):
        inf = self.info()
        nodes = np.zeros(max(1, inf.num_nodes), dtype=abi.NODE_DTYPE)
        order = np.zeros(max(1, inf.num_prims), dtype=np.int32)
        check(self.dev.L.bnx_group_download_bvh(self.h, nodes.ctypes.data, abi.iptr(order)))
        return nodes[: inf.num_nodes], order[: inf.num_prims], inf.max_depth

    def trace(self, rays: np.ndarray, any_hit: bool = False) -> np.ndarray:
        """Closest (or any) hit of host rays (n x 8 float32, bnx_ray layout)."""
        rays = np.ascontiguousarray(rays, dtype=np.float32).reshape(-1, 8)
        n = len(rays)
        hits = np.zeros(n, dtype=abi.HIT_DTYPE)
        if n == 0:
            return hits
        rb = Buffer(self.dev, data=rays)
        hb = Buffer(self.dev, nbytes=hits.nbytes)
        try:
            check(self.dev.L.bnx_trace(self.h, rb.dptr, n, hb.dptr,
                                       abi.BNX_TRACE_ANY if any_hit else abi.BNX_TRACE_CLOSEST))
            hb.download(hits)
        finally:
            rb.close()
            hb.close()
        return hits


def bvh_build_boxes(dev: Device, boxes: np.ndarray, max_leaf: int = 4):
    boxes = np.ascontiguousarray(boxes, dtype=np.float32).reshape(-1, 6)
    n = len(boxes)
    nodes = np.zeros(max(1, n), dtype=abi.NODE_DTYPE)
    order = np.zeros(max(1, n), dtype=np.int32)
    inf = abi.BvhInfo()
    b = Buffer(dev, data=boxes if n else np.zeros(6, np.float32))
    try:
        check(dev.L.bnx_bvh_build_boxes(dev.h, b.dptr, n, max_leaf, nodes.ctypes.data, abi.iptr(order),
                                        C.byref(inf)))
    finally:
        b.close()
    return nodes[: inf.num_nodes], order[:n], inf


def camera_perspective(position, direction, up, fovy_deg, aperture_radius=0.0, focus_distance=0.0) -> abi.Camera:
    """PerspectiveCamera; depth of field when aperture_radius > 0"""
    L = lib()
    cam = abi.Camera()
    p = (C.c_float * 3)(*position)
    d = (C.c_float * 3)(*direction)
    u = (C.c_float * 3)(*up)
    check(L.bnx_camera_perspective(p, d, u, fovy_deg, C.byref(cam)))
    if aperture_radius > 0.0:
        cam.aperture_radius = aperture_radius
        cam.focus_distance = focus_distance
    return cam


def camera_orthographic(position, direction, up, height, aspect) -> abi.Camera:
    L = lib()
    cam = abi.Camera()
    p, d, u = ((C.c_float * 3)(*v) for v in (position, direction, up))
    check(L.bnx_camera_orthographic(p, d, u, float(height), float(aspect), C.byref(cam)))
    return cam


def camera_omni(position, direction, up) -> abi.Camera:
    L = lib()
    cam = abi.Camera()
    p, d, u = ((C.c_float * 3)(*v) for v in (position, direction, up))
    check(L.bnx_camera_omni(p, d, u, C.byref(cam)))
    return cam


def _channel_array(channel, h, w):
    if channel == abi.BNX_FB_COLOR:
        return np.zeros((h, w, 4), np.float32)
    if channel in (abi.BNX_FB_COLOR_RGBA8, abi.BNX_FB_COLOR_SRGB8):
        return np.zeros((h, w, 4), np.uint8)
    if channel == abi.BNX_FB_NORMAL:
        return np.zeros((h, w, 3), np.float32)
    if channel == abi.BNX_FB_DEPTH:
        return np.zeros((h, w), np.float32)
    return np.zeros((h, w), np.int32)


class Volume:
    """A volume with its transfer function (MCVolumeAccel) over a structured
    field (scenes.StructuredVolumeScene) or a tet mesh (scenes.UMeshVolumeScene)."""

    def __init__(self, dev: Device, vs, user_id: int = 0):
        self.dev = dev
        h = vp()
        if hasattr(vs, "tets"):
            d, t, keep = abi.umesh_desc(vs)
            check(dev.L.bnx_volume_create_umesh(dev.h, C.byref(d), C.byref(t), user_id, C.byref(h)))
        else:
            d, t, keep = abi.structured_desc(vs)
            check(dev.L.bnx_volume_create_structured(dev.h, C.byref(d), C.byref(t), user_id, C.byref(h)))
        self.h = h

    def mc_ranges(self):
        dims = (C.c_int32 * 3)()
        check(self.dev.L.bnx_volume_mc_ranges(self.h, dims, None, None))
        lo = np.zeros(dims[0] * dims[1] * dims[2], np.float32)
        hi = np.zeros_like(lo)
        check(self.dev.L.bnx_volume_mc_ranges(self.h, dims, abi.fptr(lo), abi.fptr(hi)))
        return np.array(list(dims)), lo, hi

    def set_transfer_function(self, tf, domain=(0.0, 1.0), base_density=1.0):
        tf = np.ascontiguousarray(tf, dtype=np.float32)
        t = abi.TransferFunction(abi.fptr(tf), len(tf), (C.c_float * 2)(*domain), base_density)
        check(self.dev.L.bnx_volume_set_transfer_function(self.h, C.byref(t)))

    def majorants(self):
        dims = (C.c_int32 * 3)()
        check(self.dev.L.bnx_volume_majorants(self.h, dims, None))
        m = np.zeros(dims[0] * dims[1] * dims[2], np.float32)
        check(self.dev.L.bnx_volume_majorants(self.h, dims, abi.fptr(m)))
        return np.array(list(dims)), m


class Isosurface:
    """An isosurface over a volume's field (MCIsoSurfaceAccel), opaque matte."""

    def __init__(self, field: Volume, iso: float, color=(0.8, 0.8, 0.8, 1.0), user_id: int = 0):
        self.dev, self.field = field.dev, field
        m = abi.matte(color=tuple(color), from_attribute=False)
        h = vp()
        check(field.dev.L.bnx_isosurface_create(field.h, float(iso), C.byref(m), user_id, C.byref(h)))
        self.h = h


class World:
    def __init__(self, dev: Device, instances: Group, dir_lights=()):
        self.dev, self.instances = dev, instances
        arr = abi.dir_lights(list(dir_lights))
        h = vp()
        check(dev.L.bnx_world_create(dev.h, instances.h, arr, len(dir_lights), C.byref(h)))
        self.h = h
        self.volumes = []

    def add_volume(self, vol: Volume, user_inst_id: int = 0):
        check(self.dev.L.bnx_world_add_volume(self.h, vol.h, user_inst_id))
        self.volumes.append(vol)

    def add_isosurface(self, iso: Isosurface, user_inst_id: int = 0):
        check(self.dev.L.bnx_world_add_isosurface(self.h, iso.h, user_inst_id))
        self.volumes.append(iso)

    def set_env_map(self, radiance, direction=(1.0, 0.0, 0.0), up=(0.0, 0.0, 1.0), scale=1.0):
        """HDRI environment light from an (height, width, 3) radiance image
        (row 0 at theta = 0); None removes it"""
        if radiance is None:
            check(self.dev.L.bnx_world_set_env_map(self.h, None))
            return
        d, _keep = abi.env_map_desc(radiance, direction, up, scale)
        check(self.dev.L.bnx_world_set_env_map(self.h, C.byref(d)))

    def set_env(self, env):
        """the env map of a scene description (scenes.EnvMap or None)"""
        if env is not None:
            self.set_env_map(env.radiance, env.direction, env.up, env.scale)


class Frame:
    def __init__(self, dev: Device, w: int, h: int):
        self.dev, self.w, self.h_ = dev, w, None
        self.hgt = h
        hh = vp()
        check(dev.L.bnx_frame_create(dev.h, w, h, C.byref(hh)))
        self.h_ = hh

    @property
    def accum_id(self) -> int:
        return self.dev.L.bnx_frame_accum_id(self.h_)

    def reset(self):
        check(self.dev.L.bnx_frame_reset(self.h_))

    def read(self, channel: int) -> np.ndarray:
        out = _channel_array(channel, self.hgt, self.w)
        check(self.dev.L.bnx_frame_read(self.h_, channel, out.ctypes.data))
        return out

    def set_denoiser(self, fn, fade_out: bool = False):
        """bnx_frame_set_denoiser: fn(user, d_in_rgba, d_in_normal, d_out_rgba,
        width, height, blend, stream) -> int, called with device pointers
        (ints) when the colour channel is read; None removes the hook"""
        self._denoise = DENOISE_FN(fn) if fn is not None else None
        check(self.dev.L.bnx_frame_set_denoiser(self.h_, self._denoise if fn is not None else DENOISE_FN(),
                                                None, int(bool(fade_out))))

    def close(self):
        if self.h_:
            self.dev.L.bnx_frame_destroy(self.h_)
            self.h_ = None


def autotune(world: World, cam: abi.Camera, rend: abi.Renderer, width: int, height: int, probe_spp: int = 1):
    """bnx_autotune: choose and set the device's launch tuning for requests
    like this one; returns (tuning dict, report dict)"""
    L = world.dev.L
    t, rep = abi.Tuning(), abi.AutotuneReport()
    check(L.bnx_autotune(world.h, C.byref(cam), C.byref(rend), width, height, probe_spp, C.byref(t), C.byref(rep)))
    return ({k: getattr(t, k) for k in abi.TUNING_FIELDS},
            {"ms_before": rep.ms_before, "ms_after": rep.ms_after, "probes": rep.probes})


def render(world: World, cam: abi.Camera, rend: abi.Renderer, frame: Frame, spp: int) -> dict:
    st = abi.RenderStats()
    check(world.dev.L.bnx_render(world.h, C.byref(cam), C.byref(rend), frame.h_, spp, C.byref(st)))
    return dict(rays=st.rays_traced, shadow=st.shadow_rays, generations=st.generations,
                volume_samples=st.volume_samples, ms=st.ms_total, nodes=st.node_visits, tris=st.tri_tests,
                spheres=st.sphere_tests, trace_ms=st.trace_ms, trace_launches=st.trace_launches,
                march_ms=st.march_ms, march_launches=st.march_launches, mc_cells=st.mc_cells,
                march_rays=st.march_rays, umesh_nodes=st.umesh_nodes, umesh_tets=st.umesh_tets,
                captured_batches=st.captured_batches, uncaptured_batches=st.uncaptured_batches,
                capture_error=st.capture_error)


class Server:
    """Continuous-batching request server (bnx_server_*): requests of any
    resident world, resolution and spp share one wavefront on the device."""

    def __init__(self, dev: Device, frame_memory_budget=0, max_paths_in_flight=0, max_active=0,
                 deadline_slack=0.0):
        self.dev, self.L = dev, dev.L
        cfg = abi.ServerConfig(frame_memory_budget, max_paths_in_flight, max_active, float(deadline_slack))
        h = vp()
        check(self.L.bnx_server_create(dev.h, C.byref(cfg), C.byref(h)))
        self.h = h
        self.sizes = {}
        self.pending = set()
        self._done = (C.c_int64 * 4096)()

    def submit(self, world: World, cam: abi.Camera, rend: abi.Renderer, width: int, height: int, spp: int) -> int:
        r = abi.Request(world.h, cam, width, height, rend, spp, 0)
        t = C.c_int64()
        check(self.L.bnx_server_submit(self.h, C.byref(r), C.byref(t)))
        self.sizes[t.value] = (width, height)
        self.pending.add(t.value)
        return t.value

    def step(self, generations: int = 0) -> list:
        """one generation read back by the host (generations=0, bnx_server_step)
        or that many generations with the bookkeeping on the device and one
        read-back (bnx_server_step_n); returns the tickets completed"""
        n = C.c_int()
        if generations > 0:
            check(self.L.bnx_server_step_n(self.h, generations, self._done, len(self._done), C.byref(n)))
        else:
            check(self.L.bnx_server_step(self.h, self._done, len(self._done), C.byref(n)))
        done = [self._done[k] for k in range(min(n.value, len(self._done)))]
        self.pending.difference_update(done)
        return done

    def poll(self, ticket: int) -> dict:
        st = abi.RequestStatus()
        check(self.L.bnx_server_poll(self.h, ticket, C.byref(st)))
        return dict(state=st.state, submit_ms=st.submit_ms, admit_ms=st.admit_ms, done_ms=st.done_ms,
                    paths_started=st.paths_started, paths_total=st.paths_total)

    def read(self, ticket: int, channel: int) -> np.ndarray:
        w, h = self.sizes[ticket]
        out = _channel_array(channel, h, w)
        check(self.L.bnx_server_read(self.h, ticket, channel, out.ctypes.data))
        return out

    def release(self, ticket: int):
        check(self.L.bnx_server_release(self.h, ticket))
        self.sizes.pop(ticket, None)

    def stats(self) -> dict:
        st = abi.ServerStats()
        check(self.L.bnx_server_get_stats(self.h, C.byref(st)))
        return {k: getattr(st, k) for k, _ in abi.ServerStats._fields_}

    def run_until_idle(self, max_steps=1_000_000, generations: int = 0) -> list:
        """step until every submitted request is done; returns the tickets in completion order"""
        done = []
        for _ in range(max_steps):
            if not self.pending:
                break
            done += self.step(generations)
        return done

    def close(self):
        if self.h:
            self.L.bnx_server_destroy(self.h)
            self.h = None


class SurfaceWorld:
    """GPU world for a barney_amd.scenes.SurfaceScene, flattened like
    native/ModelSlot.cpp:128-155: user-geometry group (spheres) first, then
    the triangles group, all as instances of user instance 0.  Cylinders,
    capsules and cones get a user group each, after the spheres', in that
    order."""

    def __init__(self, dev: Device, sc):
        self.dev = dev
        self.sph = [Geom.spheres(dev, s.centers, s.radii, s.colors, user_id=s.user_id,
                                 material=s.material).set_attributes(s) for s in sc.spheres]
        self.cyl = [Geom.cylinders(dev, c.vertices, c.indices, c.radii, c.colors, user_id=c.user_id,
                                   material=c.material).set_attributes(c) for c in getattr(sc, "cylinders", [])]
        self.cap = [Geom.capsules(dev, c.vertices, c.indices, c.colors, user_id=c.user_id,
                                  material=c.material).set_attributes(c) for c in getattr(sc, "capsules", [])]
        self.cone = [Geom.cones(dev, c.vertices, c.radii, c.indices, c.colors, user_id=c.user_id,
                                material=c.material).set_attributes(c) for c in getattr(sc, "cones", [])]
        self.tri = [Geom.triangles(dev, m.vertices, m.indices, m.colors, user_id=m.user_id,
                                   material=m.material).set_attributes(m) for m in sc.meshes]
        groups = []
        if self.sph:
            groups.append(Group.geoms(dev, USER, self.sph).build())
        if self.cyl:
            groups.append(Group.geoms(dev, USER, self.cyl).build())
        if self.cap:
            groups.append(Group.geoms(dev, USER, self.cap).build())
        if self.cone:
            groups.append(Group.geoms(dev, USER, self.cone).build())
        if self.tri:
            groups.append(Group.geoms(dev, TRIANGLES, self.tri).build())
        self.groups = groups
        self.insts = Group.instances(dev, groups, [0] * len(groups))
        for slot, vals in getattr(sc, "inst_attr", {}).items():
            v = np.ascontiguousarray(vals, dtype=np.float32)
            check(dev.L.bnx_group_set_instance_attribute(self.insts.h, slot, abi.fptr(v), len(v)))
        self.insts.build()
        self.world = World(dev, self.insts, sc.dir_lights)
        self.world.set_env(getattr(sc, "env", None))


class VolumeWorld:
    """GPU world for a barney_amd.scenes.StructuredVolumeScene: no surfaces,
    one volume as user instance 0."""

    def __init__(self, dev: Device, vs):
        self.dev = dev
        self.insts = Group.instances(dev, [], []).build()
        self.volume = Volume(dev, vs, user_id=vs.user_id)
        self.world = World(dev, self.insts, vs.dir_lights)
        self.world.set_env(getattr(vs, "env", None))
        self.world.add_volume(self.volume, 0)


class UMeshWorld:
    """GPU world for a barney_amd.scenes.UMeshVolumeScene: no surfaces, the
    tet field as a volume (user instance 0, unless us.volume is false) and
    the isosurface over it (user instance 1)."""

    def __init__(self, dev: Device, us):
        self.dev = dev
        self.insts = Group.instances(dev, [], []).build()
        self.volume = Volume(dev, us, user_id=us.user_id)
        self.iso = Isosurface(self.volume, us.iso_value, us.iso_color, user_id=us.iso_user_id)
        self.world = World(dev, self.insts, us.dir_lights)
        self.world.set_env(getattr(us, "env", None))
        if us.volume:
            self.world.add_volume(self.volume, 0)
        self.world.add_isosurface(self.iso, 1)
