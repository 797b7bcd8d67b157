"""ctypes mirror of the POD records of include/bnx_rt.h."""
from __future__ import annotations

import ctypes as C

import numpy as np

BNX_OK = 0
BNX_ERR_INVALID, BNX_ERR_HIP, BNX_ERR_OOM, BNX_ERR_UNSUPPORTED, BNX_ERR_STATE = -1, -2, -3, -4, -5
BNX_MATERIAL_MATTE = 0
BNX_MATERIAL_PBR = 1
BNX_CAMERA_PERSPECTIVE = 1
BNX_RENDER_COMPAT_SPHERE_INSTID13 = 1
BNX_RENDER_RATIO_TRACKING = 2
BNX_FB_COLOR, BNX_FB_DEPTH, BNX_FB_PRIMID, BNX_FB_INSTID, BNX_FB_OBJID, BNX_FB_NORMAL = range(6)
BNX_FB_COLOR_RGBA8, BNX_FB_COLOR_SRGB8 = 6, 7
BNX_CAMERA_PERSPECTIVE, BNX_CAMERA_ORTHOGRAPHIC, BNX_CAMERA_OMNI = 1, 2, 3
BNX_TRACE_CLOSEST = 0
BNX_TRACE_ANY = 1

f3 = C.c_float * 3
f4 = C.c_float * 4
fp = C.POINTER(C.c_float)
ip = C.POINTER(C.c_int32)


class Material(C.Structure):
    _fields_ = [("type", C.c_int32), ("color_from_attribute", C.c_int32), ("color", f4),
                ("opacity", C.c_float), ("metallic", C.c_float), ("roughness", C.c_float),
                ("transmission", C.c_float), ("ior", C.c_float), ("_pad", C.c_float)]


class TrianglesDesc(C.Structure):
    _fields_ = [("vertices", fp), ("num_vertices", C.c_int32), ("num_triangles", C.c_int32),
                ("indices", ip), ("prim_colors", fp), ("user_id", C.c_int32), ("_pad", C.c_int32),
                ("material", Material)]


class SpheresDesc(C.Structure):
    _fields_ = [("centers", fp), ("radii", fp), ("colors", fp), ("num_spheres", C.c_int32),
                ("default_radius", C.c_float), ("user_id", C.c_int32), ("_pad", C.c_int32),
                ("material", Material)]


class CylindersDesc(C.Structure):
    _fields_ = [("vertices", fp), ("indices", ip), ("radii", fp), ("colors", fp), ("num_vertices", C.c_int32),
                ("num_cylinders", C.c_int32), ("default_radius", C.c_float), ("user_id", C.c_int32),
                ("material", Material)]


class CapsulesDesc(C.Structure):
    _fields_ = [("vertices", fp), ("indices", ip), ("colors", fp), ("num_vertices", C.c_int32),
                ("num_capsules", C.c_int32), ("user_id", C.c_int32), ("_pad", C.c_int32), ("material", Material)]


class ConesDesc(C.Structure):
    _fields_ = [("vertices", fp), ("radii", fp), ("indices", ip), ("colors", fp), ("num_vertices", C.c_int32),
                ("num_cones", C.c_int32), ("user_id", C.c_int32), ("_pad", C.c_int32), ("material", Material)]


class Camera(C.Structure):
    """bnx_camera: the perspective fields by name; an orthographic or omni
    camera (bnx_camera_orthographic / bnx_camera_omni) reuses the same 14
    floats as its own fields"""
    _fields_ = [("type", C.c_int32), ("dir_00", f3), ("dir_du", f3), ("dir_dv", f3), ("lens_00", f3),
                ("aperture_radius", C.c_float), ("focus_distance", C.c_float)]


class Renderer(C.Structure):
    _fields_ = [("bg_color", f4), ("ambient_radiance", C.c_float), ("cut_plane", f4), ("flags", C.c_uint32),
                ("max_generations", C.c_int32), ("_pad", C.c_int32)]


class DirLight(C.Structure):
    _fields_ = [("direction", f3), ("radiance", C.c_float), ("color", f3)]


class EnvMapDesc(C.Structure):
    _fields_ = [("radiance", fp), ("width", C.c_int32), ("height", C.c_int32), ("direction", f3), ("up", f3),
                ("scale", C.c_float), ("_pad", C.c_int32)]


def env_map_desc(radiance, direction=(1.0, 0.0, 0.0), up=(0.0, 0.0, 1.0), scale=1.0):
    """(EnvMapDesc, keep-alive array) for an (height, width, 3) float32
    radiance image (row 0 at theta = 0)."""
    rad = np.ascontiguousarray(radiance, dtype=np.float32)
    if rad.ndim != 3 or rad.shape[2] != 3:
        raise ValueError("env map radiance must be (height, width, 3)")
    d = EnvMapDesc(fptr(rad), rad.shape[1], rad.shape[0], (C.c_float * 3)(*direction), (C.c_float * 3)(*up),
                   float(scale), 0)
    return d, rad


class StructuredDesc(C.Structure):
    _fields_ = [("scalars", fp), ("dims", C.c_int32 * 3), ("cells_per_mc", C.c_int32), ("grid_origin", f3),
                ("grid_spacing", f3)]


class TransferFunction(C.Structure):
    _fields_ = [("values", fp), ("num_values", C.c_int32), ("domain", C.c_float * 2), ("base_density", C.c_float)]


def structured_desc(vs):
    """(StructuredDesc, TransferFunction, keep-alive arrays) for a
    scenes.StructuredVolumeScene."""
    sc = np.ascontiguousarray(vs.scalars, dtype=np.float32)
    tf = np.ascontiguousarray(vs.tf, dtype=np.float32)
    d = StructuredDesc(fptr(sc), (C.c_int32 * 3)(*vs.dims), vs.cells_per_mc, f3(*vs.grid_origin),
                       f3(*vs.grid_spacing))
    t = TransferFunction(fptr(tf), len(tf), (C.c_float * 2)(*vs.tf_domain), vs.base_density)
    return d, t, (sc, tf)


class UMeshDesc(C.Structure):
    _fields_ = [("vertices", fp), ("scalars", fp), ("indices", ip), ("cell_offsets", ip),
                ("cell_types", C.POINTER(C.c_uint8)), ("num_vertices", C.c_int32), ("num_indices", C.c_int32),
                ("num_cells", C.c_int32), ("mc_dims", C.c_int32 * 3)]


def umesh_desc(us):
    """(UMeshDesc, TransferFunction, keep-alive arrays) for a
    scenes.UMeshVolumeScene: its cells (indices, offsets, types) when it has
    them, else one tet per cell, cell c at index 4c with no type array."""
    v = np.ascontiguousarray(us.vertices, dtype=np.float32)
    sc = np.ascontiguousarray(us.scalars, dtype=np.float32)
    types = None
    if getattr(us, "cells", None) is not None:
        idx = np.ascontiguousarray(us.cells[0], dtype=np.int32)
        offs = np.ascontiguousarray(us.cells[1], dtype=np.int32)
        types = np.ascontiguousarray(us.cells[2], dtype=np.uint8)
    else:
        idx = np.ascontiguousarray(us.tets, dtype=np.int32).ravel()
        offs = (4 * np.arange(len(us.tets), dtype=np.int64)).astype(np.int32)
    n = len(offs)
    tf = np.ascontiguousarray(us.tf, dtype=np.float32)
    tp = types.ctypes.data_as(C.POINTER(C.c_uint8)) if types is not None else None
    d = UMeshDesc(fptr(v), fptr(sc), iptr(idx), iptr(offs), tp, len(v), len(idx), n,
                  (C.c_int32 * 3)(*us.mc_dims))
    t = TransferFunction(fptr(tf), len(tf), (C.c_float * 2)(*us.tf_domain), us.base_density)
    return d, t, (v, sc, idx, offs, tf, types)


class RenderStats(C.Structure):
    _fields_ = [("rays_traced", C.c_int64), ("shadow_rays", C.c_int64), ("generations", C.c_int64),
                ("volume_samples", C.c_int64), ("ms_total", C.c_double), ("node_visits", C.c_int64),
                ("tri_tests", C.c_int64), ("sphere_tests", C.c_int64), ("trace_ms", C.c_double),
                ("trace_launches", C.c_int64), ("march_ms", C.c_double), ("march_launches", C.c_int64),
                ("mc_cells", C.c_int64), ("march_rays", C.c_int64), ("umesh_nodes", C.c_int64),
                ("umesh_tets", C.c_int64), ("captured_batches", C.c_int64), ("uncaptured_batches", C.c_int64),
                ("capture_error", C.c_int32), ("reserved", C.c_int32)]


class BvhInfo(C.Structure):
    _fields_ = [("num_nodes", C.c_int32), ("num_prims", C.c_int32), ("max_depth", C.c_int32),
                ("root_is_leaf", C.c_int32), ("root_lower", f3), ("root_upper", f3), ("build_ms", C.c_double)]


class ServerConfig(C.Structure):
    _fields_ = [("frame_memory_budget", C.c_size_t), ("max_paths_in_flight", C.c_int64),
                ("max_active", C.c_int32), ("deadline_slack", C.c_float)]


class Tuning(C.Structure):
    _fields_ = [("raygen_tile_w", C.c_int32), ("raygen_tile_h", C.c_int32), ("trace_stack_lds", C.c_int32),
                ("trace_blocks_per_cu", C.c_int32), ("trace_refill", C.c_int32), ("trace_walk", C.c_int32),
                ("sort_rays", C.c_int32), ("shade_block", C.c_int32),
                ("trace_refill_prims", C.c_int64), ("render_paths", C.c_int64), ("graph_paths", C.c_int64)]


class AutotuneReport(C.Structure):
    _fields_ = [("ms_before", C.c_double), ("ms_after", C.c_double), ("probes", C.c_int32), ("_pad", C.c_int32)]


TUNING_FIELDS = ("raygen_tile_w", "raygen_tile_h", "trace_stack_lds", "trace_blocks_per_cu", "trace_refill", "trace_walk",
                 "sort_rays", "shade_block",
                 "trace_refill_prims", "render_paths", "graph_paths")


class Request(C.Structure):
    _fields_ = [("world", C.c_void_p), ("camera", Camera), ("width", C.c_int32), ("height", C.c_int32),
                ("renderer", Renderer), ("spp", C.c_int32), ("_pad", C.c_int32)]


class RequestStatus(C.Structure):
    _fields_ = [("state", C.c_int32), ("_pad", C.c_int32), ("submit_ms", C.c_double), ("admit_ms", C.c_double),
                ("done_ms", C.c_double), ("paths_started", C.c_int64), ("paths_total", C.c_int64)]


class ServerStats(C.Structure):
    _fields_ = [("submitted", C.c_int64), ("admitted", C.c_int64), ("done", C.c_int64), ("released", C.c_int64),
                ("rays_traced", C.c_int64), ("shadow_rays", C.c_int64), ("volume_samples", C.c_int64),
                ("generations", C.c_int64), ("peak_queue", C.c_int64), ("active", C.c_int32),
                ("queued", C.c_int32), ("frame_bytes", C.c_size_t), ("busy_ms", C.c_double),
                ("trace_ms", C.c_double), ("trace_launches", C.c_int64), ("march_ms", C.c_double),
                ("march_launches", C.c_int64), ("march_rays", C.c_int64), ("node_visits", C.c_int64),
                ("tri_tests", C.c_int64), ("sphere_tests", C.c_int64), ("mc_cells", C.c_int64)]


BNX_REQ_QUEUED, BNX_REQ_ACTIVE, BNX_REQ_DONE, BNX_REQ_FAILED = 0, 1, 2, 3

assert C.sizeof(Material) == 48 and C.sizeof(Camera) == 60 and C.sizeof(DirLight) == 28
assert C.sizeof(Tuning) == 56 and C.sizeof(AutotuneReport) == 24
assert C.sizeof(ServerConfig) == 24 and C.sizeof(Request) == 136 and C.sizeof(RequestStatus) == 48
assert C.sizeof(UMeshDesc) == 64
assert C.sizeof(ServerStats) == 168 and Request.renderer.offset == 76 and Request.spp.offset == 124

# 64-byte BVH node: child boxes lo0,hi0,lo1,hi1 then ref[2], cnt[2]
NODE_DTYPE = np.dtype([("lo0", "<f4", 3), ("hi0", "<f4", 3), ("lo1", "<f4", 3), ("hi1", "<f4", 3),
                       ("ref", "<i4", 2), ("cnt", "<i4", 2)])
assert NODE_DTYPE.itemsize == 64

HIT_DTYPE = np.dtype([("t", "<f4"), ("u", "<f4"), ("v", "<f4"), ("prim_id", "<i4"), ("geom_id", "<i4"),
                      ("inst", "<i4"), ("user_inst_id", "<i4"), ("obj_id", "<i4")])
assert HIT_DTYPE.itemsize == 32


def fptr(a: np.ndarray):
    return a.ctypes.data_as(fp) if a is not None else None


def iptr(a: np.ndarray):
    return a.ctypes.data_as(ip) if a is not None else None


# where a material's color comes from (BNX_ATTRIB_*), geometry attribute
# slots (BNX_ATTRIB_SLOT_*) and scopes (BNX_SCOPE_*)
ATTRIB_NONE, ATTRIB_COLOR, ATTRIB_0 = 0, 1, 2
ATTRIB_WORLD_POSITION, ATTRIB_OBJECT_POSITION, ATTRIB_WORLD_NORMAL, ATTRIB_OBJECT_NORMAL, ATTRIB_PRIMITIVE_ID = \
    6, 7, 8, 9, 10
ATTRIB_SLOT_0, ATTRIB_SLOT_COLOR, ATTRIB_SLOT_NORMAL = 0, 4, 5
SCOPE_NONE, SCOPE_CONSTANT, SCOPE_PER_PRIM, SCOPE_PER_VERTEX, SCOPE_FACE_VARYING = 0, 1, 2, 3, 4


def _color_source(from_attribute) -> int:
    """True: the 'color' attribute; False: the value; an int: a BNX_ATTRIB_* code"""
    if from_attribute is True:
        return ATTRIB_COLOR
    if from_attribute is False or from_attribute is None:
        return ATTRIB_NONE
    return int(from_attribute)


def matte(color=(0.8, 0.8, 0.8, 1.0), from_attribute=True, opacity=1.0) -> Material:
    m = Material()
    m.type = BNX_MATERIAL_MATTE
    m.color_from_attribute = _color_source(from_attribute)
    m.color = f4(*color)
    m.opacity = opacity
    m.metallic, m.roughness, m.transmission, m.ior = 1.0, 1.0, 0.0, 1.45
    return m


def pbr(base_color=(1.0, 1.0, 1.0, 1.0), from_attribute=True, metallic=1.0, roughness=1.0, opacity=1.0,
        transmission=0.0, ior=1.45) -> Material:
    """ANARI physicallyBased material (AnariPBR); defaults are ANARI's"""
    m = Material()
    m.type = BNX_MATERIAL_PBR
    m.color_from_attribute = _color_source(from_attribute)
    m.color = f4(*base_color)
    m.opacity = opacity
    m.metallic, m.roughness, m.transmission, m.ior = metallic, roughness, transmission, ior
    return m


def renderer(bg=(0.0, 0.0, 0.0, 1.0), ambient=1.0, flags=0, max_generations=0, cut_plane=None) -> Renderer:
    """cut_plane (nx, ny, nz, w): the renderer's 'cutPlane'; surface hits with
    dot(n, P) + w < 0 are invisible (None: no cut plane)"""
    r = Renderer()
    r.bg_color = f4(*bg)
    r.ambient_radiance = ambient
    r.cut_plane = f4(0.0, 0.0, 0.0, -1e30) if cut_plane is None else f4(*[float(x) for x in cut_plane])
    r.flags = flags
    r.max_generations = max_generations
    return r


def dir_lights(lights) -> "C.Array":
    arr = (DirLight * max(1, len(lights)))()
    for i, (d, rad, col) in enumerate(lights):
        arr[i].direction = f3(*[float(x) for x in d])
        arr[i].radiance = float(rad)
        arr[i].color = f3(*[float(x) for x in col])
    return arr
