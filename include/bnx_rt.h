/* bnx_rt.h -- C-ABI of the MI355X (gfx950) ray-tracing + volume backend.
 *
 * This is the drop-in boundary underneath barney's `rtcore` backend interface
 * (namespace BARNEY_NS::rtc).  Every entry point below names the reference
 * interface it replaces (paths relative to the barney v1.4.0 source tree).
 * All functions return BNX_OK (0) on success or a negative BNX_ERR_* code;
 * bnx_last_error() returns a message for the most recent failure on the
 * calling thread.  A C++ adapter converts codes to std::runtime_error, which
 * is what BARNEY_RAISE does in rtcore/cudaCommon/cuda-helper.h:9-42.
 *
 * Threading: one host thread drives a bnx_device; every call is enqueued on
 * that device's single HIP stream (rtcore/cudaCommon/Device.cu:109-122).
 * Pointers named d_* are device pointers (from bnx_buffer_dptr); every other
 * pointer is a host pointer.  All structs are plain-old-data with the sizes
 * asserted below.
 */
#ifndef BNX_RT_H
#define BNX_RT_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define BNX_ABI_VERSION 1

/* ---------------------------------------------------------------- errors */
#define BNX_OK               0
#define BNX_ERR_INVALID     -1  /* bad handle / argument                       */
#define BNX_ERR_HIP         -2  /* HIP runtime error                           */
#define BNX_ERR_OOM         -3  /* device allocation failed / over budget      */
#define BNX_ERR_UNSUPPORTED -4  /* feature not provided by this backend         */
#define BNX_ERR_STATE       -5  /* call out of order (e.g. trace before build) */

/* Message for the last failed call on this thread (never NULL). */
const char *bnx_last_error(void);
int bnx_abi_version(void);

/* ------------------------------------------------------------- handles */
typedef struct bnx_device_s *bnx_device;
typedef struct bnx_buffer_s *bnx_buffer;
typedef struct bnx_geom_s   *bnx_geom;
typedef struct bnx_group_s  *bnx_group;
typedef struct bnx_world_s  *bnx_world;
typedef struct bnx_frame_s  *bnx_frame;
typedef struct bnx_volume_s *bnx_volume;
typedef struct bnx_isosurface_s *bnx_isosurface;

/* ------------------------------------------------------------- records */
/* One ray of a trace batch: traceRays.dev.cu:38-50 passes (org, dir, 0, tMax). */
typedef struct bnx_ray {
  float org[3];
  float tmin;
  float dir[3];
  float tmax;
} bnx_ray;

/* Closest hit of one ray.  prim_id is local to its geometry (Triangles.dev.cu:
 * 31, Spheres.dev.cu:181); geom_id is the geometry's index inside its group;
 * inst is the rtc instance index and user_inst_id the user instance ID it maps
 * to (rtcore/cpu/TraceInterface.h:107-108); obj_id is the geometry's userID
 * (Triangles.dev.cu:57).  For triangles (u,v) are the barycentrics; for spheres
 * u is the hit distance measured from the moved origin (Spheres.dev.cu:132-166)
 * and v is 0.  A miss has prim_id == -1 and t == the ray's tmax. */
typedef struct bnx_hit {
  float t;
  float u;
  float v;
  int32_t prim_id;
  int32_t geom_id;
  int32_t inst;
  int32_t user_inst_id;
  int32_t obj_id;
} bnx_hit;

/* ----------------------------------------------------------- materials */
#define BNX_MATERIAL_MATTE 0        /* AnariMatte -> Lambertian (material/AnariMatte.h:51-76) */
#define BNX_MATERIAL_PBR   1        /* AnariPBR -> NVisii (material/AnariPBR.h:62-96,
                                       packedBSDF/NVisii.h)                        */

/* Where a material's color comes from (PossiblyMappedParameter, material/
 * HostMaterial.h:100-114, HitAttributes::get, render/HitAttributes.h:61-80):
 * its value, or a hit attribute.  Positions read (x, y, z, 1); the reference
 * leaves the normals and the primitive ID unimplemented there, reading
 * (0, 0, 0, 1). */
#define BNX_ATTRIB_NONE            0
#define BNX_ATTRIB_COLOR           1
#define BNX_ATTRIB_0               2   /* .. BNX_ATTRIB_0 + 3: attribute0..3   */
#define BNX_ATTRIB_WORLD_POSITION  6
#define BNX_ATTRIB_OBJECT_POSITION 7
#define BNX_ATTRIB_WORLD_NORMAL    8
#define BNX_ATTRIB_OBJECT_NORMAL   9
#define BNX_ATTRIB_PRIMITIVE_ID   10

typedef struct bnx_material {
  int32_t type;                     /* BNX_MATERIAL_*                              */
  int32_t color_from_attribute;     /* BNX_ATTRIB_*: the hit attribute the color
                                       (with its alpha) is read from; 0: 'color' */
  float   color[4];                 /* matte 'color' / PBR 'baseColor' when not
                                       from the attribute                          */
  float   opacity;                  /* 'opacity'                                   */
  float   metallic;                 /* PBR 'metallic' (ANARI default 1)            */
  float   roughness;                /* PBR 'roughness' (default 1), clamped to
                                       [0.001, 0.9] as AnariPBR does               */
  float   transmission;             /* PBR 'transmission' (default 0)              */
  float   ior;                      /* PBR 'ior' (default 1.45)                    */
  float   _pad;
} bnx_material;

/* ---------------------------------------------------------- geometries */
/* Triangles: rtc::Geom::setVertices/setIndices + Triangles::DD
 * (geometry/Triangles.h:26-32, native/geometry/Triangles.cpp:50-82). */
typedef struct bnx_triangles_desc {
  const float   *vertices;          /* float3[num_vertices]                        */
  int32_t        num_vertices;
  int32_t        num_triangles;
  const int32_t *indices;           /* int3[num_triangles]                         */
  const float   *prim_colors;       /* float3[num_triangles] or NULL: the "color"
                                       attribute per primitive                     */
  int32_t        user_id;           /* objID reported in hit IDs                   */
  int32_t        _pad;
  bnx_material   material;
} bnx_triangles_desc;

/* Spheres: Spheres::DD (geometry/Spheres.h:15-20). */
typedef struct bnx_spheres_desc {
  const float   *centers;           /* float3[num_spheres]                         */
  const float   *radii;             /* float[num_spheres] or NULL                  */
  const float   *colors;            /* float3[num_spheres] or NULL                 */
  int32_t        num_spheres;
  float          default_radius;
  int32_t        user_id;
  int32_t        _pad;
  bnx_material   material;
} bnx_spheres_desc;

/* Cylinders: Cylinders::DD (native/geometry/Cylinders.h:20-27), capped
 * cylinders between two vertices each, intersected as CylindersPrograms::
 * intersect does (native/geometry/Cylinders.dev.cu:107-260).  colors: the
 * "color" attribute per cylinder (bnx_geom_set_attribute sets attributes per
 * vertex, interpolated along the axis). */
typedef struct bnx_cylinders_desc {
  const float   *vertices;          /* float3[num_vertices]                        */
  const int32_t *indices;           /* int2[num_cylinders]: end points             */
  const float   *radii;             /* float[num_cylinders] or NULL                */
  const float   *colors;            /* float3[num_cylinders] or NULL               */
  int32_t        num_vertices;
  int32_t        num_cylinders;
  float          default_radius;
  int32_t        user_id;
  bnx_material   material;
} bnx_cylinders_desc;

/* Capsules: Capsules::DD (native/geometry/Capsules.h:30-34), rounded cones
 * between two vertices each, every vertex with its own radius, intersected
 * as CapsulesPrograms::intersect does (native/geometry/Capsules.dev.cu:
 * 19-300: the origin moved to 0.99 of the box entry, the body and both end
 * spheres; stochastic opacity when the material's alpha can be below 1).
 * colors: the "color" attribute per capsule. */
typedef struct bnx_capsules_desc {
  const float   *vertices;          /* float4[num_vertices]: position, radius    */
  const int32_t *indices;           /* int2[num_capsules]: end points             */
  const float   *colors;            /* float3[num_capsules] or NULL               */
  int32_t        num_vertices;
  int32_t        num_capsules;
  int32_t        user_id;
  int32_t        _pad;
  bnx_material   material;
} bnx_capsules_desc;

/* Cones: Cones::DD (native/geometry/Cones.h:20-24), capped cones between two
 * vertices each with a radius per vertex, intersected as ConesPrograms::
 * intersect does (native/geometry/Cones.dev.cu:60-266).  Their BVH boxes
 * are the boxes of both end spheres: ConesPrograms::bounds (Cones.dev.cu:
 * 21-42) takes the second end's lower corner from the first radius
 * (pb - ra), a box that need not hold the cone, and a hit outside its box
 * would be found or not depending on traversal order.  indices NULL: cone i
 * joins vertices 2i and 2i+1. */
typedef struct bnx_cones_desc {
  const float   *vertices;          /* float3[num_vertices]                       */
  const float   *radii;             /* float[num_vertices]                        */
  const int32_t *indices;           /* int2[num_cones] or NULL                    */
  const float   *colors;            /* float3[num_cones] or NULL                  */
  int32_t        num_vertices;
  int32_t        num_cones;
  int32_t        user_id;
  int32_t        _pad;
  bnx_material   material;
} bnx_cones_desc;

/* Camera::DD (native/Camera.h:30-65): a type and, by type, the fields of
 * the perspective camera (with depth of field when aperture_radius > 0),
 * the orthographic camera or the omnidirectional camera, as their commit()
 * methods write them (native/Camera.cu:82-101, 168-181, 230-242).  Fill it
 * with bnx_camera_perspective / _orthographic / _omni. */
#define BNX_CAMERA_PERSPECTIVE  1
#define BNX_CAMERA_ORTHOGRAPHIC 2
#define BNX_CAMERA_OMNI         3
typedef struct bnx_camera {
  int32_t type;
  union {
    struct {                        /* BNX_CAMERA_PERSPECTIVE                      */
      float dir_00[3];
      float dir_du[3];
      float dir_dv[3];
      float lens_00[3];
      float aperture_radius;
      float focus_distance;
    };
    struct {                        /* BNX_CAMERA_ORTHOGRAPHIC                     */
      float org_00[3];
      float org_du[3];
      float org_dv[3];
      float dir[3];
      float height;
      float aspect;
    } ortho;
    struct {                        /* BNX_CAMERA_OMNI: affine3f toWorld           */
      float to_world[9];            /* columns vx, vy, vz                          */
      float position[3];
    } omni;
  };
} bnx_camera;

/* Renderer::DD + the renderer constants of kernels/shadeRays.cu:17-34. */
#define BNX_RENDER_COMPAT_SPHERE_INSTID13 1u  /* reproduce Spheres.dev.cu:193-194 */
/* Shadow rays through volumes estimate their transmittance by ratio tracking
 * (Novak et al. 2014) and carry it as a weight, instead of the reference's
 * delta tracking (volume/Volume.h:168-203), whose 0/1 outcome has the same
 * expectation and more variance.  Off by default: with it frames differ
 * from the reference's sample by sample, not in expectation. */
#define BNX_RENDER_RATIO_TRACKING         2u
typedef struct bnx_renderer {
  float    bg_color[4];             /* 'background' (anari/Renderer.cpp:29)        */
  float    ambient_radiance;        /* 'ambientRadiance' (anari/Renderer.cpp:24)   */
  float    cut_plane[4];            /* disabled when cut_plane[3] <= -1e28         */
  uint32_t flags;                   /* BNX_RENDER_*                                */
  int32_t  max_generations;         /* host loop bound; 0 = MAX_PATH_DEPTH+1        */
  int32_t  _pad;
} bnx_renderer;

/* DirLight::DD (light/DirLight.h:13-17): 'direction' is the emission direction. */
typedef struct bnx_dir_light {
  float direction[3];
  float radiance;
  float color[3];
} bnx_dir_light;

/* HDRI environment light: EnvMapLight (light/EnvMap.h:27-90, EnvMap.cu:
 * 165-239) set up as anari/light/HDRI.cpp:21-80 does.  'radiance' holds
 * width x height float3 texels, row-major; u = column / width runs over phi
 * in [0, 2 pi) and wraps, v = row / height runs over theta in [0, pi] and
 * clamps, with bilinear filtering.  The map's frame: toWorld.vz = -up,
 * toWorld.vy = -normalize(cross(vz, direction)), toWorld.vx =
 * normalize(cross(vy, vz)).  The light replaces the renderer's ambient
 * radiance as the environment: primary rays that miss show it, light
 * sampling draws from its importance CDFs, and its pdf enters the MIS
 * weights. */
typedef struct bnx_env_map_desc {
  const float *radiance;
  int32_t width, height;
  float direction[3];               /* 'direction' (default 1, 0, 0)               */
  float up[3];                      /* 'up' (default 0, 0, 1)                      */
  float scale;                      /* 'scale' (default 1)                         */
  int32_t _pad;
} bnx_env_map_desc;

/* StructuredData (volume/StructuredData.h:30-74, set3i "dims", set3f
 * "gridOrigin"/"gridSpacing"): scalars at dims[0]*dims[1]*dims[2] grid
 * vertices, x fastest; cells_per_mc is the macrocell edge in cells
 * (StructuredData.cu:26 hard-codes 8). */
typedef struct bnx_structured_desc {
  const float *scalars;
  int32_t      dims[3];
  int32_t      cells_per_mc;
  float        grid_origin[3];
  float        grid_spacing[3];
} bnx_structured_desc;

/* TransferFunction (volume/TransferFunction.h:20-117): num_values RGBA
 * entries over [domain[0], domain[1]]; base_density = 1/unitDistance
 * (anari/Volume.cpp:101-106). */
typedef struct bnx_transfer_function {
  const float *values;              /* float4[num_values]                          */
  int32_t      num_values;
  float        domain[2];
  float        base_density;
} bnx_transfer_function;

/* UMeshField (umesh/common/UMeshField.h:33-168, anari "vertex.position",
 * "vertex.data", "index", "cell.index", "cell.type"): per-vertex scalars,
 * each cell's vertex indices starting at cell_offsets[cell].  Cell types
 * (UMeshField.h:17-29, ANARI or VTK codes; NULL cell_types means all tets):
 * tetrahedra (4 vertices), hexahedra (8), prisms (6) and pyramids (5), in
 * VTK vertex order; polyhedra (VTK 42) and other codes return
 * BNX_ERR_UNSUPPORTED.  A tet (v0, v1, v2, v3) contains P when P is on the
 * non-negative side of its four faces as UMeshField::DD::tetScalar orients
 * them (UMeshField.h:391-436); the other cells invert their isoparametric
 * map by Newton's method (ElementIntersection.h:77-476) and contain P when
 * its parametric coordinates lie in the cell's domain (within 1e-6). */
#define BNX_CELL_TET        0
#define BNX_CELL_HEX        1
#define BNX_CELL_PRISM      2
#define BNX_CELL_PYR        3
#define BNX_CELL_VTK_TET   10
#define BNX_CELL_VTK_HEX   12
#define BNX_CELL_VTK_PRISM 13
#define BNX_CELL_VTK_PYR   14
typedef struct bnx_umesh_desc {
  const float   *vertices;          /* float3[num_vertices]                        */
  const float   *scalars;           /* float[num_vertices]                         */
  const int32_t *indices;           /* int[num_indices]                            */
  const int32_t *cell_offsets;      /* int[num_cells]                              */
  const uint8_t *cell_types;        /* uint8[num_cells] or NULL                    */
  int32_t        num_vertices;
  int32_t        num_indices;
  int32_t        num_cells;
  int32_t        mc_dims[3];        /* macrocell grid; 0: 1 + size * (G-1)/maxWidth,
                                       G = 200 + sqrt(cells/1000) (UMeshField.cu:181-186) */
} bnx_umesh_desc;

/* Frame-buffer channels (fb/FrameBuffer.cu:315-394 bnFrameBufferRead). */
#define BNX_FB_COLOR   0   /* float4 per pixel: accum * 1/accumID           */
#define BNX_FB_DEPTH   1   /* float per pixel: primary-ray tMax             */
#define BNX_FB_PRIMID  2   /* int32 per pixel                               */
#define BNX_FB_INSTID  3   /* int32 per pixel                               */
#define BNX_FB_OBJID   4   /* int32 per pixel                               */
#define BNX_FB_NORMAL  5   /* float3 per pixel (first-sample normal)        */
/* the colour channel as 8-bit RGBA, one uint32 per pixel (R in the low
 * byte): BN_UFIXED8_RGBA and BN_UFIXED8_RGBA_SRGB (fb/FrameBuffer.cu:111-130,
 * 284-294): saturated to [0, 1], the sRGB curve on RGB for the second
 * (common/math.h:61-75), then min(255, int(256 x)) per channel */
#define BNX_FB_COLOR_RGBA8 6
#define BNX_FB_COLOR_SRGB8 7

/* Counters of one bnx_render call (device-side counts, read after sync). */
typedef struct bnx_render_stats {
  int64_t rays_traced;              /* sum over generations of rays entering trace */
  int64_t shadow_rays;
  int64_t generations;              /* generation launches that had rays           */
  int64_t volume_samples;           /* density evaluations (volumes)               */
  double  ms_total;                 /* wall time of the call, host clock          */
  int64_t node_visits;              /* BVH inner-node fetches (64 B each)          */
  int64_t tri_tests;                /* triangle tests (48 B records)               */
  int64_t sphere_tests;             /* sphere tests (16 B records)                 */
  double  trace_ms;                 /* summed trace-kernel time (HIP events)       */
  int64_t trace_launches;
  double  march_ms;                 /* summed volume-march kernel time             */
  int64_t march_launches;
  int64_t mc_cells;                 /* macrocells stepped by the volume DDA        */
  int64_t march_rays;               /* rays entering the volume-march kernel       */
  int64_t umesh_nodes;              /* element-BVH nodes visited by tet point queries */
  int64_t umesh_tets;               /* tets tested by tet point queries            */
  int64_t captured_batches;         /* batches whose generations replayed a captured hipGraph */
  int64_t uncaptured_batches;       /* device-driven batches enqueued directly: capture is
                                       unavailable on this device (capture_error)  */
  int32_t capture_error;            /* hipError_t of the device's failed capture, 0 if none */
  int32_t reserved;
} bnx_render_stats;

/* BVH summary of one BLAS / TLAS (for tests and the byte model). */
typedef struct bnx_bvh_info {
  int32_t num_nodes;                /* inner nodes (64 B each)                     */
  int32_t num_prims;
  int32_t max_depth;
  int32_t root_is_leaf;
  float   root_lower[3];
  float   root_upper[3];
  double  build_ms;
} bnx_bvh_info;

/* ------------------------------------------------------------- device */
/* rtc::physicalDeviceCount (rtcore/cuda/Device.cu:14-19) */
int bnx_device_count(int *count);
/* rtc::Device(gpuID) -- rtcore/cudaCommon/Device.h, rtcore/cuda/Device.h:23-56 */
int bnx_device_create(int gpu_id, bnx_device *out);
int bnx_device_destroy(bnx_device dev);
/* rtc::Device::sync (rtcore/cudaCommon/Device.cu) */
int bnx_device_sync(bnx_device dev);
/* copy bytes between host and device memory (either may be either: the
 * runtime tells them apart) on the device's stream, and wait for it (dev
 * NULL: after all work of the current device, for callers that hold no
 * device handle, such as a denoiser behind barney's bn* API); for hooks that
 * are handed device images (bnx_frame_set_denoiser).  No counterpart in the
 * reference, whose hooks own their buffers. */
int bnx_copy(bnx_device dev, void *dst, const void *src, size_t bytes);
/* free / total device memory in bytes (used by memory-based admission) */
int bnx_device_mem_info(bnx_device dev, size_t *free_bytes, size_t *total_bytes);

/* ---------------------------------------------------- device memory
 * The memory methods of rtc::CudaDeviceBase (rtcore/cudaCommon/Device.h:
 * 21-46, implemented in rtcore/cudaCommon/Device.cu): allocMem / freeMem
 * (device memory), allocHost / freeHost (pinned host memory), copyAsync
 * (any direction, on the device's stream; copy = copyAsync + sync),
 * memsetAsync, and setActive / restoreActive (the GPU made current, and
 * the one current before).  Failures return codes instead of raising. */
int bnx_mem_alloc(bnx_device dev, size_t bytes, void **d_ptr);
int bnx_mem_free(bnx_device dev, void *d_ptr);
int bnx_host_alloc(bnx_device dev, size_t bytes, void **h_ptr);
int bnx_host_free(bnx_device dev, void *h_ptr);
int bnx_mem_copy_async(bnx_device dev, void *dst, const void *src, size_t bytes);
int bnx_mem_set_async(bnx_device dev, void *d_ptr, int value, size_t bytes);
int bnx_device_set_active(bnx_device dev, int *previous_gpu);
int bnx_device_restore_active(int previous_gpu);

/* ------------------------------------------------------------- tuning
 * Launch parameters of the render path, per device.  None of them changes a
 * result: they set the order in which primary rays enter the queue, how the
 * traversal stack is split between LDS and private memory, the size of the
 * pooled trace launch and how many paths bnx_render keeps in flight.  No
 * counterpart in the reference, whose launch shapes are fixed (one thread
 * per ray in blocks of 256, kernels/traceRays.cu:20-121; all of a frame's
 * pixels per generation, native/Context.cpp:139-182); bnx_autotune chooses
 * them per scene and request shape. */
typedef struct bnx_tuning {
  int32_t raygen_tile_w;            /* primary rays in tiles of w x h pixels, each   */
  int32_t raygen_tile_h;            /* tile row-major; 0 x 0: the image row-major
                                       (w, h: both 0, or both in 1..256)             */
  int32_t trace_stack_lds;          /* traversal-stack entries per lane kept in LDS:
                                       8, 16 or 32 (of 192; the rest in private
                                       memory)                                       */
  int32_t trace_blocks_per_cu;      /* pooled trace launch: blocks per CU, 1..64     */
  int32_t trace_refill;             /* pooled trace: idle lanes of a wave that make
                                       it draw new rays, 1..64                       */
  int32_t trace_walk;               /* one-ray-per-thread walk: 0 = a node step or a
                                       leaf per trip; 1 = speculative while-while
                                       (leaves set aside until every lane has one)  */
  int32_t sort_rays;                /* bnx_render: 1 = reorder each secondary
                                       generation of 2^16 rays or more by shadow
                                       flag, direction octant and origin Morton
                                       code before tracing it (0: queue order)      */
  int32_t shade_block;              /* threads per shade block: 256 or 1024          */
  int64_t trace_refill_prims;       /* secondary generations of worlds with at least
                                       this many primitives use the pooled trace     */
  int64_t render_paths;             /* bnx_render: paths per batch of samples,
                                       2^16 .. 2^26                                  */
  int64_t graph_paths;              /* bnx_render: a batch of at most this many
                                       paths runs its bounce loop as a captured
                                       hipGraph, one read-back per batch instead
                                       of one per generation (0: never; <= 2^26)    */
} bnx_tuning;
/* a new device starts with 8 x 8 tiles, 16 LDS entries, walk 0, no reordering,
 * 1024-thread shade blocks, 8 blocks per CU,
 * refill 16 for worlds of 2^20 primitives or more, 2^23 paths per batch and
 * captured bounce loops for batches of up to 2^22 paths */

int bnx_device_get_tuning(bnx_device dev, bnx_tuning *out);
/* BNX_ERR_INVALID (and no change) for a field out of range */
int bnx_device_set_tuning(bnx_device dev, const bnx_tuning *tuning);

typedef struct bnx_autotune_report {
  double  ms_before;                /* probe render time with the tuning in place    */
  double  ms_after;                 /* probe render time with the tuning chosen      */
  int32_t probes;                   /* probe renders timed                           */
  int32_t _pad;
} bnx_autotune_report;

/* Choose the device's tuning for requests like this one: coordinate descent,
 * two passes, over the fields of bnx_tuning (raygen tile, LDS stack entries,
 * walk order, ray reordering, shade block, refill threshold, pooled-launch size, paths per batch, captured-loop
 * limit), each candidate timed on
 * probe renders of width x height at probe_spp samples into a scratch frame
 * (best of 2 to 20, so that each candidate is timed for about 10 ms; a
 * candidate must be 1% faster to replace the current one).
 * Sets the tuning chosen on the device (the one in place if nothing was
 * faster) and returns it with the timings.  No counterpart in the
 * reference. */
int bnx_autotune(bnx_world world, const bnx_camera *camera, const bnx_renderer *renderer, int width, int height,
                 int probe_spp, bnx_tuning *chosen, bnx_autotune_report *report);

/* ------------------------------------------------------------- buffers */
/* rtc::Device::createBuffer / rtc::Buffer::upload / freeBuffer */
int   bnx_buffer_create(bnx_device dev, size_t bytes, const void *init, bnx_buffer *out);
int   bnx_buffer_upload(bnx_buffer buf, const void *host, size_t bytes, size_t offset);
int   bnx_buffer_download(bnx_buffer buf, void *host, size_t bytes, size_t offset);
int   bnx_buffer_destroy(bnx_buffer buf);
void *bnx_buffer_dptr(bnx_buffer buf);
size_t bnx_buffer_size(bnx_buffer buf);

/* --------------------------------------------------------------- geoms */
/* GeomType::createGeom + Geom::setVertices/setIndices/setDD for the
 * 'Triangles' type (RTC_EXPORT_TRIANGLES_GEOM, Triangles.dev.cu:139). */
int bnx_geom_create_triangles(bnx_device dev, const bnx_triangles_desc *desc, bnx_geom *out);
/* ... and for the 'Spheres' user type (RTC_EXPORT_USER_GEOM, Spheres.dev.cu:205). */
int bnx_geom_create_spheres(bnx_device dev, const bnx_spheres_desc *desc, bnx_geom *out);
/* rtc::Device::freeGeom */
/* GeomType::createGeom + Cylinders::commit (native/geometry/Cylinders.cpp)
 * for the 'Cylinders' user type (RTC_EXPORT_USER_GEOM(Cylinders, ...),
 * Cylinders.dev.cu:266) */
int bnx_geom_create_cylinders(bnx_device dev, const bnx_cylinders_desc *desc, bnx_geom *out);
/* ... Capsules::commit (native/geometry/Capsules.cpp:18-45) for the
 * 'Capsules' user type (RTC_EXPORT_USER_GEOM(Capsules, ...),
 * Capsules.dev.cu:306) */
int bnx_geom_create_capsules(bnx_device dev, const bnx_capsules_desc *desc, bnx_geom *out);
/* ... Cones::commit (native/geometry/Cones.cpp:18-52) for the 'Cones' user
 * type (RTC_EXPORT_USER_GEOM(Cones, ...), Cones.dev.cu:270) */
int bnx_geom_create_cones(bnx_device dev, const bnx_cones_desc *desc, bnx_geom *out);
int bnx_geom_destroy(bnx_geom geom);
/* Geometry attributes (GeometryAttributes, render/GeometryAttributes.h:
 * 32-63; Geometry::DD::setHitAttributes, geometry/Geometry.h:94-139): one
 * of attribute0..3, color or normal (BNX_ATTRIB_SLOT_*) set for the geometry
 * with a scope: a constant value, or per primitive, per vertex (triangles:
 * barycentric interpolation; cylinders, capsules and cones: linear between
 * the two ends; spheres: the sphere's own) or face-varying (3 per triangle,
 * 2 per cylinder, capsule or cone) elements of
 * 1-4 floats each (AttributeArray::DD::valueAt: missing y, z read 0, a
 * missing w reads 1).  BNX_SCOPE_NONE clears it.  The 'normal' attribute
 * replaces a triangle's object normal (Triangles.dev.cu:86-111).  The data
 * are copied; a group built over the geometry sees them after its next
 * bnx_group_build.  count: elements in data. */
#define BNX_ATTRIB_SLOT_0      0
#define BNX_ATTRIB_SLOT_COLOR  4
#define BNX_ATTRIB_SLOT_NORMAL 5
#define BNX_SCOPE_NONE         0   /* GeometryAttribute::INVALID      */
#define BNX_SCOPE_CONSTANT     1
#define BNX_SCOPE_PER_PRIM     2
#define BNX_SCOPE_PER_VERTEX   3
#define BNX_SCOPE_FACE_VARYING 4
int bnx_geom_set_attribute(bnx_geom geom, int slot, int scope, int components, const float *data, int64_t count,
                           const float value[4]);
/* Triangles::DD::normals (geometry/Triangles.h:26-32): per-vertex normals,
 * interpolated and normalized at the hit (Triangles.dev.cu:68-75); NULL
 * clears them. */
int bnx_geom_set_vertex_normals(bnx_geom geom, const float *normals, int num_vertices);

/* -------------------------------------------------------------- groups */
/* rtc::Device::createTrianglesGroup (rtcore/cuda/Group.cu:261-413) */
int bnx_group_create_triangles(bnx_device dev, const bnx_geom *geoms, int num_geoms, bnx_group *out);
/* rtc::Device::createUserGeomsGroup (rtcore/cuda/Group.cu:418-562) */
int bnx_group_create_user(bnx_device dev, const bnx_geom *geoms, int num_geoms, bnx_group *out);
/* rtc::Device::createInstanceGroup(groups, instIDs, xfms) (rtcore/cuda/Group.cu:144-258).
 * xfms: num_instances affine3f, 12 floats each, column-major like owl's
 * affine3f: {vx.x,vx.y,vx.z, vy.x,vy.y,vy.z, vz.x,vz.y,vz.z, p.x,p.y,p.z}. */
int bnx_group_create_instances(bnx_device dev, const bnx_group *groups, const int32_t *inst_ids,
                               const float *xfms, int num_instances, bnx_group *out);
/* New transforms for every instance of a built instance group (12 floats
 * each, as at creation; the count must match), then the instance BVH built
 * again: rtc::InstanceGroup::setTransforms + refitAccel, which is buildAccel
 * (rtcore/cuda/Group.h:21, rtcore/cpu/Group.h:17), as bnUpdateInstanceTransforms
 * (native/barney.cu:351-361, ModelSlot::updateInstanceTransforms,
 * native/ModelSlot.cpp:157-180) uses them.  Worlds over the group render with
 * the new transforms from their next request; not while requests over it are
 * in flight on a server. */
int bnx_group_set_transforms(bnx_group instances, const float *xfms, int num_instances);
/* Instance attributes (World::DD::instanceAttributes, bnSetInstanceAttributes,
 * native/barney.cu): one float4 per instance of an instance group for
 * attribute0..3 or color (slot 0-4); a hit on a geometry that does not set
 * that attribute reads its instance's value (Geometry.h:103-110).  NULL
 * clears the slot.  Worlds over the group see it from their next request. */
int bnx_group_set_instance_attribute(bnx_group instances, int slot, const float *values, int num_instances);
/* rtc::Group::buildAccel: binned-SAH BVH built on the GPU. For an instance
 * group every referenced group must have been built first. */
int bnx_group_build(bnx_group group);
int bnx_group_bvh_info(bnx_group group, bnx_bvh_info *info);
/* Download the BVH: nodes (64 B each, num_nodes) and the prim order
 * (num_prims int32: original prim index for each BVH-order slot). */
int bnx_group_download_bvh(bnx_group group, void *nodes, int32_t *prim_order);
/* rtc::Device::freeGroup */
int bnx_group_destroy(bnx_group group);

/* Binned-SAH build over raw boxes (d_boxes: n x {lo.xyz, hi.xyz} floats).
 * Writes the same node layout as bnx_group_download_bvh. For tests. */
int bnx_bvh_build_boxes(bnx_device dev, const float *d_boxes, int n, int max_leaf,
                        void *nodes_host, int32_t *prim_order_host, bnx_bvh_info *info);

/* --------------------------------------------------------------- trace */
#define BNX_TRACE_CLOSEST 0u
#define BNX_TRACE_ANY     1u   /* occlusion: any opaque hit ends the ray */
/* TraceKernel2D::launch of native::TraceRays (kernels/traceRays.dev.cu:26-54,
 * rtcore/cuda/TraceInterface.h:24-139): closest hit for n rays, lower
 * (inst, geom, prim) wins a tie in t. Asynchronous on the device stream. */
int bnx_trace(bnx_group instances, const bnx_ray *d_rays, int n, bnx_hit *d_hits, uint32_t flags);

/* ------------------------------------------------------------- volumes */
/* Volume over a StructuredData field with its MCVolumeAccel (volume/
 * MCAccelerator.h:155-195): builds the macrocell value ranges
 * (StructuredData.cu:30-96) and majorants (MCGrid.cu:96-141) on the GPU. */
int bnx_volume_create_structured(bnx_device dev, const bnx_structured_desc *field,
                                 const bnx_transfer_function *tf, int32_t user_id, bnx_volume *out);
/* Volume::setXF + VolumeAccel::rebuildMajorantsOnly */
int bnx_volume_set_transfer_function(bnx_volume vol, const bnx_transfer_function *tf);
/* macrocell grid dims and majorants (for tests) */
int bnx_volume_majorants(bnx_volume vol, int32_t dims[3], float *majorants);
int bnx_volume_destroy(bnx_volume vol);
/* Volume over an unstructured (tet) field (UMeshField + UMeshCuBQLSampler +
 * MCVolumeAccel, umesh/mc/UMeshCuBQLSampler.cu:45-78, UMeshField.cu:168-197):
 * builds on the GPU the element BVH (binned SAH, leaf <= 4), the macrocell
 * value ranges by rasterizing each tet split 5 times on its longest edge
 * (rasterTet<5>, UMeshField.cu:28-107) and every other cell's box
 * (UMeshField.cu:149-152), and the majorants.  Sampling is a point query:
 * the first cell containing P in a depth-first walk of the element BVH
 * (child 0 first, leaf cells in order). */
int bnx_volume_create_umesh(bnx_device dev, const bnx_umesh_desc *field, const bnx_transfer_function *tf,
                            int32_t user_id, bnx_volume *out);
/* macrocell value ranges (lo, hi per cell; for tests) */
int bnx_volume_mc_ranges(bnx_volume vol, int32_t dims[3], float *lo, float *hi);

/* IsoSurface geometry (geometry/IsoSurface.h:31-42, anari "isovalue",
 * "field") over the field of a volume, with its MCIsoSurfaceAccel
 * (volume/MCAccelerator.h:325-548): DDA over the field's macrocell value
 * ranges, 10 sub-steps per overlapping cell, central-difference normal.
 * Only opaque matte materials are provided. */
int bnx_isosurface_create(bnx_volume field, float iso_value, const bnx_material *material, int32_t user_id,
                          bnx_isosurface *out);
int bnx_isosurface_destroy(bnx_isosurface iso);

/* --------------------------------------------------------------- world */
/* World::DD (render/World.h:32-47): instances + lights. */
int bnx_world_create(bnx_device dev, bnx_group instances, const bnx_dir_light *dir_lights,
                     int num_dir_lights, bnx_world *out);
/* Adds a volume as a user instance (ModelSlot.cpp:128-155 puts volume
 * groups after the geometry groups of an instance).  Volumes are marched
 * after the surface trace, in the order they were added. */
int bnx_world_add_volume(bnx_world world, bnx_volume vol, int32_t user_inst_id);
/* Adds an isosurface as a user instance; isosurfaces are intersected after
 * the surface trace and before the volumes, in the order they were added,
 * and act as closest-hit surfaces. */
int bnx_world_add_isosurface(bnx_world world, bnx_isosurface iso, int32_t user_inst_id);
/* Sets (or, with desc NULL or desc->radiance NULL, removes) the world's
 * HDRI environment light (World::DD::envMapLight, render/World.h:32-47):
 * uploads the texture and builds its importance CDFs on the device
 * (EnvMapLight::computeCDFs, light/EnvMap.cu:183-239).  A server's
 * admitted requests hold the map's device memory: do not change the map
 * while requests on this world are in flight. */
int bnx_world_set_env_map(bnx_world world, const bnx_env_map_desc *desc);
int bnx_world_destroy(bnx_world world);

/* -------------------------------------------------------------- camera */
/* PerspectiveCamera::commit (native/Camera.cu:82-101); fovy in degrees.
 * Depth of field: set out->aperture_radius and out->focus_distance
 * ('apertureRadius', 'focusDistance') afterwards. */
int bnx_camera_perspective(const float position[3], const float direction[3], const float up[3],
                           float fovy_deg, bnx_camera *out);
/* OrthographicCamera::commit (native/Camera.cu:168-181): 'height' is the
 * image plane's height in world units, 'aspect' its width over height. */
int bnx_camera_orthographic(const float position[3], const float direction[3], const float up[3], float height,
                            float aspect, bnx_camera *out);
/* OmniCamera::commit (native/Camera.cu:230-242): pixel (u, v) looks along
 * toWorld (cos(2 pi u) sin(pi v), sin(2 pi u) sin(pi v), cos(pi v))
 * (generateRays.cu:24-38, 140-146). */
int bnx_camera_omni(const float position[3], const float direction[3], const float up[3], bnx_camera *out);

/* --------------------------------------------------------------- frame */
/* FrameBuffer/TiledFB resize (fb/TiledFB.cu:249-319) */
int bnx_frame_create(bnx_device dev, int width, int height, bnx_frame *out);
/* bnAccumReset (fb/FrameBuffer.h:147) */
int bnx_frame_reset(bnx_frame fb);
/* bnFrameBufferRead (fb/FrameBuffer.cu:315-394) + linearize (fb/TiledFB.cu:12-111) */
int bnx_frame_read(bnx_frame fb, int channel, void *host);
/* Denoiser hook (rtc::Denoiser, rtcore/cpu/Denoiser.h:16-30, as
 * FrameBuffer::finalizeFrame and readColorChannel use it, fb/FrameBuffer.cu:
 * 196-290).  With a hook set, reading the colour channel (BNX_FB_COLOR or
 * the 8-bit forms) first writes the linear colour (accum x 1/accumID) to
 * in_rgba (float4 per pixel) and the first-sample normals to in_normal
 * (float3 per pixel), copies in_rgba to out_rgba (the alpha passes through,
 * as the OIDN denoiser leaves it), and calls fn with these device images of
 * width x height pixels, row-major, and the device's hipStream_t; the
 * channel is then read from out_rgba.  fn may enqueue its work on the stream
 * or finish it before returning; a nonzero return fails the read.  blend is
 * (accumID - 1) / (accumID + 100) when fade_out is set (fadeOutDenoiser),
 * else 0.  fn NULL removes the hook.  barney's upscale mode (a half-size
 * render upscaled 2x after denoising) is not provided. */
typedef int (*bnx_denoise_fn)(void *user, const float *d_in_rgba, const float *d_in_normal, float *d_out_rgba,
                              int32_t width, int32_t height, float blend, void *stream);
int bnx_frame_set_denoiser(bnx_frame fb, bnx_denoise_fn fn, void *user, int32_t fade_out);
int bnx_frame_accum_id(bnx_frame fb);
int bnx_frame_destroy(bnx_frame fb);

/* -------------------------------------------------------------- render */
/* Context::renderTiles (native/Context.cpp:139-182): spp samples per pixel,
 * each through raygen -> (trace, shade)* generations, accumulated into fb. */
int bnx_render(bnx_world world, const bnx_camera *camera, const bnx_renderer *renderer,
               bnx_frame fb, int spp, bnx_render_stats *stats);

/* bnx_render without allocation: the same frame loop and kernels, with the
 * ray queues, counters and staging in a static device array of the library
 * (24 MB, one copy per GPU, held until the process ends) and no memory
 * allocated or freed (no hipMalloc, hipFree, pinned or library allocation;
 * counters are cleared by a kernel, not a memset).  No counterpart in the
 * reference: for callers that cannot allocate (HIP graph capture set-up
 * around it, serving loops, harnesses that forbid it).  Limits: frames of at
 * most 65536 pixels (BNX_ERR_UNSUPPORTED beyond), samples in batches of up
 * to 65536 paths; no ray reordering, no captured bounce loop and no kernel
 * timing (the stats' *_ms fields are 0).  The world's tables must exist,
 * as they do once its volumes and isosurfaces are added.  Every call on a
 * GPU shares that GPU's array: calls that overlap on one GPU (from two
 * threads) race.  The results equal bnx_render's. */
int bnx_render_noalloc(bnx_world world, const bnx_camera *camera, const bnx_renderer *renderer, bnx_frame fb,
                       int spp, bnx_render_stats *stats);

/* ------------------------------------------------------------- request server
 * Continuous batching of render requests on one device.  barney renders one
 * frame per bnRender call (Context::renderTiles, native/Context.cpp:139-182,
 * one raygen + trace/shade loop per call); the server keeps worlds and BVHs
 * resident and runs ONE wavefront for all requests in flight: every
 * generation traces, marches and shades the rays of all admitted requests in
 * one launch per kernel, and tops the shared queue up with primary rays of
 * admitted requests (shortest remaining work first, or earliest deadline
 * first, see bnx_server_config).  Requests are admitted
 * in arrival order as request slots and frame memory free up.  A request's
 * result is the same as bnx_render of the same world, camera, renderer and
 * spp into a cleared frame (up to the order of fixed-point additions, which
 * does not change the sum).  The worlds named by requests must outlive them.
 * Calls on one server are not thread-safe. */
typedef struct bnx_server_s *bnx_server;

typedef struct bnx_server_config {
  size_t  frame_memory_budget;      /* bytes of request frames held at once (0: 16 GiB) */
  int64_t max_paths_in_flight;      /* paths in the shared queue (0: the device's
                                       bnx_tuning.render_paths at creation); the
                                       queue holds three times as many rays           */
  int32_t max_active;               /* requests holding a slot at once (0: 256, <= 4096) */
  float   deadline_slack;           /* order of the primary top-up: s > 0 =
                                       earliest deadline first, a request's
                                       deadline being its submit time + s x its
                                       estimated time alone (its world's measured
                                       cost per path, from completed requests, x
                                       its paths); 0: s = 2; s < 0: shortest
                                       remaining work first                      */
} bnx_server_config;

typedef struct bnx_request {
  bnx_world    world;
  bnx_camera   camera;
  int32_t      width, height;
  bnx_renderer renderer;
  int32_t      spp;                 /* 1 .. 4096 */
  int32_t      _pad;
} bnx_request;

/* FAILED: admission could not set the request up (bnx_server_step returned
 * the error); the request holds no slot or frame and can only be released */
enum { BNX_REQ_QUEUED = 0, BNX_REQ_ACTIVE = 1, BNX_REQ_DONE = 2, BNX_REQ_FAILED = 3 };

typedef struct bnx_request_status {
  int32_t state;                    /* BNX_REQ_*                                       */
  int32_t _pad;
  double  submit_ms, admit_ms, done_ms;   /* host clock, ms since server creation     */
  int64_t paths_started;            /* primary rays generated so far                   */
  int64_t paths_total;              /* width * height * spp                            */
} bnx_request_status;

typedef struct bnx_server_stats {
  int64_t submitted, admitted, done, released;
  int64_t rays_traced;              /* sum over generations of rays entering trace     */
  int64_t shadow_rays;
  int64_t volume_samples;
  int64_t generations;              /* generation launches that had rays               */
  int64_t peak_queue;               /* largest generation, rays                        */
  int32_t active, queued;
  size_t  frame_bytes;              /* frame memory held (admitted + pooled)           */
  double  busy_ms;                  /* wall time spent inside bnx_server_step          */
  double  trace_ms;                 /* summed trace-kernel time (HIP events)           */
  int64_t trace_launches;
  double  march_ms;                 /* summed volume-march kernel time                 */
  int64_t march_launches;
  int64_t march_rays;               /* rays entering the volume-march kernel           */
  int64_t node_visits, tri_tests, sphere_tests, mc_cells;
} bnx_server_stats;

int bnx_server_create(bnx_device dev, const bnx_server_config *config, bnx_server *out);
int bnx_server_destroy(bnx_server server);
/* queue a request; BNX_ERR_INVALID for bad sizes or a world of another device */
int bnx_server_submit(bnx_server server, const bnx_request *request, int64_t *ticket);
/* one generation for all requests in flight (admission, queue top-up, trace,
 * march, shade); writes up to max_done tickets that completed in it.  With
 * nothing queued or active it returns at once with *num_done = 0. */
int bnx_server_step(bnx_server server, int64_t *done_tickets, int max_done, int *num_done);
/* generations (1..16) generations for all requests in flight with one host
 * read-back at the end instead of one per generation: admission happens
 * before the first, and a kernel does each generation's bookkeeping on the
 * device (the queue top-up in the same priority order, the per-request
 * segments, the output regions).  Requests that finish inside the step are
 * reported at its end.  Results equal bnx_server_step's and bnx_render's.
 * No counterpart in the reference, whose frame loop (Context::renderTiles,
 * native/Context.cpp:139-182) reads back after every generation. */
int bnx_server_step_n(bnx_server server, int generations, int64_t *done_tickets, int max_done, int *num_done);
int bnx_server_poll(bnx_server server, int64_t ticket, bnx_request_status *status);
/* read a channel (BNX_FB_*) of a completed request, as bnx_frame_read */
int bnx_server_read(bnx_server server, int64_t ticket, int channel, void *host);
/* drop a completed request; its frame memory goes back to the server's pool */
int bnx_server_release(bnx_server server, int64_t ticket);
int bnx_server_get_stats(bnx_server server, bnx_server_stats *stats);

#ifdef __cplusplus
} /* extern "C" */
#endif

#ifdef __cplusplus
static_assert(sizeof(bnx_ray) == 32, "bnx_ray layout");
static_assert(sizeof(bnx_hit) == 32, "bnx_hit layout");
static_assert(sizeof(bnx_material) == 48, "bnx_material layout");
static_assert(sizeof(bnx_camera) == 60, "bnx_camera layout");
static_assert(sizeof(bnx_dir_light) == 28, "bnx_dir_light layout");
static_assert(sizeof(bnx_env_map_desc) == 48, "bnx_env_map_desc layout");
static_assert(sizeof(bnx_triangles_desc) == 88, "bnx_triangles_desc layout");
static_assert(sizeof(bnx_spheres_desc) == 88, "bnx_spheres_desc layout");
static_assert(sizeof(bnx_cylinders_desc) == 96, "bnx_cylinders_desc layout");
static_assert(sizeof(bnx_capsules_desc) == 88, "bnx_capsules_desc layout");
static_assert(sizeof(bnx_cones_desc) == 96, "bnx_cones_desc layout");
static_assert(sizeof(bnx_server_config) == 24, "bnx_server_config layout");
static_assert(sizeof(bnx_umesh_desc) == 64, "bnx_umesh_desc layout");
static_assert(sizeof(bnx_request) == 136, "bnx_request layout");
static_assert(sizeof(bnx_request_status) == 48, "bnx_request_status layout");
static_assert(sizeof(bnx_server_stats) == 168, "bnx_server_stats layout");
static_assert(sizeof(bnx_tuning) == 56, "bnx_tuning layout");
static_assert(sizeof(bnx_autotune_report) == 24, "bnx_autotune_report layout");
static_assert(sizeof(bnx_renderer) == 48, "bnx_renderer layout");
static_assert(sizeof(bnx_render_stats) == 152, "bnx_render_stats layout");
#endif

#endif /* BNX_RT_H */
