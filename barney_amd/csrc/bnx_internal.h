// bnx_internal.h -- host-side objects behind the C-ABI handles.
#pragma once

#include <hip/hip_runtime.h>

#include <new>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/bnx_rt.h"
#include "bnx_device.h"

namespace bnx {

void set_error(const std::string &msg);

#define BNX_HIP(call)                                                                 \
  do {                                                                                \
    hipError_t e_ = (call);                                                           \
    if (e_ != hipSuccess) {                                                           \
      ::bnx::set_error(std::string(#call) + ": " + hipGetErrorString(e_));             \
      return BNX_ERR_HIP;                                                             \
    }                                                                                 \
  } while (0)

#define BNX_CHECK(cond, code, msg)                                                    \
  do {                                                                                \
    if (!(cond)) {                                                                    \
      ::bnx::set_error(msg);                                                          \
      return code;                                                                    \
    }                                                                                 \
  } while (0)

// exception guard of every C-ABI entry point
#define GUARD_BEGIN try {
#define GUARD_END                                                                     \
  }                                                                                   \
  catch (const std::bad_alloc &) {                                                    \
    ::bnx::set_error("host allocation failed");                                       \
    return BNX_ERR_OOM;                                                               \
  }                                                                                   \
  catch (const std::exception &e) {                                                   \
    ::bnx::set_error(e.what());                                                       \
    return BNX_ERR_INVALID;                                                           \
  }

// device memory owned by a C++ object, or (borrowed) a fixed slice of
// memory owned elsewhere that is never freed or grown
struct DevMem {
  void *ptr = nullptr;
  size_t bytes = 0;
  bool borrowed = false;
  ~DevMem() { release(); }
  void release()
  {
    if (ptr && !borrowed) (void)hipFree(ptr);
    ptr = nullptr;
    bytes = 0;
  }
  // (re)allocate if smaller than needed; contents are not preserved.  A
  // borrowed slice only reports whether it is large enough.
  hipError_t ensure(size_t b)
  {
    if (borrowed) return b <= bytes ? hipSuccess : hipErrorOutOfMemory;
    if (b <= bytes && ptr) return hipSuccess;
    release();
    hipError_t e = hipMalloc(&ptr, b ? b : 16);
    if (e == hipSuccess) bytes = b;
    return e;
  }
  void borrow(void *p, size_t b)
  {
    release();
    ptr = p;
    bytes = b;
    borrowed = true;
  }
  template <typename T> T *as() const { return (T *)ptr; }
};

struct Device;

// scratch arrays of the BVH builder, kept per device and grown on demand
struct BuildScratch {
  DevMem lo, hi, order, order_tmp, slot_task, flags, scan, cub_tmp;
  DevMem tasks, tasks_next, bins, big_map, small_tasks, counters, bounds, split;
};

struct Device {
  int gpu = 0;
  hipStream_t stream = nullptr;
  int num_cus = 0;
  int cyl_worlds = 0;            // built instance groups with a cylinders BLAS (trace kernels with the cylinder test)
  int xf_worlds = 0;             // built instance groups with a non-identity instance (trace kernels that change frame)
  int translucent_worlds = 0;    // built instance groups with stochastic-opacity geometry (the EXT trace kernels)
  int wall_clock_khz = 100000;   // rate of the device wall clock (launch timing in graphs)
  // launch tuning (bnx_device_set_tuning, bnx_autotune); none of it changes results
  int raygen_tile_w = 8, raygen_tile_h = 8;   // primary-ray pixel order: tiles (0: row-major)
  int trace_stack_lds = 16;  // LDS stack entries per lane in the queue-trace kernels (8, 16, 32)
  int trace_walk = 0;        // one-ray-per-thread walk: 0 node step or leaf per trip, 1 speculative while-while
  int sort_rays = 0;         // bnx_render: reorder secondary generations by direction and origin
  int shade_block = 1024;    // threads per shade block (256 or 1024)
  int trace_refill = 16;    // secondary rays of large scenes: idle lanes that trigger a trace-pool refill
  long long trace_refill_prims = 1 << 20;   // "large": at least this many primitives
  int trace_blocks_per_cu = 8;   // trace launch size: blocks per CU
  long long render_paths = 8LL << 20;       // bnx_render: paths per batch
  long long graph_paths = 4LL << 20;        // bnx_render: batches this small run a captured bounce loop
  BuildScratch build;
  DevMem counters;      // small device counters (trace work queue etc.)
  void *queues = nullptr;   // render queues (bnx_render.h Queues)
  // bnx_render_noalloc: this device's copy of the library's static scratch
  // array (its address from hipGetSymbolAddress) and the queues laid out in it
  void *noalloc_scratch = nullptr;
  void *noalloc_queues = nullptr;
};

// at device creation, with the device current: the address of this
// device's copy of the static scratch array of bnx_render_noalloc
int noalloc_init(Device *dev);
// the device's render queues (both forms), freed with the device
void noalloc_destroy(Device *dev);

struct Buffer {
  Device *dev;
  DevMem mem;
};

// one geometry attribute as set (bnx_geom_set_attribute); w_below_1: some
// value it can give has w < 1 (a stochastic-opacity alpha source)
struct HostAttr {
  int scope = BNX_SCOPE_NONE;
  int comps = 0;
  float value[4] = {0.f, 0.f, 0.f, 1.f};
  DevMem data;
  bool w_below_1 = false;
};

struct Geom {
  Device *dev;
  int type;             // 0 triangles, 1 spheres, 2 cylinders
  int user_id;
  bnx_material mat;
  int num_prims;
  int num_vertices;
  DevMem vertices, indices, colors, spheres, cyls;
  DevMem normals;       // triangles: per-vertex normals
  HostAttr attr[6];     // attribute0..3, color, normal
};

struct Bvh {
  DevMem nodes;         // Node[num_nodes] (capacity max(1,n))
  DevMem order;         // int[num_prims]
  DevMem nodes4;        // Node4[num_nodes4]: the 4-wide traversal form (collapse_bvh4)
  int num_nodes4 = 0;
  int stack_need4 = 0;  // most traversal-stack entries a root-to-leaf walk can push
  int num_nodes = 0;
  int num_prims = 0;
  int max_depth = 0;
  float root_lo[3] = {0, 0, 0}, root_hi[3] = {0, 0, 0};
  double build_ms = 0.0;
};

enum GroupKind { GROUP_TRIANGLES = 0, GROUP_USER = 1, GROUP_INSTANCES = 2 };

struct Group {
  Device *dev;
  int kind;
  std::vector<Geom *> geoms;
  // BLAS data
  Bvh bvh;
  DevMem prims, refs;   // BVH-ordered primitive records
  bool built = false;
  // instance-group data
  std::vector<Group *> groups;
  std::vector<int> inst_ids;
  std::vector<float> xfms;   // 12 floats per instance, owl affine3f layout
  DevMem inst_attr[5];       // instance attributes (attribute0..3, color): float4 per instance
  bool inst_attr_w_below_1[5] = {false, false, false, false, false};
  DevMem d_insts, d_blas, d_geoms;
  SceneDev scene{};
  int max_blas_depth = 0;    // largest traversal-stack need of its BLASes (4-wide walk)
  long long blas_prims = 0;  // primitives in the distinct BLASes it instances
  bool has_cyl = false;      // counted in Device::cyl_worlds
  bool has_xf = false;       // counted in Device::xf_worlds
  bool has_translucent = false;   // counted in Device::translucent_worlds
};

struct Volume;
struct IsoSurface;
struct World {
  Device *dev;
  Group *instances;
  std::vector<bnx_dir_light> dir_lights;
  DevMem d_dir_lights;
  std::vector<Volume *> volumes;
  std::vector<int> volume_inst_ids;
  DevMem d_volumes;           // VolumeDev[volumes.size()]
  std::vector<IsoSurface *> isos;
  std::vector<int> iso_inst_ids;
  DevMem d_isos;              // IsoDev[isos.size()]
  // versions of the volumes (and isosurface fields) last uploaded; empty
  // until the first upload
  std::vector<uint64_t> vol_seen, iso_seen;
  // HDRI environment light (no map while env_W == 0)
  DevMem env_tex;             // float4[W * H]
  DevMem env_cdf_x, env_cdf_y;
  int env_W = 0, env_H = 0;
  float env_scale = 1.f;
  float env_to_world[9], env_to_local[9];
};

struct Frame {
  Device *dev;
  int w, h;
  int accum_id = 0;
  DevMem accum;         // int64 x 4 per pixel, 2^-32 fixed point
  DevMem depth, prim, inst, obj, normal;
  DevMem out;           // staging for reads
  // denoiser hook (bnx_frame_set_denoiser) and its device images: linear
  // colour in, normals in (float3), colour out
  bnx_denoise_fn denoise = nullptr;
  void *denoise_user = nullptr;
  bool denoise_fade = false;
  DevMem dn_in, dn_normal, dn_out;
};

// builder entry (bnx_bvh.hip): boxes in d_lo/d_hi (float4, n); returns host info
int build_bvh(Device *dev, const float4 *d_lo, const float4 *d_hi, int n, int max_leaf, Bvh &out);
// collapse the built binary BVH into 4-wide nodes (out.nodes4)
int collapse_bvh4(Device *dev, Bvh &bvh);
// trace-pool refill threshold for the secondary rays of a world: refilling
// lanes as their walks end pays off when walks are long (large BVHs); on
// small scenes whole-wave refills keep rays together and cost less
int trace_refill_for(const Device *dev, const World *w);

// traversal entry (bnx_trace.hip)
struct TraceQueue {
  float4 *org;          // org.xyz, tmax (tmax replaced by hit t on hit)
  float4 *dir;          // dir.xyz, flags (bit 0: shadow ray -> occlusion query)
  int4 *hit;            // prim, geom, inst, bits(u)
  float *hit_v;         // v
  const uint64_t *rng;  // RNGSeed state (stochastic opacity)
  const uint4 *ps;      // path state (generation, pixel: ID channels of translucent hits)
};
int launch_trace_api(Group *g, const bnx_ray *d_rays, int n, bnx_hit *d_hits, uint32_t flags);
struct SlotDev;
struct SegView;
// trace every ray of the generation's segments: one ray per thread
// (blocks256 blocks of 256 threads) or, pooled, blocksT blocks drawing rays
// from their segment's pool; cut: a request in the launch has a cut plane
// (the launch then takes the kernels that apply it)
int launch_trace_queue(Device *dev, const SlotDev *d_slots, const SegView &sv, int blocks256, int blocksT,
                       bool pooled, bool cut, const TraceQueue &q, unsigned long long *counters);
// whether a trace launch takes the extended kernels (cylinder test, cut
// plane, stochastic opacity): a built world holds cylinders or translucent
// triangles, or a request of the launch has a cut plane
inline bool ext_kernels(const Device *dev, bool cut)
{
  return dev->cyl_worlds > 0 || dev->translucent_worlds > 0 || cut;
}

} // namespace bnx
