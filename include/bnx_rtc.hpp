/* bnx_rtc.hpp -- barney's rtcore host interface (namespace BARNEY_NS::rtc,
 * rtcore/cuda/Device.h:12-69, Buffer.h, Geom.h, GeomType.h, Group.h:14-25)
 * mirrored in C++ over the C-ABI of bnx_rt.h, for the rtc adapter of an
 * "amd" backend.  Same class and method names and argument meaning as the
 * reference; every failing call throws std::runtime_error with the
 * library's message, as BARNEY_RAISE does (rtcore/cudaCommon/cuda-helper.h:
 * 9-42).  Header-only; link with libbnx.so.
 *
 * Differences that follow from the boundary: a Buffer keeps a host copy of
 * what was uploaded (the C-ABI builds geometries from host arrays), and a
 * geometry is created in the library when its group is created.  The
 * reference's program model (SBT records, device programs per GeomType) is
 * the library's closed set of types: "Triangles", "Spheres" and
 * "Cylinders". */
#pragma once

#include "bnx_rt.h"

#include <cstdint>
#include <cstring>
#include <stdexcept>
#include <string>
#include <vector>

namespace bnx {
namespace rtc {

struct vec3f {
  float x, y, z;
};
struct vec3i {
  int x, y, z;
};
// owl::common::affine3f: linear part as three column vectors, then the
// translation
struct affine3f {
  vec3f vx, vy, vz, p;
};
using AccelHandle = void *;

inline void check(int rc)
{
  if (rc != BNX_OK) throw std::runtime_error(std::string("bnx: ") + bnx_last_error());
}

// rtc::physicalDeviceCount (rtcore/cuda/Device.cu:14-19)
inline int physicalDeviceCount()
{
  int n = 0;
  check(bnx_device_count(&n));
  return n;
}

struct Device;
struct Group;

// rtc::Buffer (rtcore/cuda/Buffer.h): device memory; upload also keeps the
// bytes on the host
struct Buffer {
  Buffer(Device *device, size_t numBytes, const void *initValues);
  ~Buffer() { bnx_buffer_destroy(h); }
  void *getDD() const { return bnx_buffer_dptr(h); }
  void upload(const void *hostPtr, size_t numBytes, size_t ofs = 0)
  {
    if (ofs + numBytes > host.size()) throw std::runtime_error("bnx: Buffer::upload past the end");
    check(bnx_buffer_upload(h, hostPtr, numBytes, ofs));
    std::memcpy(host.data() + ofs, hostPtr, numBytes);
  }
  void download(void *hostPtr, size_t numBytes, size_t ofs = 0) const
  {
    check(bnx_buffer_download(h, hostPtr, numBytes, ofs));
  }
  size_t size() const { return host.size(); }
  const void *hostData() const { return host.data(); }

  Device *const device;
  bnx_buffer h = nullptr;
  std::vector<uint8_t> host;
};

struct GeomType;

// rtc::Geom (rtcore/cuda/Geom.h:14-45): setVertices/setIndices for
// triangles; setDD carries a user type's data, here the Spheres::DD fields
// (geometry/Spheres.h:15-20) as setSpheres and the Cylinders::DD fields
// (geometry/Cylinders.h:16-22) as setCylinders
struct Geom {
  explicit Geom(GeomType *gt) : gt(gt) {}
  ~Geom()
  {
    if (h) bnx_geom_destroy(h);
  }
  void setPrimCount(int primCount) { this->primCount = primCount; }
  void setVertices(Buffer *vertices, int numVertices)
  {
    this->vertices = vertices;
    this->numVertices = numVertices;
  }
  void setIndices(Buffer *indices, int numIndices)
  {
    this->indices = indices;
    this->numIndices = numIndices;
  }
  void setPrimColors(Buffer *colors) { this->colors = colors; }
  void setSpheres(Buffer *origins, Buffer *radii, Buffer *colors, int numSpheres, float defaultRadius)
  {
    this->origins = origins;
    this->radii = radii;
    this->colors = colors;
    primCount = numSpheres;
    this->defaultRadius = defaultRadius;
  }
  // vertices: vec3f[numVertices]; indices: vec2i[numCylinders] (end
  // points); radii: float[numCylinders], or null for defaultRadius
  void setCylinders(Buffer *vertices, int numVertices, Buffer *indices, int numCylinders, Buffer *radii,
                    Buffer *colors, float defaultRadius)
  {
    this->vertices = vertices;
    this->numVertices = numVertices;
    this->indices = indices;
    this->numIndices = numCylinders;
    this->radii = radii;
    this->colors = colors;
    primCount = numCylinders;
    this->defaultRadius = defaultRadius;
  }
  void setUserID(int userID) { this->userID = userID; }
  void setMaterial(const bnx_material &m) { material = m; }
  // the library's geometry, created on first use by a group
  bnx_geom handle(bnx_device dev);

  GeomType *const gt;
  bnx_geom h = nullptr;
  Buffer *vertices = nullptr, *indices = nullptr, *colors = nullptr, *origins = nullptr, *radii = nullptr;
  int numVertices = 0, numIndices = 0, primCount = 0, userID = 0;
  float defaultRadius = .01f;
  bnx_material material{BNX_MATERIAL_MATTE, 1, {.8f, .8f, .8f, 1.f}, 1.f, 1.f, 1.f, 0.f, 1.45f, 0.f};
};

// rtc::GeomType (rtcore/cuda/GeomType.h): "Triangles", "Spheres" or
// "Cylinders"
struct GeomType {
  enum Kind { TRIANGLES = 0, SPHERES = 1, CYLINDERS = 2 };
  GeomType(Device *device, Kind kind) : device(device), kind(kind) {}
  Geom *createGeom() { return new Geom(this); }
  Device *const device;
  const Kind kind;
};

// rtc::Group (rtcore/cuda/Group.h:14-25)
struct Group {
  Group(Device *device, bnx_group h, bool instances) : device(device), h(h), instances(instances) {}
  ~Group() { bnx_group_destroy(h); }
  void buildAccel() { check(bnx_group_build(h)); }
  // InstanceGroup::setTransforms keeps the transforms for the next
  // refitAccel, which builds the instance BVH again (as buildAccel does in
  // the reference's CUDA, HIP and CPU backends)
  void setTransforms(const std::vector<affine3f> &xfms)
  {
    if (instances) pendingXfms = xfms;
  }
  void refitAccel()
  {
    if (pendingXfms.empty()) {
      buildAccel();
      return;
    }
    check(bnx_group_set_transforms(h, (const float *)pendingXfms.data(), (int)pendingXfms.size()));
    pendingXfms.clear();
  }
  AccelHandle getDD() const { return (AccelHandle)h; }

  Device *const device;
  const bnx_group h;
  const bool instances;
  std::vector<affine3f> pendingXfms;
};

// rtc::Device (rtcore/cuda/Device.h:23-56)
struct Device {
  explicit Device(int physicalGPU) { check(bnx_device_create(physicalGPU, &h)); }
  ~Device() { destroy(); }
  void destroy()
  {
    if (h) bnx_device_destroy(h);
    h = nullptr;
  }
  std::string toString() const { return "bnx::rtc::Device (MI355X, gfx950)"; }
  void sync() { check(bnx_device_sync(h)); }

  GeomType *createGeomType(GeomType::Kind kind) { return new GeomType(this, kind); }
  void freeGeomType(GeomType *gt) { delete gt; }
  void freeGeom(Geom *g) { delete g; }

  Group *createTrianglesGroup(const std::vector<Geom *> &geoms)
  {
    std::vector<bnx_geom> hs;
    for (Geom *g : geoms) hs.push_back(g->handle(h));
    bnx_group out = nullptr;
    check(bnx_group_create_triangles(h, hs.data(), (int)hs.size(), &out));
    return new Group(this, out, false);
  }
  Group *createUserGeomsGroup(const std::vector<Geom *> &geoms)
  {
    std::vector<bnx_geom> hs;
    for (Geom *g : geoms) hs.push_back(g->handle(h));
    bnx_group out = nullptr;
    check(bnx_group_create_user(h, hs.data(), (int)hs.size(), &out));
    return new Group(this, out, false);
  }
  Group *createInstanceGroup(const std::vector<Group *> &groups, const std::vector<int> &instIDs,
                             const std::vector<affine3f> &xfms)
  {
    if (instIDs.size() != groups.size() || xfms.size() != groups.size())
      throw std::runtime_error("bnx: createInstanceGroup: groups, instIDs and xfms differ in size");
    std::vector<bnx_group> hs;
    for (Group *g : groups) hs.push_back(g->h);
    bnx_group out = nullptr;
    check(bnx_group_create_instances(h, hs.data(), instIDs.data(), (const float *)xfms.data(), (int)hs.size(),
                                     &out));
    return new Group(this, out, true);
  }
  void freeGroup(Group *g) { delete g; }
  Buffer *createBuffer(size_t numBytes, const void *initValues = nullptr) { return new Buffer(this, numBytes, initValues); }
  void freeBuffer(Buffer *b) { delete b; }
  // the SBT and pipeline of the reference's backends: nothing to build here
  void buildPipeline() {}
  void buildSBT() {}

  bnx_device h = nullptr;
};

static_assert(sizeof(affine3f) == 12 * sizeof(float), "affine3f is 12 floats");

inline Buffer::Buffer(Device *device, size_t numBytes, const void *initValues) : device(device), host(numBytes)
{
  check(bnx_buffer_create(device->h, numBytes, initValues, &h));
  if (initValues && numBytes) std::memcpy(host.data(), initValues, numBytes);
}

inline bnx_geom Geom::handle(bnx_device dev)
{
  if (h) return h;
  if (gt->kind == GeomType::TRIANGLES) {
    bnx_triangles_desc d{};
    d.vertices = vertices ? (const float *)vertices->hostData() : nullptr;
    d.num_vertices = numVertices;
    d.indices = indices ? (const int32_t *)indices->hostData() : nullptr;
    d.num_triangles = numIndices;
    d.prim_colors = colors ? (const float *)colors->hostData() : nullptr;
    d.user_id = userID;
    d.material = material;
    check(bnx_geom_create_triangles(dev, &d, &h));
  } else if (gt->kind == GeomType::CYLINDERS) {
    bnx_cylinders_desc d{};
    d.vertices = vertices ? (const float *)vertices->hostData() : nullptr;
    d.num_vertices = numVertices;
    d.indices = indices ? (const int32_t *)indices->hostData() : nullptr;
    d.num_cylinders = primCount;
    d.radii = radii ? (const float *)radii->hostData() : nullptr;
    d.colors = colors ? (const float *)colors->hostData() : nullptr;
    d.default_radius = defaultRadius;
    d.user_id = userID;
    d.material = material;
    check(bnx_geom_create_cylinders(dev, &d, &h));
  } else {
    bnx_spheres_desc d{};
    d.centers = origins ? (const float *)origins->hostData() : nullptr;
    d.radii = radii ? (const float *)radii->hostData() : nullptr;
    d.colors = colors ? (const float *)colors->hostData() : nullptr;
    d.num_spheres = primCount;
    d.default_radius = defaultRadius;
    d.user_id = userID;
    d.material = material;
    check(bnx_geom_create_spheres(dev, &d, &h));
  }
  return h;
}

// TraceKernel2D::launch of TraceRays (kernels/traceRays.cu:20-121) over an
// instance group: closest (or any) hit of n device-resident rays
inline void traceRays(Group *world, const bnx_ray *d_rays, int n, bnx_hit *d_hits, bool anyHit = false)
{
  check(bnx_trace(world->h, d_rays, n, d_hits, anyHit ? BNX_TRACE_ANY : BNX_TRACE_CLOSEST));
}

} // namespace rtc
} // namespace bnx
