/* bnx_rtc.hpp -- barney's rtcore host interface (namespace BARNEY_NS::rtc,
 * rtcore/cuda/Device.h:12-69, rtcore/cudaCommon/Device.h:21-46, Buffer.h,
 * Geom.h, GeomType.h, Group.h:14-25) mirrored in C++ over the C-ABI of
 * bnx_rt.h, for the rtc adapter of an "amd" backend.  Same class and method
 * names and argument meaning as the reference; every failing call throws
 * std::runtime_error with the library's message, as BARNEY_RAISE does
 * (rtcore/cudaCommon/cuda-helper.h:9-42).  Header-only; link with libbnx.so.
 *
 * Differences that follow from the boundary: a Buffer keeps a host copy of
 * what was uploaded (the C-ABI builds geometries from host arrays), and a
 * geometry is created in the library when a group is created over it.  A
 * geometry whose setters were called again, or one of whose buffers was
 * uploaded to since, is created anew for the next group (Triangles::commit,
 * native/geometry/Triangles.cpp:50-80, re-sets the buffers of the same
 * rtc::Geom and the next Group::build creates a new group); groups created
 * before keep the geometry they were built over, which lives until the last
 * of them is freed.  The reference's program model (SBT records, device
 * programs per GeomType) is the library's closed set of types:
 * "Triangles", "Spheres", "Cylinders", "Capsules" and "Cones". */
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
    version++;
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
  uint64_t version = 0;   // uploads so far (a geometry over it is rebuilt when it changes)
};

struct GeomType;

// rtc::Geom (rtcore/cuda/Geom.h:14-45): setVertices/setIndices for
// triangles; setDD carries a user type's data, here the Spheres::DD fields
// (geometry/Spheres.h:15-20) as setSpheres, the Cylinders::DD fields
// (geometry/Cylinders.h:16-22) as setCylinders, the Capsules::DD fields
// (native/geometry/Capsules.h:30-34) as setCapsules and the Cones::DD
// fields (native/geometry/Cones.h:20-24) as setCones
struct Geom {
  explicit Geom(GeomType *gt) : gt(gt) {}
  // groups must be freed before the geometries they hold (as in the
  // reference, whose native layer frees groups first)
  ~Geom()
  {
    if (h) bnx_geom_destroy(h);
    for (auto &r : retired) bnx_geom_destroy(r.first);
  }
  void setPrimCount(int primCount)
  {
    this->primCount = primCount;
    dirty = true;
  }
  void setVertices(Buffer *vertices, int numVertices)
  {
    this->vertices = vertices;
    this->numVertices = numVertices;
    dirty = true;
  }
  void setIndices(Buffer *indices, int numIndices)
  {
    this->indices = indices;
    this->numIndices = numIndices;
    dirty = true;
  }
  void setPrimColors(Buffer *colors)
  {
    this->colors = colors;
    dirty = true;
  }
  void setSpheres(Buffer *origins, Buffer *radii, Buffer *colors, int numSpheres, float defaultRadius)
  {
    dirty = true;
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
    dirty = true;
    this->vertices = vertices;
    this->numVertices = numVertices;
    this->indices = indices;
    this->numIndices = numCylinders;
    this->radii = radii;
    this->colors = colors;
    primCount = numCylinders;
    this->defaultRadius = defaultRadius;
  }
  // vertices: vec4f[numVertices] (position, radius); indices:
  // vec2i[numCapsules]
  void setCapsules(Buffer *vertices, int numVertices, Buffer *indices, int numCapsules, Buffer *colors)
  {
    dirty = true;
    this->vertices = vertices;
    this->numVertices = numVertices;
    this->indices = indices;
    this->numIndices = numCapsules;
    this->colors = colors;
    primCount = numCapsules;
  }
  // vertices: vec3f[numVertices]; radii: float[numVertices]; indices:
  // vec2i[numCones], or null (cone i joins vertices 2i, 2i + 1)
  void setCones(Buffer *vertices, int numVertices, Buffer *radii, Buffer *indices, int numCones, Buffer *colors)
  {
    dirty = true;
    this->vertices = vertices;
    this->numVertices = numVertices;
    this->radii = radii;
    this->indices = indices;
    this->numIndices = numCones;
    this->colors = colors;
    primCount = numCones;
  }
  void setUserID(int userID)
  {
    this->userID = userID;
    dirty = true;
  }
  void setMaterial(const bnx_material &m)
  {
    material = m;
    dirty = true;
  }
  // the library's geometry for a new group: created on first use, and
  // again after a setter call or an upload to one of its buffers; the group
  // holds it until freed (acquire / release)
  bnx_geom acquire(bnx_device dev);
  void release(bnx_geom g);

  GeomType *const gt;
  bnx_geom h = nullptr;
  int h_groups = 0;                                     // groups holding h
  std::vector<std::pair<bnx_geom, int>> retired;        // older geometries still held by groups
  bool dirty = true;
  uint64_t seen[5] = {0, 0, 0, 0, 0};                   // buffer versions h was built from
  Buffer *vertices = nullptr, *indices = nullptr, *colors = nullptr, *origins = nullptr, *radii = nullptr;
  int numVertices = 0, numIndices = 0, primCount = 0, userID = 0;
  float defaultRadius = .01f;
  bnx_material material{BNX_MATERIAL_MATTE, 1, {.8f, .8f, .8f, 1.f}, 1.f, 1.f, 1.f, 0.f, 1.45f, 0.f};
};

// rtc::GeomType (rtcore/cuda/GeomType.h): "Triangles", "Spheres",
// "Cylinders", "Capsules" or "Cones"
struct GeomType {
  enum Kind { TRIANGLES = 0, SPHERES = 1, CYLINDERS = 2, CAPSULES = 3, CONES = 4 };
  GeomType(Device *device, Kind kind) : device(device), kind(kind) {}
  Geom *createGeom() { return new Geom(this); }
  Device *const device;
  const Kind kind;
};

// rtc::Group (rtcore/cuda/Group.h:14-25)
struct Group {
  Group(Device *device, bnx_group h, bool instances) : device(device), h(h), instances(instances) {}
  ~Group()
  {
    bnx_group_destroy(h);
    for (auto &u : uses) u.first->release(u.second);
  }
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
  std::vector<std::pair<Geom *, bnx_geom>> uses;   // the geometries it was created over
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

  Group *createTrianglesGroup(const std::vector<Geom *> &geoms) { return geomGroup(geoms, false); }
  Group *createUserGeomsGroup(const std::vector<Geom *> &geoms) { return geomGroup(geoms, true); }
  Group *geomGroup(const std::vector<Geom *> &geoms, bool user)
  {
    std::vector<std::pair<Geom *, bnx_geom>> uses;
    std::vector<bnx_geom> hs;
    bnx_group out = nullptr;
    int rc = BNX_OK;
    try {
      for (Geom *g : geoms) {
        hs.push_back(g->acquire(h));
        uses.push_back({g, hs.back()});
      }
      rc = user ? bnx_group_create_user(h, hs.data(), (int)hs.size(), &out)
                : bnx_group_create_triangles(h, hs.data(), (int)hs.size(), &out);
    } catch (...) {
      for (auto &u : uses) u.first->release(u.second);
      throw;
    }
    if (rc != BNX_OK) {
      const std::string msg = std::string("bnx: ") + bnx_last_error();
      for (auto &u : uses) u.first->release(u.second);
      throw std::runtime_error(msg);
    }
    Group *g = new Group(this, out, false);
    g->uses = uses;
    return g;
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

  // rtc::CudaDeviceBase memory methods (rtcore/cudaCommon/Device.h:21-46)
  void copyAsync(void *dst, const void *src, size_t numBytes) { check(bnx_mem_copy_async(h, dst, src, numBytes)); }
  void copy(void *dst, const void *src, size_t numBytes)
  {
    copyAsync(dst, src, numBytes);
    sync();
  }
  void *allocHost(size_t numBytes)
  {
    void *p = nullptr;
    check(bnx_host_alloc(h, numBytes, &p));
    return p;
  }
  void freeHost(void *mem) { check(bnx_host_free(h, mem)); }
  void memsetAsync(void *mem, int value, size_t size) { check(bnx_mem_set_async(h, mem, value, size)); }
  void *allocMem(size_t numBytes)
  {
    void *p = nullptr;
    check(bnx_mem_alloc(h, numBytes, &p));
    return p;
  }
  void freeMem(void *mem) { check(bnx_mem_free(h, mem)); }
  // makes this GPU current; returns the one current before
  int setActive() const
  {
    int prev = -1;
    check(bnx_device_set_active(h, &prev));
    return prev;
  }
  void restoreActive(int oldActive) const { check(bnx_device_restore_active(oldActive)); }

  bnx_device h = nullptr;
};

// rtc::SetActiveGPU (rtcore/cudaCommon/Device.h:49-56): RAII around
// setActive / restoreActive
struct SetActiveGPU {
  explicit SetActiveGPU(const Device *device) : saved(device->setActive()) {}
  ~SetActiveGPU() { (void)bnx_device_restore_active(saved); }
  SetActiveGPU(const SetActiveGPU &) = delete;
  SetActiveGPU &operator=(const SetActiveGPU &) = delete;

private:
  int saved = -1;
};

static_assert(sizeof(affine3f) == 12 * sizeof(float), "affine3f is 12 floats");

inline Buffer::Buffer(Device *device, size_t numBytes, const void *initValues) : device(device), host(numBytes)
{
  check(bnx_buffer_create(device->h, numBytes, initValues, &h));
  if (initValues && numBytes) std::memcpy(host.data(), initValues, numBytes);
}

inline void Geom::release(bnx_geom g)
{
  if (g == h) {
    h_groups--;
    return;
  }
  for (size_t k = 0; k < retired.size(); k++)
    if (retired[k].first == g) {
      if (--retired[k].second == 0) {
        bnx_geom_destroy(g);
        retired.erase(retired.begin() + (long)k);
      }
      return;
    }
}

inline bnx_geom Geom::acquire(bnx_device dev)
{
  Buffer *bufs[5] = {vertices, indices, colors, origins, radii};
  bool stale = dirty || !h;
  for (int k = 0; k < 5; k++) stale = stale || (bufs[k] ? bufs[k]->version : 0) != seen[k];
  if (!stale) {
    h_groups++;
    return h;
  }
  bnx_geom old = h;
  h = nullptr;
  if (old) {
    if (h_groups > 0) retired.push_back({old, h_groups});
    else bnx_geom_destroy(old);
  }
  h_groups = 0;
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
  } else if (gt->kind == GeomType::CAPSULES) {
    bnx_capsules_desc d{};
    d.vertices = vertices ? (const float *)vertices->hostData() : nullptr;
    d.num_vertices = numVertices;
    d.indices = indices ? (const int32_t *)indices->hostData() : nullptr;
    d.num_capsules = primCount;
    d.colors = colors ? (const float *)colors->hostData() : nullptr;
    d.user_id = userID;
    d.material = material;
    check(bnx_geom_create_capsules(dev, &d, &h));
  } else if (gt->kind == GeomType::CONES) {
    bnx_cones_desc d{};
    d.vertices = vertices ? (const float *)vertices->hostData() : nullptr;
    d.num_vertices = numVertices;
    d.radii = radii ? (const float *)radii->hostData() : nullptr;
    d.indices = indices ? (const int32_t *)indices->hostData() : nullptr;
    d.num_cones = primCount;
    d.colors = colors ? (const float *)colors->hostData() : nullptr;
    d.user_id = userID;
    d.material = material;
    check(bnx_geom_create_cones(dev, &d, &h));
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
  for (int k = 0; k < 5; k++) seen[k] = bufs[k] ? bufs[k]->version : 0;
  dirty = false;
  h_groups = 1;
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
