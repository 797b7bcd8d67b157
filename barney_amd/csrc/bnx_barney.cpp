// bnx_barney.cpp -- barney's public bn* API (include/bnx_barney.h, after
// native/include/barney.h) implemented over the C-ABI of bnx_rt.h.
//
// Host code only.  Objects hold their staged parameters until bnCommit;
// commit turns them into the backend's records (geometry descriptors,
// materials, cameras, transfer functions), bnGroupBuild builds a group's
// BLASes, bnBuild flattens a model's instances into one instance group and a
// world (ModelSlot::build, native/ModelSlot.cpp:180-216), bnRender renders
// renderer.pathsPerPixel samples into the frame (Context::renderTiles,
// native/Context.cpp:139-182).
#include "../../include/bnx_barney.h"
#include "../../include/bnx_rt.h"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <set>
#include <stdexcept>
#include <string>
#include <vector>

namespace bnb {

thread_local std::string last_error;

void warn(const std::string &msg)
{
  last_error = msg;
  std::fprintf(stderr, "#bnx-barney: %s\n", msg.c_str());
}

struct Failure : std::runtime_error {
  using std::runtime_error::runtime_error;
};

inline void check(int rc, const char *what)
{
  if (rc != BNX_OK) throw Failure(std::string(what) + ": " + bnx_last_error());
}

// ------------------------------------------------------------------ objects
struct Context;

// intrusive reference count: the application's references (bnRelease,
// bnAddReference) and those of other objects that hold it
struct Obj {
  Context *ctx = nullptr;
  std::string kind, type;
  int refs = 1;
  virtual ~Obj() = default;
  virtual void commit() {}
};

void retain(Obj *o)
{
  if (o) o->refs++;
}
void release(Obj *o);

// a counted reference held by another object
template <typename T> struct Ref {
  T *p = nullptr;
  Ref() = default;
  Ref(T *x) : p(x) { retain(p); }
  Ref(const Ref &o) : p(o.p) { retain(p); }
  Ref &operator=(const Ref &o)
  {
    if (o.p != p) {
      retain(o.p);
      release(p);
      p = o.p;
    }
    return *this;
  }
  ~Ref() { release(p); }
  T *operator->() const { return p; }
  T &operator*() const { return *p; }
  T *get() const { return p; }
  explicit operator bool() const { return p != nullptr; }
};

struct Param {
  enum Kind { NONE, INT, FLOAT, STRING, OBJECT, XFM } kind = NONE;
  int n = 0;
  int i[4] = {0, 0, 0, 0};
  float f[16] = {0};
  std::string s;
  Ref<Obj> obj;
};

struct Object : Obj {
  std::map<std::string, Param> staged;
  bool has(const std::string &k) const { return staged.count(k) != 0; }
  const Param *get(const std::string &k, Param::Kind kind) const
  {
    auto it = staged.find(k);
    return it != staged.end() && it->second.kind == kind ? &it->second : nullptr;
  }
  float get1f(const std::string &k, float def) const
  {
    const Param *p = get(k, Param::FLOAT);
    if (p) return p->f[0];
    p = get(k, Param::INT);
    return p ? (float)p->i[0] : def;
  }
  int get1i(const std::string &k, int def) const
  {
    const Param *p = get(k, Param::INT);
    return p ? p->i[0] : def;
  }
  bool getf(const std::string &k, int n, float *out) const
  {
    const Param *p = get(k, Param::FLOAT);
    if (!p || p->n < n) return false;
    std::copy(p->f, p->f + n, out);
    return true;
  }
  bool geti(const std::string &k, int n, int *out) const
  {
    const Param *p = get(k, Param::INT);
    if (!p || p->n < n) return false;
    std::copy(p->i, p->i + n, out);
    return true;
  }
  template <typename T> T *getobj(const std::string &k, const char *kind) const
  {
    const Param *p = get(k, Param::OBJECT);
    if (!p || !p->obj || p->obj->kind != kind) return nullptr;
    return static_cast<T *>(p->obj.get());
  }
  std::string gets(const std::string &k) const
  {
    const Param *p = get(k, Param::STRING);
    return p ? p->s : std::string();
  }
};

struct Context {
  bnx_device dev = nullptr;
  std::set<Obj *> live;      // every object not yet destroyed
  bool dying = false;
  // the denoiser of the context's frame buffers (bnxContextSetDenoiser):
  // barney's device creates its own (rtc::Device::createDenoiser, OIDN or
  // OptiX); here the host hands one in
  bnx_denoise_fn denoise = nullptr;
  void *denoise_user = nullptr;
};

void release(Obj *o)
{
  if (!o) return;
  Context *c = o->ctx;
  if (c && c->dying) return;      // bnContextDestroy deletes every object itself
  if (--o->refs > 0) return;
  if (c) c->live.erase(o);
  delete o;
}

// ------------------------------------------------------------- data
size_t elem_size(BNDataType t)
{
  switch (t) {
  case BN_INT8:
  case BN_UINT8: return 1;
  case BN_INT32:
  case BN_UINT32:
  case BN_FLOAT32: return 4;
  case BN_INT32_VEC2:
  case BN_FLOAT32_VEC2: return 8;
  case BN_INT32_VEC3:
  case BN_FLOAT32_VEC3: return 12;
  case BN_INT32_VEC4:
  case BN_FLOAT32_VEC4: return 16;
  case BN_OBJECT:
  case BN_DATA:
  case BN_TEXTURE:
  case BN_TEXTURE_3D: return sizeof(void *);
  case BN_UFIXED8_RGBA:
  case BN_UFIXED8_RGBA_SRGB: return 4;
  case BN_UFIXED8: return 1;
  default: return 0;
  }
}

struct Data : Object {
  BNDataType dtype = BN_DATA_UNDEFINED;
  size_t count = 0;
  std::vector<uint8_t> bytes;
  std::vector<Ref<Obj>> objs;   // BN_OBJECT elements
  void set(size_t n, const void *items)
  {
    const size_t es = elem_size(dtype);
    if (es == 0) throw Failure("bnDataCreate: unsupported data type " + std::to_string((int)dtype));
    count = n;
    bytes.assign((const uint8_t *)items, (const uint8_t *)items + (items ? es * n : 0));
    if (!items) bytes.assign(es * n, 0);
    objs.clear();
    if (dtype == BN_OBJECT || dtype == BN_DATA || dtype == BN_TEXTURE || dtype == BN_TEXTURE_3D)
      for (size_t k = 0; k < n; k++) objs.push_back(Ref<Obj>(((Obj *const *)bytes.data())[k]));
  }
  const float *fp() const { return (const float *)bytes.data(); }
  const int32_t *ip() const { return (const int32_t *)bytes.data(); }
};

struct TextureData : Object {
  BNDataType format = BN_FLOAT32;
  int dims[3] = {0, 0, 0};
  std::vector<uint8_t> bytes;
};

struct Texture : Object {
  Ref<TextureData> data;
  BNTextureFilterMode filter = BN_TEXTURE_LINEAR;
  BNTextureAddressMode addr[2] = {BN_TEXTURE_WRAP, BN_TEXTURE_WRAP};
};

// ------------------------------------------------------------ material
struct Material : Object {
  bnx_material m{};
  void commit() override
  {
    const bool pbr = !(type == "matte" || type == "AnariMatte");
    m = bnx_material{};
    m.type = pbr ? BNX_MATERIAL_PBR : BNX_MATERIAL_MATTE;
    // AnariMatte color (.8, .8, .8), AnariPBR baseColor (1, 1, 1)
    // (material/AnariMatte.h:45-46, material/AnariPBR.h:52-58)
    const char *ck = pbr ? "baseColor" : "color";
    float c[4] = {pbr ? 1.f : .8f, pbr ? 1.f : .8f, pbr ? 1.f : .8f, 1.f};
    const Param *p = get(ck, Param::FLOAT);
    if (p) {
      for (int k = 0; k < std::min(p->n, 4); k++) c[k] = p->f[k];
      if (p->n == 3) c[3] = 1.f;
    }
    std::copy(c, c + 4, m.color);
    // a string names the attribute the colour comes from (PossiblyMappedParameter,
    // material/HostMaterial.cpp:59-64); this backend provides the per-prim "color"
    const std::string attr = gets(ck);
    m.color_from_attribute = (!attr.empty()) ? 1 : 0;
    if (!attr.empty() && attr != "color" && attr != "attribute0")
      warn("material " + type + ": colour attribute '" + attr + "' read as the geometry's per-primitive color");
    m.opacity = get1f("opacity", 1.f);
    m.metallic = get1f("metallic", 1.f);
    m.roughness = get1f("roughness", 1.f);
    m.transmission = get1f("transmission", 0.f);
    m.ior = get1f("ior", 1.45f);
  }
};

// ----------------------------------------------------------- geometry
struct ScalarField;
struct Geom : Object {
  int user_id = 0;
  Ref<Material> material;
  // the backend geometry built from the last commit, and older ones still
  // held by groups built before (with the number of groups holding each)
  bnx_geom h = nullptr;
  int h_users = 0;
  std::vector<std::pair<bnx_geom, int>> retired;
  bool valid = false;
  // iso_surface
  Ref<Obj> field;
  float iso = 0.f;
  ~Geom() override
  {
    if (h) bnx_geom_destroy(h);
    for (auto &r : retired) bnx_geom_destroy(r.first);
  }
  void drop(bnx_geom g)
  {
    if (g == h) {
      h_users--;
      return;
    }
    for (size_t k = 0; k < retired.size(); k++)
      if (retired[k].first == g && --retired[k].second == 0) {
        bnx_geom_destroy(g);
        retired.erase(retired.begin() + (long)k);
        return;
      }
  }
  bnx_geom use()
  {
    if (h) h_users++;
    return h;
  }
  void replace(bnx_geom g)
  {
    if (h) {
      if (h_users > 0) retired.push_back({h, h_users});
      else bnx_geom_destroy(h);
    }
    h = g;
    h_users = 0;
  }
  void commit() override;
  int backend_kind() const   // 0 triangles, 1 spheres, 2 cylinders, 3 capsules, 4 cones, 5 iso_surface
  {
    if (type == "triangles") return 0;
    if (type == "spheres") return 1;
    if (type == "cylinders") return 2;
    if (type == "capsules") return 3;
    if (type == "cones") return 4;
    return 5;
  }
};

const Data *data_param(const Object *o, const char *name)
{
  const Param *p = o->get(name, Param::OBJECT);
  if (!p || !p->obj || p->obj->kind != "data") return nullptr;
  return static_cast<const Data *>(p->obj.get());
}

void Geom::commit()
{
  user_id = get1i("userID", 0);
  material = Ref<Material>(getobj<Material>("material", "material"));
  bnx_material mat{};
  if (material) mat = material->m;
  else {
    Material def;
    def.type = "AnariMatte";
    def.commit();
    mat = def.m;
  }
  valid = false;
  const Data *colors = data_param(this, "primitive.color");
  if (!colors) colors = data_param(this, "colors");
  std::vector<float> col3;
  if (colors) {
    // per-primitive colour as float3 (float4 data keeps its rgb)
    const size_t n = colors->count;
    const int comps = colors->dtype == BN_FLOAT32_VEC4 ? 4 : 3;
    col3.resize(3 * n);
    for (size_t k = 0; k < n; k++)
      for (int c = 0; c < 3; c++) col3[3 * k + c] = colors->fp()[comps * k + c];
  }
  bnx_device dev = ctx->dev;
  bnx_geom g = nullptr;
  if (type == "triangles") {
    const Data *v = data_param(this, "vertices"), *ix = data_param(this, "indices");
    if (!v || !ix) return warn("triangles: 'vertices' and 'indices' are required");
    if (data_param(this, "normals") || data_param(this, "vertex.normal"))
      warn("triangles: vertex normals are not used (geometric normals)");
    bnx_triangles_desc d{};
    d.vertices = v->fp();
    d.num_vertices = (int)v->count;
    d.indices = ix->ip();
    d.num_triangles = (int)ix->count;
    d.prim_colors = colors ? col3.data() : nullptr;
    d.user_id = user_id;
    d.material = mat;
    check(bnx_geom_create_triangles(dev, &d, &g), "triangles");
  } else if (type == "spheres") {
    const Data *o = data_param(this, "origins"), *r = data_param(this, "radii");
    if (!o) return warn("spheres: 'origins' is required");
    bnx_spheres_desc d{};
    d.centers = o->fp();
    d.num_spheres = (int)o->count;
    d.radii = r ? r->fp() : nullptr;
    d.colors = colors ? col3.data() : nullptr;
    d.default_radius = get1f("radius", .1f);
    d.user_id = user_id;
    d.material = mat;
    check(bnx_geom_create_spheres(dev, &d, &g), "spheres");
  } else if (type == "cylinders") {
    const Data *v = data_param(this, "vertices"), *ix = data_param(this, "indices"), *r = data_param(this, "radii");
    if (!v || !ix) return warn("cylinders: 'vertices' and 'indices' are required");
    bnx_cylinders_desc d{};
    d.vertices = v->fp();
    d.num_vertices = (int)v->count;
    d.indices = ix->ip();
    d.num_cylinders = (int)(ix->dtype == BN_INT32_VEC2 ? ix->count : ix->count / 2);
    d.radii = r ? r->fp() : nullptr;
    d.colors = colors ? col3.data() : nullptr;
    d.default_radius = get1f("radius", .1f);
    d.user_id = user_id;
    d.material = mat;
    check(bnx_geom_create_cylinders(dev, &d, &g), "cylinders");
  } else if (type == "capsules") {
    // Capsules::setData (native/geometry/Capsules.cpp:49-65): 'vertices'
    // (vec4f: position, radius), 'indices' (vec2i)
    const Data *v = data_param(this, "vertices"), *ix = data_param(this, "indices");
    if (!v || !ix) return warn("capsules: 'vertices' and 'indices' are required");
    if (v->dtype != BN_FLOAT32_VEC4) return warn("capsules: 'vertices' must be float4 (position, radius)");
    bnx_capsules_desc d{};
    d.vertices = v->fp();
    d.num_vertices = (int)v->count;
    d.indices = ix->ip();
    d.num_capsules = (int)(ix->dtype == BN_INT32_VEC2 ? ix->count : ix->count / 2);
    d.colors = colors ? col3.data() : nullptr;
    d.user_id = user_id;
    d.material = mat;
    check(bnx_geom_create_capsules(dev, &d, &g), "capsules");
  } else if (type == "cones") {
    // Cones::setData / commit (native/geometry/Cones.cpp:18-73): 'vertices'
    // (vec3f), 'radii' (per vertex), optional 'indices' (vec2i; without
    // them cone i joins vertices 2i and 2i + 1)
    const Data *v = data_param(this, "vertices"), *ix = data_param(this, "indices"), *r = data_param(this, "radii");
    if (!v || !r) return warn("cones: 'vertices' and 'radii' are required");
    bnx_cones_desc d{};
    d.vertices = v->fp();
    d.num_vertices = (int)v->count;
    d.radii = r->fp();
    d.indices = ix ? ix->ip() : nullptr;
    d.num_cones = ix ? (int)(ix->dtype == BN_INT32_VEC2 ? ix->count : ix->count / 2) : (int)(v->count / 2);
    d.colors = colors ? col3.data() : nullptr;
    d.user_id = user_id;
    d.material = mat;
    check(bnx_geom_create_cones(dev, &d, &g), "cones");
  } else if (type == "iso_surface") {
    // IsoSurface (geometry/IsoSurface.cpp:44-90): 'isoValue', 'scalarField'
    const Param *p = get("scalarField", Param::OBJECT);
    field = p ? p->obj : Ref<Obj>();
    iso = get1f("isoValue", 0.f);
    if (!field || field->kind != "scalarfield") return warn("iso_surface: 'scalarField' is required");
    valid = true;
    return;
  } else {
    return warn("geometry type '" + type + "' is not provided by this backend");
  }
  replace(g);
  valid = true;
}

// ------------------------------------------------------- scalar fields
struct ScalarField : Object {
  bool ready = false;
  std::vector<float> scalars, verts;
  std::vector<int32_t> indices, offsets;
  std::vector<uint8_t> types;
  int dims[3] = {0, 0, 0};
  float origin[3] = {0, 0, 0}, spacing[3] = {1, 1, 1};
  uint64_t version = 0;
  void commit() override
  {
    ready = false;
    version++;
    if (type == "structured") {
      // StructuredData (volume/StructuredData.cu:125-181): textureData, dims,
      // gridOrigin, gridSpacing
      const TextureData *td = getobj<TextureData>("textureData", "texturedata");
      if (!td) return warn("structured field: 'textureData' is required");
      if (td->format != BN_FLOAT32) return warn("structured field: only float texels are provided");
      int d[3] = {td->dims[0], td->dims[1], td->dims[2]};
      geti("dims", 3, d);
      std::copy(d, d + 3, dims);
      getf("gridOrigin", 3, origin);
      getf("gridSpacing", 3, spacing);
      const size_t n = (size_t)dims[0] * dims[1] * dims[2];
      if (td->bytes.size() < 4 * n) return warn("structured field: textureData smaller than dims");
      scalars.assign((const float *)td->bytes.data(), (const float *)td->bytes.data() + n);
      ready = true;
    } else if (type == "unstructured") {
      // UMeshField (umesh/common/UMeshField.cu:200-232)
      const Data *pos = data_param(this, "vertex.position"), *val = data_param(this, "vertex.data");
      const Data *idx = data_param(this, "index"), *ci = data_param(this, "cell.index");
      const Data *ct = data_param(this, "cell.type");
      if (!pos || !val || !idx || !ci) return warn("unstructured field: vertex.position, vertex.data, index and "
                                                   "cell.index are required");
      verts.assign(pos->fp(), pos->fp() + 3 * pos->count);
      scalars.assign(val->fp(), val->fp() + val->count);
      indices.assign(idx->ip(), idx->ip() + idx->count);
      offsets.assign(ci->ip(), ci->ip() + ci->count);
      types.clear();
      if (ct) types.assign(ct->bytes.begin(), ct->bytes.begin() + ct->count);
      ready = true;
    } else {
      warn("scalar field type '" + type + "' is not provided by this backend");
    }
  }
};

struct Volume : Object {
  Ref<ScalarField> field;
  std::vector<float> tf;
  float domain[2] = {0.f, 1.f};
  float density = 1.f;
  int user_id = 0;
  bnx_volume h = nullptr;
  uint64_t built_version = ~0ull;
  bool tf_dirty = true;
  ~Volume() override
  {
    if (h) bnx_volume_destroy(h);
  }
  void commit() override { user_id = get1i("userID", 0); }
  // the backend volume for the field's current state
  bnx_volume ensure()
  {
    if (!field || !field->ready) throw Failure("volume: its scalar field is not committed");
    if (tf.empty()) throw Failure("volume: no transfer function (bnVolumeSetXF)");
    bnx_transfer_function t{};
    t.values = tf.data();
    t.num_values = (int)(tf.size() / 4);
    t.domain[0] = domain[0];
    t.domain[1] = domain[1];
    t.base_density = density;
    if (h && built_version == field->version) {
      if (tf_dirty) check(bnx_volume_set_transfer_function(h, &t), "bnVolumeSetXF");
      tf_dirty = false;
      return h;
    }
    if (h) {
      bnx_volume_destroy(h);
      h = nullptr;
    }
    ScalarField &f = *field;
    if (f.type == "structured") {
      bnx_structured_desc d{};
      d.scalars = f.scalars.data();
      std::copy(f.dims, f.dims + 3, d.dims);
      d.cells_per_mc = 8;
      std::copy(f.origin, f.origin + 3, d.grid_origin);
      std::copy(f.spacing, f.spacing + 3, d.grid_spacing);
      check(bnx_volume_create_structured(ctx->dev, &d, &t, user_id, &h), "structured volume");
    } else {
      bnx_umesh_desc d{};
      d.vertices = f.verts.data();
      d.scalars = f.scalars.data();
      d.indices = f.indices.data();
      d.cell_offsets = f.offsets.data();
      d.cell_types = f.types.empty() ? nullptr : f.types.data();
      d.num_vertices = (int)f.scalars.size();
      d.num_indices = (int)f.indices.size();
      d.num_cells = (int)f.offsets.size();
      check(bnx_volume_create_umesh(ctx->dev, &d, &t, user_id, &h), "unstructured volume");
    }
    built_version = f.version;
    tf_dirty = false;
    return h;
  }
};

// ---------------------------------------------------------------- lights
struct Light : Object {
  float dir[3] = {0.f, 0.f, 1.f};   // emission direction
  float color[3] = {1.f, 1.f, 1.f};
  float radiance = 1.f;
  float up[3] = {0.f, 0.f, 1.f};
  float scale = 1.f;
  Ref<Texture> texture;
  void commit() override
  {
    // DirLight (light/DirLight.cu:24-60), EnvMapLight (light/EnvMap.cu:280-335)
    getf("color", 3, color);
    if (type == "directional") {
      getf("direction", 3, dir);
      radiance = get1f("radiance", get1f("irradiance", 1.f));
    } else if (type == "envmap") {
      float d[3] = {1.f, 0.f, 0.f};
      getf("direction", 3, d);
      std::copy(d, d + 3, dir);
      getf("up", 3, up);
      scale = get1f("scale", 1.f);
      texture = Ref<Texture>(getobj<Texture>("texture", "texture"));
    } else {
      warn("light type '" + type + "' is not provided by this backend");
    }
  }
};

// ------------------------------------------------------------------ group
struct Group : Object {
  std::vector<Ref<Geom>> geoms;
  std::vector<Ref<Volume>> volumes;
  std::vector<Ref<Light>> lights;
  // built: one backend group per geometry kind (spheres, cylinders,
  // capsules, cones, triangles), with the geometries they were built over
  static constexpr int KINDS = 5;
  bnx_group g[KINDS] = {nullptr, nullptr, nullptr, nullptr, nullptr};
  std::vector<std::pair<Geom *, bnx_geom>> held;
  std::vector<Ref<Geom>> held_refs;
  std::vector<Ref<Geom>> isos;
  bool built = false;
  void free_built()
  {
    for (bnx_group &x : g)
      if (x) {
        bnx_group_destroy(x);
        x = nullptr;
      }
    for (auto &u : held) u.first->drop(u.second);
    held.clear();
    held_refs.clear();
    isos.clear();
    built = false;
  }
  ~Group() override { free_built(); }
  void commit() override
  {
    lights.clear();
    if (const Data *l = data_param(this, "lights"))
      for (const Ref<Obj> &o : l->objs)
        if (o && o->kind == "light") lights.push_back(Ref<Light>(static_cast<Light *>(o.get())));
  }
  void build()
  {
    free_built();
    // spheres, cylinders, capsules, cones (user geometries, one type per
    // backend group), triangles
    std::vector<bnx_geom> per[KINDS];
    for (Ref<Geom> &ge : geoms) {
      if (!ge || !ge->valid) continue;
      const int k = ge->backend_kind();
      if (k == 5) {
        isos.push_back(ge);
        continue;
      }
      const int slot = k == 0 ? KINDS - 1 : k - 1;
      bnx_geom h = ge->use();
      per[slot].push_back(h);
      held.push_back({ge.get(), h});
      held_refs.push_back(ge);
    }
    for (int s = 0; s < KINDS; s++) {
      if (per[s].empty()) continue;
      const int rc = s == KINDS - 1 ? bnx_group_create_triangles(ctx->dev, per[s].data(), (int)per[s].size(), &g[s])
                            : bnx_group_create_user(ctx->dev, per[s].data(), (int)per[s].size(), &g[s]);
      check(rc, "bnGroupBuild");
      check(bnx_group_build(g[s]), "bnGroupBuild");
    }
    built = true;
  }
};

// ------------------------------------------------------------------ model
struct Model : Object {
  std::vector<Ref<Group>> groups;
  std::vector<BNTransform> xfms;
  bnx_group inst = nullptr;
  bnx_world world = nullptr;
  std::vector<Ref<Group>> built_groups;   // what inst / world refer to
  std::vector<Ref<Volume>> built_vols;
  std::vector<bnx_isosurface> isos_made;
  int flat_count = 0;
  ~Model() override { free_built(); }
  void free_built()
  {
    if (world) bnx_world_destroy(world);
    for (bnx_isosurface h : isos_made) bnx_isosurface_destroy(h);
    if (inst) bnx_group_destroy(inst);
    world = nullptr;
    inst = nullptr;
    isos_made.clear();
    built_groups.clear();
    built_vols.clear();
  }
  static std::vector<float> affine(const BNTransform &t)
  {
    return {t.l.vx.x, t.l.vx.y, t.l.vx.z, t.l.vy.x, t.l.vy.y, t.l.vy.z,
            t.l.vz.x, t.l.vz.y, t.l.vz.z, t.p.x,    t.p.y,    t.p.z};
  }
  static bool is_identity(const BNTransform &t)
  {
    const float I[12] = {1, 0, 0, 0, 1, 0, 0, 0, 1, 0, 0, 0};
    const std::vector<float> a = affine(t);
    return std::equal(a.begin(), a.end(), I);
  }
  // flattened instance list (ModelSlot::flattenInstancesForDevice): per
  // instance its spheres, cylinders, capsules, cones and triangles groups
  void flatten(std::vector<bnx_group> *gs, std::vector<int32_t> *ids, std::vector<float> &xf) const
  {
    for (size_t i = 0; i < groups.size(); i++) {
      const Group *g = groups[i].get();
      if (!g) continue;
      for (int s = 0; s < Group::KINDS; s++)
        if (g->g[s]) {
          if (gs) gs->push_back(g->g[s]);
          if (ids) ids->push_back((int32_t)i);
          const std::vector<float> a = affine(xfms[i]);
          xf.insert(xf.end(), a.begin(), a.end());
        }
    }
  }
  void build()
  {
    for (Ref<Group> &g : groups)
      if (g && !g->built) g->build();
    free_built();
    std::vector<bnx_group> gs;
    std::vector<int32_t> ids;
    std::vector<float> xf;
    flatten(&gs, &ids, xf);
    flat_count = (int)gs.size();
    check(bnx_group_create_instances(ctx->dev, gs.data(), ids.data(), xf.data(), (int)gs.size(), &inst),
          "bnBuild");
    check(bnx_group_build(inst), "bnBuild");
    // lights of every instance (ModelSlot::updateWorldLightsFromInstances,
    // native/ModelSlot.cpp:88-125): directional lights transformed by their
    // instance, the last environment map
    std::vector<bnx_dir_light> dl;
    const Light *env = nullptr;
    BNTransform env_xf{};
    for (size_t i = 0; i < groups.size(); i++) {
      if (!groups[i]) continue;
      for (const Ref<Light> &l : groups[i]->lights) {
        if (l->type == "directional") {
          // DirLight::getDD (light/DirLight.cu:11-21): normalize(xfmVector(xfm, direction))
          const BNTransform &t = xfms[i];
          float d[3];
          d[0] = (t.l.vx.x * l->dir[0] + t.l.vy.x * l->dir[1]) + t.l.vz.x * l->dir[2];
          d[1] = (t.l.vx.y * l->dir[0] + t.l.vy.y * l->dir[1]) + t.l.vz.y * l->dir[2];
          d[2] = (t.l.vx.z * l->dir[0] + t.l.vy.z * l->dir[1]) + t.l.vz.z * l->dir[2];
          const float s = 1.f / std::sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
          bnx_dir_light x{};
          for (int k = 0; k < 3; k++) {
            x.direction[k] = d[k] * s;
            x.color[k] = l->color[k];
          }
          x.radiance = l->radiance;
          dl.push_back(x);
        } else if (l->type == "envmap") {
          env = l.get();
          env_xf = xfms[i];
        }
      }
    }
    check(bnx_world_create(ctx->dev, inst, dl.data(), (int)dl.size(), &world), "bnBuild");
    if (env && env->texture && env->texture->data) {
      if (!is_identity(env_xf)) warn("envmap light under a non-identity instance: its transform is not applied");
      const TextureData &td = *env->texture->data;
      if (td.format != BN_FLOAT32_VEC4 && td.format != BN_FLOAT32_VEC3)
        throw Failure("envmap: only float3 / float4 texels are provided");
      const int W = td.dims[0], H = td.dims[1];
      const int comps = td.format == BN_FLOAT32_VEC4 ? 4 : 3;
      std::vector<float> rgb((size_t)3 * W * H);
      const float *src = (const float *)td.bytes.data();
      for (size_t k = 0; k < (size_t)W * H; k++)
        for (int c = 0; c < 3; c++) rgb[3 * k + c] = src[comps * k + c];
      bnx_env_map_desc e{};
      e.radiance = rgb.data();
      e.width = W;
      e.height = H;
      std::copy(env->dir, env->dir + 3, e.direction);
      std::copy(env->up, env->up + 3, e.up);
      e.scale = env->scale;
      check(bnx_world_set_env_map(world, &e), "bnBuild (envmap)");
    }
    // isosurfaces, then volumes, per instance (user instance ID = index)
    for (size_t i = 0; i < groups.size(); i++) {
      Group *g = groups[i].get();
      if (!g) continue;
      const bool ident = is_identity(xfms[i]);
      for (Ref<Geom> &iso : g->isos) {
        if (!ident) throw Failure("bnBuild: an iso_surface under a non-identity instance transform");
        // the iso_surface's field as a volume of its own (not rendered)
        Volume *fv = nullptr;
        for (Ref<Volume> &v : built_vols)
          if (v->field.get() == iso->field.get()) fv = v.get();
        if (!fv) {
          Volume *nv = new Volume;
          nv->ctx = ctx;
          nv->kind = "volume";
          nv->field = Ref<ScalarField>(static_cast<ScalarField *>(iso->field.get()));
          nv->tf = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
          nv->refs = 0;
          built_vols.push_back(Ref<Volume>(nv));
          fv = nv;
        }
        bnx_material m = iso->material ? iso->material->m : bnx_material{};
        if (!iso->material) {
          m.type = BNX_MATERIAL_MATTE;
          m.color[0] = m.color[1] = m.color[2] = .8f;
          m.color[3] = 1.f;
          m.opacity = 1.f;
        }
        bnx_isosurface h = nullptr;
        check(bnx_isosurface_create(fv->ensure(), iso->iso, &m, iso->user_id, &h), "bnBuild (iso_surface)");
        check(bnx_world_add_isosurface(world, h, (int32_t)i), "bnBuild (iso_surface)");
        isos_made.push_back(h);
      }
      for (Ref<Volume> &v : g->volumes) {
        if (!ident) throw Failure("bnBuild: a volume under a non-identity instance transform");
        check(bnx_world_add_volume(world, v->ensure(), (int32_t)i), "bnBuild (volume)");
        built_vols.push_back(v);
      }
    }
    built_groups = groups;
  }
};

// ----------------------------------------------------------------- camera
struct Camera : Object {
  bnx_camera cam{};
  bool ok = false;
  void commit() override
  {
    float pos[3] = {0, 0, 0}, dir[3] = {0, 0, 1}, up[3] = {0, 1, 0};
    getf("position", 3, pos);
    getf("direction", 3, dir);
    getf("up", 3, up);
    // set3f("direction") normalizes (native/Camera.cu:68-71, 150-153, 226-229)
    const float l = 1.f / std::sqrt(dir[0] * dir[0] + dir[1] * dir[1] + dir[2] * dir[2]);
    for (float &x : dir) x *= l;
    ok = false;
    if (type == "perspective" || type == "pinhole") {
      check(bnx_camera_perspective(pos, dir, up, get1f("fovy", 60.f), &cam), "perspective camera");
      cam.aperture_radius = get1f("apertureRadius", 0.f);
      cam.focus_distance = get1f("focusDistance", 0.f);
    } else if (type == "orthographic") {
      check(bnx_camera_orthographic(pos, dir, up, get1f("height", 1.f), get1f("aspect", 1.f), &cam),
            "orthographic camera");
    } else if (type == "omni") {
      check(bnx_camera_omni(pos, dir, up, &cam), "omni camera");
    } else {
      return warn("camera type '" + type + "' is not provided by this backend");
    }
    ok = true;
  }
};

// --------------------------------------------------------------- renderer
struct Renderer : Object {
  bnx_renderer r{};
  int paths_per_pixel = 1;
  Renderer()
  {
    r.bg_color[3] = 1.f;
    r.ambient_radiance = 1.f;
    r.cut_plane[3] = -1e30f;
  }
  void commit() override
  {
    // Renderer::commit (native/render/Renderer.cpp:25-110)
    float bg[4] = {0.f, 0.f, 0.f, 1.f};
    getf("bgColor", 4, bg);
    std::copy(bg, bg + 4, r.bg_color);
    r.ambient_radiance = get1f("ambientRadiance", 1.f);
    paths_per_pixel = std::max(1, get1i("pathsPerPixel", 1));
    float cp[4] = {0.f, 0.f, 0.f, -1e30f};
    getf("cutPlane", 4, cp);
    // a zero normal with w >= 0 hides nothing: the same as no plane
    if (cp[0] == 0.f && cp[1] == 0.f && cp[2] == 0.f && cp[3] >= 0.f) cp[3] = -1e30f;
    std::copy(cp, cp + 4, r.cut_plane);
    if (has("bgTexture") && get("bgTexture", Param::OBJECT) && get("bgTexture", Param::OBJECT)->obj)
      warn("renderer: background textures are not provided by this backend");
  }
};

// ----------------------------------------------------------- frame buffer
struct FrameBuffer : Object {
  bnx_frame f = nullptr;
  int w = 0, h = 0;
  BNDataType format = BN_FLOAT32_VEC4;
  ~FrameBuffer() override
  {
    if (f) bnx_frame_destroy(f);
  }
};

// --------------------------------------------------------------- helpers
Context *ctx_of(BNContext c)
{
  if (!c) throw Failure("null context");
  return reinterpret_cast<Context *>(c);
}
Object *obj_of(BNObject o)
{
  if (!o) throw Failure("null object");
  return static_cast<Object *>(reinterpret_cast<Obj *>(o));
}
template <typename T> T *as(BNObject o, const char *kind)
{
  Object *x = obj_of(o);
  if (x->kind != kind) throw Failure(std::string("object is a ") + x->kind + ", not a " + kind);
  return static_cast<T *>(x);
}
template <typename T> T *make(Context *c, const char *kind, const std::string &type)
{
  T *o = new T;
  o->ctx = c;
  o->kind = kind;
  o->type = type;
  c->live.insert(o);
  return o;
}
template <typename H> H handle(Obj *o) { return reinterpret_cast<H>(o); }

#define BNB_TRY try {
#define BNB_CATCH(where_, ret)                                                                     \
  }                                                                                                \
  catch (const std::exception &e)                                                                  \
  {                                                                                                \
    warn(std::string(where_) + ": " + e.what());                                                   \
    return ret;                                                                                    \
  }

void set_param(BNObject target, const char *name, const Param &p)
{
  Object *o = obj_of(target);
  o->staged[name] = p;
}

} // namespace bnb

using namespace bnb;

extern "C" {

const char *bnGetLastError(void) { return last_error.c_str(); }

int bnPhysicalDeviceCount(void)
{
  int n = 0;
  return bnx_device_count(&n) == BNX_OK ? n : 0;
}

BNContext bnContextCreate(int numDataGroups, const int *, int numGPUs, const int *gpuIDs)
{
  BNB_TRY
  if (numDataGroups > 1) warn("bnContextCreate: one data group on one GPU (the rest are ignored)");
  const int gpu = (numGPUs > 0 && gpuIDs) ? gpuIDs[0] : 0;
  Context *c = new Context;
  if (bnx_device_create(gpu, &c->dev) != BNX_OK) {
    const std::string msg = bnx_last_error();
    delete c;
    throw Failure(msg);
  }
  return reinterpret_cast<BNContext>(c);
  BNB_CATCH("bnContextCreate", nullptr)
}

void bnContextDestroy(BNContext context)
{
  if (!context) return;
  Context *c = reinterpret_cast<Context *>(context);
  // every object goes, holders before what they hold (models hold worlds
  // over groups' BLASes and volumes, groups hold geometries' backend
  // handles); references between them are not counted down meanwhile
  c->dying = true;
  static const char *order[] = {"model", "group", "volume", "geometry", "light", "camera", "renderer",
                                "framebuffer", "scalarfield", "material", "texture", "texturedata", "data",
                                "sampler"};
  std::vector<Obj *> all(c->live.begin(), c->live.end());
  auto rank = [](const Obj *o) {
    for (int k = 0; k < (int)(sizeof(order) / sizeof(order[0])); k++)
      if (o->kind == order[k]) return k;
    return 100;
  };
  std::stable_sort(all.begin(), all.end(), [&](const Obj *a, const Obj *b) { return rank(a) < rank(b); });
  for (Obj *o : all) delete o;
  c->live.clear();
  bnx_device_destroy(c->dev);
  delete c;
}

void bnRelease(BNObject object)
{
  if (object) release(reinterpret_cast<Obj *>(object));
}

void bnAddReference(BNObject object)
{
  if (object) retain(reinterpret_cast<Obj *>(object));
}

BNCamera bnCameraCreate(BNContext context, const char *type)
{
  BNB_TRY
  return handle<BNCamera>(make<bnb::Camera>(ctx_of(context), "camera", type ? type : ""));
  BNB_CATCH("bnCameraCreate", nullptr)
}

BNFrameBuffer bnFrameBufferCreate(BNContext context, int)
{
  BNB_TRY
  return handle<BNFrameBuffer>(make<bnb::FrameBuffer>(ctx_of(context), "framebuffer", ""));
  BNB_CATCH("bnFrameBufferCreate", nullptr)
}

BNModel bnModelCreate(BNContext ctx)
{
  BNB_TRY
  return handle<BNModel>(make<bnb::Model>(ctx_of(ctx), "model", ""));
  BNB_CATCH("bnModelCreate", nullptr)
}

BNRenderer bnRendererCreate(BNContext ctx, const char *type)
{
  BNB_TRY
  return handle<BNRenderer>(make<bnb::Renderer>(ctx_of(ctx), "renderer", type ? type : "default"));
  BNB_CATCH("bnRendererCreate", nullptr)
}

BNData bnDataCreate(BNContext context, int, BNDataType dataType, size_t numItems, const void *items)
{
  BNB_TRY
  bnb::Data *d = make<bnb::Data>(ctx_of(context), "data", "");
  d->dtype = dataType;
  try {
    d->set(numItems, items);
  } catch (...) {
    release(d);
    throw;
  }
  return handle<BNData>(d);
  BNB_CATCH("bnDataCreate", nullptr)
}

void bnDataSet(BNData data, size_t numItems, const void *items)
{
  BNB_TRY
  as<bnb::Data>(data, "data")->set(numItems, items);
  BNB_CATCH("bnDataSet", )
}

static BNTextureData texdata(BNContext context, BNDataType fmt, int w, int h, int d, const void *items)
{
  const size_t es = elem_size(fmt);
  if (es == 0) throw Failure("unsupported texel format " + std::to_string((int)fmt));
  if (w <= 0 || h <= 0 || d <= 0) throw Failure("texture dimensions must be positive");
  bnb::TextureData *t = make<bnb::TextureData>(ctx_of(context), "texturedata", "");
  t->format = fmt;
  t->dims[0] = w;
  t->dims[1] = h;
  t->dims[2] = d;
  const size_t n = es * (size_t)w * h * d;
  if (items) t->bytes.assign((const uint8_t *)items, (const uint8_t *)items + n);
  else t->bytes.assign(n, 0);
  return handle<BNTextureData>(t);
}

BNTextureData bnTextureData2DCreate(BNContext context, int, BNDataType texelFormat, int width, int height,
                                    const void *items)
{
  BNB_TRY
  return texdata(context, texelFormat, width, height, 1, items);
  BNB_CATCH("bnTextureData2DCreate", nullptr)
}

BNTextureData bnTextureData3DCreate(BNContext context, int, BNDataType texelFormat, int width, int height, int depth,
                                    const void *items)
{
  BNB_TRY
  return texdata(context, texelFormat, width, height, depth, items);
  BNB_CATCH("bnTextureData3DCreate", nullptr)
}

BNTexture2D bnTexture2DCreate(BNContext context, int slot, BNDataType texelFormat, uint32_t size_x, uint32_t size_y,
                              const void *texels, BNTextureFilterMode filterMode, BNTextureAddressMode ax,
                              BNTextureAddressMode ay, BNTextureColorSpace)
{
  BNB_TRY
  BNTextureData td = texdata(context, texelFormat, (int)size_x, (int)size_y, 1, texels);
  bnb::Texture *t = make<bnb::Texture>(ctx_of(context), "texture", "2D");
  t->data = Ref<bnb::TextureData>(static_cast<bnb::TextureData *>(obj_of(td)));
  release(obj_of(td));   // the texture holds the data
  t->filter = filterMode;
  t->addr[0] = ax;
  t->addr[1] = ay;
  (void)slot;
  return handle<BNTexture2D>(t);
  BNB_CATCH("bnTexture2DCreate", nullptr)
}

BNLight bnLightCreate(BNContext context, int, const char *type)
{
  BNB_TRY
  return handle<BNLight>(make<bnb::Light>(ctx_of(context), "light", type ? type : ""));
  BNB_CATCH("bnLightCreate", nullptr)
}

BNGroup bnGroupCreate(BNContext context, int, BNGeom *geoms, int numGeoms, BNVolume *volumes, int numVolumes)
{
  BNB_TRY
  bnb::Group *g = make<bnb::Group>(ctx_of(context), "group", "");
  for (int k = 0; k < numGeoms; k++)
    if (geoms[k]) g->geoms.push_back(Ref<bnb::Geom>(as<bnb::Geom>(geoms[k], "geometry")));
  for (int k = 0; k < numVolumes; k++)
    if (volumes[k]) g->volumes.push_back(Ref<bnb::Volume>(as<bnb::Volume>(volumes[k], "volume")));
  return handle<BNGroup>(g);
  BNB_CATCH("bnGroupCreate", nullptr)
}

void bnGroupBuild(BNGroup group)
{
  BNB_TRY
  as<bnb::Group>(group, "group")->build();
  BNB_CATCH("bnGroupBuild", )
}

BNGeom bnGeometryCreate(BNContext context, int, const char *type)
{
  BNB_TRY
  const std::string t = type ? type : "";
  if (t != "triangles" && t != "spheres" && t != "cylinders" && t != "capsules" && t != "cones" &&
      t != "iso_surface")
    warn("bnGeometryCreate: geometry type '" + t + "' is not provided by this backend");
  return handle<BNGeom>(make<bnb::Geom>(ctx_of(context), "geometry", t));
  BNB_CATCH("bnGeometryCreate", nullptr)
}

BNScalarField bnScalarFieldCreate(BNContext context, int, const char *type)
{
  BNB_TRY
  return handle<BNScalarField>(make<bnb::ScalarField>(ctx_of(context), "scalarfield", type ? type : ""));
  BNB_CATCH("bnScalarFieldCreate", nullptr)
}

BNMaterial bnMaterialCreate(BNContext context, int, const char *type)
{
  BNB_TRY
  bnb::Material *m = make<bnb::Material>(ctx_of(context), "material", type ? type : "matte");
  m->commit();
  return handle<BNMaterial>(m);
  BNB_CATCH("bnMaterialCreate", nullptr)
}

BNSampler bnSamplerCreate(BNContext context, int, const char *type)
{
  BNB_TRY
  warn(std::string("bnSamplerCreate: samplers ('") + (type ? type : "") + "') are not provided by this backend");
  return handle<BNSampler>(make<bnb::Object>(ctx_of(context), "sampler", type ? type : ""));
  BNB_CATCH("bnSamplerCreate", nullptr)
}

BNVolume bnVolumeCreate(BNContext context, int, BNScalarField sf)
{
  BNB_TRY
  bnb::Volume *v = make<bnb::Volume>(ctx_of(context), "volume", "");
  v->field = Ref<bnb::ScalarField>(as<bnb::ScalarField>(sf, "scalarfield"));
  return handle<BNVolume>(v);
  BNB_CATCH("bnVolumeCreate", nullptr)
}

void bnVolumeSetXF(BNVolume volume, bn_float2 domain, const bn_float4 *colorMap, int numColorMapValues,
                   float densityAt1)
{
  BNB_TRY
  bnb::Volume *v = as<bnb::Volume>(volume, "volume");
  if (!colorMap || numColorMapValues < 2) throw Failure("at least two colour map values are needed");
  v->tf.assign((const float *)colorMap, (const float *)colorMap + 4 * (size_t)numColorMapValues);
  v->domain[0] = domain.x;
  v->domain[1] = domain.y;
  v->density = densityAt1;
  v->tf_dirty = true;
  BNB_CATCH("bnVolumeSetXF", )
}

void bnCommit(BNObject target)
{
  BNB_TRY
  obj_of(target)->commit();
  BNB_CATCH("bnCommit", )
}

void bnSetString(BNObject target, const char *name, const char *value)
{
  BNB_TRY
  Param p;
  p.kind = Param::STRING;
  p.s = value ? value : "";
  set_param(target, name, p);
  BNB_CATCH("bnSetString", )
}

void bnSetObject(BNObject target, const char *name, BNObject value)
{
  BNB_TRY
  Param p;
  p.kind = Param::OBJECT;
  p.obj = Ref<Obj>(value ? reinterpret_cast<Obj *>(value) : nullptr);
  set_param(target, name, p);
  BNB_CATCH("bnSetObject", )
}

void bnSetData(BNObject target, const char *name, BNData value) { bnSetObject(target, name, value); }
void bnSetLight(BNObject target, const char *name, BNLight value) { bnSetObject(target, name, value); }

static void seti(BNObject t, const char *name, int n, int a, int b, int c, int d)
{
  BNB_TRY
  Param p;
  p.kind = Param::INT;
  p.n = n;
  p.i[0] = a;
  p.i[1] = b;
  p.i[2] = c;
  p.i[3] = d;
  set_param(t, name, p);
  BNB_CATCH("bnSet*i", )
}
static void setf(BNObject t, const char *name, int n, float a, float b, float c, float d)
{
  BNB_TRY
  Param p;
  p.kind = Param::FLOAT;
  p.n = n;
  p.f[0] = a;
  p.f[1] = b;
  p.f[2] = c;
  p.f[3] = d;
  set_param(t, name, p);
  BNB_CATCH("bnSet*f", )
}

void bnSet1i(BNObject t, const char *n, int v) { seti(t, n, 1, v, 0, 0, 0); }
void bnSet2i(BNObject t, const char *n, int x, int y) { seti(t, n, 2, x, y, 0, 0); }
void bnSet3i(BNObject t, const char *n, int x, int y, int z) { seti(t, n, 3, x, y, z, 0); }
void bnSet4i(BNObject t, const char *n, int x, int y, int z, int w) { seti(t, n, 4, x, y, z, w); }
void bnSet1f(BNObject t, const char *n, float v) { setf(t, n, 1, v, 0, 0, 0); }
void bnSet2f(BNObject t, const char *n, float x, float y) { setf(t, n, 2, x, y, 0, 0); }
void bnSet3f(BNObject t, const char *n, float x, float y, float z) { setf(t, n, 3, x, y, z, 0); }
void bnSet4f(BNObject t, const char *n, float x, float y, float z, float w) { setf(t, n, 4, x, y, z, w); }

void bnSet4x3fv(BNObject target, const char *name, const BNTransform *m)
{
  BNB_TRY
  if (!m) throw Failure("null transform");
  Param p;
  p.kind = Param::XFM;
  p.n = 12;
  std::memcpy(p.f, m, sizeof(BNTransform));
  set_param(target, name, p);
  BNB_CATCH("bnSet4x3fv", )
}

void bnSet4x4fv(BNObject target, const char *name, const bn_float4 *m)
{
  BNB_TRY
  if (!m) throw Failure("null matrix");
  Param p;
  p.kind = Param::XFM;
  p.n = 16;
  std::memcpy(p.f, m, 16 * sizeof(float));
  set_param(target, name, p);
  BNB_CATCH("bnSet4x4fv", )
}

void bnSetInstances(BNModel model, int slot, BNGroup *groups, BNTransform *xfms, int n)
{
  BNB_TRY
  if (slot != 0) throw Failure("one data slot (0) on this backend");
  bnb::Model *m = as<bnb::Model>(model, "model");
  m->groups.clear();
  m->xfms.clear();
  for (int k = 0; k < n; k++) {
    m->groups.push_back(Ref<bnb::Group>(groups[k] ? as<bnb::Group>(groups[k], "group") : nullptr));
    BNTransform I{};
    I.l.vx.x = I.l.vy.y = I.l.vz.z = 1.f;
    m->xfms.push_back(xfms ? xfms[k] : I);
  }
  BNB_CATCH("bnSetInstances", )
}

void bnUpdateInstanceTransforms(BNModel model, int slot, BNTransform *xfms, int n)
{
  BNB_TRY
  if (slot != 0) throw Failure("one data slot (0) on this backend");
  bnb::Model *m = as<bnb::Model>(model, "model");
  if (n != (int)m->groups.size()) {
    warn("bnUpdateInstanceTransforms: ignoring an update with a mismatched instance count");
    return;
  }
  std::copy(xfms, xfms + n, m->xfms.begin());
  if (!m->inst) return;
  std::vector<float> xf;
  m->flatten(nullptr, nullptr, xf);
  check(bnx_group_set_transforms(m->inst, xf.data(), (int)(xf.size() / 12)), "bnUpdateInstanceTransforms");
  BNB_CATCH("bnUpdateInstanceTransforms", )
}

void bnSetInstanceAttributes(BNModel, int, const char *attributeName, BNData)
{
  warn(std::string("bnSetInstanceAttributes('") + (attributeName ? attributeName : "") +
       "'): instance attributes are not provided by this backend");
}

void bnBuild(BNModel model, int slot)
{
  BNB_TRY
  if (slot != 0) throw Failure("one data slot (0) on this backend");
  as<bnb::Model>(model, "model")->build();
  BNB_CATCH("bnBuild", )
}

void bnAccumReset(BNFrameBuffer fb)
{
  BNB_TRY
  bnb::FrameBuffer *f = as<bnb::FrameBuffer>(fb, "framebuffer");
  if (f->f) check(bnx_frame_reset(f->f), "bnAccumReset");
  BNB_CATCH("bnAccumReset", )
}

void bnFrameBufferResize(BNFrameBuffer fb, BNDataType colorFormat, int sizeX, int sizeY, uint32_t)
{
  BNB_TRY
  bnb::FrameBuffer *f = as<bnb::FrameBuffer>(fb, "framebuffer");
  if (sizeX <= 0 || sizeY <= 0) throw Failure("frame size must be positive");
  f->format = colorFormat;
  if (f->f && f->w == sizeX && f->h == sizeY) {
    check(bnx_frame_reset(f->f), "bnFrameBufferResize");
    return;
  }
  if (f->f) bnx_frame_destroy(f->f);
  f->f = nullptr;
  check(bnx_frame_create(f->ctx->dev, sizeX, sizeY, &f->f), "bnFrameBufferResize");
  f->w = sizeX;
  f->h = sizeY;
  BNB_CATCH("bnFrameBufferResize", )
}

void bnFrameBufferGetSize(BNFrameBuffer fb, int *sizeX, int *sizeY)
{
  BNB_TRY
  bnb::FrameBuffer *f = as<bnb::FrameBuffer>(fb, "framebuffer");
  if (sizeX) *sizeX = f->w;
  if (sizeY) *sizeY = f->h;
  BNB_CATCH("bnFrameBufferGetSize", )
}

void bnFrameBufferRead(BNFrameBuffer fb, BNFrameBufferChannel channel, void *out, BNDataType format)
{
  BNB_TRY
  bnb::FrameBuffer *f = as<bnb::FrameBuffer>(fb, "framebuffer");
  if (!out) return;
  if (!f->f) throw Failure("frame buffer not sized (bnFrameBufferResize)");
  int ch;
  switch (channel) {
  case BN_FB_COLOR:
    if (format == BN_FLOAT32_VEC4) ch = BNX_FB_COLOR;
    else if (format == BN_UFIXED8_RGBA) ch = BNX_FB_COLOR_RGBA8;
    else if (format == BN_UFIXED8_RGBA_SRGB) ch = BNX_FB_COLOR_SRGB8;
    else throw Failure("colour format " + std::to_string((int)format) + " not supported");
    break;
  case BN_FB_DEPTH: ch = BNX_FB_DEPTH; break;
  case BN_FB_PRIMID: ch = BNX_FB_PRIMID; break;
  case BN_FB_INSTID: ch = BNX_FB_INSTID; break;
  case BN_FB_OBJID: ch = BNX_FB_OBJID; break;
  case BN_FB_NORMAL: ch = BNX_FB_NORMAL; break;
  default: throw Failure("unknown channel");
  }
  if (channel == BN_FB_COLOR) {
    // FrameBuffer::set1i (fb/FrameBuffer.cu:60-75) applies "enableDenoising"
    // and "fadeOutDenoiser" as they are set; the context's denoiser runs on
    // the colour read when enabled ("enableUpscaling" is not provided)
    const bool on = f->get1i("enableDenoising", 0) != 0 && f->ctx->denoise;
    check(bnx_frame_set_denoiser(f->f, on ? f->ctx->denoise : nullptr, f->ctx->denoise_user,
                                 f->get1i("fadeOutDenoiser", 0)),
          "bnFrameBufferRead");
  }
  check(bnx_frame_read(f->f, ch, out), "bnFrameBufferRead");
  BNB_CATCH("bnFrameBufferRead", )
}

void bnxContextSetDenoiser(BNContext context, bnx_denoise_fn fn, void *user)
{
  BNB_TRY
  bnb::Context *c = bnb::ctx_of(context);
  c->denoise = fn;
  c->denoise_user = fn ? user : nullptr;
  BNB_CATCH("bnxContextSetDenoiser", )
}

void bnRender(BNRenderer renderer, BNModel model, BNCamera camera, BNFrameBuffer fb)
{
  BNB_TRY
  bnb::Renderer *r = as<bnb::Renderer>(renderer, "renderer");
  bnb::Model *m = as<bnb::Model>(model, "model");
  bnb::Camera *c = as<bnb::Camera>(camera, "camera");
  bnb::FrameBuffer *f = as<bnb::FrameBuffer>(fb, "framebuffer");
  if (!m->world) throw Failure("model not built (bnBuild)");
  if (!c->ok) throw Failure("camera not committed");
  if (!f->f) throw Failure("frame buffer not sized (bnFrameBufferResize)");
  check(bnx_render(m->world, &c->cam, &r->r, f->f, r->paths_per_pixel, nullptr), "bnRender");
  BNB_CATCH("bnRender", )
}

} // extern "C"
