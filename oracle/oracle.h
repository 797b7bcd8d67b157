// oracle.h -- CPU oracle for barney's single-GPU ray-tracing hot path.
//
// TEST INFRASTRUCTURE ONLY.  This is a scalar C++ restatement of the
// reference CPU (embree/cuBQL) path, used by tests/ and by bench.py's
// cpu_baseline leg as the checker.  The product (barney_amd/) never
// includes, links or calls anything in this directory.
//
// Every function cites the reference file:line it restates (paths relative
// to the barney v1.4.0 tree).  Arithmetic is written operation by operation
// in the order the reference writes it and is compiled with
// -ffp-contract=off, so results are reproducible bit for bit.
//
// Shared with the product: only barney_amd/csrc/bnx_numeric.h (sin/cos with
// a fixed operation sequence, pinned against libm by tests/test_numeric.py)
// and the POD layouts of include/bnx_rt.h.
//
// Parity status of each part is listed in the table at the top of
// oracle.cpp.
#pragma once

#include <stdint.h>
#include <string.h>
#include <math.h>
#include <vector>
#include <atomic>
#include <thread>

#include "../include/bnx_rt.h"
#include "../barney_amd/csrc/bnx_numeric.h"

namespace orc {

// ------------------------------------------------------------------ vec3
struct vec3 {
  float x, y, z;
  vec3() : x(0), y(0), z(0) {}
  vec3(float a) : x(a), y(a), z(a) {}
  vec3(float a, float b, float c) : x(a), y(b), z(c) {}
  float operator[](int i) const { return i == 0 ? x : (i == 1 ? y : z); }
  float &operator[](int i) { return i == 0 ? x : (i == 1 ? y : z); }
};
inline vec3 operator+(vec3 a, vec3 b) { return vec3(a.x + b.x, a.y + b.y, a.z + b.z); }
inline vec3 operator-(vec3 a, vec3 b) { return vec3(a.x - b.x, a.y - b.y, a.z - b.z); }
inline vec3 operator-(vec3 a) { return vec3(-a.x, -a.y, -a.z); }
inline vec3 operator*(vec3 a, vec3 b) { return vec3(a.x * b.x, a.y * b.y, a.z * b.z); }
inline vec3 operator*(vec3 a, float s) { return vec3(a.x * s, a.y * s, a.z * s); }
inline vec3 operator*(float s, vec3 a) { return vec3(s * a.x, s * a.y, s * a.z); }
inline vec3 operator/(vec3 a, float s) { return vec3(a.x / s, a.y / s, a.z / s); }
inline bool operator==(vec3 a, vec3 b) { return a.x == b.x && a.y == b.y && a.z == b.z; }
inline bool operator!=(vec3 a, vec3 b) { return !(a == b); }
inline float fmn(float a, float b) { return a < b ? a : b; }
inline float fmx(float a, float b) { return a > b ? a : b; }
inline vec3 vmin(vec3 a, vec3 b) { return vec3(fmn(a.x, b.x), fmn(a.y, b.y), fmn(a.z, b.z)); }
inline vec3 vmax(vec3 a, vec3 b) { return vec3(fmx(a.x, b.x), fmx(a.y, b.y), fmx(a.z, b.z)); }
inline float dot(vec3 a, vec3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline vec3 cross(vec3 a, vec3 b)
{
  return vec3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
inline float length(vec3 a) { return sqrtf(dot(a, a)); }
inline vec3 normalize(vec3 a) { return a * (1.f / sqrtf(dot(a, a))); }
inline float reduce_max(vec3 a) { return fmx(fmx(a.x, a.y), a.z); }
inline float reduce_min(vec3 a) { return fmn(fmn(a.x, a.y), a.z); }
inline vec3 vabs(vec3 a) { return vec3(fabsf(a.x), fabsf(a.y), fabsf(a.z)); }
inline vec3 load3(const float *p) { return vec3(p[0], p[1], p[2]); }

inline uint32_t fbits(float f) { uint32_t u; memcpy(&u, &f, 4); return u; }

static const float ONE_PI = (float)M_PI;
static const float TWO_PI = 2.f * ONE_PI;
static const float FOUR_PI = 4.f * ONE_PI;
static const float ONE_OVER_PI = 1.f / ONE_PI;
static const float ONE_OVER_TWO_PI = 1.f / TWO_PI;
static const float ONE_OVER_FOUR_PI = 1.f / FOUR_PI;

// ------------------------------------------------------------------- RNG
// RNGSeed / Random2 -- native/common/barney-common.h:21-67
struct RNGSeed {
  uint64_t state = 0xcafef00dd15ea5e5ull;
  void seed(uint32_t a, uint32_t b)               // :22-28
  {
    state ^= (uint64_t)(uint32_t)(2u * a + 1u);  // 32-bit unsigned arithmetic, then widened
    next();
    state ^= (uint64_t)(uint32_t)(2u * b);
    next();
  }
  uint32_t next()                                 // :30-38
  {
    uint64_t x = state;
    unsigned count = (unsigned)(x >> 61);
    state = x * 6364136223846793005ull;
    x ^= x >> 22;
    return (uint32_t)(x >> (22 + count));
  }
  uint32_t next_h(uint64_t hash)                  // :43-49
  {
    state ^= (hash << 1);
    return next();
  }
};
struct Random2 : RNGSeed {
  Random2(uint32_t a, uint32_t b) { seed(a, b); } // :53-55
  Random2(RNGSeed &parent, uint32_t b)            // :56-58
  {
    state = parent.state;
    parent.next();
    next_h((uint64_t)(b ? b : 290374u));
  }
  // Random2(RNGSeed &, uint64_t) (:59-61): a 64-bit hash
  static Random2 from64(RNGSeed &parent, uint64_t b)
  {
    Random2 r(0u, 0u);
    r.state = parent.state;
    parent.next();
    r.next_h(b ? b : 290374ull);
    return r;
  }
  float operator()() { return (float)(next() & 0x00FFFFFF) * (1.f / (float)0x01000000); } // :63-66
};

// FNV-1 style hash -- barney-common.h:88-108
inline uint64_t fnv_hash(uint32_t v) { return 0xcbf29ce484222325ull ^ (0x100000001b3ull * v); }
inline uint64_t fnv_hash(uint64_t h, uint32_t v) { return h * 0x100000001b3ull ^ v; }
inline uint64_t fnv_hash3(uint32_t a, uint32_t b, uint32_t c) { return fnv_hash(fnv_hash(fnv_hash(a), b), c); }

// ---------------------------------------------------------------- half
// vec3h storage of Ray::N and PathState::throughput (render/Ray.h:17-23,76).
// IEEE round-to-nearest-even, the device's conversion (common/half.h:48-52).
uint16_t f32_to_f16(float f);
float f16_to_f32(uint16_t h);
inline vec3 quantize_half(vec3 v)
{
  return vec3(f16_to_f32(f32_to_f16(v.x)), f16_to_f32(f32_to_f16(v.y)), f16_to_f32(f32_to_f16(v.z)));
}

// ------------------------------------------------------------------ box
struct Box {
  vec3 lo, hi;
  Box() : lo(INFINITY), hi(-INFINITY) {}
  Box(vec3 a, vec3 b) : lo(a), hi(b) {}
  void grow(const Box &b) { lo = vmin(lo, b.lo); hi = vmax(hi, b.hi); }
  void grow(vec3 p) { lo = vmin(lo, p); hi = vmax(hi, p); }
};
inline float half_area(const Box &b)
{
  vec3 d = b.hi - b.lo;
  return (d.x * d.y + d.y * d.z) + d.z * d.x;
}

// ------------------------------------------------------------------ BVH
// Node layout shared with the GPU builder: two child boxes + refs.  A child
// with count > 0 is a leaf over prim slots [ref, ref+count); count == 0 is
// the inner node 'ref'; count < 0 is an empty child.
struct Node {
  float lo0[3], hi0[3], lo1[3], hi1[3];
  int32_t ref[2];
  int32_t cnt[2];
};
static_assert(sizeof(Node) == 64, "node layout");

struct Bvh {
  std::vector<Node> nodes;
  std::vector<int32_t> order;   // BVH slot -> original prim index
  Box root;
  int max_depth = 0;
};

enum { SAH_BINS = 16, MAX_SAH_DEPTH = 48 };
void build_bvh(const std::vector<Box> &boxes, int max_leaf, Bvh &out);

// --------------------------------------------------------------- scene
struct Material {
  int type = BNX_MATERIAL_MATTE;
  int color_from_attribute = BNX_ATTRIB_NONE;   // the hit attribute the color comes from
  float color[4] = {0.8f, 0.8f, 0.8f, 1.f};
  float opacity = 1.f;
  float metallic = 1.f, roughness = 1.f, transmission = 0.f, ior = 1.45f;   // AnariPBR
};

// NVisii parameters of an AnariPBR hit, fp16-rounded as the packed BSDF
// stores them (packedBSDF/NVisii.h:1006-1024)
struct PbrParams {
  vec3 base;
  float metallic = 0.f, roughness = 0.f, ior = 0.f;
};
PbrParams pbr_params(const Material &mat, const float base[4]);
// NVisii::eval / pdf / scatter for geometric normal Ng (not flipped toward
// wo) and wo = -ray direction; scatter returns true for the diffuse lobe
void pbr_eval(const PbrParams &p, vec3 Ng, vec3 wo, vec3 wi, vec3 &value, float &pdf);
float pbr_pdf(const PbrParams &p, vec3 Ng, vec3 wo, vec3 wi);
bool pbr_scatter(const PbrParams &p, vec3 Ng, vec3 wo, Random2 &rng, vec3 &dir, float &pdf, vec3 &f_r);

// GeometryAttribute (render/GeometryAttributes.h:22-38) as resolved for the
// device: a scope (BNX_SCOPE_*) and a constant or an array of comps floats
// per element
struct Attribute {
  int scope = BNX_SCOPE_NONE;
  int comps = 0;
  float value[4] = {0.f, 0.f, 0.f, 1.f};
  std::vector<float> data;
};

struct Geom {
  int type = 0;                     // 0 triangles, 1 spheres, 2 cylinders, 3 capsules, 4 cones
  int user_id = 0;
  Material mat;
  std::vector<vec3> vertices;
  std::vector<int32_t> indices;     // 3 per triangle, 2 per cylinder, capsule or cone
  std::vector<vec3> centers;
  std::vector<float> radii;         // spheres, cylinders: per primitive; capsules, cones: per vertex
  float default_radius = 0.1f;
  std::vector<vec3> colors;         // spheres' own colors (Spheres::DD::colors), may be empty
  std::vector<vec3> normals;        // triangles' vertex normals (Triangles::DD::normals), may be empty
  Attribute attr[6];                // attribute0..3, color, normal
  int num_prims() const
  {
    return type == 0 ? (int)indices.size() / 3 : (type >= 2 ? (int)indices.size() / 2 : (int)centers.size());
  }
};

// CylindersPrograms::intersect (native/geometry/Cylinders.dev.cu:107-203):
// the capped cylinder v0-v1 in the ray's object frame; the front hit when it
// lies in [tmin, tmax], else the back hit; *N the unnormalized object normal
bool cylinder_intersect(vec3 ray_org, vec3 ray_dir, vec3 v0, vec3 v1, float radius, float tmin, float tmax,
                        float &t, vec3 *N);

// CapsulesPrograms::intersect (native/geometry/Capsules.dev.cu:154-212)
// with intersectRoundedCone (:19-98): the rounded cone between the end
// spheres (pa, ra), (pb, rb) in the ray's object frame, the origin moved to
// 0.99 of the entry distance into the box of both end spheres; the nearest
// of body and end spheres below tmax.  *N: the normalized object normal,
// *P: the object point (moved origin + moved distance)
bool capsule_intersect(vec3 ro, vec3 rd, vec3 pa, float ra, vec3 pb, float rb, float tmax, float &t, vec3 *N,
                       vec3 *P);
// ConesPrograms::intersect (native/geometry/Cones.dev.cu:60-266): the cap on
// the origin's side first, then the body's nearer root, in (tmin, tmax).
// *N: the unnormalized object normal; *lerp: weight of the second end
bool cone_intersect(vec3 ro, vec3 rd, vec3 p0, float ra, vec3 p1, float rb, float tmin, float tmax, float &t,
                    vec3 *N, float *lerp);

struct Group {
  int type = 0;                     // 0 triangles group, 1 user group
  std::vector<int> geoms;           // indices into Scene::geoms
  std::vector<int32_t> prim_geom;   // flattened prim list: geom index in group
  std::vector<int32_t> prim_id;     //                      prim index in geom
  Bvh bvh;
};

struct Affine {                       // owl affine3f: linear (columns vx,vy,vz) + p
  vec3 vx{1, 0, 0}, vy{0, 1, 0}, vz{0, 0, 1}, p{0, 0, 0};
  bool identity = true;
};

inline vec3 xfm_point(const Affine &a, vec3 p) { return ((a.vx * p.x + a.vy * p.y) + a.vz * p.z) + a.p; }
inline vec3 xfm_vector(const Affine &a, vec3 v) { return (a.vx * v.x + a.vy * v.y) + a.vz * v.z; }

struct Instance {
  int group;
  int user_id;
  Affine o2w, w2o;
};

// StructuredData volume with its macrocell majorants (volume/StructuredData.*,
// volume/MCGrid.*, volume/TransferFunction.h)
enum { FIELD_STRUCTURED = 0, FIELD_UMESH = 1 };
struct Volume {
  int field = FIELD_STRUCTURED;
  bool rendered = true;             // false: only the field of isosurfaces
  int dims[3] = {0, 0, 0};
  int cells_per_mc = 8;
  vec3 origin, spacing{1, 1, 1};
  std::vector<float> scalars;
  std::vector<float> tf;            // rgba per entry
  float dom0 = 0.f, dom1 = 1.f, base_density = 1.f;
  int user_id = 0, user_inst = 0;
  // derived by build_volume
  int ncells[3];
  Box bounds;
  int mc[3];
  vec3 mc_origin, mc_spacing;
  std::vector<float> mc_lo, mc_hi, majorant;
  // unstructured field: UMeshField (umesh/common/UMeshField.h)
  std::vector<float> verts;         // x, y, z, scalar per vertex
  std::vector<int32_t> tets;        // vertex indices 0..3 per cell; element-BVH slot order after build
  // other cells than tets: per cell its type (CELL_*) and vertex indices
  // 4..7 (-1 where unused); both empty when every cell is a tet
  std::vector<uint8_t> types;
  std::vector<int32_t> ext;
  Bvh ebvh;                         // element BVH over the cells' boxes
  int mc_req[3] = {0, 0, 0};        // requested macrocell grid (0: the reference's formula)
};

// IsoSurface geometry (geometry/IsoSurface.h:31-42) over a volume's field
struct IsoSurface {
  int field = 0;                    // index of the volume whose field it cuts
  float iso = 0.f;
  float color[3] = {0.8f, 0.8f, 0.8f};
  int user_id = 0, user_inst = 0;
};
void build_volume(Volume &v);
void build_majorants(Volume &v);
float volume_sample(const Volume &v, vec3 P);
void tf_map(const Volume &v, float s, float out[4]);
float tf_majorant(const Volume &v, float lo, float hi);
// MCVolumeAccel::isProg for one volume; 'dir' is the fixed trace direction.
// Returns true on a scattering event (t_out, rgb); advances 'parent' once when
// the ray overlaps the volume's bounds.
bool volume_march(const Volume &v, uint32_t rtc_inst, vec3 org, vec3 dir, float tmin, float tmax, RNGSeed &parent,
                  float &t_out, float rgb[3], uint64_t *samples);
float fast_log(float f);
// ratio tracking of a shadow ray through one volume (opt-in,
// BNX_RENDER_RATIO_TRACKING): the transmittance estimate in [0, 1], with the
// walk, RNG child and tentative collisions of volume_march
float ratio_track(const Volume &v, uint32_t rtc_inst, vec3 org, vec3 dir, float tmin, float tmax, RNGSeed &parent,
                  uint64_t *samples);
// tets: UMeshField commit + buildInitialMacroCells (UMeshField.cu:168-197,
// 320-379) + element BVH (UMeshCuBQLSampler::build, umesh/mc/
// UMeshCuBQLSampler.cu:45-78, with this build's binned-SAH builder)
void build_umesh(Volume &v);
// UMeshCuBQLSampler::DD::sample (umesh/mc/UMeshCuBQLSampler.h:46-62): the
// first tet that contains P in a depth-first walk of the element BVH
// (child 0 first, leaf slots in order); NaN if none.  'visits' counts
// element-BVH nodes and tet tests when non-null.
float umesh_sample(const Volume &v, vec3 P, uint64_t *visits = nullptr);
// tetScalar (UMeshField.h:391-436) of the tet in slot 'slot'
bool tet_scalar(const Volume &v, int slot, vec3 P, float &out);
// cell types (UMeshField.h:17-22, the ANARI codes)
enum { CELL_TET = 0, CELL_HEX = 1, CELL_PRISM = 2, CELL_PYR = 3 };
// eltScalar (UMeshField.h:372-389) of the cell in slot 'slot': tetScalar,
// or pyrScalar / prismScalar / hexScalar (:438-495) with the Newton
// inversions of ElementIntersection.h:77-476
bool cell_scalar(const Volume &v, int slot, vec3 P, float &out);
// the Newton inversion of one pyramid (5 vertices), prism (6) or hexahedron
// (8), vertices as x, y, z, scalar in VTK order
bool cell_invert(int type, const float *verts4, vec3 P, float &out);
// MCIsoSurfaceAccel::isProg (volume/MCAccelerator.h:325-548): t of the
// surface f = iso along the ray within [tmin, tmax), with the unnormalized
// central-difference gradient normal; advances 'parent' once when the ray
// overlaps the field's bounds.
bool iso_march(const Volume &field, const IsoSurface &iso, uint32_t rtc_inst, vec3 org, vec3 dir, float tmin,
               float tmax, RNGSeed &parent, float &t_out, vec3 &normal);

// EnvMapLight (light/EnvMap.h:27-90) over its HDRI texture: float4 texels,
// bilinear, WRAP in u and CLAMP in v (anari/light/HDRI.cpp:43-78)
struct EnvMap {
  bool present = false;
  int W = 0, H = 0;
  std::vector<float> tex;           // rgba per texel, row-major, row 0 at theta = 0
  float scale = 1.f;
  float to_world[9], to_local[9];   // columns vx, vy, vz
  std::vector<float> cdf_x;         // W per row (allCDFs_x)
  std::vector<float> cdf_y;         // H
};
// EnvMapLight::computeCDFs (light/EnvMap.cu:61-138, 183-239)
void env_build(EnvMap &e);
// rtc::tex2D of the env texture (rtcore/cpu/Texture.cpp:24-89, 294-321)
void env_tex(const EnvMap &e, float u, float v, float out[4]);
// EnvMapLight::DD::pdf (light/EnvMap.h:135-160); 1/(4 pi) without a texture
float env_pdf(const EnvMap &e, vec3 dir);
// EnvMapLight::DD::sample (light/EnvMap.h:211-250): four draws of r
void env_sample(const EnvMap &e, Random2 &r, vec3 &direction, vec3 &radiance, float &pdf);
// radianceFromEnv with a texture (kernels/shadeRays.cu:435-454)
vec3 env_radiance(const EnvMap &e, vec3 ray_dir);
// EnvMapLight::DD::uvToWorld (light/EnvMap.h:172-182)
vec3 env_uv_to_world(const EnvMap &e, float fx, float fy);

struct Scene {
  std::vector<Geom> geoms;
  std::vector<Group> groups;
  std::vector<Instance> insts;
  Bvh tlas;
  EnvMap env;
  std::vector<bnx_dir_light> dir_lights;
  std::vector<Volume> volumes;      // marched after the surface trace, in order
  std::vector<IsoSurface> isos;     // tested after the surface trace, before the volumes
  // World::DD::instanceAttributes (attribute0..3, color): rgba per instance, or empty
  std::vector<float> inst_attr[5];
};

// Geometry::DD::setHitAttributes (geometry/Geometry.h:94-139) for attribute
// slot k at a hit: elements vi (per vertex) / fi (face-varying) weighted by w
// (n of them; n == 1: a sphere reads its own element for both)
struct Corners {
  int n = 1;
  int vi[3] = {0, 0, 0}, fi[3] = {0, 0, 0};
  float w[3] = {1.f, 0.f, 0.f};
};
void hit_attribute(const Scene &s, const Geom &g, int k, int inst, int prim, const Corners &c, const float dflt[4],
                   float out[4]);
// PossiblyMappedParameter::DD::eval (material/HostMaterial.h:100-114) of the
// material color; color0: the hit's color before the attribute system
void material_color(const Scene &s, const Geom &g, int inst, int prim, const Corners &c, const float color0[4],
                    vec3 P, vec3 osP, float out[4]);
// the BSDF's opacity (Lambertian / NVisii alpha, AnariMatte.h:64, AnariPBR.h:87-90)
float material_alpha(const Material &m, const float base[4]);
Corners tri_corners(const int32_t *ix, int prim, float u, float v);
// the two ends of a cylinder, capsule or cone at attribute weight lerp
Corners pair_corners(const Geom &g, int prim, float lerp);
// the hit attributes of a capsule or cone hit at distance t_hit, found again
// in the instance's object frame (o, d: the object ray): object point and
// normal, world point (capsules: moved off the surface) and normal, and the
// attribute weight of the second end
struct PairHit {
  vec3 osP, osN, P, Nw;
  float lerp = 0.f;
};
void pair_hit(const Instance &in, const Geom &g, int prim, vec3 o, vec3 d, float t_hit, PairHit &h);

void build_scene(Scene &s);        // BLAS per group + TLAS

// ------------------------------------------------------------- tracing
// direction fix: traceRays.dev.cu:41-43 (exact zero -> 1e-6), then
// rtcore/cuda/TraceInterface.h:37-39 (|d| < 1e-6 -> 1e-6)
inline vec3 fix_dir(vec3 d)
{
  for (int k = 0; k < 3; k++) {
    if (d[k] == 0.f) d[k] = 1e-6f;
    if (fabsf(d[k]) < 1e-6f) d[k] = 1e-6f;
  }
  return d;
}

struct TraceCounters {
  uint64_t nodes = 0, tris = 0, spheres = 0, rays = 0;
};

struct Hit {
  float t;
  float u, v;
  int prim = -1, geom = -1, inst = -1;
};

// closest hit (any_hit=false) or occlusion (any_hit=true); with cut_plane
// (nx, ny, nz, w; off when w <= -1e28) surface hits on its negative side are
// rejected (OptixGlobals::hitOnInvisibleSide, native/render/OptixGlobals.h:36-45)
//
// Stochastic opacity (Triangles.dev.cu:120-133): with seed (the ray's RNG
// state) a triangle candidate whose material alpha is below 1 is kept with
// probability alpha, its draw taken from a copy of the seed advanced by the
// bits of its object-space position (so that the kept set does not depend on
// traversal order); seed null: every surface is opaque.  rej (when given)
// receives the closest candidate the opacity test rejected (prim -1: none).
Hit trace_bvh(const Scene &s, vec3 org, vec3 dir, float tmin, float tmax, bool any_hit, TraceCounters *ctr,
              const float *cut_plane = nullptr, const uint64_t *seed = nullptr, Hit *rej = nullptr);
Hit trace_brute(const Scene &s, vec3 org, vec3 dir, float tmin, float tmax, const float *cut_plane = nullptr,
                const uint64_t *seed = nullptr, Hit *rej = nullptr);

// ------------------------------------------------------------- render
struct Frame {
  int w = 0, h = 0, accum_id = 0;
  std::vector<int64_t> accum;      // 4 fixed-point channels per pixel (2^-32 units)
  std::vector<float> depth;
  std::vector<int32_t> prim, inst, obj;
  std::vector<float> normal;       // 3 per pixel
  void resize(int w, int h);
  void reset();
};

struct RenderStats {
  uint64_t rays = 0, shadow = 0, generations = 0, volume_samples = 0;
  TraceCounters ctr;
  // input: trace shadow rays as occlusion (any-hit) queries, as the GPU
  // path does; the image is the same, only the traversal counters change
  bool shadow_any_hit = false;
};

void render(const Scene &s, const bnx_camera &cam, const bnx_renderer &r, Frame &fb, int spp,
            int num_threads, RenderStats *stats);

// thread pool over [0,n) in chunks (rtcore/cpu/Compute.cpp:17-80: std::thread
// workers pulling task indices from an atomic counter)
template <typename F> void parallel_for(int n, int num_threads, F &&f)
{
  if (num_threads <= 1 || n < 256) {
    for (int i = 0; i < n; i++) f(i);
    return;
  }
  std::atomic<int> next{0};
  const int chunk = 64;
  auto worker = [&]() {
    while (true) {
      int b = next.fetch_add(chunk);
      if (b >= n) break;
      int e = b + chunk < n ? b + chunk : n;
      for (int i = b; i < e; i++) f(i);
    }
  };
  std::vector<std::thread> ts;
  for (int t = 0; t < num_threads - 1; t++) ts.emplace_back(worker);
  worker();
  for (auto &t : ts) t.join();
}

} // namespace orc
