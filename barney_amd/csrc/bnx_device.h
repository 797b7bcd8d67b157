// bnx_device.h -- device-side types and math of the gfx950 path.
//
// Float arithmetic is written operation by operation (the library is built
// with -ffp-contract=off and correctly rounded division / square root), so
// every hit distance, position and normal is bit-identical to a scalar CPU
// evaluation of the same expressions.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <type_traits>

#include "../../include/bnx_rt.h"
#include "bnx_numeric.h"

#define BNX_D __device__ __forceinline__

// Pointer fields of the device tables (scene, BLAS, geometry, volume,
// request-slot and frame records) point into global memory.  In the device
// pass they carry the global address space, so loads, stores and atomics
// through them and through pointers derived from them are global-memory
// instructions; a pointer loaded from memory as a plain pointer would be
// generic, and flat accesses through it also wait on the LDS counter.  The
// host pass sees plain pointers of the same size and assigns them with
// BNX_SET.
#if defined(__HIP_DEVICE_COMPILE__)
#define BNX_G __attribute__((address_space(1)))
#else
#define BNX_G
#endif
#define BNX_SET(field, value) ((field) = (std::remove_reference_t<decltype(field)>)(value))

namespace bnx {

// ------------------------------------------------------------------ vec3
struct v3 {
  float x, y, z;
};
BNX_D v3 mk(float a, float b, float c) { return v3{a, b, c}; }
BNX_D v3 splat(float a) { return v3{a, a, a}; }
BNX_D v3 operator+(v3 a, v3 b) { return v3{a.x + b.x, a.y + b.y, a.z + b.z}; }
BNX_D v3 operator-(v3 a, v3 b) { return v3{a.x - b.x, a.y - b.y, a.z - b.z}; }
BNX_D v3 operator-(v3 a) { return v3{-a.x, -a.y, -a.z}; }
BNX_D v3 operator*(v3 a, v3 b) { return v3{a.x * b.x, a.y * b.y, a.z * b.z}; }
BNX_D v3 operator*(v3 a, float s) { return v3{a.x * s, a.y * s, a.z * s}; }
BNX_D v3 operator*(float s, v3 a) { return v3{s * a.x, s * a.y, s * a.z}; }
BNX_D v3 operator/(v3 a, float s) { return v3{a.x / s, a.y / s, a.z / s}; }
BNX_D bool operator==(v3 a, v3 b) { return a.x == b.x && a.y == b.y && a.z == b.z; }
BNX_D float fmn(float a, float b) { return a < b ? a : b; }
BNX_D float fmx(float a, float b) { return a > b ? a : b; }
BNX_D v3 vmin(v3 a, v3 b) { return v3{fmn(a.x, b.x), fmn(a.y, b.y), fmn(a.z, b.z)}; }
BNX_D float dot(v3 a, v3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
BNX_D v3 cross(v3 a, v3 b) { return v3{a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
BNX_D float length(v3 a) { return sqrtf(dot(a, a)); }
BNX_D v3 normalize(v3 a) { return a * (1.f / sqrtf(dot(a, a))); }
BNX_D float reduce_max(v3 a) { return fmx(fmx(a.x, a.y), a.z); }
BNX_D v3 vabs(v3 a) { return v3{fabsf(a.x), fabsf(a.y), fabsf(a.z)}; }
BNX_D v3 xyz(float4 f) { return v3{f.x, f.y, f.z}; }
BNX_D v3 ld3(const float *p) { return v3{p[0], p[1], p[2]}; }

// direction fix of traceRays.dev.cu:41-43 and rtcore/cuda/TraceInterface.h:37-39
BNX_D float fix1(float d)
{
  if (d == 0.f) d = 1e-6f;
  if (fabsf(d) < 1e-6f) d = 1e-6f;
  return d;
}
BNX_D v3 fix_dir(v3 d) { return v3{fix1(d.x), fix1(d.y), fix1(d.z)}; }

// --------------------------------------------------------------- half
// vec3h storage (render/Ray.h:17-23,76): hardware IEEE RNE conversion
BNX_D uint32_t f2h(float f)
{
  // The empty asm pins f as a float32 register value: without it the
  // compiler fuses a preceding multiply into v_fma_mix*_f16, rounding once to
  // fp16 instead of to float32 and then to fp16 as the scalar code does.
  asm volatile("" : "+v"(f));
  _Float16 h = (_Float16)f;
  return (uint32_t)__builtin_bit_cast(unsigned short, h);
}
BNX_D float h2f(uint32_t b)
{
  _Float16 h = __builtin_bit_cast(_Float16, (unsigned short)(b & 0xffffu));
  return (float)h;
}
BNX_D v3 quant_half(v3 v) { return v3{h2f(f2h(v.x)), h2f(f2h(v.y)), h2f(f2h(v.z))}; }

// ------------------------------------------------------------------ RNG
// RNGSeed / Random2 (native/common/barney-common.h:21-67)
BNX_D uint32_t rng_next(uint64_t &state)
{
  uint64_t x = state;
  unsigned count = (unsigned)(x >> 61);
  state = x * 6364136223846793005ull;
  x ^= x >> 22;
  return (uint32_t)(x >> (22 + count));
}
BNX_D void rng_seed(uint64_t &state, uint32_t a, uint32_t b)
{
  state ^= (uint64_t)(uint32_t)(2u * a + 1u);
  rng_next(state);
  state ^= (uint64_t)(uint32_t)(2u * b);
  rng_next(state);
}
BNX_D uint32_t rng_next_h(uint64_t &state, uint64_t h)
{
  state ^= (h << 1);
  return rng_next(state);
}
BNX_D uint64_t fnv3(uint32_t a, uint32_t b, uint32_t c)
{
  // hash(v0,v1,v2) of native/common/barney-common.h:88-105
  const uint64_t prime = 0x100000001b3ull;
  uint64_t h = 0xcbf29ce484222325ull ^ (prime * (uint64_t)a);
  h = h * prime ^ (uint64_t)b;
  h = h * prime ^ (uint64_t)c;
  return h;
}

struct Random2 {
  uint64_t state;
  BNX_D Random2(uint32_t a, uint32_t b) : state(0xcafef00dd15ea5e5ull) { rng_seed(state, a, b); }
  BNX_D Random2(uint64_t &parent, uint32_t b)
  {
    state = parent;
    rng_next(parent);
    rng_next_h(state, (uint64_t)(b ? b : 290374u));
  }
  BNX_D float operator()() { return (float)(rng_next(state) & 0x00FFFFFFu) * (1.f / (float)0x01000000); }
};

// ------------------------------------------------------------------ BVH
// 64-byte node: boxes of both children plus refs (see bnx_group_download_bvh)
struct __align__(16) Node {
  float lo0x, lo0y, lo0z, hi0x;
  float hi0y, hi0z, lo1x, lo1y;
  float lo1z, hi1x, hi1y, hi1z;
  int ref0, ref1, cnt0, cnt1;
};
static_assert(sizeof(Node) == 64, "node");

// traversal encoding of a child: inner node -> its index (< 2^30);
// leaf -> 0x80000000 | first << 3 | count
constexpr uint32_t LEAF_BIT = 0x80000000u;
constexpr uint32_t LEAVE_BLAS = 0x7fffffffu;
BNX_D uint32_t enc_child(int ref, int cnt)
{
  return cnt > 0 ? (LEAF_BIT | ((uint32_t)ref << 3) | (uint32_t)cnt) : (uint32_t)ref;
}

// 128-byte 4-wide node, collapsed from the binary BVH for traversal: child
// k's box is component k of the six axis vectors; child codes as
// enc_child, EMPTY4 for an unused slot
struct __align__(128) Node4 {
  float4 lox, hix, loy, hiy, loz, hiz;
  uint4 child;
  uint4 pad;
};
static_assert(sizeof(Node4) == 128, "node4");
constexpr uint32_t EMPTY4 = 0x7ffffffeu;

// ----------------------------------------------------------- scene data
// BLAS types; the user types with two float4 records per primitive
// (cylinders, capsules, cones) share the trace kernels' pair leaf test
enum { BLAS_TRIANGLES = 0, BLAS_SPHERES = 1, BLAS_CYLINDERS = 2, BLAS_CAPSULES = 3, BLAS_CONES = 4 };

struct BlasDev {
  const Node4 BNX_G *nodes;
  const float4 BNX_G *prims;    // triangles: 3 per slot (v0|prim, e1|geom, e2); spheres: (c, r)
  const int2 BNX_G *refs;  // spheres: (geom, prim) per slot
  int type;
  int num_prims;
  int translucent;         // a geometry of it has stochastic opacity
  int _pad;
};

struct InstDev {
  float4 w2o[3];          // rows of the world->object affine (x,y,z,p)
  float4 o2w[3];          // rows of the object->world affine
  int blas;
  int user_id;
  int identity;
  int geom_base;          // first entry of this instance's group in the geometry table
};

// one geometry attribute (GeometryAttribute::DD, render/GeometryAttributes.h:
// 22-38): a constant, or elements of comps floats per prim / vertex / face
// corner (scope BNX_SCOPE_*; BNX_SCOPE_NONE: not set)
struct AttrDev {
  float4 value;
  const float BNX_G *data;
  int scope;
  int comps;
};
enum { ATTR_SLOTS = 6 };   // attribute0..3, color, normal (BNX_ATTRIB_SLOT_*)

// per-geometry shading data
struct GeomDev {
  int type;               // 0 triangles, 1 spheres, 2 cylinders, 3 capsules, 4 cones
  int user_id;
  int color_from_attr;    // BNX_ATTRIB_*: the hit attribute the color comes from
  int num_prims;
  float4 color;           // constant material color (matte 'color', PBR 'baseColor')
  int material;           // BNX_MATERIAL_MATTE or BNX_MATERIAL_PBR
  float metallic, roughness, ior;   // PBR parameters as given (rounded to fp16 per hit)
  float opacity, transmission;      // 'opacity'; PBR 'transmission'
  int translucent;        // triangles whose alpha can be below 1: stochastic opacity
  int _pad;
  const float BNX_G *vertices;   // triangles: float3; cylinders, capsules, cones: float3
  const int BNX_G *indices;      // triangles: int3; cylinders, capsules, cones: int2
  const float BNX_G *colors;     // spheres' own per-sphere float3 colors, or null
  const float BNX_G *normals;    // triangles: per-vertex float3 normals, or null
  const float4 BNX_G *spheres;   // spheres: (center, radius)
  const float4 BNX_G *cyls;      // two records per primitive: cylinders (v0, radius), (v1, 0);
                                 // capsules and cones (v0, r0), (v1, r1)
  AttrDev attr[ATTR_SLOTS];
};

struct SceneDev {
  const Node4 BNX_G *tlas_nodes;
  const int BNX_G *tlas_order;
  const InstDev BNX_G *insts;
  const BlasDev BNX_G *blas;
  const GeomDev BNX_G *geoms;
  // World::DD::instanceAttributes: a float4 per rtc instance for
  // attribute0..3 and color, or null
  const float4 BNX_G *inst_attr[5];
  int num_insts;
};

// AttributeArray::DD::valueAt (render/GeometryAttributes.h:68-91)
BNX_D float4 attr_value_at(const AttrDev &a, int i)
{
  const float *p = a.data + (long long)a.comps * i;
  switch (a.comps) {
  case 1: return make_float4(p[0], 0.f, 0.f, 1.f);
  case 2: return make_float4(p[0], p[1], 0.f, 1.f);
  case 3: return make_float4(p[0], p[1], p[2], 1.f);
  case 4: return make_float4(p[0], p[1], p[2], p[3]);
  default: return make_float4(0.f, 0.f, 0.f, 0.f);
  }
}

// The elements a hit interpolates per-vertex (vi) or face-varying (fi)
// attributes from, with weights w: a triangle's three corners at
// barycentrics (1-u-v, u, v) (Triangles.dev.cu:96-111), a cylinder's two
// ends at (1-lerp_t, lerp_t) (Cylinders.dev.cu:208-222); a sphere reads its
// own element for both (n = 1, Spheres.dev.cu:94-100).
struct Corners {
  int n;
  int vi[3], fi[3];
  float w[3];
};

BNX_D float4 scale4(float s, float4 a) { return make_float4(s * a.x, s * a.y, s * a.z, s * a.w); }
BNX_D float4 add4(float4 a, float4 b) { return make_float4(a.x + b.x, a.y + b.y, a.z + b.z, a.w + b.w); }

BNX_D float4 attr_interp(const AttrDev &a, const Corners &c, bool face_varying)
{
  const int *ix = face_varying ? c.fi : c.vi;
  if (c.n == 1) return attr_value_at(a, ix[0]);
  float4 r = add4(scale4(c.w[0], attr_value_at(a, ix[0])), scale4(c.w[1], attr_value_at(a, ix[1])));
  if (c.n == 3) r = add4(r, scale4(c.w[2], attr_value_at(a, ix[2])));
  return r;
}

// the value of geometry attribute slot k at a hit, as Geometry::DD::
// setHitAttributes sets it (geometry/Geometry.h:94-139): the geometry's own
// (constant, per primitive, per vertex, face-varying), else its instance's
// (slots 0-4, indexed by the rtc instance), else dflt
BNX_D float4 hit_attribute(const SceneDev &S, const GeomDev &G, int k, int inst, int prim, const Corners &c,
                           float4 dflt)
{
  const AttrDev &a = G.attr[k];
  switch (a.scope) {
  case BNX_SCOPE_CONSTANT: return a.value;
  case BNX_SCOPE_PER_PRIM: return attr_value_at(a, prim);
  case BNX_SCOPE_PER_VERTEX: return attr_interp(a, c, false);
  case BNX_SCOPE_FACE_VARYING: return attr_interp(a, c, true);
  default: break;
  }
  if (k < 5 && S.inst_attr[k]) return S.inst_attr[k][inst];
  return dflt;
}

// The material color: PossiblyMappedParameter::DD::eval (material/
// HostMaterial.h:100-114) of the matte 'color' / PBR 'baseColor', a value
// (NaN: unset, (0,0,0,1)) or a hit attribute (HitAttributes::get, render/
// HitAttributes.h:61-80).  color0: the hit's color before the attribute
// system (spheres' own colors, else (0,0,0,1)); P / osP: world / object
// position of the hit.
BNX_D float4 material_color(const SceneDev &S, const GeomDev &G, int inst, int prim, const Corners &c, float4 color0,
                            v3 P, v3 osP)
{
  const int k = G.color_from_attr;
  const float4 none = make_float4(0.f, 0.f, 0.f, 1.f);
  if (k == BNX_ATTRIB_NONE) return G.color.x != G.color.x ? none : G.color;
  if (k == BNX_ATTRIB_COLOR) return hit_attribute(S, G, 4, inst, prim, c, color0);
  if (k >= BNX_ATTRIB_0 && k < BNX_ATTRIB_0 + 4) return hit_attribute(S, G, k - BNX_ATTRIB_0, inst, prim, c, none);
  if (k == BNX_ATTRIB_WORLD_POSITION) return make_float4(P.x, P.y, P.z, 1.f);
  if (k == BNX_ATTRIB_OBJECT_POSITION) return make_float4(osP.x, osP.y, osP.z, 1.f);
  return none;
}

// the BSDF's opacity: Lambertian alpha = color.w * opacity (AnariMatte.h:
// 64); NVisii alpha = (1 - transmission) * baseColor.w * opacity, stored as
// half (AnariPBR.h:87-90, NVisii.h:1082)
BNX_D float material_alpha(const GeomDev &G, float4 base)
{
  if (G.material == BNX_MATERIAL_PBR) return h2f(f2h(((1.f - G.transmission) * base.w) * G.opacity));
  return base.w * G.opacity;
}

// the two ends of a cylinder, capsule or cone at attribute weight lerp
BNX_D Corners pair_corners(const GeomDev &G, int prim, float lerp)
{
  Corners c;
  c.n = 2;
  c.vi[0] = G.indices[2 * prim];
  c.vi[1] = G.indices[2 * prim + 1];
  c.fi[0] = 2 * prim;
  c.fi[1] = 2 * prim + 1;
  c.w[0] = 1.f - lerp;
  c.w[1] = lerp;
  return c;
}

// a triangle hit's corners for attribute interpolation
BNX_D Corners tri_corners(const int *ix, int prim, float u, float v)
{
  Corners c;
  c.n = 3;
  c.vi[0] = ix[0];
  c.vi[1] = ix[1];
  c.vi[2] = ix[2];
  c.fi[0] = 3 * prim;
  c.fi[1] = 3 * prim + 1;
  c.fi[2] = 3 * prim + 2;
  c.w[0] = (1.f - u) - v;
  c.w[1] = u;
  c.w[2] = v;
  return c;
}

BNX_D v3 xfm_point_rows(const float4 *m, v3 p)
{
  // rows (a,b,c,d): x' = a*px + b*py + c*pz + d, grouping as owl xfmPoint:
  // ((vx*px + vy*py) + vz*pz) + p per component
  return v3{((m[0].x * p.x + m[0].y * p.y) + m[0].z * p.z) + m[0].w,
            ((m[1].x * p.x + m[1].y * p.y) + m[1].z * p.z) + m[1].w,
            ((m[2].x * p.x + m[2].y * p.y) + m[2].z * p.z) + m[2].w};
}
BNX_D v3 xfm_vector_rows(const float4 *m, v3 v)
{
  return v3{(m[0].x * v.x + m[0].y * v.y) + m[0].z * v.z, (m[1].x * v.x + m[1].y * v.y) + m[1].z * v.z,
            (m[2].x * v.x + m[2].y * v.y) + m[2].z * v.z};
}

// CylindersPrograms::intersect (native/geometry/Cylinders.dev.cu:107-203),
// the capped cylinder from v0 to v1 of the given radius, in the ray's
// object frame: the front hit t0 when it lies in [tmin, tmax], else the back
// hit t1 when that does; *N (if asked) the unnormalized object normal of the
// reference (for the back hit it keeps the reference's test t0 == cap_t1).
// Written operation by operation as the reference evaluates it.
BNX_D bool cyl_intersect(v3 ray_org, v3 ray_dir, v3 v0, v3 v1, float radius, float tmin, float tmax, float &t,
                         v3 *N)
{
  const v3 d = ray_dir;
  const v3 s = v1 - v0;
  const v3 sxd = cross(s, d);
  const float a = dot(sxd, sxd);
  if (a == 0.f) return false;
  const v3 f = v0 - ray_org;
  const v3 sxf = cross(s, f);
  const float ra = 1.f / a;
  const float ts = dot(sxd, sxf) * ra;
  const v3 fp = f - ts * d;
  const float s2 = dot(s, s);
  const v3 perp = cross(s, fp);
  const float c = (radius * radius) * s2 - dot(perp, perp);
  if (c < 0.f) return false;
  float td = sqrtf(c * ra);
  const float tube_t0 = ts - td;
  const float tube_t1 = ts + td;
  const float sf = dot(s, f);
  const float sd = dot(s, d);
  float cap_t0 = -1e20f, cap_t1 = 1e20f;
  float cap_t_v0 = cap_t0;
  if (sd == 0.f) {
    if (dot(ray_org - v0, v1 - v0) < 0.f) return false;
    if (dot(ray_org - v1, v0 - v1) < 0.f) return false;
  } else {
    const float rsd = 1.f / sd;
    cap_t_v0 = sf * rsd;
    const float cap_t_v1 = cap_t_v0 + s2 * rsd;
    cap_t0 = fmn(cap_t_v0, cap_t_v1);
    cap_t1 = fmx(cap_t_v0, cap_t_v1);
  }
  const float t0 = fmx(cap_t0, tube_t0);
  const float t1 = fmn(cap_t1, tube_t1);
  if (t0 > t1) return false;
  bool front;
  if (tmin <= t0 && t0 <= tmax) front = true;
  else if (tmin <= t1 && t1 <= tmax) front = false;
  else return false;
  t = front ? t0 : t1;
  if (N) {
    if (front) td = -td;
    const float u = ((t * sd - sf) * 1.f) / s2;
    const bool cap = front ? (t0 == cap_t0) : (t0 == cap_t1);
    if (cap) *N = (cap_t0 == cap_t_v0) ? v3{-s.x, -s.y, -s.z} : s;
    else *N = (td * d - fp) - u * s;
  }
  return true;
}

// CapsulesPrograms::intersect (native/geometry/Capsules.dev.cu:154-212) with
// intersectRoundedCone (:19-98): the rounded cone between the end spheres
// (A.xyz, A.w) and (B.xyz, B.w), in the ray's object frame.  The origin
// first moves to 0.99 of the ray's entry distance into the box of both end
// spheres; then the nearest of the body and the two end spheres below the
// bound tmax is taken (each candidate must lie beyond 1e-6 from the moved
// origin).  t: the distance from the original origin; *N: the object normal
// (normalized as the reference normalizes it); *P: the object-space point
// (moved origin + moved distance, as the reference computes it).  Written
// operation by operation as the reference evaluates it.
BNX_D bool capsule_intersect(v3 ro, v3 rd, float4 A, float4 B, float tmax, float &t, v3 *N, v3 *P)
{
  const v3 pa = xyz(A), pb = xyz(B);
  const float ra = A.w, rb = B.w;
  // getBounds (:100-110) and Ray.h's boxTest from t0 = 0, t1 = tmax
  const v3 blo{fmn(pa.x - ra, pb.x - rb), fmn(pa.y - ra, pb.y - rb), fmn(pa.z - ra, pb.z - rb)};
  const v3 bhi{fmx(pa.x + ra, pb.x + rb), fmx(pa.y + ra, pb.y + rb), fmx(pa.z + ra, pb.z + rb)};
  const v3 rcp{1.f / rd.x, 1.f / rd.y, 1.f / rd.z};
  const v3 tlo{(blo.x - ro.x) * rcp.x, (blo.y - ro.y) * rcp.y, (blo.z - ro.z) * rcp.z};
  const v3 thi{(bhi.x - ro.x) * rcp.x, (bhi.y - ro.y) * rcp.y, (bhi.z - ro.z) * rcp.z};
  const float t0 = fmx(0.f, fmx(fmx(fmn(tlo.x, thi.x), fmn(tlo.y, thi.y)), fmn(tlo.z, thi.z)));
  const float t1 = fmn(tmax, fmn(fmn(fmx(tlo.x, thi.x), fmx(tlo.y, thi.y)), fmx(tlo.z, thi.z)));
  if (!(t0 <= t1)) return false;
  const float t_move = .99f * t0;
  const v3 o = ro + t_move * rd;
  const float bound = tmax - t_move;
  float hit_t = bound;
  // intersectRoundedCone: no capsule when one end sphere holds the other
  const float dist_ab = length(pa - pb);
  if (dist_ab + ra <= rb) return false;
  if (dist_ab + rb <= ra) return false;
  const v3 ba = pb - pa, oa = o - pa, ob = o - pb;
  const float rr = ra - rb;
  const float m0 = dot(ba, ba), m1 = dot(ba, oa), m2 = dot(ba, rd), m3 = dot(rd, oa);
  const float m5 = dot(oa, oa), m6 = dot(ob, rd), m7 = dot(ob, ob);
  const float d2 = m0 - rr * rr;
  const float k2 = d2 - m2 * m2;
  if (k2 == 0.f) return false;
  const float k1 = (d2 * m3 - m1 * m2) + (m2 * rr) * ra;
  const float k0 = ((d2 * m5 - m1 * m1) + ((m1 * rr) * ra) * 2.f) - (m0 * ra) * ra;
  const float h = k1 * k1 - k0 * k2;
  if (h < 0.f) return false;
  bool had = false;
  v3 n{0.f, 0.f, 0.f};
  float tc = (-sqrtf(h) - k1) / k2;
  const float y = (m1 - ra * rr) + tc * m2;
  if (y > 0.f && y < d2 && tc > 1e-6f && tc < hit_t) {
    hit_t = tc;
    n = d2 * (oa + tc * rd) - ba * y;
    had = true;
  }
  // the end spheres
  const float h1 = (m3 * m3 - m5) + ra * ra;
  if (h1 > 0.f) {
    tc = -m3 - sqrtf(h1);
    if (tc > 1e-6f && tc < hit_t) {
      hit_t = tc;
      const v3 q = (o + tc * rd) - pa;
      n = v3{q.x / ra, q.y / ra, q.z / ra};
      had = true;
    }
  }
  const float h2 = (m6 * m6 - m7) + rb * rb;
  if (h2 > 0.f) {
    tc = -m6 - sqrtf(h2);
    if (tc > 1e-6f && tc < hit_t) {
      hit_t = tc;
      const v3 q = (o + tc * rd) - pb;
      n = v3{q.x / rb, q.y / rb, q.z / rb};
      had = true;
    }
  }
  if (!had) return false;
  if (hit_t < 0.f || hit_t > bound) return false;
  if (N) *N = normalize(n);
  if (P) *P = o + hit_t * rd;
  t = hit_t + t_move;
  return true;
}

// ConesPrograms::intersect (native/geometry/Cones.dev.cu:60-266): the
// capped cone from (p0, ra) to (p1, rb) in the ray's object frame.  The cap
// on the side of the origin first (a hit there in (tmin, tmax) is taken
// without looking at the body), then the body's nearer root.  *N: the
// unnormalized object normal; *lerp: the attribute weight of the far end
// (0 at the first cap, 1 at the second, y / m0 clamped on the body).
BNX_D bool cone_intersect(v3 ro, v3 rd, float4 A, float4 B, float tmin, float tmax, float &t, v3 *N, float *lerp)
{
  const v3 p0 = xyz(A), p1 = xyz(B);
  const float ra = A.w, rb = B.w;
  const v3 ba = p1 - p0, oa = ro - p0, ob = ro - p1;
  const float m0 = dot(ba, ba), m1 = dot(oa, ba), m2 = dot(ob, ba), m3 = dot(rd, ba);
  // inversesqrt of the reference: 1. / sqrtf(f), a double division rounded to float
  const float isq = (float)(1.0 / (double)sqrtf(m0));
  if (m1 < 0.f) {
    const v3 q = oa * m3 - rd * m1;
    if (dot(q, q) < ((ra * ra) * m3) * m3) {
      const float tc = -m1 / m3;
      if (tc > tmin && tc < tmax) {
        t = tc;
        if (N) *N = -ba * isq;
        if (lerp) *lerp = 0.f;
        return true;
      }
    }
  } else if (m2 > 0.f) {
    const v3 q = ob * m3 - rd * m2;
    if (dot(q, q) < ((rb * rb) * m3) * m3) {
      const float tc = -m2 / m3;
      if (tc > tmin && tc < tmax) {
        t = tc;
        if (N) *N = ba * isq;
        if (lerp) *lerp = 1.f;
        return true;
      }
    }
  }
  const float m4 = dot(rd, oa), m5 = dot(oa, oa);
  const float rr = ra - rb;
  const float hy = m0 + rr * rr;
  const float k2 = m0 * m0 - (m3 * m3) * hy;
  const float k1 = ((m0 * m0) * m4 - (m1 * m3) * hy) + (m0 * ra) * ((rr * m3) * 1.f);
  const float k0 = ((m0 * m0) * m5 - (m1 * m1) * hy) + (m0 * ra) * ((rr * m1) * 2.f - m0 * ra);
  const float h = k1 * k1 - k2 * k0;
  if (h < 0.f) return false;
  const float tc = (-k1 - sqrtf(h)) / k2;
  const float y = m1 + tc * m3;
  if (!(y > 0.f && y < m0 && tc > tmin && tc < tmax)) return false;
  t = tc;
  if (N) *N = m0 * (m0 * (oa + tc * rd) + (rr * ba) * ra) - (ba * hy) * y;
  if (lerp) *lerp = fmx(0.f, fmn(1.f, y / m0));
  return true;
}

// transformNormalFromObjectToWorldSpace: the transpose of the world-to-
// object linear part applied to n
BNX_D v3 xfm_normal_rows(const float4 *w2o, v3 n)
{
  return v3{(w2o[0].x * n.x + w2o[1].x * n.y) + w2o[2].x * n.z, (w2o[0].y * n.x + w2o[1].y * n.y) + w2o[2].y * n.z,
            (w2o[0].z * n.x + w2o[1].z * n.y) + w2o[2].z * n.z};
}

// The hit attributes of a capsule or cone hit at distance t_hit, found
// again from the geometry's records in the instance's object frame (o, d:
// the object ray): the object point and normal, the world point and normal,
// and the weight of the second end for attribute interpolation.
struct PairHit {
  v3 osP, osN, P, Nw;
  float lerp;
};

// Capsules.dev.cu:212-262: the object point of capsule_intersect, world
// point moved off the surface by max(1, |P|_inf) * 1e-5 along the world
// normal, and the attribute weight the reference computes from the angle
// between v0->P and v0->v1 (clamped to [0, 1])
BNX_D void capsule_hit(const InstDev &I, float4 A, float4 B, v3 o, v3 d, float t_hit, PairHit &h)
{
  float t = t_hit;
  v3 N{1.f, 0.f, 0.f}, objP = o + t_hit * d;
  (void)capsule_intersect(o, d, A, B, t_hit * 1.0001f, t, &N, &objP);
  const v3 v0 = xyz(A), v1 = xyz(B);
  const float l01 = length(v1 - v0), lp0 = length(objP - v0);
  float lerp = 0.f;
  if (!(l01 < 1e-8f || lp0 < 1e-8f)) lerp = fmx(0.f, fmn(1.f, dot(objP - v0, v1 - v0) / (lp0 * l01)));
  h.osP = objP;
  h.osN = normalize(N);
  if (I.identity) {
    h.P = objP;
    h.Nw = normalize(N);
  } else {
    h.P = xfm_point_rows(I.o2w, objP);
    h.Nw = normalize(xfm_normal_rows(I.w2o, N));
  }
  float eps = 1.f;
  eps = fmx(eps, fabsf(h.P.x));
  eps = fmx(eps, fabsf(h.P.y));
  eps = fmx(eps, fabsf(h.P.z));
  eps *= 1e-5f;
  h.P = h.P + eps * h.Nw;
  h.lerp = lerp;
}

// Cones.dev.cu:117-262: P = o + t d (not moved off the surface), the
// normals normalized, the attribute weight of cone_intersect
BNX_D void cone_hit(const InstDev &I, float4 A, float4 B, v3 o, v3 d, float t_hit, PairHit &h)
{
  float t = t_hit, lerp = 0.f;
  v3 N{1.f, 0.f, 0.f};
  (void)cone_intersect(o, d, A, B, 0.f, t_hit * 1.0001f, t, &N, &lerp);
  const v3 objP = o + t_hit * d;
  h.osP = objP;
  h.osN = normalize(N);
  if (I.identity) {
    h.P = objP;
    h.Nw = normalize(h.osN);
  } else {
    h.P = xfm_point_rows(I.o2w, objP);
    h.Nw = normalize(xfm_normal_rows(I.w2o, h.osN));
  }
  h.lerp = lerp;
}

// ------------------------------------------------------- ordered floats
BNX_D int f2o(float f)
{
  int i = __float_as_int(f);
  return i >= 0 ? i : (i ^ 0x7fffffff);
}
BNX_D float o2f(int i) { return __int_as_float(i >= 0 ? i : (i ^ 0x7fffffff)); }

// ----------------------------------------------------------- wave helpers
BNX_D int lane_id() { return __lane_id(); }
BNX_D uint64_t lanemask_lt() { return (1ull << __lane_id()) - 1ull; }
BNX_D uint32_t wave_sum(uint32_t x)
{
  for (int off = 32; off > 0; off >>= 1) x += __shfl_xor(x, off);
  return x;
}

// Statistics counters, spread over STAT_SLOTS addresses per counter so that
// the per-wave atomics of a launch do not serialise on one L2 line; the host
// sums the slots.
constexpr int STAT_SLOTS = 256;
enum { STAT_SHADOW = 0, STAT_NODES, STAT_TRIS, STAT_SPHERES, STAT_SAMPLES, STAT_CELLS, STAT_UNODES, STAT_UTETS,
       STAT_COUNT };
// call with the wave-wide total from lane 0
BNX_D void stat_add(unsigned long long *stats, int idx, uint32_t v)
{
  if (!stats || v == 0) return;
  uint32_t slot = (blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6)) & (STAT_SLOTS - 1);
  atomicAdd(stats + idx * STAT_SLOTS + slot, (unsigned long long)v);
}

} // namespace bnx
