// oracle.cpp -- CPU oracle: BVH spec, traversal, ray generation, shading.
//
// TEST INFRASTRUCTURE ONLY (see oracle.h).
//
// Parity status per function (how each is pinned; tests/ holds the checks):
//   RNGSeed/Random2, fnv_hash ....... closed form: tests/test_oracle_cpu.py
//                                     re-derives the sequences in Python
//                                     integers from barney-common.h:21-108.
//   f32_to_f16 / f16_to_f32 ......... numpy.float16 (IEEE RNE) on all
//                                     65536 halves and random floats.
//   camera (Camera.cu:82-101) ....... closed form: centre pixel looks
//                                     along -z; corner directions from the
//                                     fovy formula.
//   tri_test (hiprt/TraceInterface.h:38-76) ... closed forms (axis-aligned
//                                     quads, t = 4 on the unit icosphere at
//                                     the centre pixel), brute force.
//   sphere_test (Spheres.dev.cu:114-166) ..... analytic roots.
//   build_bvh ....................... invariants (each prim in exactly one
//                                     leaf, child box == union of its prims,
//                                     leaf size), SAH cost <= median split.
//   trace_bvh ....................... == trace_brute bit for bit on random
//                                     scenes, including constructed ties.
//   shade (shadeRays.cu:474-880) .... white-furnace and closed-form
//                                     expectations on analytic scenes; no
//                                     reference-held vectors exist for it.
//   owl::common::frame(N) ........... parity unpinned against the reference
//                                     (un-vendored OWL); restated as Duff et
//                                     al. 2017 and pinned by orthonormality.
//   env map (env.cpp) ............... tests/test_env_cpu.py: texture fetch
//                                     against a float64 bilinear lookup (texel
//                                     centres exact, seam, poles); CDFs
//                                     against closed forms for row-only and
//                                     column-only maps; samples map back to
//                                     their texel with pdf_x pdf_y /
//                                     (2 pi^2 sin theta); the pdf integrates
//                                     to 1; E[L / pdf] = integral of L; a
//                                     sphere's pixels match the MIS
//                                     expectation integral.  atan2 / acos
//                                     (bnx_numeric.h) within 3 ulp of libm.
#include "oracle.h"

#include <algorithm>
#include <thread>
#include <cstdio>

namespace orc {

// ======================================================================
// half <-> float, IEEE round to nearest even
// ======================================================================
uint16_t f32_to_f16(float f)
{
  uint32_t x = fbits(f);
  uint32_t sign = (x >> 16) & 0x8000u;
  uint32_t mag = x & 0x7fffffffu;
  if (mag >= 0x7f800000u) {                       // inf / nan
    return (uint16_t)(sign | 0x7c00u | (mag > 0x7f800000u ? 0x200u : 0u));
  }
  if (mag >= 0x477ff000u) return (uint16_t)(sign | 0x7c00u);   // rounds to inf
  if (mag < 0x38800000u) {                        // result is subnormal or zero
    if (mag < 0x33000000u) return (uint16_t)sign; // < half of smallest subnormal
    uint32_t e = mag >> 23;
    uint32_t m = (mag & 0x7fffffu) | 0x800000u;
    uint32_t shift = 126u - e;                    // 14 + (113 - e) - ... see below
    // value = m * 2^(e-150); subnormal unit = 2^-24 -> q = m * 2^(e-126)
    uint32_t q = m >> shift;
    uint32_t rem = m & ((1u << shift) - 1u);
    uint32_t halfway = 1u << (shift - 1u);
    if (rem > halfway || (rem == halfway && (q & 1u))) q++;
    return (uint16_t)(sign | q);
  }
  uint32_t e = (mag >> 23) - 112u;                // rebias 127 -> 15
  uint32_t m = mag & 0x7fffffu;
  uint32_t q = (e << 10) | (m >> 13);
  uint32_t rem = m & 0x1fffu;
  if (rem > 0x1000u || (rem == 0x1000u && (q & 1u))) q++;
  return (uint16_t)(sign | q);
}

float f16_to_f32(uint16_t h)
{
  uint32_t sign = ((uint32_t)h & 0x8000u) << 16;
  uint32_t e = (h >> 10) & 0x1fu;
  uint32_t m = h & 0x3ffu;
  uint32_t x;
  if (e == 0) {
    if (m == 0) x = sign;
    else {                                        // subnormal: normalize
      int k = 0;
      while (!(m & 0x400u)) { m <<= 1; k++; }
      m &= 0x3ffu;
      x = sign | ((uint32_t)(113 - k) << 23) | (m << 13);
    }
  } else if (e == 31) {
    x = sign | 0x7f800000u | (m << 13);
  } else {
    x = sign | ((e + 112u) << 23) | (m << 13);
  }
  float f;
  memcpy(&f, &x, 4);
  return f;
}

// ======================================================================
// binned-SAH BVH (the build specification the GPU builder implements)
//
// Reference: cuBQL gpuBuilder with leaf <= 4 (rtcore/cuda/Group.cu:382,
// 503-504) and leaf == 1 for instances (:198); cuBQL's own algorithm is
// un-vendored, so this is the build's own binned SAH: 16 bins per axis over
// the task's centroid bounds, cost = area(L)*|L| + area(R)*|R|, lowest cost
// wins (ties: lower axis, then lower split), stable partition; median split
// when no axis has a valid plane or the depth reaches MAX_SAH_DEPTH.
// ======================================================================
namespace {

struct BuildCtx {
  const std::vector<Box> *boxes;
  std::vector<vec3> cent;
  std::vector<int32_t> order, tmp;
  int max_leaf;
  Bvh *out;
};

void task_bounds(const BuildCtx &c, int b, int e, Box &box, Box &cbox)
{
  box = Box();
  cbox = Box();
  for (int i = b; i < e; i++) {
    int p = c.order[i];
    box.grow((*c.boxes)[p]);
    cbox.grow(c.cent[p]);
  }
}

inline int bin_of(float c, float lo, float scale)
{
  int b = (int)((c - lo) * scale);
  return b > SAH_BINS - 1 ? SAH_BINS - 1 : b;
}

// returns split position (b < mid < e)
int split_task(BuildCtx &c, int b, int e, int depth, const Box &cbox)
{
  int n = e - b;
  int best_axis = -1, best_bin = -1;
  float best_cost = INFINITY;
  if (depth < MAX_SAH_DEPTH) {
    for (int a = 0; a < 3; a++) {
      float ext = cbox.hi[a] - cbox.lo[a];
      if (!(ext >= 1e-30f)) continue;
      float scale = (float)SAH_BINS / ext;
      Box bb[SAH_BINS];
      int bc[SAH_BINS] = {0};
      for (int i = b; i < e; i++) {
        int p = c.order[i];
        int k = bin_of(c.cent[p][a], cbox.lo[a], scale);
        bc[k]++;
        bb[k].grow((*c.boxes)[p]);
      }
      float areaL[SAH_BINS], areaR[SAH_BINS];
      int cntL[SAH_BINS], cntR[SAH_BINS];
      Box acc;
      int cn = 0;
      for (int j = 0; j < SAH_BINS - 1; j++) {
        acc.grow(bb[j]);
        cn += bc[j];
        areaL[j] = half_area(acc);
        cntL[j] = cn;
      }
      acc = Box();
      cn = 0;
      for (int j = SAH_BINS - 1; j > 0; j--) {
        acc.grow(bb[j]);
        cn += bc[j];
        areaR[j - 1] = half_area(acc);
        cntR[j - 1] = cn;
      }
      for (int j = 0; j < SAH_BINS - 1; j++) {
        if (cntL[j] == 0 || cntR[j] == 0) continue;
        float cost = areaL[j] * (float)cntL[j] + areaR[j] * (float)cntR[j];
        if (cost < best_cost) {
          best_cost = cost;
          best_axis = a;
          best_bin = j;
        }
      }
    }
  }
  if (best_axis < 0) return b + n / 2;           // median split in current order
  float ext = cbox.hi[best_axis] - cbox.lo[best_axis];
  float scale = (float)SAH_BINS / ext;
  int nl = 0, nr = 0;
  for (int i = b; i < e; i++) {                   // stable partition
    int p = c.order[i];
    if (bin_of(c.cent[p][best_axis], cbox.lo[best_axis], scale) <= best_bin) c.order[b + nl++] = p;
    else c.tmp[nr++] = p;
  }
  for (int i = 0; i < nr; i++) c.order[b + nl + i] = c.tmp[i];
  return b + nl;
}

void set_child(Node &nd, int side, const Box &box, int ref, int cnt)
{
  float *lo = side == 0 ? nd.lo0 : nd.lo1;
  float *hi = side == 0 ? nd.hi0 : nd.hi1;
  for (int k = 0; k < 3; k++) { lo[k] = box.lo[k]; hi[k] = box.hi[k]; }
  nd.ref[side] = ref;
  nd.cnt[side] = cnt;
}

void build_rec(BuildCtx &c, int b, int e, int depth, const Box &cbox, int node)
{
  int mid = split_task(c, b, e, depth, cbox);
  int rng[2][2] = {{b, mid}, {mid, e}};
  for (int s = 0; s < 2; s++) {
    int cb = rng[s][0], ce = rng[s][1];
    Box box, cb_box;
    task_bounds(c, cb, ce, box, cb_box);
    if (ce - cb <= c.max_leaf) {
      set_child(c.out->nodes[node], s, box, cb, ce - cb);
      c.out->max_depth = std::max(c.out->max_depth, depth + 1);
    } else {
      int idx = (int)c.out->nodes.size();
      c.out->nodes.push_back(Node());
      set_child(c.out->nodes[node], s, box, idx, 0);
      build_rec(c, cb, ce, depth + 1, cb_box, idx);
    }
  }
}

} // namespace

void build_bvh(const std::vector<Box> &boxes, int max_leaf, Bvh &out)
{
  int n = (int)boxes.size();
  out.nodes.clear();
  out.max_depth = 0;
  out.nodes.push_back(Node());
  BuildCtx c;
  c.boxes = &boxes;
  c.max_leaf = max_leaf;
  c.out = &out;
  c.cent.resize(n);
  for (int i = 0; i < n; i++)
    c.cent[i] = (boxes[i].lo + boxes[i].hi) * 0.5f;
  c.order.resize(n);
  c.tmp.resize(n);
  for (int i = 0; i < n; i++) c.order[i] = i;
  Box box, cbox;
  task_bounds(c, 0, n, box, cbox);
  out.root = box;
  Node &root = out.nodes[0];
  set_child(root, 0, Box(), 0, -1);
  set_child(root, 1, Box(), 0, -1);
  if (n == 0) {
  } else if (n <= max_leaf) {
    set_child(out.nodes[0], 0, box, 0, n);
    out.max_depth = 1;
  } else {
    build_rec(c, 0, n, 0, cbox, 0);
  }
  out.order = c.order;
}

// ======================================================================
// scene build: BLAS per group, TLAS over instances
// (ModelSlot.cpp:128-155 flattening; rtcore/cuda/Group.cu:144-562)
// ======================================================================

void build_scene(Scene &s)
{
  for (auto &g : s.groups) {
    g.prim_geom.clear();
    g.prim_id.clear();
    std::vector<Box> boxes;
    for (int gi = 0; gi < (int)g.geoms.size(); gi++) {
      const Geom &geom = s.geoms[g.geoms[gi]];
      int np = geom.num_prims();
      for (int p = 0; p < np; p++) {
        g.prim_geom.push_back(gi);
        g.prim_id.push_back(p);
        Box b;
        if (geom.type == 0) {
          // createTriangleBounds (rtcore/cuda/Group.cu:91-112)
          for (int k = 0; k < 3; k++) b.grow(geom.vertices[geom.indices[3 * p + k]]);
        } else if (geom.type == 1) {
          // SpheresPrograms::bounds (Spheres.dev.cu:29-40)
          vec3 ctr = geom.centers[p];
          float r = geom.radii.empty() ? geom.default_radius : geom.radii[p];
          b = Box(ctr - vec3(r), ctr + vec3(r));
        } else if (geom.type == 2) {
          // CylindersPrograms::bounds (native/geometry/Cylinders.dev.cu:69-85)
          vec3 a = geom.vertices[geom.indices[2 * p]], c = geom.vertices[geom.indices[2 * p + 1]];
          float r = geom.radii.empty() ? geom.default_radius : geom.radii[p];
          b = Box(vmin(a - vec3(r), c - vec3(r)), vmax(a + vec3(r), c + vec3(r)));
        } else {
          // capsules: getBounds (Capsules.dev.cu:100-110), the boxes of both
          // end spheres; cones the same (ConesPrograms::bounds,
          // Cones.dev.cu:21-42, builds the second box as (pb - ra, pb + rb),
          // which need not hold the cone; a hit outside its primitive's box
          // would depend on traversal order)
          const int ia = geom.indices[2 * p], ib = geom.indices[2 * p + 1];
          const vec3 a = geom.vertices[ia], c = geom.vertices[ib];
          const float ra = geom.radii[ia], rb = geom.radii[ib];
          b = Box(vmin(a - vec3(ra), c - vec3(rb)), vmax(a + vec3(ra), c + vec3(rb)));
        }
        boxes.push_back(b);
      }
    }
    build_bvh(boxes, 4, g.bvh);
  }
  // computeInstanceBounds (rtcore/cuda/Group.cu:125-137)
  std::vector<Box> ib;
  for (auto &in : s.insts) {
    const Box &rb = s.groups[in.group].bvh.root;
    Box b;
    if (in.o2w.identity) b = rb;
    else if (rb.lo.x <= rb.hi.x) {
      for (int k = 0; k < 8; k++) {
        vec3 c((k & 1) ? rb.hi.x : rb.lo.x, (k & 2) ? rb.hi.y : rb.lo.y, (k & 4) ? rb.hi.z : rb.lo.z);
        b.grow(xfm_point(in.o2w, c));
      }
    }
    ib.push_back(b);
  }
  build_bvh(ib, 1, s.tlas);
}

// ======================================================================
// hit attributes
// ======================================================================
namespace {

// AttributeArray::DD::valueAt (render/GeometryAttributes.h:68-91)
void value_at(const Attribute &a, int i, float out[4])
{
  const float *p = &a.data[(size_t)a.comps * i];
  out[0] = out[1] = out[2] = 0.f;
  out[3] = 1.f;
  for (int k = 0; k < a.comps && k < 4; k++) out[k] = p[k];
}

// (w0 a + w1 b) + w2 c, component by component (vec4f operators)
void interpolate(const Attribute &a, const Corners &c, bool face_varying, float out[4])
{
  const int *ix = face_varying ? c.fi : c.vi;
  if (c.n == 1) {
    value_at(a, ix[0], out);
    return;
  }
  float va[4], vb[4], vc[4];
  value_at(a, ix[0], va);
  value_at(a, ix[1], vb);
  for (int k = 0; k < 4; k++) out[k] = c.w[0] * va[k] + c.w[1] * vb[k];
  if (c.n == 3) {
    value_at(a, ix[2], vc);
    for (int k = 0; k < 4; k++) out[k] = out[k] + c.w[2] * vc[k];
  }
}

} // namespace

void hit_attribute(const Scene &s, const Geom &g, int k, int inst, int prim, const Corners &c, const float dflt[4],
                   float out[4])
{
  const Attribute &a = g.attr[k];
  switch (a.scope) {
  case BNX_SCOPE_CONSTANT: memcpy(out, a.value, 4 * sizeof(float)); return;
  case BNX_SCOPE_PER_PRIM: value_at(a, prim, out); return;
  case BNX_SCOPE_PER_VERTEX: interpolate(a, c, false, out); return;
  case BNX_SCOPE_FACE_VARYING: interpolate(a, c, true, out); return;
  default: break;
  }
  // no geometry attribute: the instance's value (Geometry.h:103-110)
  if (k < 5 && !s.inst_attr[k].empty()) {
    memcpy(out, &s.inst_attr[k][4 * (size_t)inst], 4 * sizeof(float));
    return;
  }
  memcpy(out, dflt, 4 * sizeof(float));
}

void material_color(const Scene &s, const Geom &g, int inst, int prim, const Corners &c, const float color0[4],
                    vec3 P, vec3 osP, float out[4])
{
  const float none[4] = {0.f, 0.f, 0.f, 1.f};
  const int k = g.mat.color_from_attribute;
  if (k == BNX_ATTRIB_NONE) {
    // a value; NaN: unset (HostMaterial.h:104-106)
    memcpy(out, std::isnan(g.mat.color[0]) ? none : g.mat.color, 4 * sizeof(float));
  } else if (k == BNX_ATTRIB_COLOR) {
    hit_attribute(s, g, 4, inst, prim, c, color0, out);
  } else if (k >= BNX_ATTRIB_0 && k < BNX_ATTRIB_0 + 4) {
    hit_attribute(s, g, k - BNX_ATTRIB_0, inst, prim, c, none, out);
  } else if (k == BNX_ATTRIB_WORLD_POSITION || k == BNX_ATTRIB_OBJECT_POSITION) {
    // HitAttributes::get (render/HitAttributes.h:61-80)
    const vec3 p = k == BNX_ATTRIB_WORLD_POSITION ? P : osP;
    out[0] = p.x;
    out[1] = p.y;
    out[2] = p.z;
    out[3] = 1.f;
  } else {
    memcpy(out, none, sizeof(none));
  }
}

float material_alpha(const Material &m, const float base[4])
{
  if (m.type == BNX_MATERIAL_PBR) return f16_to_f32(f32_to_f16(((1.f - m.transmission) * base[3]) * m.opacity));
  return base[3] * m.opacity;
}

Corners tri_corners(const int32_t *ix, int prim, float u, float v)
{
  Corners c;
  c.n = 3;
  for (int k = 0; k < 3; k++) {
    c.vi[k] = ix[k];
    c.fi[k] = 3 * prim + k;
  }
  c.w[0] = (1.f - u) - v;
  c.w[1] = u;
  c.w[2] = v;
  return c;
}

// ======================================================================
// traversal
// ======================================================================
namespace {


// lexicographic (inst, geom, prim) order: the lower ID wins a tie in t
inline bool id_less(int i0, int g0, int p0, int i1, int g1, int p1)
{
  if (i0 != i1) return i0 < i1;
  if (g0 != g1) return g0 < g1;
  return p0 < p1;
}

inline bool accept_t(float t, const Hit &best, int inst, int geom, int prim)
{
  if (t < best.t) return true;
  if (t > best.t) return false;
  return best.prim < 0 ? false : id_less(inst, geom, prim, best.inst, best.geom, best.prim);
}

// barney's triangle test (rtcore/hiprt/TraceInterface.h:38-76), with the
// acceptance bound replaced by the tie rule above
bool tri_test(vec3 org, vec3 dir, float tmin, vec3 v0, vec3 v1, vec3 v2, Hit &best, int inst, int geom,
              int prim)
{
  vec3 e1 = v1 - v0;
  vec3 e2 = v2 - v0;
  vec3 N = cross(e1, e2);
  float dn = dot(dir, N);
  if (fabsf(dn) < 1e-12f) return false;
  float t = -dot(org - v0, N) / dn;
  if (t <= tmin) return false;
  if (!accept_t(t, best, inst, geom, prim)) return false;
  vec3 P = (org - v0) + t * dir;
  float e1u, e2u, Pu, e1v, e2v, Pv;
  if (fabsf(N.x) >= fmx(fabsf(N.y), fabsf(N.z))) {
    e1u = e1.y; e2u = e2.y; Pu = P.y;
    e1v = e1.z; e2v = e2.z; Pv = P.z;
  } else if (fabsf(N.y) > fabsf(N.z)) {
    e1u = e1.x; e2u = e2.x; Pu = P.x;
    e1v = e1.z; e2v = e2.z; Pv = P.z;
  } else {
    e1u = e1.x; e2u = e2.x; Pu = P.x;
    e1v = e1.y; e2v = e2.y; Pv = P.y;
  }
  float den = e1u * e2v - e1v * e2u;
  if (den == 0.f) return false;
  float u = (Pu * e2v - Pv * e2u) / den;
  float v = (e1u * Pv - e1v * Pu) / den;
  if (u < 0.f || v < 0.f || (u + v) >= 1.f) return false;
  best.t = t;
  best.u = u;
  best.v = v;
  best.inst = inst;
  best.geom = geom;
  best.prim = prim;
  return true;
}

// SpheresPrograms::intersect (Spheres.dev.cu:114-166), "move the origin";
// the nearest root above tmin is accepted under the tie rule
bool sphere_test(vec3 org, vec3 dir, float tmin, vec3 center, float radius, Hit &best, int inst, int geom,
                 int prim)
{
  float t_move = fmx(0.f, length(center - org) - 3.f * radius);
  vec3 o = org + t_move * dir;
  float tmin_l = fmx(0.f, tmin - t_move);
  vec3 oc = o - center;
  float a = dot(dir, dir);
  float b = dot(oc, dir);
  float c = dot(oc, oc) - radius * radius;
  float disc = b * b - a * c;
  if (disc < 0.f) return false;
  float sq = sqrtf(disc);
  float r1 = (-b - sq) / a;
  float r2 = (-b + sq) / a;
  float cand;
  if (r1 > tmin_l) cand = r1;
  else if (r2 > tmin_l) cand = r2;
  else return false;
  float t = cand + t_move;
  if (!accept_t(t, best, inst, geom, prim)) return false;
  best.t = t;
  best.u = cand;
  best.v = 0.f;
  best.inst = inst;
  best.geom = geom;
  best.prim = prim;
  return true;
}

// The reference's cylinder: s the axis, f = v0 - org; the tube's roots
// around the ray's closest approach to the axis (ts +- td), clipped to the
// slab between the end caps; unnormalized normal of the tube (td d - f' - u s)
// or of a cap (+-s).
bool cylinder_test_geom(vec3 ray_org, vec3 ray_dir, vec3 v0, vec3 v1, float radius, float tmin, float tmax,
                        float &t, vec3 *N)
{
  const vec3 d = ray_dir;
  const vec3 s = v1 - v0;
  const vec3 sxd = cross(s, d);
  const float a = dot(sxd, sxd);
  if (a == 0.f) return false;
  const vec3 f = v0 - ray_org;
  const vec3 sxf = cross(s, f);
  const float ra = 1.f / a;
  const float ts = dot(sxd, sxf) * ra;
  const vec3 fp = f - ts * d;
  const float s2 = dot(s, s);
  const vec3 perp = cross(s, fp);
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
    // the reference tests t0 == cap_t1 for the back hit (Cylinders.dev.cu:195)
    const bool cap = front ? (t0 == cap_t0) : (t0 == cap_t1);
    if (cap) *N = (cap_t0 == cap_t_v0) ? vec3(-s.x, -s.y, -s.z) : s;
    else *N = (td * d - fp) - u * s;
  }
  return true;
}

// Capsules.dev.cu:154-212 + intersectRoundedCone (:19-98), step by step
bool capsule_test_geom(vec3 ro, vec3 rd, vec3 pa, float ra, vec3 pb, float rb, float tmax, float &t, vec3 *N,
                       vec3 *P)
{
  // getBounds (:100-110), then Ray.h's boxTest with t0 = 0, t1 = tmax
  const Box bb(vmin(pa - vec3(ra), pb - vec3(rb)), vmax(pa + vec3(ra), pb + vec3(rb)));
  const vec3 rcp(1.f / rd.x, 1.f / rd.y, 1.f / rd.z);
  const vec3 tlo = (bb.lo - ro) * rcp, thi = (bb.hi - ro) * rcp;
  const float t0 = fmx(0.f, reduce_max(vmin(tlo, thi)));
  const float t1 = fmn(tmax, reduce_min(vmax(tlo, thi)));
  if (!(t0 <= t1)) return false;
  // move the origin to just before the box entry (:191-196)
  const float t_move = .99f * t0;
  const vec3 o = ro + t_move * rd;
  const float bound = tmax - t_move;
  float hit_t = bound;
  // one end sphere inside the other: no capsule (:30-36)
  const float dist_ab = length(pa - pb);
  if (dist_ab + ra <= rb) return false;
  if (dist_ab + rb <= ra) return false;
  const vec3 ba = pb - pa, oa = o - pa, ob = o - pb;
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
  vec3 n(0.f);
  // the body (:62-70)
  float tc = (-sqrtf(h) - k1) / k2;
  const float y = (m1 - ra * rr) + tc * m2;
  if (y > 0.f && y < d2 && tc > 1e-6f && tc < hit_t) {
    hit_t = tc;
    n = d2 * (oa + tc * rd) - ba * y;
    had = true;
  }
  // the two end spheres (:72-95)
  const float h1 = (m3 * m3 - m5) + ra * ra;
  if (h1 > 0.f) {
    tc = -m3 - sqrtf(h1);
    if (tc > 1e-6f && tc < hit_t) {
      hit_t = tc;
      n = ((o + tc * rd) - pa) / ra;
      had = true;
    }
  }
  const float h2 = (m6 * m6 - m7) + rb * rb;
  if (h2 > 0.f) {
    tc = -m6 - sqrtf(h2);
    if (tc > 1e-6f && tc < hit_t) {
      hit_t = tc;
      n = ((o + tc * rd) - pb) / rb;
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

// Cones.dev.cu:60-266, step by step; inversesqrt (:14) is 1. / sqrtf(f),
// a double division rounded to float
bool cone_test_geom(vec3 ro, vec3 rd, vec3 p0, float ra, vec3 p1, float rb, float tmin, float tmax, float &t,
                    vec3 *N, float *lerp)
{
  const vec3 ba = p1 - p0, oa = ro - p0, ob = ro - p1;
  const float m0 = dot(ba, ba), m1 = dot(oa, ba), m2 = dot(ob, ba), m3 = dot(rd, ba);
  const float isq = (float)(1.0 / (double)sqrtf(m0));
  if (m1 < 0.f) {
    // the first cap (:117-160)
    const vec3 q = oa * m3 - rd * m1;
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
    // the second cap (:162-205)
    const vec3 q = ob * m3 - rd * m2;
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
  // the body (:209-264)
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

// slab test, conservative by a relative 1e-5 on the far distance
inline bool box_test(vec3 org, vec3 inv, const float *lo, const float *hi, float tmin, float tmax, float &tnear)
{
  float tn = tmin, tf = tmax;
  for (int k = 0; k < 3; k++) {
    float t0 = (lo[k] - org[k]) * inv[k];
    float t1 = (hi[k] - org[k]) * inv[k];
    tn = fmx(tn, fmn(t0, t1));
    tf = fmn(tf, fmx(t0, t1));
  }
  tnear = tn;
  return tn <= tf * 1.00001f;
}

// OptixGlobals::hitOnInvisibleSide (native/render/OptixGlobals.h:36-45): a
// hit at t whose world position org + t dir lies on the negative side of
// the cut plane (nx, ny, nz, w) is rejected by every surface program
// (Triangles.dev.cu:39, Spheres.dev.cu:170, Cylinders.dev.cu:258); the
// plane is off when w <= -1e28
struct CutRay {
  const float *plane;   // null: no cut plane
  vec3 org, dir;        // the world ray (dir after the zero fix)
};
inline bool cut_rejects(const CutRay &c, float t)
{
  if (!c.plane || c.plane[3] <= -1e28f) return false;
  vec3 P = c.org + t * c.dir;
  return dot(load3(c.plane), P) + c.plane[3] < 0.f;
}

// whether a triangle candidate survives its opacity draw (Triangles.dev.cu:
// 112-133: osP from the vertices and barycentrics, the material's alpha from
// its hit attributes; the draw from a copy of the ray's state)
bool opacity_keeps(const Scene &s, const Geom &geom, int inst, int prim, float u, float v, uint64_t seed)
{
  const int32_t *ix = &geom.indices[3 * prim];
  const vec3 v0 = geom.vertices[ix[0]], v1 = geom.vertices[ix[1]], v2 = geom.vertices[ix[2]];
  const vec3 osP = (((1.f - u) - v) * v0 + u * v1) + v * v2;
  const Instance &in = s.insts[inst];
  const vec3 P = in.o2w.identity ? osP : xfm_point(in.o2w, osP);
  const float color0[4] = {0.f, 0.f, 0.f, 1.f};
  float base[4];
  material_color(s, geom, inst, prim, tri_corners(ix, prim, u, v), color0, P, osP, base);
  const float alpha = material_alpha(geom.mat, base);
  if (alpha < 1.f) {
    RNGSeed r;
    r.state = seed;
    uint32_t b;
    memcpy(&b, &osP.x, 4);
    r.next_h(b);
    memcpy(&b, &osP.y, 4);
    r.next_h(b);
    memcpy(&b, &osP.z, 4);
    r.next_h(b);
    Random2 rng(r, 290374u);
    if (rng() > alpha) return false;
  }
  return true;
}

// whether a capsule candidate survives its opacity draw (Capsules.dev.cu:
// 266-280): alpha from the material at the hit's attributes; the draw of
// Random(ray seed, hash(rtc instance, geometry, primitive)) on a copy of the
// ray's state
bool capsule_keeps(const Scene &s, const Geom &geom, int inst, int gi, int prim, vec3 org, vec3 dir, float t,
                   uint64_t seed)
{
  PairHit h;
  pair_hit(s.insts[inst], geom, prim, org, dir, t, h);
  const float color0[4] = {0.f, 0.f, 0.f, 1.f};
  float base[4];
  material_color(s, geom, inst, prim, pair_corners(geom, prim, h.lerp), color0, h.P, h.osP, base);
  const float alpha = material_alpha(geom.mat, base);
  if (alpha < 1.f) {
    RNGSeed copy;
    copy.state = seed;
    Random2 rng = Random2::from64(copy, fnv_hash3((uint32_t)inst, (uint32_t)gi, (uint32_t)prim));
    if (rng() > alpha) return false;
  }
  return true;
}

bool prim_test(const Scene &s, const Group &g, int slot, vec3 org, vec3 dir, float tmin, Hit &best, int inst,
               TraceCounters *ctr, const CutRay &cut, const uint64_t *seed, Hit *rej)
{
  int pidx = g.bvh.order[slot];
  int gi = g.prim_geom[pidx];
  int pid = g.prim_id[pidx];
  const Geom &geom = s.geoms[g.geoms[gi]];
  if (geom.type == 0) {
    if (ctr) ctr->tris++;
    const int32_t *ix = &geom.indices[3 * pid];
    Hit h = best;
    if (!tri_test(org, dir, tmin, geom.vertices[ix[0]], geom.vertices[ix[1]], geom.vertices[ix[2]], h, inst, gi,
                  pid))
      return false;
    if (cut_rejects(cut, h.t)) return false;
    if (seed && !opacity_keeps(s, geom, inst, pid, h.u, h.v, *seed)) {
      if (rej && (h.t < rej->t || (h.t == rej->t && (rej->prim < 0 || id_less(inst, gi, pid, rej->inst, rej->geom,
                                                                                rej->prim)))))
        *rej = h;
      return false;
    }
    best = h;
    return true;
  } else if (geom.type == 1) {
    if (ctr) ctr->spheres++;
    float r = geom.radii.empty() ? geom.default_radius : geom.radii[pid];
    Hit h = best;
    if (!sphere_test(org, dir, tmin, geom.centers[pid], r, h, inst, gi, pid)) return false;
    if (cut_rejects(cut, h.t)) return false;
    best = h;
    return true;
  } else {
    if (ctr) ctr->spheres++;
    float t;
    const int ia = geom.indices[2 * pid], ib = geom.indices[2 * pid + 1];
    if (geom.type == 2) {
      float r = geom.radii.empty() ? geom.default_radius : geom.radii[pid];
      if (!cylinder_intersect(org, dir, geom.vertices[ia], geom.vertices[ib], r, tmin, best.t, t, nullptr))
        return false;
    } else if (geom.type == 3) {
      // the reference tests against the closest hit so far and keeps only
      // strictly nearer candidates; the bound just above it lets a
      // candidate at exactly that distance reach the tie rule
      if (!capsule_intersect(org, dir, geom.vertices[ia], geom.radii[ia], geom.vertices[ib], geom.radii[ib],
                             best.t * 1.0001f, t, nullptr, nullptr))
        return false;
    } else {
      if (!cone_intersect(org, dir, geom.vertices[ia], geom.radii[ia], geom.vertices[ib], geom.radii[ib], tmin,
                          best.t * 1.0001f, t, nullptr, nullptr))
        return false;
    }
    if (!accept_t(t, best, inst, gi, pid)) return false;
    if (geom.type == 3 && seed && !capsule_keeps(s, geom, inst, gi, pid, org, dir, t, *seed)) return false;
    if (cut_rejects(cut, t)) return false;
    best.t = t;
    best.u = t;
    best.v = 0.f;
    best.inst = inst;
    best.geom = gi;
    best.prim = pid;
    return true;
  }
}

} // namespace

bool cylinder_intersect(vec3 ray_org, vec3 ray_dir, vec3 v0, vec3 v1, float radius, float tmin, float tmax,
                        float &t, vec3 *N)
{
  return cylinder_test_geom(ray_org, ray_dir, v0, v1, radius, tmin, tmax, t, N);
}

bool capsule_intersect(vec3 ro, vec3 rd, vec3 pa, float ra, vec3 pb, float rb, float tmax, float &t, vec3 *N,
                       vec3 *P)
{
  return capsule_test_geom(ro, rd, pa, ra, pb, rb, tmax, t, N, P);
}

bool cone_intersect(vec3 ro, vec3 rd, vec3 p0, float ra, vec3 p1, float rb, float tmin, float tmax, float &t,
                    vec3 *N, float *lerp)
{
  return cone_test_geom(ro, rd, p0, ra, p1, rb, tmin, tmax, t, N, lerp);
}

Corners pair_corners(const Geom &g, int prim, float lerp)
{
  Corners c;
  c.n = 2;
  c.vi[0] = g.indices[2 * prim];
  c.vi[1] = g.indices[2 * prim + 1];
  c.fi[0] = 2 * prim;
  c.fi[1] = 2 * prim + 1;
  c.w[0] = 1.f - lerp;
  c.w[1] = lerp;
  return c;
}

// transformNormalFromObjectToWorldSpace: the transpose of the world-to-
// object linear part
static vec3 xfm_normal(const Affine &w, vec3 n)
{
  return vec3(w.vx.x * n.x + w.vx.y * n.y + w.vx.z * n.z, w.vy.x * n.x + w.vy.y * n.y + w.vy.z * n.z,
              w.vz.x * n.x + w.vz.y * n.y + w.vz.z * n.z);
}

void pair_hit(const Instance &in, const Geom &g, int prim, vec3 o, vec3 d, float t_hit, PairHit &h)
{
  const int ia = g.indices[2 * prim], ib = g.indices[2 * prim + 1];
  const vec3 v0 = g.vertices[ia], v1 = g.vertices[ib];
  const float r0 = g.radii[ia], r1 = g.radii[ib];
  float t = t_hit;
  vec3 N(1.f, 0.f, 0.f);
  if (g.type == 3) {
    // Capsules.dev.cu:212-262: the point of the moved ray, the angle-based
    // attribute weight, the world point moved off the surface
    vec3 objP = o + t_hit * d;
    (void)capsule_test_geom(o, d, v0, r0, v1, r1, t_hit * 1.0001f, t, &N, &objP);
    const float l01 = length(v1 - v0), lp0 = length(objP - v0);
    float lerp = 0.f;
    if (!(l01 < 1e-8f || lp0 < 1e-8f)) lerp = fmx(0.f, fmn(1.f, dot(objP - v0, v1 - v0) / (lp0 * l01)));
    h.osP = objP;
    h.osN = normalize(N);
    if (in.o2w.identity) {
      h.P = objP;
      h.Nw = normalize(N);
    } else {
      h.P = xfm_point(in.o2w, objP);
      h.Nw = normalize(xfm_normal(in.w2o, N));
    }
    float eps = 1.f;
    eps = fmx(eps, fabsf(h.P.x));
    eps = fmx(eps, fabsf(h.P.y));
    eps = fmx(eps, fabsf(h.P.z));
    eps *= 1e-5f;
    h.P = h.P + eps * h.Nw;
    h.lerp = lerp;
  } else {
    // Cones.dev.cu:117-262: P = o + t d, the normal normalized once
    float lerp = 0.f;
    (void)cone_test_geom(o, d, v0, r0, v1, r1, 0.f, t_hit * 1.0001f, t, &N, &lerp);
    const vec3 objP = o + t_hit * d;
    h.osP = objP;
    h.osN = normalize(N);
    if (in.o2w.identity) {
      h.P = objP;
      h.Nw = normalize(h.osN);
    } else {
      h.P = xfm_point(in.o2w, objP);
      h.Nw = normalize(xfm_normal(in.w2o, h.osN));
    }
    h.lerp = lerp;
  }
}

Hit trace_bvh(const Scene &s, vec3 org, vec3 dir0, float tmin, float tmax, bool any_hit, TraceCounters *ctr,
              const float *cut_plane, const uint64_t *seed, Hit *rej)
{
  if (rej) *rej = Hit{INFINITY, 0.f, 0.f, -1, -1, -1};
  vec3 dir = fix_dir(dir0);
  const CutRay cut{cut_plane, org, dir};
  Hit best;
  best.t = tmax;
  best.u = best.v = 0.f;
  if (ctr) ctr->rays++;
  vec3 winv(1.f / dir.x, 1.f / dir.y, 1.f / dir.z);
  // stack entries: cnt == 0 inner node, cnt > 0 leaf [ref, ref+cnt),
  // ENTER: instance 'ref' (TLAS leaf), LEAVE: back to the TLAS
  struct Entry { int32_t ref, cnt; };
  const int32_t LEAVE = -2, ENTER = -3;
  Entry stack[512];
  int sp = 0;
  stack[sp++] = {0, 0};
  bool in_blas = false;
  int inst = -1;
  vec3 o = org, d = dir, inv = winv;
  while (sp > 0) {
    Entry en = stack[--sp];
    if (en.cnt == LEAVE) {
      in_blas = false;
      o = org; d = dir; inv = winv;
      continue;
    }
    if (en.cnt == ENTER) {
      // enterBlas (rtcore/cuda/TraceInterface.h:107-122)
      inst = en.ref;
      in_blas = true;
      const Instance &in = s.insts[inst];
      if (in.w2o.identity) {
        o = org; d = dir; inv = winv;
      } else {
        o = xfm_point(in.w2o, org);
        d = fix_dir(xfm_vector(in.w2o, dir));
        inv = vec3(1.f / d.x, 1.f / d.y, 1.f / d.z);
      }
      stack[sp++] = {0, LEAVE};
      stack[sp++] = {0, 0};
      continue;
    }
    const Bvh &bvh = in_blas ? s.groups[s.insts[inst].group].bvh : s.tlas;
    if (en.cnt == 0) {
      const Node &nd = bvh.nodes[en.ref];
      if (ctr) ctr->nodes++;
      float tn[2] = {0.f, 0.f};
      bool h[2];
      h[0] = nd.cnt[0] >= 0 && box_test(o, inv, nd.lo0, nd.hi0, tmin, best.t, tn[0]);
      h[1] = nd.cnt[1] >= 0 && box_test(o, inv, nd.lo1, nd.hi1, tmin, best.t, tn[1]);
      if (h[0] && h[1]) {
        int nr = tn[1] < tn[0] ? 1 : 0;
        stack[sp++] = {nd.ref[1 - nr], nd.cnt[1 - nr]};
        stack[sp++] = {nd.ref[nr], nd.cnt[nr]};
      } else if (h[0]) {
        stack[sp++] = {nd.ref[0], nd.cnt[0]};
      } else if (h[1]) {
        stack[sp++] = {nd.ref[1], nd.cnt[1]};
      }
      continue;
    }
    if (!in_blas) {
      for (int k = en.ref + en.cnt - 1; k >= en.ref; k--) stack[sp++] = {s.tlas.order[k], ENTER};
      continue;
    }
    const Group &g = s.groups[s.insts[inst].group];
    for (int k = en.ref; k < en.ref + en.cnt; k++) {
      bool hit = prim_test(s, g, k, o, d, tmin, best, inst, ctr, cut, seed, rej);
      if (hit && any_hit) return best;
    }
  }
  return best;
}

Hit trace_brute(const Scene &s, vec3 org, vec3 dir0, float tmin, float tmax, const float *cut_plane,
                const uint64_t *seed, Hit *rej)
{
  if (rej) *rej = Hit{INFINITY, 0.f, 0.f, -1, -1, -1};
  vec3 dir = fix_dir(dir0);
  const CutRay cut{cut_plane, org, dir};
  Hit best;
  best.t = tmax;
  best.u = best.v = 0.f;
  for (int inst = 0; inst < (int)s.insts.size(); inst++) {
    const Instance &in = s.insts[inst];
    vec3 o = org, d = dir;
    if (!in.w2o.identity) {
      o = xfm_point(in.w2o, org);
      d = fix_dir(xfm_vector(in.w2o, dir));
    }
    const Group &g = s.groups[in.group];
    for (int k = 0; k < (int)g.bvh.order.size(); k++)
      prim_test(s, g, k, o, d, tmin, best, inst, nullptr, cut, seed, rej);
  }
  return best;
}

// ======================================================================
}  // namespace orc
