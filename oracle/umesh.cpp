// umesh.cpp -- CPU oracle of barney's unstructured (tetrahedral) field and
// of the isosurface intersection program.
//
// TEST INFRASTRUCTURE ONLY (see oracle.h).
#include "oracle.h"
#include "dda.h"

#include <algorithm>

namespace orc {

namespace {

struct vec4 {
  float x, y, z, w;
};
inline vec4 v4(const Volume &v, int i)
{
  const float *p = &v.verts[4 * (size_t)i];
  return vec4{p[0], p[1], p[2], p[3]};
}
inline vec3 pos(vec4 a) { return vec3(a.x, a.y, a.z); }
// 0.5f * (a + b) on vec4f
inline vec4 mid(vec4 a, vec4 b)
{
  return vec4{0.5f * (a.x + b.x), 0.5f * (a.y + b.y), 0.5f * (a.z + b.z), 0.5f * (a.w + b.w)};
}
inline float length3(vec4 a, vec4 b) { return length(pos(b) - pos(a)); }   // length3(b - a)

struct Box4 {
  vec4 lo{INFINITY, INFINITY, INFINITY, INFINITY}, hi{-INFINITY, -INFINITY, -INFINITY, -INFINITY};
  void extend(vec4 p)
  {
    lo = vec4{fmn(lo.x, p.x), fmn(lo.y, p.y), fmn(lo.z, p.z), fmn(lo.w, p.w)};
    hi = vec4{fmx(hi.x, p.x), fmx(hi.y, p.y), fmx(hi.z, p.z), fmx(hi.w, p.w)};
  }
};

// rasterBox(grid, primBounds4) (volume/MCGrid.cuh:95-122): every cell the
// 3D box touches gets min/max of the box's value range; boxes that are flat
// in any dimension are skipped (>= tests)
void raster_box(Volume &v, const Box4 &b, vec3 rsp)
{
  if (b.lo.x >= b.hi.x) return;
  if (b.lo.y >= b.hi.y) return;
  if (b.lo.z >= b.hi.z) return;
  int lo[3], hi[3];
  const float bl[3] = {b.lo.x, b.lo.y, b.lo.z}, bh[3] = {b.hi.x, b.hi.y, b.hi.z};
  for (int k = 0; k < 3; k++) {
    lo[k] = (int)((bl[k] - v.mc_origin[k]) * rsp[k]);
    hi[k] = (int)((bh[k] - v.mc_origin[k]) * rsp[k]);
    lo[k] = std::min(std::max(lo[k], 0), v.mc[k] - 1);
    hi[k] = std::min(std::max(hi[k], 0), v.mc[k] - 1);
  }
  for (int iz = lo[2]; iz <= hi[2]; iz++)
    for (int iy = lo[1]; iy <= hi[1]; iy++)
      for (int ix = lo[0]; ix <= hi[0]; ix++) {
        const size_t c = (size_t)ix + (size_t)iy * v.mc[0] + (size_t)iz * v.mc[0] * v.mc[1];
        v.mc_lo[c] = fmn(v.mc_lo[c], b.lo.w);
        v.mc_hi[c] = fmx(v.mc_hi[c], b.hi.w);
      }
}

// rasterTet<D> (UMeshField.cu:28-107): split on the longest edge D times,
// then raster the boxes of the pieces
void raster_tet(Volume &v, int D, vec4 a, vec4 b, vec4 c, vec4 d, vec3 rsp)
{
  if (D == 0) {
    Box4 bb;
    bb.extend(a);
    bb.extend(b);
    bb.extend(c);
    bb.extend(d);
    raster_box(v, bb, rsp);
    return;
  }
  const float lab = length3(a, b), lac = length3(a, c), lad = length3(a, d);
  const float lbc = length3(b, c), lbd = length3(b, d), lcd = length3(c, d);
  const float maxLen = fmx(fmx(fmx(fmx(fmx(lab, lac), lad), lbc), lbd), lcd);
  vec4 oa, ob, oc, od0, od1;
  if (lab >= maxLen) { oa = mid(a, b); ob = c; oc = d; od0 = a; od1 = b; }
  else if (lac >= maxLen) { oa = mid(a, c); ob = b; oc = d; od0 = a; od1 = c; }
  else if (lad >= maxLen) { oa = mid(a, d); ob = b; oc = c; od0 = a; od1 = d; }
  else if (lbc >= maxLen) { oa = mid(b, c); ob = a; oc = d; od0 = b; od1 = c; }
  else if (lbd >= maxLen) { oa = mid(b, d); ob = a; oc = c; od0 = b; od1 = d; }
  else { oa = mid(c, d); ob = a; oc = b; od0 = c; od1 = d; }
  raster_tet(v, D - 1, oa, ob, oc, od0, rsp);
  raster_tet(v, D - 1, oa, ob, oc, od1, rsp);
}

// evalToImplicitPlane (UMeshField.h:365-371): unnormalized distance of P
// to the plane (a, b, c), positive on the side cross(b - a, c - a) points to
inline float plane(vec3 P, vec4 a, vec4 b, vec4 c)
{
  vec3 N = cross(pos(b) - pos(a), pos(c) - pos(a));
  return dot(P - pos(a), N);
}

} // namespace

void build_umesh(Volume &v)
{
  const int ntet = (int)(v.tets.size() / 4);
  // computeWorldBounds (UMeshField.cu:304-318): union of the cells' boxes
  Box wb;
  std::vector<Box> boxes((size_t)ntet);
  for (int t = 0; t < ntet; t++) {
    Box b;
    for (int q = 0; q < 4; q++) b.grow(pos(v4(v, v.tets[4 * (size_t)t + q])));
    boxes[t] = b;
    wb.grow(b);
  }
  v.bounds = wb;
  // buildInitialMacroCells (UMeshField.cu:168-197)
  const vec3 size = wb.hi - wb.lo;
  if (v.mc_req[0] > 0) {
    for (int k = 0; k < 3; k++) v.mc[k] = v.mc_req[k];
  } else {
    const float maxWidth = reduce_max(size);
    const int G = 200 + (int)sqrtf((float)ntet / 1000.f);
    const float s = (float)(G - 1) / maxWidth;
    for (int k = 0; k < 3; k++) v.mc[k] = 1 + (int)(size[k] * s);
  }
  v.mc_origin = wb.lo;
  v.mc_spacing = size * vec3(1.f / (float)v.mc[0], 1.f / (float)v.mc[1], 1.f / (float)v.mc[2]);
  const size_t nmc = (size_t)v.mc[0] * v.mc[1] * v.mc[2];
  v.mc_lo.assign(nmc, INFINITY);
  v.mc_hi.assign(nmc, -INFINITY);
  const vec3 rsp(1.f / v.mc_spacing.x, 1.f / v.mc_spacing.y, 1.f / v.mc_spacing.z);
  // umeshRasterCells (UMeshField.cu:109-161), tets only
  for (int t = 0; t < ntet; t++) {
    const int32_t *ix = &v.tets[4 * (size_t)t];
    raster_tet(v, 5, v4(v, ix[0]), v4(v, ix[1]), v4(v, ix[2]), v4(v, ix[3]), rsp);
  }
  build_majorants(v);
  // element BVH; the tets are stored in its slot order
  build_bvh(boxes, 4, v.ebvh);
  std::vector<int32_t> sorted(v.tets.size());
  for (int s = 0; s < ntet; s++)
    for (int q = 0; q < 4; q++) sorted[4 * (size_t)s + q] = v.tets[4 * (size_t)v.ebvh.order[s] + q];
  v.tets.swap(sorted);
}

bool tet_scalar(const Volume &v, int slot, vec3 P, float &out)
{
  const int32_t *ix = &v.tets[4 * (size_t)slot];
  const vec4 v0 = v4(v, ix[0]), v1 = v4(v, ix[1]), v2 = v4(v, ix[2]), v3 = v4(v, ix[3]);
  const float t3 = plane(P, v0, v1, v2);
  if (t3 < 0.f) return false;
  const float t2 = plane(P, v0, v3, v1);
  if (t2 < 0.f) return false;
  const float t1 = plane(P, v0, v2, v3);
  if (t1 < 0.f) return false;
  const float t0 = plane(P, v1, v3, v2);
  if (t0 < 0.f) return false;
  const float scale = 1.f / (((t0 + t1) + t2) + t3);
  out = scale * (((t0 * v0.w + t1 * v1.w) + t2 * v2.w) + t3 * v3.w);
  return true;
}

float umesh_sample(const Volume &v, vec3 P, uint64_t *visits)
{
  // fixed-box query with a point box (cuBQL fixedBoxQuery::forEachPrim):
  // closed box tests, depth first, child 0 before child 1
  struct Entry { int32_t ref, cnt; };
  Entry stack[128];
  int sp = 0;
  stack[sp++] = {0, 0};
  auto inside = [&](const float *lo, const float *hi) {
    return lo[0] <= P.x && P.x <= hi[0] && lo[1] <= P.y && P.y <= hi[1] && lo[2] <= P.z && P.z <= hi[2];
  };
  while (sp > 0) {
    const Entry e = stack[--sp];
    if (e.cnt == 0) {
      const Node &nd = v.ebvh.nodes[e.ref];
      if (visits) (*visits)++;
      const bool h0 = nd.cnt[0] >= 0 && inside(nd.lo0, nd.hi0);
      const bool h1 = nd.cnt[1] >= 0 && inside(nd.lo1, nd.hi1);
      if (h1) stack[sp++] = {nd.ref[1], nd.cnt[1]};
      if (h0) stack[sp++] = {nd.ref[0], nd.cnt[0]};
    } else {
      for (int k = 0; k < e.cnt; k++) {
        float f;
        if (visits) (*visits)++;
        if (tet_scalar(v, e.ref + k, P, f)) return f;
      }
    }
  }
  return NAN;
}

bool iso_march(const Volume &field, const IsoSurface &iso, uint32_t rtc_inst, vec3 org, vec3 dir, float tmin,
               float tmax, RNGSeed &parent, float &t_out, vec3 &normal)
{
  // MCIsoSurfaceAccel::isProg (volume/MCAccelerator.h:325-548)
  float r0 = tmin, r1 = tmax;
  {
    // boxTest (render/Ray.h:169-217)
    vec3 rd(1.f / dir.x, 1.f / dir.y, 1.f / dir.z);
    vec3 t_lo = (field.bounds.lo - org) * rd;
    vec3 t_hi = (field.bounds.hi - org) * rd;
    r0 = fmx(r0, reduce_max(vmin(t_lo, t_hi)));
    r1 = fmn(r1, reduce_min(vmax(t_lo, t_hi)));
    if (!(r0 <= r1)) return false;
  }
  vec3 rsp(1.f / field.mc_spacing.x, 1.f / field.mc_spacing.y, 1.f / field.mc_spacing.z);
  vec3 dorg = (org - field.mc_origin) * rsp;
  vec3 ddir = dir * rsp;
  // Random rng(ray.rngSeed, hash(inst, geom, prim)): only drawn from for
  // opacity < 1 (not provided), but constructing it advances the ray's seed
  Random2 rng = Random2::from64(parent, fnv_hash3(rtc_inst, 0u, 0u));
  (void)rng;
  const float ray_tmax = tmax;
  float tHit = ray_tmax;
  const float isoValue = iso.iso;
  dda3(dorg, ddir, r1, field.mc, [&](const int cell[3], float t0, float t1) -> bool {
    if (!(t0 < fmn(t1, ray_tmax))) return true;
    const size_t c = mc_index(field, cell);
    if (!(isoValue >= field.mc_lo[c] && isoValue <= field.mc_hi[c])) return true;
    float tt1 = t0;
    float ff1 = volume_sample(field, org + tt1 * dir);
    for (int i = 1; i <= 10; i++) {
      const float tt0 = tt1;
      const float ff0 = ff1;
      tt1 = (1.f - (float)i / 10.f) * t0 + ((float)i / 10.f) * t1;   // lerp_l(i/10, t0, t1)
      ff1 = volume_sample(field, org + tt1 * dir);
      if (std::isnan(ff0) || std::isnan(ff1)) continue;
      const float lo = fmn(ff0, ff1), hi = fmx(ff0, ff1);
      if (isoValue >= lo && isoValue <= hi) {
        float t = (isoValue - ff0) / (ff1 - ff0);
        t = (1.f - t) * tt0 + t * tt1;
        // min(tHit, t): a NaN candidate (ff0 == ff1 == iso) is ignored
        if (t < tHit) tHit = t;
        if (tHit < ray_tmax) return false;
      }
    }
    return true;
  });
  if (!(tHit < ray_tmax)) return false;
  // central-difference normal (:464-489)
  const vec3 osP = org + tHit * dir;
  const vec3 bs = field.bounds.hi - field.bounds.lo;
  const float delta = length(bs) * .1f / (float)(field.mc[0] + field.mc[1] + field.mc[2]);
  const float fP = volume_sample(field, osP);
  float fx0 = volume_sample(field, osP + vec3(-delta, 0.f, 0.f));
  float fx1 = volume_sample(field, osP + vec3(+delta, 0.f, 0.f));
  float fy0 = volume_sample(field, osP + vec3(0.f, -delta, 0.f));
  float fy1 = volume_sample(field, osP + vec3(0.f, +delta, 0.f));
  float fz0 = volume_sample(field, osP + vec3(0.f, 0.f, -delta));
  float fz1 = volume_sample(field, osP + vec3(0.f, 0.f, +delta));
  float dx = 2.f, dy = 2.f, dz = 2.f;
  if (std::isnan(fx0)) { dx -= 1.f; fx0 = fP; }
  if (std::isnan(fx1)) { dx -= 1.f; fx1 = fP; }
  if (std::isnan(fy0)) { dy -= 1.f; fy0 = fP; }
  if (std::isnan(fy1)) { dy -= 1.f; fy1 = fP; }
  if (std::isnan(fz0)) { dz -= 1.f; fz0 = fP; }
  if (std::isnan(fz1)) { dz -= 1.f; fz1 = fP; }
  vec3 osN(dx == 0.f ? 0.f : (fx1 - fx0) / dx, dy == 0.f ? 0.f : (fy1 - fy0) / dy,
           dz == 0.f ? 0.f : (fz1 - fz0) / dz);
  if (osN == vec3(0.f)) osN = -normalize(dir);
  t_out = tHit;
  normal = normalize(osN);
  return true;
}

} // namespace orc
