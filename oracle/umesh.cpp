// umesh.cpp -- CPU oracle of barney's unstructured field (tets, pyramids,
// prisms, hexahedra) and of the isosurface intersection program.
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

// rasterBox (volume/MCGrid.cuh:66-122): every cell the 3D box touches gets
// min/max of the box's value range.  strict: boxes flat in any dimension are
// skipped (>=, the two-argument form, for tet pieces); otherwise only
// inverted ones (>, the three-argument form, for the other cells)
void raster_box(Volume &v, const Box4 &b, vec3 rsp, bool strict = true)
{
  if (strict ? b.lo.x >= b.hi.x : b.lo.x > b.hi.x) return;
  if (strict ? b.lo.y >= b.hi.y : b.lo.y > b.hi.y) return;
  if (strict ? b.lo.z >= b.hi.z : b.lo.z > b.hi.z) return;
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

int nverts_of(int type) { return type == CELL_HEX ? 8 : type == CELL_PRISM ? 6 : type == CELL_PYR ? 5 : 4; }
int cell_type(const Volume &v, int c) { return v.types.empty() ? CELL_TET : (int)v.types[c]; }
// vertex k of cell c (k < nverts_of(type))
int cell_vertex(const Volume &v, int c, int k)
{
  return k < 4 ? v.tets[4 * (size_t)c + k] : v.ext[4 * (size_t)c + (k - 4)];
}

// ---- isoparametric cells: shape functions N_i(r, s, t) and their partial
// derivatives in VTK vertex order (ElementIntersection.h:37-75, 171-211,
// 286-347), with the same arithmetic
void shape_hex(float r, float s, float t, float *w, float *dr, float *ds, float *dt)
{
  const float rm = 1.f - r, sm = 1.f - s, tm = 1.f - t;
  const float W[8] = {rm * sm * tm, r * sm * tm, r * s * tm, rm * s * tm, rm * sm * t, r * sm * t, r * s * t, rm * s * t};
  const float R[8] = {-sm * tm, sm * tm, s * tm, -s * tm, -sm * t, sm * t, s * t, -s * t};
  const float S[8] = {-rm * tm, -r * tm, r * tm, rm * tm, -rm * t, -r * t, r * t, rm * t};
  const float T[8] = {-rm * sm, -r * sm, -r * s, -rm * s, rm * sm, r * sm, r * s, rm * s};
  for (int i = 0; i < 8; i++) {
    w[i] = W[i];
    dr[i] = R[i];
    ds[i] = S[i];
    dt[i] = T[i];
  }
}
void shape_prism(float r, float s, float t, float *w, float *dr, float *ds, float *dt)
{
  const float W[6] = {(1.f - r - s) * (1.f - t), r * (1.f - t), s * (1.f - t), (1.f - r - s) * t, r * t, s * t};
  const float R[6] = {-1.f + t, 1.f - t, 0.f, -t, t, 0.f};
  const float S[6] = {-1.f + t, 0.f, 1.f - t, -t, 0.f, t};
  const float T[6] = {-1.f + r + s, -r, -s, 1.f - r - s, r, s};
  for (int i = 0; i < 6; i++) {
    w[i] = W[i];
    dr[i] = R[i];
    ds[i] = S[i];
    dt[i] = T[i];
  }
}
void shape_pyr(float r, float s, float t, float *w, float *dr, float *ds, float *dt)
{
  const float rm = 1.f - r, sm = 1.f - s, tm = 1.f - t;
  const float W[5] = {rm * sm * tm, r * sm * tm, r * s * tm, rm * s * tm, t};
  const float R[5] = {-(s - 1.f) * (t - 1.f), (s - 1.f) * (t - 1.f), s - s * t, s * (t - 1.f), 0.f};
  const float S[5] = {-(r - 1.f) * (t - 1.f), r * (t - 1.f), r - r * t, (r - 1.f) * (t - 1.f), 0.f};
  const float T[5] = {-(r - 1.f) * (s - 1.f), r * (s - 1.f), -r * s, (r - 1.f) * s, 1.f};
  for (int i = 0; i < 5; i++) {
    w[i] = W[i];
    dr[i] = R[i];
    ds[i] = S[i];
    dt[i] = T[i];
  }
}
// det of the matrix with columns a, b, c (LinearSpace3f det: a . (b x c))
float det3(vec3 a, vec3 b, vec3 c) { return dot(a, cross(b, c)); }

} // namespace

bool cell_invert(int type, const float *V, vec3 P, float &out)
{
  const int n = nverts_of(type);
  float tol = 1e-6f;                 // pyramids and prisms
  if (type == CELL_HEX) {
    // hexahedra: 1e-10 x the diagonal of the vertices' box
    vec3 lo(V[0], V[1], V[2]), hi = lo;
    for (int i = 1; i < 8; i++) {
      const vec3 p(V[4 * i], V[4 * i + 1], V[4 * i + 2]);
      lo = vmin(lo, p);
      hi = vmax(hi, p);
    }
    tol = 1e-10f * length(hi - lo);
  }
  float pc[3] = {.5f, .5f, .5f};
  float w[8], dr[8], ds[8], dt[8];
  bool converged = false;
  for (int it = 0; it < 10; it++) {
    if (type == CELL_HEX) shape_hex(pc[0], pc[1], pc[2], w, dr, ds, dt);
    else if (type == CELL_PRISM) shape_prism(pc[0], pc[1], pc[2], w, dr, ds, dt);
    else shape_pyr(pc[0], pc[1], pc[2], w, dr, ds, dt);
    vec3 f(0.f), cr(0.f), cs(0.f), ct(0.f);
    for (int i = 0; i < n; i++) {
      const vec3 X(V[4 * i], V[4 * i + 1], V[4 * i + 2]);
      f = f + X * w[i];
      cr = cr + X * dr[i];
      cs = cs + X * ds[i];
      ct = ct + X * dt[i];
    }
    f = f - P;
    const float d = det3(cr, cs, ct);
    if (fabsf(d) < tol) return false;
    const float d0 = det3(f, cs, ct) / d;
    const float d1 = det3(cr, f, ct) / d;
    const float d2 = det3(cr, cs, f) / d;
    pc[0] = pc[0] - d0;
    pc[1] = pc[1] - d1;
    pc[2] = pc[2] - d2;
    if (fabsf(d0) < 1e-4f && fabsf(d1) < 1e-4f && fabsf(d2) < 1e-4f) {
      converged = true;
      break;
    }
    if (fabsf(pc[0]) > 1e6f || fabsf(pc[1]) > 1e6f || fabsf(pc[2]) > 1e6f) return false;
  }
  if (!converged) return false;
  const float lo = 0.f - 1e-6f, hi = 1.f + 1e-6f;
  for (int k = 0; k < 3; k++)
    if (!(pc[k] >= lo && pc[k] <= hi)) return false;
  if (type == CELL_PRISM && !(pc[0] + pc[1] <= hi)) return false;
  float val = 0.f;
  for (int i = 0; i < n; i++) val += w[i] * V[4 * i + 3];
  out = val;
  return true;
}

bool cell_scalar(const Volume &v, int slot, vec3 P, float &out)
{
  const int type = cell_type(v, slot);
  if (type == CELL_TET) return tet_scalar(v, slot, P, out);
  float V[32];
  for (int k = 0; k < nverts_of(type); k++) {
    const vec4 p = v4(v, cell_vertex(v, slot, k));
    V[4 * k] = p.x;
    V[4 * k + 1] = p.y;
    V[4 * k + 2] = p.z;
    V[4 * k + 3] = p.w;
  }
  return cell_invert(type, V, P, out);
}

void build_umesh(Volume &v)
{
  const int ntet = (int)(v.tets.size() / 4);
  // computeWorldBounds (UMeshField.cu:304-318): union of the cells' boxes
  Box wb;
  std::vector<Box> boxes((size_t)ntet);
  for (int t = 0; t < ntet; t++) {
    Box b;
    for (int q = 0; q < nverts_of(cell_type(v, t)); q++) b.grow(pos(v4(v, cell_vertex(v, t, q))));
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
  // umeshRasterCells (UMeshField.cu:109-161): tets split 5 times, every
  // other cell by its box (position and value range)
  for (int t = 0; t < ntet; t++) {
    const int type = cell_type(v, t);
    if (type == CELL_TET) {
      const int32_t *ix = &v.tets[4 * (size_t)t];
      raster_tet(v, 5, v4(v, ix[0]), v4(v, ix[1]), v4(v, ix[2]), v4(v, ix[3]), rsp);
    } else {
      Box4 bb;
      for (int q = 0; q < nverts_of(type); q++) bb.extend(v4(v, cell_vertex(v, t, q)));
      raster_box(v, bb, rsp, false);
    }
  }
  build_majorants(v);
  // element BVH; the cells are stored in its slot order
  build_bvh(boxes, 4, v.ebvh);
  std::vector<int32_t> sorted(v.tets.size()), sorted_ext(v.ext.size());
  std::vector<uint8_t> sorted_types(v.types.size());
  for (int s = 0; s < ntet; s++) {
    const int c = v.ebvh.order[s];
    for (int q = 0; q < 4; q++) sorted[4 * (size_t)s + q] = v.tets[4 * (size_t)c + q];
    if (!v.types.empty()) {
      for (int q = 0; q < 4; q++) sorted_ext[4 * (size_t)s + q] = v.ext[4 * (size_t)c + q];
      sorted_types[s] = v.types[c];
    }
  }
  v.tets.swap(sorted);
  v.ext.swap(sorted_ext);
  v.types.swap(sorted_types);
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
        if (cell_scalar(v, e.ref + k, P, f)) return f;
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
