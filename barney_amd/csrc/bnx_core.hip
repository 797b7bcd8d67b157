// bnx_core.hip -- the C-ABI of include/bnx_rt.h: devices, buffers,
// geometries, groups (BLAS/TLAS builds), worlds, frames and the frame loop.
#include "bnx_internal.h"
#include "bnx_render.h"
#include "bnx_volume.h"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <new>

namespace bnx {

static thread_local std::string g_last_error = "no error";
void set_error(const std::string &msg) { g_last_error = msg; }

namespace {

inline int nblk(long n, int b) { return (int)((n + b - 1) / b); }

// ------------------------------------------------------------ build prep
struct GeomBuild {
  int type;
  int base;             // first flattened prim of this geometry
  int num_prims;
  int pad;
  const float *vertices;
  const int *indices;
  const float4 *spheres;
  const float4 *cyls;   // two records per prim: cylinders (v0, radius), (v1, 0); capsules, cones (v0, r0), (v1, r1)
};

__global__ void k_prim_boxes(const GeomBuild *gb, int num_geoms, int n, float4 *lo, float4 *hi, int *pgeom)
{
  int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  int g = 0;
  while (g + 1 < num_geoms && gb[g + 1].base <= p) g++;
  const GeomBuild G = gb[g];
  int prim = p - G.base;
  pgeom[p] = g;
  if (G.type == 0) {
    // createTriangleBounds (rtcore/cuda/Group.cu:91-112)
    const int *ix = G.indices + 3 * prim;
    float l[3] = {INFINITY, INFINITY, INFINITY}, h[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (int k = 0; k < 3; k++) {
      const float *v = G.vertices + 3 * ix[k];
      for (int a = 0; a < 3; a++) {
        l[a] = fmn(l[a], v[a]);
        h[a] = fmx(h[a], v[a]);
      }
    }
    lo[p] = make_float4(l[0], l[1], l[2], 0.f);
    hi[p] = make_float4(h[0], h[1], h[2], 0.f);
  } else if (G.type == 1) {
    // SpheresPrograms::bounds (Spheres.dev.cu:29-40)
    float4 s = G.spheres[prim];
    lo[p] = make_float4(s.x - s.w, s.y - s.w, s.z - s.w, 0.f);
    hi[p] = make_float4(s.x + s.w, s.y + s.w, s.z + s.w, 0.f);
  } else if (G.type == 2) {
    // CylindersPrograms::bounds (native/geometry/Cylinders.dev.cu:69-85):
    // the boxes of both end spheres
    const float4 a = G.cyls[2 * prim], b = G.cyls[2 * prim + 1];
    const float r = a.w;
    lo[p] = make_float4(fmn(a.x - r, b.x - r), fmn(a.y - r, b.y - r), fmn(a.z - r, b.z - r), 0.f);
    hi[p] = make_float4(fmx(a.x + r, b.x + r), fmx(a.y + r, b.y + r), fmx(a.z + r, b.z + r), 0.f);
  } else {
    // capsules and cones: the boxes of both end spheres (getBounds,
    // Capsules.dev.cu:100-110).  ConesPrograms::bounds (Cones.dev.cu:21-42)
    // takes the second end's lower corner from the first radius (pb - ra),
    // a box that can miss part of the cone; a hit outside its box is found
    // or not depending on traversal order, so cones take the box that holds
    // them
    const float4 a = G.cyls[2 * prim], b = G.cyls[2 * prim + 1];
    lo[p] = make_float4(fmn(a.x - a.w, b.x - b.w), fmn(a.y - a.w, b.y - b.w), fmn(a.z - a.w, b.z - b.w), 0.f);
    hi[p] = make_float4(fmx(a.x + a.w, b.x + b.w), fmx(a.y + a.w, b.y + b.w), fmx(a.z + a.w, b.z + b.w), 0.f);
  }
}

// BVH-ordered primitive records (reorderPrims, rtcore/cuda/Group.cu:114-123)
__global__ void k_prim_records(const GeomBuild *gb, const int *order, const int *pgeom, int n, float4 *prims,
                               int2 *refs)
{
  int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n) return;
  int p = order[s];
  int g = pgeom[p];
  const GeomBuild G = gb[g];
  int prim = p - G.base;
  if (G.type == 0) {
    const int *ix = G.indices + 3 * prim;
    const float *a = G.vertices + 3 * ix[0], *b = G.vertices + 3 * ix[1], *c = G.vertices + 3 * ix[2];
    prims[3 * s + 0] = make_float4(a[0], a[1], a[2], __int_as_float(prim));
    prims[3 * s + 1] = make_float4(b[0] - a[0], b[1] - a[1], b[2] - a[2], __int_as_float(g));
    prims[3 * s + 2] = make_float4(c[0] - a[0], c[1] - a[1], c[2] - a[2], 0.f);
  } else if (G.type == 1) {
    prims[s] = G.spheres[prim];
    refs[s] = make_int2(g, prim);
  } else {
    prims[2 * s] = G.cyls[2 * prim];
    prims[2 * s + 1] = G.cyls[2 * prim + 1];
    refs[s] = make_int2(g, prim);
  }
}

__global__ void k_boxes_from_floats(const float *b, int n, float4 *lo, float4 *hi)
{
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  lo[i] = make_float4(b[6 * i + 0], b[6 * i + 1], b[6 * i + 2], 0.f);
  hi[i] = make_float4(b[6 * i + 3], b[6 * i + 4], b[6 * i + 5], 0.f);
}

template <typename T> int upload(Device *dev, DevMem &m, const T *host, size_t count)
{
  BNX_HIP(m.ensure(sizeof(T) * (count ? count : 1)));
  if (count) BNX_HIP(hipMemcpyAsync(m.ptr, host, sizeof(T) * count, hipMemcpyHostToDevice, dev->stream));
  return BNX_OK;
}

int build_blas(Group *g)
{
  Device *dev = g->dev;
  std::vector<GeomBuild> gb;
  int n = 0;
  for (Geom *geom : g->geoms) {
    GeomBuild b{};
    b.type = geom->type;
    b.base = n;
    b.num_prims = geom->num_prims;
    b.vertices = geom->vertices.as<float>();
    b.indices = geom->indices.as<int>();
    b.spheres = geom->spheres.as<float4>();
    b.cyls = geom->cyls.as<float4>();
    gb.push_back(b);
    n += geom->num_prims;
  }
  DevMem d_gb, lo, hi, pgeom;
  int rc;
  if ((rc = upload(dev, d_gb, gb.data(), gb.size())) != BNX_OK) return rc;
  BNX_HIP(lo.ensure(sizeof(float4) * (size_t)(n ? n : 1)));
  BNX_HIP(hi.ensure(sizeof(float4) * (size_t)(n ? n : 1)));
  BNX_HIP(pgeom.ensure(sizeof(int) * (size_t)(n ? n : 1)));
  if (n) k_prim_boxes<<<nblk(n, 256), 256, 0, dev->stream>>>(d_gb.as<GeomBuild>(), (int)gb.size(), n, lo.as<float4>(),
                                                             hi.as<float4>(), pgeom.as<int>());
  if ((rc = build_bvh(dev, lo.as<float4>(), hi.as<float4>(), n, 4, g->bvh)) != BNX_OK) return rc;
  if ((rc = collapse_bvh4(dev, g->bvh)) != BNX_OK) return rc;
  // record float4s per primitive: triangles 3, spheres 1, cylinders,
  // capsules and cones 2 (a user group holds one user type)
  const int utype = g->geoms.empty() ? 1 : g->geoms[0]->type;
  size_t per = g->kind == GROUP_TRIANGLES ? 3 : (utype >= 2 ? 2 : 1);
  BNX_HIP(g->prims.ensure(sizeof(float4) * per * (size_t)(n ? n : 1)));
  BNX_HIP(g->refs.ensure(sizeof(int2) * (size_t)(n ? n : 1)));
  if (n)
    k_prim_records<<<nblk(n, 256), 256, 0, dev->stream>>>(d_gb.as<GeomBuild>(), g->bvh.order.as<int>(),
                                                          pgeom.as<int>(), n, g->prims.as<float4>(),
                                                          g->refs.as<int2>());
  BNX_HIP(hipStreamSynchronize(dev->stream));
  BNX_HIP(hipGetLastError());
  g->built = true;
  return BNX_OK;
}

// inverse of an affine (columns vx, vy, vz, p) in double precision
void affine_inverse(const float *a, float *out)
{
  double m[3][3], p[3];
  for (int c = 0; c < 3; c++)
    for (int r = 0; r < 3; r++) m[r][c] = a[3 * c + r];
  for (int r = 0; r < 3; r++) p[r] = a[9 + r];
  double det = m[0][0] * (m[1][1] * m[2][2] - m[1][2] * m[2][1]) - m[0][1] * (m[1][0] * m[2][2] - m[1][2] * m[2][0]) +
               m[0][2] * (m[1][0] * m[2][1] - m[1][1] * m[2][0]);
  double inv[3][3];
  inv[0][0] = (m[1][1] * m[2][2] - m[1][2] * m[2][1]) / det;
  inv[0][1] = (m[0][2] * m[2][1] - m[0][1] * m[2][2]) / det;
  inv[0][2] = (m[0][1] * m[1][2] - m[0][2] * m[1][1]) / det;
  inv[1][0] = (m[1][2] * m[2][0] - m[1][0] * m[2][2]) / det;
  inv[1][1] = (m[0][0] * m[2][2] - m[0][2] * m[2][0]) / det;
  inv[1][2] = (m[0][2] * m[1][0] - m[0][0] * m[1][2]) / det;
  inv[2][0] = (m[1][0] * m[2][1] - m[1][1] * m[2][0]) / det;
  inv[2][1] = (m[0][1] * m[2][0] - m[0][0] * m[2][1]) / det;
  inv[2][2] = (m[0][0] * m[1][1] - m[0][1] * m[1][0]) / det;
  for (int c = 0; c < 3; c++)
    for (int r = 0; r < 3; r++) out[3 * c + r] = (float)inv[r][c];
  for (int r = 0; r < 3; r++) {
    double s = -(inv[r][0] * p[0] + inv[r][1] * p[1] + inv[r][2] * p[2]);
    out[9 + r] = (float)s;
  }
}

void to_rows(const float *a, float4 *rows)
{
  // columns (vx, vy, vz, p) -> rows (x: vx.x vy.x vz.x p.x, ...)
  rows[0] = make_float4(a[0], a[3], a[6], a[9]);
  rows[1] = make_float4(a[1], a[4], a[7], a[10]);
  rows[2] = make_float4(a[2], a[5], a[8], a[11]);
}

bool is_identity(const float *a)
{
  static const float I[12] = {1, 0, 0, 0, 1, 0, 0, 0, 1, 0, 0, 0};
  for (int k = 0; k < 12; k++)
    if (a[k] != I[k]) return false;
  return true;
}

// whether a triangle geometry's alpha (Lambertian color.w * opacity;
// NVisii (1 - transmission) * baseColor.w * opacity) can be below 1 under
// instance group g, so that its candidates take the opacity test
bool can_be_translucent(const Geom &gm, const Group *g)
{
  const bnx_material &m = gm.mat;
  const int k = m.color_from_attribute;
  bool w_below_1 = false;
  if (k == BNX_ATTRIB_NONE) {
    w_below_1 = !(m.color[0] != m.color[0]) && !(m.color[3] >= 1.f);
  } else if (k == BNX_ATTRIB_COLOR || (k >= BNX_ATTRIB_0 && k < BNX_ATTRIB_0 + 4)) {
    const int slot = k == BNX_ATTRIB_COLOR ? 4 : k - BNX_ATTRIB_0;
    w_below_1 = gm.attr[slot].scope != BNX_SCOPE_NONE ? gm.attr[slot].w_below_1 : g->inst_attr_w_below_1[slot];
  }
  bool below = w_below_1 || !(m.opacity >= 1.f);
  if (m.type == BNX_MATERIAL_PBR) below = below || m.transmission != 0.f;
  return below;
}

int build_tlas(Group *g)
{
  Device *dev = g->dev;
  // the group is rebuilt in place: it counts as unbuilt (traces, renders and
  // new worlds refuse it) until every step below, the stack-depth check
  // included, has succeeded
  g->built = false;
  const int n = (int)g->groups.size();
  std::vector<Group *> blas_groups;
  std::vector<InstDev> insts(n);
  std::vector<float4> lo(n ? n : 1), hi(n ? n : 1);
  std::vector<GeomDev> geoms;
  std::vector<int> geom_base;
  g->max_blas_depth = 0;
  g->blas_prims = 0;
  for (int i = 0; i < n; i++) {
    Group *bg = g->groups[i];
    if (!bg || !bg->built || bg->kind == GROUP_INSTANCES) {
      set_error("bnx_group_build: instance refers to a group that is not a built geometry group");
      return BNX_ERR_STATE;
    }
    int bi = -1;
    for (int k = 0; k < (int)blas_groups.size(); k++)
      if (blas_groups[k] == bg) bi = k;
    if (bi < 0) {
      bi = (int)blas_groups.size();
      blas_groups.push_back(bg);
      geom_base.push_back((int)geoms.size());
      for (Geom *gm : bg->geoms) {
        GeomDev gd{};
        gd.type = gm->type;
        gd.user_id = gm->user_id;
        gd.color_from_attr = gm->mat.color_from_attribute;
        gd.num_prims = gm->num_prims;
        gd.color = make_float4(gm->mat.color[0], gm->mat.color[1], gm->mat.color[2], gm->mat.color[3]);
        gd.material = gm->mat.type;
        gd.metallic = gm->mat.metallic;
        gd.roughness = gm->mat.roughness;
        gd.ior = gm->mat.ior;
        gd.opacity = gm->mat.opacity;
        gd.transmission = gm->mat.transmission;
        // stochastic opacity: triangles and capsules (spheres, cylinders
        // and cones have no opacity test in the reference)
        gd.translucent = (gm->type == 0 || gm->type == 3) && can_be_translucent(*gm, g) ? 1 : 0;
        BNX_SET(gd.vertices, gm->vertices.as<float>());
        BNX_SET(gd.indices, gm->indices.as<int>());
        BNX_SET(gd.colors, gm->colors.ptr ? gm->colors.as<float>() : nullptr);
        BNX_SET(gd.normals, gm->normals.ptr ? gm->normals.as<float>() : nullptr);
        BNX_SET(gd.spheres, gm->spheres.as<float4>());
        BNX_SET(gd.cyls, gm->cyls.as<float4>());
        for (int k = 0; k < ATTR_SLOTS; k++) {
          const HostAttr &a = gm->attr[k];
          AttrDev &o = gd.attr[k];
          o.scope = a.scope;
          o.comps = a.comps;
          o.value = make_float4(a.value[0], a.value[1], a.value[2], a.value[3]);
          BNX_SET(o.data, a.data.ptr ? a.data.as<float>() : nullptr);
        }
        geoms.push_back(gd);
      }
      g->max_blas_depth = std::max(g->max_blas_depth, bg->bvh.stack_need4);
      g->blas_prims += bg->bvh.num_prims;
    }
    const float *x = &g->xfms[12 * i];
    InstDev &I = insts[i];
    I.blas = bi;
    I.user_id = g->inst_ids[i];
    I.identity = is_identity(x) ? 1 : 0;
    I.geom_base = geom_base[bi];
    float inv[12];
    affine_inverse(x, inv);
    to_rows(x, I.o2w);
    to_rows(inv, I.w2o);
    // computeInstanceBounds (rtcore/cuda/Group.cu:125-137)
    const float *rl = bg->bvh.root_lo, *rh = bg->bvh.root_hi;
    float bl[3] = {INFINITY, INFINITY, INFINITY}, bh[3] = {-INFINITY, -INFINITY, -INFINITY};
    if (I.identity) {
      for (int k = 0; k < 3; k++) { bl[k] = rl[k]; bh[k] = rh[k]; }
    } else if (rl[0] <= rh[0]) {
      for (int c = 0; c < 8; c++) {
        float pc[3] = {(c & 1) ? rh[0] : rl[0], (c & 2) ? rh[1] : rl[1], (c & 4) ? rh[2] : rl[2]};
        for (int r = 0; r < 3; r++) {
          // ((vx*px + vy*py) + vz*pz) + p
          float v = ((x[r] * pc[0] + x[3 + r] * pc[1]) + x[6 + r] * pc[2]) + x[9 + r];
          bl[r] = std::min(bl[r], v);
          bh[r] = std::max(bh[r], v);
        }
      }
    }
    lo[i] = make_float4(bl[0], bl[1], bl[2], 0.f);
    hi[i] = make_float4(bh[0], bh[1], bh[2], 0.f);
  }
  std::vector<BlasDev> blas(blas_groups.size());
  for (size_t k = 0; k < blas_groups.size(); k++) {
    Group *bg = blas_groups[k];
    BNX_SET(blas[k].nodes, bg->bvh.nodes4.as<Node4>());
    BNX_SET(blas[k].prims, bg->prims.as<float4>());
    BNX_SET(blas[k].refs, bg->refs.as<int2>());
    // a user group holds one user type: spheres, cylinders, capsules or cones
    blas[k].type = bg->kind == GROUP_TRIANGLES ? BLAS_TRIANGLES
                                               : (bg->geoms.empty() ? BLAS_SPHERES : bg->geoms[0]->type);
    blas[k].num_prims = bg->bvh.num_prims;
    blas[k].translucent = 0;
    for (size_t j = 0; j < bg->geoms.size(); j++) blas[k].translucent |= geoms[geom_base[k] + j].translucent;
  }
  bool has_cyl = false;
  for (const BlasDev &b : blas) has_cyl |= b.type >= BLAS_CYLINDERS;
  if (has_cyl != g->has_cyl) {
    dev->cyl_worlds += has_cyl ? 1 : -1;
    g->has_cyl = has_cyl;
  }
  bool has_translucent = false;
  for (const BlasDev &b : blas) has_translucent |= b.translucent != 0;
  if (has_translucent != g->has_translucent) {
    dev->translucent_worlds += has_translucent ? 1 : -1;
    g->has_translucent = has_translucent;
  }
  bool has_xf = false;
  for (const InstDev &I : insts) has_xf |= !I.identity;
  if (has_xf != g->has_xf) {
    dev->xf_worlds += has_xf ? 1 : -1;
    g->has_xf = has_xf;
  }
  DevMem d_lo, d_hi;
  int rc;
  if ((rc = upload(dev, d_lo, lo.data(), n)) != BNX_OK) return rc;
  if ((rc = upload(dev, d_hi, hi.data(), n)) != BNX_OK) return rc;
  if ((rc = build_bvh(dev, d_lo.as<float4>(), d_hi.as<float4>(), n, 1, g->bvh)) != BNX_OK) return rc;
  if ((rc = collapse_bvh4(dev, g->bvh)) != BNX_OK) return rc;
  if ((rc = upload(dev, g->d_insts, insts.data(), insts.size())) != BNX_OK) return rc;
  if ((rc = upload(dev, g->d_blas, blas.data(), blas.size())) != BNX_OK) return rc;
  if ((rc = upload(dev, g->d_geoms, geoms.data(), geoms.size())) != BNX_OK) return rc;
  BNX_HIP(hipStreamSynchronize(dev->stream));
  // a walk pushes at most the TLAS path's entries, the LEAVE marker and a BLAS path's entries
  if (g->bvh.stack_need4 + 1 + g->max_blas_depth + 1 > trace_stack_capacity()) {
    set_error("bnx_group_build: BVH deeper than the traversal stack");
    return BNX_ERR_UNSUPPORTED;
  }
  BNX_SET(g->scene.tlas_nodes, g->bvh.nodes4.as<Node4>());
  BNX_SET(g->scene.tlas_order, g->bvh.order.as<int>());
  BNX_SET(g->scene.insts, g->d_insts.as<InstDev>());
  BNX_SET(g->scene.blas, g->d_blas.as<BlasDev>());
  BNX_SET(g->scene.geoms, g->d_geoms.as<GeomDev>());
  for (int k = 0; k < 5; k++)
    BNX_SET(g->scene.inst_attr[k], g->inst_attr[k].ptr ? g->inst_attr[k].as<float4>() : nullptr);
  g->scene.num_insts = n;
  g->built = true;
  return BNX_OK;
}

} // namespace

int world_volumes(World *w, VolumeList &vl)
{
  bool dirty = w->vol_seen.size() != w->volumes.size();
  for (size_t k = 0; !dirty && k < w->volumes.size(); k++) dirty = w->vol_seen[k] != w->volumes[k]->version;
  if (dirty) {
    std::vector<VolumeDev> vd;
    const int nsurf = w->instances ? (int)w->instances->groups.size() : 0;
    w->vol_seen.clear();
    for (size_t k = 0; k < w->volumes.size(); k++) {
      VolumeDev d = w->volumes[k]->dd;
      d.rtc_inst = nsurf + (int)k;
      d.user_inst = w->volume_inst_ids[k];
      vd.push_back(d);
      w->vol_seen.push_back(w->volumes[k]->version);
    }
    int rc = upload(w->dev, w->d_volumes, vd.data(), vd.size());
    if (rc != BNX_OK) return rc;
    BNX_HIP(hipStreamSynchronize(w->dev->stream));
  }
  BNX_SET(vl.vols, w->d_volumes.as<VolumeDev>());
  vl.count = (int)w->volumes.size();
  return BNX_OK;
}

int world_isos(World *w, const IsoDev BNX_G *&isos, int &count)
{
  bool dirty = w->iso_seen.size() != w->isos.size();
  for (size_t k = 0; !dirty && k < w->isos.size(); k++) dirty = w->iso_seen[k] != w->isos[k]->field->version;
  if (dirty) {
    std::vector<IsoDev> id;
    const int nsurf = w->instances ? (int)w->instances->groups.size() : 0;
    w->iso_seen.clear();
    for (size_t k = 0; k < w->isos.size(); k++) {
      const IsoSurface *s = w->isos[k];
      IsoDev d{};
      d.field = s->field->dd;
      d.iso = s->iso;
      for (int c = 0; c < 3; c++) d.color[c] = s->color[c];
      d.rtc_inst = nsurf + (int)w->volumes.size() + (int)k;
      d.user_inst = w->iso_inst_ids[k];
      d.user_id = s->user_id;
      id.push_back(d);
      w->iso_seen.push_back(s->field->version);
    }
    int rc = upload(w->dev, w->d_isos, id.data(), id.size());
    if (rc != BNX_OK) return rc;
    BNX_HIP(hipStreamSynchronize(w->dev->stream));
  }
  BNX_SET(isos, w->d_isos.ptr);
  count = (int)w->isos.size();
  return BNX_OK;
}

} // namespace bnx

using namespace bnx;

struct bnx_device_s : Device {};
struct bnx_buffer_s : Buffer {};
struct bnx_geom_s : Geom {};
struct bnx_group_s : Group {};
struct bnx_world_s : World {};
struct bnx_frame_s : Frame {};
struct bnx_volume_s : Volume {};
struct bnx_isosurface_s : IsoSurface {};

extern "C" {

const char *bnx_last_error(void) { return g_last_error.c_str(); }
int bnx_abi_version(void) { return BNX_ABI_VERSION; }

int bnx_device_count(int *count)
{
  BNX_CHECK(count, BNX_ERR_INVALID, "bnx_device_count: null out");
  *count = 0;
  BNX_HIP(hipGetDeviceCount(count));
  return BNX_OK;
}

int bnx_device_create(int gpu_id, bnx_device *out)
{
  GUARD_BEGIN
  BNX_CHECK(out, BNX_ERR_INVALID, "bnx_device_create: null out");
  int count = 0;
  BNX_HIP(hipGetDeviceCount(&count));
  BNX_CHECK(gpu_id >= 0 && gpu_id < count, BNX_ERR_INVALID, "bnx_device_create: no such GPU");
  BNX_HIP(hipSetDevice(gpu_id));
  bnx_device d = new bnx_device_s;
  d->gpu = gpu_id;
  BNX_HIP(hipStreamCreateWithFlags(&d->stream, hipStreamNonBlocking));
  hipDeviceProp_t prop;
  BNX_HIP(hipGetDeviceProperties(&prop, gpu_id));
  d->num_cus = prop.multiProcessorCount > 0 ? prop.multiProcessorCount : 256;
  int khz = 0;
  if (hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, gpu_id) == hipSuccess && khz > 0)
    d->wall_clock_khz = khz;
  // tuning overrides of the trace launch (pool refill threshold, blocks per CU)
  if (const char *e = getenv("BNX_TRACE_REFILL")) d->trace_refill = std::min(64, std::max(1, atoi(e)));
  if (const char *e = getenv("BNX_TRACE_REFILL_PRIMS")) d->trace_refill_prims = std::max(0LL, atoll(e));
  if (const char *e = getenv("BNX_TRACE_BLOCKS_PER_CU")) d->trace_blocks_per_cu = std::min(1 << 20, std::max(1, atoi(e)));
  int rc = noalloc_init(d);
  if (rc != BNX_OK) {
    (void)hipStreamDestroy(d->stream);
    delete d;
    return rc;
  }
  *out = d;
  return BNX_OK;
  GUARD_END
}

int bnx_device_destroy(bnx_device d)
{
  BNX_CHECK(d, BNX_ERR_INVALID, "bnx_device_destroy: null device");
  (void)hipStreamSynchronize(d->stream);
  wavefront_destroy(d);
  noalloc_destroy(d);
  (void)hipStreamDestroy(d->stream);
  delete d;
  return BNX_OK;
}

int bnx_device_sync(bnx_device d)
{
  BNX_CHECK(d, BNX_ERR_INVALID, "bnx_device_sync: null device");
  BNX_HIP(hipStreamSynchronize(d->stream));
  return BNX_OK;
}

int bnx_device_mem_info(bnx_device d, size_t *free_bytes, size_t *total_bytes)
{
  BNX_CHECK(d && free_bytes && total_bytes, BNX_ERR_INVALID, "bnx_device_mem_info: null argument");
  BNX_HIP(hipSetDevice(d->gpu));
  BNX_HIP(hipMemGetInfo(free_bytes, total_bytes));
  return BNX_OK;
}

// ------------------------------------------------------- device memory
int bnx_mem_alloc(bnx_device d, size_t bytes, void **d_ptr)
{
  BNX_CHECK(d && d_ptr, BNX_ERR_INVALID, "bnx_mem_alloc: null argument");
  *d_ptr = nullptr;
  BNX_HIP(hipSetDevice(d->gpu));
  const hipError_t e = hipMalloc(d_ptr, bytes ? bytes : 16);
  if (e != hipSuccess) {
    *d_ptr = nullptr;
    (void)hipGetLastError();
    set_error(std::string("bnx_mem_alloc: ") + hipGetErrorString(e));
    return BNX_ERR_OOM;
  }
  return BNX_OK;
}

int bnx_mem_free(bnx_device d, void *d_ptr)
{
  BNX_CHECK(d, BNX_ERR_INVALID, "bnx_mem_free: null device");
  if (!d_ptr) return BNX_OK;
  BNX_HIP(hipStreamSynchronize(d->stream));
  BNX_HIP(hipFree(d_ptr));
  return BNX_OK;
}

int bnx_host_alloc(bnx_device d, size_t bytes, void **h_ptr)
{
  BNX_CHECK(d && h_ptr, BNX_ERR_INVALID, "bnx_host_alloc: null argument");
  *h_ptr = nullptr;
  const hipError_t e = hipHostMalloc(h_ptr, bytes ? bytes : 16, hipHostMallocDefault);
  if (e != hipSuccess) {
    *h_ptr = nullptr;
    (void)hipGetLastError();
    set_error(std::string("bnx_host_alloc: ") + hipGetErrorString(e));
    return BNX_ERR_OOM;
  }
  return BNX_OK;
}

int bnx_host_free(bnx_device d, void *h_ptr)
{
  BNX_CHECK(d, BNX_ERR_INVALID, "bnx_host_free: null device");
  if (!h_ptr) return BNX_OK;
  BNX_HIP(hipStreamSynchronize(d->stream));
  BNX_HIP(hipHostFree(h_ptr));
  return BNX_OK;
}

int bnx_mem_copy_async(bnx_device d, void *dst, const void *src, size_t bytes)
{
  BNX_CHECK(d && (bytes == 0 || (dst && src)), BNX_ERR_INVALID, "bnx_mem_copy_async: null argument");
  if (bytes) BNX_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDefault, d->stream));
  return BNX_OK;
}

int bnx_mem_set_async(bnx_device d, void *d_ptr, int value, size_t bytes)
{
  BNX_CHECK(d && (bytes == 0 || d_ptr), BNX_ERR_INVALID, "bnx_mem_set_async: null argument");
  if (bytes) BNX_HIP(hipMemsetAsync(d_ptr, value, bytes, d->stream));
  return BNX_OK;
}

int bnx_device_set_active(bnx_device d, int *previous_gpu)
{
  BNX_CHECK(d && previous_gpu, BNX_ERR_INVALID, "bnx_device_set_active: null argument");
  BNX_HIP(hipGetDevice(previous_gpu));
  BNX_HIP(hipSetDevice(d->gpu));
  return BNX_OK;
}

int bnx_device_restore_active(int previous_gpu)
{
  if (previous_gpu < 0) return BNX_OK;
  BNX_HIP(hipSetDevice(previous_gpu));
  return BNX_OK;
}

// ------------------------------------------------------------------ buffers
int bnx_buffer_create(bnx_device d, size_t bytes, const void *init, bnx_buffer *out)
{
  GUARD_BEGIN
  BNX_CHECK(d && out, BNX_ERR_INVALID, "bnx_buffer_create: null argument");
  bnx_buffer b = new bnx_buffer_s;
  b->dev = d;
  hipError_t e = b->mem.ensure(bytes);
  if (e != hipSuccess) {
    delete b;
    set_error(std::string("bnx_buffer_create: ") + hipGetErrorString(e));
    return BNX_ERR_OOM;
  }
  if (init && bytes) BNX_HIP(hipMemcpy(b->mem.ptr, init, bytes, hipMemcpyHostToDevice));
  *out = b;
  return BNX_OK;
  GUARD_END
}

int bnx_buffer_upload(bnx_buffer b, const void *host, size_t bytes, size_t offset)
{
  BNX_CHECK(b && host, BNX_ERR_INVALID, "bnx_buffer_upload: null argument");
  BNX_CHECK(offset + bytes <= b->mem.bytes, BNX_ERR_INVALID, "bnx_buffer_upload: out of range");
  BNX_HIP(hipMemcpyAsync((char *)b->mem.ptr + offset, host, bytes, hipMemcpyHostToDevice, b->dev->stream));
  BNX_HIP(hipStreamSynchronize(b->dev->stream));
  return BNX_OK;
}

int bnx_copy(bnx_device d, void *dst, const void *src, size_t bytes)
{
  BNX_CHECK(dst && src, BNX_ERR_INVALID, "bnx_copy: null argument");
  if (bytes == 0) return BNX_OK;
  if (!d) {
    // no device handle: after all work of the current device
    BNX_HIP(hipDeviceSynchronize());
    BNX_HIP(hipMemcpy(dst, src, bytes, hipMemcpyDefault));
    return BNX_OK;
  }
  BNX_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDefault, d->stream));
  BNX_HIP(hipStreamSynchronize(d->stream));
  return BNX_OK;
}

int bnx_buffer_download(bnx_buffer b, void *host, size_t bytes, size_t offset)
{
  BNX_CHECK(b && host, BNX_ERR_INVALID, "bnx_buffer_download: null argument");
  BNX_CHECK(offset + bytes <= b->mem.bytes, BNX_ERR_INVALID, "bnx_buffer_download: out of range");
  BNX_HIP(hipMemcpyAsync(host, (char *)b->mem.ptr + offset, bytes, hipMemcpyDeviceToHost, b->dev->stream));
  BNX_HIP(hipStreamSynchronize(b->dev->stream));
  return BNX_OK;
}

int bnx_buffer_destroy(bnx_buffer b)
{
  BNX_CHECK(b, BNX_ERR_INVALID, "bnx_buffer_destroy: null buffer");
  (void)hipStreamSynchronize(b->dev->stream);
  delete b;
  return BNX_OK;
}

void *bnx_buffer_dptr(bnx_buffer b) { return b ? b->mem.ptr : nullptr; }
size_t bnx_buffer_size(bnx_buffer b) { return b ? b->mem.bytes : 0; }

// ------------------------------------------------------------------- geoms
static void set_mat(Geom *g, const bnx_material &m) { g->mat = m; }

// AnariMatte, or AnariPBR creating the NVisii BSDF (AnariPBR.h:62-96).
// Not provided: the Glass BSDF (PBR with ior != 1 and transmission >=
// 1e-3).  Triangles whose alpha is below 1 take the stochastic opacity test
// (Triangles.dev.cu:120-133); spheres and cylinders have no any-hit program
// and ignore opacity.
static int check_material(const bnx_material &m, bool)
{
  BNX_CHECK(m.type == BNX_MATERIAL_MATTE || m.type == BNX_MATERIAL_PBR, BNX_ERR_UNSUPPORTED,
            "only matte and physically-based materials are provided");
  if (m.type == BNX_MATERIAL_PBR)
    BNX_CHECK(m.ior == 1.f || m.transmission < 1e-3f, BNX_ERR_UNSUPPORTED,
              "the Glass BSDF (PBR transmission with ior != 1) is not provided");
  BNX_CHECK(m.color_from_attribute >= BNX_ATTRIB_NONE && m.color_from_attribute <= BNX_ATTRIB_PRIMITIVE_ID,
            BNX_ERR_INVALID, "material: unknown color attribute");
  return BNX_OK;
}

// the "color" attribute per primitive from a float3 array of a geometry desc
static int set_prim_colors(Geom *g, const float *colors, int n)
{
  HostAttr &a = g->attr[BNX_ATTRIB_SLOT_COLOR];
  a.scope = BNX_SCOPE_PER_PRIM;
  a.comps = 3;
  a.w_below_1 = false;
  return upload(g->dev, a.data, colors, 3 * (size_t)n);
}

int bnx_geom_create_triangles(bnx_device d, const bnx_triangles_desc *desc, bnx_geom *out)
{
  GUARD_BEGIN
  BNX_CHECK(d && desc && out, BNX_ERR_INVALID, "bnx_geom_create_triangles: null argument");
  BNX_CHECK(desc->num_triangles >= 0 && desc->num_vertices >= 0, BNX_ERR_INVALID, "negative count");
  BNX_CHECK(desc->num_triangles == 0 || (desc->vertices && desc->indices), BNX_ERR_INVALID,
            "bnx_geom_create_triangles: missing vertices or indices");
  if (int rc = check_material(desc->material, true)) return rc;
  for (int i = 0; i < 3 * desc->num_triangles; i++)
    BNX_CHECK(desc->indices[i] >= 0 && desc->indices[i] < desc->num_vertices, BNX_ERR_INVALID,
              "bnx_geom_create_triangles: index out of range");
  bnx_geom g = new bnx_geom_s;
  g->dev = d;
  g->type = 0;
  g->user_id = desc->user_id;
  set_mat(g, desc->material);
  g->num_prims = desc->num_triangles;
  g->num_vertices = desc->num_vertices;
  int rc;
  if ((rc = upload(d, g->vertices, desc->vertices, 3 * (size_t)desc->num_vertices)) != BNX_OK ||
      (rc = upload(d, g->indices, desc->indices, 3 * (size_t)desc->num_triangles)) != BNX_OK) {
    delete g;
    return rc;
  }
  if (desc->prim_colors && (rc = set_prim_colors(g, desc->prim_colors, desc->num_triangles)) != BNX_OK) {
    delete g;
    return rc;
  }
  BNX_HIP(hipStreamSynchronize(d->stream));
  *out = g;
  return BNX_OK;
  GUARD_END
}

int bnx_geom_create_cylinders(bnx_device d, const bnx_cylinders_desc *desc, bnx_geom *out)
{
  GUARD_BEGIN
  BNX_CHECK(d && desc && out, BNX_ERR_INVALID, "bnx_geom_create_cylinders: null argument");
  BNX_CHECK(desc->num_cylinders >= 0 && desc->num_vertices >= 0, BNX_ERR_INVALID, "negative count");
  BNX_CHECK(desc->num_cylinders == 0 || (desc->vertices && desc->indices), BNX_ERR_INVALID,
            "bnx_geom_create_cylinders: missing vertices or indices");
  if (int rc = check_material(desc->material, false)) return rc;
  std::vector<float4> rec(2 * (size_t)desc->num_cylinders);
  for (int i = 0; i < desc->num_cylinders; i++) {
    const int a = desc->indices[2 * i], b = desc->indices[2 * i + 1];
    BNX_CHECK(a >= 0 && a < desc->num_vertices && b >= 0 && b < desc->num_vertices, BNX_ERR_INVALID,
              "bnx_geom_create_cylinders: index out of range");
    const float r = desc->radii ? desc->radii[i] : desc->default_radius;
    const float *va = desc->vertices + 3 * (size_t)a, *vb = desc->vertices + 3 * (size_t)b;
    rec[2 * i] = make_float4(va[0], va[1], va[2], r);
    rec[2 * i + 1] = make_float4(vb[0], vb[1], vb[2], 0.f);
  }
  bnx_geom g = new bnx_geom_s;
  g->dev = d;
  g->type = 2;
  g->user_id = desc->user_id;
  set_mat(g, desc->material);
  g->num_prims = desc->num_cylinders;
  g->num_vertices = desc->num_vertices;
  int rc;
  if ((rc = upload(d, g->cyls, rec.data(), rec.size())) != BNX_OK ||
      (rc = upload(d, g->indices, desc->indices, 2 * (size_t)desc->num_cylinders)) != BNX_OK) {
    delete g;
    return rc;
  }
  if (desc->colors && (rc = set_prim_colors(g, desc->colors, desc->num_cylinders)) != BNX_OK) {
    delete g;
    return rc;
  }
  BNX_HIP(hipStreamSynchronize(d->stream));
  *out = g;
  return BNX_OK;
  GUARD_END
}

// a user geometry with two records per primitive (capsules, cones) from
// its end-point records and index pairs
static int create_pair_geom(bnx_device d, int type, const std::vector<float4> &rec, const std::vector<int> &idx,
                            int num_prims, int num_vertices, const float *colors, int32_t user_id,
                            const bnx_material &material, bnx_geom *out)
{
  bnx_geom g = new bnx_geom_s;
  g->dev = d;
  g->type = type;
  g->user_id = user_id;
  set_mat(g, material);
  g->num_prims = num_prims;
  g->num_vertices = num_vertices;
  int rc;
  if ((rc = upload(d, g->cyls, rec.data(), rec.size())) != BNX_OK ||
      (rc = upload(d, g->indices, idx.data(), idx.size())) != BNX_OK) {
    delete g;
    return rc;
  }
  if (colors && (rc = set_prim_colors(g, colors, num_prims)) != BNX_OK) {
    delete g;
    return rc;
  }
  BNX_HIP(hipStreamSynchronize(d->stream));
  *out = g;
  return BNX_OK;
}

int bnx_geom_create_capsules(bnx_device d, const bnx_capsules_desc *desc, bnx_geom *out)
{
  GUARD_BEGIN
  BNX_CHECK(d && desc && out, BNX_ERR_INVALID, "bnx_geom_create_capsules: null argument");
  BNX_CHECK(desc->num_capsules >= 0 && desc->num_vertices >= 0, BNX_ERR_INVALID, "negative count");
  BNX_CHECK(desc->num_capsules == 0 || (desc->vertices && desc->indices), BNX_ERR_INVALID,
            "bnx_geom_create_capsules: missing vertices or indices");
  if (int rc = check_material(desc->material, false)) return rc;
  const int n = desc->num_capsules;
  std::vector<float4> rec(2 * (size_t)n);
  std::vector<int> idx(2 * (size_t)n);
  for (int i = 0; i < n; i++) {
    const int a = desc->indices[2 * i], b = desc->indices[2 * i + 1];
    BNX_CHECK(a >= 0 && a < desc->num_vertices && b >= 0 && b < desc->num_vertices, BNX_ERR_INVALID,
              "bnx_geom_create_capsules: index out of range");
    const float *va = desc->vertices + 4 * (size_t)a, *vb = desc->vertices + 4 * (size_t)b;
    rec[2 * i] = make_float4(va[0], va[1], va[2], va[3]);
    rec[2 * i + 1] = make_float4(vb[0], vb[1], vb[2], vb[3]);
    idx[2 * i] = a;
    idx[2 * i + 1] = b;
  }
  return create_pair_geom(d, 3, rec, idx, n, desc->num_vertices, desc->colors, desc->user_id, desc->material, out);
  GUARD_END
}

int bnx_geom_create_cones(bnx_device d, const bnx_cones_desc *desc, bnx_geom *out)
{
  GUARD_BEGIN
  BNX_CHECK(d && desc && out, BNX_ERR_INVALID, "bnx_geom_create_cones: null argument");
  BNX_CHECK(desc->num_cones >= 0 && desc->num_vertices >= 0, BNX_ERR_INVALID, "negative count");
  BNX_CHECK(desc->num_cones == 0 || (desc->vertices && desc->radii), BNX_ERR_INVALID,
            "bnx_geom_create_cones: missing vertices or radii");
  if (int rc = check_material(desc->material, false)) return rc;
  const int n = desc->num_cones;
  std::vector<float4> rec(2 * (size_t)n);
  std::vector<int> idx(2 * (size_t)n);
  for (int i = 0; i < n; i++) {
    // Cones.dev.cu:81-84: without indices cone i joins vertices 2i, 2i + 1
    const int a = desc->indices ? desc->indices[2 * i] : 2 * i;
    const int b = desc->indices ? desc->indices[2 * i + 1] : 2 * i + 1;
    BNX_CHECK(a >= 0 && a < desc->num_vertices && b >= 0 && b < desc->num_vertices, BNX_ERR_INVALID,
              "bnx_geom_create_cones: index out of range");
    const float *va = desc->vertices + 3 * (size_t)a, *vb = desc->vertices + 3 * (size_t)b;
    rec[2 * i] = make_float4(va[0], va[1], va[2], desc->radii[a]);
    rec[2 * i + 1] = make_float4(vb[0], vb[1], vb[2], desc->radii[b]);
    idx[2 * i] = a;
    idx[2 * i + 1] = b;
  }
  return create_pair_geom(d, 4, rec, idx, n, desc->num_vertices, desc->colors, desc->user_id, desc->material, out);
  GUARD_END
}

int bnx_geom_create_spheres(bnx_device d, const bnx_spheres_desc *desc, bnx_geom *out)
{
  GUARD_BEGIN
  BNX_CHECK(d && desc && out, BNX_ERR_INVALID, "bnx_geom_create_spheres: null argument");
  BNX_CHECK(desc->num_spheres >= 0, BNX_ERR_INVALID, "negative count");
  BNX_CHECK(desc->num_spheres == 0 || desc->centers, BNX_ERR_INVALID, "bnx_geom_create_spheres: missing centers");
  if (int rc = check_material(desc->material, false)) return rc;
  bnx_geom g = new bnx_geom_s;
  g->dev = d;
  g->type = 1;
  g->user_id = desc->user_id;
  set_mat(g, desc->material);
  g->num_prims = desc->num_spheres;
  g->num_vertices = 0;
  std::vector<float4> sp((size_t)desc->num_spheres);
  for (int i = 0; i < desc->num_spheres; i++)
    sp[i] = make_float4(desc->centers[3 * i], desc->centers[3 * i + 1], desc->centers[3 * i + 2],
                        desc->radii ? desc->radii[i] : desc->default_radius);
  int rc;
  if ((rc = upload(d, g->spheres, sp.data(), sp.size())) != BNX_OK) {
    delete g;
    return rc;
  }
  if (desc->colors && (rc = upload(d, g->colors, desc->colors, 3 * (size_t)desc->num_spheres)) != BNX_OK) {
    delete g;
    return rc;
  }
  BNX_HIP(hipStreamSynchronize(d->stream));
  *out = g;
  return BNX_OK;
  GUARD_END
}

int bnx_geom_set_attribute(bnx_geom g, int slot, int scope, int comps, const float *data, int64_t count,
                           const float value[4])
{
  GUARD_BEGIN
  BNX_CHECK(g, BNX_ERR_INVALID, "bnx_geom_set_attribute: null geom");
  BNX_CHECK(slot >= 0 && slot < ATTR_SLOTS, BNX_ERR_INVALID, "bnx_geom_set_attribute: slot out of range");
  BNX_CHECK(scope >= BNX_SCOPE_NONE && scope <= BNX_SCOPE_FACE_VARYING, BNX_ERR_INVALID,
            "bnx_geom_set_attribute: unknown scope");
  HostAttr &a = g->attr[slot];
  if (scope == BNX_SCOPE_NONE) {
    a.scope = BNX_SCOPE_NONE;
    a.comps = 0;
    a.data.release();
    a.w_below_1 = false;
    return BNX_OK;
  }
  if (scope == BNX_SCOPE_CONSTANT) {
    BNX_CHECK(value, BNX_ERR_INVALID, "bnx_geom_set_attribute: constant without a value");
    a.scope = scope;
    a.comps = 4;
    for (int k = 0; k < 4; k++) a.value[k] = value[k];
    a.data.release();
    a.w_below_1 = !(value[3] >= 1.f);
    return BNX_OK;
  }
  BNX_CHECK(comps >= 1 && comps <= 4, BNX_ERR_INVALID, "bnx_geom_set_attribute: 1 to 4 floats per element");
  // elements a hit can read: per primitive, per vertex, or per corner
  // (3 per triangle, 2 per cylinder; a sphere reads its own element)
  int64_t need = g->num_prims;
  if (scope == BNX_SCOPE_PER_VERTEX) need = g->type == 1 ? g->num_prims : g->num_vertices;
  if (scope == BNX_SCOPE_FACE_VARYING) need = g->type == 0 ? 3LL * g->num_prims
                                             : (g->type >= 2 ? 2LL * g->num_prims : g->num_prims);
  BNX_CHECK(data && count >= need, BNX_ERR_INVALID,
            "bnx_geom_set_attribute: fewer elements than the scope needs (" + std::to_string(need) + ")");
  int rc;
  if ((rc = upload(g->dev, a.data, data, (size_t)comps * (size_t)count)) != BNX_OK) return rc;
  BNX_HIP(hipStreamSynchronize(g->dev->stream));
  a.scope = scope;
  a.comps = comps;
  a.w_below_1 = false;
  if (comps == 4)
    for (int64_t i = 0; i < count && !a.w_below_1; i++) a.w_below_1 = !(data[4 * i + 3] >= 1.f);
  return BNX_OK;
  GUARD_END
}

int bnx_geom_set_vertex_normals(bnx_geom g, const float *normals, int n)
{
  GUARD_BEGIN
  BNX_CHECK(g && g->type == 0, BNX_ERR_INVALID, "bnx_geom_set_vertex_normals: not a triangles geometry");
  if (!normals) {
    g->normals.release();
    return BNX_OK;
  }
  BNX_CHECK(n == g->num_vertices, BNX_ERR_INVALID, "bnx_geom_set_vertex_normals: one normal per vertex");
  int rc;
  if ((rc = upload(g->dev, g->normals, normals, 3 * (size_t)n)) != BNX_OK) return rc;
  BNX_HIP(hipStreamSynchronize(g->dev->stream));
  return BNX_OK;
  GUARD_END
}

int bnx_geom_destroy(bnx_geom g)
{
  BNX_CHECK(g, BNX_ERR_INVALID, "bnx_geom_destroy: null geom");
  (void)hipStreamSynchronize(g->dev->stream);
  delete g;
  return BNX_OK;
}

// ------------------------------------------------------------------ groups
static int create_geom_group(bnx_device d, const bnx_geom *geoms, int n, int kind, bnx_group *out)
{
  GUARD_BEGIN
  BNX_CHECK(d && out && (n == 0 || geoms), BNX_ERR_INVALID, "group create: null argument");
  bnx_group g = new bnx_group_s;
  g->dev = d;
  g->kind = kind;
  for (int i = 0; i < n; i++) {
    const bool ok = geoms[i] && (kind == GROUP_TRIANGLES ? geoms[i]->type == 0 : geoms[i]->type != 0);
    if (!ok) {
      delete g;
      set_error("group create: geometry type does not match the group kind");
      return BNX_ERR_INVALID;
    }
    if (i > 0 && geoms[i]->type != geoms[0]->type) {
      delete g;
      set_error("group create: a user-geometry group holds one type (spheres, cylinders, capsules or cones)");
      return BNX_ERR_UNSUPPORTED;
    }
    g->geoms.push_back(geoms[i]);
  }
  *out = g;
  return BNX_OK;
  GUARD_END
}

int bnx_group_create_triangles(bnx_device d, const bnx_geom *geoms, int n, bnx_group *out)
{
  return create_geom_group(d, geoms, n, GROUP_TRIANGLES, out);
}

int bnx_group_create_user(bnx_device d, const bnx_geom *geoms, int n, bnx_group *out)
{
  return create_geom_group(d, geoms, n, GROUP_USER, out);
}

int bnx_group_create_instances(bnx_device d, const bnx_group *groups, const int32_t *inst_ids, const float *xfms,
                               int n, bnx_group *out)
{
  GUARD_BEGIN
  BNX_CHECK(d && out && (n == 0 || groups), BNX_ERR_INVALID, "bnx_group_create_instances: null argument");
  bnx_group g = new bnx_group_s;
  g->dev = d;
  g->kind = GROUP_INSTANCES;
  for (int i = 0; i < n; i++) {
    g->groups.push_back(groups[i]);
    g->inst_ids.push_back(inst_ids ? inst_ids[i] : i);
    static const float I[12] = {1, 0, 0, 0, 1, 0, 0, 0, 1, 0, 0, 0};
    const float *x = xfms ? xfms + 12 * i : I;
    g->xfms.insert(g->xfms.end(), x, x + 12);
  }
  *out = g;
  return BNX_OK;
  GUARD_END
}

int bnx_group_build(bnx_group g)
{
  GUARD_BEGIN
  BNX_CHECK(g, BNX_ERR_INVALID, "bnx_group_build: null group");
  BNX_HIP(hipSetDevice(g->dev->gpu));
  if (g->kind == GROUP_INSTANCES) return build_tlas(g);
  return build_blas(g);
  GUARD_END
}

int bnx_group_set_instance_attribute(bnx_group g, int slot, const float *values, int n)
{
  GUARD_BEGIN
  BNX_CHECK(g && g->kind == GROUP_INSTANCES, BNX_ERR_INVALID,
            "bnx_group_set_instance_attribute: not an instance group");
  BNX_CHECK(slot >= 0 && slot < 5, BNX_ERR_INVALID, "bnx_group_set_instance_attribute: slot out of range");
  if (!values) {
    g->inst_attr[slot].release();
    g->inst_attr_w_below_1[slot] = false;
    return BNX_OK;
  }
  BNX_CHECK(n == (int)g->groups.size(), BNX_ERR_INVALID,
            "bnx_group_set_instance_attribute: one value per instance");
  int rc;
  if ((rc = upload(g->dev, g->inst_attr[slot], (const float4 *)values, (size_t)n)) != BNX_OK) return rc;
  BNX_HIP(hipStreamSynchronize(g->dev->stream));
  bool below = false;
  for (int i = 0; i < n && !below; i++) below = !(values[4 * i + 3] >= 1.f);
  g->inst_attr_w_below_1[slot] = below;
  return BNX_OK;
  GUARD_END
}

int bnx_group_set_transforms(bnx_group g, const float *xfms, int n)
{
  GUARD_BEGIN
  BNX_CHECK(g && xfms, BNX_ERR_INVALID, "bnx_group_set_transforms: null argument");
  BNX_CHECK(g->kind == GROUP_INSTANCES, BNX_ERR_INVALID, "bnx_group_set_transforms: not an instance group");
  BNX_CHECK(n == (int)g->groups.size(), BNX_ERR_INVALID,
            "bnx_group_set_transforms: instance count differs from the group's");
  for (int i = 0; i < 12 * n; i++)
    BNX_CHECK(std::isfinite(xfms[i]), BNX_ERR_INVALID, "bnx_group_set_transforms: transform not finite");
  std::copy(xfms, xfms + 12 * (size_t)n, g->xfms.begin());
  BNX_HIP(hipSetDevice(g->dev->gpu));
  // refitAccel is buildAccel in the reference's CUDA, HIP and CPU backends
  return build_tlas(g);
  GUARD_END
}

int bnx_group_bvh_info(bnx_group g, bnx_bvh_info *info)
{
  BNX_CHECK(g && info, BNX_ERR_INVALID, "bnx_group_bvh_info: null argument");
  BNX_CHECK(g->built, BNX_ERR_STATE, "bnx_group_bvh_info: group not built");
  info->num_nodes = g->bvh.num_nodes;
  info->num_prims = g->bvh.num_prims;
  info->max_depth = g->bvh.max_depth;
  info->root_is_leaf = g->bvh.num_prims <= (g->kind == GROUP_INSTANCES ? 1 : 4) ? 1 : 0;
  for (int k = 0; k < 3; k++) {
    info->root_lower[k] = g->bvh.root_lo[k];
    info->root_upper[k] = g->bvh.root_hi[k];
  }
  info->build_ms = g->bvh.build_ms;
  return BNX_OK;
}

int bnx_group_download_bvh(bnx_group g, void *nodes, int32_t *order)
{
  BNX_CHECK(g, BNX_ERR_INVALID, "bnx_group_download_bvh: null group");
  BNX_CHECK(g->built, BNX_ERR_STATE, "bnx_group_download_bvh: group not built");
  if (nodes)
    BNX_HIP(hipMemcpy(nodes, g->bvh.nodes.ptr, sizeof(Node) * (size_t)g->bvh.num_nodes, hipMemcpyDeviceToHost));
  if (order && g->bvh.num_prims)
    BNX_HIP(hipMemcpy(order, g->bvh.order.ptr, sizeof(int) * (size_t)g->bvh.num_prims, hipMemcpyDeviceToHost));
  return BNX_OK;
}

int bnx_group_destroy(bnx_group g)
{
  BNX_CHECK(g, BNX_ERR_INVALID, "bnx_group_destroy: null group");
  (void)hipStreamSynchronize(g->dev->stream);
  if (g->has_cyl) g->dev->cyl_worlds--;
  if (g->has_xf) g->dev->xf_worlds--;
  if (g->has_translucent) g->dev->translucent_worlds--;
  delete g;
  return BNX_OK;
}

int bnx_bvh_build_boxes(bnx_device d, const float *d_boxes, int n, int max_leaf, void *nodes_host,
                        int32_t *order_host, bnx_bvh_info *info)
{
  GUARD_BEGIN
  BNX_CHECK(d && (n == 0 || d_boxes) && n >= 0 && max_leaf >= 1 && max_leaf <= 7, BNX_ERR_INVALID,
            "bnx_bvh_build_boxes: bad argument");
  DevMem lo, hi;
  BNX_HIP(lo.ensure(sizeof(float4) * (size_t)(n ? n : 1)));
  BNX_HIP(hi.ensure(sizeof(float4) * (size_t)(n ? n : 1)));
  if (n) k_boxes_from_floats<<<nblk(n, 256), 256, 0, d->stream>>>(d_boxes, n, lo.as<float4>(), hi.as<float4>());
  Bvh bvh;
  int rc = build_bvh(d, lo.as<float4>(), hi.as<float4>(), n, max_leaf, bvh);
  if (rc != BNX_OK) return rc;
  if (nodes_host)
    BNX_HIP(hipMemcpy(nodes_host, bvh.nodes.ptr, sizeof(Node) * (size_t)bvh.num_nodes, hipMemcpyDeviceToHost));
  if (order_host && n) BNX_HIP(hipMemcpy(order_host, bvh.order.ptr, sizeof(int) * (size_t)n, hipMemcpyDeviceToHost));
  if (info) {
    info->num_nodes = bvh.num_nodes;
    info->num_prims = n;
    info->max_depth = bvh.max_depth;
    info->root_is_leaf = n <= max_leaf;
    for (int k = 0; k < 3; k++) {
      info->root_lower[k] = bvh.root_lo[k];
      info->root_upper[k] = bvh.root_hi[k];
    }
    info->build_ms = bvh.build_ms;
  }
  return BNX_OK;
  GUARD_END
}

// ------------------------------------------------------------------- trace
int bnx_trace(bnx_group g, const bnx_ray *d_rays, int n, bnx_hit *d_hits, uint32_t flags)
{
  BNX_CHECK(g && n >= 0 && (n == 0 || (d_rays && d_hits)), BNX_ERR_INVALID, "bnx_trace: bad argument");
  BNX_CHECK(g->kind == GROUP_INSTANCES && g->built, BNX_ERR_STATE, "bnx_trace: needs a built instance group");
  if (n == 0) return BNX_OK;
  return launch_trace_api(g, d_rays, n, d_hits, flags);
}

// ----------------------------------------------------------------- volumes
static int set_tf(Volume *v, const bnx_transfer_function *tf)
{
  BNX_CHECK(tf && tf->values && tf->num_values >= 2, BNX_ERR_INVALID, "transfer function needs >= 2 values");
  BNX_CHECK(tf->domain[1] > tf->domain[0], BNX_ERR_INVALID, "transfer function domain is empty");
  int rc = upload(v->dev, v->tf, (const float4 *)tf->values, (size_t)tf->num_values);
  if (rc != BNX_OK) return rc;
  BNX_SET(v->dd.tf, v->tf.as<float4>());
  v->dd.ntf = tf->num_values;
  v->dd.dom0 = tf->domain[0];
  v->dd.dom1 = tf->domain[1];
  v->dd.base_density = tf->base_density;
  v->version++;
  return BNX_OK;
}

int bnx_volume_create_structured(bnx_device d, const bnx_structured_desc *f, const bnx_transfer_function *tf,
                                 int32_t user_id, bnx_volume *out)
{
  GUARD_BEGIN
  BNX_CHECK(d && f && tf && out && f->scalars, BNX_ERR_INVALID, "bnx_volume_create_structured: null argument");
  BNX_CHECK(f->dims[0] >= 2 && f->dims[1] >= 2 && f->dims[2] >= 2, BNX_ERR_INVALID,
            "bnx_volume_create_structured: dims must be >= 2");
  BNX_CHECK(f->grid_spacing[0] > 0.f && f->grid_spacing[1] > 0.f && f->grid_spacing[2] > 0.f, BNX_ERR_INVALID,
            "bnx_volume_create_structured: spacing must be positive");
  bnx_volume v = new bnx_volume_s;
  v->dev = d;
  VolumeDev &dd = v->dd;
  const size_t n = (size_t)f->dims[0] * f->dims[1] * f->dims[2];
  int rc = upload(d, v->scalars, f->scalars, n);
  if (rc != BNX_OK) {
    delete v;
    return rc;
  }
  BNX_SET(dd.scalars, v->scalars.as<float>());
  const int c = f->cells_per_mc > 0 ? f->cells_per_mc : 8;
  dd.cells_per_mc = c;
  for (int k = 0; k < 3; k++) {
    dd.dims[k] = f->dims[k];
    dd.ncells[k] = f->dims[k] - 1;
    dd.origin[k] = f->grid_origin[k];
    dd.rcp_spacing[k] = 1.f / f->grid_spacing[k];
    // StructuredData::commit (StructuredData.cu:177-181)
    dd.lo[k] = f->grid_origin[k];
    dd.hi[k] = f->grid_origin[k] + f->grid_spacing[k] * (float)dd.ncells[k];
    // buildMCs (StructuredData.cu:71-78)
    dd.mc[k] = (dd.ncells[k] + c - 1) / c;
    dd.mc_origin[k] = dd.lo[k];
    dd.mc_rcp_spacing[k] = 1.f / ((float)c * f->grid_spacing[k]);
  }
  dd.user_id = user_id;
  // the sampler reads the linear array unless BNX_VOLUME_BRICK asks for a
  // bricked copy (bricks of 2^shift cells a side with a one-texel apron); a
  // layout only, the values and results are the same.  On the 256^3 field
  // (1024^2, 16 spp) bricks of 4^3..16^3 cells cut the march's L2 misses by
  // a third but ran 4-5% slower than the linear array: the march is bound by
  // the latency of its dependent gathers, not by their bytes
  int shift = 0;
  if (const char *e = getenv("BNX_VOLUME_BRICK")) shift = std::min(5, std::max(0, atoi(e)));
  if ((rc = set_tf(v, tf)) != BNX_OK || (rc = volume_build(v)) != BNX_OK || (rc = volume_brick(v, shift)) != BNX_OK) {
    delete v;
    return rc;
  }
  *out = v;
  return BNX_OK;
  GUARD_END
}

int bnx_volume_set_transfer_function(bnx_volume v, const bnx_transfer_function *tf)
{
  GUARD_BEGIN
  BNX_CHECK(v, BNX_ERR_INVALID, "bnx_volume_set_transfer_function: null volume");
  int rc = set_tf(v, tf);
  if (rc != BNX_OK) return rc;
  return volume_rebuild_majorants(v);
  GUARD_END
}

int bnx_volume_majorants(bnx_volume v, int32_t dims[3], float *maj)
{
  BNX_CHECK(v && dims, BNX_ERR_INVALID, "bnx_volume_majorants: null argument");
  for (int k = 0; k < 3; k++) dims[k] = v->dd.mc[k];
  if (maj)
    BNX_HIP(hipMemcpy(maj, v->majorants.ptr, sizeof(float) * (size_t)dims[0] * dims[1] * dims[2],
                      hipMemcpyDeviceToHost));
  return BNX_OK;
}

int bnx_volume_destroy(bnx_volume v)
{
  BNX_CHECK(v, BNX_ERR_INVALID, "bnx_volume_destroy: null volume");
  (void)hipStreamSynchronize(v->dev->stream);
  delete v;
  return BNX_OK;
}

int bnx_volume_create_umesh(bnx_device d, const bnx_umesh_desc *f, const bnx_transfer_function *tf, int32_t user_id,
                            bnx_volume *out)
{
  GUARD_BEGIN
  BNX_CHECK(d && f && tf && out && f->vertices && f->scalars && f->indices && f->cell_offsets, BNX_ERR_INVALID,
            "bnx_volume_create_umesh: null argument");
  BNX_CHECK(f->num_vertices > 0 && f->num_cells > 0 && f->num_indices > 0, BNX_ERR_INVALID,
            "bnx_volume_create_umesh: empty mesh");
  const bool auto_mc = f->mc_dims[0] == 0 && f->mc_dims[1] == 0 && f->mc_dims[2] == 0;
  BNX_CHECK(auto_mc || (f->mc_dims[0] > 0 && f->mc_dims[1] > 0 && f->mc_dims[2] > 0 && f->mc_dims[0] <= 1024 &&
                        f->mc_dims[1] <= 1024 && f->mc_dims[2] <= 1024),
            BNX_ERR_INVALID, "bnx_volume_create_umesh: macrocell dims must be all 0 or all in 1..1024");
  const int nc = f->num_cells, nv = f->num_vertices;
  // cell types (UMeshField.h:17-29): ANARI and VTK codes of tets, hexahedra,
  // prisms and pyramids, kept as the ANARI code
  std::vector<uint8_t> types((size_t)nc, (uint8_t)CELL_TET);
  bool mixed = false;
  for (int c = 0; f->cell_types && c < nc; c++) {
    const int t = f->cell_types[c];
    int k;
    if (t == BNX_CELL_TET || t == BNX_CELL_VTK_TET) k = CELL_TET;
    else if (t == BNX_CELL_HEX || t == BNX_CELL_VTK_HEX) k = CELL_HEX;
    else if (t == BNX_CELL_PRISM || t == BNX_CELL_VTK_PRISM) k = CELL_PRISM;
    else if (t == BNX_CELL_PYR || t == BNX_CELL_VTK_PYR) k = CELL_PYR;
    else {
      set_error("bnx_volume_create_umesh: cell type " + std::to_string(t) +
                " not provided (tets, hexahedra, prisms and pyramids are)");
      return BNX_ERR_UNSUPPORTED;
    }
    types[c] = (uint8_t)k;
    mixed |= k != CELL_TET;
  }
  std::vector<int4> tets4((size_t)nc), ext4(mixed ? (size_t)nc : 0);
  const int nverts_of[4] = {4, 8, 6, 5};
  for (int c = 0; c < nc; c++) {
    const int n = nverts_of[types[c]];
    const long long o = f->cell_offsets[c];
    BNX_CHECK(o >= 0 && o + n <= (long long)f->num_indices, BNX_ERR_INVALID,
              "bnx_volume_create_umesh: cell offset out of range");
    const int32_t *ix = f->indices + o;
    int q8[8] = {-1, -1, -1, -1, -1, -1, -1, -1};
    for (int q = 0; q < n; q++) {
      BNX_CHECK(ix[q] >= 0 && ix[q] < nv, BNX_ERR_INVALID, "bnx_volume_create_umesh: vertex index out of range");
      q8[q] = ix[q];
    }
    tets4[c] = make_int4(q8[0], q8[1], q8[2], q8[3]);
    if (mixed) ext4[c] = make_int4(q8[4], q8[5], q8[6], q8[7]);
  }
  std::vector<float4> verts4((size_t)nv);
  for (int i = 0; i < nv; i++)
    verts4[i] = make_float4(f->vertices[3 * (size_t)i], f->vertices[3 * (size_t)i + 1],
                            f->vertices[3 * (size_t)i + 2], f->scalars[i]);
  bnx_volume v = new bnx_volume_s;
  v->dev = d;
  VolumeDev &dd = v->dd;
  dd.field = FIELD_UMESH;
  // computeWorldBounds (UMeshField.cu:304-318): union of the cells' boxes
  float lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
  for (int c = 0; c < nc; c++) {
    const int n = nverts_of[types[c]];
    const int q8[8] = {tets4[c].x, tets4[c].y, tets4[c].z, tets4[c].w,
                       mixed ? ext4[c].x : -1, mixed ? ext4[c].y : -1, mixed ? ext4[c].z : -1, mixed ? ext4[c].w : -1};
    for (int q = 0; q < n; q++) {
      const float4 p = verts4[q8[q]];
      const float pc[3] = {p.x, p.y, p.z};
      for (int k = 0; k < 3; k++) {
        lo[k] = std::min(lo[k], pc[k]);
        hi[k] = std::max(hi[k], pc[k]);
      }
    }
  }
  // buildInitialMacroCells (UMeshField.cu:168-197), host float math
  const float size[3] = {hi[0] - lo[0], hi[1] - lo[1], hi[2] - lo[2]};
  if (auto_mc) {
    const float maxWidth = std::max(std::max(size[0], size[1]), size[2]);
    const int G = 200 + (int)sqrtf((float)nc / 1000.f);
    const float s = (float)(G - 1) / maxWidth;
    for (int k = 0; k < 3; k++) dd.mc[k] = 1 + (int)(size[k] * s);
  } else {
    for (int k = 0; k < 3; k++) dd.mc[k] = f->mc_dims[k];
  }
  for (int k = 0; k < 3; k++) {
    dd.lo[k] = lo[k];
    dd.hi[k] = hi[k];
    dd.mc_origin[k] = lo[k];
    const float spacing = size[k] * (1.f / (float)dd.mc[k]);
    dd.mc_rcp_spacing[k] = 1.f / spacing;
  }
  dd.user_id = user_id;
  int rc;
  if ((rc = set_tf(v, tf)) != BNX_OK ||
      (rc = volume_build_umesh(v, tets4.data(), mixed ? ext4.data() : nullptr, mixed ? types.data() : nullptr, nc,
                               verts4.data(), nv)) != BNX_OK) {
    delete v;
    return rc;
  }
  v->version++;
  *out = v;
  return BNX_OK;
  GUARD_END
}

int bnx_volume_mc_ranges(bnx_volume v, int32_t dims[3], float *lo, float *hi)
{
  BNX_CHECK(v && dims, BNX_ERR_INVALID, "bnx_volume_mc_ranges: null argument");
  for (int k = 0; k < 3; k++) dims[k] = v->dd.mc[k];
  const size_t n = (size_t)dims[0] * dims[1] * dims[2];
  if (lo || hi) {
    std::vector<float2> r(n);
    BNX_HIP(hipMemcpy(r.data(), v->ranges.ptr, sizeof(float2) * n, hipMemcpyDeviceToHost));
    for (size_t i = 0; i < n; i++) {
      if (lo) lo[i] = r[i].x;
      if (hi) hi[i] = r[i].y;
    }
  }
  return BNX_OK;
}

int bnx_isosurface_create(bnx_volume field, float iso_value, const bnx_material *m, int32_t user_id,
                          bnx_isosurface *out)
{
  GUARD_BEGIN
  BNX_CHECK(field && out, BNX_ERR_INVALID, "bnx_isosurface_create: null argument");
  BNX_CHECK(!m || (m->type == BNX_MATERIAL_MATTE && m->opacity >= 1.f), BNX_ERR_UNSUPPORTED,
            "bnx_isosurface_create: only opaque matte materials are provided");
  bnx_isosurface s = new bnx_isosurface_s;
  s->dev = field->dev;
  s->field = field;
  s->iso = iso_value;
  const float def[3] = {0.8f, 0.8f, 0.8f};
  for (int k = 0; k < 3; k++) s->color[k] = m ? m->color[k] : def[k];
  s->user_id = user_id;
  *out = s;
  return BNX_OK;
  GUARD_END
}

int bnx_isosurface_destroy(bnx_isosurface s)
{
  BNX_CHECK(s, BNX_ERR_INVALID, "bnx_isosurface_destroy: null isosurface");
  delete s;
  return BNX_OK;
}

// ------------------------------------------------------------------- world
int bnx_world_add_isosurface(bnx_world w, bnx_isosurface s, int32_t user_inst_id)
{
  GUARD_BEGIN
  BNX_CHECK(w && s, BNX_ERR_INVALID, "bnx_world_add_isosurface: null argument");
  BNX_CHECK(w->dev == s->dev, BNX_ERR_INVALID, "bnx_world_add_isosurface: isosurface lives on another device");
  w->isos.push_back(s);
  w->iso_inst_ids.push_back(user_inst_id);
  w->iso_seen.clear();
  // the world's table now (renders only refresh it in place)
  const IsoDev BNX_G *isos = nullptr;
  int count = 0;
  return world_isos(w, isos, count);
  GUARD_END
}

int bnx_world_add_volume(bnx_world w, bnx_volume v, int32_t user_inst_id)
{
  GUARD_BEGIN
  BNX_CHECK(w && v, BNX_ERR_INVALID, "bnx_world_add_volume: null argument");
  BNX_CHECK(w->dev == v->dev, BNX_ERR_INVALID, "bnx_world_add_volume: volume lives on another device");
  w->volumes.push_back(v);
  w->volume_inst_ids.push_back(user_inst_id);
  w->vol_seen.clear();
  w->iso_seen.clear();                 // isosurface rtc instances follow the volumes
  // the world's tables now (renders only refresh them in place)
  VolumeList vl{};
  int rc = world_volumes(w, vl);
  if (rc != BNX_OK) return rc;
  const IsoDev BNX_G *isos = nullptr;
  int count = 0;
  return world_isos(w, isos, count);
  GUARD_END
}

int bnx_world_create(bnx_device d, bnx_group instances, const bnx_dir_light *lights, int n, bnx_world *out)
{
  GUARD_BEGIN
  BNX_CHECK(d && instances && out && n >= 0 && (n == 0 || lights), BNX_ERR_INVALID, "bnx_world_create: bad argument");
  BNX_CHECK(instances->kind == GROUP_INSTANCES && instances->built, BNX_ERR_STATE,
            "bnx_world_create: needs a built instance group");
  bnx_world w = new bnx_world_s;
  w->dev = d;
  w->instances = instances;
  w->dir_lights.assign(lights, lights + n);
  int rc = upload(d, w->d_dir_lights, w->dir_lights.data(), w->dir_lights.size());
  if (rc != BNX_OK) {
    delete w;
    return rc;
  }
  BNX_HIP(hipStreamSynchronize(d->stream));
  *out = w;
  return BNX_OK;
  GUARD_END
}

int bnx_world_destroy(bnx_world w)
{
  BNX_CHECK(w, BNX_ERR_INVALID, "bnx_world_destroy: null world");
  delete w;
  return BNX_OK;
}

// ------------------------------------------------------------------ camera
int bnx_camera_perspective(const float position[3], const float direction[3], const float up[3], float fovy_deg,
                           bnx_camera *out)
{
  BNX_CHECK(position && direction && up && out, BNX_ERR_INVALID, "bnx_camera_perspective: null argument");
  // PerspectiveCamera::commit (native/Camera.cu:82-101), host float math
  auto nrm = [](const float *v, float *o) {
    float l = 1.f / sqrtf(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
    o[0] = v[0] * l;
    o[1] = v[1] * l;
    o[2] = v[2] * l;
  };
  auto crs = [](const float *a, const float *b, float *o) {
    o[0] = a[1] * b[2] - a[2] * b[1];
    o[1] = a[2] * b[0] - a[0] * b[2];
    o[2] = a[0] * b[1] - a[1] * b[0];
  };
  float d00[3], du[3], dv[3], t[3];
  nrm(direction, d00);
  crs(d00, up, t);
  nrm(t, du);
  crs(du, d00, t);
  nrm(t, dv);
  float half_angle = ((0.5f * fovy_deg) * (float)M_PI) / 180.0f;
  float sc = (float)(1.f / (2.0f * tanf(half_angle)));
  memset(out, 0, sizeof(*out));
  out->type = BNX_CAMERA_PERSPECTIVE;
  for (int k = 0; k < 3; k++) {
    out->dir_00[k] = d00[k] * sc;
    out->dir_du[k] = du[k];
    out->dir_dv[k] = dv[k];
    out->lens_00[k] = position[k];
  }
  out->aperture_radius = 0.f;
  out->focus_distance = 0.f;   // PerspectiveCamera::focusDistance default
  return BNX_OK;
}

namespace {
void cam_normalize(const float *v, float *o)
{
  float l = 1.f / sqrtf(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
  o[0] = v[0] * l;
  o[1] = v[1] * l;
  o[2] = v[2] * l;
}
void cam_cross(const float *a, const float *b, float *o)
{
  o[0] = a[1] * b[2] - a[2] * b[1];
  o[1] = a[2] * b[0] - a[0] * b[2];
  o[2] = a[0] * b[1] - a[1] * b[0];
}
} // namespace

int bnx_camera_orthographic(const float position[3], const float direction[3], const float up[3], float height,
                            float aspect, bnx_camera *out)
{
  BNX_CHECK(position && direction && up && out, BNX_ERR_INVALID, "bnx_camera_orthographic: null argument");
  // OrthographicCamera::commit (native/Camera.cu:168-181), host float math
  float d00[3], du[3], dv[3], t[3];
  cam_normalize(direction, d00);
  cam_cross(d00, up, t);
  cam_normalize(t, du);
  cam_cross(du, d00, t);
  cam_normalize(t, dv);
  memset(out, 0, sizeof(*out));
  out->type = BNX_CAMERA_ORTHOGRAPHIC;
  for (int k = 0; k < 3; k++) {
    out->ortho.dir[k] = d00[k];
    out->ortho.org_00[k] = position[k];
    out->ortho.org_du[k] = du[k];
    out->ortho.org_dv[k] = dv[k];
  }
  out->ortho.height = height;
  out->ortho.aspect = aspect;
  return BNX_OK;
}

int bnx_camera_omni(const float position[3], const float direction[3], const float up[3], bnx_camera *out)
{
  BNX_CHECK(position && direction && up && out, BNX_ERR_INVALID, "bnx_camera_omni: null argument");
  // OmniCamera::commit (native/Camera.cu:230-242): the frame of the
  // environment map's construction (EnvMap.cu:165-180)
  float to_local[9];
  memset(out, 0, sizeof(*out));
  out->type = BNX_CAMERA_OMNI;
  bnx_num::env_frame(direction, up, out->omni.to_world, to_local);
  for (int k = 0; k < 3; k++) out->omni.position[k] = position[k];
  return BNX_OK;
}

// ------------------------------------------------------------------- frame
int bnx_frame_create(bnx_device d, int w, int h, bnx_frame *out)
{
  GUARD_BEGIN
  BNX_CHECK(d && out && w > 0 && h > 0, BNX_ERR_INVALID, "bnx_frame_create: bad argument");
  bnx_frame f = new bnx_frame_s;
  f->dev = d;
  f->w = w;
  f->h = h;
  size_t np = (size_t)w * h;
  hipError_t e;
  if ((e = f->accum.ensure(32 * np)) != hipSuccess || (e = f->depth.ensure(4 * np)) != hipSuccess ||
      (e = f->prim.ensure(4 * np)) != hipSuccess || (e = f->inst.ensure(4 * np)) != hipSuccess ||
      (e = f->obj.ensure(4 * np)) != hipSuccess || (e = f->normal.ensure(16 * np)) != hipSuccess) {
    delete f;
    set_error(std::string("bnx_frame_create: ") + hipGetErrorString(e));
    return BNX_ERR_OOM;
  }
  int rc = frame_clear(f);
  if (rc != BNX_OK) {
    delete f;
    return rc;
  }
  BNX_HIP(hipStreamSynchronize(d->stream));
  *out = f;
  return BNX_OK;
  GUARD_END
}

int bnx_frame_reset(bnx_frame f)
{
  BNX_CHECK(f, BNX_ERR_INVALID, "bnx_frame_reset: null frame");
  return frame_clear(f);
}

int bnx_frame_read(bnx_frame f, int channel, void *host)
{
  BNX_CHECK(f && host, BNX_ERR_INVALID, "bnx_frame_read: null argument");
  return frame_read(f, channel, host);
}

int bnx_frame_set_denoiser(bnx_frame f, bnx_denoise_fn fn, void *user, int32_t fade_out)
{
  BNX_CHECK(f, BNX_ERR_INVALID, "bnx_frame_set_denoiser: null frame");
  f->denoise = fn;
  f->denoise_user = fn ? user : nullptr;
  f->denoise_fade = fn && fade_out != 0;
  if (!fn) {
    f->dn_in.release();
    f->dn_out.release();
    f->dn_normal.release();
  }
  return BNX_OK;
}

int bnx_frame_accum_id(bnx_frame f) { return f ? f->accum_id : -1; }

int bnx_frame_destroy(bnx_frame f)
{
  BNX_CHECK(f, BNX_ERR_INVALID, "bnx_frame_destroy: null frame");
  (void)hipStreamSynchronize(f->dev->stream);
  delete f;
  return BNX_OK;
}

// ------------------------------------------------------------------ render
int bnx_render(bnx_world w, const bnx_camera *cam, const bnx_renderer *r, bnx_frame f, int spp,
               bnx_render_stats *stats)
{
  GUARD_BEGIN
  BNX_CHECK(w && cam && r && f && spp >= 0, BNX_ERR_INVALID, "bnx_render: bad argument");
  BNX_CHECK(f->dev == w->dev, BNX_ERR_INVALID, "bnx_render: frame and world live on different devices");
  BNX_CHECK(w->instances && w->instances->built, BNX_ERR_STATE,
            "bnx_render: the world's instance group is not built (a rebuild failed)");
  BNX_CHECK(cam->type >= BNX_CAMERA_PERSPECTIVE && cam->type <= BNX_CAMERA_OMNI, BNX_ERR_INVALID,
            "bnx_render: unknown camera type");
  return render_frame(w, *cam, *r, f, spp, stats);
  GUARD_END
}

} // extern "C"
