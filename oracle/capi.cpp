// capi.cpp -- C entry points of the CPU oracle, loaded by tests/ and by
// bench.py's cpu_baseline leg through ctypes.
//
// TEST INFRASTRUCTURE ONLY (see oracle.h).
#include "oracle.h"
#include "dda.h"

#include <chrono>
#include <cstdio>

namespace orc {
void camera_perspective(const float *pos, const float *dir, const float *up, float fovy, bnx_camera &out);
void camera_orthographic(const float *pos, const float *dir, const float *up, float height, float aspect,
                         bnx_camera &out);
void camera_omni(const float *pos, const float *dir, const float *up, bnx_camera &out);
float linear_to_srgb(float x);
void primary_rays(const bnx_camera &cam, int W, int H, int accumID, bnx_ray *out);
uint32_t make_8bit(float f);
}

using namespace orc;

struct orc_scene {
  Scene s;
};
struct orc_frame {
  Frame f;
};

static void set_material(Material &m, const bnx_material &d)
{
  m.type = d.type;
  m.color_from_attribute = d.color_from_attribute;
  memcpy(m.color, d.color, sizeof(m.color));
  m.opacity = d.opacity;
  m.metallic = d.metallic;
  m.roughness = d.roughness;
  m.transmission = d.transmission;
  m.ior = d.ior;
}

extern "C" {

orc_scene *orc_scene_create(void) { return new orc_scene; }
void orc_scene_destroy(orc_scene *s) { delete s; }

int orc_scene_add_triangles(orc_scene *sc, const bnx_triangles_desc *d)
{
  Geom g;
  g.type = 0;
  g.user_id = d->user_id;
  set_material(g.mat, d->material);
  for (int i = 0; i < d->num_vertices; i++) g.vertices.push_back(load3(d->vertices + 3 * i));
  g.indices.assign(d->indices, d->indices + 3 * (size_t)d->num_triangles);
  if (d->prim_colors) {
    // the "color" attribute per primitive
    Attribute &a = g.attr[BNX_ATTRIB_SLOT_COLOR];
    a.scope = BNX_SCOPE_PER_PRIM;
    a.comps = 3;
    a.data.assign(d->prim_colors, d->prim_colors + 3 * (size_t)d->num_triangles);
  }
  sc->s.geoms.push_back(std::move(g));
  return (int)sc->s.geoms.size() - 1;
}

int orc_scene_add_spheres(orc_scene *sc, const bnx_spheres_desc *d)
{
  Geom g;
  g.type = 1;
  g.user_id = d->user_id;
  set_material(g.mat, d->material);
  g.default_radius = d->default_radius;
  for (int i = 0; i < d->num_spheres; i++) g.centers.push_back(load3(d->centers + 3 * i));
  if (d->radii) g.radii.assign(d->radii, d->radii + d->num_spheres);
  if (d->colors)
    for (int i = 0; i < d->num_spheres; i++) g.colors.push_back(load3(d->colors + 3 * i));
  sc->s.geoms.push_back(std::move(g));
  return (int)sc->s.geoms.size() - 1;
}

int orc_scene_add_cylinders(orc_scene *sc, const bnx_cylinders_desc *d)
{
  Geom g;
  g.type = 2;
  g.user_id = d->user_id;
  set_material(g.mat, d->material);
  g.default_radius = d->default_radius;
  for (int i = 0; i < d->num_vertices; i++) g.vertices.push_back(load3(d->vertices + 3 * i));
  g.indices.assign(d->indices, d->indices + 2 * (size_t)d->num_cylinders);
  if (d->radii) g.radii.assign(d->radii, d->radii + d->num_cylinders);
  if (d->colors) {
    Attribute &a = g.attr[BNX_ATTRIB_SLOT_COLOR];
    a.scope = BNX_SCOPE_PER_PRIM;
    a.comps = 3;
    a.data.assign(d->colors, d->colors + 3 * (size_t)d->num_cylinders);
  }
  sc->s.geoms.push_back(std::move(g));
  return (int)sc->s.geoms.size() - 1;
}

int orc_scene_add_capsules(orc_scene *sc, const bnx_capsules_desc *d)
{
  Geom g;
  g.type = 3;
  g.user_id = d->user_id;
  set_material(g.mat, d->material);
  for (int i = 0; i < d->num_vertices; i++) {
    g.vertices.push_back(load3(d->vertices + 4 * i));
    g.radii.push_back(d->vertices[4 * i + 3]);
  }
  g.indices.assign(d->indices, d->indices + 2 * (size_t)d->num_capsules);
  if (d->colors) {
    Attribute &a = g.attr[BNX_ATTRIB_SLOT_COLOR];
    a.scope = BNX_SCOPE_PER_PRIM;
    a.comps = 3;
    a.data.assign(d->colors, d->colors + 3 * (size_t)d->num_capsules);
  }
  sc->s.geoms.push_back(std::move(g));
  return (int)sc->s.geoms.size() - 1;
}

int orc_scene_add_cones(orc_scene *sc, const bnx_cones_desc *d)
{
  Geom g;
  g.type = 4;
  g.user_id = d->user_id;
  set_material(g.mat, d->material);
  for (int i = 0; i < d->num_vertices; i++) g.vertices.push_back(load3(d->vertices + 3 * i));
  g.radii.assign(d->radii, d->radii + d->num_vertices);
  // Cones.dev.cu:81-84: without indices cone i joins vertices 2i, 2i + 1
  for (int i = 0; i < d->num_cones; i++) {
    g.indices.push_back(d->indices ? d->indices[2 * i] : 2 * i);
    g.indices.push_back(d->indices ? d->indices[2 * i + 1] : 2 * i + 1);
  }
  if (d->colors) {
    Attribute &a = g.attr[BNX_ATTRIB_SLOT_COLOR];
    a.scope = BNX_SCOPE_PER_PRIM;
    a.comps = 3;
    a.data.assign(d->colors, d->colors + 3 * (size_t)d->num_cones);
  }
  sc->s.geoms.push_back(std::move(g));
  return (int)sc->s.geoms.size() - 1;
}

// GeometryAttributes of geometry 'geom' (render/GeometryAttributes.cpp:11-46
// resolves one scope per attribute); count elements of comps floats
int orc_geom_set_attribute(orc_scene *sc, int geom, int slot, int scope, int comps, const float *data, int64_t count,
                           const float *value)
{
  if (geom < 0 || geom >= (int)sc->s.geoms.size() || slot < 0 || slot >= 6) return -1;
  Attribute &a = sc->s.geoms[geom].attr[slot];
  a = Attribute{};
  a.scope = scope;
  if (scope == BNX_SCOPE_CONSTANT) {
    a.comps = 4;
    memcpy(a.value, value, sizeof(a.value));
  } else if (scope != BNX_SCOPE_NONE) {
    a.comps = comps;
    a.data.assign(data, data + (size_t)comps * count);
  }
  return 0;
}

int orc_geom_set_vertex_normals(orc_scene *sc, int geom, const float *normals, int n)
{
  if (geom < 0 || geom >= (int)sc->s.geoms.size()) return -1;
  Geom &g = sc->s.geoms[geom];
  g.normals.clear();
  for (int i = 0; normals && i < n; i++) g.normals.push_back(load3(normals + 3 * i));
  return 0;
}

// World::DD::instanceAttributes slot (attribute0..3, color): rgba per instance
int orc_scene_set_instance_attribute(orc_scene *sc, int slot, const float *values, int n)
{
  if (slot < 0 || slot >= 5) return -1;
  sc->s.inst_attr[slot].assign(values ? values : nullptr, values ? values + 4 * (size_t)n : nullptr);
  return 0;
}

int orc_scene_add_group(orc_scene *sc, int type, const int *geoms, int n)
{
  Group g;
  g.type = type;
  g.geoms.assign(geoms, geoms + n);
  sc->s.groups.push_back(std::move(g));
  return (int)sc->s.groups.size() - 1;
}

// xfm: 12 floats (vx, vy, vz, p) or NULL for identity; w2o is its inverse
int orc_scene_add_instance(orc_scene *sc, int group, int user_id, const float *o2w, const float *w2o)
{
  Instance in;
  in.group = group;
  in.user_id = user_id;
  if (o2w && w2o) {
    Affine a, b;
    a.vx = load3(o2w); a.vy = load3(o2w + 3); a.vz = load3(o2w + 6); a.p = load3(o2w + 9);
    b.vx = load3(w2o); b.vy = load3(w2o + 3); b.vz = load3(w2o + 6); b.p = load3(w2o + 9);
    a.identity = b.identity = false;
    in.o2w = a;
    in.w2o = b;
  }
  sc->s.insts.push_back(in);
  return (int)sc->s.insts.size() - 1;
}

// structured volume as user instance 'user_inst' (after all surface instances)
int orc_scene_add_volume(orc_scene *sc, const bnx_structured_desc *f, const bnx_transfer_function *tf, int user_id,
                         int user_inst)
{
  Volume v;
  for (int k = 0; k < 3; k++) v.dims[k] = f->dims[k];
  v.cells_per_mc = f->cells_per_mc > 0 ? f->cells_per_mc : 8;
  v.origin = load3(f->grid_origin);
  v.spacing = load3(f->grid_spacing);
  v.scalars.assign(f->scalars, f->scalars + (size_t)f->dims[0] * f->dims[1] * f->dims[2]);
  v.tf.assign(tf->values, tf->values + 4 * (size_t)tf->num_values);
  v.dom0 = tf->domain[0];
  v.dom1 = tf->domain[1];
  v.base_density = tf->base_density;
  v.user_id = user_id;
  v.user_inst = user_inst;
  build_volume(v);
  sc->s.volumes.push_back(std::move(v));
  return (int)sc->s.volumes.size() - 1;
}

// unstructured (tet) volume; rendered == 0 keeps it only as the field of
// isosurfaces (it is not marched as a volume)
int orc_scene_add_umesh(orc_scene *sc, const bnx_umesh_desc *f, const bnx_transfer_function *tf, int user_id,
                        int user_inst, int rendered)
{
  Volume v;
  v.field = FIELD_UMESH;
  v.rendered = rendered != 0;
  v.verts.resize(4 * (size_t)f->num_vertices);
  for (int i = 0; i < f->num_vertices; i++) {
    for (int q = 0; q < 3; q++) v.verts[4 * (size_t)i + q] = f->vertices[3 * (size_t)i + q];
    v.verts[4 * (size_t)i + 3] = f->scalars[i];
  }
  // cell types: the ANARI and VTK codes of tets, hexahedra, prisms and
  // pyramids (UMeshField.h:17-29); NULL: every cell is a tet
  const int nc = f->num_cells;
  std::vector<uint8_t> types((size_t)nc, (uint8_t)CELL_TET);
  bool mixed = false;
  for (int c = 0; f->cell_types && c < nc; c++) {
    const int t = f->cell_types[c];
    const int k = (t == 0 || t == 10) ? CELL_TET : (t == 1 || t == 12) ? CELL_HEX : (t == 2 || t == 13) ? CELL_PRISM
                : (t == 3 || t == 14) ? CELL_PYR : -1;
    if (k < 0) return -1;
    types[c] = (uint8_t)k;
    mixed |= k != CELL_TET;
  }
  v.tets.assign(4 * (size_t)nc, -1);
  if (mixed) v.ext.assign(4 * (size_t)nc, -1);
  for (int c = 0; c < nc; c++) {
    const int n = types[c] == CELL_HEX ? 8 : types[c] == CELL_PRISM ? 6 : types[c] == CELL_PYR ? 5 : 4;
    for (int q = 0; q < n; q++) {
      const int32_t i = f->indices[f->cell_offsets[c] + q];
      if (q < 4) v.tets[4 * (size_t)c + q] = i;
      else v.ext[4 * (size_t)c + q - 4] = i;
    }
  }
  if (mixed) v.types.swap(types);
  for (int k = 0; k < 3; k++) v.mc_req[k] = f->mc_dims[k];
  v.tf.assign(tf->values, tf->values + 4 * (size_t)tf->num_values);
  v.dom0 = tf->domain[0];
  v.dom1 = tf->domain[1];
  v.base_density = tf->base_density;
  v.user_id = user_id;
  v.user_inst = user_inst;
  build_volume(v);
  sc->s.volumes.push_back(std::move(v));
  return (int)sc->s.volumes.size() - 1;
}

// isosurface over the field of volume 'field' as user instance 'user_inst'
int orc_scene_add_isosurface(orc_scene *sc, int field, float iso, const float *color, int user_id, int user_inst)
{
  IsoSurface s;
  s.field = field;
  s.iso = iso;
  for (int k = 0; k < 3; k++) s.color[k] = color[k];
  s.user_id = user_id;
  s.user_inst = user_inst;
  sc->s.isos.push_back(s);
  return (int)sc->s.isos.size() - 1;
}

int orc_volume_mc_ranges(orc_scene *sc, int k, int32_t *dims, float *lo, float *hi)
{
  const Volume &v = sc->s.volumes[k];
  for (int i = 0; i < 3; i++) dims[i] = v.mc[i];
  if (lo) memcpy(lo, v.mc_lo.data(), v.mc_lo.size() * sizeof(float));
  if (hi) memcpy(hi, v.mc_hi.data(), v.mc_hi.size() * sizeof(float));
  return 0;
}

// element BVH of an unstructured volume: info = {num_nodes, num_prims, max_depth}
int orc_volume_ebvh(orc_scene *sc, int k, void *nodes_out, int32_t *order_out, int32_t *info)
{
  const Volume &v = sc->s.volumes[k];
  info[0] = (int32_t)v.ebvh.nodes.size();
  info[1] = (int32_t)v.ebvh.order.size();
  info[2] = v.ebvh.max_depth;
  if (nodes_out) memcpy(nodes_out, v.ebvh.nodes.data(), v.ebvh.nodes.size() * sizeof(Node));
  if (order_out) memcpy(order_out, v.ebvh.order.data(), v.ebvh.order.size() * sizeof(int32_t));
  return 0;
}

// the Newton inversion of one cell (type: CELL_*, verts: x, y, z, scalar per
// vertex in VTK order) at n points: inside[i] and value[i]
void orc_cell_invert_n(int type, const float *verts, int n, const float *P, int32_t *inside, float *value)
{
  for (int i = 0; i < n; i++) {
    float f = NAN;
    inside[i] = cell_invert(type, verts, load3(P + 3 * (size_t)i), f) ? 1 : 0;
    value[i] = f;
  }
}

// point query with the number of element-BVH nodes and tets visited
float orc_umesh_sample(orc_scene *sc, int k, const float *P, uint64_t *visits)
{
  uint64_t n = 0;
  float f = umesh_sample(sc->s.volumes[k], load3(P), &n);
  if (visits) *visits = n;
  return f;
}

// isosurface k along one ray with a parent seed (a, b): t (or -1) and normal
float orc_iso_march(orc_scene *sc, int k, const float *org, const float *dir, float tmax, uint32_t a, uint32_t b,
                    float *normal)
{
  RNGSeed p;
  p.seed(a, b);
  const IsoSurface &iso = sc->s.isos[k];
  float t = -1.f;
  vec3 N;
  if (!iso_march(sc->s.volumes[iso.field], iso, 0u, load3(org), fix_dir(load3(dir)), 0.f, tmax, p, t, N))
    return -1.f;
  normal[0] = N.x;
  normal[1] = N.y;
  normal[2] = N.z;
  return t;
}

int orc_volume_majorants(orc_scene *sc, int k, int32_t *dims, float *maj)
{
  const Volume &v = sc->s.volumes[k];
  for (int i = 0; i < 3; i++) dims[i] = v.mc[i];
  if (maj) memcpy(maj, v.majorant.data(), v.majorant.size() * sizeof(float));
  return 0;
}

float orc_volume_sample(orc_scene *sc, int k, const float *P) { return volume_sample(sc->s.volumes[k], load3(P)); }
void orc_volume_tf_map(orc_scene *sc, int k, float s, float *out) { tf_map(sc->s.volumes[k], s, out); }
float orc_fast_log(float f) { return fast_log(f); }

// dda3 (volume/DDA.h:60-201) of one ray in grid space over a dims grid:
// writes up to max_cells visited cells (x, y, z) and their parametric
// ranges (t0, t1); returns the number of cells the walk visits (all of
// them, even past max_cells)
int orc_dda_cells(const float *dorg, const float *ddir, float dda_tmax, const int32_t *dims, int max_cells,
                  int32_t *cells, float *ts)
{
  const int d[3] = {dims[0], dims[1], dims[2]};
  int n = 0;
  dda3(load3(dorg), load3(ddir), dda_tmax, d, [&](const int c[3], float t0, float t1) -> bool {
    if (n < max_cells) {
      cells[3 * n] = c[0];
      cells[3 * n + 1] = c[1];
      cells[3 * n + 2] = c[2];
      ts[2 * n] = t0;
      ts[2 * n + 1] = t1;
    }
    n++;
    return n < (1 << 20);
  });
  return n;
}

// volume k marched by n rays (bnx_ray layout, from tmin 0 to the ray's
// tmax), ray i with the parent seed (seed_a + i, seed_b): event t, or -1
void orc_volume_march_n(orc_scene *sc, int k, const bnx_ray *rays, int n, uint32_t seed_a, uint32_t seed_b,
                        float *t_out, int threads)
{
  parallel_for(n, threads, [&](int i) {
    RNGSeed p;
    p.seed(seed_a + (uint32_t)i, seed_b);
    float t = -1.f, rgb[3];
    if (!volume_march(sc->s.volumes[k], (uint32_t)(sc->s.insts.size() + k), load3(rays[i].org),
                      fix_dir(load3(rays[i].dir)), 0.f, rays[i].tmax, p, t, rgb, nullptr))
      t = -1.f;
    t_out[i] = t;
  });
}

// ratio-tracking transmittance estimate of each ray through volume k; ray
// i with the parent seed (seed_a + i, seed_b); samples: tentative
// collisions per ray (may be null)
void orc_volume_ratio_n(orc_scene *sc, int k, const bnx_ray *rays, int n, uint32_t seed_a, uint32_t seed_b,
                        float *T_out, uint32_t *samples, int threads)
{
  parallel_for(n, threads, [&](int i) {
    RNGSeed p;
    p.seed(seed_a + (uint32_t)i, seed_b);
    uint64_t ns = 0;
    T_out[i] = ratio_track(sc->s.volumes[k], (uint32_t)(sc->s.insts.size() + k), load3(rays[i].org),
                           fix_dir(load3(rays[i].dir)), 0.f, rays[i].tmax, p, &ns);
    if (samples) samples[i] = (uint32_t)ns;
  });
}

// march one ray through volume k with a parent seed (a, b): event t or -1
float orc_volume_march(orc_scene *sc, int k, const float *org, const float *dir, float tmax, uint32_t a, uint32_t b,
                       float *rgb, uint64_t *samples)
{
  RNGSeed p;
  p.seed(a, b);
  float t = -1.f;
  uint64_t ns = 0;
  if (!volume_march(sc->s.volumes[k], (uint32_t)(sc->s.insts.size() + k), load3(org), fix_dir(load3(dir)), 0.f,
                    tmax, p, t, rgb, &ns))
    t = -1.f;
  if (samples) *samples = ns;
  return t;
}

void orc_scene_set_dir_lights(orc_scene *sc, const bnx_dir_light *l, int n)
{
  sc->s.dir_lights.assign(l, l + n);
}

// HDRI light as bnx_world_set_env_map sets it; radiance NULL removes it
int orc_scene_set_env_map(orc_scene *sc, const bnx_env_map_desc *d)
{
  EnvMap &e = sc->s.env;
  e = EnvMap();
  if (!d || !d->radiance) return 0;
  if (d->width < 1 || d->height < 1) return -1;
  e.present = true;
  e.W = d->width;
  e.H = d->height;
  e.scale = d->scale;
  const size_t n = (size_t)e.W * e.H;
  e.tex.assign(4 * n, 0.f);
  for (size_t i = 0; i < n; i++)
    for (int k = 0; k < 3; k++) e.tex[4 * i + k] = d->radiance[3 * i + k];
  bnx_num::env_frame(d->direction, d->up, e.to_world, e.to_local);
  env_build(e);
  return 0;
}
void orc_env_cdfs(orc_scene *sc, float *cdf_x, float *cdf_y, float *to_world, float *to_local)
{
  const EnvMap &e = sc->s.env;
  memcpy(cdf_x, e.cdf_x.data(), e.cdf_x.size() * 4);
  memcpy(cdf_y, e.cdf_y.data(), e.cdf_y.size() * 4);
  memcpy(to_world, e.to_world, 36);
  memcpy(to_local, e.to_local, 36);
}
void orc_env_tex(orc_scene *sc, const float *uv, int n, float *out)
{
  for (int i = 0; i < n; i++) env_tex(sc->s.env, uv[2 * i], uv[2 * i + 1], out + 4 * i);
}
void orc_env_pdf(orc_scene *sc, const float *dirs, int n, float *out)
{
  for (int i = 0; i < n; i++) out[i] = env_pdf(sc->s.env, load3(dirs + 3 * i));
}
void orc_env_radiance(orc_scene *sc, const float *dirs, int n, float *out)
{
  for (int i = 0; i < n; i++) {
    vec3 r = env_radiance(sc->s.env, load3(dirs + 3 * i));
    out[3 * i] = r.x; out[3 * i + 1] = r.y; out[3 * i + 2] = r.z;
  }
}
// n light samples drawn in sequence from Random2(a, b); also returns the
// sampled uv of each (uv = (ix + r) / W, (iy + r) / H)
void orc_env_sample(orc_scene *sc, uint32_t a, uint32_t b, int n, float *dirs, float *rads, float *pdfs)
{
  Random2 r(a, b);
  for (int i = 0; i < n; i++) {
    vec3 d, rad;
    float pdf;
    env_sample(sc->s.env, r, d, rad, pdf);
    dirs[3 * i] = d.x; dirs[3 * i + 1] = d.y; dirs[3 * i + 2] = d.z;
    rads[3 * i] = rad.x; rads[3 * i + 1] = rad.y; rads[3 * i + 2] = rad.z;
    pdfs[i] = pdf;
  }
}
void orc_env_uv_to_world(orc_scene *sc, const float *uv, int n, float *out)
{
  for (int i = 0; i < n; i++) {
    vec3 d = env_uv_to_world(sc->s.env, uv[2 * i], uv[2 * i + 1]);
    out[3 * i] = d.x; out[3 * i + 1] = d.y; out[3 * i + 2] = d.z;
  }
}

double orc_scene_build(orc_scene *sc)
{
  auto t0 = std::chrono::steady_clock::now();
  build_scene(sc->s);
  return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
}

// mode 0: BVH closest, 1: brute force closest, 2: BVH any hit.
// counters: nodes, tris, spheres, rays.  cut_plane: 4 floats or null.
// closest hits with stochastic opacity: ray i's RNG state seeds[i]; rej[i]
// the closest candidate the opacity test rejected; mode 0 the BVH walk,
// 1 brute force
int orc_trace_opacity(orc_scene *sc, const bnx_ray *rays, const uint64_t *seeds, int n, bnx_hit *hits, bnx_hit *rej,
                      int mode)
{
  auto out = [&](const Hit &h, bnx_hit &o) {
    o.t = h.t;
    o.u = h.prim >= 0 ? h.u : 0.f;
    o.v = h.prim >= 0 ? h.v : 0.f;
    o.prim_id = h.prim;
    o.geom_id = h.prim >= 0 ? h.geom : -1;
    o.inst = h.prim >= 0 ? h.inst : -1;
    o.user_inst_id = h.prim >= 0 ? sc->s.insts[h.inst].user_id : -1;
    o.obj_id = h.prim >= 0 ? sc->s.geoms[sc->s.groups[sc->s.insts[h.inst].group].geoms[h.geom]].user_id : -1;
  };
  for (int i = 0; i < n; i++) {
    const bnx_ray &r = rays[i];
    Hit rj;
    Hit h = mode == 1 ? trace_brute(sc->s, load3(r.org), load3(r.dir), r.tmin, r.tmax, nullptr, seeds + i, &rj)
                      : trace_bvh(sc->s, load3(r.org), load3(r.dir), r.tmin, r.tmax, false, nullptr, nullptr,
                                  seeds + i, &rj);
    out(h, hits[i]);
    out(rj, rej[i]);
  }
  return 0;
}

int orc_trace(orc_scene *sc, const bnx_ray *rays, int n, bnx_hit *hits, int mode, uint64_t *counters, int threads,
              const float *cut_plane)
{
  std::atomic<uint64_t> cn{0}, ct{0}, cs{0};
  parallel_for(n, threads, [&](int i) {
    const bnx_ray &r = rays[i];
    Hit h;
    TraceCounters c;
    if (mode == 1) h = trace_brute(sc->s, load3(r.org), load3(r.dir), r.tmin, r.tmax, cut_plane);
    else h = trace_bvh(sc->s, load3(r.org), load3(r.dir), r.tmin, r.tmax, mode == 2, &c, cut_plane);
    bnx_hit &o = hits[i];
    o.t = h.t;
    o.u = h.prim >= 0 ? h.u : 0.f;
    o.v = h.prim >= 0 ? h.v : 0.f;
    o.prim_id = h.prim;
    o.geom_id = h.prim >= 0 ? h.geom : -1;
    o.inst = h.prim >= 0 ? h.inst : -1;
    o.user_inst_id = h.prim >= 0 ? sc->s.insts[h.inst].user_id : -1;
    o.obj_id = h.prim >= 0 ? sc->s.geoms[sc->s.groups[sc->s.insts[h.inst].group].geoms[h.geom]].user_id : -1;
    cn += c.nodes;
    ct += c.tris;
    cs += c.spheres;
  });
  if (counters) {
    counters[0] = cn;
    counters[1] = ct;
    counters[2] = cs;
    counters[3] = (uint64_t)n;
  }
  return 0;
}

// BVH over raw boxes (n x 6 floats); nodes_out holds max(1, n) nodes.
int orc_bvh_build_boxes(const float *boxes, int n, int max_leaf, void *nodes_out, int32_t *order_out, int32_t *info)
{
  std::vector<Box> b(n);
  for (int i = 0; i < n; i++) b[i] = Box(load3(boxes + 6 * i), load3(boxes + 6 * i + 3));
  Bvh bvh;
  build_bvh(b, max_leaf, bvh);
  memcpy(nodes_out, bvh.nodes.data(), bvh.nodes.size() * sizeof(Node));
  if (n) memcpy(order_out, bvh.order.data(), (size_t)n * sizeof(int32_t));
  info[0] = (int32_t)bvh.nodes.size();
  info[1] = bvh.max_depth;
  return 0;
}

int orc_scene_bvh(orc_scene *sc, int group, void *nodes_out, int32_t *order_out, int32_t *info)
{
  const Bvh &bvh = group < 0 ? sc->s.tlas : sc->s.groups[group].bvh;
  if (nodes_out) memcpy(nodes_out, bvh.nodes.data(), bvh.nodes.size() * sizeof(Node));
  if (order_out && !bvh.order.empty()) memcpy(order_out, bvh.order.data(), bvh.order.size() * sizeof(int32_t));
  info[0] = (int32_t)bvh.nodes.size();
  info[1] = bvh.max_depth;
  info[2] = (int32_t)bvh.order.size();
  return 0;
}

void orc_camera_perspective(const float *pos, const float *dir, const float *up, float fovy, bnx_camera *out)
{
  camera_perspective(pos, dir, up, fovy, *out);
}
void orc_camera_orthographic(const float *pos, const float *dir, const float *up, float height, float aspect,
                             bnx_camera *out)
{
  camera_orthographic(pos, dir, up, height, aspect, *out);
}
void orc_camera_omni(const float *pos, const float *dir, const float *up, bnx_camera *out)
{
  camera_omni(pos, dir, up, *out);
}
void orc_primary_rays(const bnx_camera *cam, int w, int h, int accum_id, bnx_ray *out)
{
  primary_rays(*cam, w, h, accum_id, out);
}

orc_frame *orc_frame_create(int w, int h)
{
  orc_frame *f = new orc_frame;
  f->f.resize(w, h);
  return f;
}
void orc_frame_destroy(orc_frame *f) { delete f; }
void orc_frame_reset(orc_frame *f) { f->f.reset(); }

// stats: rays, shadow rays, generations, volume samples, nodes, tris, spheres.
// flags bit 0: count shadow rays as occlusion queries (RenderStats::shadow_any_hit)
double orc_render(orc_scene *sc, const bnx_camera *cam, const bnx_renderer *r, orc_frame *fb, int spp, int threads,
                  uint64_t *stats, int flags)
{
  RenderStats st;
  st.shadow_any_hit = (flags & 1) != 0;
  auto t0 = std::chrono::steady_clock::now();
  render(sc->s, *cam, *r, fb->f, spp, threads, stats ? &st : nullptr);
  double ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  if (stats) {
    stats[0] = st.rays;
    stats[1] = st.shadow;
    stats[2] = st.generations;
    stats[3] = st.volume_samples;
    stats[4] = st.ctr.nodes;
    stats[5] = st.ctr.tris;
    stats[6] = st.ctr.spheres;
  }
  return ms;
}

// channel layout as bnx_frame_read
int orc_frame_read(orc_frame *fb, int channel, void *host)
{
  const Frame &f = fb->f;
  size_t np = (size_t)f.w * f.h;
  if (channel == BNX_FB_COLOR) {
    float *o = (float *)host;
    float scale = 1.f / (float)f.accum_id;
    for (size_t i = 0; i < 4 * np; i++) o[i] = (float)((double)f.accum[i] * 0x1p-32) * scale;
  } else if (channel == BNX_FB_DEPTH) {
    memcpy(host, f.depth.data(), np * 4);
  } else if (channel == BNX_FB_PRIMID) {
    memcpy(host, f.prim.data(), np * 4);
  } else if (channel == BNX_FB_INSTID) {
    memcpy(host, f.inst.data(), np * 4);
  } else if (channel == BNX_FB_OBJID) {
    memcpy(host, f.obj.data(), np * 4);
  } else if (channel == BNX_FB_NORMAL) {
    memcpy(host, f.normal.data(), np * 12);
  } else if (channel == BNX_FB_COLOR_RGBA8 || channel == BNX_FB_COLOR_SRGB8) {
    // LinearToFixed8 (fb/FrameBuffer.cu:111-130): saturate, sRGB on RGB,
    // make_rgba (common/math.h:103-125)
    uint32_t *o = (uint32_t *)host;
    float scale = 1.f / (float)f.accum_id;
    for (size_t i = 0; i < np; i++) {
      float c[4];
      for (int k = 0; k < 4; k++) {
        c[k] = (float)((double)f.accum[4 * i + k] * 0x1p-32) * scale;
        c[k] = std::min(std::max(c[k], 0.f), 1.f);
        if (k < 3 && channel == BNX_FB_COLOR_SRGB8) c[k] = linear_to_srgb(c[k]);
      }
      o[i] = (make_8bit(c[0]) << 0) + (make_8bit(c[1]) << 8) + (make_8bit(c[2]) << 16) + (make_8bit(c[3]) << 24);
    }
  } else {
    return -1;
  }
  return 0;
}

// --- numeric helpers for the known-answer tests
// one capped cylinder against n rays: ray k is org (3 floats), dir (3
// floats); cyl = v0, v1, radius; out k = t, then the object normal (0 where
// missed); returns the number of hits
int orc_cylinder_intersect_n(const float *rays, int n, const float *cyl, float tmin, float tmax, float *out)
{
  int hits = 0;
  const vec3 v0 = load3(cyl), v1 = load3(cyl + 3);
  for (int k = 0; k < n; k++) {
    float t = 0.f;
    vec3 N(0.f, 0.f, 0.f);
    const bool h = cylinder_intersect(load3(rays + 6 * k), load3(rays + 6 * k + 3), v0, v1, cyl[6], tmin, tmax, t, &N);
    out[4 * k] = h ? t : -1.f;
    out[4 * k + 1] = N.x;
    out[4 * k + 2] = N.y;
    out[4 * k + 3] = N.z;
    hits += h;
  }
  return hits;
}

// one capsule (v0, r0, v1, r1 in cap) against n rays (org, dir): out k = t
// (-1 where missed), the normalized object normal, the object point
int orc_capsule_intersect_n(const float *rays, int n, const float *cap, float tmax, float *out)
{
  int hits = 0;
  for (int k = 0; k < n; k++) {
    float t = 0.f;
    vec3 N(0.f), P(0.f);
    const bool h = capsule_intersect(load3(rays + 6 * k), load3(rays + 6 * k + 3), load3(cap), cap[3], load3(cap + 4),
                                     cap[7], tmax, t, &N, &P);
    float *o = out + 7 * k;
    o[0] = h ? t : -1.f;
    o[1] = N.x; o[2] = N.y; o[3] = N.z;
    o[4] = P.x; o[5] = P.y; o[6] = P.z;
    hits += h;
  }
  return hits;
}

// one cone (p0, ra, p1, rb in cone) against n rays: out k = t (-1 where
// missed), the unnormalized object normal, the attribute weight
int orc_cone_intersect_n(const float *rays, int n, const float *cone, float tmin, float tmax, float *out)
{
  int hits = 0;
  for (int k = 0; k < n; k++) {
    float t = 0.f, lerp = -1.f;
    vec3 N(0.f);
    const bool h = cone_intersect(load3(rays + 6 * k), load3(rays + 6 * k + 3), load3(cone), cone[3],
                                  load3(cone + 4), cone[7], tmin, tmax, t, &N, &lerp);
    float *o = out + 5 * k;
    o[0] = h ? t : -1.f;
    o[1] = N.x; o[2] = N.y; o[3] = N.z;
    o[4] = lerp;
    hits += h;
  }
  return hits;
}

void orc_sincos(float x, float *s, float *c) { bnx_num::sincos(x, *s, *c); }
void orc_atan2_n(const float *y, const float *x, int n, float *out)
{
  for (int i = 0; i < n; i++) out[i] = bnx_num::atan2(y[i], x[i]);
}
void orc_acos_n(const float *x, int n, float *out)
{
  for (int i = 0; i < n; i++) out[i] = bnx_num::acos(x[i]);
}
// NVisii of an AnariPBR material with base color, metallic, roughness, ior
static PbrParams pbr_of(const float *mat4)
{
  Material m;
  m.type = BNX_MATERIAL_PBR;
  m.metallic = mat4[3];
  m.roughness = mat4[4];
  m.ior = mat4[5];
  float base[4] = {mat4[0], mat4[1], mat4[2], 1.f};
  return pbr_params(m, base);
}
// n (Ng, wo, wi) triples -> value (3) and pdf; mat = base rgb, metallic, roughness, ior
void orc_pbr_eval(const float *mat, const float *vecs, int n, float *value, float *pdf)
{
  const PbrParams p = pbr_of(mat);
  for (int i = 0; i < n; i++) {
    const float *v = vecs + 9 * i;
    vec3 f;
    pbr_eval(p, load3(v), load3(v + 3), load3(v + 6), f, pdf[i]);
    value[3 * i] = f.x; value[3 * i + 1] = f.y; value[3 * i + 2] = f.z;
  }
}
void orc_pbr_pdf(const float *mat, const float *vecs, int n, float *pdf)
{
  const PbrParams p = pbr_of(mat);
  for (int i = 0; i < n; i++) {
    const float *v = vecs + 9 * i;
    pdf[i] = pbr_pdf(p, load3(v), load3(v + 3), load3(v + 6));
  }
}
// n scatter samples for one (Ng, wo) from a Random2(a, b) stream: dir (3),
// pdf, f_r (3), lobe flag (1 diffuse)
void orc_pbr_scatter(const float *mat, const float *ng, const float *wo, uint32_t a, uint32_t b, int n,
                     float *dirs, float *pdfs, float *frs, int32_t *diffuse)
{
  const PbrParams p = pbr_of(mat);
  Random2 r(a, b);
  for (int i = 0; i < n; i++) {
    vec3 d(0.f), f(0.f);
    float pdf = 0.f;
    diffuse[i] = pbr_scatter(p, load3(ng), load3(wo), r, d, pdf, f) ? 1 : 0;
    dirs[3 * i] = d.x; dirs[3 * i + 1] = d.y; dirs[3 * i + 2] = d.z;
    frs[3 * i] = f.x; frs[3 * i + 1] = f.y; frs[3 * i + 2] = f.z;
    pdfs[i] = pdf;
  }
}
void orc_env_frame(const float *dir, const float *up, float *to_world, float *to_local)
{
  bnx_num::env_frame(dir, up, to_world, to_local);
}
uint16_t orc_f32_to_f16(float f) { return f32_to_f16(f); }
float orc_f16_to_f32(uint16_t h) { return f16_to_f32(h); }
// RNG: seed(a,b) from the default state, then n outputs of next()
void orc_rng_seq(uint32_t a, uint32_t b, int n, uint32_t *out)
{
  RNGSeed r;
  r.seed(a, b);
  for (int i = 0; i < n; i++) out[i] = r.next();
}
// Random2(parent seeded (a,b), hash) -> n floats; also returns parent state
void orc_random2_seq(uint32_t a, uint32_t b, uint32_t hash, int n, float *out, uint64_t *parent_state)
{
  RNGSeed p;
  p.seed(a, b);
  Random2 r(p, hash);
  for (int i = 0; i < n; i++) out[i] = r();
  *parent_state = p.state;
}

} // extern "C"
