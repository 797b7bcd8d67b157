// render.cpp -- CPU oracle of barney's wavefront frame loop: ray
// generation, hit programs, path-tracing shade and accumulation.
//
// TEST INFRASTRUCTURE ONLY (see oracle.h).
#include "oracle.h"

#include <algorithm>

namespace orc {

namespace {

const float BNX_INF = INFINITY;
enum { MAX_PATH_DEPTH = 10, MAX_DIFFUSE_BOUNCES = 1 };   // shadeRays.cu:17,34
const float CLAMP_F_R = 13.f;                            // shadeRays.cu:27

// Ray + PathState (render/Ray.h:17-93), scalar form
struct PathRay {
  vec3 org, dir;
  float tmax;
  RNGSeed rng;
  bool had_hit = false;
  vec3 P, N;                 // N holds the fp16-rounded normal (vec3h)
  vec3 albedo;               // Lambertian::albedo or SimplePhase::albedo
  int bsdf = 0;              // 1 Lambertian, 2 SimplePhase (volume hit), 3 NVisii (AnariPBR)
  PbrParams pbr;
  bool is_shadow = false, is_specular = false, in_medium = false;
  float miss[4];
  vec3 tp;                   // PathState::throughput, fp16-rounded (vec3h)
  int pixel = 0;
  float mis = 0.f;
  int ndiff = 0;
};

struct HitIDs {
  float depth;
  int prim, inst, obj;
};

// Owl's frame(N) is un-vendored: orthonormal basis of Duff, Burgess,
// Christensen, Hery, Kensler, Liani, Villemin, "Building an Orthonormal
// Basis, Revisited", JCGT 6(1), 2017 (listing 3).
void onb(vec3 n, vec3 &b1, vec3 &b2)
{
  float sign = copysignf(1.0f, n.z);
  float a = -1.0f / (sign + n.z);
  float b = (n.x * n.y) * a;
  b1 = vec3(1.0f + ((sign * n.x) * n.x) * a, sign * b, -sign * n.x);
  b2 = vec3(b, sign + (n.y * n.y) * a, -n.y);
}

// randomDirection (render/DG.h:186-196)
vec3 random_direction(Random2 &rng)
{
  while (true) {
    vec3 v;
    v.x = 1.f - 2.f * rng();
    v.y = 1.f - 2.f * rng();
    v.z = 1.f - 2.f * rng();
    if (dot(v, v) <= 1.f) return normalize(v);
  }
}

struct LightSample {
  vec3 direction;
  float distance;
  vec3 radiance;
  float pdf;
};

// sampleEnvLight (shadeRays.cu:265-291): the HDRI map when there is one,
// else uniform ambient radiance
bool sample_env(LightSample &ls, const EnvMap &env, float ambient, vec3 N, Random2 &random)
{
  if (env.present) {
    env_sample(env, random, ls.direction, ls.radiance, ls.pdf);
    ls.distance = BNX_INF;
    return true;
  }
  ls.direction = random_direction(random);
  ls.distance = BNX_INF;
  if (N != vec3(0.f)) {
    ls.radiance = vec3(TWO_PI * ambient);
    if (dot(ls.direction, N) < 0.f) ls.direction = -ls.direction;
    ls.pdf = ONE_OVER_TWO_PI;
  } else {
    ls.radiance = vec3(FOUR_PI * ambient);
    ls.pdf = ONE_OVER_FOUR_PI;
  }
  return true;
}

// sampleDirLights (shadeRays.cu:138-189), reservoir of 2
bool sample_dir(LightSample &ls, const std::vector<bnx_dir_light> &lights, vec3 N, Random2 &random)
{
  int n = (int)lights.size();
  if (n == 0) return false;
  int lID[2];
  float weights[2];
  float sum = 0.f;
  for (int i = 0; i < 2; i++) {
    lID[i] = std::min((int)(random() * (float)n), n - 1);
    weights[i] = 0.f;
    const bnx_dir_light &L = lights[lID[i]];
    vec3 rad = load3(L.color) * L.radiance;
    vec3 ldir = -load3(L.direction);
    float w = (N == vec3(0.f)) ? 1.f : dot(ldir, N);
    if (w <= 1e-3f) continue;
    w *= reduce_max(rad);
    if (w <= 1e-3f) continue;
    weights[i] = w;
    sum += w;
  }
  if (sum == 0.f) return false;
  float r = random() * sum;
  int i = 0;
  while (i < 2 && r >= weights[i]) {
    r -= weights[i];
    ++i;
  }
  if (i == 2) return false;
  const bnx_dir_light &L = lights[lID[i]];
  ls.direction = -load3(L.direction);
  ls.distance = BNX_INF;
  float denom = (weights[i] / sum) * (2.f / (float)n);
  ls.radiance = ((load3(L.color) * L.radiance) * 1.f) / denom;
  ls.pdf = BNX_INF;
  return weights[i] != 0.f;
}

// sampleLights (shadeRays.cu:293-402); no quad or point lights here
bool sample_lights(LightSample &ls, const Scene &s, const bnx_renderer &r, vec3 N, Random2 &random,
                   bool &needs_mis, bool &is_dir)
{
  needs_mis = false;
  LightSample els, dls;
  float elsW = sample_env(els, s.env, r.ambient_radiance, N, random) ? reduce_max(els.radiance) : 0.f;
  float alsW = 0.f;
  float dlsW = sample_dir(dls, s.dir_lights, N, random) ? reduce_max(dls.radiance) : 0.f;
  float plsW = 0.f;
  float sum = ((alsW + dlsW) + elsW) + plsW;
  if (sum == 0.f) return false;
  elsW *= 1.f / sum;
  alsW *= 1.f / sum;
  dlsW *= 1.f / sum;
  plsW *= 1.f / sum;
  float rr = random();
  if (rr <= alsW) {
    return false;                         // no area lights in these scenes
  } else if (rr <= alsW + elsW) {
    ls = els;
    ls.radiance = ls.radiance * (1.f / elsW);
    needs_mis = true;
  } else if (rr <= (alsW + elsW) + plsW) {
    return false;                         // no point lights
  } else {
    ls = dls;
    ls.radiance = ls.radiance * (1.f / dlsW);
    is_dir = true;
  }
  if (std::isnan(ls.pdf) || ls.pdf <= 0.f) return false;
  return true;
}

// Lambertian (packedBSDF/Lambertian.h:46-85)
struct DGeo {
  vec3 Ng, Ns, wo, P;
};
float lambert_pdf(const DGeo &dg, vec3 wi)
{
  vec3 N = dg.Ng;
  if (dot(wi, N) < 0.f) N = -N;
  float c = fmx(dot(wi, N), 0.f);
  return c * ONE_OVER_PI;
}
void lambert_eval(vec3 albedo, const DGeo &dg, vec3 wi, vec3 &value, float &pdf)
{
  float c = fmx(dot(wi, dg.Ns), 0.f);
  pdf = c * ONE_OVER_PI;
  value = (albedo * ONE_OVER_PI) * c;
}
void lambert_scatter(vec3 albedo, const DGeo &dg, Random2 &random, vec3 &dir, float &pdf, vec3 &f_r)
{
  vec3 N = dg.Ng;
  if (dot(dg.wo, N) < 0.f) N = -N;
  float sx = random();
  float sy = random();
  // cosineSampleHemisphere (render/DG.h:158-164)
  float phi = TWO_PI * sx;
  float cosT = sqrtf(sy);
  float sinT = sqrtf(1.0f - sy);
  float sp, cp;
  bnx_num::sincos(phi, sp, cp);
  vec3 local(cp * sinT, sp * sinT, cosT);
  vec3 b1, b2;
  onb(N, b1, b2);
  dir = (b1 * local.x + b2 * local.y) + N * local.z;
  pdf = local.z * ONE_OVER_PI;
  f_r = albedo;
}

// bounce() (shadeRays.cu:474-880) for surfaces with a Lambertian BSDF
void bounce(const Scene &s, const bnx_renderer &rr, vec3 &frag, PathRay &ray, PathRay &shadow, int depth)
{
  const float EPS = 1e-4f;
  const bool no_hit = !ray.had_hit;
  const vec3 in_tp = ray.tp;
  if (ray.is_shadow) {
    if (no_hit) frag = ray.tp * ray.mis;
    shadow.tmax = -1.f;
    ray.tmax = -1.f;
    return;
  }
  vec3 Ng = ray.N;
  const bool is_vol = (Ng == vec3(0.f));
  if (!is_vol) Ng = normalize(Ng);
  const bool front = dot(ray.dir, Ng) < 0.f;
  vec3 Ngff = front ? Ng : -Ng;
  if (no_hit) {
    // primaryRayMissColor / radianceFromEnv (shadeRays.cu:435-470, 558-571)
    if (depth == 0) frag = s.env.present ? env_radiance(s.env, ray.dir) : vec3(ray.miss[0], ray.miss[1], ray.miss[2]);
    else frag = (ray.tp * (s.env.present ? env_radiance(s.env, ray.dir) : vec3(rr.ambient_radiance))) * ray.mis;
    ray.tmax = -1.f;
    return;
  }
  Random2 random(ray.rng, fbits(ray.tmax));
  DGeo dg;
  dg.P = ray.P;
  dg.Ng = Ng;
  dg.Ns = Ng;
  dg.wo = -normalize(ray.dir);
  const float off_eps = fmx(EPS, 1e-6f * reduce_max(vabs(dg.P)));   // safe_eps (:36-41)
  vec3 ff_off = is_vol ? dg.wo : Ngff;

  LightSample ls;
  bool needs_mis = false, is_dir = false;
  if (sample_lights(ls, s, rr, Ngff, random, needs_mis, is_dir)) {
    vec3 fv;
    float fpdf;
    if (ray.bsdf == 2) {
      // SimplePhase::eval (packedBSDF/Phase.h:147-152)
      fpdf = ONE_OVER_FOUR_PI;
      fv = ONE_OVER_FOUR_PI * ray.albedo;
    } else if (ray.bsdf == 3) {
      pbr_eval(ray.pbr, dg.Ng, dg.wo, ls.direction, fv, fpdf);
    } else {
      lambert_eval(ray.albedo, dg, ls.direction, fv, fpdf);
    }
    if (!(fpdf > 0.f) || reduce_max(fv) < 1e-4f) {
      shadow.tmax = -1.f;
    } else {
      fv = vmin(fv, vec3(CLAMP_F_R));
      float s1 = std::isinf(ls.pdf) ? 1.f : 1.f / ls.pdf;
      float s2 = is_vol ? 1.f : ONE_OVER_PI * fabsf(dot(dg.Ng, ls.direction));
      vec3 tp_sr = (((in_tp * s1) * fv) * ls.radiance) * s2;
      // makeShadowRay (render/Ray.h:146-157)
      shadow.had_hit = false;
      shadow.is_shadow = true;
      shadow.dir = ls.direction;
      shadow.org = dg.P + off_eps * ff_off;
      shadow.tmax = ls.distance * (1.f - 2.f * off_eps);
      shadow.tp = quantize_half(tp_sr);
      shadow.rng = ray.rng;
      ray.rng.next_h((uint64_t)fbits(ray.tmax));
      shadow.pixel = ray.pixel;
      shadow.mis = 1.f;
      if (!is_dir && needs_mis) {
        float pl = env_pdf(s.env, ls.direction);     // EnvMapLight::DD::pdf (light/EnvMap.h:135-160)
        float ps = ray.bsdf == 2   ? ONE_OVER_FOUR_PI
                   : ray.bsdf == 3 ? pbr_pdf(ray.pbr, dg.Ng, dg.wo, ls.direction)
                                   : lambert_pdf(dg, ls.direction);
        shadow.mis = pl / ((pl + ps) + 1e-10f);
        if (shadow.mis < 1e-5f) shadow.tmax = -1.f;
      }
    }
  }
  if (depth >= MAX_PATH_DEPTH) {
    ray.tmax = -1.f;
    return;
  }
  ray.had_hit = false;
  ray.tmax = BNX_INF;
  if (ray.bsdf == 2) {
    // SimplePhase::scatter (Phase.h:155-170) draws two numbers and returns a
    // VOLUME scatter, which SCI_VIS_MODE ends (shadeRays.cu:779-789)
    (void)random();
    (void)random();
    ray.tmax = -1.f;
    return;
  }
  vec3 sdir, f_r;
  float spdf;
  bool diffuse = true;
  if (ray.bsdf == 3) diffuse = pbr_scatter(ray.pbr, dg.Ng, dg.wo, random, sdir, spdf, f_r);
  else lambert_scatter(ray.albedo, dg, random, sdir, spdf, f_r);
  if (!(spdf > 0.f) || spdf <= 1e-6f) return;
  // a DIFFUSE scatter counts against the diffuse bounce limit; GLOSSY does not
  if (diffuse) {
    if (ray.ndiff >= MAX_DIFFUSE_BOUNCES) {
      ray.tmax = -1.f;
      return;
    }
    ray.ndiff = ray.ndiff + 1;
  }
  ray.is_specular = false;
  ray.org = dg.P + ((1.f * off_eps) * ff_off);
  ray.dir = normalize(sdir);
  float cosf_ = is_vol ? 1.f : fabsf(dot(dg.Ng, normalize(ray.dir)));
  float den = std::isinf(spdf) ? 1.f : spdf + 1e-10f;
  vec3 sf = (f_r * cosf_) / den;
  sf = vmin(sf, vec3(2.5f));
  ray.tp = quantize_half(ray.tp * sf);
  if (needs_mis && !std::isinf(spdf)) {
    float ps = ray.bsdf == 3 ? pbr_pdf(ray.pbr, dg.Ng, dg.wo, ray.dir) : lambert_pdf(dg, ray.dir);
    float pl = env_pdf(s.env, ray.dir);
    ray.mis = ps / (ps + pl);
  } else {
    ray.mis = 1.f;
  }
}

// Hit programs: TrianglesPrograms::anyHit (Triangles.dev.cu:19-135) and
// SpheresPrograms::closestHit (Spheres.dev.cu:42-107) ending in
// DeviceMaterial::setHit (material/DeviceMaterial.h:64-72) with AnariMatte.
void apply_hit(const Scene &s, PathRay &r, const Hit &h, HitIDs *ids, bool compat13)
{
  if (h.prim < 0) return;
  const Instance &in = s.insts[h.inst];
  const Group &g = s.groups[in.group];
  const Geom &geom = s.geoms[g.geoms[h.geom]];
  vec3 P, Nw, osP;
  float color0[4] = {0.f, 0.f, 0.f, 1.f};        // HitAttributes default (render/HitAttributes.h:49-58)
  Corners cn;
  if (geom.type == 0) {
    const int32_t *ix = &geom.indices[3 * h.prim];
    vec3 v0 = geom.vertices[ix[0]], v1 = geom.vertices[ix[1]], v2 = geom.vertices[ix[2]];
    vec3 n = normalize(cross(v1 - v0, v2 - v0));
    if (!geom.normals.empty()) {
      // interpolated vertex normals (Triangles.dev.cu:68-75)
      vec3 Ns = (((1.f - h.u) - h.v) * geom.normals[ix[0]] + h.u * geom.normals[ix[1]]) + h.v * geom.normals[ix[2]];
      n = normalize(Ns);
    }
    vec3 osN = normalize(n);
    osP = (((1.f - h.u) - h.v) * v0 + h.u * v1) + h.v * v2;
    cn = tri_corners(ix, h.prim, h.u, h.v);
    // the 'normal' attribute replaces the object normal (Geometry.h:134),
    // transformed to world space unnormalized (Triangles.dev.cu:112-114)
    if (geom.attr[5].scope != BNX_SCOPE_NONE) {
      const float d[4] = {osN.x, osN.y, osN.z, 0.f};
      float an[4];
      hit_attribute(s, geom, 5, h.inst, h.prim, cn, d, an);
      osN = vec3(an[0], an[1], an[2]);
    }
    if (in.o2w.identity) {
      P = osP;
      Nw = osN;
    } else {
      P = xfm_point(in.o2w, osP);
      // normal: transpose of the world-to-object linear part
      const Affine &w = in.w2o;
      Nw = vec3(w.vx.x * osN.x + w.vx.y * osN.y + w.vx.z * osN.z,
                w.vy.x * osN.x + w.vy.y * osN.y + w.vy.z * osN.z,
                w.vz.x * osN.x + w.vz.y * osN.y + w.vz.z * osN.z);
    }
  } else if (geom.type == 2) {
    // CylindersPrograms::intersect's hit attributes (native/geometry/
    // Cylinders.dev.cu:205-236): object normal of the hit, the point moved
    // 1e-4 along it, world normal = normalize(transformNormal(normalize(N)));
    // attributes interpolated between the ends by the clamped position of
    // the point along the axis (:208-222)
    vec3 o = r.org, d = fix_dir(r.dir);
    if (!in.w2o.identity) {
      o = xfm_point(in.w2o, r.org);
      d = fix_dir(xfm_vector(in.w2o, d));
    }
    vec3 v0 = geom.vertices[geom.indices[2 * h.prim]], v1 = geom.vertices[geom.indices[2 * h.prim + 1]];
    float rad = geom.radii.empty() ? geom.default_radius : geom.radii[h.prim];
    float t = h.u;
    vec3 objN(1.f, 0.f, 0.f);
    (void)cylinder_intersect(o, d, v0, v1, rad, 0.f, h.u, t, &objN);
    vec3 objP = o + h.u * d;
    objP = objP + 1e-4f * normalize(objN);
    const vec3 on = normalize(objN);
    osP = objP;
    float lerp_t = dot(objP - v0, v1 - v0) / dot(v1 - v0, v1 - v0);
    lerp_t = fmx(0.f, fmn(1.f, lerp_t));
    cn.n = 2;
    cn.vi[0] = geom.indices[2 * h.prim];
    cn.vi[1] = geom.indices[2 * h.prim + 1];
    cn.fi[0] = 2 * h.prim;
    cn.fi[1] = 2 * h.prim + 1;
    cn.w[0] = 1.f - lerp_t;
    cn.w[1] = lerp_t;
    if (in.o2w.identity) {
      P = objP;
      Nw = normalize(on);
    } else {
      P = xfm_point(in.o2w, objP);
      const Affine &w = in.w2o;
      Nw = normalize(vec3(w.vx.x * on.x + w.vx.y * on.y + w.vx.z * on.z,
                          w.vy.x * on.x + w.vy.y * on.y + w.vy.z * on.z,
                          w.vz.x * on.x + w.vz.y * on.y + w.vz.z * on.z));
    }
  } else if (geom.type == 3 || geom.type == 4) {
    // CapsulesPrograms / ConesPrograms::intersect's hit attributes: the hit
    // found again at the traced distance
    vec3 o = r.org, d = fix_dir(r.dir);
    if (!in.w2o.identity) {
      o = xfm_point(in.w2o, r.org);
      d = fix_dir(xfm_vector(in.w2o, d));
    }
    PairHit ph;
    pair_hit(in, geom, h.prim, o, d, h.u, ph);
    osP = ph.osP;
    P = ph.P;
    Nw = ph.Nw;
    cn = pair_corners(geom, h.prim, ph.lerp);
  } else {
    vec3 o = r.org, d = fix_dir(r.dir);
    if (!in.w2o.identity) {
      o = xfm_point(in.w2o, r.org);
      d = fix_dir(xfm_vector(in.w2o, d));
    }
    vec3 ctr = geom.centers[h.prim];
    float rad = geom.radii.empty() ? geom.default_radius : geom.radii[h.prim];
    float t_move = fmx(0.f, length(ctr - o) - 3.f * rad);
    vec3 om = o + t_move * d;
    vec3 objP = om + h.u * d;
    P = in.o2w.identity ? objP : xfm_point(in.o2w, objP);
    vec3 objN = (objP == ctr) ? vec3(1.f, 0.f, 0.f) : (objP - ctr);
    vec3 wN = in.o2w.identity ? objN : xfm_vector(in.o2w, objN);
    Nw = normalize(wN);
    // the object position moved off the sphere along the unnormalized
    // offset (Spheres.dev.cu:73-78)
    float eps = fmx(1e-6f, 1e-6f * fabsf(rad));
    eps = fmx(eps, 1e-6f * fmx(fmx(fabsf(objP.x), fabsf(objP.y)), fabsf(objP.z)));
    osP = ctr + (rad + eps) * objN;
    if (!geom.colors.empty()) {
      vec3 c = geom.colors[h.prim];
      color0[0] = c.x; color0[1] = c.y; color0[2] = c.z;
    }
    cn.n = 1;
    cn.vi[0] = cn.fi[0] = h.prim;
  }
  const Material &m = geom.mat;
  float base[4];
  material_color(s, geom, h.inst, h.prim, cn, color0, P, osP, base);
  r.had_hit = true;
  r.tmax = h.t;
  r.P = P;
  r.N = quantize_half(Nw);
  if (m.type == BNX_MATERIAL_PBR) {
    r.pbr = pbr_params(m, base);
    r.bsdf = 3;
  } else {
    r.albedo = (0.85f * vec3(base[0], base[1], base[2])) * ONE_OVER_PI;   // AnariMatte.h:58-63
    r.bsdf = 1;
  }
  if (ids) {
    ids->depth = h.t;
    ids->prim = h.prim;
    ids->inst = (geom.type == 1 && compat13) ? 13 : in.user_id;
    ids->obj = geom.user_id;
  }
}

// Ray::setVolumeHit (render/Ray.h:137-144): SimplePhase(0.7 * rgb), N = 0
void apply_volume_hit(const Scene &s, PathRay &r, int k, float t, const float *rgb, HitIDs *ids)
{
  const Volume &v = s.volumes[k];
  vec3 d = fix_dir(r.dir);
  r.had_hit = true;
  r.tmax = t;
  r.P = r.org + t * d;
  r.N = vec3(0.f);
  r.albedo = .7f * vec3(rgb[0], rgb[1], rgb[2]);
  r.bsdf = 2;
  if (ids) {
    ids->depth = t;
    ids->prim = 0;
    ids->inst = v.user_inst;
    ids->obj = v.user_id;
  }
}

// MCIsoSurfaceAccel::isProg hit (volume/MCAccelerator.h:493-547): matte
// material (AnariMatte.h:58-63) at the crossing, normal = the normalized
// gradient; primID 0 and the isosurface's user IDs
void apply_iso_hit(const Scene &s, PathRay &r, int k, float t, vec3 N, HitIDs *ids)
{
  const IsoSurface &iso = s.isos[k];
  vec3 d = fix_dir(r.dir);
  r.had_hit = true;
  r.tmax = t;
  r.P = r.org + t * d;
  r.N = quantize_half(N);
  r.albedo = (0.85f * vec3(iso.color[0], iso.color[1], iso.color[2])) * ONE_OVER_PI;
  r.bsdf = 1;
  if (ids) {
    ids->depth = t;
    ids->prim = 0;
    ids->inst = iso.user_inst;
    ids->obj = iso.user_id;
  }
}

inline void atomic_add64(int64_t *p, int64_t v) { __atomic_fetch_add(p, v, __ATOMIC_RELAXED); }
inline int64_t to_fixed(float f) { return (int64_t)(f * 0x1p32f); }

} // namespace

void Frame::resize(int w_, int h_)
{
  w = w_;
  h = h_;
  accum.assign((size_t)w * h * 4, 0);
  depth.assign((size_t)w * h, BNX_INF);
  prim.assign((size_t)w * h, -1);
  inst.assign((size_t)w * h, -1);
  obj.assign((size_t)w * h, -1);
  normal.assign((size_t)w * h * 3, 0.f);
  accum_id = 0;
}

void Frame::reset() { resize(w, h); }

// the camera part of _generateRays (generateRays.cu:89-146): pixel
// (ix, iy) of a W x H frame at sample accumID; 'rand' is the pixel's
// Random(ix + W accumID, iy + H accumID), already past the pixel jitter
void primary_ray(const bnx_camera &cam, int W, int H, int accumID, float image_u, float image_v, Random2 &rand,
                 vec3 &org, vec3 &dir)
{
  const vec3 dir00 = load3(cam.dir_00), du = load3(cam.dir_du), dv = load3(cam.dir_dv), lens = load3(cam.lens_00);
  const vec3 o00 = load3(cam.ortho.org_00), odu = load3(cam.ortho.org_du), odv = load3(cam.ortho.org_dv);
  float aspect = (float)W / (float)H;
  if (cam.type == BNX_CAMERA_ORTHOGRAPHIC) {
    // generateRays.cu:130-139
    dir = normalize(load3(cam.ortho.dir));
    org = (o00 + (((image_u - .5f) * cam.ortho.aspect) * cam.ortho.height) * odu) +
          ((image_v - .5f) * cam.ortho.height) * odv;
  } else if (cam.type == BNX_CAMERA_OMNI) {
    // uvToWorld (generateRays.cu:24-38, 140-146); not normalized
    const float phi = TWO_PI * image_u;
    const float theta = ONE_PI * image_v;
    float st, ct, sp, cp;
    bnx_num::sincos(theta, st, ct);
    bnx_num::sincos(phi, sp, cp);
    const float *m = cam.omni.to_world;
    dir = vec3((m[0] * (cp * st) + m[3] * (sp * st)) + m[6] * ct,
               (m[1] * (cp * st) + m[4] * (sp * st)) + m[7] * ct,
               (m[2] * (cp * st) + m[5] * (sp * st)) + m[8] * ct);
    org = load3(cam.omni.position);
  } else {
    org = lens;
    vec3 d = (dir00 + ((1.f * aspect) * (image_u - .5f)) * du) + (1.f * (image_v - .5f)) * dv;
    if (cam.aperture_radius > 0.f) {
      // depth of field (generateRays.cu:101-126)
      const vec3 lens_du = normalize(du), lens_dv = normalize(dv);
      const vec3 lens_n = cross(lens_du, lens_dv);
      const vec3 D = normalize(d);
      const vec3 on_plane = D * (cam.focus_distance / fabsf(dot(D, lens_n)));
      float lu = 0.f, lv = 0.f;
      if (accumID != 0) {
        while (true) {
          lu = 2.f * rand() - 1.f;
          lv = 2.f * rand() - 1.f;
          if (lu * lu + lv * lv > 1.f) continue;
          break;
        }
      }
      const vec3 off = (cam.aperture_radius * lu) * lens_du + (cam.aperture_radius * lv) * lens_dv;
      org = org + off;
      dir = normalize(on_plane - off);
    } else {
      dir = normalize(d);
    }
  }
}

// the primary rays of sample accumID of a W x H frame, pixel (ix, iy) at
// ray ix + W iy (bnx_ray layout, tmin 0, tmax 1e30, no cut plane)
void primary_rays(const bnx_camera &cam, int W, int H, int accumID, bnx_ray *out)
{
  for (int iy = 0; iy < H; iy++)
    for (int ix = 0; ix < W; ix++) {
      Random2 rand((uint32_t)(ix + W * accumID), (uint32_t)(iy + H * accumID));
      float pu = (accumID == 0) ? .5f : rand();
      float pv = (accumID == 0) ? .5f : rand();
      vec3 org, dir;
      primary_ray(cam, W, H, accumID, ((float)ix + pu) / (float)W, ((float)iy + pv) / (float)H, rand, org, dir);
      bnx_ray &r = out[ix + (size_t)W * iy];
      r.org[0] = org.x; r.org[1] = org.y; r.org[2] = org.z;
      r.dir[0] = dir.x; r.dir[1] = dir.y; r.dir[2] = dir.z;
      r.tmin = 0.f;
      r.tmax = 1e30f;
    }
}

void render(const Scene &s, const bnx_camera &cam, const bnx_renderer &rr, Frame &fb, int spp, int threads,
            RenderStats *stats)
{
  const int W = fb.w, H = fb.h;
  const bool compat13 = (rr.flags & BNX_RENDER_COMPAT_SPHERE_INSTID13) != 0;
  for (int sidx = 0; sidx < spp; sidx++) {
    const int accumID = fb.accum_id;
    // ---------------------------------------------- _generateRays (generateRays.cu:42-245)
    std::vector<PathRay> queue((size_t)W * H);
    parallel_for(W * H, threads, [&](int pix) {
      int ix = pix % W, iy = pix / W;
      PathRay &ray = queue[pix];
      ray.mis = 0.f;
      ray.pixel = pix;
      Random2 rand((uint32_t)(ix + W * accumID), (uint32_t)(iy + H * accumID));
      ray.rng.seed((uint32_t)(ix + accumID * W), (uint32_t)iy);
      float pu = (accumID == 0) ? .5f : rand();
      float pv = (accumID == 0) ? .5f : rand();
      float image_u = ((float)ix + pu) / (float)W;
      float image_v = ((float)iy + pv) / (float)H;
      vec3 org, dir;
      primary_ray(cam, W, H, accumID, image_u, image_v, rand, org, dir);
      ray.org = org;
      ray.dir = dir;
      ray.had_hit = false;
      ray.is_shadow = false;
      ray.in_medium = false;
      ray.tmax = 1e30f;
      if (rr.cut_plane[3] > -1e28f) {
        vec3 N = load3(rr.cut_plane);
        float dist_org = dot(N, ray.org) + rr.cut_plane[3];
        float denom = dot(N, ray.dir);
        if (dist_org >= 0.f) {
          if (denom < 0.f) ray.tmax = fmn(ray.tmax, -dist_org / denom);
        } else {
          if (denom > 0.f) ray.org = ray.org + (-dist_org / denom) * ray.dir;
          else ray.tmax = 0.f;
        }
      }
      ray.ndiff = 0;
      float t = ((float)iy + .5f) / (float)H;
      if (rr.bg_color[3] >= 0.f) {
        memcpy(ray.miss, rr.bg_color, sizeof(ray.miss));
      } else {
        float a[4] = {0.9f, 0.9f, 0.9f, 1.f}, b[4] = {0.15f, 0.25f, .8f, 1.f};
        for (int k = 0; k < 4; k++) ray.miss[k] = (1.0f - t) * a[k] + t * b[k];
      }
      ray.tp = vec3(1.f);
    });
    for (int gen = 0;; gen++) {
      if (queue.empty()) break;
      if (rr.max_generations > 0 && gen >= rr.max_generations) break;
      const int n = (int)queue.size();
      std::vector<HitIDs> ids(gen == 0 ? n : 0);
      std::vector<TraceCounters> ctrs(threads > 0 ? 1 : 1);
      std::atomic<uint64_t> nn{0}, nt{0}, ns{0}, nsh{0}, nvs{0};
      // ------------------------------------------ trace (traceRays.dev.cu:26-51)
      parallel_for(n, threads, [&](int i) {
        PathRay &r = queue[i];
        TraceCounters c;
        const bool occlusion = r.is_shadow && stats && stats->shadow_any_hit;
        Hit rej;
        Hit h = trace_bvh(s, r.org, r.dir, 0.f, r.tmax, occlusion, stats ? &c : nullptr, rr.cut_plane, &r.rng.state,
                          &rej);
        if (gen == 0) ids[i] = {BNX_INF, -1, -1, -1};
        // isosurfaces, then volumes, after the surface trace, each clipped
        // to the closest hit so far
        float tcur = h.prim >= 0 ? h.t : r.tmax;
        const vec3 fd = fix_dir(r.dir);
        int nvol = 0;
        for (const Volume &v : s.volumes) nvol += v.rendered ? 1 : 0;
        int ik = -1;
        float it = 0.f;
        vec3 iN;
        for (int k = 0; k < (int)s.isos.size(); k++) {
          float t;
          vec3 N;
          const IsoSurface &iso = s.isos[k];
          if (iso_march(s.volumes[iso.field], iso, (uint32_t)(s.insts.size() + nvol + k), r.org, fd, 0.f, tcur,
                        r.rng, t, N)) {
            ik = k;
            it = t;
            iN = N;
            tcur = t;
          }
        }
        int vk = -1;
        float vt = 0.f, vrgb[3] = {0.f, 0.f, 0.f};
        uint64_t vs = 0;
        // ratio tracking (opt-in): a shadow ray's transmittance estimate
        // scales its MIS weight; zero occludes it
        const bool ratio = r.is_shadow && ik < 0 && (rr.flags & BNX_RENDER_RATIO_TRACKING) != 0;
        float T = 1.f;
        // a shadow ray is occluded by its first event: once an isosurface or
        // a volume stopped it, the remaining volumes are not marched (their
        // outcome would not change it; only the work counted differs)
        for (int k = 0, rk = 0; k < (int)s.volumes.size() && !(r.is_shadow && (ik >= 0 || vk >= 0)); k++) {
          if (!s.volumes[k].rendered) continue;
          float t, rgb[3];
          if (ratio) {
            T *= ratio_track(s.volumes[k], (uint32_t)(s.insts.size() + rk), r.org, fd, 0.f, tcur, r.rng, &vs);
            rk++;
            if (T == 0.f) {
              vk = k;
              vt = tcur;
              break;
            }
            continue;
          }
          if (volume_march(s.volumes[k], (uint32_t)(s.insts.size() + rk), r.org, fd, 0.f, tcur, r.rng, t, rgb, &vs)) {
            vk = k;
            vt = t;
            tcur = t;
            memcpy(vrgb, rgb, sizeof(vrgb));
          }
          rk++;
        }
        if (ratio && vk < 0 && T < 1.f) r.mis = r.mis * T;
        if (vk >= 0) apply_volume_hit(s, r, vk, vt, vrgb, gen == 0 ? &ids[i] : nullptr);
        else if (ik >= 0) apply_iso_hit(s, r, ik, it, iN, gen == 0 ? &ids[i] : nullptr);
        else apply_hit(s, r, h, gen == 0 ? &ids[i] : nullptr, compat13);
        if (gen == 0 && rej.prim >= 0 && rej.t < h.t) {
          // the ID channels report the closest candidate whatever its
          // opacity (Triangles.dev.cu:45-60)
          const Instance &in = s.insts[rej.inst];
          ids[i].prim = rej.prim;
          ids[i].inst = in.user_id;
          ids[i].obj = s.geoms[s.groups[in.group].geoms[rej.geom]].user_id;
        }
        if (stats) nvs += vs;
        if (stats) {
          nn += c.nodes;
          nt += c.tris;
          ns += c.spheres;
          if (r.is_shadow) nsh++;
        }
      });
      if (stats) {
        stats->rays += n;
        stats->shadow += nsh;
        stats->generations++;
        stats->ctr.nodes += nn;
        stats->ctr.tris += nt;
        stats->ctr.spheres += ns;
        stats->ctr.rays += n;
        stats->volume_samples += nvs;
      }
      // ------------------------------------------ _shadeRays (shadeRays.cu:883-1041)
      std::vector<PathRay> out((size_t)2 * n);
      std::vector<char> live((size_t)2 * n, 0);
      const float clampMax = 10.f * (float)(1 + accumID);
      parallel_for(n, threads, [&](int i) {
        PathRay ray = queue[i];
        float alpha = (gen == 0) ? (ray.had_hit ? 1.f : ray.miss[3]) : 0.f;
        float inZ = ray.tmax;
        vec3 inN = ray.had_hit ? ray.N : vec3(0.f);
        vec3 frag(0.f);
        PathRay sh;
        sh.tmax = -1.f;
        bounce(s, rr, frag, ray, sh, gen);
        if (sh.tmax > 0.f) { out[2 * i] = sh; live[2 * i] = 1; }
        if (ray.tmax > 0.f) { out[2 * i + 1] = ray; live[2 * i + 1] = 1; }
        frag = vmin(frag, vec3(clampMax));
        int pix = ray.pixel;
        if (gen == 0) {
          if (accumID == 0) {
            fb.depth[pix] = inZ;
            fb.prim[pix] = ids[i].prim;
            fb.inst[pix] = ids[i].inst;
            fb.obj[pix] = ids[i].obj;
            fb.normal[3 * pix + 0] = inN.x;
            fb.normal[3 * pix + 1] = inN.y;
            fb.normal[3 * pix + 2] = inN.z;
          } else if (inZ < fb.depth[pix]) {
            fb.depth[pix] = inZ;
          }
          if (alpha > 0.f) atomic_add64(&fb.accum[4 * (size_t)pix + 3], to_fixed(alpha));
        }
        if (frag.x > 0.f) atomic_add64(&fb.accum[4 * (size_t)pix + 0], to_fixed(frag.x));
        if (frag.y > 0.f) atomic_add64(&fb.accum[4 * (size_t)pix + 1], to_fixed(frag.y));
        if (frag.z > 0.f) atomic_add64(&fb.accum[4 * (size_t)pix + 2], to_fixed(frag.z));
      });
      std::vector<PathRay> next;
      next.reserve(n);
      for (size_t k = 0; k < out.size(); k++)
        if (live[k]) next.push_back(out[k]);
      queue.swap(next);
    }
    fb.accum_id++;
  }
}

// PerspectiveCamera::commit (native/Camera.cu:82-101), fovy in degrees
void camera_perspective(const float *pos, const float *dir, const float *up, float fovy, bnx_camera &out)
{
  vec3 d00 = normalize(load3(dir));
  vec3 ddu = normalize(cross(d00, load3(up)));
  vec3 ddv = normalize(cross(ddu, d00));
  float sc = (float)(1.f / (2.0f * tanf(((0.5f * fovy) * (float)M_PI) / 180.0f)));
  d00 = d00 * sc;
  memset(&out, 0, sizeof(out));
  out.type = BNX_CAMERA_PERSPECTIVE;
  for (int k = 0; k < 3; k++) {
    out.dir_00[k] = d00[k];
    out.dir_du[k] = ddu[k];
    out.dir_dv[k] = ddv[k];
    out.lens_00[k] = pos[k];
  }
  out.aperture_radius = 0.f;
  out.focus_distance = 0.f;
}

// OrthographicCamera::commit (native/Camera.cu:168-181)
void camera_orthographic(const float *pos, const float *dir, const float *up, float height, float aspect,
                         bnx_camera &out)
{
  vec3 d00 = normalize(load3(dir));
  vec3 ddu = normalize(cross(d00, load3(up)));
  vec3 ddv = normalize(cross(ddu, d00));
  memset(&out, 0, sizeof(out));
  out.type = BNX_CAMERA_ORTHOGRAPHIC;
  for (int k = 0; k < 3; k++) {
    out.ortho.dir[k] = d00[k];
    out.ortho.org_00[k] = pos[k];
    out.ortho.org_du[k] = ddu[k];
    out.ortho.org_dv[k] = ddv[k];
  }
  out.ortho.height = height;
  out.ortho.aspect = aspect;
}

// OmniCamera::commit (native/Camera.cu:230-242): toWorld vz = -normalize(up),
// vy = -normalize(cross(vz, direction)), vx = normalize(cross(vy, vz))
void camera_omni(const float *pos, const float *dir, const float *up, bnx_camera &out)
{
  memset(&out, 0, sizeof(out));
  out.type = BNX_CAMERA_OMNI;
  vec3 vz = -normalize(load3(up));
  vec3 vy = -normalize(cross(vz, load3(dir)));
  vec3 vx = normalize(cross(vy, vz));
  for (int k = 0; k < 3; k++) {
    out.omni.to_world[k] = vx[k];
    out.omni.to_world[3 + k] = vy[k];
    out.omni.to_world[6 + k] = vz[k];
    out.omni.position[k] = pos[k];
  }
}

// linear_to_srgb (common/math.h:61-67) and make_8bit (:103-106)
float linear_to_srgb(float x)
{
  if (x <= 0.0031308f) return 12.92f * x;
  return 1.055f * powf(x, 1.f / 2.4f) - 0.055f;
}
uint32_t make_8bit(float f) { return (uint32_t)std::min(255, std::max(0, (int)(f * 256.f))); }

} // namespace orc
