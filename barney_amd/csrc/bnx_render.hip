// bnx_render.hip -- wavefront frame loop on gfx950: ray generation, shade
// (hit programs + path-tracing bounce), accumulation and frame read-back.
//
// Replaces Context::renderTiles (native/Context.cpp:139-182) with
// _generateRays (kernels/generateRays.cu:42-245), _shadeRays
// (kernels/shadeRays.cu:883-1041) and linearizeColorAndNormalKernel
// (fb/TiledFB.cu:12-52).
//
// Layout: the ray queue is structure-of-arrays (16-byte records per
// field), so each wave reads and writes whole cache lines; several samples
// per pixel, and several requests, are in flight in one queue (a run of
// primary rays holds consecutive path indices sample * pixels + pixel).
// The queue is cut into segments of one request each, and every block lies
// inside one segment, so the request's state (scene, lights, camera,
// renderer, frame: a slot table in HBM) is block-uniform.  Shade appends
// each request's shadow and continuation rays to that request's region of
// the next queue with one atomic per block (ballot + popcount prefix inside
// each wave, LDS scan across waves).  The frame accumulates in
// 2^-32 fixed point with 64-bit integer atomics, so a pixel's sum does not
// depend on the order in which rays of different samples or requests
// arrive.
#include "bnx_internal.h"
#include "bnx_render.h"
#include "bnx_volume.h"
#include "bnx_pbr.h"

#include <chrono>

namespace bnx {
namespace {

constexpr float BNX_INF = __builtin_huge_valf();
constexpr int MAX_PATH_DEPTH = 10;      // shadeRays.cu:34
constexpr int MAX_DIFFUSE_BOUNCES = 1;  // shadeRays.cu:17
constexpr float CLAMP_F_R = 13.f;       // shadeRays.cu:27
constexpr float ONE_PI = 3.14159265358979323846f;
constexpr float TWO_PI = 2.f * ONE_PI;
constexpr float FOUR_PI = 4.f * ONE_PI;
constexpr float ONE_OVER_PI = 1.f / ONE_PI;
constexpr float ONE_OVER_TWO_PI = 1.f / TWO_PI;
constexpr float ONE_OVER_FOUR_PI = 1.f / FOUR_PI;

// ---------------------------------------------------------------- raygen
// One thread per (request, sample, pixel) of the batch.  The chunk table
// lists each request's first queue index in ascending order; a thread finds
// its request by binary search over it.
// pixel of the pix-th path of a sample: row-major when tw == 0; otherwise
// bands of th rows, each cut into tiles tw pixels wide (the last band and
// tile of a row may be narrower), a tile's pixels row-major.  A bijection on
// [0, W H): only the order of the primary rays in the queue changes, never a
// pixel's rays or random numbers.
BNX_D void tile_pixel(int pix, int W, int H, int tw, int th, int &ix, int &iy)
{
  if (tw <= 0) {
    ix = pix % W;
    iy = pix / W;
    return;
  }
  const int y0 = (int)(pix / ((long long)th * W)) * th;
  const int bh = min(th, H - y0);
  const long long pb = pix - (long long)y0 * W;
  const int x0 = (int)(pb / (tw * bh)) * tw;
  const int w = min(tw, W - x0);
  const int loc = (int)(pb - (long long)x0 * bh);
  ix = x0 + loc % w;
  iy = y0 + loc / w;
}

// range: q_begin and total from device memory instead of the arguments
__global__ void k_raygen(const RaygenChunk *chunks, int num_chunks, const SlotDev *slots, long long q_begin,
                         long long total, const long long *range, RayQueue q, int tile_w, int tile_h)
{
  if (range) {
    q_begin = range[0];
    total = range[1];
  }
  const long long idx = q_begin + (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= q_begin + total) return;
  int lo = 0, hi = num_chunks - 1;
  while (lo < hi) {
    int mid = (lo + hi + 1) >> 1;
    if (chunks[mid].begin <= idx) lo = mid;
    else hi = mid - 1;
  }
  const RaygenChunk ch = chunks[lo];
  const SlotDev &SL = slots[ch.slot];
  const bnx_camera &cam = SL.cam;
  const bnx_renderer &rr = SL.rr;
  const int W = SL.W, H = SL.H;
  const int np = W * H;
  const long long path = ch.first + (idx - ch.begin);
  const int s = (int)(path / np), pix = (int)(path - (long long)s * np);
  const int accum_base = SL.accum_base;
  int ix, iy;
  tile_pixel(pix, W, H, tile_w, tile_h, ix, iy);
  int accumID = accum_base + s;
  Random2 rnd((uint32_t)(ix + W * accumID), (uint32_t)(iy + H * accumID));
  uint64_t seed = 0xcafef00dd15ea5e5ull;
  rng_seed(seed, (uint32_t)(ix + accumID * W), (uint32_t)iy);
  float pu = (accumID == 0) ? .5f : rnd();
  float pv = (accumID == 0) ? .5f : rnd();
  float image_u = ((float)ix + pu) / (float)W;
  float image_v = ((float)iy + pv) / (float)H;
  float aspect = (float)W / (float)H;
  v3 org, dir;
  if (cam.type == BNX_CAMERA_ORTHOGRAPHIC) {
    // generateRays.cu:130-139
    dir = normalize(ld3(cam.ortho.dir));
    org = (ld3(cam.ortho.org_00) + (((image_u - .5f) * cam.ortho.aspect) * cam.ortho.height) * ld3(cam.ortho.org_du)) +
          ((image_v - .5f) * cam.ortho.height) * ld3(cam.ortho.org_dv);
  } else if (cam.type == BNX_CAMERA_OMNI) {
    // generateRays.cu:24-38, 140-146: uvToWorld, not normalized
    const float phi = TWO_PI * image_u;
    const float theta = ONE_PI * image_v;
    float st, ct, sp, cp;
    bnx_num::sincos(theta, st, ct);
    bnx_num::sincos(phi, sp, cp);
    const float *m = cam.omni.to_world;
    const v3 l{cp * st, sp * st, ct};
    dir = v3{(m[0] * l.x + m[3] * l.y) + m[6] * l.z, (m[1] * l.x + m[4] * l.y) + m[7] * l.z,
             (m[2] * l.x + m[5] * l.y) + m[8] * l.z};
    org = ld3(cam.omni.position);
  } else {
    org = ld3(cam.lens_00);
    v3 d = (ld3(cam.dir_00) + ((1.f * aspect) * (image_u - .5f)) * ld3(cam.dir_du)) +
           (1.f * (image_v - .5f)) * ld3(cam.dir_dv);
    if (cam.aperture_radius > 0.f) {
      // depth of field (generateRays.cu:101-126)
      const v3 lens_du = normalize(ld3(cam.dir_du));
      const v3 lens_dv = normalize(ld3(cam.dir_dv));
      const v3 lens_n = cross(lens_du, lens_dv);
      const v3 D = normalize(d);
      const v3 on_plane = D * (cam.focus_distance / fabsf(dot(D, lens_n)));
      float lu = 0.f, lv = 0.f;
      if (accumID != 0) {
        while (true) {
          lu = 2.f * rnd() - 1.f;
          lv = 2.f * rnd() - 1.f;
          if (lu * lu + lv * lv > 1.f) continue;
          break;
        }
      }
      const v3 off = (cam.aperture_radius * lu) * lens_du + (cam.aperture_radius * lv) * lens_dv;
      org = org + off;
      dir = normalize(on_plane - off);
    } else {
      dir = normalize(d);
    }
  }
  float tmax = 1e30f;
  if (rr.cut_plane[3] > -1e28f) {
    v3 N = ld3(rr.cut_plane);
    float dist_org = dot(N, org) + rr.cut_plane[3];
    float denom = dot(N, dir);
    if (dist_org >= 0.f) {
      if (denom < 0.f) tmax = fmn(tmax, -dist_org / denom);
    } else {
      if (denom > 0.f) org = org + (-dist_org / denom) * dir;
      else tmax = 0.f;
    }
  }
  q.org[idx] = make_float4(org.x, org.y, org.z, tmax);
  q.dir[idx] = make_float4(dir.x, dir.y, dir.z, __uint_as_float(mk_flags(false, s)));
  q.rng[idx] = seed;
  uint32_t one = f2h(1.f);
  q.ps[idx] = make_uint4(one | (one << 16), one, (uint32_t)(ix + iy * W), __float_as_uint(0.f));   // generation 0
}

// ----------------------------------------------------------------- shade
struct SurfHit {
  v3 P, N;          // N: fp16-rounded world normal (zero for a volume event)
  v3 albedo;        // Lambertian albedo or SimplePhase albedo
  bool phase;       // volume event: SimplePhase (packedBSDF/Phase.h:119-172)
  bool pbr;         // AnariPBR surface: NVisii BSDF with parameters pp
  pbr::Params pp;
};

BNX_D void onb(v3 n, v3 &b1, v3 &b2)
{
  // Duff et al., "Building an Orthonormal Basis, Revisited", JCGT 6(1) 2017
  float sign = copysignf(1.0f, n.z);
  float a = -1.0f / (sign + n.z);
  float b = (n.x * n.y) * a;
  b1 = v3{1.0f + ((sign * n.x) * n.x) * a, sign * b, -sign * n.x};
  b2 = v3{b, sign + (n.y * n.y) * a, -n.y};
}

BNX_D v3 random_direction(Random2 &rng)
{
  while (true) {
    v3 v;
    v.x = 1.f - 2.f * rng();
    v.y = 1.f - 2.f * rng();
    v.z = 1.f - 2.f * rng();
    if (dot(v, v) <= 1.f) return normalize(v);
  }
}

struct LightSample {
  v3 direction;
  float distance;
  v3 radiance;
  float pdf;
};

BNX_D bool sample_dir(LightSample &ls, const bnx_dir_light *lights, int n, v3 N, Random2 &random)
{
  if (n == 0) return false;
  int lID[2];
  float weights[2];
  float sum = 0.f;
  for (int i = 0; i < 2; i++) {
    int l = (int)(random() * (float)n);
    lID[i] = l < n - 1 ? l : n - 1;
    weights[i] = 0.f;
    const bnx_dir_light &L = lights[lID[i]];
    v3 rad = ld3(L.color) * L.radiance;
    v3 ldir = -ld3(L.direction);
    float w = (N == splat(0.f)) ? 1.f : dot(ldir, N);
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
  ls.direction = -ld3(L.direction);
  ls.distance = BNX_INF;
  float denom = (weights[i] / sum) * (2.f / (float)n);
  ls.radiance = ((ld3(L.color) * L.radiance) * 1.f) / denom;
  ls.pdf = BNX_INF;
  return weights[i] != 0.f;
}

BNX_D bool sample_lights(LightSample &ls, const SlotDev &P, v3 N, Random2 &random, bool &needs_mis,
                         bool &is_dir)
{
  needs_mis = false;
  LightSample els, dls;
  // sampleEnvLight (shadeRays.cu:265-291): the HDRI map, else ambient
  els.distance = BNX_INF;
  if (P.env.tex) {
    env_sample(P.env, random, els.direction, els.radiance, els.pdf);
  } else {
    els.direction = random_direction(random);
    if (!(N == splat(0.f))) {
      els.radiance = splat(TWO_PI * P.rr.ambient_radiance);
      if (dot(els.direction, N) < 0.f) els.direction = -els.direction;
      els.pdf = ONE_OVER_TWO_PI;
    } else {
      els.radiance = splat(FOUR_PI * P.rr.ambient_radiance);
      els.pdf = ONE_OVER_FOUR_PI;
    }
  }
  float elsW = reduce_max(els.radiance);
  float alsW = 0.f;
  float dlsW = sample_dir(dls, P.dir_lights, P.num_dir_lights, N, random) ? reduce_max(dls.radiance) : 0.f;
  float plsW = 0.f;
  float sum = ((alsW + dlsW) + elsW) + plsW;
  if (sum == 0.f) return false;
  elsW *= 1.f / sum;
  alsW *= 1.f / sum;
  dlsW *= 1.f / sum;
  plsW *= 1.f / sum;
  float rr = random();
  if (rr <= alsW) {
    return false;
  } else if (rr <= alsW + elsW) {
    ls = els;
    ls.radiance = ls.radiance * (1.f / elsW);
    needs_mis = true;
  } else if (rr <= (alsW + elsW) + plsW) {
    return false;
  } else {
    ls = dls;
    ls.radiance = ls.radiance * (1.f / dlsW);
    is_dir = true;
  }
  if (isnan(ls.pdf) || ls.pdf <= 0.f) return false;
  return true;
}

BNX_D float lambert_pdf(v3 Ng, v3 wi)
{
  v3 N = Ng;
  if (dot(wi, N) < 0.f) N = -N;
  float c = fmx(dot(wi, N), 0.f);
  return c * ONE_OVER_PI;
}

// hit programs: Triangles.dev.cu:19-135 / Spheres.dev.cu:42-107 /
// Cylinders.dev.cu:205-247, each with Geometry::DD::setHitAttributes
// (geometry/Geometry.h:94-139) and the material's createBSDF
BNX_D SurfHit eval_hit(const SceneDev &S, int4 hit, float hv, v3 rorg, v3 rdir)
{
  SurfHit sh;
  const InstDev &I = S.insts[hit.z];
  const GeomDev &G = S.geoms[I.geom_base + hit.y];
  const int prim = hit.x;
  const float hu = __int_as_float(hit.w);
  float4 color0 = make_float4(0.f, 0.f, 0.f, 1.f);   // HitAttributes() (render/HitAttributes.h:49-58)
  v3 P, Nw, osP;
  Corners cn;
  if (G.type == 0) {
    const int *ix = G.indices + 3 * prim;
    const float *vx = G.vertices;
    v3 v0 = ld3(vx + 3 * ix[0]), v1 = ld3(vx + 3 * ix[1]), v2 = ld3(vx + 3 * ix[2]);
    v3 n = normalize(cross(v1 - v0, v2 - v0));
    if (G.normals) {
      // interpolated vertex normals (Triangles.dev.cu:68-75)
      const v3 Ns = (((1.f - hu) - hv) * ld3(G.normals + 3 * ix[0]) + hu * ld3(G.normals + 3 * ix[1])) +
                    hv * ld3(G.normals + 3 * ix[2]);
      n = normalize(Ns);
    }
    v3 osN = normalize(n);
    osP = (((1.f - hu) - hv) * v0 + hu * v1) + hv * v2;
    cn = tri_corners(ix, prim, hu, hv);
    // the 'normal' attribute replaces the object normal (Geometry.h:134),
    // which goes to world space unnormalized (Triangles.dev.cu:112-114)
    if (G.attr[5].scope != BNX_SCOPE_NONE) {
      const float4 an = hit_attribute(S, G, 5, hit.z, prim, cn, make_float4(osN.x, osN.y, osN.z, 0.f));
      osN = v3{an.x, an.y, an.z};
    }
    if (I.identity) {
      P = osP;
      Nw = osN;
    } else {
      P = xfm_point_rows(I.o2w, osP);
      // transpose of the world-to-object linear part: columns of the rows
      Nw = v3{(I.w2o[0].x * osN.x + I.w2o[1].x * osN.y) + I.w2o[2].x * osN.z,
              (I.w2o[0].y * osN.x + I.w2o[1].y * osN.y) + I.w2o[2].y * osN.z,
              (I.w2o[0].z * osN.x + I.w2o[1].z * osN.y) + I.w2o[2].z * osN.z};
    }
  } else if (G.type == 2) {
    // CylindersPrograms::intersect's hit attributes (Cylinders.dev.cu:
    // 205-236): the object normal of the hit (front or back, tube or cap),
    // the point moved 1e-4 along it, the world normal normalized twice;
    // attributes interpolated between the two ends by the point's
    // clamped position along the axis
    v3 o = rorg, d = fix_dir(rdir);
    if (!I.identity) {
      o = xfm_point_rows(I.w2o, rorg);
      d = fix_dir(xfm_vector_rows(I.w2o, d));
    }
    const float4 A = G.cyls[2 * prim], B = G.cyls[2 * prim + 1];
    float t = hu;
    v3 objN{1.f, 0.f, 0.f};
    (void)cyl_intersect(o, d, xyz(A), xyz(B), A.w, 0.f, hu, t, &objN);
    v3 objP = o + hu * d;
    objP = objP + 1e-4f * normalize(objN);
    const v3 on = normalize(objN);
    osP = objP;
    const v3 v0 = xyz(A), v1 = xyz(B);
    float lerp_t = dot(objP - v0, v1 - v0) / dot(v1 - v0, v1 - v0);
    lerp_t = fmx(0.f, fmn(1.f, lerp_t));
    cn.n = 2;
    cn.vi[0] = G.indices[2 * prim];
    cn.vi[1] = G.indices[2 * prim + 1];
    cn.fi[0] = 2 * prim;
    cn.fi[1] = 2 * prim + 1;
    cn.w[0] = 1.f - lerp_t;
    cn.w[1] = lerp_t;
    if (I.identity) {
      P = objP;
      Nw = normalize(on);
    } else {
      P = xfm_point_rows(I.o2w, objP);
      Nw = normalize(v3{(I.w2o[0].x * on.x + I.w2o[1].x * on.y) + I.w2o[2].x * on.z,
                        (I.w2o[0].y * on.x + I.w2o[1].y * on.y) + I.w2o[2].y * on.z,
                        (I.w2o[0].z * on.x + I.w2o[1].z * on.y) + I.w2o[2].z * on.z});
    }
  } else if (G.type == 3 || G.type == 4) {
    // CapsulesPrograms / ConesPrograms::intersect's hit attributes: the hit
    // found again from the records at the traced distance
    v3 o = rorg, d = fix_dir(rdir);
    if (!I.identity) {
      o = xfm_point_rows(I.w2o, rorg);
      d = fix_dir(xfm_vector_rows(I.w2o, d));
    }
    PairHit ph;
    if (G.type == 3) capsule_hit(I, G.cyls[2 * prim], G.cyls[2 * prim + 1], o, d, hu, ph);
    else cone_hit(I, G.cyls[2 * prim], G.cyls[2 * prim + 1], o, d, hu, ph);
    osP = ph.osP;
    P = ph.P;
    Nw = ph.Nw;
    cn = pair_corners(G, prim, ph.lerp);
  } else {
    v3 o = rorg, d = fix_dir(rdir);
    if (!I.identity) {
      o = xfm_point_rows(I.w2o, rorg);
      d = fix_dir(xfm_vector_rows(I.w2o, d));
    }
    float4 sp = G.spheres[prim];
    v3 ctr = xyz(sp);
    float rad = sp.w;
    float t_move = fmx(0.f, length(ctr - o) - 3.f * rad);
    v3 om = o + t_move * d;
    v3 objP = om + hu * d;
    P = I.identity ? objP : xfm_point_rows(I.o2w, objP);
    v3 objN = (objP == ctr) ? v3{1.f, 0.f, 0.f} : (objP - ctr);
    v3 wN = I.identity ? objN : xfm_vector_rows(I.o2w, objN);
    Nw = normalize(wN);
    // the object position moved off the sphere (Spheres.dev.cu:73-78),
    // along the unnormalized offset as the reference does
    float eps = fmx(1e-6f, 1e-6f * fabsf(rad));
    eps = fmx(eps, 1e-6f * fmx(fmx(fabsf(objP.x), fabsf(objP.y)), fabsf(objP.z)));
    osP = ctr + (rad + eps) * objN;
    if (G.colors) {
      const float *cp = G.colors + 3 * prim;
      color0 = make_float4(cp[0], cp[1], cp[2], 1.f);
    }
    cn.n = 1;
    cn.vi[0] = cn.fi[0] = prim;
  }
  const float4 base = material_color(S, G, hit.z, prim, cn, color0, P, osP);
  sh.pbr = G.material == BNX_MATERIAL_PBR;
  if (sh.pbr) sh.pp = pbr::params(v3{base.x, base.y, base.z}, G.metallic, G.roughness, G.ior);
  else sh.albedo = (0.85f * v3{base.x, base.y, base.z}) * ONE_OVER_PI;
  sh.P = P;
  sh.N = quant_half(Nw);
  sh.phase = false;
  return sh;
}

// Ray::setVolumeHit (render/Ray.h:137-144): SimplePhase(0.7 * rgb), N = 0
BNX_D SurfHit eval_volume_hit(float t, float4 rgb, v3 rorg, v3 rdir)
{
  SurfHit sh;
  v3 d = fix_dir(rdir);
  sh.P = rorg + t * d;
  sh.N = splat(0.f);
  sh.albedo = .7f * v3{rgb.x, rgb.y, rgb.z};
  sh.phase = true;
  sh.pbr = false;
  return sh;
}

// MCIsoSurfaceAccel::isProg hit (volume/MCAccelerator.h:493-547): opaque
// matte at the crossing, normal = the normalized field gradient (fp16 in the
// ray, like every surface normal)
BNX_D SurfHit eval_iso_hit(float t, float4 n, v3 rorg, v3 rdir, const IsoDev &iso)
{
  SurfHit sh;
  v3 d = fix_dir(rdir);
  sh.P = rorg + t * d;
  sh.N = quant_half(v3{n.x, n.y, n.z});
  sh.albedo = (0.85f * v3{iso.color[0], iso.color[1], iso.color[2]}) * ONE_OVER_PI;
  sh.phase = false;
  sh.pbr = false;
  return sh;
}

// primary-ray miss color (generateRays.cu:205-216): background or gradient
BNX_D float4 miss_color(const SlotDev &P, int pixel)
{
  if (P.rr.bg_color[3] >= 0.f) return make_float4(P.rr.bg_color[0], P.rr.bg_color[1], P.rr.bg_color[2], P.rr.bg_color[3]);
  int iy = pixel / P.W;
  float t = ((float)iy + .5f) / (float)P.H;
  return make_float4((1.0f - t) * 0.9f + t * 0.15f, (1.0f - t) * 0.9f + t * 0.25f, (1.0f - t) * 0.9f + t * .8f,
                     (1.0f - t) * 1.f + t * 1.f);
}

BNX_D int64_t to_fixed(float f) { return (int64_t)(f * 0x1p32f); }
BNX_D void acc_add(unsigned long long *p, float f) { atomicAdd(p, (unsigned long long)to_fixed(f)); }

struct OutRay {
  float4 org, dir;
  uint64_t rng;
  uint4 ps;
};

BNX_D void pack_ps(uint4 &ps, v3 tp, int ndiff, int gen, int pixel, float mis)
{
  ps.x = f2h(tp.x) | (f2h(tp.y) << 16);
  ps.y = f2h(tp.z) | ((uint32_t)ndiff << 16) | ((uint32_t)gen << 24);
  ps.z = (uint32_t)pixel;
  ps.w = __float_as_uint(mis);
}

// Each ray carries its own generation, so one launch can shade rays of
// different requests at different depths of their paths.  Each block lies
// inside one request's segment: the request's parameters are block-uniform
// (scalar registers) and the block appends to that request's output region.
// SHADE_BLOCK is 1024 or 256 threads (a tuning field): a block of 1024 takes
// one atomic per 1024 rays on its request's counter but fits one block per
// CU at this kernel's register count; 256 fits more waves per SIMD.
template <int SHADE_BLOCK>
__global__ __launch_bounds__(SHADE_BLOCK) void k_shade(const SlotDev *__restrict__ slots, SegView sv, RayQueue in,
                                                       RayQueue out, SlotOut so, unsigned long long *stats)
{
  __shared__ int wave_base[SHADE_BLOCK / 64];
  __shared__ int wave_paths[SHADE_BLOCK / 64];
  __shared__ int block_base;
  int i;
  bool valid;
  const Segment sg = segment_of<SHADE_BLOCK>(sv, i, valid);
  const int slot_id = sg.slot;
  const SlotDev &P = slots[slot_id];
  bool w_sh = false, w_ray = false;
  OutRay o_sh, o_ray;
  if (valid) {
    float4 ro = in.org[i], rd = in.dir[i];
    uint64_t rng = in.rng[i];
    uint4 ps = in.ps[i];
    int4 hit = in.hit[i];
    float hv = in.hit_v[i];
    const uint32_t flags = __float_as_uint(rd.w);
    const bool is_shadow = (flags & FLAG_SHADOW) != 0;
    const int sample = flag_sample(flags);
    const SceneDev &S = P.scene;
    const FrameDev &fb = P.fb;
    const int accumID = P.accum_base + sample;
    v3 tp = v3{h2f(ps.x), h2f(ps.x >> 16), h2f(ps.y)};
    int ndiff = (int)((ps.y >> 16) & 0xffu);
    const int gen = (int)(ps.y >> 24);
    const int pixel = (int)ps.z;
    const float mis = __uint_as_float(ps.w);
    v3 rorg = xyz(ro), rdir = xyz(rd);
    float tmax = ro.w;
    const bool had_hit = hit.x >= 0 && !is_shadow;
    const bool shadow_hit = hit.x >= 0;

    SurfHit sh{};
    const bool vol_hit = had_hit && hit.y >= VOLUME_GEOM_BASE && hit.y < ISO_GEOM_BASE;
    const bool iso_hit = had_hit && hit.y >= ISO_GEOM_BASE;
    if (vol_hit) sh = eval_volume_hit(tmax, in.vcol[i], rorg, rdir);
    else if (iso_hit) sh = eval_iso_hit(tmax, in.vcol[i], rorg, rdir, P.isos[hit.y - ISO_GEOM_BASE]);
    else if (had_hit) sh = eval_hit(S, hit, hv, rorg, rdir);

    const float4 mc = miss_color(P, pixel);
    float alpha = (gen == 0) ? (had_hit ? 1.f : mc.w) : 0.f;
    float inZ = tmax;
    v3 inN = had_hit ? sh.N : splat(0.f);
    v3 frag = splat(0.f);

    // ------------------------------------------------ bounce (shadeRays.cu:474-880)
    const float EPS = 1e-4f;
    float sh_tmax = -1.f, ray_tmax = tmax;
    if (is_shadow) {
      if (!shadow_hit) frag = tp * mis;
      ray_tmax = -1.f;
    } else {
      v3 Ng = sh.N;
      const bool is_vol = (Ng == splat(0.f));
      if (!is_vol) Ng = normalize(Ng);
      const bool front = dot(rdir, Ng) < 0.f;
      v3 Ngff = front ? Ng : -Ng;
      if (!had_hit) {
        // primaryRayMissColor / radianceFromEnv (shadeRays.cu:435-470, 558-571)
        const v3 env_l = P.env.tex ? env_radiance(P.env, rdir) : splat(P.rr.ambient_radiance);
        if (gen == 0) {
          frag = P.env.tex ? env_l : xyz(mc);
        } else {
          frag = (tp * env_l) * mis;
        }
        ray_tmax = -1.f;
      } else {
        Random2 random(rng, __float_as_uint(tmax));
        v3 dgP = sh.P;
        v3 wo = -normalize(rdir);
        const float off_eps = fmx(EPS, 1e-6f * reduce_max(vabs(dgP)));
        v3 ff_off = is_vol ? wo : Ngff;
        LightSample ls;
        bool needs_mis = false, is_dir = false;
        if (sample_lights(ls, P, Ngff, random, needs_mis, is_dir)) {
          float fpdf;
          v3 fv;
          if (sh.phase) {
            fpdf = ONE_OVER_FOUR_PI;
            fv = ONE_OVER_FOUR_PI * sh.albedo;
          } else if (sh.pbr) {
            pbr_eval(sh.pp, Ng, wo, ls.direction, fv, fpdf);
          } else {
            float c = fmx(dot(ls.direction, Ng), 0.f);
            fpdf = c * ONE_OVER_PI;
            fv = (sh.albedo * ONE_OVER_PI) * c;
          }
          if (!(fpdf > 0.f) || reduce_max(fv) < 1e-4f) {
            sh_tmax = -1.f;
          } else {
            fv = vmin(fv, splat(CLAMP_F_R));
            float s1 = isinf(ls.pdf) ? 1.f : 1.f / ls.pdf;
            float s2 = is_vol ? 1.f : ONE_OVER_PI * fabsf(dot(Ng, ls.direction));
            v3 tp_sr = (((tp * s1) * fv) * ls.radiance) * s2;
            v3 so = dgP + off_eps * ff_off;
            sh_tmax = ls.distance * (1.f - 2.f * off_eps);
            o_sh.org = make_float4(so.x, so.y, so.z, sh_tmax);
            o_sh.dir = make_float4(ls.direction.x, ls.direction.y, ls.direction.z,
                                   __uint_as_float(mk_flags(true, sample)));
            o_sh.rng = rng;
            rng_next_h(rng, (uint64_t)__float_as_uint(tmax));
            float smis = 1.f;
            if (!is_dir && needs_mis) {
              float pl = env_pdf(P.env, ls.direction);
              float psd = sh.phase ? ONE_OVER_FOUR_PI
                          : sh.pbr ? pbr_pdf(sh.pp, Ng, wo, ls.direction)
                                   : lambert_pdf(Ng, ls.direction);
              smis = pl / ((pl + psd) + 1e-10f);
              if (smis < 1e-5f) sh_tmax = -1.f;
            }
            pack_ps(o_sh.ps, quant_half(tp_sr), 0, gen + 1, pixel, smis);
          }
        }
        if (gen >= MAX_PATH_DEPTH || sh.phase) {
          // SimplePhase scatters VOLUME, which SCI_VIS_MODE ends (shadeRays.cu:779-789)
          ray_tmax = -1.f;
        } else {
          ray_tmax = BNX_INF;
          v3 sdir, f_r;
          float spdf;
          bool diffuse = true;
          if (sh.pbr) {
            diffuse = pbr_scatter(sh.pp, Ng, wo, random, sdir, spdf, f_r);
          } else {
            // Lambertian::scatter (packedBSDF/Lambertian.h:70-85)
            v3 Nsc = Ng;
            if (dot(wo, Nsc) < 0.f) Nsc = -Nsc;
            float sx = random();
            float sy = random();
            float phi = TWO_PI * sx;
            float cosT = sqrtf(sy);
            float sinT = sqrtf(1.0f - sy);
            float sp, cp;
            bnx_num::sincos(phi, sp, cp);
            v3 local = v3{cp * sinT, sp * sinT, cosT};
            v3 b1, b2;
            onb(Nsc, b1, b2);
            sdir = (b1 * local.x + b2 * local.y) + Nsc * local.z;
            spdf = local.z * ONE_OVER_PI;
            f_r = sh.albedo;
          }
          if (!(spdf > 0.f) || spdf <= 1e-6f) {
            // invalid scatter: the ray goes on unchanged (shadeRays.cu:774-775)
            ps.y = (ps.y & 0x00ffffffu) | ((uint32_t)(gen + 1) << 24);
          } else if (diffuse && ndiff >= MAX_DIFFUSE_BOUNCES) {
            // DIFFUSE scatters count against the diffuse bounce limit; GLOSSY ones do not
            ray_tmax = -1.f;
          } else {
            if (diffuse) ndiff = ndiff + 1;
            rorg = dgP + ((1.f * off_eps) * ff_off);
            rdir = normalize(sdir);
            float cf = is_vol ? 1.f : fabsf(dot(Ng, normalize(rdir)));
            float den = isinf(spdf) ? 1.f : spdf + 1e-10f;
            v3 sf = (f_r * cf) / den;
            sf = vmin(sf, splat(2.5f));
            tp = quant_half(tp * sf);
            float nmis;
            if (needs_mis && !isinf(spdf)) {
              float psd = sh.pbr ? pbr_pdf(sh.pp, Ng, wo, rdir) : lambert_pdf(Ng, rdir);
              float pl = env_pdf(P.env, rdir);
              nmis = psd / (psd + pl);
            } else {
              nmis = 1.f;
            }
            pack_ps(ps, tp, ndiff, gen + 1, pixel, nmis);
          }
          if (ray_tmax > 0.f) {
            o_ray.org = make_float4(rorg.x, rorg.y, rorg.z, ray_tmax);
            o_ray.dir = make_float4(rdir.x, rdir.y, rdir.z, __uint_as_float(mk_flags(false, sample)));
            o_ray.rng = rng;
            o_ray.ps = ps;
          }
        }
      }
    }
    w_sh = sh_tmax > 0.f;
    w_ray = ray_tmax > 0.f;
    // renderer's generation limit: rays of generation max_generations are
    // never traced, so they are not written
    if (P.rr.max_generations > 0 && gen + 1 >= P.rr.max_generations) w_sh = w_ray = false;

    // ------------------------------------------------ accumulate (shadeRays.cu:993-1040)
    const float clampMax = 10.f * (float)(1 + accumID);
    frag = vmin(frag, splat(clampMax));
    unsigned long long *acc = fb.accum + 4 * (size_t)pixel;
    if (gen == 0) {
      if (accumID == 0 && !(flags & FLAG_IDS)) {
        fb.prim[pixel] = hit.x >= 0 ? hit.x : -1;
        int inst = -1, obj = -1;
        if (iso_hit) {
          const IsoDev &I = P.isos[hit.y - ISO_GEOM_BASE];
          inst = I.user_inst;
          obj = I.user_id;
        } else if (vol_hit) {
          const VolumeDev &V = P.vols[hit.y - VOLUME_GEOM_BASE];
          inst = V.user_inst;
          obj = V.user_id;
        } else if (hit.x >= 0) {
          const InstDev &I = S.insts[hit.z];
          const GeomDev &G = S.geoms[I.geom_base + hit.y];
          inst = (G.type == 1 && (P.rr.flags & BNX_RENDER_COMPAT_SPHERE_INSTID13)) ? 13 : I.user_id;
          obj = G.user_id;
        }
        fb.inst[pixel] = inst;
        fb.obj[pixel] = obj;
      }
      if (accumID == 0) fb.normal[pixel] = make_float4(inN.x, inN.y, inN.z, 0.f);
      atomicMin(fb.depth + pixel, f2o(inZ));
      if (alpha > 0.f) acc_add(acc + 3, alpha);
    }
    if (frag.x > 0.f) acc_add(acc + 0, frag.x);
    if (frag.y > 0.f) acc_add(acc + 1, frag.y);
    if (frag.z > 0.f) acc_add(acc + 2, frag.z);
  }
  // ------------------------------------------------ compacted append
  // wave ballot + popcount gives each lane its offset inside the wave, the
  // wave totals are scanned in LDS, and one atomic per block reserves the
  // block's range in its request's output region
  const uint64_t b1 = __ballot(w_sh);
  const uint64_t b2 = __ballot(w_ray);
  const uint64_t lt = lanemask_lt();
  const int pre = __popcll(b1 & lt) + __popcll(b2 & lt);
  const int tot = __popcll(b1) + __popcll(b2);
  const int wid = threadIdx.x >> 6;
  if (lane_id() == 0) {
    wave_base[wid] = tot;
    wave_paths[wid] = __popcll(b2);
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    int acc = 0, paths = 0;
    for (int w = 0; w < SHADE_BLOCK / 64; w++) {
      int c = wave_base[w];
      wave_base[w] = acc;
      acc += c;
      paths += wave_paths[w];
    }
    block_base = so.base[slot_id] + (acc > 0 ? atomicAdd(so.cnt + 2 * slot_id, acc) : 0);
    if (paths > 0) atomicAdd(so.cnt + 2 * slot_id + 1, paths);
  }
  __syncthreads();
  int o = block_base + wave_base[wid] + pre;
  if (w_sh) {
    out.org[o] = o_sh.org;
    out.dir[o] = o_sh.dir;
    out.rng[o] = o_sh.rng;
    out.ps[o] = o_sh.ps;
    o++;
  }
  if (w_ray) {
    out.org[o] = o_ray.org;
    out.dir[o] = o_ray.dir;
    out.rng[o] = o_ray.rng;
    out.ps[o] = o_ray.ps;
  }
  if (lane_id() == 0) stat_add(stats, STAT_SHADOW, (uint32_t)__popcll(b1));
}

// block -> segment maps: one thread per block of either launch width
__global__ void k_segment_blocks(SegView sv, int nsegs, int blocks256, int blocks1024, int blocksT)
{
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  const Segment *segs = sv.segs;
  // the pool of segment b starts after its static part: 256 rays (one chunk
  // per wave) for each of its trace blocks
  if (b < nsegs) {
    const int nb = (b + 1 < nsegs ? segs[b + 1].blockT : blocksT) - segs[b].blockT;
    sv.pool[b] = min(segs[b].count, 256 * nb);
  }
  if (b < blocks256) {
    int lo = 0, hi = nsegs - 1;
    while (lo < hi) {
      const int mid = (lo + hi + 1) >> 1;
      if (segs[mid].block256 <= b) lo = mid;
      else hi = mid - 1;
    }
    ((int *)sv.blk256)[b] = lo;
  }
  if (b < blocks1024) {
    int lo = 0, hi = nsegs - 1;
    while (lo < hi) {
      const int mid = (lo + hi + 1) >> 1;
      if (segs[mid].block1024 <= b) lo = mid;
      else hi = mid - 1;
    }
    ((int *)sv.blk1024)[b] = lo;
  }
  if (b < blocksT) {
    int lo = 0, hi = nsegs - 1;
    while (lo < hi) {
      const int mid = (lo + hi + 1) >> 1;
      if (segs[mid].blockT <= b) lo = mid;
      else hi = mid - 1;
    }
    ((int *)sv.blkT)[b] = lo;
  }
}

// One generation's set-up for a device-driven bounce loop (one request,
// one segment, no read-back between generations): the segment takes its
// count from the previous shade's output counter (count0 for the primary
// generation), the counter is cleared for this generation's shade, and the
// count is kept in gen_counts[g] for the host.
__global__ void k_gen_setup(Segment *segs, int *out_cnt, long long *gen_counts, int g, int count0, int refill)
{
  if (threadIdx.x != 0) return;
  const int count = g == 0 ? count0 : out_cnt[0];
  segs[0] = Segment{0, 0, count, 0, 0, 0, refill, 0};
  out_cnt[0] = 0;
  out_cnt[1] = 0;
  gen_counts[g] = count;
}

// the device's constant-rate wall clock, for timing launches inside a
// captured graph (where event timestamps are not available)
__global__ void k_stamp(unsigned long long *t)
{
  if (threadIdx.x == 0) *t = wall_clock64();
}

// ----------------------------------------------------------- frame ops
__global__ void k_frame_clear(FrameDev fb, int np)
{
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= np) return;
  for (int k = 0; k < 4; k++) fb.accum[4 * (size_t)i + k] = 0ull;
  fb.depth[i] = f2o(BNX_INF);
  fb.prim[i] = -1;
  fb.inst[i] = -1;
  fb.obj[i] = -1;
  fb.normal[i] = make_float4(0.f, 0.f, 0.f, 0.f);
}

// linearizeColorAndNormalKernel (fb/TiledFB.cu:12-52): accum * (1/accumID)
__global__ void k_frame_color(FrameDev fb, int np, float scale, float4 *out)
{
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= np) return;
  const unsigned long long *a = fb.accum + 4 * (size_t)i;
  float c[4];
  for (int k = 0; k < 4; k++) c[k] = (float)((double)(long long)a[k] * 0x1p-32) * scale;
  out[i] = make_float4(c[0], c[1], c[2], c[3]);
}

// LinearToFixed8 (fb/FrameBuffer.cu:111-130): saturate, the sRGB curve on
// RGB (common/math.h:61-67; powf of the device library), make_rgba
// (common/math.h:103-125: min(255, max(0, int(256 f))) per channel)
BNX_D uint32_t make_8bit(float f) { return (uint32_t)min(255, max(0, (int)(f * 256.f))); }
BNX_D float linear_to_srgb(float x)
{
  if (x <= 0.0031308f) return 12.92f * x;
  return 1.055f * powf(x, 1.f / 2.4f) - 0.055f;
}
__global__ void k_frame_rgba8(FrameDev fb, int np, float scale, bool srgb, uint32_t *out)
{
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= np) return;
  const unsigned long long *a = fb.accum + 4 * (size_t)i;
  float c[4];
  for (int k = 0; k < 4; k++) {
    c[k] = (float)((double)(long long)a[k] * 0x1p-32) * scale;
    c[k] = fminf(fmaxf(c[k], 0.f), 1.f);
    if (srgb && k < 3) c[k] = linear_to_srgb(c[k]);
  }
  out[i] = (make_8bit(c[0]) << 0) + (make_8bit(c[1]) << 8) + (make_8bit(c[2]) << 16) + (make_8bit(c[3]) << 24);
}

// 8-bit RGBA of linear float4 colours (a denoiser's output)
__global__ void k_rgba8_from(const float4 *in, int np, bool srgb, uint32_t *out)
{
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= np) return;
  const float4 v = in[i];
  float c[4] = {v.x, v.y, v.z, v.w};
  for (int k = 0; k < 4; k++) {
    c[k] = fminf(fmaxf(c[k], 0.f), 1.f);
    if (srgb && k < 3) c[k] = linear_to_srgb(c[k]);
  }
  out[i] = (make_8bit(c[0]) << 0) + (make_8bit(c[1]) << 8) + (make_8bit(c[2]) << 16) + (make_8bit(c[3]) << 24);
}

// the denoiser's inputs (finalizeFrame's gather, fb/FrameBuffer.cu:196-240):
// linear colour to in_rgba and out_rgba, normals to in_normal (float3)
__global__ void k_denoise_inputs(FrameDev fb, int np, float scale, float4 *in_rgba, float4 *out_rgba, float *in_normal)
{
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= np) return;
  const unsigned long long *a = fb.accum + 4 * (size_t)i;
  float c[4];
  for (int k = 0; k < 4; k++) c[k] = (float)((double)(long long)a[k] * 0x1p-32) * scale;
  const float4 v = make_float4(c[0], c[1], c[2], c[3]);
  in_rgba[i] = v;
  out_rgba[i] = v;
  const float4 n = fb.normal[i];
  in_normal[3 * (size_t)i + 0] = n.x;
  in_normal[3 * (size_t)i + 1] = n.y;
  in_normal[3 * (size_t)i + 2] = n.z;
}

__global__ void k_frame_depth(FrameDev fb, int np, float *out)
{
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < np) out[i] = o2f(fb.depth[i]);
}

inline int nblk(long n, int b) { return (int)((n + b - 1) / b); }

} // namespace

int frame_clear(Frame *f)
{
  int np = f->w * f->h;
  if (np > 0) k_frame_clear<<<nblk(np, 256), 256, 0, f->dev->stream>>>(frame_dev(f), np);
  BNX_HIP(hipGetLastError());
  f->accum_id = 0;
  return BNX_OK;
}

// run the frame's denoiser hook: its output colour in f->dn_out
int frame_denoise(Frame *f)
{
  const int np = f->w * f->h;
  hipStream_t st = f->dev->stream;
  BNX_HIP(f->dn_in.ensure(sizeof(float4) * (size_t)np));
  BNX_HIP(f->dn_out.ensure(sizeof(float4) * (size_t)np));
  BNX_HIP(f->dn_normal.ensure(sizeof(float) * 3 * (size_t)np));
  const float scale = f->accum_id > 0 ? 1.f / (float)f->accum_id : 0.f;
  k_denoise_inputs<<<nblk(np, 256), 256, 0, st>>>(frame_dev(f), np, scale, f->dn_in.as<float4>(),
                                                  f->dn_out.as<float4>(), f->dn_normal.as<float>());
  BNX_HIP(hipGetLastError());
  // readColorChannel (fb/FrameBuffer.cu:257): fade the denoiser out as
  // samples accumulate
  const float blend = f->denoise_fade ? (float)(f->accum_id - 1) / ((float)f->accum_id + 100.f) : 0.f;
  const int rc = f->denoise(f->denoise_user, f->dn_in.as<float>(), f->dn_normal.as<float>(), f->dn_out.as<float>(),
                            f->w, f->h, blend, (void *)st);
  BNX_CHECK(rc == 0, BNX_ERR_INVALID, "bnx_frame_read: the denoiser hook failed (" + std::to_string(rc) + ")");
  return BNX_OK;
}

int frame_read(Frame *f, int channel, void *host)
{
  int np = f->w * f->h;
  hipStream_t st = f->dev->stream;
  FrameDev fd = frame_dev(f);
  if (np == 0) return BNX_OK;
  const bool color = channel == BNX_FB_COLOR || channel == BNX_FB_COLOR_RGBA8 || channel == BNX_FB_COLOR_SRGB8;
  if (color && f->denoise) {
    int rc = frame_denoise(f);
    if (rc != BNX_OK) return rc;
    if (channel == BNX_FB_COLOR) {
      BNX_HIP(hipMemcpyAsync(host, f->dn_out.ptr, sizeof(float4) * (size_t)np, hipMemcpyDeviceToHost, st));
    } else {
      BNX_HIP(f->out.ensure(sizeof(uint32_t) * (size_t)np));
      k_rgba8_from<<<nblk(np, 256), 256, 0, st>>>(f->dn_out.as<float4>(), np, channel == BNX_FB_COLOR_SRGB8,
                                                  f->out.as<uint32_t>());
      BNX_HIP(hipMemcpyAsync(host, f->out.ptr, sizeof(uint32_t) * (size_t)np, hipMemcpyDeviceToHost, st));
    }
    BNX_HIP(hipStreamSynchronize(st));
    return BNX_OK;
  }
  switch (channel) {
  case BNX_FB_COLOR: {
    BNX_HIP(f->out.ensure(sizeof(float4) * (size_t)np));
    float scale = f->accum_id > 0 ? 1.f / (float)f->accum_id : 0.f;
    k_frame_color<<<nblk(np, 256), 256, 0, st>>>(fd, np, scale, f->out.as<float4>());
    BNX_HIP(hipMemcpyAsync(host, f->out.ptr, sizeof(float4) * (size_t)np, hipMemcpyDeviceToHost, st));
    break;
  }
  case BNX_FB_COLOR_RGBA8:
  case BNX_FB_COLOR_SRGB8: {
    BNX_HIP(f->out.ensure(sizeof(uint32_t) * (size_t)np));
    float scale = f->accum_id > 0 ? 1.f / (float)f->accum_id : 0.f;
    k_frame_rgba8<<<nblk(np, 256), 256, 0, st>>>(fd, np, scale, channel == BNX_FB_COLOR_SRGB8,
                                                 f->out.as<uint32_t>());
    BNX_HIP(hipMemcpyAsync(host, f->out.ptr, sizeof(uint32_t) * (size_t)np, hipMemcpyDeviceToHost, st));
    break;
  }
  case BNX_FB_DEPTH:
    BNX_HIP(f->out.ensure(sizeof(float) * (size_t)np));
    k_frame_depth<<<nblk(np, 256), 256, 0, st>>>(fd, np, f->out.as<float>());
    BNX_HIP(hipMemcpyAsync(host, f->out.ptr, sizeof(float) * (size_t)np, hipMemcpyDeviceToHost, st));
    break;
  case BNX_FB_PRIMID:
    BNX_HIP(hipMemcpyAsync(host, fd.prim, 4 * (size_t)np, hipMemcpyDeviceToHost, st));
    break;
  case BNX_FB_INSTID:
    BNX_HIP(hipMemcpyAsync(host, fd.inst, 4 * (size_t)np, hipMemcpyDeviceToHost, st));
    break;
  case BNX_FB_OBJID:
    BNX_HIP(hipMemcpyAsync(host, fd.obj, 4 * (size_t)np, hipMemcpyDeviceToHost, st));
    break;
  case BNX_FB_NORMAL: {
    std::vector<float4> tmp((size_t)np);
    BNX_HIP(hipMemcpyAsync(tmp.data(), fd.normal, sizeof(float4) * (size_t)np, hipMemcpyDeviceToHost, st));
    BNX_HIP(hipStreamSynchronize(st));
    float *o = (float *)host;
    for (int i = 0; i < np; i++) {
      o[3 * i + 0] = tmp[i].x;
      o[3 * i + 1] = tmp[i].y;
      o[3 * i + 2] = tmp[i].z;
    }
    return BNX_OK;
  }
  default:
    set_error("bnx_frame_read: unknown channel");
    return BNX_ERR_INVALID;
  }
  BNX_HIP(hipStreamSynchronize(st));
  return BNX_OK;
}

int launch_raygen(Device *dev, const RaygenChunk *chunks, int num_chunks, const SlotDev *slots, long long q_begin,
                  long long total, const RayQueue &q)
{
  if (total > 0 && num_chunks > 0)
    k_raygen<<<nblk(total, 256), 256, 0, dev->stream>>>(chunks, num_chunks, slots, q_begin, total, nullptr, q,
                                                         dev->raygen_tile_w, dev->raygen_tile_h);
  BNX_HIP(hipGetLastError());
  return BNX_OK;
}

int launch_raygen_dev(Device *dev, const RaygenChunk *chunks, int num_chunks, const SlotDev *slots,
                      const long long *range, long long max_total, const RayQueue &q)
{
  if (max_total > 0 && num_chunks > 0)
    k_raygen<<<nblk(max_total, 256), 256, 0, dev->stream>>>(chunks, num_chunks, slots, 0, 0, range, q,
                                                             dev->raygen_tile_w, dev->raygen_tile_h);
  BNX_HIP(hipGetLastError());
  return BNX_OK;
}

int launch_shade(Device *dev, const SlotDev *slots, const SegView &sv, int blocks256, int blocks1024,
                 const RayQueue &in, const RayQueue &out, const SlotOut &so, unsigned long long *stats)
{
  if (dev->shade_block == 256) {
    if (blocks256 > 0) k_shade<256><<<blocks256, 256, 0, dev->stream>>>(slots, sv, in, out, so, stats);
  } else if (blocks1024 > 0) {
    k_shade<1024><<<blocks1024, 1024, 0, dev->stream>>>(slots, sv, in, out, so, stats);
  }
  BNX_HIP(hipGetLastError());
  return BNX_OK;
}

int launch_gen_setup(Device *dev, Segment *segs, int *out_cnt, long long *gen_counts, int g, int count0, int refill)
{
  k_gen_setup<<<1, 64, 0, dev->stream>>>(segs, out_cnt, gen_counts, g, count0, refill);
  BNX_HIP(hipGetLastError());
  return BNX_OK;
}

int launch_stamp(Device *dev, unsigned long long *t)
{
  k_stamp<<<1, 64, 0, dev->stream>>>(t);
  BNX_HIP(hipGetLastError());
  return BNX_OK;
}

int launch_segment_blocks(Device *dev, const SegView &sv, int nsegs, int blocks256, int blocks1024, int blocksT)
{
  const int n = std::max(std::max(blocks256, blocksT), nsegs);
  if (nsegs > 0 && n > 0)
    k_segment_blocks<<<nblk(n, 256), 256, 0, dev->stream>>>(sv, nsegs, blocks256, blocks1024, blocksT);
  BNX_HIP(hipGetLastError());
  return BNX_OK;
}

} // namespace bnx
