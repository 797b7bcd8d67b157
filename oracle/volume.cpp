// volume.cpp -- CPU oracle of barney's volume path: structured-field
// sampling, transfer function, macrocell majorants, and the DDA + Woodcock
// (delta) tracking that volumes of every field type share.
//
// TEST INFRASTRUCTURE ONLY (see oracle.h).
#include "oracle.h"
#include "dda.h"

#include <algorithm>

namespace orc {

// CPU texture addressing for LINEAR + CLAMP, unnormalized coordinates
// (rtcore/cpu/Texture.cpp:324-336)
static void lerp_coord(int N, float tc, int &i0, int &i1, float &f)
{
  if (tc - .5f <= 0.f) { i0 = i1 = 0; f = 0.f; return; }
  if (tc - .5f >= (float)(N - 1)) { i0 = i1 = N - 1; f = 0.f; return; }
  f = (tc - .5f) - (float)(int)(tc - .5f);
  i0 = (int)(tc - .5f);
  i1 = i0 + 1;
}

static inline float lerp_l(float f, float a, float b) { return (1.f - f) * a + f * b; }

// StructuredDataSampler::DD::sample (volume/StructuredData.h:101-112) with
// the CPU tex3D trilinear filter (rtcore/cpu/Texture.cpp:339-412)
float volume_sample(const Volume &v, vec3 P)
{
  if (v.field == FIELD_UMESH) return umesh_sample(v, P);
  vec3 rel = (P - v.origin) * vec3(1.f / v.spacing.x, 1.f / v.spacing.y, 1.f / v.spacing.z);
  if (rel.x < 0.f || rel.y < 0.f || rel.z < 0.f) return NAN;
  if (rel.x >= (float)v.ncells[0] || rel.y >= (float)v.ncells[1] || rel.z >= (float)v.ncells[2]) return NAN;
  int x0, x1, y0, y1, z0, z1;
  float fx, fy, fz;
  lerp_coord(v.dims[0], rel.x + .5f, x0, x1, fx);
  lerp_coord(v.dims[1], rel.y + .5f, y0, y1, fy);
  lerp_coord(v.dims[2], rel.z + .5f, z0, z1, fz);
  const int Nx = v.dims[0], Ny = v.dims[1];
  auto at = [&](int x, int y, int z) { return v.scalars[(size_t)x + (size_t)Nx * ((size_t)y + (size_t)Ny * z)]; };
  float v00 = lerp_l(fx, at(x0, y0, z0), at(x1, y0, z0));
  float v01 = lerp_l(fx, at(x0, y1, z0), at(x1, y1, z0));
  float v10 = lerp_l(fx, at(x0, y0, z1), at(x1, y0, z1));
  float v11 = lerp_l(fx, at(x0, y1, z1), at(x1, y1, z1));
  float v0 = lerp_l(fy, v00, v01);
  float v1 = lerp_l(fy, v10, v11);
  return lerp_l(fz, v0, v1);
}

// TransferFunction::DD::map (volume/TransferFunction.h:67-87)
void tf_map(const Volume &v, float s, float out[4])
{
  const int N = (int)v.tf.size() / 4;
  float f = (s - v.dom0) / (v.dom1 - v.dom0);
  f = fmn(fmx(f, 0.f), 1.f);
  f *= (float)(N - 1);
  int idx = std::min(std::max((int)f, 0), N - 2);
  f -= (float)idx;
  for (int k = 0; k < 4; k++) out[k] = (1.f - f) * v.tf[4 * idx + k] + f * v.tf[4 * (idx + 1) + k];
  out[3] *= v.base_density;
}

// TransferFunction::DD::majorant (volume/TransferFunction.h:89-117)
float tf_majorant(const Volume &v, float lo, float hi)
{
  if (lo > hi) return 0.f;
  const int N = (int)v.tf.size() / 4;
  const float span = v.dom1 - v.dom0;
  float f_lo = (lo - v.dom0) / span;
  float f_hi = (hi - v.dom0) / span;
  f_lo = fmn(fmx(f_lo, 0.f), 1.f);
  f_hi = fmn(fmx(f_hi, 0.f), 1.f);
  f_lo *= (float)(N - 1);
  f_hi *= (float)(N - 1);
  int i_lo = std::min(std::max((int)f_lo, 0), N - 2);
  int i_hi = std::min(std::max((int)f_hi, 0), N - 2);
  f_lo -= (float)i_lo;
  f_hi -= (float)i_hi;
  float l_lo = (1.f - f_lo) * v.tf[4 * i_lo + 3] + f_lo * v.tf[4 * (i_lo + 1) + 3];
  float l_hi = (1.f - f_hi) * v.tf[4 * i_hi + 3] + f_hi * v.tf[4 * (i_hi + 1) + 3];
  float m = 0.f;
  m = fmx(m, l_lo);
  m = fmx(m, l_hi);
  for (int i = i_lo + 1; i <= i_hi; i++) m = fmx(m, v.tf[4 * i + 3]);
  return m * v.base_density;
}

// StructuredData::commit (:177-181), buildMCs + StructuredData_computeMCs
// (StructuredData.cu:30-96), then majorants
void build_volume(Volume &v)
{
  if (v.field == FIELD_UMESH) {
    build_umesh(v);
    return;
  }
  for (int k = 0; k < 3; k++) v.ncells[k] = v.dims[k] - 1;
  v.bounds.lo = v.origin;
  v.bounds.hi = v.origin + v.spacing * vec3((float)v.ncells[0], (float)v.ncells[1], (float)v.ncells[2]);
  const int c = v.cells_per_mc;
  for (int k = 0; k < 3; k++) v.mc[k] = (v.ncells[k] + c - 1) / c;
  v.mc_origin = v.bounds.lo;
  v.mc_spacing = vec3((float)c) * v.spacing;
  const size_t nmc = (size_t)v.mc[0] * v.mc[1] * v.mc[2];
  v.mc_lo.assign(nmc, INFINITY);
  v.mc_hi.assign(nmc, -INFINITY);
  for (int mz = 0; mz < v.mc[2]; mz++)
    for (int my = 0; my < v.mc[1]; my++)
      for (int mx = 0; mx < v.mc[0]; mx++) {
        float lo = INFINITY, hi = -INFINITY;
        for (int iz = 0; iz <= c; iz++)
          for (int iy = 0; iy <= c; iy++)
            for (int ix = 0; ix <= c; ix++) {
              int sx = mx * c + ix, sy = my * c + iy, sz = mz * c + iz;
              if (sx >= v.dims[0] || sy >= v.dims[1] || sz >= v.dims[2]) continue;
              float f = v.scalars[(size_t)sx + (size_t)v.dims[0] * ((size_t)sy + (size_t)v.dims[1] * sz)];
              lo = fmn(lo, f);
              hi = fmx(hi, f);
            }
        size_t i = (size_t)mx + (size_t)v.mc[0] * ((size_t)my + (size_t)v.mc[1] * mz);
        v.mc_lo[i] = lo;
        v.mc_hi[i] = hi;
      }
  build_majorants(v);
}

// mapMCs (volume/MCGrid.cu:96-113)
void build_majorants(Volume &v)
{
  v.majorant.resize(v.mc_lo.size());
  for (size_t i = 0; i < v.mc_lo.size(); i++) v.majorant[i] = tf_majorant(v, v.mc_lo[i], v.mc_hi[i]);
}

// fastLog (volume/Volume.h:154-165): 3-term atanh series of log
float fast_log(float f)
{
  f = (f - 1.f) / (f + 1.f);
  float f2 = f * f;
  float s = f;
  f *= f2;
  s += (1.f / 3.f) * f;
  f *= f2;
  s += (1.f / 5.f) * f;
  return s + s;
}

static bool box_clip(vec3 org, vec3 dir, const Box &b, float &t0, float &t1)
{
  // boxTest (render/Ray.h:169-217)
  vec3 rd(1.f / dir.x, 1.f / dir.y, 1.f / dir.z);
  vec3 t_lo = (b.lo - org) * rd;
  vec3 t_hi = (b.hi - org) * rd;
  vec3 t_nr = vmin(t_lo, t_hi), t_fr = vmax(t_lo, t_hi);
  t0 = fmx(t0, reduce_max(t_nr));
  t1 = fmn(t1, reduce_min(t_fr));
  return t0 <= t1;
}

bool volume_march(const Volume &v, uint32_t rtc_inst, vec3 org, vec3 dir, float tmin, float tmax, RNGSeed &parent,
                  float &t_out, float rgb[3], uint64_t *samples)
{
  // MCVolumeAccel::isProg (volume/MCAccelerator.h:550-689)
  float r0 = tmin, r1 = tmax;
  if (!box_clip(org, dir, v.bounds, r0, r1)) return false;
  vec3 rsp(1.f / v.mc_spacing.x, 1.f / v.mc_spacing.y, 1.f / v.mc_spacing.z);
  vec3 dorg = (org - v.mc_origin) * rsp;
  vec3 ddir = dir * rsp;
  Random2 rng = Random2::from64(parent, fnv_hash3(rtc_inst, 0u, 0u));
  const float ray_tmax = tmax;   // ray.tMax seen by the lambda (:603)
  bool found = false;
  // dda3 over the macrocell grid, clipped to the bounds overlap
  dda3(dorg, ddir, r1, v.mc, [&](const int cell[3], float c0, float c1) -> bool {
    // lambda (:597-686)
    const float maj = v.majorant[mc_index(v, cell)];
    if (maj == 0.f) return true;
    // Woodcock::sampleRange (volume/Volume.h:168-203)
    float lower = c0, upper = fmn(c1, ray_tmax);
    float t = lower;
    while (true) {
      float r = rng();
      float dt = -fast_log(1.f - r) / maj;
      t += dt;
      if (t >= upper) return true;
      vec3 P = org + t * dir;
      float f = volume_sample(v, P);
      if (samples) (*samples)++;
      float s[4] = {0.f, 0.f, 0.f, 0.f};
      if (!std::isnan(f)) tf_map(v, f, s);
      if (s[3] >= rng() * maj) {
        t_out = t;
        rgb[0] = s[0];
        rgb[1] = s[1];
        rgb[2] = s[2];
        found = true;
        return false;
      }
    }
  });
  return found;
}

// Ratio tracking (Novak, Selle and Jarosz, SIGGRAPH Asia 2014, section 3)
// of a shadow ray through one volume: the walk, RNG child and tentative
// collisions of volume_march (one draw per step), the estimate multiplied by
// 1 - sigma / majorant at each.  Opt-in (BNX_RENDER_RATIO_TRACKING); the
// reference has delta tracking only (volume/Volume.h:168-203).
float ratio_track(const Volume &v, uint32_t rtc_inst, vec3 org, vec3 dir, float tmin, float tmax, RNGSeed &parent,
                  uint64_t *samples)
{
  float r0 = tmin, r1 = tmax;
  if (!box_clip(org, dir, v.bounds, r0, r1)) return 1.f;
  vec3 rsp(1.f / v.mc_spacing.x, 1.f / v.mc_spacing.y, 1.f / v.mc_spacing.z);
  vec3 dorg = (org - v.mc_origin) * rsp;
  vec3 ddir = dir * rsp;
  Random2 rng = Random2::from64(parent, fnv_hash3(rtc_inst, 0u, 0u));
  const float ray_tmax = tmax;
  float T = 1.f;
  dda3(dorg, ddir, r1, v.mc, [&](const int cell[3], float c0, float c1) -> bool {
    const float maj = v.majorant[mc_index(v, cell)];
    if (maj == 0.f) return true;
    const float upper = fmn(c1, ray_tmax);
    float t = c0;
    while (true) {
      const float r = rng();
      t += -fast_log(1.f - r) / maj;
      if (t >= upper) return true;
      const vec3 P = org + t * dir;
      const float f = volume_sample(v, P);
      if (samples) (*samples)++;
      float s[4] = {0.f, 0.f, 0.f, 0.f};
      if (!std::isnan(f)) tf_map(v, f, s);
      T *= fmx(0.f, 1.f - s[3] / maj);
      if (T == 0.f) return false;
    }
  });
  return T;
}

} // namespace orc
