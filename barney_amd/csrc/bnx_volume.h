// bnx_volume.h -- structured-volume device data and the per-ray march.
#pragma once

#include "bnx_internal.h"
#include "bnx_render.h"

namespace bnx {

// hit-record geometry index of a volume event: VOLUME_GEOM_BASE + volume slot
constexpr int VOLUME_GEOM_BASE = 0x40000000;

struct VolumeDev {
  const float BNX_G *scalars;   // dims[0]*dims[1]*dims[2], x fastest
  int dims[3];
  int ncells[3];
  float origin[3];
  float rcp_spacing[3];       // 1/gridSpacing (host float division)
  float lo[3], hi[3];         // world bounds (StructuredData::commit)
  int mc[3];
  int cells_per_mc;
  float mc_origin[3];
  float mc_rcp_spacing[3];    // 1/(cells_per_mc * gridSpacing)
  const float BNX_G *majorants;
  const float4 BNX_G *tf;
  int ntf;
  float dom0, dom1, base_density;
  int rtc_inst;               // rtc instance index (seeds the march RNG)
  int user_inst;
  int user_id;
};

struct VolumeList {
  const VolumeDev BNX_G *vols;  // device array
  int count;
};

struct Volume {
  Device *dev;
  VolumeDev dd{};
  DevMem scalars, ranges, majorants, tf;
};

int volume_build(Volume *v);
int volume_rebuild_majorants(Volume *v);
// upload the world's volume table (if changed) and return it
int world_volumes(World *w, VolumeList &vl);
// march every ray of the generation's segments (blocks256 blocks of 256
// threads); surfaces: false when the trace kernel did not run, so the hit
// records are not written yet
int launch_volume_march(Device *dev, const SlotDev *d_slots, const SegView &sv, int blocks256, const RayQueue &q,
                        bool surfaces, unsigned long long *stats);

// ---------------------------------------------------------------- device
BNX_D uint64_t fnv3(uint32_t a, uint32_t b, uint32_t c)
{
  // hash(v0,v1,v2) of native/common/barney-common.h:88-105
  const uint64_t prime = 0x100000001b3ull;
  uint64_t h = 0xcbf29ce484222325ull ^ (prime * (uint64_t)a);
  h = h * prime ^ (uint64_t)b;
  h = h * prime ^ (uint64_t)c;
  return h;
}

// CPU texture LINEAR/CLAMP addressing (rtcore/cpu/Texture.cpp:324-336)
BNX_D void lerp_coord(int N, float tc, int &i0, int &i1, float &f)
{
  if (tc - .5f <= 0.f) { i0 = i1 = 0; f = 0.f; return; }
  if (tc - .5f >= (float)(N - 1)) { i0 = i1 = N - 1; f = 0.f; return; }
  f = (tc - .5f) - (float)(int)(tc - .5f);
  i0 = (int)(tc - .5f);
  i1 = i0 + 1;
}

BNX_D float lerp_l(float f, float a, float b) { return (1.f - f) * a + f * b; }

// StructuredDataSampler::DD::sample (volume/StructuredData.h:101-112)
BNX_D float volume_sample(const VolumeDev &v, v3 P)
{
  float rx = (P.x - v.origin[0]) * v.rcp_spacing[0];
  float ry = (P.y - v.origin[1]) * v.rcp_spacing[1];
  float rz = (P.z - v.origin[2]) * v.rcp_spacing[2];
  if (rx < 0.f || ry < 0.f || rz < 0.f) return __builtin_nanf("");
  if (rx >= (float)v.ncells[0] || ry >= (float)v.ncells[1] || rz >= (float)v.ncells[2]) return __builtin_nanf("");
  int x0, x1, y0, y1, z0, z1;
  float fx, fy, fz;
  lerp_coord(v.dims[0], rx + .5f, x0, x1, fx);
  lerp_coord(v.dims[1], ry + .5f, y0, y1, fy);
  lerp_coord(v.dims[2], rz + .5f, z0, z1, fz);
  const size_t Nx = (size_t)v.dims[0], NxNy = (size_t)v.dims[0] * (size_t)v.dims[1];
  const float *sc = v.scalars;
  const float *b00 = sc + (size_t)y0 * Nx + (size_t)z0 * NxNy;
  const float *b01 = sc + (size_t)y1 * Nx + (size_t)z0 * NxNy;
  const float *b10 = sc + (size_t)y0 * Nx + (size_t)z1 * NxNy;
  const float *b11 = sc + (size_t)y1 * Nx + (size_t)z1 * NxNy;
  float v00 = lerp_l(fx, b00[x0], b00[x1]);
  float v01 = lerp_l(fx, b01[x0], b01[x1]);
  float v10 = lerp_l(fx, b10[x0], b10[x1]);
  float v11 = lerp_l(fx, b11[x0], b11[x1]);
  float v0 = lerp_l(fy, v00, v01);
  float v1 = lerp_l(fy, v10, v11);
  return lerp_l(fz, v0, v1);
}

// TransferFunction::DD::map (volume/TransferFunction.h:67-87)
BNX_D float4 tf_map(const VolumeDev &v, float s)
{
  const int N = v.ntf;
  float f = (s - v.dom0) / (v.dom1 - v.dom0);
  f = fmn(fmx(f, 0.f), 1.f);
  f *= (float)(N - 1);
  int idx = (int)f;
  idx = idx < 0 ? 0 : (idx > N - 2 ? N - 2 : idx);
  f -= (float)idx;
  const float4 *tf = v.tf;
  float4 a = tf[idx], b = tf[idx + 1];
  float4 r;
  r.x = (1.f - f) * a.x + f * b.x;
  r.y = (1.f - f) * a.y + f * b.y;
  r.z = (1.f - f) * a.z + f * b.z;
  r.w = (1.f - f) * a.w + f * b.w;
  r.w *= v.base_density;
  return r;
}

// TransferFunction::DD::majorant (volume/TransferFunction.h:89-117)
BNX_D float tf_majorant(const VolumeDev &v, float lo, float hi)
{
  if (lo > hi) return 0.f;
  const int N = v.ntf;
  const float span = v.dom1 - v.dom0;
  float f_lo = (lo - v.dom0) / span;
  float f_hi = (hi - v.dom0) / span;
  f_lo = fmn(fmx(f_lo, 0.f), 1.f);
  f_hi = fmn(fmx(f_hi, 0.f), 1.f);
  f_lo *= (float)(N - 1);
  f_hi *= (float)(N - 1);
  int i_lo = (int)f_lo, i_hi = (int)f_hi;
  i_lo = i_lo < 0 ? 0 : (i_lo > N - 2 ? N - 2 : i_lo);
  i_hi = i_hi < 0 ? 0 : (i_hi > N - 2 ? N - 2 : i_hi);
  f_lo -= (float)i_lo;
  f_hi -= (float)i_hi;
  const float4 *tf = v.tf;
  float l_lo = (1.f - f_lo) * tf[i_lo].w + f_lo * tf[i_lo + 1].w;
  float l_hi = (1.f - f_hi) * tf[i_hi].w + f_hi * tf[i_hi + 1].w;
  float m = 0.f;
  m = fmx(m, l_lo);
  m = fmx(m, l_hi);
  for (int i = i_lo + 1; i <= i_hi; i++) m = fmx(m, tf[i].w);
  return m * v.base_density;
}

// fastLog (volume/Volume.h:154-165)
BNX_D float fast_log(float f)
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

// MCVolumeAccel::isProg (volume/MCAccelerator.h:550-689): box clip, RNG
// child seeded by hash(rtc instance, geometry 0, 0), dda3 over macrocells,
// Woodcock sampling per non-empty cell.  'dir' is the fixed trace direction.
BNX_D bool volume_march(const VolumeDev &v, v3 org, v3 dir, float tmin, float tmax, uint64_t &parent, float &t_out,
                        float rgb[3], uint32_t &nsamples, uint32_t &ncells)
{
  // boxTest (render/Ray.h:169-217)
  float r0 = tmin, r1 = tmax;
  {
    v3 rd = v3{1.f / dir.x, 1.f / dir.y, 1.f / dir.z};
    float tl[3] = {(v.lo[0] - org.x) * rd.x, (v.lo[1] - org.y) * rd.y, (v.lo[2] - org.z) * rd.z};
    float th[3] = {(v.hi[0] - org.x) * rd.x, (v.hi[1] - org.y) * rd.y, (v.hi[2] - org.z) * rd.z};
    float nr = fmx(fmx(fmn(tl[0], th[0]), fmn(tl[1], th[1])), fmn(tl[2], th[2]));
    float fr = fmn(fmn(fmx(tl[0], th[0]), fmx(tl[1], th[1])), fmx(tl[2], th[2]));
    r0 = fmx(r0, nr);
    r1 = fmn(r1, fr);
    if (!(r0 <= r1)) return false;
  }
  const v3 dorg = v3{(org.x - v.mc_origin[0]) * v.mc_rcp_spacing[0], (org.y - v.mc_origin[1]) * v.mc_rcp_spacing[1],
                     (org.z - v.mc_origin[2]) * v.mc_rcp_spacing[2]};
  const v3 ddir = v3{dir.x * v.mc_rcp_spacing[0], dir.y * v.mc_rcp_spacing[1], dir.z * v.mc_rcp_spacing[2]};
  // Random(ray.rngSeed, hash(inst, geom, 0)) -- barney-common.h:59-61
  uint64_t rs = parent;
  rng_next(parent);
  rng_next_h(rs, fnv3((uint32_t)v.rtc_inst, 0u, 0u));
  auto rnd = [&]() -> float { return (float)(rng_next(rs) & 0x00FFFFFFu) * (1.f / (float)0x01000000); };
  const float ray_tmax = tmax;

  // dda3 (volume/DDA.h:60-201)
  const float dda_tmax = r1;
  const float fs[3] = {(float)v.mc[0], (float)v.mc[1], (float)v.mc[2]};
  const float o3[3] = {dorg.x, dorg.y, dorg.z}, d3[3] = {ddir.x, ddir.y, ddir.z};
  float rcp[3], tn[3], tf[3];
  for (int k = 0; k < 3; k++) {
    rcp[k] = 1.f / d3[k];
    float a = (0.f - o3[k]) * rcp[k];
    float b = (fs[k] - o3[k]) * rcp[k];
    tn[k] = fmn(a, b);
    tf[k] = fmx(a, b);
    if (d3[k] == 0.f) {
      if (o3[k] < 0.f || o3[k] > fs[k]) return false;
      tn[k] = -INFINITY;
      tf[k] = INFINITY;
    }
  }
  float ray_t0 = fmx(0.f, fmx(fmx(tn[0], tn[1]), tn[2]));
  float ray_t1 = fmn(dda_tmax, fmn(fmn(tf[0], tf[1]), tf[2]));
  if (ray_t0 > ray_t1) return false;
  float t_step[3], t_next[3];
  int stop[3], delta[3], cell[3];
  for (int k = 0; k < 3; k++) {
    float oiv = o3[k] + ray_t0 * d3[k];
    float fc = fmx(0.f, fmn(fs[k] - 1.f, floorf(oiv)));
    float fe = d3[k] > 0.f ? fc + 1.f : fc;
    t_step[k] = fabsf(rcp[k]);
    t_next[k] = (d3[k] == 0.f) ? INFINITY : fabsf(fe - oiv) * t_step[k];
    stop[k] = d3[k] > 0.f ? v.mc[k] : -1;
    delta[k] = d3[k] > 0.f ? 1 : -1;
    cell[k] = (int)fc;
  }
  float next_begin = 0.f;
  while (true) {
    const float t_closest = fmn(fmn(t_next[0], t_next[1]), t_next[2]);
    const float c0 = ray_t0 + next_begin;
    const float c1 = ray_t0 + fmn(t_closest, dda_tmax);
    const float maj = v.majorants[(size_t)cell[0] + (size_t)v.mc[0] * ((size_t)cell[1] + (size_t)v.mc[1] * cell[2])];
    ncells++;
    if (maj != 0.f) {
      // Woodcock::sampleRange (volume/Volume.h:168-203)
      const float upper = fmn(c1, ray_tmax);
      float t = c0;
      while (true) {
        float r = rnd();
        float dt = -fast_log(1.f - r) / maj;
        t += dt;
        if (t >= upper) break;
        v3 P = org + t * dir;
        float f = volume_sample(v, P);
        nsamples++;
        float4 s = isnan(f) ? make_float4(0.f, 0.f, 0.f, 0.f) : tf_map(v, f);
        if (s.w >= rnd() * maj) {
          t_out = t;
          rgb[0] = s.x;
          rgb[1] = s.y;
          rgb[2] = s.z;
          return true;
        }
      }
    }
    next_begin = t_closest;
    if (t_next[0] == t_closest) {
      t_next[0] += t_step[0];
      cell[0] += delta[0];
      if (cell[0] == stop[0]) return false;
    }
    if (t_next[1] == t_closest) {
      t_next[1] += t_step[1];
      cell[1] += delta[1];
      if (cell[1] == stop[1]) return false;
    }
    if (t_next[2] == t_closest) {
      t_next[2] += t_step[2];
      cell[2] += delta[2];
      if (cell[2] == stop[2]) return false;
    }
  }
}

} // namespace bnx
