// bnx_render.h -- wavefront queue, shading parameters and frame views.
#pragma once

#include "bnx_internal.h"

namespace bnx {

// structure-of-arrays ray queue (Ray + PathState + hit of render/Ray.h:17-93)
struct RayQueue {
  float4 *org;          // org.xyz, tMax (hit t after trace)
  float4 *dir;          // dir.xyz, flags: bit0 shadow, bits 16.. sample index in the batch
  uint64_t *rng;        // RNGSeed state
  uint4 *ps;            // throughput (3 x fp16), diffuse bounces, pixel, misWeight
  int4 *hit;            // prim, geom, inst, bits(u)
  float *hit_v;
  float4 *vcol;         // rgb of a volume event
};

struct VolumeDev;
struct ShadeParams {
  bnx_renderer rr;
  const VolumeDev *vols;       // world volumes (slot = hit geom - VOLUME_GEOM_BASE)
  int num_vols;
  const bnx_dir_light *dir_lights;
  int num_dir_lights;
  int W, H;
  int accum_base;
  int generation;
  unsigned long long *stats;   // statistics array (STAT_* x STAT_SLOTS)
};

struct FrameDev {
  unsigned long long *accum;   // 4 x int64 per pixel, 2^-32 fixed point
  int *depth;                  // ordered-int float
  int *prim, *inst, *obj;
  float4 *normal;
};

inline FrameDev frame_dev(Frame *f)
{
  FrameDev d;
  d.accum = f->accum.as<unsigned long long>();
  d.depth = f->depth.as<int>();
  d.prim = f->prim.as<int>();
  d.inst = f->inst.as<int>();
  d.obj = f->obj.as<int>();
  d.normal = f->normal.as<float4>();
  return d;
}

// per-device ray queues (two, ping-pong) and counters
struct Queues {
  DevMem org[2], dir[2], rng[2], ps[2], hit[2], hit_v[2], vcol[2];
  DevMem counts;      // [0], [1]: queue counts
  DevMem stats;       // STAT_COUNT x STAT_SLOTS 64-bit counters
  size_t cap = 0;
  hipEvent_t ev_begin = nullptr, ev_end = nullptr, ev_vbegin = nullptr, ev_vend = nullptr;
  int ensure(size_t n)
  {
    if (n <= cap) return BNX_OK;
    for (int k = 0; k < 2; k++) {
      BNX_HIP(org[k].ensure(sizeof(float4) * n));
      BNX_HIP(dir[k].ensure(sizeof(float4) * n));
      BNX_HIP(rng[k].ensure(sizeof(uint64_t) * n));
      BNX_HIP(ps[k].ensure(sizeof(uint4) * n));
      BNX_HIP(hit[k].ensure(sizeof(int4) * n));
      BNX_HIP(hit_v[k].ensure(sizeof(float) * n));
      BNX_HIP(vcol[k].ensure(sizeof(float4) * n));
    }
    BNX_HIP(counts.ensure(128));
    BNX_HIP(stats.ensure(sizeof(unsigned long long) * STAT_COUNT * STAT_SLOTS));
    cap = n;
    return BNX_OK;
  }
  RayQueue q(int k)
  {
    RayQueue r;
    r.org = org[k].as<float4>();
    r.dir = dir[k].as<float4>();
    r.rng = rng[k].as<uint64_t>();
    r.ps = ps[k].as<uint4>();
    r.hit = hit[k].as<int4>();
    r.hit_v = hit_v[k].as<float>();
    r.vcol = vcol[k].as<float4>();
    return r;
  }
};

int frame_clear(Frame *f);
int frame_read(Frame *f, int channel, void *host);
int launch_raygen(Device *dev, const bnx_camera &cam, const bnx_renderer &rr, int W, int H, int accum_base, int S,
                  const RayQueue &q);
int launch_shade(Device *dev, const SceneDev &S, const ShadeParams &P, const RayQueue &in, const int *in_count,
                 int max_in, const RayQueue &out, int *out_count, const FrameDev &fb);
int trace_stack_capacity();

} // namespace bnx
