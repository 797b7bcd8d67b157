// bnx_render.h -- wavefront queue, shading parameters and frame views.
#pragma once

#include "bnx_internal.h"

namespace bnx {

// structure-of-arrays ray queue (Ray + PathState + hit of render/Ray.h:17-93)
struct RayQueue {
  float4 *org;          // org.xyz, tMax (hit t after trace)
  float4 *dir;          // dir.xyz, flags (see mk_flags)
  uint64_t *rng;        // RNGSeed state
  uint4 *ps;            // x: throughput.xy (fp16 x 2); y: throughput.z (fp16) | diffuse bounces << 16 |
                        // generation << 24; z: pixel; w: misWeight
  int4 *hit;            // prim, geom, inst, bits(u)
  float *hit_v;
  float4 *vcol;         // rgb of a volume event
};

// ray flags in dir.w: bit 0 shadow ray, bits 4..15 sample index inside the
// current batch of its request, bits 16..31 request slot
constexpr uint32_t FLAG_SHADOW = 1u;
constexpr int MAX_BATCH_SAMPLES = 4096;
constexpr int MAX_SLOTS = 65536;
BNX_D uint32_t mk_flags(bool shadow, int sample, int slot)
{
  return (shadow ? FLAG_SHADOW : 0u) | ((uint32_t)sample << 4) | ((uint32_t)slot << 16);
}
BNX_D int flag_sample(uint32_t f) { return (int)((f >> 4) & 0xfffu); }
BNX_D int flag_slot(uint32_t f) { return (int)(f >> 16); }

struct FrameDev {
  unsigned long long *accum;   // 4 x int64 per pixel, 2^-32 fixed point
  int *depth;                  // ordered-int float
  int *prim, *inst, *obj;
  float4 *normal;
};

struct VolumeDev;
// One request in flight: everything its rays need to be generated, traced,
// marched, shaded and accumulated.  Rays of different requests (different
// worlds, cameras, resolutions, frames) share one queue and one launch per
// kernel; each ray finds its request through the slot index in its flags.
struct SlotDev {
  SceneDev scene;
  const VolumeDev *vols;       // world volumes (volume k: hit geom VOLUME_GEOM_BASE + k)
  int num_vols;
  int num_dir_lights;
  const bnx_dir_light *dir_lights;
  bnx_renderer rr;
  bnx_camera cam;
  int W, H;
  int accum_base;              // accumID of batch sample 0
  int pad;
  FrameDev fb;
};

// A run of primary rays of one request: queue entries [begin, next chunk's
// begin) hold the request's path indices first, first + 1, ... where path
// index = sample * W * H + pixel and sample counts from the slot's accum_base.
struct RaygenChunk {
  long long begin;             // first queue index
  long long first;             // first path index
  int slot;
  int pad;
};

// per-request count of rays written by shade, spread over LIVE_SPREAD
// counters per slot so that one generation's atomics do not serialize;
// [slot][0][.]: all rays, [slot][1][.]: path (non-shadow) rays
constexpr int LIVE_SPREAD = 32;

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
// primary rays for total queue entries starting at queue index q_begin
int launch_raygen(Device *dev, const RaygenChunk *d_chunks, int num_chunks, const SlotDev *d_slots,
                  long long q_begin, long long total, const RayQueue &q);
// shade in[0, *in_count), appending to out at *out_count; live (optional):
// per-slot counters of written rays, see LIVE_SPREAD
int launch_shade(Device *dev, const SlotDev *d_slots, const RayQueue &in, const int *in_count, int max_in,
                 const RayQueue &out, int *out_count, unsigned int *live, unsigned long long *stats);
// sum and clear the live counters of the listed slots: out[2k], out[2k+1] =
// rays and path rays of slot slots[k]
int launch_live_reduce(Device *dev, unsigned int *live, const int *d_slot_list, int n, unsigned int *out);
int trace_stack_capacity();

// the per-request frame loop behind bnx_render
int render_frame(World *w, const bnx_camera &cam, const bnx_renderer &rr, Frame *f, int spp,
                 bnx_render_stats *stats);
// free the device's render queues
void wavefront_destroy(Device *dev);

} // namespace bnx
