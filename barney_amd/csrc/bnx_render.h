// bnx_render.h -- wavefront queue, request slots, queue segments and frame
// views shared by the raygen, trace, volume-march and shade kernels.
#pragma once

#include "bnx_internal.h"
#include "bnx_env.h"

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
  float4 *vcol;         // rgb of a volume event, or the normal of an isosurface hit
};

// ray flags in dir.w: bit 0 shadow ray, bits 4..15 sample index inside the
// current batch of its request
constexpr uint32_t FLAG_SHADOW = 1u;
// bit 1: the trace wrote the ray's ID channels (a translucent candidate in
// front of its hit), so shade leaves them
constexpr uint32_t FLAG_IDS = 2u;
constexpr int MAX_BATCH_SAMPLES = 4096;
BNX_D uint32_t mk_flags(bool shadow, int sample)
{
  return (shadow ? FLAG_SHADOW : 0u) | ((uint32_t)sample << 4);
}
BNX_D int flag_sample(uint32_t f) { return (int)((f >> 4) & 0xfffu); }

struct FrameDev {
  unsigned long long BNX_G *accum;   // 4 x int64 per pixel, 2^-32 fixed point
  int BNX_G *depth;                  // ordered-int float
  int BNX_G *prim, *inst, *obj;
  float4 BNX_G *normal;
};

struct VolumeDev;
struct IsoDev;
// One request in flight: everything its rays need to be generated, traced,
// marched, shaded and accumulated.  Rays of different requests (different
// worlds, cameras, resolutions, frames) share one queue and one launch per
// kernel; the queue is cut into segments of one request each (Segment).
struct SlotDev {
  SceneDev scene;
  const VolumeDev BNX_G *vols; // world volumes (volume k: hit geom VOLUME_GEOM_BASE + k)
  const IsoDev BNX_G *isos;    // world isosurfaces (isosurface k: hit geom ISO_GEOM_BASE + k)
  int num_vols;
  int num_isos;
  int num_dir_lights;
  int pad0;
  const bnx_dir_light BNX_G *dir_lights;
  bnx_renderer rr;
  bnx_camera cam;
  int W, H;
  int accum_base;              // accumID of batch sample 0
  int pad;
  FrameDev fb;
  EnvDev env;                  // HDRI light (env.tex null: ambient radiance)
};

// the renderer's cut plane (nx, ny, nz, w; off when w <= -1e28)
BNX_D float4 slot_cut_plane(const SlotDev &s)
{
  return make_float4(s.rr.cut_plane[0], s.rr.cut_plane[1], s.rr.cut_plane[2], s.rr.cut_plane[3]);
}
inline bool has_cut_plane(const bnx_renderer &r) { return r.cut_plane[3] > -1e28f; }

// the device view of a world's env map
EnvDev env_dev(const World *w);

// A run of queue entries that all belong to one request slot.  Launches are
// laid out so that every block lies inside one segment: a block reads its
// segment (and through it its request's scene, volumes, lights, camera and
// frame) with scalar loads, and the request's parameters sit in scalar
// registers for the whole block.  Segment k covers blocks
// [block256, block256 + ceil(count / 256)) of a 256-thread launch,
// [block1024, ...) of a 1024-thread launch, and [blockT, ...) of the trace
// launch, whose blocks draw the segment's rays from a shared pool.
struct Segment {
  int slot;
  int begin;                   // first queue index
  int count;                   // rays
  int block256;
  int block1024;
  int blockT;
  int refill;                  // trace pool: idle lanes per wave that trigger a refill
  int pad;
};
static_assert(sizeof(Segment) == 32, "segment");

// kernel-side view of one generation's input: segments and the block ->
// segment maps of both launch widths
struct SegView {
  const Segment *segs;
  const int *blk256;           // segment of each 256-thread block
  const int *blk1024;          // segment of each 1024-thread block
  const int *blkT;             // segment of each trace block
  int *pool;                   // per segment: next ray of the trace pool
};

// shade output of one generation: request slot s appends to its own region
// of the output queue, starting at base[s]; cnt[2s] counts the rays written
// and cnt[2s + 1] the path (non-shadow) rays among them
struct SlotOut {
  const int *base;
  int *cnt;
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

inline FrameDev frame_dev(Frame *f)
{
  FrameDev d;
  BNX_SET(d.accum, f->accum.ptr);
  BNX_SET(d.depth, f->depth.ptr);
  BNX_SET(d.prim, f->prim.ptr);
  BNX_SET(d.inst, f->inst.ptr);
  BNX_SET(d.obj, f->obj.ptr);
  BNX_SET(d.normal, f->normal.ptr);
  return d;
}

// ray queues (two, ping-pong) and statistics counters
struct Queues {
  DevMem org[2], dir[2], rng[2], ps[2], hit[2], hit_v[2], vcol[2];
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
// the same with q_begin = d_range[0] and total = d_range[1] read on the
// device (a generation laid out by an earlier kernel); max_total bounds the
// launch
int launch_raygen_dev(Device *dev, const RaygenChunk *d_chunks, int num_chunks, const SlotDev *d_slots,
                      const long long *d_range, long long max_total, const RayQueue &q);
// a device-driven generation's single segment from the previous shade's
// output counter (count0 when g == 0); clears the counter, keeps the count
// in gen_counts[g]
int launch_gen_setup(Device *dev, Segment *segs, int *out_cnt, long long *gen_counts, int g, int count0, int refill);
// write the device wall clock (ticks of wall_clock_khz) to *t once the
// stream reaches this point
int launch_stamp(Device *dev, unsigned long long *t);

// scratch of the ray reordering (keys, indices, radix-sort temporaries)
struct SortScratch {
  DevMem keys[2], idx[2], tmp;
};
// the n rays of src (org, dir, rng, ps) into dst in the order of a key of
// shadow flag, direction octant and the origin's Morton code in [lo, hi]
int sort_rays(Device *dev, SortScratch &S, const RayQueue &src, const RayQueue &dst, int n, const float lo[3],
              const float hi[3]);
// fill the block -> segment maps of nsegs segments (blocks256 / blocks1024 /
// blocksT blocks in all) and reset the segments' trace pools
int launch_segment_blocks(Device *dev, const SegView &sv, int nsegs, int blocks256, int blocks1024, int blocksT);
// shade every ray of the segments (blocks256 blocks of 256 threads or
// blocks1024 of 1024, by the device's shade_block),
// appending each request's shadow and continuation rays to its region of out
int launch_shade(Device *dev, const SlotDev *d_slots, const SegView &sv, int blocks256, int blocks1024, const RayQueue &in,
                 const RayQueue &out, const SlotOut &so, unsigned long long *stats);
int trace_stack_capacity();

// the per-request frame loop behind bnx_render
int render_frame(World *w, const bnx_camera &cam, const bnx_renderer &rr, Frame *f, int spp,
                 bnx_render_stats *stats);
// free the device's render queues
void wavefront_destroy(Device *dev);

// The segment of this block of a B-thread launch, and the ray this thread
// takes (i, valid).  Every thread of a block gets the same segment.
template <int B> BNX_D Segment segment_of(const SegView &sv, int &i, bool &valid)
{
  const int k = (B == 256 ? sv.blk256 : sv.blk1024)[blockIdx.x];
  const Segment sg = sv.segs[k];
  const int local = ((int)blockIdx.x - (B == 256 ? sg.block256 : sg.block1024)) * B + (int)threadIdx.x;
  valid = local < sg.count;
  i = sg.begin + local;
  return sg;
}

} // namespace bnx
