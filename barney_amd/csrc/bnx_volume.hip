// bnx_volume.hip -- structured volumes on gfx950: macrocell value ranges,
// majorants, and the DDA + Woodcock (delta tracking) march.
//
// Replaces MCVolumeAccel<StructuredDataSampler> (volume/MCAccelerator.h:
// 155-195, 550-689), StructuredData_computeMCs (volume/StructuredData.cu:
// 30-62), mapMCs (volume/MCGrid.cu:96-113), dda3 (volume/DDA.h:60-201) and
// Woodcock::sampleRange (volume/Volume.h:154-203).
//
// Design: volumes are marched in their own wavefront kernel after the
// surface trace (one thread per queued ray, persistent waves), clipped to
// the surface hit.  The march is long, data-dependent and texture-bound, so
// keeping it out of the BVH traversal kernel keeps that kernel's registers
// and divergence low.  Trilinear filtering is done in software on the raw
// float grid (hardware texture filtering uses 8-bit fixed-point weights);
// a 256^3 field is 64 MiB and stays resident in the 256 MiB Infinity Cache.
#include "bnx_internal.h"
#include "bnx_render.h"
#include "bnx_volume.h"

namespace bnx {
namespace {

inline int nblk(long n, int b) { return (int)((n + b - 1) / b); }

// StructuredData_computeMCs: min/max over the (cpm+1)^3 vertices of a macrocell
__global__ void k_mc_ranges(VolumeDev v, float2 *ranges)
{
  int tid = blockIdx.x * blockDim.x + threadIdx.x;
  int n = v.mc[0] * v.mc[1] * v.mc[2];
  if (tid >= n) return;
  int mx = tid % v.mc[0], my = (tid / v.mc[0]) % v.mc[1], mz = tid / (v.mc[0] * v.mc[1]);
  const int c = v.cells_per_mc;
  float lo = INFINITY, hi = -INFINITY;
  for (int iz = 0; iz <= c; iz++) {
    int sz = mz * c + iz;
    if (sz >= v.dims[2]) continue;
    for (int iy = 0; iy <= c; iy++) {
      int sy = my * c + iy;
      if (sy >= v.dims[1]) continue;
      const float *row = v.scalars + (size_t)v.dims[0] * ((size_t)sy + (size_t)v.dims[1] * sz);
      for (int ix = 0; ix <= c; ix++) {
        int sx = mx * c + ix;
        if (sx >= v.dims[0]) continue;
        float f = row[sx];
        lo = fmn(lo, f);
        hi = fmx(hi, f);
      }
    }
  }
  ranges[tid] = make_float2(lo, hi);
}

// mapMCs + TransferFunction::DD::majorant
__global__ void k_majorants(VolumeDev v, const float2 *ranges, float *maj)
{
  int tid = blockIdx.x * blockDim.x + threadIdx.x;
  int n = v.mc[0] * v.mc[1] * v.mc[2];
  if (tid >= n) return;
  float2 r = ranges[tid];
  maj[tid] = tf_majorant(v, r.x, r.y);
}

// march every ray of the generation's segments through its request's
// volumes, one ray per thread; each block lies inside one request's segment,
// so the volume table is block-uniform.  With no surface instances in the
// batch the trace kernel is skipped and every ray enters here as a surface
// miss (surfaces == 0).
__global__ __launch_bounds__(256) void k_volume_march(const SlotDev *__restrict__ slots, SegView sv, RayQueue q,
                                                      int surfaces, unsigned long long *stats)
{
  uint32_t ns = 0, nc = 0;
  int i;
  bool valid;
  const Segment sg = segment_of<256>(sv, i, valid);
  const VolumeList vl{slots[sg.slot].vols, slots[sg.slot].num_vols};
  if (valid) {
    float4 o = q.org[i];
    float4 d = q.dir[i];
    int4 hit = surfaces ? q.hit[i] : make_int4(-1, -1, -1, 0);
    const bool shadow = (__float_as_uint(d.w) & FLAG_SHADOW) != 0;
    if (!(shadow && hit.x >= 0)) {                 // an occluded shadow ray is done
      const v3 org = xyz(o), dir = fix_dir(xyz(d));
      float tcur = o.w;
      uint64_t rng = q.rng[i];
      const uint64_t rng0 = rng;
      int vk = -1;
      float vt = 0.f, rgb[3] = {0.f, 0.f, 0.f};
      for (int k = 0; k < vl.count; k++) {
        float t, c[3];
        if (volume_march(vl.vols[k], org, dir, 0.f, tcur, rng, t, c, ns, nc)) {
          vk = k;
          vt = t;
          tcur = t;
          rgb[0] = c[0]; rgb[1] = c[1]; rgb[2] = c[2];
          if (shadow) break;
        }
      }
      if (rng != rng0) q.rng[i] = rng;
      if (vk >= 0) {
        q.org[i].w = vt;
        hit = make_int4(0, VOLUME_GEOM_BASE + vk, vl.vols[vk].rtc_inst, 0);
        q.vcol[i] = make_float4(rgb[0], rgb[1], rgb[2], 0.f);
        q.hit[i] = hit;
      } else if (!surfaces) {
        q.hit[i] = hit;
      }
    } else if (!surfaces) {
      q.hit[i] = hit;
    }
  }
  uint32_t s = wave_sum(ns), c = wave_sum(nc);
  if (lane_id() == 0) {
    stat_add(stats, STAT_SAMPLES, s);
    stat_add(stats, STAT_CELLS, c);
  }
}

} // namespace

int volume_build(Volume *v)
{
  Device *dev = v->dev;
  VolumeDev &d = v->dd;
  const size_t nmc = (size_t)d.mc[0] * d.mc[1] * d.mc[2];
  BNX_HIP(v->ranges.ensure(sizeof(float2) * nmc));
  BNX_HIP(v->majorants.ensure(sizeof(float) * nmc));
  BNX_SET(d.majorants, v->majorants.ptr);
  k_mc_ranges<<<nblk((long)nmc, 64), 64, 0, dev->stream>>>(d, v->ranges.as<float2>());
  k_majorants<<<nblk((long)nmc, 256), 256, 0, dev->stream>>>(d, v->ranges.as<float2>(), v->majorants.as<float>());
  BNX_HIP(hipStreamSynchronize(dev->stream));
  BNX_HIP(hipGetLastError());
  return BNX_OK;
}

int volume_rebuild_majorants(Volume *v)
{
  Device *dev = v->dev;
  const size_t nmc = (size_t)v->dd.mc[0] * v->dd.mc[1] * v->dd.mc[2];
  k_majorants<<<nblk((long)nmc, 256), 256, 0, dev->stream>>>(v->dd, v->ranges.as<float2>(),
                                                             v->majorants.as<float>());
  BNX_HIP(hipStreamSynchronize(dev->stream));
  BNX_HIP(hipGetLastError());
  return BNX_OK;
}

int launch_volume_march(Device *dev, const SlotDev *slots, const SegView &sv, int blocks256, const RayQueue &q,
                        bool surfaces, unsigned long long *stats)
{
  if (blocks256 > 0)
    k_volume_march<<<blocks256, 256, 0, dev->stream>>>(slots, sv, q, surfaces ? 1 : 0, stats);
  BNX_HIP(hipGetLastError());
  return BNX_OK;
}

} // namespace bnx
