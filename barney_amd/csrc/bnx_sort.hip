// bnx_sort.hip -- reordering of a generation's rays before they are traced:
// rays that start close together and head the same way walk much of the
// same BVH, and a wave pays for the union of its lanes' walks.  The rays of
// one request's queue region are sorted by a 31-bit key (shadow flag,
// direction octant, Morton code of the origin inside the world's bounds)
// and gathered into the other queue.  Only the order of rays in the queue
// changes; every ray keeps its own state.  Requested by bnx_tuning's
// sort_rays; bnx_render's host loop applies it to secondary generations.
#include "bnx_internal.h"
#include "bnx_render.h"

#include <hipcub/hipcub.hpp>

namespace bnx {
namespace {

// the 9 low bits of v spread to every third bit
BNX_D uint32_t spread3(uint32_t v)
{
  v &= 0x1ffu;
  v = (v | (v << 16)) & 0x030000ffu;
  v = (v | (v << 8)) & 0x0300f00fu;
  v = (v | (v << 4)) & 0x030c30c3u;
  v = (v | (v << 2)) & 0x09249249u;
  return v;
}

BNX_D uint32_t cell9(float x, float lo, float scale)
{
  const float c = (x - lo) * scale;
  return (uint32_t)fminf(511.f, fmaxf(0.f, c));
}

__global__ void k_ray_keys(RayQueue q, int n, float3 lo, float3 scale, uint32_t *keys, uint32_t *idx)
{
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float4 o = q.org[i], d = q.dir[i];
  const uint32_t shadow = (__float_as_uint(d.w) & FLAG_SHADOW) != 0;
  const uint32_t oct = ((d.x < 0.f) << 2) | ((d.y < 0.f) << 1) | (uint32_t)(d.z < 0.f);
  const uint32_t m = spread3(cell9(o.x, lo.x, scale.x)) | (spread3(cell9(o.y, lo.y, scale.y)) << 1) |
                     (spread3(cell9(o.z, lo.z, scale.z)) << 2);
  keys[i] = (shadow << 30) | (oct << 27) | m;
  idx[i] = (uint32_t)i;
}

__global__ void k_ray_gather(RayQueue src, RayQueue dst, const uint32_t *idx, int n)
{
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= n) return;
  const uint32_t i = idx[j];
  dst.org[j] = src.org[i];
  dst.dir[j] = src.dir[i];
  dst.rng[j] = src.rng[i];
  dst.ps[j] = src.ps[i];
}

} // namespace

int sort_rays(Device *dev, SortScratch &S, const RayQueue &src, const RayQueue &dst, int n, const float lo[3],
              const float hi[3])
{
  if (n <= 1) return BNX_OK;
  size_t tmp_bytes = 0;
  BNX_HIP(hipcub::DeviceRadixSort::SortPairs(nullptr, tmp_bytes, (const uint32_t *)nullptr, (uint32_t *)nullptr,
                                             (const uint32_t *)nullptr, (uint32_t *)nullptr, n, 0, 31,
                                             dev->stream));
  const size_t nb = sizeof(uint32_t) * (size_t)n;
  BNX_HIP(S.keys[0].ensure(nb));
  BNX_HIP(S.keys[1].ensure(nb));
  BNX_HIP(S.idx[0].ensure(nb));
  BNX_HIP(S.idx[1].ensure(nb));
  BNX_HIP(S.tmp.ensure(tmp_bytes));
  float3 flo, fscale;
  float *l = &flo.x, *sc = &fscale.x;
  for (int k = 0; k < 3; k++) {
    const float ext = hi[k] - lo[k];
    l[k] = lo[k];
    sc[k] = ext > 0.f ? 512.f / ext : 0.f;
  }
  const int blocks = (n + 255) / 256;
  k_ray_keys<<<blocks, 256, 0, dev->stream>>>(src, n, flo, fscale, S.keys[0].as<uint32_t>(),
                                              S.idx[0].as<uint32_t>());
  BNX_HIP(hipGetLastError());
  BNX_HIP(hipcub::DeviceRadixSort::SortPairs(S.tmp.ptr, tmp_bytes, S.keys[0].as<uint32_t>(),
                                             S.keys[1].as<uint32_t>(), S.idx[0].as<uint32_t>(),
                                             S.idx[1].as<uint32_t>(), n, 0, 31, dev->stream));
  k_ray_gather<<<blocks, 256, 0, dev->stream>>>(src, dst, S.idx[1].as<uint32_t>(), n);
  BNX_HIP(hipGetLastError());
  return BNX_OK;
}

} // namespace bnx
