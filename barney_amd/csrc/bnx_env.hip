// bnx_env.hip -- HDRI environment light: texture upload and the importance
// CDFs built on the device.
//
// Replaces EnvMapLight::commit / computeCDFs (light/EnvMap.cu:165-239) with
// its three kernels ComputeWeights_xy, ComputeCDFs_doLine and
// Normalize_cdf_y (EnvMap.cu:61-138), and HDRILight::setBarneyParameters
// (anari/light/HDRI.cpp:43-78), which widens the float3 radiance to float4
// texels.  The row sums run in the reference's sequential order (one thread
// per row, one thread for the row weights), so the CDFs match the CPU
// evaluation bit for bit; the build runs once per map, not per frame.
#include "bnx_render.h"

#include <vector>

namespace bnx {
namespace {

inline int nblk(long n, int b) { return (int)((n + b - 1) / b); }

// importance of each texel: the largest channel over a 3 x 3 grid of
// half-texel offsets (bilinear lookups), one thread per texel
__global__ void k_env_weights(EnvDev e, float *cdf_x)
{
  const int ix = blockIdx.x * blockDim.x + threadIdx.x;
  const int iy = blockIdx.y * blockDim.y + threadIdx.y;
  if (ix >= e.W || iy >= e.H) return;
  float weight = 0.f;
  for (int iiy = 0; iiy <= 2; iiy++)
    for (int iix = 0; iix <= 2; iix++) {
      const float4 t = env::tex(e, ((float)ix + (float)iix * .5f) / (float)e.W,
                                ((float)iy + (float)iiy * .5f) / (float)e.H);
      weight = fmx(weight, fmx(fmx(t.x, t.y), t.z));
    }
  cdf_x[(size_t)ix + (size_t)e.W * iy] = weight;
}

// each row's weights -> that row's CDF; the row's weight (sum * sin(theta))
// goes to cdf_y
__global__ void k_env_rows(float *cdf_x, float *cdf_y, int W, int H)
{
  const int y = blockIdx.x * blockDim.x + threadIdx.x;
  if (y >= H) return;
  float *line = cdf_x + (size_t)y * W;
  float sum = 0.f;
  for (int ix = 0; ix < W; ix++) sum += line[ix];
  const float rcp_sum = 1.f / sum;
  sum = 0.f;
  for (int ix = 0; ix < W; ix++) {
    sum += line[ix];
    line[ix] = sum * rcp_sum;
  }
  line[W - 1] = 1.f;
  const float rel_y = ((float)y + .5f) / (float)H;
  cdf_y[y] = sum * bnx_num::sin(env::PI * rel_y);
}

// row weights -> CDF over rows (one thread)
__global__ void k_env_norm_y(float *cdf_y, int H)
{
  if (blockIdx.x != 0 || threadIdx.x != 0) return;
  float sum = 0.f;
  for (int i = 0; i < H; i++) sum += cdf_y[i];
  const float rcp_sum = 1.f / sum;
  sum = 0.f;
  for (int i = 0; i < H; i++) {
    sum += cdf_y[i];
    cdf_y[i] = sum * rcp_sum;
  }
  cdf_y[H - 1] = 1.f;
}

} // namespace

EnvDev env_dev(const World *w)
{
  EnvDev e{};
  if (w->env_W > 0) {
    BNX_SET(e.tex, w->env_tex.as<float4>());
    BNX_SET(e.cdf_x, w->env_cdf_x.as<float>());
    BNX_SET(e.cdf_y, w->env_cdf_y.as<float>());
    e.W = w->env_W;
    e.H = w->env_H;
    e.scale = w->env_scale;
    for (int k = 0; k < 9; k++) {
      e.to_world[k] = w->env_to_world[k];
      e.to_local[k] = w->env_to_local[k];
    }
  }
  return e;
}

} // namespace bnx

using namespace bnx;

extern "C" int bnx_world_set_env_map(bnx_world wh, const bnx_env_map_desc *d)
{
  GUARD_BEGIN
  World *w = (World *)wh;
  BNX_CHECK(w, BNX_ERR_INVALID, "bnx_world_set_env_map: null world");
  Device *dev = w->dev;
  if (!d || !d->radiance) {
    BNX_HIP(hipStreamSynchronize(dev->stream));
    w->env_W = w->env_H = 0;
    w->env_tex.release();
    w->env_cdf_x.release();
    w->env_cdf_y.release();
    return BNX_OK;
  }
  BNX_CHECK(d->width >= 1 && d->height >= 1 && (int64_t)d->width * d->height <= (int64_t)1 << 28, BNX_ERR_INVALID,
            "bnx_world_set_env_map: width and height must be >= 1 (at most 2^28 texels)");
  const size_t n = (size_t)d->width * d->height;
  std::vector<float4> texels(n);
  for (size_t i = 0; i < n; i++)
    texels[i] = make_float4(d->radiance[3 * i], d->radiance[3 * i + 1], d->radiance[3 * i + 2], 0.f);
  BNX_HIP(hipStreamSynchronize(dev->stream));    // no frame in flight reads the old map
  w->env_W = w->env_H = 0;                        // no map until the new one is complete
  BNX_HIP(w->env_tex.ensure(sizeof(float4) * n));
  BNX_HIP(w->env_cdf_x.ensure(sizeof(float) * n));
  BNX_HIP(w->env_cdf_y.ensure(sizeof(float) * d->height));
  BNX_HIP(hipMemcpyAsync(w->env_tex.ptr, texels.data(), sizeof(float4) * n, hipMemcpyHostToDevice, dev->stream));
  w->env_scale = d->scale;
  bnx_num::env_frame(d->direction, d->up, w->env_to_world, w->env_to_local);
  w->env_W = d->width;
  w->env_H = d->height;
  const EnvDev e = env_dev(w);
  dim3 bs(16, 16), nb(nblk(d->width, 16), nblk(d->height, 16));
  k_env_weights<<<nb, bs, 0, dev->stream>>>(e, w->env_cdf_x.as<float>());
  k_env_rows<<<nblk(d->height, 64), 64, 0, dev->stream>>>(w->env_cdf_x.as<float>(), w->env_cdf_y.as<float>(),
                                                          d->width, d->height);
  k_env_norm_y<<<1, 1, 0, dev->stream>>>(w->env_cdf_y.as<float>(), d->height);
  hipError_t err = hipGetLastError();
  if (err == hipSuccess) err = hipStreamSynchronize(dev->stream);
  if (err != hipSuccess) w->env_W = w->env_H = 0;
  BNX_HIP(err);
  return BNX_OK;
  GUARD_END
}
