// bnx_env.h -- HDRI environment light on the device: the texture lookup,
// importance sampling from the row/column CDFs, its pdf and the radiance
// seen by rays that leave the scene.
//
// Restates EnvMapLight::DD (light/EnvMap.h:93-250), radianceFromEnv
// (kernels/shadeRays.cu:435-454) and the CPU backend's bilinear texture
// fetch (rtcore/cpu/Texture.cpp:24-89, 294-321: float4 texels, WRAP in u,
// CLAMP in v), operation by operation, so results match bit for bit under
// -ffp-contract=off.
#pragma once

#include "bnx_device.h"
#include "bnx_numeric.h"

namespace bnx {

// Device view of the world's env map; tex == null: no map (the renderer's
// ambient radiance lights the scene).
struct EnvDev {
  const float4 BNX_G *tex;     // W x H rgba texels, row-major, row 0 at theta = 0
  const float BNX_G *cdf_x;    // W per row: each row's CDF over u
  const float BNX_G *cdf_y;    // H: CDF over rows
  int W, H;
  float scale;
  float to_world[9];           // columns vx, vy, vz
  float to_local[9];
  int pad;
};

namespace env {

constexpr float PI = 3.14159265358979323846f;
constexpr float TWO_PI = 2.f * PI;

BNX_D v3 xfm(const float *m, v3 v)
{
  return v3{(m[0] * v.x + m[3] * v.y) + m[6] * v.z, (m[1] * v.x + m[4] * v.y) + m[7] * v.z,
            (m[2] * v.x + m[5] * v.y) + m[8] * v.z};
}

// texel pair and weight along one axis; wrap: u, clamp: v
BNX_D void addr_wrap(float tc, int N, int &i0, int &i1, float &f)
{
  const float fN = (float)N;
  tc = tc * fN - .5f;
  tc = fmodf(tc, fN);
  if (tc < 0.f) tc += fN;
  const int ic = (int)floorf(tc);
  f = tc - (float)ic;
  // a tiny negative tc wraps to exactly N after rounding: texel 0
  i0 = ic < N ? ic : ic - N;
  i1 = (int)fmodf(tc + 1.f, fN);
}
BNX_D void addr_clamp(float tc, int N, int &i0, int &i1, float &f)
{
  tc = tc * (float)N;
  const float t = tc - .5f;
  if (t <= 0.f) { i0 = i1 = 0; f = 0.f; }
  else if (t >= (float)(N - 1)) { i0 = i1 = N - 1; f = 0.f; }
  else {
    i0 = (int)t;
    f = t - (float)i0;
    i1 = i0 + 1;
  }
}

BNX_D float4 lerp4(float f, float4 a, float4 b)
{
  const float g = 1.f - f;
  return make_float4(g * a.x + f * b.x, g * a.y + f * b.y, g * a.z + f * b.z, g * a.w + f * b.w);
}

BNX_D float4 tex(const EnvDev &e, float u, float v)
{
  int x0, x1, y0, y1;
  float fx, fy;
  addr_wrap(u, e.W, x0, x1, fx);
  addr_clamp(v, e.H, y0, y1, fy);
  const float4 BNX_G *r0 = e.tex + (size_t)y0 * e.W;
  const float4 BNX_G *r1 = e.tex + (size_t)y1 * e.W;
  return lerp4(fy, lerp4(fx, r0[x0], r0[x1]), lerp4(fx, r1[x0], r1[x1]));
}

BNX_D float theta_of(v3 d) { return bnx_num::acos(fmx(fmn(d.z, 1.f), -1.f)); }
BNX_D float phi_of(v3 d)
{
  const float p = bnx_num::atan2(d.y, d.x);
  return (p < 0.f) ? (p + TWO_PI) : p;
}

BNX_D float cdf_pdf(int pos, const float BNX_G *cdf, int N)
{
  const float at = cdf[pos];
  const float before = pos ? cdf[pos - 1] : 0.f;
  return (float)N * (at - before);
}

BNX_D int sample_cdf(const float BNX_G *cdf, int N, float v, float &pdf)
{
  int begin = 0, end = N;
  while ((end - begin) > 1) {
    const int mid = (begin + end) / 2;
    if (v < cdf[mid - 1]) end = mid;
    else begin = mid;
  }
  pdf = cdf_pdf(begin, cdf, N);
  return begin;
}

} // namespace env

// pdf of the light sampling direction 'dir' (world space)
BNX_D float env_pdf(const EnvDev &e, v3 dir)
{
  if (!e.tex) return 1.f / (4.f * env::PI);
  const v3 local = env::xfm(e.to_local, dir);
  const float theta = env::theta_of(local);
  const float phi = env::phi_of(local);
  const float u = phi * (1.f / env::TWO_PI), v = theta * (1.f / env::PI);
  const int ix = (int)fmx(fmn(u * (float)e.W, (float)e.W - 1.f), 0.f);
  const int iy = (int)fmx(fmn(v * (float)e.H, (float)e.H - 1.f), 0.f);
  const float pdf_y = env::cdf_pdf(iy, e.cdf_y, e.H);
  const float pdf_x = env::cdf_pdf(ix, e.cdf_x + (size_t)iy * e.W, e.W);
  const float theta_c = env::PI * (((float)iy + .5f) / (float)e.H);
  return ((pdf_x * pdf_y) * 1.f) / ((env::TWO_PI * env::PI) * bnx_num::sin(theta_c));
}

// one light sample: row from cdf_y, column from that row's CDF, a uniform
// point inside the texel (four draws of r)
BNX_D void env_sample(const EnvDev &e, Random2 &r, v3 &direction, v3 &radiance, float &pdf)
{
  const float r_y = r();
  const float r_x = r();
  float pdf_y, pdf_x;
  const int iy = env::sample_cdf(e.cdf_y, e.H, r_y, pdf_y);
  const int ix = env::sample_cdf(e.cdf_x + (size_t)e.W * iy, e.W, r_x, pdf_x);
  const float sx = ((float)ix + r()) / (float)e.W;
  const float sy = ((float)iy + r()) / (float)e.H;
  const float4 t = env::tex(e, sx, sy);
  radiance = e.scale * v3{t.x, t.y, t.z};
  const float phi = env::TWO_PI * sx;
  const float theta = env::PI * sy;
  float st, ct, sp, cp;
  bnx_num::sincos(theta, st, ct);
  bnx_num::sincos(phi, sp, cp);
  direction = env::xfm(e.to_world, v3{cp * st, sp * st, ct});
  pdf = ((pdf_x * pdf_y) * 1.f) / ((env::TWO_PI * env::PI) * st);
}

// radiance arriving along a ray that leaves the scene in direction ray_dir
BNX_D v3 env_radiance(const EnvDev &e, v3 ray_dir)
{
  const v3 d = env::xfm(e.to_local, normalize(ray_dir));
  const float theta = env::theta_of(d);
  const float phi = env::phi_of(d);
  const float4 t = env::tex(e, phi * (1.f / env::TWO_PI), theta * (1.f / env::PI));
  return e.scale * v3{t.x, t.y, t.z};
}

} // namespace bnx
