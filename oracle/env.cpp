// env.cpp -- CPU oracle of barney's HDRI environment light (EnvMapLight)
// with the CPU backend's 2D texture lookup.
//
// TEST INFRASTRUCTURE ONLY (see oracle.h).
#include "oracle.h"

namespace orc {

namespace {

// computeAddress (rtcore/cpu/Texture.cpp:24-89) for WRAP and CLAMP
void address(bool wrap, float tc, int N, int &i0, int &i1, float &f)
{
  if (wrap) {
    tc = tc * (float)N - .5f;
    tc = fmodf(tc, (float)N);
    if (tc < 0.f) tc += (float)N;
    const float fc = floorf(tc);
    const int ic = (int)fc;
    f = tc - (float)ic;
    // a tiny negative tc wraps to tc + N == N after rounding: that is texel
    // 0 (the reference's CPU backend would read one texel past the row)
    i0 = ic < N ? ic : ic - N;
    i1 = (int)fmodf(tc + 1.f, (float)N);
  } else {
    tc = tc * (float)N;
    if (tc - .5f <= 0.f) { i0 = i1 = 0; f = 0.f; }
    else if (tc - .5f >= (float)(N - 1)) { i0 = i1 = N - 1; f = 0.f; }
    else {
      f = (tc - .5f) - (float)(int)(tc - .5f);
      i0 = (int)(tc - .5f);
      i1 = i0 + 1;
    }
  }
}

inline vec3 xfm(const float m[9], vec3 v)
{
  // xfmVector(linear3f, v) = vx * v.x + vy * v.y + vz * v.z
  return vec3((m[0] * v.x + m[3] * v.y) + m[6] * v.z, (m[1] * v.x + m[4] * v.y) + m[7] * v.z,
              (m[2] * v.x + m[5] * v.y) + m[8] * v.z);
}

// sphericalTheta / sphericalPhi (render/DG.h:97-106)
inline float spherical_theta(vec3 v) { return bnx_num::acos(fmx(fmn(v.z, 1.f), -1.f)); }
inline float spherical_phi(vec3 v)
{
  float p = bnx_num::atan2(v.y, v.x);
  return (p < 0.f) ? (p + 2.f * ONE_PI) : p;
}

// cdfGetPDF / sampleCDF (light/EnvMap.h:98-133)
inline float cdf_pdf(int pos, const float *cdf, int N)
{
  const float at = cdf[pos];
  const float before = pos ? cdf[pos - 1] : 0.f;
  return (float)N * (at - before);
}
int sample_cdf(const float *cdf, int N, float v, float &pdf)
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

} // namespace

// tex2D of the HDRI texture: float4, LINEAR, WRAP in u, CLAMP in v
// (anari/light/HDRI.cpp:43-63, rtcore/cpu/Texture.cpp:294-321)
void env_tex(const EnvMap &e, float u, float v, float out[4])
{
  int x0, x1, y0, y1;
  float fx, fy;
  address(true, u, e.W, x0, x1, fx);
  address(false, v, e.H, y0, y1, fy);
  const float *t00 = &e.tex[4 * ((size_t)x0 + (size_t)y0 * e.W)];
  const float *t01 = &e.tex[4 * ((size_t)x1 + (size_t)y0 * e.W)];
  const float *t10 = &e.tex[4 * ((size_t)x0 + (size_t)y1 * e.W)];
  const float *t11 = &e.tex[4 * ((size_t)x1 + (size_t)y1 * e.W)];
  for (int k = 0; k < 4; k++) {
    const float v0 = (1.f - fx) * t00[k] + fx * t01[k];
    const float v1 = (1.f - fx) * t10[k] + fx * t11[k];
    out[k] = (1.f - fy) * v0 + fy * v1;
  }
}

// EnvMapLight::computeCDFs (light/EnvMap.cu:61-138, 183-239)
void env_build(EnvMap &e)
{
  const int W = e.W, H = e.H;
  e.cdf_x.assign((size_t)W * H, 0.f);
  e.cdf_y.assign((size_t)H, 0.f);
  // ComputeWeights_xy: max over a 3x3 half-texel neighbourhood
  for (int iy = 0; iy < H; iy++)
    for (int ix = 0; ix < W; ix++) {
      float weight = 0.f;
      for (int iiy = 0; iiy <= 2; iiy++)
        for (int iix = 0; iix <= 2; iix++) {
          float t[4];
          env_tex(e, ((float)ix + (float)iix * .5f) / (float)W, ((float)iy + (float)iiy * .5f) / (float)H, t);
          weight = fmx(weight, fmx(fmx(t[0], t[1]), t[2]));
        }
      e.cdf_x[(size_t)ix + (size_t)W * iy] = weight;
    }
  // ComputeCDFs_doLine: per-row CDF, row weight = sum * sin(theta)
  for (int y = 0; y < H; y++) {
    float *line = &e.cdf_x[(size_t)y * W];
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
    e.cdf_y[y] = sum * bnx_num::sin(ONE_PI * rel_y);
  }
  // Normalize_cdf_y
  float sum = 0.f;
  for (int i = 0; i < H; i++) sum += e.cdf_y[i];
  const float rcp_sum = 1.f / sum;
  sum = 0.f;
  for (int i = 0; i < H; i++) {
    sum += e.cdf_y[i];
    e.cdf_y[i] = sum * rcp_sum;
  }
  e.cdf_y[H - 1] = 1.f;
}

vec3 env_uv_to_world(const EnvMap &e, float fx, float fy)
{
  // EnvMapLight::DD::uvToWorld (light/EnvMap.h:172-182)
  const float phi = (2.f * ONE_PI) * fx;
  const float theta = ONE_PI * fy;
  float st, ct, sp, cp;
  bnx_num::sincos(theta, st, ct);
  bnx_num::sincos(phi, sp, cp);
  return xfm(e.to_world, vec3(cp * st, sp * st, ct));
}

float env_pdf(const EnvMap &e, vec3 dir)
{
  // EnvMapLight::DD::pdf + worldToPixel (light/EnvMap.h:135-160)
  if (!e.present) return 1.f / (4.f * ONE_PI);
  const vec3 local = xfm(e.to_local, dir);
  const float theta = spherical_theta(local);
  const float phi = spherical_phi(local);
  const float u = phi * (1.f / (2.f * ONE_PI)), v = theta * (1.f / ONE_PI);
  const int ix = (int)fmx(fmn(u * (float)e.W, (float)e.W - 1.f), 0.f);
  const int iy = (int)fmx(fmn(v * (float)e.H, (float)e.H - 1.f), 0.f);
  const float pdf_y = cdf_pdf(iy, e.cdf_y.data(), e.H);
  const float pdf_x = cdf_pdf(ix, &e.cdf_x[(size_t)iy * e.W], e.W);
  const float rel_y = ((float)iy + .5f) / (float)e.H;
  const float theta_c = ONE_PI * rel_y;
  return ((pdf_x * pdf_y) * 1.f) / (((2.f * ONE_PI) * ONE_PI) * bnx_num::sin(theta_c));
}

void env_sample(const EnvMap &e, Random2 &r, vec3 &direction, vec3 &radiance, float &pdf)
{
  // EnvMapLight::DD::sample (light/EnvMap.h:211-250)
  const float r_y = r();
  const float r_x = r();
  float pdf_y, pdf_x;
  const int iy = sample_cdf(e.cdf_y.data(), e.H, r_y, pdf_y);
  const int ix = sample_cdf(&e.cdf_x[(size_t)e.W * iy], e.W, r_x, pdf_x);
  const float sx = ((float)ix + r()) / (float)e.W;
  const float sy = ((float)iy + r()) / (float)e.H;
  float t[4];
  env_tex(e, sx, sy, t);
  radiance = e.scale * vec3(t[0], t[1], t[2]);
  direction = env_uv_to_world(e, sx, sy);
  const float theta = ONE_PI * sy;
  pdf = ((pdf_x * pdf_y) * 1.f) / (((2.f * ONE_PI) * ONE_PI) * bnx_num::sin(theta));
}

vec3 env_radiance(const EnvMap &e, vec3 ray_dir)
{
  // radianceFromEnv (kernels/shadeRays.cu:435-454)
  const vec3 d = xfm(e.to_local, normalize(ray_dir));
  const float theta = spherical_theta(d);
  const float phi = spherical_phi(d);
  float t[4];
  env_tex(e, phi * (1.f / (2.f * ONE_PI)), theta * (1.f / ONE_PI), t);
  return e.scale * vec3(t[0], t[1], t[2]);
}

} // namespace orc
