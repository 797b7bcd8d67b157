// bnx_pbr.h -- the physically-based material on the device: AnariPBR
// (material/AnariPBR.h:62-96) creating the NVisii Disney BSDF
// (packedBSDF/NVisii.h) with its eval, pdf and scatter, as the path tracer
// calls them (kernels/shadeRays.cu:680-744, 757-879).
//
// Written operation by operation in the order of the reference so that,
// under -ffp-contract=off, every result matches a scalar CPU evaluation bit
// for bit.  AnariPBR fixes anisotropy, sheen, clearcoat, specular
// transmission and flatness at zero (NVisii::setDefaults, NVisii.h:960-1003):
// the terms they scale are exact zeros and are left out, except the
// subsurface term, which can carry a NaN into the flatness blend.  The
// clearcoat lobe is never sampled (after the glossy test fails,
// (r - w_d) - w_g >= 0 = w_c).  schlick_weight evaluates powf(x, 5) as the
// file's own pow5 (pow2(pow2(x)) * x), and a lobe sample draws its two
// random numbers x first, then y.
#pragma once

#include "bnx_device.h"
#include "bnx_numeric.h"

namespace bnx {
namespace pbr {

constexpr float PI = 3.14159265358979323846f;
constexpr float ONE_OVER_PI = (float)0.318309886183790671538;   // (float)M_1_PI
constexpr float MIN_ALPHA = .002f;
constexpr float SMALL_EPSILON = 2e-10f;

// NVisii parameters of a hit (fp16-rounded as the packed BSDF stores them)
struct Params {
  v3 base;
  float metallic, roughness, ior;
};

// constant parameters of an AnariPBR material
struct Material {
  float metallic, roughness, ior;
};

BNX_D float h16(float f) { return h2f(f2h(f)); }

BNX_D Params params(const v3 &base, float metallic, float roughness, float ior)
{
  Params p;
  p.base = quant_half(base);
  p.metallic = h16(metallic);
  p.roughness = h16(fmn(fmx(roughness, .001f), 1.f - .1f));
  p.ior = h16(ior);
  return p;
}

BNX_D float pow2(float f) { return f * f; }
BNX_D float saturate(float f) { return fmx(0.f, fmn(f, 1.f)); }
BNX_D float lerp_r(float a, float b, float f) { return (1.f - f) * a + f * b; }
BNX_D v3 lerp_r(v3 a, v3 b, float f)
{
  return v3{(1.f - f) * a.x + f * b.x, (1.f - f) * a.y + f * b.y, (1.f - f) * a.z + f * b.z};
}
BNX_D float luminance(v3 c) { return (0.212671f * c.x + 0.715160f * c.y) + 0.072169f * c.z; }
BNX_D v3 reflect(v3 v, v3 n) { return v - (2.f * dot(v, n)) * n; }
BNX_D v3 refract(v3 v, v3 n, float eta)
{
  const float dv = dot(n, v);
  const float k = 1.f - (eta * eta) * (1.f - dv * dv);
  if (!(k >= 0.f)) return splat(0.f);
  const float a = eta * dv;
  const float sk = sqrtf(k);
  return eta * v - v3{a + sk * n.x, a + sk * n.y, a + sk * n.z};
}
BNX_D bool same_hemisphere(v3 wo, v3 wi, v3 n) { return dot(wo, n) * dot(wi, n) > 0.f; }
BNX_D float schlick_weight(float c)
{
  const float x = saturate(1.f - c);
  return pow2(pow2(x)) * x;
}
BNX_D float gtr_2(float cos_h, float alpha)
{
  const float a2 = alpha * alpha;
  const float den1 = 1.f + ((a2 - 1.f) * cos_h) * cos_h;
  const float den2 = fmx(den1, SMALL_EPSILON);
  return (ONE_OVER_PI * 1.f) / den2;
}
BNX_D float smith_ggx(float n_dot_o, float alpha_g)
{
  const float a = alpha_g * alpha_g;
  const float b = n_dot_o * n_dot_o;
  return 1.f / (n_dot_o + sqrtf((a + b) - a * b));
}
BNX_D v3 to_frame(v3 h, v3 vx, v3 vy, v3 n) { return (h.x * vx + h.y * vy) + h.z * n; }

// concentric-disk cosine sample (NVisii.h:68-86)
BNX_D v3 cos_sample_hemisphere(float ux, float uy)
{
  const float sx = 2.f * ux - 1.f, sy = 2.f * uy - 1.f;
  float radius = 0.f, theta = 0.f;
  if (!(sx == 0.f && sy == 0.f)) {
    if (fabsf(sx) > fabsf(sy)) {
      radius = sx;
      theta = (PI / 4.f) * (sy / sx);
    } else {
      radius = sy;
      theta = (PI / 2.f) - (PI / 4.f) * (sx / sy);
    }
  }
  float st, ct;
  bnx_num::sincos(theta, st, ct);
  const float dx = radius * ct, dy = radius * st;
  return v3{dx, dy, sqrtf(fmx(0.f, (1.f - dx * dx) - dy * dy))};
}

// GGX half-vector sample (NVisii.h:180-188); the "sine" is 1 - cos^2
BNX_D v3 sample_gtr_2_h(v3 n, v3 vx, v3 vy, float alpha, float ux, float uy)
{
  const float phi_h = (2.f * PI) * ux;
  const float c2 = (1.f - uy) / (1.f + (alpha * alpha - 1.f) * uy);
  const float c = sqrtf(c2);
  const float s = 1.f - c2;
  float sp, cp;
  bnx_num::sincos(phi_h, sp, cp);
  return to_frame(normalize(v3{s * cp, s * sp, c}), vx, vy, n);
}

// disney_microfacet_isotropic (NVisii.h:463-490), specular 0.5, tint 0
BNX_D v3 microfacet(const Params &p, v3 n, v3 wo, v3 wi, v3 wh)
{
  const float lum = luminance(p.base);
  const v3 tint = lum > 0.f ? p.base / lum : splat(1.f);
  const v3 spec = lerp_r((.5f * 0.08f) * lerp_r(splat(1.f), tint, 0.f), p.base, p.metallic);
  const float alpha = fmx(MIN_ALPHA, p.roughness * p.roughness);
  const float d = gtr_2(fabsf(dot(n, wh)), alpha);
  const float fw = schlick_weight(fabsf(dot(wi, wh))) * lerp_r(.5f, 1.f, fmx(p.metallic, alpha));
  const v3 f = lerp_r(spec, splat(1.f), fw);
  const float g = smith_ggx(fabsf(dot(n, wi)), alpha) * smith_ggx(fabsf(dot(n, wo)), alpha);
  return (d * f) * g;
}

// disney_brdf (NVisii.h:639-690) without the transmission branch
BNX_D v3 brdf(const Params &p, v3 n, v3 wo, v3 wi, v3 wh)
{
  const float n_dot_o = fabsf(dot(wo, n)), n_dot_i = fabsf(dot(wi, n));
  const float i_dot_h = dot(wi, wh);
  const float fd90 = 0.5f + ((1.0f * p.roughness) * i_dot_h) * i_dot_h;
  const float fi = schlick_weight(n_dot_i), fo = schlick_weight(n_dot_o);
  const float diffuse = lerp_r(1.f, fd90, fi) * lerp_r(1.f, fd90, fo);
  const float fss90 = (i_dot_h * i_dot_h) * p.roughness;
  const float fss = lerp_r(1.0f, fss90, fi) * lerp_r(1.0f, fss90, fo);
  const float ss = 1.25f * (fss * (1.f / (n_dot_i + n_dot_o) - .5f) + .5f);
  const float subsurface = ONE_OVER_PI * ss;
  const v3 gloss = microfacet(p, n, wo, wi, wh);
  const v3 a = diffuse * p.base, b = subsurface * splat(h16(.8f));
  const v3 flat = lerp_r(a, b, 0.f);
  v3 r = (flat * (1.f - p.metallic)) * (1.f - 0.f);
  r = r + splat(0.f);
  r = r + splat(0.f);
  return r + gloss;
}

// disney_pdf (NVisii.h:705-810)
BNX_D float pdf(const Params &p, v3 n, v3 wo, v3 wi, v3 wh)
{
  const float dw = (1.f - p.metallic) * (1.f - 0.f);
  const float gw = (((1.f + p.metallic) + 0.f) + .5f) + p.roughness;
  const float alpha = fmx(MIN_ALPHA, p.roughness * p.roughness);
  const float dd = dot(wi, n);
  const float diffuse = dd > 0.f ? dd * ONE_OVER_PI : 0.f;
  float mf = 0.f;
  if (same_hemisphere(wo, wi, n)) {
    const float cos_h = fabsf(dot(n, wh));
    const float d = gtr_2(cos_h, alpha);
    mf = (d * cos_h) / (4.f * fabsf(dot(wo, wh)));
  }
  return dw * diffuse + gw * mf;
}

} // namespace pbr

// NVisii::eval (NVisii.h:1176-1203): Ng un-flipped, wo = -ray direction
BNX_D void pbr_eval(const pbr::Params &p, v3 Ng, v3 wo, v3 wi, v3 &value, float &pdf)
{
  const v3 wh = normalize(wi + wo);
  value = pbr::brdf(p, Ng, wo, wi, wh);
  pdf = pbr::pdf(p, Ng, wo, wi, wh);
}

BNX_D float pbr_pdf(const pbr::Params &p, v3 Ng, v3 wo, v3 wi) { return pbr::pdf(p, Ng, wo, wi, normalize(wi + wo)); }

// NVisii::scatter + sample_disney_brdf (NVisii.h:825-924, 1112-1174);
// returns true for the diffuse lobe
BNX_D bool pbr_scatter(const pbr::Params &p, v3 Ng, v3 wo, Random2 &rng, v3 &dir, float &spdf, v3 &f_r)
{
  const v3 n = Ng;
  const v3 vy = normalize(cross(Ng, wo));
  const v3 vx = normalize(cross(Ng, vy));
  float dw = (1.f - p.metallic) * (1.f - 0.f);
  float gw = (((1.f + p.metallic) + 0.f) + .5f) + p.roughness;
  const float sum = ((dw + gw) + 0.f) + 0.f;
  spdf = 0.f;
  const float scale = 1.f / sum;
  dw *= scale;
  gw *= scale;
  const float type_rng = rng();
  const int lobe = (type_rng < dw) ? 0 : ((type_rng - dw) < gw ? 1 : 3);
  const float ux = rng();
  const float uy = rng();
  v3 wi;
  if (lobe == 0) {
    wi = pbr::to_frame(normalize(pbr::cos_sample_hemisphere(ux, uy)), vx, vy, n);
  } else if (lobe == 1) {
    const float alpha = fmx(pbr::MIN_ALPHA, p.roughness * p.roughness);
    wi = pbr::reflect(-wo, pbr::sample_gtr_2_h(n, vx, vy, alpha, ux, uy));
    if (!pbr::same_hemisphere(wo, wi, n)) return false;
  } else {
    // specular transmission lobe: reached only when rounding leaves
    // dw + gw just below 1
    const float tr = pbr::h16(.04f);
    const float alpha = fmx(pbr::MIN_ALPHA, tr * tr);
    const v3 wh = pbr::sample_gtr_2_h(n, vx, vy, alpha, ux, uy);
    const bool entering = dot(wo, wh) > 0.f;
    const float eta_i = entering ? 1.f : p.ior;
    const float eta_o = entering ? p.ior : 1.f;
    const v3 hh = entering ? wh : -wh;
    wi = pbr::refract(-wo, hh, eta_i / eta_o);
    if (wi.x == 0.f && wi.y == 0.f && wi.z == 0.f) {
      wi = pbr::reflect(-wo, hh);
      spdf = 1.f;
      f_r = splat(1.f);
      dir = normalize(wi);
      return false;
    }
  }
  const v3 wh = normalize(wi + wo);
  spdf = pbr::pdf(p, n, wo, wi, wh);
  f_r = pbr::brdf(p, n, wo, wi, wh);
  dir = normalize(wi);
  return lobe == 0;
}

} // namespace bnx
