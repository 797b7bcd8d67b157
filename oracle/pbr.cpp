// pbr.cpp -- CPU oracle of barney's physically-based material: AnariPBR
// (material/AnariPBR.h:62-96) creating the NVisii Disney BSDF
// (packedBSDF/NVisii.h) and its eval / pdf / scatter as the path tracer
// calls them.
//
// TEST INFRASTRUCTURE ONLY (see oracle.h).
//
// AnariPBR fixes anisotropy, sheen, clearcoat, specular transmission and
// flatness at zero (NVisii::setDefaults, NVisii.h:960-1003).  The terms
// they scale are then exactly zero (a zero factor times a finite value),
// except where a NaN or infinity could pass through; those terms are kept.
// The clearcoat lobe is never sampled: after the glossy test fails,
// (r - w_d) - w_g >= 0 = w_c.  schlick_weight restates powf(x, 5) as the
// file's own pow5 (pow2(pow2(x)) * x); the two random numbers of a lobe
// sample are drawn x first, then y (the reference passes both draws as
// constructor arguments, whose order C++ leaves unspecified).
#include "oracle.h"

namespace orc {

namespace {

const float M_1_PI_F = (float)M_1_PI;
const float MIN_ALPHA = .002f;
const float SMALL_EPSILON = 2e-10f;

inline float pow2(float f) { return f * f; }
inline float saturate(float f) { return fmx(0.f, fmn(f, 1.f)); }
inline float lerp_r(float a, float b, float f) { return (1.f - f) * a + f * b; }
inline vec3 lerp_r(vec3 a, vec3 b, float f)
{
  return vec3((1.f - f) * a.x + f * b.x, (1.f - f) * a.y + f * b.y, (1.f - f) * a.z + f * b.z);
}
inline float luminance(vec3 c) { return (0.212671f * c.x + 0.715160f * c.y) + 0.072169f * c.z; }
inline vec3 reflect(vec3 v, vec3 n) { return v - (2.f * dot(v, n)) * n; }
inline vec3 refract(vec3 v, vec3 n, float eta)
{
  const float dv = dot(n, v);
  const float k = 1.f - (eta * eta) * (1.f - dv * dv);
  if (!(k >= 0.f)) return vec3(0.f);
  const float a = eta * dv;
  const float sk = sqrtf(k);
  return eta * v - vec3(a + sk * n.x, a + sk * n.y, a + sk * n.z);
}
inline bool same_hemisphere(vec3 wo, vec3 wi, vec3 n) { return dot(wo, n) * dot(wi, n) > 0.f; }

// schlick_weight (NVisii.h:103-106)
inline float schlick_weight(float c)
{
  const float x = saturate(1.f - c);
  return pow2(pow2(x)) * x;
}

// gtr_2 (NVisii.h:131-140)
inline float gtr_2(float cos_h, float alpha)
{
  const float a2 = alpha * alpha;
  const float den1 = 1.f + ((a2 - 1.f) * cos_h) * cos_h;
  const float den2 = fmx(den1, SMALL_EPSILON);
  return (M_1_PI_F * 1.f) / den2;
}

// smith_shadowing_ggx (NVisii.h:149-154)
inline float smith_ggx(float n_dot_o, float alpha_g)
{
  const float a = alpha_g * alpha_g;
  const float b = n_dot_o * n_dot_o;
  return 1.f / (n_dot_o + sqrtf((a + b) - a * b));
}

// cos_sample_hemisphere (NVisii.h:68-86): concentric disk, lifted
vec3 cos_sample_hemisphere(float ux, float uy)
{
  const float sx = 2.f * ux - 1.f, sy = 2.f * uy - 1.f;
  float radius = 0.f, theta = 0.f;
  if (!(sx == 0.f && sy == 0.f)) {
    if (fabsf(sx) > fabsf(sy)) {
      radius = sx;
      theta = (ONE_PI / 4.f) * (sy / sx);
    } else {
      radius = sy;
      theta = (ONE_PI / 2.f) - (ONE_PI / 4.f) * (sx / sy);
    }
  }
  float st, ct;
  bnx_num::sincos(theta, st, ct);
  const float dx = radius * ct, dy = radius * st;
  return vec3(dx, dy, sqrtf(fmx(0.f, (1.f - dx * dx) - dy * dy)));
}

inline vec3 to_frame(vec3 h, vec3 vx, vec3 vy, vec3 n) { return (h.x * vx + h.y * vy) + h.z * n; }

// sample_gtr_2_h (NVisii.h:180-188)
vec3 sample_gtr_2_h(vec3 n, vec3 vx, vec3 vy, float alpha, float ux, float uy)
{
  const float phi_h = (2.f * ONE_PI) * ux;
  const float c2 = (1.f - uy) / (1.f + (alpha * alpha - 1.f) * uy);
  const float c = sqrtf(c2);
  const float s = 1.f - c2;
  float sp, cp;
  bnx_num::sincos(phi_h, sp, cp);
  const vec3 hemi = normalize(vec3(s * cp, s * sp, c));
  return to_frame(hemi, vx, vy, n);
}

// the material with its fp16-packed parameters unpacked (NVisii.h:926-946)
struct Disney {
  vec3 base, subsurface;
  float metallic, specular, roughness, specular_tint, sheen, sheen_tint, clearcoat, clearcoat_gloss;
  float ior, specular_transmission, transmission_roughness, flatness;
};

inline float h16(float f) { return f16_to_f32(f32_to_f16(f)); }

Disney unpack(const PbrParams &p)
{
  Disney m;
  m.base = p.base;
  m.subsurface = vec3(h16(.8f));
  m.metallic = p.metallic;
  m.specular = h16(.5f);
  m.roughness = p.roughness;
  m.specular_tint = 0.f;
  m.sheen = 0.f;
  m.sheen_tint = h16(.5f);
  m.clearcoat = 0.f;
  m.clearcoat_gloss = h16(1.f - .03f * .03f);
  m.ior = p.ior;
  m.specular_transmission = 0.f;
  m.transmission_roughness = h16(.04f);
  m.flatness = 0.f;
  return m;
}

// disney_microfacet_isotropic (NVisii.h:463-490)
vec3 microfacet_isotropic(const Disney &m, vec3 n, vec3 wo, vec3 wi, vec3 wh)
{
  const float lum = luminance(m.base);
  const vec3 tint = lum > 0.f ? m.base / lum : vec3(1.f);
  const vec3 spec = lerp_r((m.specular * 0.08f) * lerp_r(vec3(1.f), tint, m.specular_tint), m.base, m.metallic);
  const float alpha = fmx(MIN_ALPHA, m.roughness * m.roughness);
  const float d = gtr_2(fabsf(dot(n, wh)), alpha);
  const float fw = schlick_weight(fabsf(dot(wi, wh))) * lerp_r(.5f, 1.f, fmx(m.metallic, alpha));
  const vec3 f = lerp_r(spec, vec3(1.f), fw);
  const float g = smith_ggx(fabsf(dot(n, wi)), alpha) * smith_ggx(fabsf(dot(n, wo)), alpha);
  return (d * f) * g;
}

// disney_brdf (NVisii.h:639-690) without the transmission branch
vec3 disney_brdf(const Disney &m, vec3 n, vec3 wo, vec3 wi, vec3 wh)
{
  // disney_diffuse (:330-343)
  const float n_dot_o = fabsf(dot(wo, n)), n_dot_i = fabsf(dot(wi, n));
  const float i_dot_h = dot(wi, wh);
  const float fd90 = 0.5f + ((1.0f * m.roughness) * i_dot_h) * i_dot_h;
  const float fi = schlick_weight(n_dot_i), fo = schlick_weight(n_dot_o);
  const float diffuse = lerp_r(1.f, fd90, fi) * lerp_r(1.f, fd90, fo);
  // disney_subsurface (:346-358)
  const float fss90 = (i_dot_h * i_dot_h) * m.roughness;
  const float fss = lerp_r(1.0f, fss90, fi) * lerp_r(1.0f, fss90, fo);
  const float ss = 1.25f * (fss * (1.f / (n_dot_i + n_dot_o) - .5f) + .5f);
  const float subsurface = M_1_PI_F * ss;
  const vec3 gloss = microfacet_isotropic(m, n, wo, wi, wh);
  const vec3 a = diffuse * m.base, b = subsurface * m.subsurface;
  const vec3 flat = lerp_r(a, b, m.flatness);
  vec3 r = (flat * (1.f - m.metallic)) * (1.f - m.specular_transmission);
  r = r + vec3(0.f);   // sheen
  r = r + vec3(0.f);   // clearcoat
  return r + gloss;
}

// disney_pdf (NVisii.h:705-810)
float disney_pdf(const Disney &m, vec3 n, vec3 wo, vec3 wi, vec3 wh)
{
  const float dw = (1.f - m.metallic) * (1.f - m.specular_transmission);
  const float gw = (((1.f + m.metallic) + m.sheen) + m.specular) + m.roughness;
  const float sum = ((dw + gw) + m.clearcoat) + 0.f;
  if (sum == 0.f) return 0.f;
  const float alpha = fmx(MIN_ALPHA, m.roughness * m.roughness);
  const float dd = dot(wi, n);
  const float diffuse = dd > 0.f ? dd * M_1_PI_F : 0.f;
  float microfacet = 0.f;
  if (same_hemisphere(wo, wi, n)) {
    const float cos_h = fabsf(dot(n, wh));
    const float d = gtr_2(cos_h, alpha);
    microfacet = (d * cos_h) / (4.f * fabsf(dot(wo, wh)));
  }
  return dw * diffuse + gw * microfacet;
}

} // namespace

PbrParams pbr_params(const Material &mat, const float base[4])
{
  // AnariPBR::DD::createBSDF (AnariPBR.h:62-96) packing into NVisii's halves
  PbrParams p;
  p.base = quantize_half(vec3(base[0], base[1], base[2]));
  p.metallic = h16(mat.metallic);
  p.roughness = h16(fmn(fmx(mat.roughness, .001f), 1.f - .1f));
  p.ior = h16(mat.ior);
  return p;
}

// NVisii::eval (NVisii.h:1176-1203)
void pbr_eval(const PbrParams &p, vec3 Ng, vec3 wo, vec3 wi, vec3 &value, float &pdf)
{
  const Disney m = unpack(p);
  const vec3 wh = normalize(wi + wo);
  value = disney_brdf(m, Ng, wo, wi, wh);
  pdf = disney_pdf(m, Ng, wo, wi, wh);
}

// NVisii::pdf (NVisii.h:1205-1224)
float pbr_pdf(const PbrParams &p, vec3 Ng, vec3 wo, vec3 wi)
{
  const Disney m = unpack(p);
  return disney_pdf(m, Ng, wo, wi, normalize(wi + wo));
}

// NVisii::scatter + sample_disney_brdf (NVisii.h:825-924, 1112-1174)
bool pbr_scatter(const PbrParams &p, vec3 Ng, vec3 wo, Random2 &rng, vec3 &dir, float &pdf, vec3 &f_r)
{
  const Disney m = unpack(p);
  const vec3 n = Ng;
  const vec3 vy = normalize(cross(Ng, wo));
  const vec3 vx = normalize(cross(Ng, vy));
  float dw = (1.f - m.metallic) * (1.f - m.specular_transmission);
  float gw = (((1.f + m.metallic) + m.sheen) + m.specular) + m.roughness;
  const float sum = ((dw + gw) + m.clearcoat) + 0.f;
  bool diffuse_lobe = false;
  pdf = 0.f;
  if (sum == 0.f) return false;
  const float scale = 1.f / sum;
  dw *= scale;
  gw *= scale;
  const float type_rng = rng();
  int lobe;                               // 0 diffuse, 1 glossy, 3 transmission
  if (type_rng < dw) lobe = 0;
  else if ((type_rng - dw) < gw) lobe = 1;
  else lobe = 3;
  const float ux = rng();
  const float uy = rng();
  vec3 wi;
  if (lobe == 0) {
    wi = to_frame(normalize(cos_sample_hemisphere(ux, uy)), vx, vy, n);
    diffuse_lobe = true;
  } else if (lobe == 1) {
    const float alpha = fmx(MIN_ALPHA, m.roughness * m.roughness);
    const vec3 wh = sample_gtr_2_h(n, vx, vy, alpha, ux, uy);
    wi = reflect(-wo, wh);
    if (!same_hemisphere(wo, wi, n)) {
      pdf = 0.f;
      return false;
    }
  } else {
    const float alpha = fmx(MIN_ALPHA, m.transmission_roughness * m.transmission_roughness);
    const vec3 wh = sample_gtr_2_h(n, vx, vy, alpha, ux, uy);
    const bool entering = dot(wo, wh) > 0.f;
    const float eta_i = entering ? 1.f : m.ior;
    const float eta_o = entering ? m.ior : 1.f;
    const vec3 hh = entering ? wh : -wh;
    wi = refract(-wo, hh, eta_i / eta_o);
    if (wi.x == 0.f && wi.y == 0.f && wi.z == 0.f) {
      wi = reflect(-wo, hh);
      pdf = 1.f;
      f_r = vec3(1.f);
      dir = normalize(wi);
      return false;
    }
  }
  const vec3 wh = normalize(wi + wo);
  pdf = disney_pdf(m, n, wo, wi, wh);
  f_r = disney_brdf(m, n, wo, wi, wh);
  dir = normalize(wi);
  return diffuse_lobe;
}

} // namespace orc
