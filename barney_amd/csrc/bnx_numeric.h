// bnx_numeric.h -- transcendental functions with a fixed float32 operation
// sequence, compiled identically for gfx950 device code and x86-64 host code.
//
// Why: the reference reseeds its per-ray RNG from the raw bits of hit
// distances (kernels/shadeRays.cu:608,720), so every float that feeds a
// later hit distance has to be bit-identical between the GPU path and the
// CPU oracle.  Library sinf/cosf differ in the last ulp between the GPU math
// library and glibc; these functions are the one definition both sides use.
//
// Method: Cody & Waite argument reduction (W. J. Cody, W. Waite, "Software
// Manual for the Elementary Functions", 1980) by k*pi/2 with pi/2 split into
// three float32 parts whose leading two have 12 significant bits, so k*P1 and
// k*P2 are exact for |k| < 2^12 (|x| < ~6400).  On the reduced interval
// |r| <= pi/4 the Taylor series of sin (to r^11) and cos (to r^12) are
// evaluated in Horner form; the truncation error is below 1e-11, so the
// result is within a few ulp of the correctly rounded value (checked against
// libm by the tests).  Compile with -ffp-contract=off on both sides: no
// fused multiply-add may change the rounding of these steps.
#pragma once

#if defined(__HIPCC__)
#  define BNX_HD __host__ __device__
#else
#  define BNX_HD
#endif

#include <math.h>
#include <stdint.h>

namespace bnx_num {

// pi/2 = P1 + P2 + P3 (P1, P2: 12 significant bits; P3 rounded)
static constexpr float PIO2_1 = 0x1.92p+0f;          // 1.5703125
static constexpr float PIO2_2 = 0x1.fb4p-12f;        // 4.8375129699707031e-04
static constexpr float PIO2_3 = 0x1.4442d2p-24f;     // 7.5497901264043322e-08
static constexpr float TWO_OVER_PI = 0x1.45f306p-1f; // 0.63661974668502808

// 1/n! rounded to float32
static constexpr float INV_F2  = 0x1p-1f;
static constexpr float INV_F3  = 0x1.555556p-3f;
static constexpr float INV_F4  = 0x1.555556p-5f;
static constexpr float INV_F5  = 0x1.111112p-7f;
static constexpr float INV_F6  = 0x1.6c16c2p-10f;
static constexpr float INV_F7  = 0x1.a01a02p-13f;
static constexpr float INV_F8  = 0x1.a01a02p-16f;
static constexpr float INV_F9  = 0x1.71de3ap-19f;
static constexpr float INV_F10 = 0x1.27e4fcp-22f;
static constexpr float INV_F11 = 0x1.ae6456p-26f;
static constexpr float INV_F12 = 0x1.1eed8ep-29f;

// sin on |r| <= pi/4 (Taylor to r^11, Horner in r^2)
BNX_HD inline float sin_kernel(float r)
{
  float r2 = r * r;
  float p = -INV_F11;
  p = p * r2;  p = p + INV_F9;
  p = p * r2;  p = p - INV_F7;
  p = p * r2;  p = p + INV_F5;
  p = p * r2;  p = p - INV_F3;
  p = p * r2;
  float rp = r * p;
  return r + rp;
}

// cos on |r| <= pi/4 (Taylor to r^12, Horner in r^2)
BNX_HD inline float cos_kernel(float r)
{
  float r2 = r * r;
  float p = INV_F12;
  p = p * r2;  p = p - INV_F10;
  p = p * r2;  p = p + INV_F8;
  p = p * r2;  p = p - INV_F6;
  p = p * r2;  p = p + INV_F4;
  p = p * r2;  p = p - INV_F2;
  p = p * r2;
  return 1.0f + p;
}

// x = k*pi/2 + r, |r| <= ~pi/4; returns r and k (quadrant)
BNX_HD inline float reduce_pio2(float x, int &k)
{
  float fk = rintf(x * TWO_OVER_PI);
  k = (int)fk;
  float r = x - fk * PIO2_1;
  r = r - fk * PIO2_2;
  r = r - fk * PIO2_3;
  return r;
}

BNX_HD inline void sincos(float x, float &s, float &c)
{
  int k;
  float r = reduce_pio2(x, k);
  float sr = sin_kernel(r);
  float cr = cos_kernel(r);
  switch (k & 3) {
  case 0:  s = sr;  c = cr;  break;
  case 1:  s = cr;  c = -sr; break;
  case 2:  s = -sr; c = -cr; break;
  default: s = -cr; c = sr;  break;
  }
}

BNX_HD inline float sin(float x) { float s, c; sincos(x, s, c); return s; }
BNX_HD inline float cos(float x) { float s, c; sincos(x, s, c); return c; }

// ------------------------------------------------------------- atan, acos
// atan: reduce |x| > 1 by atan(x) = pi/2 - atan(1/x), then t > tan(pi/12)
// by atan(t) = pi/6 + atan((t*sqrt(3) - 1) / (t + sqrt(3))), leaving
// |u| <= tan(pi/12) ~ 0.268, where the Taylor series to u^17 has a
// truncation error below 1e-11.  acos(x) = atan2(sqrt((1-x)(1+x)), x).
static constexpr float PI_F      = 0x1.921fb6p+1f;   // 3.14159274
static constexpr float PIO2_F    = 0x1.921fb6p+0f;   // 1.57079637
static constexpr float PIO6_F    = 0x1.0c1524p-1f;   // 0.52359879
static constexpr float SQRT3_F   = 0x1.bb67aep+0f;   // 1.73205078
static constexpr float TAN_PI12  = 0x1.126146p-2f;   // 0.26794919
// 1/(2k+1) rounded to float32, k = 1..8
static constexpr float INV_3  = 0x1.555556p-2f;
static constexpr float INV_5  = 0x1.99999ap-3f;
static constexpr float INV_7  = 0x1.24924ap-3f;
static constexpr float INV_9  = 0x1.c71c72p-4f;
static constexpr float INV_11 = 0x1.745d18p-4f;
static constexpr float INV_13 = 0x1.3b13b2p-4f;
static constexpr float INV_15 = 0x1.111112p-4f;
static constexpr float INV_17 = 0x1.e1e1e2p-5f;

// atan on |u| <= tan(pi/12): u - u^3/3 + u^5/5 - ... - u^15/15 + u^17/17
BNX_HD inline float atan_kernel(float u)
{
  float u2 = u * u;
  float p = INV_17;
  p = p * u2;  p = p - INV_15;
  p = p * u2;  p = p + INV_13;
  p = p * u2;  p = p - INV_11;
  p = p * u2;  p = p + INV_9;
  p = p * u2;  p = p - INV_7;
  p = p * u2;  p = p + INV_5;
  p = p * u2;  p = p - INV_3;
  p = p * u2;
  float up = u * p;
  return u + up;
}

BNX_HD inline float atan(float x)
{
  float a = fabsf(x);
  const bool inv = a > 1.f;
  if (inv) a = 1.f / a;
  float r;
  if (a > TAN_PI12) {
    float num = a * SQRT3_F - 1.f;
    float den = a + SQRT3_F;
    r = PIO6_F + atan_kernel(num / den);
  } else {
    r = atan_kernel(a);
  }
  if (inv) r = PIO2_F - r;
  return x < 0.f ? -r : r;
}

// atan2(y, x) in [-pi, pi]; (0, 0) gives 0
BNX_HD inline float atan2(float y, float x)
{
  if (x > 0.f) return atan(y / x);
  if (x < 0.f) {
    float a = atan(y / x);
    return y >= 0.f ? a + PI_F : a - PI_F;
  }
  if (y > 0.f) return PIO2_F;
  if (y < 0.f) return -PIO2_F;
  return 0.f;
}

BNX_HD inline float acos(float x)
{
  float s = sqrtf((1.f - x) * (1.f + x));
  return atan2(s, x);
}

// ------------------------------------------------------- env-map frame
// EnvMapLight::commit (light/EnvMap.cu:165-180): toWorld columns
// vz = -normalize(up), vy = -normalize(cross(vz, direction)),
// vx = normalize(cross(vy, vz)); toLocal = rcp(toWorld) as owl computes a
// 3x3 inverse: the transposed cofactors over the determinant.  Matrices are
// 9 floats, column vectors vx, vy, vz.
BNX_HD inline void env_frame(const float dir[3], const float up[3], float to_world[9], float to_local[9])
{
  auto nrm = [](float v[3]) {
    float l = 1.f / sqrtf(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
    v[0] = v[0] * l;
    v[1] = v[1] * l;
    v[2] = v[2] * l;
  };
  auto cross3 = [](const float a[3], const float b[3], float o[3]) {
    o[0] = a[1] * b[2] - a[2] * b[1];
    o[1] = a[2] * b[0] - a[0] * b[2];
    o[2] = a[0] * b[1] - a[1] * b[0];
  };
  float vz[3] = {up[0], up[1], up[2]};
  nrm(vz);
  vz[0] = -vz[0]; vz[1] = -vz[1]; vz[2] = -vz[2];
  float vy[3];
  cross3(vz, dir, vy);
  nrm(vy);
  vy[0] = -vy[0]; vy[1] = -vy[1]; vy[2] = -vy[2];
  float vx[3];
  cross3(vy, vz, vx);
  nrm(vx);
  for (int k = 0; k < 3; k++) {
    to_world[k] = vx[k];
    to_world[3 + k] = vy[k];
    to_world[6 + k] = vz[k];
  }
  // rows of the inverse: cross(vy, vz), cross(vz, vx), cross(vx, vy), / det
  float r0[3], r1[3], r2[3];
  cross3(vy, vz, r0);
  cross3(vz, vx, r1);
  cross3(vx, vy, r2);
  const float det = (vx[0] * r0[0] + vx[1] * r0[1]) + vx[2] * r0[2];
  const float inv = 1.f / det;
  // column j of toLocal = (r0[j], r1[j], r2[j]) * inv
  for (int j = 0; j < 3; j++) {
    to_local[3 * j + 0] = r0[j] * inv;
    to_local[3 * j + 1] = r1[j] * inv;
    to_local[3 * j + 2] = r2[j] * inv;
  }
}

} // namespace bnx_num
