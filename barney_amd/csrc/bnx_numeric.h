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

} // namespace bnx_num
