// Shared device helpers for libfwant: DRP/opt MacCormack taps, source-time
// functions, field indices.  gfx950 only.
#pragma once
#include <hip/hip_runtime.h>
#include <hip/hip_fp16.h>
#include <stdint.h>

namespace fw {

enum { VX = 0, VY, VZ, TXX, TYY, TZZ, TXY, TXZ, TYZ, NF };
constexpr int LENFD = 3;

// DRP/opt MacCormack 5-tap one-sided coefficients, code/srcF/mod_macdrp.h:19
constexpr float MCA_1 = -0.30874f, MCA0 = -0.6326f, MCA1 = 1.233f, MCA2 = -0.3334f, MCA3 = 0.04168f;
// 2-4 and 2-2 MacCormack operators, mod_macdrp.h:230-268
constexpr float M24A0 = -7.0f / 6.0f, M24A1 = 8.0f / 6.0f, M24A2 = -1.0f / 6.0f;

// One-sided operator along a stride s: forward uses taps -1..+3, backward
// is the mirrored, negated stencil on taps -3..+1 (mod_macdrp.h:69-85).
template <int FWD>
__device__ __forceinline__ float dop(const float* __restrict__ p, long s) {
  if constexpr (FWD)
    return MCA_1 * p[-s] + MCA0 * p[0] + MCA1 * p[s] + MCA2 * p[2 * s] + MCA3 * p[3 * s];
  else
    return -MCA3 * p[-3 * s] - MCA2 * p[-2 * s] - MCA1 * p[-s] - MCA0 * p[0] - MCA_1 * p[s];
}
// Same operator on a 7-slot register window w[0..6] = offsets -3..+3.
template <int FWD>
__device__ __forceinline__ float dopw(const float* w) {
  if constexpr (FWD)
    return MCA_1 * w[2] + MCA0 * w[3] + MCA1 * w[4] + MCA2 * w[5] + MCA3 * w[6];
  else
    return -MCA3 * w[0] - MCA2 * w[1] - MCA1 * w[2] - MCA0 * w[3] - MCA_1 * w[4];
}
template <int FWD>
__device__ __forceinline__ float d24w(const float* w) {
  if constexpr (FWD)
    return M24A0 * w[3] + M24A1 * w[4] + M24A2 * w[5];
  else
    return -M24A2 * w[1] - M24A1 * w[2] - M24A0 * w[3];
}
template <int FWD>
__device__ __forceinline__ float d22w(const float* w) {
  if constexpr (FWD)
    return -w[3] + w[4];
  else
    return -w[2] + w[3];
}

// ---------------------------------------------------------------------------
// Source-time functions, code/srcF/mod_src.F90:438-622.  Terms that involve
// PI are formed in double (PI is double precision in the reference,
// mod_constants.F90:12) and rounded to float.
// ---------------------------------------------------------------------------
enum {
  SVF_BELL = 100, SVF_TRIANGLE = 110, SVF_GAUSS = 120, STF_RICKER = 200, STF_BSHIFT = 210,
  STF_GAUSS = 220, STF_BELL = 230, STF_STEP = 240, STF_DELTA = 250
};
constexpr double PI_D = 3.141592653589793;
constexpr float SEIS_ZERO = 1.0e-20f;

__host__ __device__ inline float fun_gauss(float t, float a, float t0) {
  if (fabsf(t0) >= SEIS_ZERO && (t <= 0.f || t >= 2.0f * t0)) return 0.f;
  return (float)(exp((double)(-((t - t0) * (t - t0)) / (a * a))) / (sqrt(PI_D) * (double)a));
}
__host__ __device__ inline float fun_gauss_deriv(float t, float a, float t0) {
  if (fabsf(t0) >= SEIS_ZERO && (t <= 0.f || t >= 2.0f * t0)) return 0.f;
  return (float)(exp((double)(-((t - t0) * (t - t0)) / (a * a))) / (sqrt(PI_D) * (double)a) *
                 (double)(-2.f * (t - t0) / (a * a)));
}
__host__ __device__ inline float fun_ricker(float t, float fc, float t0) {
  if (t <= 0.f) return 0.f;
  const float f0 = (float)(sqrt(PI_D) / 2.0);
  const float u = (float)((double)((t - t0) * 2.0f) * PI_D * (double)fc);
  return (u * u / 4.0f - 0.5f) * (float)exp((double)(-(u * u) / 4.0f)) * f0;
}
__host__ __device__ inline float fun_ricker_deriv(float t, float fc, float t0) {
  if (t <= 0.f) return 0.f;
  const float f0 = (float)(sqrt(PI_D) / 2.0);
  const float u = (float)((double)((t - t0) * 2.0f) * PI_D * (double)fc);
  return (float)((double)(u * (1.5f - u * u / 4.0f) * (float)exp((double)(-(u * u) / 4.0f)) * f0) * PI_D *
                 (double)fc);
}
__host__ __device__ inline float fun_bell(float t, float r) {
  return (t > 0.f && t < r) ? (float)((1.0 - cos(2 * PI_D * (double)t / (double)r)) / (double)r) : 0.f;
}
__host__ __device__ inline float fun_bell_deriv(float t, float r) {
  return (t > 0.f && t < r) ? (float)(2.0 * PI_D * sin(2 * PI_D * (double)t / (double)r) / (double)r) : 0.f;
}
__host__ __device__ inline float fun_bell_int(float t, float r) {
  if (t <= 0.f) return 0.f;
  if (t < r) return (float)((double)(t / r) - sin(2 * PI_D * (double)t / (double)r) / (2.0 * PI_D));
  return 1.f;
}
__host__ __device__ inline float fun_triangle(float t, float r) {
  const float t0 = r / 2.0f;
  if (t > r) return 0.f;
  if (t > t0) return 2.0f / t0 - t / (t0 * t0);
  if (t > 0.f) return t / (t0 * t0);
  return 0.f;
}
// src_stf (which=0) / src_svf (which=1), mod_src.F90:475-533
__host__ __device__ inline float stf_eval(int which, int id, float t, float t0, float f0, int ntime, float stept) {
  if (which) {
    float r;
    switch (id) {
      case SVF_GAUSS: r = fun_gauss(t, f0, t0); break;
      case STF_GAUSS: r = fun_gauss_deriv(t, f0, t0); break;
      case STF_RICKER: r = fun_ricker_deriv(t, f0, t0); break;
      case SVF_BELL: r = fun_bell(t - t0, f0); break;
      case STF_BELL: r = fun_bell_deriv(t - t0, f0); break;
      case SVF_TRIANGLE: r = fun_triangle(t - t0, f0); break;
      case STF_STEP: r = (ntime == 0) ? 1.0f / stept : 0.f; break;
      default: r = 0.f;
    }
    return fabsf(r) < SEIS_ZERO ? 0.f : r;
  }
  switch (id) {
    case STF_GAUSS: return fun_gauss(t, f0, t0);
    case STF_RICKER: return fun_ricker(t, f0, t0);
    case STF_BELL: return fun_bell(t - t0, f0);
    case SVF_BELL: return fun_bell_int(t - t0, f0);
    case STF_STEP: return 1.f;
    case STF_DELTA: return (ntime == 0) ? 1.0f / stept : 0.f;
    default: return 0.f;
  }
}

// ---------------------------------------------------------------------------
// Mode C (terrain-following grid: theta = x(i), phi = y(j), r = r(i, j, k);
// no counterpart in the reference, whose grid is a spherical shell).  At a
// free-surface cell with slopes p, q (N = (p, q, 1)) the radial derivatives
// a = zeta_r F(V) of the velocities follow from the traction-rate condition
// sigma_t . N = S, which is linear in a:
//   (mu n2 I + (lam + mu) N N^T) a = R,   n2 = N . N,
//   R = S - sigma_t(a = 0) . N,
// solved in closed form (Sherman-Morrison):
//   a = (R - (lam + mu) / ((lam + 2 mu) n2) (N . R) N) / (mu n2).
// The strain rates then take the zeta parts
//   E11 += p aX, E22 += q aY, E33 = aZ, E12 += (p aY + q aX) / 2,
//   E13 += (aX + p aZ) / 2, E23 += (aY + q aZ) / 2.
// ---------------------------------------------------------------------------
__device__ __forceinline__ void surf_solve(float lam, float miu, float p, float q, float R0, float R1, float R2,
                                           float& aX, float& aY, float& aZ) {
  const float n2 = 1.0f + p * p + q * q;
  const float g = (lam + miu) / ((lam + 2.0f * miu) * n2) * (p * R0 + q * R1 + R2);
  const float mn = miu * n2;
  aX = (R0 - g * p) / mn;
  aY = (R1 - g * q) / mn;
  aZ = (R2 - g) / mn;
}

// ---------------------------------------------------------------------------
// Reduced-precision SGT storage: a value x of a 16x16x16 block with largest
// magnitude amax is stored as a 16-bit code of x/amax (BF16) or of
// x*32768/amax (FP16), both rounded to nearest even; the block keeps the FP32
// scale amax (BF16) or amax/32768 (FP16), and x is restored as code * scale.
// ---------------------------------------------------------------------------
enum { SGT_FP32 = 0, SGT_BF16 = 1, SGT_FP16 = 2 };
constexpr int SGT_B = 16;  // block edge in snapshot points

__device__ __forceinline__ uint16_t bf16_rne(float x) {
  uint32_t u = __float_as_uint(x);
  if ((u & 0x7fffffffu) > 0x7f800000u) return 0x7fc0;  // NaN stays NaN
  u += 0x7fffu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}
__device__ __forceinline__ float bf16_val(uint16_t c) { return __uint_as_float((uint32_t)c << 16); }
__device__ __forceinline__ uint16_t fp16_rne(float x) { return __half_as_ushort(__float2half_rn(x)); }
__device__ __forceinline__ float fp16_val(uint16_t c) { return __half2float(__ushort_as_half(c)); }
// the factor applied before rounding, and the stored scale, for a block
// maximum amax.  A block whose maximum is below SGT_TINY = 2^-100 is stored
// as zeros with scale 0: 32768 / amax (or 1 / amax) would overflow for the
// smallest maxima (the numerical precursors ahead of a wavefront reach
// subnormal values), and x * inf gives NaN codes.
constexpr float SGT_TINY = 7.888609052210118e-31f;  // 2^-100
__device__ __forceinline__ float sgt_inv(int fmt, float amax) {
  if (!(amax >= SGT_TINY)) return 0.f;
  return fmt == SGT_FP16 ? 32768.0f / amax : 1.0f / amax;
}
__device__ __forceinline__ float sgt_scale(int fmt, float amax) {
  if (!(amax >= SGT_TINY)) return 0.f;
  return fmt == SGT_FP16 ? amax * (1.0f / 32768.0f) : amax;
}
__device__ __forceinline__ uint16_t sgt_code(int fmt, float x, float inv) {
  return fmt == SGT_FP16 ? fp16_rne(x * inv) : bf16_rne(x * inv);
}
__device__ __forceinline__ float sgt_value(int fmt, uint16_t c, float scale) {
  return (fmt == SGT_FP16 ? fp16_val(c) : bf16_val(c)) * scale;
}

}  // namespace fw
