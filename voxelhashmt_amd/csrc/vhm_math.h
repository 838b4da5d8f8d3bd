// Scalar math shared by the HIP kernels (and compiled for the host by the
// exhaustive expf check).  Every function performs the reference CPU path's
// float operations in the reference's order, so that with FP contraction off
// the GPU reproduces the reference bit for bit:
//   * vector ops, normalized, dot, cross          math_types.h:44-101
//   * std::max / std::min / clampf semantics      math_types.h:12-14
//   * pinhole back-projection / projection        math_types.h:228-244,
//                                                 voxel_hash_tsdf.h:586-593
//   * octahedral normal packing                   voxel_hash_tsdf.h:268-309
//   * exp of the Gaussian weights                 voxel_hash_tsdf.h:552-554,
//                                                 reconstruction_backend.cpp:34-36
//
// exp: the reference calls std::exp(float), i.e. glibc's expf.  vhm_expf is
// an implementation of the published table-driven method glibc uses (Szabolcs
// Nagy's expf from ARM optimized-routines, in glibc since 2.28): x*32/ln2 =
// k + r, 2^(k/32) from a 32-entry table of 2^(i/32), a degree-3 polynomial in
// r, all in double with fused multiply-adds, rounded once to float.  The table
// below is 2^(i/32) rounded to double with i<<47 subtracted from its bit
// pattern (it is recomputed from first principles in the tests).  The check
// against the host's glibc expf over every float in [-103, 88] (2.24e9
// inputs) found no difference.

#pragma once

#include <stdint.h>
#include <math.h>

#pragma clang fp contract(off)

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define VHM_HD __host__ __device__ __forceinline__
#define VHM_TABLE __device__ __constant__
#else
#define VHM_HD inline
#define VHM_TABLE
#endif

namespace vhm {

VHM_TABLE static const uint64_t kExp2Tab32[32] = {
    0x3ff0000000000000ull, 0x3fefd9b0d3158574ull, 0x3fefb5586cf9890full, 0x3fef9301d0125b51ull,
    0x3fef72b83c7d517bull, 0x3fef54873168b9aaull, 0x3fef387a6e756238ull, 0x3fef1e9df51fdee1ull,
    0x3fef06fe0a31b715ull, 0x3feef1a7373aa9cbull, 0x3feedea64c123422ull, 0x3feece086061892dull,
    0x3feebfdad5362a27ull, 0x3feeb42b569d4f82ull, 0x3feeab07dd485429ull, 0x3feea47eb03a5585ull,
    0x3feea09e667f3bcdull, 0x3fee9f75e8ec5f74ull, 0x3feea11473eb0187ull, 0x3feea589994cce13ull,
    0x3feeace5422aa0dbull, 0x3feeb737b0cdc5e5ull, 0x3feec49182a3f090ull, 0x3feed503b23e255dull,
    0x3feee89f995ad3adull, 0x3feeff76f2fb5e47ull, 0x3fef199bdd85529cull, 0x3fef3720dcef9069ull,
    0x3fef5818dcfba487ull, 0x3fef7c97337b9b5full, 0x3fefa4afa2a490daull, 0x3fefd0765b6e4540ull};

VHM_HD double bitsToDouble(uint64_t u) { return __builtin_bit_cast(double, u); }
VHM_HD uint64_t doubleToBits(double d) { return __builtin_bit_cast(uint64_t, d); }

// exp(x) for float x, identical to glibc 2.35 expf on an FMA-capable x86-64;
// tab is kExp2Tab32 or a copy of it (on the GPU, one in LDS).
VHM_HD float expf_glibc_tab(float x, const uint64_t* tab) {
  if (!(x == x)) return x + x;
  if (x < -0x1.9fe368p6f) return 0.0f;                 // below log(2^-150), and -inf
  if (x > 0x1.62e42ep6f) return __builtin_inff();      // above log(2^128)
  const double kInvLn2N = 0x1.71547652b82fep+0 * 32.0;
  const double kShift = 0x1.8p+52;
  const double kC0 = 0x1.c6af84b912394p-5 / 32.0 / 32.0 / 32.0;
  const double kC1 = 0x1.ebfce50fac4f3p-3 / 32.0 / 32.0;
  const double kC2 = 0x1.62e42ff0c52d6p-1 / 32.0;
  const double xd = static_cast<double>(x);
  double kd = __builtin_fma(kInvLn2N, xd, kShift);     // round(x*32/ln2) in the low bits
  const uint64_t ki = doubleToBits(kd);
  kd -= kShift;
  const double r = __builtin_fma(kInvLn2N, xd, -kd);
  const uint64_t t = tab[ki & 31u] + (ki << 47);
  const double s = bitsToDouble(t);
  const double z = __builtin_fma(kC0, r, kC1);
  const double r2 = r * r;
  double y = __builtin_fma(kC2, r, 1.0);
  y = __builtin_fma(z, r2, y);
  y = y * s;
  return static_cast<float>(y);
}
VHM_HD float expf_glibc(float x) { return expf_glibc_tab(x, kExp2Tab32); }

// ----------------------------------------------------------------- vectors --

struct F3 {
  float x, y, z;
};

VHM_HD F3 add3(F3 a, F3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
VHM_HD F3 sub3(F3 a, F3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
VHM_HD F3 mul3(F3 a, float s) { return {a.x * s, a.y * s, a.z * s}; }
VHM_HD F3 div3(F3 a, float s) { return {a.x / s, a.y / s, a.z / s}; }
VHM_HD float dot3(F3 a, F3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
VHM_HD F3 cross3(F3 a, F3 b) { return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
VHM_HD float norm3(F3 v) { return sqrtf(dot3(v, v)); }

// norm3(v) <= 1e-6f exactly when dot3(v, v) <= kNormSq1em6: sqrtf is
// correctly rounded and monotone, and this is the largest float whose
// rounded square root does not exceed 1e-6f (found by bisection over the
// float bit patterns; checked at the boundary in the CPU tests).
constexpr float kNormSq1em6 = 0x1.19799ap-40f;
VHM_HD bool normAtMost1em6(F3 v) { return dot3(v, v) <= kNormSq1em6; }
VHM_HD F3 normalized3(F3 v) {
  const float len = norm3(v);
  if (len <= 1e-8f) return {0.0f, 0.0f, 0.0f};
  return div3(v, len);
}

// std::max(a, b) / std::min(a, b) / clampf exactly (no NaN/-0 reinterpretation).
VHM_HD float smax(float a, float b) { return (a < b) ? b : a; }
VHM_HD float smin(float a, float b) { return (b < a) ? b : a; }
VHM_HD float sclamp(float v, float lo, float hi) { return smax(lo, smin(hi, v)); }
VHM_HD int imin(int a, int b) { return (b < a) ? b : a; }

// std::lround for the non-huge floats used here (round half away from zero).
VHM_HD int lroundi(float x) { return static_cast<int>(roundf(x)); }

// Row-major rotation + translation (Pose, math_types.h:180-198).
struct Pose34 {
  float R[9];
  float t[3];
};

VHM_HD F3 rotate(const Pose34& P, F3 v) {
  return {P.R[0] * v.x + P.R[1] * v.y + P.R[2] * v.z, P.R[3] * v.x + P.R[4] * v.y + P.R[5] * v.z,
          P.R[6] * v.x + P.R[7] * v.y + P.R[8] * v.z};
}
VHM_HD F3 transformPoint(const Pose34& P, F3 p) { return add3(rotate(P, p), F3{P.t[0], P.t[1], P.t[2]}); }

// Pose::inverse (math_types.h:192-197): R^T and R^T * (t * -1).
VHM_HD Pose34 inversePose(const Pose34& P) {
  Pose34 o;
  o.R[0] = P.R[0]; o.R[1] = P.R[3]; o.R[2] = P.R[6];
  o.R[3] = P.R[1]; o.R[4] = P.R[4]; o.R[5] = P.R[7];
  o.R[6] = P.R[2]; o.R[7] = P.R[5]; o.R[8] = P.R[8];
  const F3 mt{P.t[0] * -1.0f, P.t[1] * -1.0f, P.t[2] * -1.0f};
  const F3 r = rotate(o, mt);
  o.t[0] = r.x;
  o.t[1] = r.y;
  o.t[2] = r.z;
  return o;
}

struct Intrinsics {
  int width, height;
  float fx, fy, cx, cy;
};

VHM_HD F3 backProjectPixel(int u, int v, float d, const Intrinsics& K) {
  return {(static_cast<float>(u) - K.cx) * d / K.fx, (static_cast<float>(v) - K.cy) * d / K.fy, d};
}
VHM_HD F3 backProjectCont(float u, float v, float d, const Intrinsics& K) {
  return {(u - K.cx) * d / K.fx, (v - K.cy) * d / K.fy, d};
}
VHM_HD bool projectPoint(F3 p, const Intrinsics& K, float& u, float& v) {
  if (p.z <= 1e-6f) return false;
  u = K.fx * p.x / p.z + K.cx;
  v = K.fy * p.y / p.z + K.cy;
  return u >= 0.0f && u <= static_cast<float>(K.width - 1) && v >= 0.0f && v <= static_cast<float>(K.height - 1);
}

VHM_HD float discontinuityThreshold(float d, float base, float rel) { return smax(base, rel * smax(d, 0.0f)); }

// orientNormalTowardCamera (voxel_hash_tsdf.h:230-237).  Only the sign of
// dot(n, p / |p|) matters.  Rounding moves that dot product by at most
// 4.01 u sum|n_i p_i| (u = 2^-24: the three quotients and the dot's own
// roundings; the common positive factor 1 / |p| does not change its sign),
// so when the plain dot(n, p) exceeds 1e-5 sum|n_i p_i| -- far above that
// bound plus its own rounding -- the signs agree and the divisions are
// skipped; otherwise the reference sequence runs as written.
VHM_HD F3 orientTowardCamera(F3 n, F3 p) {
  if (normAtMost1em6(n) || normAtMost1em6(p)) return n;
  const float d1 = dot3(n, p);
  const float t = fabsf(n.x * p.x) + fabsf(n.y * p.y) + fabsf(n.z * p.z);
  if (t > 1e-30f && fabsf(d1) > 1e-5f * t) return d1 > 0.0f ? mul3(n, -1.0f) : n;
  const F3 ray = normalized3(p);
  return dot3(n, ray) > 0.0f ? mul3(n, -1.0f) : n;
}

// x / d for a divisor d fixed over many calls, from inv = RN(1 / d): the
// product x * inv corrected once by the exact residual fma(-q, d, x).  Equal
// to the correctly rounded x / d for every x whenever
// reciprocalDivisionExact(d) holds (checked once per divisor on the host);
// tiny |x| takes the plain division (the check covers normal quotients).
VHM_HD float divByConst(float x, float d, float inv) {
  if (fabsf(x) < 1e-30f) return x / d;
  const float q = x * inv;
  const float e = __builtin_fmaf(-q, d, x);
  return e == 0.0f ? q : __builtin_fmaf(e, inv, q);
}

// Host check for divByConst: every mantissa m in [1, 2) (the quotient of
// m * 2^k is the quotient of m scaled by 2^k while it stays normal, and the
// sign is symmetric).  About 8.4 M divisions; callers cache the answer.
inline bool reciprocalDivisionExact(float d) {
  if (!(d == d) || d == 0.0f || d * 0.0f != 0.0f) return false;  // NaN, zero, inf
  const float inv = 1.0f / d;
  for (uint32_t m = 0; m < (1u << 23); ++m) {
    const float x = __builtin_bit_cast(float, 0x3f800000u | m);
    if (__builtin_bit_cast(uint32_t, divByConst(x, d, inv)) != __builtin_bit_cast(uint32_t, x / d)) return false;
  }
  return true;
}

VHM_HD float gaussianWeight(float sq, float inv_two_sigma_sq) { return expf_glibc(-sq * inv_two_sigma_sq); }
VHM_HD float gaussianWeightTab(float sq, float inv_two_sigma_sq, const uint64_t* tab) {
  return expf_glibc_tab(-sq * inv_two_sigma_sq, tab);
}

// ------------------------------------------------------- octahedral normals --

VHM_HD float signNotZero(float v) { return v >= 0.0f ? 1.0f : -1.0f; }

VHM_HD uint32_t packNormalOct(F3 raw) {
  const F3 n = normalized3(raw);
  const float s = fabsf(n.x) + fabsf(n.y) + fabsf(n.z);
  if (s <= 1e-6f) return 0u;
  F3 p = div3(n, s);
  if (p.z < 0.0f) {
    const float px = (1.0f - fabsf(p.y)) * signNotZero(p.x);
    const float py = (1.0f - fabsf(p.x)) * signNotZero(p.y);
    p.x = px;
    p.y = py;
  }
  const float qx = sclamp(p.x * 0.5f + 0.5f, 0.0f, 1.0f);
  const float qy = sclamp(p.y * 0.5f + 0.5f, 0.0f, 1.0f);
  return static_cast<uint32_t>(lroundi(qx * 65535.0f)) | (static_cast<uint32_t>(lroundi(qy * 65535.0f)) << 16u);
}

// (c / 65535.0f) * 2 - 1 for a 16-bit code c.  The quotient is formed as
// q = c * RN(1/65535), corrected once with the exact residual
// fma(-q, 65535, c): equal to the correctly rounded c / 65535.0f for all
// 65,536 codes (checked exhaustively against the division on the host).
VHM_HD float dequantOct(uint32_t c) {
  const float x = static_cast<float>(c & 0xffffu);
  constexpr float kInv = 1.0f / 65535.0f;  // RN(1/65535)
  const float q = x * kInv;
  const float quot = __builtin_fmaf(__builtin_fmaf(-q, 65535.0f, x), kInv, q);
  return quot * 2.0f - 1.0f;
}

// unpackNormalOctahedral before its final normalisation: a vector with
// |x| + |y| + |z| = 1 (up to rounding), so its length is at least 1/sqrt(3).
VHM_HD F3 octDirection(uint32_t packed) {
  F3 n{dequantOct(packed), dequantOct(packed >> 16u), 0.0f};
  n.z = 1.0f - fabsf(n.x) - fabsf(n.y);
  if (n.z < 0.0f) {
    const float px = (1.0f - fabsf(n.y)) * signNotZero(n.x);
    const float py = (1.0f - fabsf(n.x)) * signNotZero(n.y);
    n.x = px;
    n.y = py;
  }
  return n;
}

VHM_HD F3 unpackNormalOct(uint32_t packed) { return normalized3(octDirection(packed)); }

// 1/sqrt(x) to within 2 ulp: the hardware reciprocal square root on the
// GPU (1 ulp), two correctly rounded operations on the host.
VHM_HD float approxRsqrt(float x) {
#if defined(__HIP_DEVICE_COMPILE__)
  return __builtin_amdgcn_rsqf(x);
#else
  return 1.0f / sqrtf(x);
#endif
}

// dot3(unpackNormalOct(packed), n) < min_cos, exactly, for |n| <= 1 + 2^-20.
// The rounded reference value R (normalise with three correctly rounded
// divisions, then a rounded dot product) and the estimate A = dot3(m, n) *
// rsqrt(dot3(m, m)) of the unnormalised direction m both lie within 8 ulp of
// 1 of the exact m.n / |m| (R: 2.5 ulp in the length, 1 in each quotient, 3
// in the dot product; A: 3 in each dot product, 2 in the rsqrt, 1 in the
// product): each is within 2^-20 of it, so |A - R| < 2^-19.  Outside a
// 2^-17 band around min_cos (a 4x margin) the estimate decides; inside it the
// reference computation does.  `rsqrt` is any 1/sqrt within 2 ulp.
template <class Rsqrt>
VHM_HD bool octNormalCosBelowWith(uint32_t packed, F3 n, float min_cos, Rsqrt rsqrt) {
  const F3 m = octDirection(packed);
  const float est = dot3(m, n) * rsqrt(dot3(m, m));
  if (fabsf(est - min_cos) > 0x1p-17f) return est < min_cos;
  return dot3(normalized3(m), n) < min_cos;
}

VHM_HD bool octNormalCosBelow(uint32_t packed, F3 n, float min_cos) {
  return octNormalCosBelowWith(packed, n, min_cos, [](float x) { return approxRsqrt(x); });
}

// ------------------------------------------------------------ block keys --

// floorDiv / positiveMod (voxel_hash_tsdf.h:531-546).
VHM_HD int floorDiv(int value, int divisor) {
  int q = value / divisor;
  const int r = value % divisor;
  if (r != 0 && ((r < 0) != (divisor < 0))) --q;
  return q;
}
VHM_HD int positiveMod(int value, int divisor) {
  int r = value % divisor;
  if (r < 0) r += divisor;
  return r;
}

// Packed block key: 21 bits per coordinate with bias 2^20 (the reference's
// device key packing, cuda_block_allocator.cu:19-22, 161-174); all ones = empty.
constexpr uint64_t kEmptyKey = ~0ull;
constexpr int kKeyBits = 21;
constexpr int kKeyBias = 1 << (kKeyBits - 1);
constexpr uint64_t kKeyMask = (1ull << kKeyBits) - 1ull;

VHM_HD bool keyInRange(int x, int y, int z) {
  return x >= -kKeyBias && x < kKeyBias && y >= -kKeyBias && y < kKeyBias && z >= -kKeyBias && z < kKeyBias;
}
VHM_HD uint64_t packKey(int x, int y, int z) {
  return static_cast<uint64_t>(x + kKeyBias) | (static_cast<uint64_t>(y + kKeyBias) << kKeyBits) |
         (static_cast<uint64_t>(z + kKeyBias) << (2 * kKeyBits));
}
VHM_HD int keyX(uint64_t k) { return static_cast<int>(k & kKeyMask) - kKeyBias; }
VHM_HD int keyY(uint64_t k) { return static_cast<int>((k >> kKeyBits) & kKeyMask) - kKeyBias; }
VHM_HD int keyZ(uint64_t k) { return static_cast<int>((k >> (2 * kKeyBits)) & kKeyMask) - kKeyBias; }

// Fibonacci hashing of the packed key onto a power-of-two table.
VHM_HD uint32_t hashSlot(uint64_t key, uint32_t log2cap) {
  return static_cast<uint32_t>((key * 0x9E3779B97F4A7C15ull) >> (64u - log2cap));
}

}  // namespace vhm
