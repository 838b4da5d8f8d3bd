// Synthetic depth frames: analytic room + spheres + axis-aligned boxes,
// ray-cast in double precision, with seeded sensor noise.  No reference
// counterpart beyond the demo renderer's idea (main.cpp:177-293); this is the
// benchmark/test input generator.
//
// Per pixel (u, v): ray d_c = ((u - cx)/fx, (v - cy)/fy, 1) in the camera,
// d_w = R d_c, origin t; because d_c.z = 1 the hit parameter is the camera
// depth z.  The room interior is left through its nearest exit face; spheres
// by the smaller positive root of |o + s d - c|^2 = r^2; boxes by the slab
// test.  Noise: z + sigma(z) * N(0,1) with sigma(z) = a + b (z - c)^2,
// N from Box-Muller on two Philox4x32-10 uniforms, dropout when a third
// uniform < p; rounded to millimetres.
#include <cmath>
#include <cstring>

#include "vhm_internal.h"

namespace vhm {

// Philox4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11).
__host__ __device__ void philox4x32_10(const uint32_t in_ctr[4], const uint32_t in_key[2], uint32_t out[4]) {
  uint32_t c0 = in_ctr[0], c1 = in_ctr[1], c2 = in_ctr[2], c3 = in_ctr[3];
  uint32_t k0 = in_key[0], k1 = in_key[1];
  for (int r = 0; r < 10; ++r) {
    const uint64_t p0 = static_cast<uint64_t>(0xD2511F53u) * c0;
    const uint64_t p1 = static_cast<uint64_t>(0xCD9E8D57u) * c2;
    const uint32_t hi0 = static_cast<uint32_t>(p0 >> 32), lo0 = static_cast<uint32_t>(p0);
    const uint32_t hi1 = static_cast<uint32_t>(p1 >> 32), lo1 = static_cast<uint32_t>(p1);
    const uint32_t n0 = hi1 ^ c1 ^ k0, n1 = lo1, n2 = hi0 ^ c3 ^ k1, n3 = lo0;
    c0 = n0; c1 = n1; c2 = n2; c3 = n3;
    k0 += 0x9E3779B9u;
    k1 += 0xBB67AE85u;
  }
  out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

namespace {

VHM_HD double roomExit(const double o[3], const double d[3], const float mn[3], const float mx[3]) {
  double t = 1e300;
  for (int i = 0; i < 3; ++i) {
    if (d[i] > 0.0) {
      const double s = (static_cast<double>(mx[i]) - o[i]) / d[i];
      if (s < t) t = s;
    } else if (d[i] < 0.0) {
      const double s = (static_cast<double>(mn[i]) - o[i]) / d[i];
      if (s < t) t = s;
    }
  }
  return t;
}

VHM_HD double hitPrimitive(const float* p, const double o[3], const double d[3]) {
  const double eps = 1e-6;
  if (p[0] < 0.5f) {  // sphere
    const double oc[3] = {o[0] - p[1], o[1] - p[2], o[2] - p[3]};
    const double A = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
    const double B = oc[0] * d[0] + oc[1] * d[1] + oc[2] * d[2];
    const double C = oc[0] * oc[0] + oc[1] * oc[1] + oc[2] * oc[2] - static_cast<double>(p[4]) * p[4];
    const double disc = B * B - A * C;
    if (disc < 0.0) return -1.0;
    const double sq = sqrt(disc);
    const double s0 = (-B - sq) / A, s1 = (-B + sq) / A;
    if (s0 > eps) return s0;
    if (s1 > eps) return s1;
    return -1.0;
  }
  double t0 = -1e300, t1 = 1e300;  // box slab test
  for (int i = 0; i < 3; ++i) {
    const double lo = static_cast<double>(p[1 + i]) - p[5 + i], hi = static_cast<double>(p[1 + i]) + p[5 + i];
    if (d[i] == 0.0) {
      if (o[i] < lo || o[i] > hi) return -1.0;
      continue;
    }
    double a = (lo - o[i]) / d[i], b = (hi - o[i]) / d[i];
    if (a > b) {
      const double tmp = a;
      a = b;
      b = tmp;
    }
    if (a > t0) t0 = a;
    if (b < t1) t1 = b;
    if (t1 < t0) return -1.0;
  }
  if (t0 > eps) return t0;
  if (t1 > eps) return t1;
  return -1.0;
}

VHM_HD uint16_t renderPixel(const SynthDeviceScene& sc, const Intrinsics& K, const Pose34& P, const SynthNoise& nz,
                            uint32_t frame, int u, int v) {
  const double dc[3] = {(static_cast<double>(u) - K.cx) / K.fx, (static_cast<double>(v) - K.cy) / K.fy, 1.0};
  double dw[3];
  for (int i = 0; i < 3; ++i) dw[i] = P.R[3 * i] * dc[0] + P.R[3 * i + 1] * dc[1] + P.R[3 * i + 2] * dc[2];
  const double o[3] = {P.t[0], P.t[1], P.t[2]};
  double t = roomExit(o, dw, sc.room_min, sc.room_max);
  for (int i = 0; i < sc.num_primitives; ++i) {
    const double s = hitPrimitive(sc.prims + 8 * i, o, dw);
    if (s > 0.0 && s < t) t = s;
  }
  if (!(t > 0.0 && t < 1e299)) return 0;
  double z = t;
  if (nz.enabled) {
    const uint32_t ctr[4] = {static_cast<uint32_t>(v * K.width + u), frame, 0u, 0u};
    const uint32_t key[2] = {nz.key0, nz.key1};
    uint32_t r[4];
    philox4x32_10(ctr, key, r);
    if ((static_cast<double>(r[2]) + 0.5) * (1.0 / 4294967296.0) < nz.dropout) return 0;
    const double u1 = (static_cast<double>(r[0]) + 0.5) * (1.0 / 4294967296.0);
    const double u2 = (static_cast<double>(r[1]) + 0.5) * (1.0 / 4294967296.0);
    const double g = sqrt(-2.0 * log(u1)) * cos(6.283185307179586 * u2);
    const double dz = z - nz.c;
    z = z + (nz.a + nz.b * dz * dz) * g;
  }
  if (nz.max_depth > 0.0f && z > nz.max_depth) return 0;
  const double mm = floor(z * 1000.0 + 0.5);
  if (!(mm > 0.0)) return 0;
  return mm > 65535.0 ? static_cast<uint16_t>(65535) : static_cast<uint16_t>(mm);
}

__global__ __launch_bounds__(256) void synthKernel(SynthDeviceScene sc, Intrinsics K, const Pose34* __restrict__ poses,
                                                   uint32_t first_frame, SynthNoise nz, uint16_t* __restrict__ out) {
  __shared__ float prims[8 * 256];
  const int n = sc.num_primitives < 256 ? sc.num_primitives : 256;
  for (int i = threadIdx.x; i < 8 * n; i += 256) prims[i] = sc.prims[i];
  __syncthreads();
  SynthDeviceScene local = sc;
  local.prims = prims;
  local.num_primitives = n;
  const int f = blockIdx.z;
  const int u = blockIdx.x * 16 + (threadIdx.x & 15), v = blockIdx.y * 16 + (threadIdx.x >> 4);
  if (u >= K.width || v >= K.height) return;
  out[(static_cast<size_t>(f) * K.height + v) * K.width + u] =
      renderPixel(local, K, poses[f], nz, first_frame + static_cast<uint32_t>(f), u, v);
}

}  // namespace

hipError_t launchSynthRender(const SynthDeviceScene& sc, const Intrinsics& K, const Pose34* poses, int n_frames,
                             uint32_t first_frame, const SynthNoise& nz, uint16_t* out, hipStream_t s) {
  const dim3 grid((K.width + 15) / 16, (K.height + 15) / 16, n_frames);
  synthKernel<<<grid, 256, 0, s>>>(sc, K, poses, first_frame, nz, out);
  return hipGetLastError();
}

void synthRenderHost(const SynthDeviceScene& sc, const Intrinsics& K, const Pose34* poses, int n_frames,
                     uint32_t first_frame, const SynthNoise& nz, uint16_t* out) {
  for (int f = 0; f < n_frames; ++f)
    for (int v = 0; v < K.height; ++v)
      for (int u = 0; u < K.width; ++u)
        out[(static_cast<size_t>(f) * K.height + v) * K.width + u] =
            renderPixel(sc, K, poses[f], nz, first_frame + static_cast<uint32_t>(f), u, v);
}

}  // namespace vhm
