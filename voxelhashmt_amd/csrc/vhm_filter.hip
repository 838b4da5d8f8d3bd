// Depth bilateral filter: per-pixel normals and the 5x5 normal-aware
// bilateral filter fused in one LDS-tiled kernel.
//
// Reference: estimateSurfaceNormalAtPixel (reconstruction_backend.cpp:57-94)
// and applyDepthBilateralFilter (reconstruction_backend.cpp:96-189); the CUDA
// twin is two kernels with a global round trip of the normals
// (cuda_tsdf_integrator.cu:629-725).
//
// Tiling: a workgroup of 256 lanes (4 wavefronts) filters a TX x (256 / TX)
// pixel tile (32x8 by default; 16x16 and 64x4 for the batched step's
// autotuner).  LDS holds the tile's depth with a 3-pixel apron (38x14 at
// 32x8, converted from uint16 once) and the normals of the tile with a
// 2-pixel apron (36x12), so
// every depth sample is read from HBM once per tile (plus the apron) and the
// normals never leave the CU.  Spatial Gaussian weights are precomputed on the
// host; the range and angular Gaussians use expf_glibc (its table in LDS).
#include "vhm_internal.h"

#pragma clang fp contract(off)

namespace vhm {
namespace {

constexpr int kR = 2;
constexpr int kDefaultTX = 32;

template <typename In>
__device__ __forceinline__ float loadDepth(const In* in, size_t i, float scale);
template <>
__device__ __forceinline__ float loadDepth<float>(const float* in, size_t i, float) {
  return in[i];
}
template <>
__device__ __forceinline__ float loadDepth<uint16_t>(const uint16_t* in, size_t i, float scale) {
  // dataset_io.h:342-356: 0 stays 0, otherwise raw * depth_scale (kept if > 0).
  const uint16_t raw = in[i];
  if (raw == 0u) return 0.0f;
  const float d = static_cast<float>(raw) * scale;
  return d > 0.0f ? d : 0.0f;
}

}  // namespace

// One TX x (256 / TX) output tile (tile_x, tile_y) of the filter; the whole
// workgroup (256 lanes) takes part.
template <typename In, int kTX>
__device__ void filterTile(const In* __restrict__ in, float* __restrict__ out, const FilterArgs& a, int tile_x,
                           int tile_y) {
  static_assert(kTX == 16 || kTX == 32 || kTX == 64, "tile widths 16, 32, 64");
  constexpr int kTY = 256 / kTX;
  constexpr int kDW = kTX + 2 * (kR + 1), kDH = kTY + 2 * (kR + 1);  // depth tile (38x14 at 32x8)
  constexpr int kNW = kTX + 2 * kR, kNH = kTY + 2 * kR;              // normal tile (36x12 at 32x8)
  __shared__ float sd[kDH][kDW];
  __shared__ float snx[kNH][kNW], sny[kNH][kNW], snz[kNH][kNW];
  __shared__ uint8_t sok[kNH][kNW];
  // expf_glibc's table in LDS: a divergent 8-byte LDS read per exp instead
  // of a load from global memory
  __shared__ uint64_t exp_tab[32];

  const int W = a.K.width, H = a.K.height;
  const int tid = threadIdx.x;
  const int x0 = tile_x * kTX, y0 = tile_y * kTY;
  const int dx0 = x0 - (kR + 1), dy0 = y0 - (kR + 1);
  if (tid < 32) exp_tab[tid] = kExp2Tab32[tid];

  for (int i = tid; i < kDW * kDH; i += 256) {
    const int lx = i % kDW, ly = i / kDW;
    const int gx = dx0 + lx, gy = dy0 + ly;
    float d = 0.0f;
    if (gx >= 0 && gx < W && gy >= 0 && gy < H) d = loadDepth<In>(in, static_cast<size_t>(gy) * W + gx, a.scale);
    sd[ly][lx] = d;
  }
  __syncthreads();

  // Normals for the tile plus a 2-pixel apron (reconstruction_backend.cpp:57-94).
  for (int i = tid; i < kNW * kNH; i += 256) {
    const int lx = i % kNW, ly = i / kNW;
    const int u = x0 - kR + lx, v = y0 - kR + ly;
    const int cx = lx + 1, cy = ly + 1;  // position in the depth tile
    bool ok = false;
    F3 n{0.0f, 0.0f, 0.0f};
    if (!(u <= 0 || v <= 0 || u + 1 >= W || v + 1 >= H)) {
      const float c = sd[cy][cx], l = sd[cy][cx - 1], r = sd[cy][cx + 1], up = sd[cy - 1][cx], dn = sd[cy + 1][cx];
      if (!(c <= 0.0f || l <= 0.0f || r <= 0.0f || up <= 0.0f || dn <= 0.0f)) {
        const float thr = discontinuityThreshold(c, a.disc_base, a.disc_rel);
        if (!(fabsf(l - c) > thr || fabsf(r - c) > thr || fabsf(up - c) > thr || fabsf(dn - c) > thr)) {
          const F3 pl = backProjectCont(static_cast<float>(u - 1), static_cast<float>(v), l, a.K);
          const F3 pr = backProjectCont(static_cast<float>(u + 1), static_cast<float>(v), r, a.K);
          const F3 pu = backProjectCont(static_cast<float>(u), static_cast<float>(v - 1), up, a.K);
          const F3 pd = backProjectCont(static_cast<float>(u), static_cast<float>(v + 1), dn, a.K);
          n = normalized3(cross3(sub3(pd, pu), sub3(pr, pl)));
          if (!normAtMost1em6(n)) {
            const F3 cp = backProjectCont(static_cast<float>(u), static_cast<float>(v), c, a.K);
            n = orientTowardCamera(n, cp);
            ok = !normAtMost1em6(n);
          }
        }
      }
    }
    snx[ly][lx] = n.x;
    sny[ly][lx] = n.y;
    snz[ly][lx] = n.z;
    sok[ly][lx] = ok ? 1 : 0;
  }
  __syncthreads();

  const int tx = tid % kTX, ty = tid / kTX;
  const int u = x0 + tx, v = y0 + ty;
  if (u >= W || v >= H) return;  // last use of the shared tile: no barrier follows
  const int cdx = tx + kR + 1, cdy = ty + kR + 1;  // centre in the depth tile
  const int cnx = tx + kR, cny = ty + kR;          // centre in the normal tile
  const float c = sd[cdy][cdx];
  float result = 0.0f;
  if (c > 0.0f) {
    const bool has_n = sok[cny][cnx] != 0;
    const F3 cn = has_n ? F3{snx[cny][cnx], sny[cny][cnx], snz[cny][cnx]} : F3{0.0f, 0.0f, 0.0f};
    const float thr = discontinuityThreshold(c, a.disc_base, a.disc_rel);
    float wd = 0.0f, ws = 0.0f;
    for (int dv = -kR; dv <= kR; ++dv) {
      const int sv = v + dv;
      if (sv < 0 || sv >= H) continue;
      for (int du = -kR; du <= kR; ++du) {
        const int su = u + du;
        if (su < 0 || su >= W) continue;
        const float s = sd[cdy + dv][cdx + du];
        if (s <= 0.0f) continue;
        if (fabsf(s - c) > thr) continue;
        const float dd = s - c;
        float w = a.spatial_w[(dv + kR) * (2 * kR + 1) + (du + kR)] * gaussianWeightTab(dd * dd, a.inv_r, exp_tab);
        if (has_n) {
          if (sok[cny + dv][cnx + du] == 0) continue;
          const F3 sn{snx[cny + dv][cnx + du], sny[cny + dv][cnx + du], snz[cny + dv][cnx + du]};
          const float cosv = sclamp(dot3(cn, sn), -1.0f, 1.0f);
          if (cosv < a.min_cos) continue;
          const float ad = 1.0f - cosv;
          w *= gaussianWeightTab(ad * ad, a.inv_a, exp_tab);
        }
        wd += w * s;
        ws += w;
      }
    }
    result = ws > 1e-6f ? (wd / ws) : c;
  }
  out[static_cast<size_t>(v) * W + u] = result;
}

namespace {

template <typename In>
__global__ __launch_bounds__(256) void filterKernel(const In* __restrict__ in, float* __restrict__ out, FilterArgs a) {
  // opening the frame here saves the frame-begin launch before allocation
  if (a.begin_hdr && blockIdx.x == 0 && blockIdx.y == 0 && threadIdx.x == 0) beginFrame(a.begin_hdr, a.begin_frame);
  filterTile<In, kDefaultTX>(in, out, a, blockIdx.x, blockIdx.y);
}

// Batched: workgroup i filters tile i of the concatenation of every scene's
// tiles (tile_prefix[s] = first tile of scene s), frame `step` of each scene.
template <int kTX>
__global__ __launch_bounds__(256) void filterBatchKernel(const BatchScene* __restrict__ scenes, int S,
                                                         const uint32_t* __restrict__ tile_prefix,
                                                         const BatchState* __restrict__ st) {
  const uint32_t step = st->step;
  const int s = findSegment(tile_prefix, S, blockIdx.x);
  const BatchScene& sc = scenes[s];
  if (step >= sc.n_frames) return;
  const uint32_t t = blockIdx.x - tile_prefix[s];
  const uint32_t tiles_x = (sc.fa.K.width + kTX - 1) / kTX;
  filterTile<uint16_t, kTX>(sc.frames + static_cast<size_t>(step) * sc.frame_elems, sc.filtered, sc.fa,
                            t % tiles_x, t / tiles_x);
}

}  // namespace

uint32_t filterTiles(const Intrinsics& K, int tile_w) {
  const int th = 256 / tile_w;
  return ((K.width + tile_w - 1) / tile_w) * ((K.height + th - 1) / th);
}

hipError_t launchFilterBatch(const BatchScene* scenes, int S, const uint32_t* tile_prefix, uint32_t total_tiles,
                             const BatchState* st, int tile_w, hipStream_t s) {
  if (tile_w == 16)
    filterBatchKernel<16><<<total_tiles, 256, 0, s>>>(scenes, S, tile_prefix, st);
  else if (tile_w == 64)
    filterBatchKernel<64><<<total_tiles, 256, 0, s>>>(scenes, S, tile_prefix, st);
  else if (tile_w == 32)
    filterBatchKernel<32><<<total_tiles, 256, 0, s>>>(scenes, S, tile_prefix, st);
  else
    return hipErrorInvalidValue;
  return hipGetLastError();
}

hipError_t launchFilterF32(const float* in, float* out, const FilterArgs& a, hipStream_t s) {
  constexpr int kTY = 256 / kDefaultTX;
  const dim3 grid((a.K.width + kDefaultTX - 1) / kDefaultTX, (a.K.height + kTY - 1) / kTY);
  filterKernel<float><<<grid, 256, 0, s>>>(in, out, a);
  return hipGetLastError();
}

hipError_t launchFilterU16(const uint16_t* in, float* out, const FilterArgs& a, hipStream_t s) {
  constexpr int kTY = 256 / kDefaultTX;
  const dim3 grid((a.K.width + kDefaultTX - 1) / kDefaultTX, (a.K.height + kTY - 1) / kTY);
  filterKernel<uint16_t><<<grid, 256, 0, s>>>(in, out, a);
  return hipGetLastError();
}

}  // namespace vhm
