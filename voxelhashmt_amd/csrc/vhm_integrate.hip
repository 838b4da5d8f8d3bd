// Directional multi-layer TSDF integration of one frame into the frame's
// active blocks.
//
// Reference: VoxelHashTSDF::integrateBlocks (voxel_hash_tsdf.h:985-1060) with
// sampleDepthBilinear (684-748), estimateSurfaceNormalCamera (750-840),
// computeObservationWeight (556-584), dominantDirectionIndex (239-254) and
// integrateDirectionalSample (479-529, with findLayerSlotForObservation
// 357-392 and sortDirectionalLayers 341-355).  CUDA twin:
// integrateActiveBlocksKernel (cuda_tsdf_integrator.cu:834-903), which
// reads/writes 288-B AoS voxels.
//
// Work decomposition: a persistent grid of 256-lane workgroups strides over
// the frame's active list (its length is read on the device, so the host never
// waits for allocation).  A workgroup owns one block at a time, 2 voxels per
// lane.  A voxel reads and writes only the 3 slots x 4 fields of the direction
// its normal selects; each (field, slot) row of a block is 2 KiB contiguous in
// the SoA pool, so lanes of a wavefront touch consecutive words.  The observed
// bitmask is updated per wavefront from a ballot (each wavefront owns 64
// consecutive voxels of the block), which keeps observedVoxelCount O(1).
#include "vhm_internal.h"

#pragma clang fp contract(off)

namespace vhm {
namespace {

struct Depth {
  const float* d;
  int W, H;
  __device__ __forceinline__ float at(int u, int v) const { return d[static_cast<size_t>(v) * W + u]; }
};

// voxel_hash_tsdf.h:673-682
__device__ __forceinline__ bool sampleNearest(const Depth& D, float u, float v, float& out) {
  const int ui = lroundi(u), vi = lroundi(v);
  if (ui < 0 || ui >= D.W || vi < 0 || vi >= D.H) return false;
  out = D.at(ui, vi);
  return out > 0.0f;
}

// voxel_hash_tsdf.h:684-748
__device__ bool sampleBilinear(const Depth& D, const IntegrateArgs& a, float u, float v, float& out) {
  if (D.W <= 0 || D.H <= 0) return false;
  if (u < 0.0f || v < 0.0f || u > static_cast<float>(D.W - 1) || v > static_cast<float>(D.H - 1)) return false;
  const int x0 = static_cast<int>(floorf(u)), y0 = static_cast<int>(floorf(v));
  const int x1 = imin(x0 + 1, D.W - 1), y1 = imin(y0 + 1, D.H - 1);
  const float tx = u - static_cast<float>(x0), ty = v - static_cast<float>(y0);
  const float d00 = D.at(x0, y0), d10 = D.at(x1, y0), d01 = D.at(x0, y1), d11 = D.at(x1, y1);
  if (d00 > 0.0f && d10 > 0.0f && d01 > 0.0f && d11 > 0.0f) {
    float ref = 0.0f;
    if (!sampleNearest(D, u, v, ref)) ref = d00;
    const float thr = discontinuityThreshold(ref, a.disc_base, a.disc_rel);
    const float mn = smin(smin(d00, d10), smin(d01, d11));
    const float mx = smax(smax(d00, d10), smax(d01, d11));
    const float top = d00 + (d10 - d00) * tx;
    const float bottom = d01 + (d11 - d01) * tx;
    if (mx - mn <= thr) {
      out = top + (bottom - top) * ty;
      return true;
    }
    const float nd[4] = {d00, d10, d01, d11};
    const float bw[4] = {(1.0f - tx) * (1.0f - ty), tx * (1.0f - ty), (1.0f - tx) * ty, tx * ty};
    float wds = 0.0f, ws = 0.0f;
    for (int i = 0; i < 4; ++i) {
      if (fabsf(nd[i] - ref) > thr) continue;
      wds += bw[i] * nd[i];
      ws += bw[i];
    }
    if (ws > 1e-6f) {
      out = wds / ws;
      return true;
    }
  }
  return sampleNearest(D, u, v, out);
}

// voxel_hash_tsdf.h:750-771
__device__ bool baseNormal(const Depth& D, const IntegrateArgs& a, float u, float v, F3& n) {
  float dl = 0.0f, dr = 0.0f, du = 0.0f, dd = 0.0f;
  if (!sampleBilinear(D, a, u - 1.0f, v, dl) || !sampleBilinear(D, a, u + 1.0f, v, dr) ||
      !sampleBilinear(D, a, u, v - 1.0f, du) || !sampleBilinear(D, a, u, v + 1.0f, dd))
    return false;
  const F3 pl = backProjectCont(u - 1.0f, v, dl, a.K), pr = backProjectCont(u + 1.0f, v, dr, a.K);
  const F3 pu = backProjectCont(u, v - 1.0f, du, a.K), pd = backProjectCont(u, v + 1.0f, dd, a.K);
  n = normalized3(cross3(sub3(pd, pu), sub3(pr, pl)));
  return norm3(n) > 1e-6f;
}

// voxel_hash_tsdf.h:773-840.  `cd` is sampleDepthBilinear(u, v), which the
// caller has already computed with identical arguments.
__device__ bool surfaceNormal(const Depth& D, const IntegrateArgs& a, float u, float v, float cd, F3& n) {
  F3 cn;
  if (!baseNormal(D, a, u, v, cn)) return false;
  const F3 cp = backProjectCont(u, v, cd, a.K);
  cn = orientTowardCamera(cn, cp);
  F3 acc{0.0f, 0.0f, 0.0f};
  float ws = 0.0f;
  for (int dv = -1; dv <= 1; ++dv)
    for (int du = -1; du <= 1; ++du) {
      const float su = u + static_cast<float>(du), sv = v + static_cast<float>(dv);
      F3 sn;
      if (!baseNormal(D, a, su, sv, sn)) continue;
      float sdep = 0.0f;
      if (!sampleBilinear(D, a, su, sv, sdep)) continue;
      const F3 sp = backProjectCont(su, sv, sdep, a.K);
      sn = orientTowardCamera(sn, sp);
      const float delta = sdep - cd;
      const float c = sclamp(dot3(cn, sn), -1.0f, 1.0f);
      const float ad = 1.0f - c;
      const float w = a.spatial_w[du * du + dv * dv] * gaussianWeight(delta * delta, a.inv_r) *
                      gaussianWeight(ad * ad, a.inv_a);
      acc = add3(acc, mul3(sn, w));
      ws += w;
    }
  if (ws <= 1e-6f) {
    n = cn;
    return true;
  }
  n = normalized3(mul3(acc, 1.0f / ws));
  n = orientTowardCamera(n, cp);
  if (norm3(n) <= 1e-6f) n = cn;
  return norm3(n) > 1e-6f;
}

// voxel_hash_tsdf.h:556-584
__device__ __forceinline__ float observationWeight(float depth, float sd, float trunc, float vaw, float min_view) {
  if (depth <= 0.0f || trunc <= 1e-6f) return 0.0f;
  const float ratio = depth / 1.5f;
  const float dw = 1.0f / (1.0f + ratio * ratio);
  const float r = fabsf(sd) / trunc;
  float robust = 1.0f;
  if (r > 0.35f) robust = 0.35f / r;
  const float band = 1.0f - 0.5f * sclamp(r, 0.0f, 1.0f);
  const float base = smax(0.05f, dw * robust * band);
  return base * sclamp(vaw, min_view, 1.0f);
}

// voxel_hash_tsdf.h:239-254
__device__ __forceinline__ int dominantDirection(F3 n) {
  const float nx = fabsf(n.x), ny = fabsf(n.y), nz = fabsf(n.z);
  if (nx <= 1e-6f && ny <= 1e-6f && nz <= 1e-6f) return -1;
  if (nx >= ny && nx >= nz) return n.x >= 0.0f ? 0 : 1;
  if (ny >= nx && ny >= nz) return n.y >= 0.0f ? 2 : 3;
  return n.z >= 0.0f ? 4 : 5;
}

// coordinateAlongDirection = dot(p, unit axis) (voxel_hash_tsdf.h:311-330).
__device__ __forceinline__ float coordinateAlong(F3 p, int d) {
  const float s = (d & 1) ? -1.0f : 1.0f;
  const F3 e = d < 2 ? F3{s, 0.0f, 0.0f} : (d < 4 ? F3{0.0f, s, 0.0f} : F3{0.0f, 0.0f, s});
  return dot3(p, e);
}

struct Slots {
  float w[3], sc[3], po[3];
  uint32_t pn[3];
};

__device__ __forceinline__ void swapSlot(Slots& L, int i, int j) {
  const float w = L.w[i], sc = L.sc[i], po = L.po[i];
  const uint32_t pn = L.pn[i];
  L.w[i] = L.w[j]; L.sc[i] = L.sc[j]; L.po[i] = L.po[j]; L.pn[i] = L.pn[j];
  L.w[j] = w; L.sc[j] = sc; L.po[j] = po; L.pn[j] = pn;
}

// integrateDirectionalSample (voxel_hash_tsdf.h:479-529) on the direction's
// three slots held in registers.  Returns true when a slot was written.
__device__ bool integrateSample(Slots& L, float coord, F3 pw, F3 nw, float w, const IntegrateArgs& a) {
  if (w <= 0.0f) return false;
  const F3 n = normalized3(nw);
  if (norm3(n) <= 1e-6f) return false;
  const float offset = dot3(n, pw);
  // findLayerSlotForObservation (357-392)
  int empty = -1, best = -1;
  float best_delta = 3.402823466e+38f;
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    if (!(L.w[k] > 0.0f)) {
      if (empty < 0) empty = k;
      continue;
    }
    if (dot3(unpackNormalOct(L.pn[k]), n) < a.min_cos) continue;
    const float delta = fabsf(L.sc[k] - coord);
    if (delta < best_delta) {
      best_delta = delta;
      best = k;
    }
  }
  const int k = (best >= 0 && best_delta <= a.merge_tol) ? best : empty;
  if (k < 0) return false;
  float pw0 = 0.0f, sc0 = 0.0f, po0 = 0.0f;
  uint32_t pn0 = 0;
#pragma unroll
  for (int j = 0; j < 3; ++j)
    if (j == k) {
      pw0 = L.w[j]; sc0 = L.sc[j]; po0 = L.po[j]; pn0 = L.pn[j];
    }
  const float ew = smin(w, smax(0.0f, a.max_weight - pw0));
  if (ew <= 0.0f) return false;
  const float nw1 = pw0 + ew;
  F3 blended = n;
  const float sc1 = pw0 > 0.0f ? ((sc0 * pw0 + coord * ew) / nw1) : coord;
  const float po1 = pw0 > 0.0f ? ((po0 * pw0 + offset * ew) / nw1) : offset;
  if (pw0 > 0.0f) {
    const F3 prev = unpackNormalOct(pn0);
    blended = normalized3(div3(add3(mul3(prev, pw0), mul3(n, ew)), nw1));
    if (norm3(blended) <= 1e-6f) blended = n;
  }
  const uint32_t pn1 = packNormalOct(blended);
#pragma unroll
  for (int j = 0; j < 3; ++j)
    if (j == k) {
      L.w[j] = nw1; L.sc[j] = sc1; L.po[j] = po1; L.pn[j] = pn1;
    }
  // sortDirectionalLayers (341-355): bubble sort, observed first, ascending coord.
#pragma unroll
  for (int pass = 0; pass < 2; ++pass)
#pragma unroll
    for (int i = 0; i < 2 - pass; ++i) {
      const bool lv = L.w[i] > 0.0f, rv = L.w[i + 1] > 0.0f;
      if ((!lv && rv) || (lv && rv && L.sc[i] > L.sc[i + 1])) swapSlot(L, i, i + 1);
    }
  return true;
}

__global__ __launch_bounds__(256) void integrateKernel(const float* __restrict__ depth, VolumeView V,
                                                       IntegrateArgs a) {
  const Depth D{depth, a.K.width, a.K.height};
  const uint32_t n_active = V.hdr->active_count;
  const int tid = threadIdx.x;
  const int lane = tid & 63;
  unsigned long long n_updates = 0, n_new_obs = 0, n_visited = 0;

  for (uint32_t i = blockIdx.x; i < n_active; i += gridDim.x) {
    const uint32_t slot = V.active[i];
    const uint32_t p = V.vals[slot];
    if (p >= V.max_blocks) continue;  // pool exhausted for this key (error already flagged)
    const uint64_t key = V.pool_keys[p];
    const int bx = keyX(key) * kBlockSide, by = keyY(key) * kBlockSide, bz = keyZ(key) * kBlockSide;
    uint32_t* blk = V.pool + static_cast<size_t>(p) * kBlockWords;
    bool any_write = false;
    if (tid == 0) n_visited += kBlockVoxels;

    for (int half = 0; half < 2; ++half) {
      const int vi = tid + 256 * half;
      const int lx = vi & 7, ly = (vi >> 3) & 7, lz = vi >> 6;
      bool reached = false, wrote = false;
      // voxel_hash_tsdf.h:1001-1046
      const F3 pw{(static_cast<float>(bx + lx) + 0.5f) * a.voxel, (static_cast<float>(by + ly) + 0.5f) * a.voxel,
                  (static_cast<float>(bz + lz) + 0.5f) * a.voxel};
      const F3 pc = transformPoint(a.T_cw, pw);
      float u = 0.0f, v = 0.0f, dep = 0.0f;
      if (!(pc.z <= 0.0f) && projectPoint(pc, a.K, u, v) && sampleBilinear(D, a, u, v, dep)) {
        const float sd = dep - pc.z;
        if (!(sd <= -a.trunc)) {
          const F3 ray = normalized3(pc);
          float vaw = 1.0f;
          F3 nc = mul3(ray, -1.0f);
          if (surfaceNormal(D, a, u, v, dep, nc)) vaw = fabsf(dot3(nc, ray));
          const float w = observationWeight(dep, sd, a.trunc, vaw, a.min_view);
          if (!(w <= 0.0f)) {
            const F3 spw = transformPoint(a.T_wc, backProjectCont(u, v, dep, a.K));
            const F3 nw = normalized3(rotate(a.T_wc, nc));
            const int d = dominantDirection(nw);
            if (d >= 0) {
              reached = true;
              const float coord = coordinateAlong(spw, d);
              Slots L;
#pragma unroll
              for (int k = 0; k < 3; ++k) {
                L.w[k] = __uint_as_float(blk[fieldOffset(d, 0, k) + vi]);
                L.sc[k] = __uint_as_float(blk[fieldOffset(d, 1, k) + vi]);
                L.po[k] = __uint_as_float(blk[fieldOffset(d, 2, k) + vi]);
                L.pn[k] = blk[fieldOffset(d, 3, k) + vi];
              }
              if (integrateSample(L, coord, spw, nw, w, a)) {
                wrote = true;
#pragma unroll
                for (int k = 0; k < 3; ++k) {
                  blk[fieldOffset(d, 0, k) + vi] = __float_as_uint(L.w[k]);
                  blk[fieldOffset(d, 1, k) + vi] = __float_as_uint(L.sc[k]);
                  blk[fieldOffset(d, 2, k) + vi] = __float_as_uint(L.po[k]);
                  blk[fieldOffset(d, 3, k) + vi] = L.pn[k];
                }
              }
            }
          }
        }
      }
      // Per-wavefront bookkeeping: this wavefront owns voxels [vi - lane, vi - lane + 63].
      const unsigned long long reached_mask = __ballot(reached);
      const unsigned long long wrote_mask = __ballot(wrote);
      if (lane == 0) {
        n_updates += __popcll(reached_mask);
        if (wrote_mask != 0ull) {
          uint32_t* ow = V.obs + static_cast<size_t>(p) * 16u + static_cast<uint32_t>(vi >> 5);
          const unsigned long long old = static_cast<unsigned long long>(ow[0]) |
                                         (static_cast<unsigned long long>(ow[1]) << 32);
          const unsigned long long now = old | wrote_mask;
          n_new_obs += __popcll(now & ~old);
          ow[0] = static_cast<uint32_t>(now);
          ow[1] = static_cast<uint32_t>(now >> 32);
        }
      }
      any_write = any_write || (wrote_mask != 0ull);
    }
    if (any_write && lane == 0) V.dirty[p] = 1u;
  }
  if (lane == 0) {
    if (n_updates) atomicAdd(&V.hdr->updates, n_updates);
    if (n_new_obs) atomicAdd(&V.hdr->observed, n_new_obs);
    if (n_visited) atomicAdd(&V.hdr->visited, n_visited);
  }
}

// Copies blocks into the reference's AoS Voxel layout (voxel_hash_tsdf.h:18-31):
// out[block][voxel][dir][slot][field].  One workgroup per requested key.
__global__ __launch_bounds__(256) void gatherBlocksAoSKernel(const int32_t* __restrict__ keys, VolumeView V,
                                                             uint32_t* __restrict__ out) {
  const uint32_t b = blockIdx.x;
  const int x = keys[3 * b], y = keys[3 * b + 1], z = keys[3 * b + 2];
  uint32_t p = kInvalid;
  if (keyInRange(x, y, z)) {
    const uint64_t key = packKey(x, y, z);
    uint32_t slot = hashSlot(key, V.log2cap);
    const int bound = (V.mask + 1u) < static_cast<uint32_t>(kMaxProbe) ? static_cast<int>(V.mask + 1u) : kMaxProbe;
    for (int probe = 0; probe < bound; ++probe) {
      const uint64_t cur = V.keys[slot];
      if (cur == key) {
        p = V.vals[slot];
        break;
      }
      if (cur == kEmptyKey) break;
      slot = (slot + 1u) & V.mask;
    }
  }
  uint32_t* dst = out + static_cast<size_t>(b) * kBlockWords;
  const uint32_t* src = (p < V.max_blocks) ? V.pool + static_cast<size_t>(p) * kBlockWords : nullptr;
  for (uint32_t i = threadIdx.x; i < kBlockWords; i += 256) {
    // destination index i = ((voxel * 6 + dir) * 3 + slot) * 4 + field
    const uint32_t field = i & 3u, rest = i >> 2;
    const uint32_t slotk = rest % 3u, rest2 = rest / 3u;
    const uint32_t dir = rest2 % 6u, voxel = rest2 / 6u;
    dst[i] = src ? src[fieldOffset(dir, field, slotk) + voxel] : 0u;
  }
}

}  // namespace

hipError_t launchIntegrate(const float* depth, const VolumeView& V, const IntegrateArgs& a, int grid,
                           hipStream_t s) {
  integrateKernel<<<grid, 256, 0, s>>>(depth, V, a);
  return hipGetLastError();
}

hipError_t launchGatherBlocksAoS(const int32_t* keys, uint32_t n, const VolumeView& V, uint32_t* out,
                                 hipStream_t s) {
  if (n == 0) return hipSuccess;
  gatherBlocksAoSKernel<<<n, 256, 0, s>>>(keys, V, out);
  return hipGetLastError();
}

}  // namespace vhm
