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
#include "vhm_integrate_core.h"
#include "vhm_internal.h"

#pragma clang fp contract(off)

namespace vhm {
namespace {

__device__ __forceinline__ NormalParams normalParams(const IntegrateArgs& a) {
  NormalParams np;
  np.K = a.K;
  np.disc_base = a.disc_base;
  np.disc_rel = a.disc_rel;
  np.spatial_w[0] = a.spatial_w[0];
  np.spatial_w[1] = a.spatial_w[1];
  np.spatial_w[2] = a.spatial_w[2];
  np.inv_r = a.inv_r;
  np.inv_a = a.inv_a;
  return np;
}

// Integrates pool block p with the whole workgroup (2 voxels per lane).
// Lane 0 of every wavefront accumulates its updates / newly observed voxels.
__device__ void integrateOneBlock(const DepthView<float>& D, const NormalParams& np, const VolumeView& V,
                                  const IntegrateArgs& a, uint32_t p, unsigned long long& n_updates,
                                  unsigned long long& n_new_obs) {
  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const uint64_t key = V.pool_keys[p];
  const int bx = keyX(key) * kBlockSide, by = keyY(key) * kBlockSide, bz = keyZ(key) * kBlockSide;
  uint32_t* blk = V.pool + static_cast<size_t>(p) * kBlockWords;
  bool any_write = false;
  uint32_t dirs_written = 0;  // wavefront-uniform after the ballots below
  for (int half = 0; half < 2; ++half) {
    const int vi = tid + 256 * half;
    const int lx = vi & 7, ly = (vi >> 3) & 7, lz = vi >> 6;
    bool reached = false, wrote = false;
    // voxel centre, voxelToWorld (voxel_hash_tsdf.h:603-609)
    const F3 pw{(static_cast<float>(bx + lx) + 0.5f) * a.voxel, (static_cast<float>(by + ly) + 0.5f) * a.voxel,
                (static_cast<float>(bz + lz) + 0.5f) * a.voxel};
    const VoxelObservation o = observeVoxel(D, np, a.T_wc, a.T_cw, pw, a.trunc, a.min_view);
    if (o.dir >= 0) {
      reached = true;
      const int d = o.dir;
      Slots L;
#pragma unroll
      for (int k = 0; k < 3; ++k) {
        L.w[k] = __uint_as_float(blk[fieldOffset(d, 0, k) + vi]);
        L.sc[k] = __uint_as_float(blk[fieldOffset(d, 1, k) + vi]);
        L.po[k] = __uint_as_float(blk[fieldOffset(d, 2, k) + vi]);
        L.pn[k] = blk[fieldOffset(d, 3, k) + vi];
      }
      if (integrateSample(L, o.coord, o.pw, o.nw, o.weight, a.merge_tol, a.min_cos, a.max_weight)) {
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
    // Per-wavefront bookkeeping: this wavefront owns voxels [vi - lane, vi - lane + 63].
    const unsigned long long reached_mask = __ballot(reached);
    const unsigned long long wrote_mask = __ballot(wrote);
    if (wrote_mask != 0ull)
#pragma unroll
      for (int d = 0; d < kDirs; ++d)
        if (__ballot(wrote && o.dir == d) != 0ull) dirs_written |= 1u << d;
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
  if (any_write && lane == 0) {
    V.dirty[p] = 1u;
    if ((V.dirmask[p] & dirs_written) != dirs_written) atomicOr(&V.dirmask[p], dirs_written);
  }
}

__global__ __launch_bounds__(256) void integrateKernel(const float* __restrict__ depth, VolumeView V,
                                                       IntegrateArgs a) {
  const DepthView<float> D{depth, a.K.width, a.K.height};
  const NormalParams np = normalParams(a);
  const uint32_t n_active = V.hdr->active_count;
  const int lane = threadIdx.x & 63;
  unsigned long long n_updates = 0, n_new_obs = 0, n_visited = 0;
  for (uint32_t i = blockIdx.x; i < n_active; i += gridDim.x) {
    const uint32_t p = V.vals[V.active[i]];
    if (p >= V.max_blocks) continue;  // pool exhausted for this key (error already flagged)
    if (threadIdx.x == 0) n_visited += kBlockVoxels;
    integrateOneBlock(D, np, V, a, p, n_updates, n_new_obs);
  }
  if (lane == 0) {
    if (n_updates) atomicAdd(&V.hdr->updates, n_updates);
    if (n_new_obs) atomicAdd(&V.hdr->observed, n_new_obs);
    if (n_visited) atomicAdd(&V.hdr->visited, n_visited);
  }
}

// Batched: a persistent grid strides over the concatenation of every
// scene's active list (prefix[s] = first item of scene s).
__global__ __launch_bounds__(256) void integrateBatchKernel(const BatchScene* __restrict__ scenes, int S,
                                                            const uint32_t* __restrict__ prefix,
                                                            const BatchState* __restrict__ st) {
  const uint32_t step = st->step;
  const uint32_t total = prefix[S];
  const int lane = threadIdx.x & 63;
  for (uint32_t i = blockIdx.x; i < total; i += gridDim.x) {
    const int s = findSegment(prefix, S, i);
    const BatchScene& sc = scenes[s];
    const VolumeView V = sc.V;
    const uint32_t p = V.vals[V.active[i - prefix[s]]];
    if (p >= V.max_blocks) continue;
    IntegrateArgs a = sc.ia;
    a.T_wc = sc.poses[step];
    a.T_cw = sc.poses_inv[step];
    const DepthView<float> D{sc.filtered, a.K.width, a.K.height};
    const NormalParams np = normalParams(a);
    unsigned long long n_updates = 0, n_new_obs = 0;
    integrateOneBlock(D, np, V, a, p, n_updates, n_new_obs);
    if (lane == 0) {
      if (n_updates) atomicAdd(&V.hdr->updates, n_updates);
      if (n_new_obs) atomicAdd(&V.hdr->observed, n_new_obs);
      if (threadIdx.x == 0) atomicAdd(&V.hdr->visited, static_cast<unsigned long long>(kBlockVoxels));
    }
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

hipError_t launchIntegrateBatch(const BatchScene* scenes, int S, const uint32_t* prefix, const BatchState* st,
                                int grid, hipStream_t s) {
  integrateBatchKernel<<<grid, 256, 0, s>>>(scenes, S, prefix, st);
  return hipGetLastError();
}

hipError_t launchGatherBlocksAoS(const int32_t* keys, uint32_t n, const VolumeView& V, uint32_t* out,
                                 hipStream_t s) {
  if (n == 0) return hipSuccess;
  gatherBlocksAoSKernel<<<n, 256, 0, s>>>(keys, V, out);
  return hipGetLastError();
}

}  // namespace vhm
