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
// Work decomposition (main pass): every wavefront takes blocks of the
// frame's active list by ticket (the list's length is read on the device,
// so the host never waits for allocation), runs the cheap half of each
// voxel's observation over its block 64 voxels at a time and the expensive
// half over the voxels that passed, 64 at a time from its own LDS queue.
// The deferred pass (border voxels) and the batched kernels share one queue
// per workgroup instead.  A voxel reads and writes only the 48-byte record
// (3 slots x 4 fields) of the direction its normal selects.  A per-block
// observed bitmask, set with an atomic OR the first time a voxel is
// observed, keeps observedVoxelCount O(1).
#include "vhm_integrate_core.h"
#include "vhm_internal.h"

#pragma clang fp contract(off)

namespace vhm {
namespace {

// The exp table of the normal's Gaussian weights, copied into LDS by each
// workgroup: a divergent 8-byte LDS read per exp instead of a global load
// through a GOT-relocated address.
__device__ __forceinline__ const uint64_t* expTableLds(uint64_t* tab) {
  if (threadIdx.x < 32) tab[threadIdx.x] = kExp2Tab32[threadIdx.x];
  __syncthreads();
  return tab;
}

__device__ __forceinline__ NormalParams normalParams(const IntegrateArgs& a, const uint64_t* exp_tab) {
  NormalParams np;
  np.exp_tab = exp_tab;
  np.K = a.K;
  np.disc_base = a.disc_base;
  np.disc_rel = a.disc_rel;
  np.spatial_w[0] = a.spatial_w[0];
  np.spatial_w[1] = a.spatial_w[1];
  np.spatial_w[2] = a.spatial_w[2];
  np.inv_r = a.inv_r;
  np.inv_a = a.inv_a;
  np.inv_fx = a.inv_fx;
  np.inv_fy = a.inv_fy;
  np.exact_fx = a.exact_fx;
  np.exact_fy = a.exact_fy;
  return np;
}

// Lane compaction.  Only about a quarter of a visited block's voxels pass
// the depth test (observeDepth) and go on to the expensive part (the
// 21-tap surface normal, the weight and the 3-slot update); run in place,
// that part would execute with most lanes of every wavefront idle.  So each
// workgroup first runs observeDepth over whole blocks (2 voxels per lane),
// appends the passing voxels -- with the projection and depth they need --
// to a queue in LDS, and runs the expensive part 256 queued voxels at a
// time, one per lane.  Voxels are independent within a frame (every block
// is in the active list once), so the order of the updates does not matter.
template <int kCap>
struct VoxelQueue {
  uint32_t p[kCap];
  uint32_t vi[kCap];
  float u[kCap], v[kCap], dep[kCap];
  float pcx[kCap], pcy[kCap], pcz[kCap];
  uint32_t n;
};
using IntegrateQueue = VoxelQueue<768>;  // workgroup queue: < 256 left over + 512 appended per block
using WaveQueue = VoxelQueue<128>;       // wavefront queue: < 64 left over + 64 appended per round

struct IntegrateTally {
  unsigned long long updates = 0, new_obs = 0, visited = 0;
};

// Which voxels of a block a pass integrates.  The surface normal of a voxel
// projecting within 2 px of the top or left image border (or so far out
// that u + 2 is inexact) needs estimateSurfaceNormalCamera's general path,
// 45 separate bilinear samples instead of 21 shared ones
// (fastNormalApplies).  Inline, that path would run for a whole wavefront
// whenever one lane needs it and would bloat the main kernel's code; so the
// main pass (kFast) integrates every other voxel and lists such blocks, and
// a second kernel (kDeferred) integrates just their border voxels.
enum class Pass { kFast, kDeferred };
constexpr int kDeferredGrid = 512;  // the deferred list is a few percent of the active blocks

// Expensive part for queue entry i (one lane).
template <Pass P, class Queue>
__device__ __forceinline__ void integrateQueued(const Queue& Q, int i, const DepthView<float>& D,
                                                const NormalParams& np, const VolumeView& V, const IntegrateArgs& a,
                                                IntegrateTally& t) {
  const uint32_t p = Q.p[i], vi = Q.vi[i];
  // bookkeeping words, loaded early: only this voxel's own update can set its
  // observed bit this frame (every block is in the active list once)
  const uint32_t obs_word = V.obs[static_cast<size_t>(p) * 16u + (vi >> 5)];
  const uint32_t dmask = V.dirmask[p];
  const uint32_t dirty_word = V.dirty[p];
  const F3 pc{Q.pcx[i], Q.pcy[i], Q.pcz[i]};
  const VoxelObservation o = observeSurface<DepthView<float>, P == Pass::kFast>(D, np, a.T_wc, pc, Q.u[i], Q.v[i],
                                                                                Q.dep[i], a.trunc, a.min_view);
  if (o.dir < 0) return;
  ++t.updates;
  const int d = o.dir;
  // the voxel's 3 slots of direction d: 48 contiguous bytes, 3 x 16-byte loads
  uint4* rec = reinterpret_cast<uint4*>(V.pool + static_cast<size_t>(p) * kBlockWords + poolWord(d, 0, 0, vi));
  uint4 old[3];
  Slots L;
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    old[k] = rec[k];
    L.w[k] = __uint_as_float(old[k].x);
    L.sc[k] = __uint_as_float(old[k].y);
    L.po[k] = __uint_as_float(old[k].z);
    L.pn[k] = old[k].w;
  }
  F3 n_new;
  float po_new;
  if (!integrateSample(L, o.coord, o.pw, o.nw, o.weight, a.merge_tol, a.min_cos, a.max_weight, &n_new, &po_new))
    return;
  // write back only the slots that changed (usually the one updated slot)
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    const uint4 now = make_uint4(__float_as_uint(L.w[k]), __float_as_uint(L.sc[k]), __float_as_uint(L.po[k]), L.pn[k]);
    if (now.x != old[k].x || now.y != old[k].y || now.z != old[k].z || now.w != old[k].w) rec[k] = now;
  }
  // observed bitmask (keeps observedVoxelCount O(1)), and the block's dirty
  // and observed direction masks for incremental meshing
  const uint32_t bit = 1u << (vi & 31u);
  if (!(obs_word & bit)) {
    atomicOr(&V.obs[static_cast<size_t>(p) * 16u + (vi >> 5)], bit);
    ++t.new_obs;
  }
  if (!(dirty_word & (1u << d))) atomicOr(&V.dirty[p], 1u << d);
  // The written slot's plane distance at the voxel centre, from the normal
  // before packing and a centre re-formed from the camera-space point: within
  // (|c|_1 + 1) 2^-12 of the distance marching cubes forms from the packed
  // normal (the 16-bit octahedral code moves a unit normal by < 1e-4).
  const F3 c = transformPoint(a.T_wc, pc);
  const float c1 = fabsf(c.x) + fabsf(c.y) + fabsf(c.z);
  const float slack = (c1 + 1.0f) * 0x1.0p-12f;
  const float vp = dot3(n_new, c) - po_new;
  uint32_t want = 1u << d;
  if (!(fabsf(vp) > a.near_base + (c1 + 4.0f * a.voxel) * 0x1.0p-18f + slack)) {  // NaN sets both, as in staging
    if (!(vp >= slack)) want |= 1u << (kSignBelow + d);
    if (!(vp <= -slack)) want |= 1u << (kSignAbove + d);
  }
  if (want & ~dmask) atomicOr(&V.dirmask[p], want);
}

// Runs the expensive part on queue entries [Q.n - m, Q.n) with m = min(Q.n, 256)
// and removes them.  Workgroup-wide.
template <Pass P>
__device__ __forceinline__ void drainQueue(IntegrateQueue& Q, const DepthView<float>& D, const NormalParams& np,
                                           const VolumeView& V, const IntegrateArgs& a, IntegrateTally& t) {
  const uint32_t n = Q.n;
  const uint32_t m = n < 256u ? n : 256u;
  if (threadIdx.x < m) integrateQueued<P>(Q, static_cast<int>(n - m + threadIdx.x), D, np, V, a, t);
  __syncthreads();
  if (threadIdx.x == 0) Q.n = n - m;
  __syncthreads();
}

// Cheap part for all 512 voxels of pool block p: append the voxels that pass
// observeDepth (and belong to pass P) to the queue.  Requires Q.n < 256 on
// entry.  Workgroup-wide.
template <Pass P>
__device__ __forceinline__ void enqueueBlock(IntegrateQueue& Q, const DepthView<float>& D, const NormalParams& np,
                                             const VolumeView& V, const IntegrateArgs& a, uint32_t p) {
  const int tid = threadIdx.x, lane = tid & 63;
  const uint64_t key = V.pool_keys[p];
  const int bx = keyX(key) * kBlockSide, by = keyY(key) * kBlockSide, bz = keyZ(key) * kBlockSide;
#pragma unroll
  for (int half = 0; half < 2; ++half) {
    const int vi = tid + 256 * half;
    const int lx = vi & 7, ly = (vi >> 3) & 7, lz = vi >> 6;
    // voxel centre, voxelToWorld (voxel_hash_tsdf.h:603-609)
    const F3 pw{(static_cast<float>(bx + lx) + 0.5f) * a.voxel, (static_cast<float>(by + ly) + 0.5f) * a.voxel,
                (static_cast<float>(bz + lz) + 0.5f) * a.voxel};
    F3 pc;
    float u, v, dep;
    bool pass = observeDepth(D, np, a.T_cw, pw, a.trunc, pc, u, v, dep);
    const bool border = pass && !fastNormalApplies(u, v);
    if constexpr (P == Pass::kFast) {
      // list the block once (the flag is cleared by the deferred pass)
      if (__any(border) && lane == 0 && atomicOr(&V.defer_flags[p], 1u) == 0u)
        V.defer_list[atomicAdd(&V.hdr->defer_count[a.defer_slot], 1u)] = p;
      pass = pass && !border;
    } else {
      pass = border;
    }
    const unsigned long long mask = __ballot(pass);
    uint32_t base = 0;
    if (lane == 0 && mask != 0ull) base = atomicAdd(&Q.n, static_cast<uint32_t>(__popcll(mask)));
    base = __shfl(base, 0);
    if (pass) {
      const uint32_t i = base + static_cast<uint32_t>(__popcll(mask & ((1ull << lane) - 1ull)));
      Q.p[i] = p;
      Q.vi[i] = static_cast<uint32_t>(vi);
      Q.u[i] = u;
      Q.v[i] = v;
      Q.dep[i] = dep;
      Q.pcx[i] = pc.x;
      Q.pcy[i] = pc.y;
      Q.pcz[i] = pc.z;
    }
  }
  __syncthreads();
}

// Integrates active-list entries [begin, end) of one volume (then drains
// the queue completely); for kDeferred, deferred-list entries.
// Workgroup-wide; Q.n == 0 on entry and exit.
template <Pass P>
__device__ __forceinline__ void integrateRange(IntegrateQueue& Q, const DepthView<float>& D, const NormalParams& np,
                               const VolumeView& V, const IntegrateArgs& a, uint32_t begin, uint32_t end,
                               IntegrateTally& t) {
  for (uint32_t i = begin; i < end; ++i) {
    uint32_t p;
    if constexpr (P == Pass::kDeferred) {
      p = V.defer_list[i];
    } else {
      p = V.vals[V.active[i]];
      if (p >= V.max_blocks) continue;  // pool exhausted for this key (error already flagged)
      if (threadIdx.x == 0) t.visited += kBlockVoxels;
    }
    enqueueBlock<P>(Q, D, np, V, a, p);
    if constexpr (P == Pass::kDeferred) {
      if (threadIdx.x == 0) V.defer_flags[p] = 0u;
    }
    while (Q.n >= 256u) drainQueue<P>(Q, D, np, V, a, t);
  }
  while (Q.n > 0u) drainQueue<P>(Q, D, np, V, a, t);
}

// Orders one wavefront's LDS accesses across its lanes (no workgroup barrier).
__device__ __forceinline__ void waveSync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// The main pass, one wavefront on its own: takes active-list entries by
// ticket until the list is exhausted, runs the cheap part over a block 64
// voxels at a time, queues the passing voxels in its own LDS queue and runs
// the expensive part whenever 64 are queued (then on the remainder).  No
// workgroup barrier, so a wavefront never waits for its neighbours' slow
// lanes, and the tickets balance dense and sparse blocks over the whole grid.
__device__ __forceinline__ void integrateWave(WaveQueue& Q, const DepthView<float>& D, const NormalParams& np,
                                              const VolumeView& V, const IntegrateArgs& a, uint32_t n_items,
                                              uint32_t* ticket, IntegrateTally& t) {
  const int lane = threadIdx.x & 63;
  const unsigned long long below = (1ull << lane) - 1ull;
  uint32_t qn = 0;  // wavefront-uniform
  uint32_t p = 0;
  int bx = 0, by = 0, bz = 0;
  int round = 8;  // 64-voxel round of block p; 8: take the next block
  bool more = true;
  for (;;) {
    while (more && qn < 64u) {
      if (round == 8) {
        uint32_t item = 0;
        if (lane == 0) item = atomicAdd(ticket, 1u);
        item = __builtin_amdgcn_readfirstlane(item);
        if (item >= n_items) {
          more = false;
          break;
        }
        p = V.vals[V.active[item]];
        if (p >= V.max_blocks) continue;  // pool exhausted for this key (error already flagged)
        if (lane == 0) t.visited += kBlockVoxels;
        const uint64_t key = V.pool_keys[p];
        bx = keyX(key) * kBlockSide;
        by = keyY(key) * kBlockSide;
        bz = keyZ(key) * kBlockSide;
        round = 0;
      }
      const int vi = round * 64 + lane;
      const int lx = vi & 7, ly = (vi >> 3) & 7, lz = vi >> 6;
      // voxel centre, voxelToWorld (voxel_hash_tsdf.h:603-609)
      const F3 pw{(static_cast<float>(bx + lx) + 0.5f) * a.voxel, (static_cast<float>(by + ly) + 0.5f) * a.voxel,
                  (static_cast<float>(bz + lz) + 0.5f) * a.voxel};
      F3 pc;
      float u, v, dep;
      bool pass = observeDepth(D, np, a.T_cw, pw, a.trunc, pc, u, v, dep);
      const bool border = pass && !fastNormalApplies(u, v);
      // list the block once for the deferred pass (which clears the flag)
      if (__any(border) && lane == 0 && atomicOr(&V.defer_flags[p], 1u) == 0u)
        V.defer_list[atomicAdd(&V.hdr->defer_count[a.defer_slot], 1u)] = p;
      pass = pass && !border;
      const unsigned long long mask = __ballot(pass);
      if (pass) {
        const uint32_t i = qn + static_cast<uint32_t>(__popcll(mask & below));
        Q.p[i] = p;
        Q.vi[i] = static_cast<uint32_t>(vi);
        Q.u[i] = u;
        Q.v[i] = v;
        Q.dep[i] = dep;
        Q.pcx[i] = pc.x;
        Q.pcy[i] = pc.y;
        Q.pcz[i] = pc.z;
      }
      qn += static_cast<uint32_t>(__popcll(mask));
      ++round;
    }
    if (qn == 0u) break;  // and no block left
    waveSync();
    const uint32_t m = qn < 64u ? qn : 64u;
    if (static_cast<uint32_t>(lane) < m)
      integrateQueued<Pass::kFast>(Q, static_cast<int>(qn - m) + lane, D, np, V, a, t);
    qn -= m;
    waveSync();  // the drained entries are rewritten by the next rounds
  }
}

__device__ __forceinline__ void flushTally(const VolumeView& V, IntegrateTally& t) {
  // one atomic per wavefront: reduce the lanes' counts first
  unsigned long long u = t.updates, o = t.new_obs;
  for (int off = 32; off > 0; off >>= 1) {
    u += __shfl_down(u, off);
    o += __shfl_down(o, off);
  }
  if ((threadIdx.x & 63) == 0) {
    if (u) atomicAdd(&V.hdr->updates, u);
    if (o) atomicAdd(&V.hdr->observed, o);
    if (t.visited) atomicAdd(&V.hdr->visited, t.visited);  // counted by lane 0 of a wavefront
  }
  t = IntegrateTally{};
}

// Workgroup w integrates the contiguous chunk of the active list
// [w * c, (w + 1) * c), c = ceil(n / grid).
// 4 wavefronts per SIMD: the allocator lands just above 128 VGPRs without the hint.
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4, 4))) void integrateKernel(const float* __restrict__ depth, VolumeView V,
                                                       IntegrateArgs a) {
  __shared__ WaveQueue Q[4];
  const DepthView<float> D{depth, a.K.width, a.K.height};
  __shared__ uint64_t exp_tab[32];
  const NormalParams np = normalParams(a, expTableLds(exp_tab));
  IntegrateTally t;
  integrateWave(Q[threadIdx.x >> 6], D, np, V, a, V.hdr->active_count, &V.hdr->integrate_ticket, t);
  flushTally(V, t);
}

// The deferred pass: the border voxels of the blocks the main pass listed.
__global__ __launch_bounds__(256) void integrateDeferredKernel(const float* __restrict__ depth, VolumeView V,
                                                               IntegrateArgs a) {
  // A few percent of the blocks, each with few border voxels: a workgroup
  // per block keeps the latency of a block at two 256-voxel rounds.
  __shared__ IntegrateQueue Q;
  if (threadIdx.x == 0) Q.n = 0;
  __syncthreads();
  const DepthView<float> D{depth, a.K.width, a.K.height};
  __shared__ uint64_t exp_tab[32];
  const NormalParams np = normalParams(a, expTableLds(exp_tab));
  const uint32_t n = V.hdr->defer_count[a.defer_slot];  // the next frame resets the other slot
  const uint32_t chunk = (n + gridDim.x - 1) / gridDim.x;
  const uint32_t begin = blockIdx.x * chunk;
  const uint32_t end = begin + chunk < n ? begin + chunk : n;
  IntegrateTally t;
  if (begin < end) integrateRange<Pass::kDeferred>(Q, D, np, V, a, begin, end, t);
  flushTally(V, t);
}

// Batched: the same over the concatenation of every scene's active list
// (prefix[s] = first entry of scene s); a range crossing scenes is split per
// scene, so queued voxels always belong to the scene being processed.
// Ranges: chunk 0, one contiguous range of ceil(total / grid) entries per
// workgroup; chunk > 0, chunk entries at a time from an atomic ticket.
__device__ __forceinline__ void integrateBatchRange(IntegrateQueue& Q, const BatchScene* __restrict__ scenes, int S,
                                                    const uint32_t* __restrict__ prefix, uint32_t step, uint32_t i,
                                                    uint32_t end, uint64_t* exp_tab) {
  while (i < end) {
    const int s = __builtin_amdgcn_readfirstlane(findSegment(prefix, S, i));  // uniform: scene fields in SGPRs
    const BatchScene& sc = scenes[s];
    const uint32_t seg_end = prefix[s + 1] < end ? prefix[s + 1] : end;
    IntegrateArgs a = sc.ia;
    a.T_wc = sc.poses[step];
    a.T_cw = sc.poses_inv[step];
    const DepthView<float> D{sc.filtered, a.K.width, a.K.height};
    const NormalParams np = normalParams(a, expTableLds(exp_tab));
    IntegrateTally t;
    integrateRange<Pass::kFast>(Q, D, np, sc.V, a, i - prefix[s], seg_end - prefix[s], t);
    flushTally(sc.V, t);
    i = seg_end;
  }
}

__global__ __launch_bounds__(256) void integrateBatchKernel(const BatchScene* __restrict__ scenes, int S,
                                                            const uint32_t* __restrict__ prefix,
                                                            BatchState* __restrict__ st, uint32_t chunk) {
  __shared__ IntegrateQueue Q;
  __shared__ uint64_t exp_tab[32];
  __shared__ uint32_t ticket;
  if (threadIdx.x == 0) Q.n = 0;
  __syncthreads();
  const uint32_t step = st->step;
  const uint32_t total = prefix[S];
  if (chunk == 0u) {
    const uint32_t c = (total + gridDim.x - 1) / gridDim.x;
    const uint32_t i = blockIdx.x * c;
    integrateBatchRange(Q, scenes, S, prefix, step, i, i + c < total ? i + c : total, exp_tab);
    return;
  }
  for (;;) {
    if (threadIdx.x == 0) ticket = atomicAdd(&st->integrate_ticket, chunk);
    __syncthreads();
    const uint32_t i = ticket;
    __syncthreads();  // ticket is rewritten by the next round
    if (i >= total) break;  // workgroup-uniform
    integrateBatchRange(Q, scenes, S, prefix, step, i, i + chunk < total ? i + chunk : total, exp_tab);
  }
}

// Batched deferred pass: workgroup column y takes scene y's deferred list.
__global__ __launch_bounds__(256) void integrateDeferredBatchKernel(const BatchScene* __restrict__ scenes,
                                                                    const BatchState* __restrict__ st) {
  const BatchScene& sc = scenes[blockIdx.y];
  const uint32_t n = sc.V.hdr->defer_count[sc.ia.defer_slot];  // reset by the next step's frame begin
  const uint32_t chunk = (n + gridDim.x - 1) / gridDim.x;
  const uint32_t begin = blockIdx.x * chunk;
  const uint32_t end = begin + chunk < n ? begin + chunk : n;
  if (begin >= end) return;  // workgroup-uniform
  __shared__ IntegrateQueue Q;
  if (threadIdx.x == 0) Q.n = 0;
  __syncthreads();
  const uint32_t step = st->step;
  IntegrateArgs a = sc.ia;
  a.T_wc = sc.poses[step];
  a.T_cw = sc.poses_inv[step];
  const DepthView<float> D{sc.filtered, a.K.width, a.K.height};
  __shared__ uint64_t exp_tab[32];
  const NormalParams np = normalParams(a, expTableLds(exp_tab));
  IntegrateTally t;
  integrateRange<Pass::kDeferred>(Q, D, np, sc.V, a, begin, end, t);
  flushTally(sc.V, t);
}

}  // namespace

hipError_t launchIntegrate(const float* depth, const VolumeView& V, const IntegrateArgs& a, int grid,
                           hipStream_t s) {
  integrateKernel<<<grid, 256, 0, s>>>(depth, V, a);
  return hipGetLastError();
}

hipError_t launchIntegrateDeferred(const float* depth, const VolumeView& V, const IntegrateArgs& a, hipStream_t s) {
  integrateDeferredKernel<<<kDeferredGrid, 256, 0, s>>>(depth, V, a);
  return hipGetLastError();
}

hipError_t launchIntegrateBatch(const BatchScene* scenes, int S, const uint32_t* prefix, BatchState* st, int grid,
                                uint32_t chunk, hipStream_t s) {
  integrateBatchKernel<<<grid, 256, 0, s>>>(scenes, S, prefix, st, chunk);
  integrateDeferredBatchKernel<<<dim3(64, S), 256, 0, s>>>(scenes, st);
  return hipGetLastError();
}

}  // namespace vhm
