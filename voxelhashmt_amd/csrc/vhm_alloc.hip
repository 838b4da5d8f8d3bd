// Hash-block allocation for one frame.
//
// Reference rule (collectBlocksForFrame, voxel_hash_tsdf.h:638-671): for every
// valid pixel (stride s) back-project, take the surface point and the points
// one truncation distance in front of / behind it along the camera ray, and
// activate the 3x3x3 block neighbourhood of each point's block
// (floorDiv(floor(x / voxel), 8), voxel_hash_tsdf.h:595-617).  The CUDA twin
// (cuda_block_allocator.cu:231-300) issues all 81 candidate inserts per pixel
// as global CAS into a table rebuilt (and memset) every frame.
//
// Here one workgroup covers a 256-sample tile (16x16; 32x8 and 64x4 for the
// batched step's autotuner) and deduplicates twice in LDS
// before touching HBM: the tile's (<= 768) centre keys, then their 27-block
// neighbourhoods.  Each surviving key is inserted into the persistent
// open-addressing table with a 64-bit CAS (linear probing).  A per-key frame
// stamp (atomic exchange) elects one finder per key per frame, which appends
// the key's slot to the frame's active list; inserting lanes take a pool index
// with one atomic add.  LDS overflow falls back to direct global inserts, so
// the resulting key set equals the reference's for every input.  A second
// kernel (numberBlocksKernel) then renumbers the frame's new blocks in key
// order, so pool order does not depend on the order lanes arrived in.
#include "vhm_hash.h"
#include "vhm_internal.h"

#pragma clang fp contract(off)

namespace vhm {
namespace {

constexpr int kTileW = 16;  // default tile: 16 x 16 samples
constexpr int kCentreSlots = 1024;  // >= 3 * 256 centre keys
constexpr int kNbrSlots = 4096;
constexpr int kNbrProbe = 32;

__device__ __forceinline__ uint32_t ldsSlot(uint64_t key, int bits) {
  return static_cast<uint32_t>((key * 0x9E3779B97F4A7C15ull) >> (64 - bits));
}

// Inserts key (when new) and elects one finder per key per frame: true, with
// the key's slot, for the lane that must append the slot to the frame's
// active list.  Neighbouring tiles activate many of the same keys: a plain
// load that already sees this frame's stamp skips the contended exchange
// (stamps only ever move to the current frame, so a stale read merely
// falls through to the exchange).
__device__ __forceinline__ bool claimKey(const VolumeView& V, uint64_t key, uint32_t frame, uint32_t& slot) {
  slot = insertBlock(V, key);
  if (slot == kInvalid) return false;
  return V.stamp[slot] != frame && atomicExch(&V.stamp[slot], frame) != frame;
}

__device__ void activateKey(const VolumeView& V, uint64_t key, uint32_t frame) {
  uint32_t slot;
  if (claimKey(V, key, frame, slot)) {
    const uint32_t i = atomicAdd(&V.hdr->active_count, 1u);
    V.active[i] = slot;  // i < 2^k: at most one entry per table slot per frame
  }
}

__device__ __forceinline__ int worldToVoxel1(float x, float voxel) {
  return static_cast<int>(floorf(x / voxel));
}

// One TW x (256 / TW) sample tile of the allocation rule for one frame.
template <int TW>
__device__ void allocTile(const float* __restrict__ depth, const VolumeView& V, const AllocArgs& a, int tile_x,
                          int tile_y) {
  static_assert(TW == 16 || TW == 32 || TW == 64, "tile widths 16, 32, 64");
  constexpr int kTH = 256 / TW;
  __shared__ uint64_t centres[kCentreSlots];
  __shared__ uint64_t centre_list[3 * 256];
  __shared__ uint64_t nbrs[kNbrSlots];
  __shared__ uint32_t n_centres, n_nbrs;
  static_assert(kCentreSlots * sizeof(uint64_t) >= kNbrSlots * sizeof(uint16_t), "set list fits the centre set");

  const int tid = threadIdx.x, lane = tid & 63;
  for (int i = tid; i < kCentreSlots; i += 256) centres[i] = kEmptyKey;
  for (int i = tid; i < kNbrSlots; i += 256) nbrs[i] = kEmptyKey;
  if (tid == 0) n_centres = n_nbrs = 0;
  __syncthreads();

  // 1. centre keys of this lane's pixel (voxel_hash_tsdf.h:645-661).
  const int sx = tile_x * TW + tid % TW, sy = tile_y * kTH + tid / TW;
  const int u = sx * a.stride, v = sy * a.stride;
  if (u < a.K.width && v < a.K.height) {
    const float d = depth[static_cast<size_t>(v) * a.K.width + u];
    if (d > 0.0f) {
      const F3 pw = transformPoint(a.T_wc, backProjectPixel(u, v, d, a.K));
      const F3 origin{a.T_wc.t[0], a.T_wc.t[1], a.T_wc.t[2]};
      const F3 ray = normalized3(sub3(pw, origin));
      const F3 pts[3] = {pw, sub3(pw, mul3(ray, a.trunc)), add3(pw, mul3(ray, a.trunc))};
      for (int j = 0; j < 3; ++j) {
        const int bx = floorDiv(worldToVoxel1(pts[j].x, a.voxel), kBlockSide);
        const int by = floorDiv(worldToVoxel1(pts[j].y, a.voxel), kBlockSide);
        const int bz = floorDiv(worldToVoxel1(pts[j].z, a.voxel), kBlockSide);
        if (!keyInRange(bx - 1, by - 1, bz - 1) || !keyInRange(bx + 1, by + 1, bz + 1)) {
          atomicOr(&V.hdr->error, kErrKeyRange);
          continue;
        }
        const uint64_t key = packKey(bx, by, bz);
        uint32_t s = ldsSlot(key, 10);
        for (;;) {  // never full: at most 768 keys in 1024 slots
          const unsigned long long prev =
              atomicCAS(reinterpret_cast<unsigned long long*>(&centres[s]), static_cast<unsigned long long>(kEmptyKey),
                        static_cast<unsigned long long>(key));
          if (prev == kEmptyKey) {
            centre_list[atomicAdd(&n_centres, 1u)] = key;
            break;
          }
          if (prev == key) break;
          s = (s + 1u) & (kCentreSlots - 1);
        }
      }
    }
  }
  __syncthreads();

  // 2. 27-neighbourhoods of the unique centres (voxel_hash_tsdf.h:626-636).
  const uint32_t work = n_centres * 27u;
  for (uint32_t w = tid; w < work; w += 256) {
    const uint64_t c = centre_list[w / 27u];
    const int o = static_cast<int>(w % 27u);
    const int dx = o % 3 - 1, dy = (o / 3) % 3 - 1, dz = o / 9 - 1;
    const uint64_t key = packKey(keyX(c) + dx, keyY(c) + dy, keyZ(c) + dz);
    uint32_t s = ldsSlot(key, 12);
    bool placed = false;
    for (int p = 0; p < kNbrProbe; ++p) {
      const unsigned long long prev =
          atomicCAS(reinterpret_cast<unsigned long long*>(&nbrs[s]), static_cast<unsigned long long>(kEmptyKey),
                    static_cast<unsigned long long>(key));
      if (prev == kEmptyKey || prev == key) {
        placed = true;
        break;
      }
      s = (s + 1u) & (kNbrSlots - 1);
    }
    if (!placed) activateKey(V, key, a.frame);  // LDS set crowded: go global directly
  }
  __syncthreads();

  // 3. compact the set (a few hundred keys in 4096 slots, usually) into a
  // list of slot indices, kept in the centre set's storage (dead by now), so
  // that the global activations below run in one or two rounds of dependent
  // memory round trips instead of 16 sparse ones.
  uint16_t* const list = reinterpret_cast<uint16_t*>(centres);
  const unsigned long long below = (1ull << lane) - 1ull;
  for (int i = tid; i < kNbrSlots; i += 256) {
    const bool has = nbrs[i] != kEmptyKey;
    const unsigned long long m = __ballot(has);
    uint32_t first = 0;
    if (lane == 0 && m != 0ull) first = atomicAdd(&n_nbrs, static_cast<uint32_t>(__popcll(m)));
    first = __shfl(first, 0);
    if (has) list[first + static_cast<uint32_t>(__popcll(m & below))] = static_cast<uint16_t>(i);
  }
  __syncthreads();

  // 4. persistent insert + frame activation of the unique neighbourhood keys;
  // one append counter update per wavefront.
  const uint32_t n = n_nbrs;
  for (uint32_t b = 0; b < n; b += 256) {
    const uint32_t j = b + static_cast<uint32_t>(tid);
    uint32_t slot = kInvalid;
    const bool append = j < n && claimKey(V, nbrs[list[j]], a.frame, slot);
    const unsigned long long m = __ballot(append);
    uint32_t first = 0;
    if (lane == 0 && m != 0ull) first = atomicAdd(&V.hdr->active_count, static_cast<uint32_t>(__popcll(m)));
    first = __shfl(first, 0);
    if (append) V.active[first + static_cast<uint32_t>(__popcll(m & below))] = slot;
  }
  __syncthreads();  // the sets and the list may be reset by a next tile of the same workgroup
}

__global__ __launch_bounds__(256) void allocateKernel(const float* __restrict__ depth, VolumeView V, AllocArgs a) {
  allocTile<kTileW>(depth, V, a, blockIdx.x, blockIdx.y);
}

__global__ void frameBeginKernel(DevHeader* hdr, uint32_t frame) { beginFrame(hdr, frame); }

__global__ void numberBeginKernel(DevHeader* hdr) { hdr->number_base = hdr->num_blocks; }

// ---- deterministic pool order ----
//
// insertBlock hands out pool indices in lane arrival order.  The blocks one
// pass created hold pool indices [number_base, num_blocks) and nothing has
// been written to them yet (zeroed pool, zeroed per-block words), so they can
// be renumbered by rewriting only pool_keys and the table's pool index.  One
// workgroup sorts the pass's keys (bitonic, in LDS) and gives the i-th
// smallest key index number_base + i.  More keys than one LDS sort holds
// (frames that open a scene) are sorted in runs into sort_tmp; a key's rank
// is then its rank in its run plus, per other run, the number of that run's
// keys below it (binary search; keys are distinct).

constexpr int kNumberThreads = 1024;
constexpr uint32_t kNumberRun = 8192;  // keys per LDS sort (64 KiB)

__device__ void bitonicSortLds(uint64_t* s, uint32_t P) {
  for (uint32_t k = 2; k <= P; k <<= 1)
    for (uint32_t j = k >> 1; j > 0; j >>= 1) {
      for (uint32_t i = threadIdx.x; i < P; i += kNumberThreads) {
        const uint32_t ixj = i ^ j;
        if (ixj > i) {
          const uint64_t a = s[i], b = s[ixj];
          if ((a > b) == ((i & k) == 0u)) {
            s[i] = b;
            s[ixj] = a;
          }
        }
      }
      __syncthreads();
    }
}

// Sorts src[0..n) (n <= kNumberRun) into LDS s[0..n), padding with kEmptyKey.
__device__ void sortRunLds(uint64_t* s, const uint64_t* src, uint32_t n) {
  uint32_t P = 2;
  while (P < n) P <<= 1;
  for (uint32_t i = threadIdx.x; i < P; i += kNumberThreads) s[i] = i < n ? src[i] : kEmptyKey;
  __syncthreads();
  bitonicSortLds(s, P);
}

__device__ void assignNumber(const VolumeView& V, uint64_t key, uint32_t p) {
  V.pool_keys[p] = key;
  const uint32_t slot = findSlot(V, key);
  if (slot != kInvalid) V.vals[slot] = p;
}

__device__ void numberBlocks(const VolumeView& V) {
  __shared__ uint64_t s[kNumberRun];
  const uint32_t base = V.hdr->number_base;
  const uint32_t nb = V.hdr->num_blocks;
  const uint32_t end = nb < V.max_blocks ? nb : V.max_blocks;
  if (end <= base + 1u) return;  // zero or one new block: nothing to order
  const uint32_t n = end - base;
  const uint64_t* src = V.pool_keys + base;
  if (n <= kNumberRun) {
    sortRunLds(s, src, n);
    for (uint32_t i = threadIdx.x; i < n; i += kNumberThreads) assignNumber(V, s[i], base + i);
    return;
  }
  const uint32_t runs = (n + kNumberRun - 1u) / kNumberRun;
  for (uint32_t r = 0; r < runs; ++r) {
    const uint32_t first = r * kNumberRun, len = (n - first) < kNumberRun ? (n - first) : kNumberRun;
    sortRunLds(s, src + first, len);
    for (uint32_t i = threadIdx.x; i < len; i += kNumberThreads) V.sort_tmp[first + i] = s[i];
    __syncthreads();  // s is reloaded by the next run
  }
  __threadfence();
  __syncthreads();
  for (uint32_t i = threadIdx.x; i < n; i += kNumberThreads) {
    const uint64_t key = V.sort_tmp[i];
    const uint32_t own = i / kNumberRun;
    uint32_t rank = i - own * kNumberRun;
    for (uint32_t r = 0; r < runs; ++r) {
      if (r == own) continue;
      const uint64_t* run = V.sort_tmp + r * kNumberRun;
      uint32_t lo = 0, hi = (n - r * kNumberRun) < kNumberRun ? (n - r * kNumberRun) : kNumberRun;
      while (lo < hi) {  // first element >= key
        const uint32_t mid = (lo + hi) >> 1;
        if (run[mid] < key) lo = mid + 1u; else hi = mid;
      }
      rank += lo;
    }
    assignNumber(V, key, base + rank);
  }
}

__global__ __launch_bounds__(kNumberThreads) void numberBlocksKernel(VolumeView V) { numberBlocks(V); }

// ---- batched steps ----

__global__ void batchFrameBeginKernel(const BatchScene* __restrict__ scenes, int S, BatchState* __restrict__ st) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s == 0) st->integrate_ticket = 0u;
  if (s >= S) return;
  scenes[s].V.hdr->active_count = 0u;
  scenes[s].V.hdr->defer_count[scenes[s].ia.defer_slot] = 0u;
  scenes[s].V.hdr->integrate_ticket = 0u;
  scenes[s].V.hdr->frame = st->step + 1u;
  scenes[s].V.hdr->number_base = scenes[s].V.hdr->num_blocks;
}

// One workgroup per scene.
__global__ __launch_bounds__(kNumberThreads) void numberBlocksBatchKernel(const BatchScene* __restrict__ scenes,
                                                                          const BatchState* __restrict__ st) {
  const BatchScene& sc = scenes[blockIdx.x];
  if (st->step >= sc.n_frames) return;
  numberBlocks(sc.V);
}

template <int TW>
__global__ __launch_bounds__(256) void allocateBatchKernel(const BatchScene* __restrict__ scenes, int S,
                                                           const uint32_t* __restrict__ tile_prefix,
                                                           const BatchState* __restrict__ st) {
  const uint32_t step = st->step;
  const int s = findSegment(tile_prefix, S, blockIdx.x);
  const BatchScene& sc = scenes[s];
  if (step >= sc.n_frames) return;
  AllocArgs a = sc.aa;
  a.T_wc = sc.poses[step];
  a.frame = step + 1u;
  const uint32_t t = blockIdx.x - tile_prefix[s];
  const int sw = (a.K.width + a.stride - 1) / a.stride;
  const uint32_t tiles_x = (sw + TW - 1) / TW;
  allocTile<TW>(sc.filtered, sc.V, a, t % tiles_x, t / tiles_x);
}

// prefix[s] = number of active blocks of scenes < s this step (S + 1 entries).
__global__ void activePrefixKernel(const BatchScene* __restrict__ scenes, int S, const BatchState* __restrict__ st,
                                   uint32_t* __restrict__ prefix) {
  if (threadIdx.x != 0) return;
  const uint32_t step = st->step;
  uint32_t acc = 0;
  for (int s = 0; s < S; ++s) {
    prefix[s] = acc;
    if (step < scenes[s].n_frames) acc += scenes[s].V.hdr->active_count;
  }
  prefix[S] = acc;
}

__global__ void stepIncrementKernel(BatchState* st) { st->step += 1u; }

}  // namespace

uint32_t allocTiles(const Intrinsics& K, int stride, int tile_w) {
  const int sw = (K.width + stride - 1) / stride, sh = (K.height + stride - 1) / stride;
  const int th = 256 / tile_w;
  return static_cast<uint32_t>(((sw + tile_w - 1) / tile_w) * ((sh + th - 1) / th));
}

hipError_t launchBatchFrameBegin(const BatchScene* scenes, int S, BatchState* st, hipStream_t s) {
  batchFrameBeginKernel<<<(S + 63) / 64, 64, 0, s>>>(scenes, S, st);
  return hipGetLastError();
}

hipError_t launchAllocateBatch(const BatchScene* scenes, int S, const uint32_t* tile_prefix, uint32_t total_tiles,
                               const BatchState* st, int tile_w, hipStream_t s) {
  if (tile_w == 16)
    allocateBatchKernel<16><<<total_tiles, 256, 0, s>>>(scenes, S, tile_prefix, st);
  else if (tile_w == 32)
    allocateBatchKernel<32><<<total_tiles, 256, 0, s>>>(scenes, S, tile_prefix, st);
  else if (tile_w == 64)
    allocateBatchKernel<64><<<total_tiles, 256, 0, s>>>(scenes, S, tile_prefix, st);
  else
    return hipErrorInvalidValue;
  return hipGetLastError();
}

hipError_t launchActivePrefix(const BatchScene* scenes, int S, const BatchState* st, uint32_t* prefix,
                              hipStream_t s) {
  activePrefixKernel<<<1, 64, 0, s>>>(scenes, S, st, prefix);
  return hipGetLastError();
}

hipError_t launchNumberBlocksBatch(const BatchScene* scenes, int S, const BatchState* st, hipStream_t s) {
  if (S <= 0) return hipSuccess;
  numberBlocksBatchKernel<<<S, kNumberThreads, 0, s>>>(scenes, st);
  return hipGetLastError();
}

hipError_t launchNumberBegin(DevHeader* hdr, hipStream_t s) {
  numberBeginKernel<<<1, 1, 0, s>>>(hdr);
  return hipGetLastError();
}

hipError_t launchNumberBlocks(const VolumeView& V, hipStream_t s) {
  numberBlocksKernel<<<1, kNumberThreads, 0, s>>>(V);
  return hipGetLastError();
}

hipError_t launchStepIncrement(BatchState* st, hipStream_t s) {
  stepIncrementKernel<<<1, 1, 0, s>>>(st);
  return hipGetLastError();
}

hipError_t launchFrameBegin(DevHeader* hdr, uint32_t frame, hipStream_t s) {
  frameBeginKernel<<<1, 1, 0, s>>>(hdr, frame);
  return hipGetLastError();
}

hipError_t launchAllocate(const float* depth, const VolumeView& V, const AllocArgs& a, hipStream_t s) {
  const int sw = (a.K.width + a.stride - 1) / a.stride, sh = (a.K.height + a.stride - 1) / a.stride;
  const dim3 grid((sw + kTileW - 1) / kTileW, (sh + 256 / kTileW - 1) / (256 / kTileW));
  allocateKernel<<<grid, 256, 0, s>>>(depth, V, a);
  return hipGetLastError();
}

}  // namespace vhm

namespace vhm {
namespace {

__global__ __launch_bounds__(256) void activeKeysKernel(VolumeView V, int32_t* __restrict__ out, uint32_t max_n) {
  const uint32_t n = V.hdr->active_count < max_n ? V.hdr->active_count : max_n;
  for (uint32_t i = blockIdx.x * 256u + threadIdx.x; i < n; i += gridDim.x * 256u) {
    const uint64_t key = V.keys[V.active[i]];
    out[3 * i] = keyX(key);
    out[3 * i + 1] = keyY(key);
    out[3 * i + 2] = keyZ(key);
  }
}

}  // namespace

// Keys of the current frame's active list (at most max_n of them).
hipError_t launchActiveKeys(const VolumeView& V, int32_t* out, hipStream_t s, uint32_t max_n) {
  const uint32_t grid = (max_n + 255u) / 256u < 1024u ? (max_n + 255u) / 256u : 1024u;
  if (grid == 0) return hipSuccess;
  activeKeysKernel<<<grid, 256, 0, s>>>(V, out, max_n);
  return hipGetLastError();
}

}  // namespace vhm
