// Device-side data layout and kernel launchers shared by the HIP translation
// units of libvhm.so.  Not part of the public C-ABI (include/vhm.h).
//
// HBM layout of one volume (one context):
//   hash table   keys[2^k] u64 packed block key (empty = ~0)
//                vals[2^k] u32 pool index of the block (~0 = none)
//                stamp[2^k] u32 last frame id that touched the key
//   block pool   pool[max_blocks][6 dir][512 voxel][3 slot][4 field]
//                (f32/u32; poolWord below): field 0 weight, 1 surface_coord,
//                2 plane_offset, 3 packed_normal.  A voxel's 3 slots of one
//                direction are 48 contiguous bytes.  147,456 B per block, as
//                the reference's Voxel array.
//                pool_keys[max_blocks] u64, obs[max_blocks][16] u32 observed
//                bitmask, dirty[max_blocks] u32, dirmask[max_blocks] u32
//                (directions with observations, sign hints).  Pool indices
//                are handed out in creation order: pass by pass (frame,
//                point-cloud chunk), and within a pass in packed-key order,
//                so the pool order of a volume depends only on its inputs.
//   per frame    active[2^k] u32 hash slots touched this frame,
//                depth staging, filtered depth (W x H f32).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/vhm.h"
#include "vhm_math.h"

namespace vhm {

constexpr int kBlockSide = 8;
constexpr int kBlockVoxels = 512;
constexpr int kDirs = 6;
constexpr int kSlots = 3;
constexpr int kFields = 4;
constexpr size_t kBlockWords = static_cast<size_t>(kDirs) * kFields * kSlots * kBlockVoxels;  // 36864
constexpr uint32_t kInvalid = 0xffffffffu;
constexpr int kMaxProbe = 1024;

// Sticky device error bits.
constexpr uint32_t kErrPoolFull = 1u;
constexpr uint32_t kErrHashFull = 2u;
constexpr uint32_t kErrKeyRange = 4u;
constexpr uint32_t kErrSampleBound = 8u;  // point cloud: more samples on a ray than the host bound

struct DevHeader {
  uint32_t num_blocks;    // pool slots handed out (can pass max_blocks on overflow)
  uint32_t active_count;  // keys touched by the current frame
  uint32_t error;         // sticky kErr* bits
  uint32_t frame;
  unsigned long long observed;  // voxels with any layer weight > 0
  unsigned long long updates;   // voxel updates (weight > 0, valid direction)
  unsigned long long visited;   // voxels visited (512 per integrated block)
  unsigned long long mesh_items;  // (block, direction) items marching cubes staged (each 9^3 x 3 slots)
  uint32_t defer_count[2];  // blocks appended to the integrator's deferred list, per frame parity: the
                            // deferred pass of frame f reads [f & 1] while frame f + 1 opens
  uint32_t integrate_ticket;  // integrator: next active-list entry to hand out this frame
  uint32_t number_base;       // num_blocks when the current insertion pass began
};

struct VolumeView {
  uint64_t* keys;
  uint32_t* vals;
  uint32_t* stamp;
  uint32_t log2cap;
  uint32_t mask;
  uint64_t* pool_keys;
  uint32_t* pool;  // float fields are stored as their bit patterns
  uint32_t* obs;
  uint32_t* dirty;
  uint32_t* dirmask;  // per block: bit d set once any voxel has an observed slot in direction d;
                      // bits 8 + d / 16 + d: kSignBelow / kSignAbove (below)
  uint32_t max_blocks;
  uint32_t* active;
  DevHeader* hdr;
  // integrator: blocks with voxels projecting within 2 px of the top/left
  // image border (estimateSurfaceNormalCamera's general path), handled by a
  // second kernel; defer_flags[p] is 1 while p is listed, 0 otherwise.
  // Null in volumes driven only by the batched step.
  uint32_t* defer_flags;
  uint32_t* defer_list;
  // launchNumberBlocks: the sorted runs of a pass's new keys when there are
  // more than one LDS sort holds (max_blocks entries)
  uint64_t* sort_tmp;
};

// Pool word of (direction, field, slot) of voxel vi within a block.  Layout
// [dir][voxel][slot][field]: the 3 slots of one direction of a voxel are 48
// contiguous, 16-byte aligned bytes (fields: weight, surface_coord,
// plane_offset, packed_normal), so a voxel update reads and writes whole
// 16-byte slots, and a direction of a block is one 24 KiB run.
VHM_HD size_t poolWord(int dir, int field, int slot, int vi) {
  return ((static_cast<size_t>(dir) * kBlockVoxels + vi) * kSlots + slot) * kFields + field;
}

struct FilterArgs {
  Intrinsics K;
  float scale;          // u16 -> metres
  float spatial_w[25];  // exp(-r^2 * inv_s) for the 5x5 taps, row-major dv, du
  float inv_r, inv_a;
  float disc_base, disc_rel, min_cos;
  DevHeader* begin_hdr;  // not null: the filter launch also opens frame begin_frame (beginFrame)
  uint32_t begin_frame;
};

// Opens a frame: the per-frame counters of the device header (one thread;
// everything of the previous frame has finished on the stream).
__device__ __forceinline__ void beginFrame(DevHeader* hdr, uint32_t frame) {
  hdr->active_count = 0u;
  hdr->defer_count[frame & 1u] = 0u;
  hdr->integrate_ticket = 0u;
  hdr->frame = frame;
  hdr->number_base = hdr->num_blocks;
}

struct AllocArgs {
  Intrinsics K;
  Pose34 T_wc;
  float voxel, trunc;
  int stride;
  uint32_t frame;
};

// Sticky per-block hints for marching cubes at iso 0, in dirmask: bit
// kSignBelow + d (kSignAbove + d) is set once a slot of direction d was
// written whose plane distance at its voxel's centre may lie within the
// staging's near window of 0 and below 0 (at or above 0).  The rule is over
// the item's 8 source blocks together: when the OR of their direction masks
// lacks the below bit or the above bit of the item's direction, no cube of
// the item can have corners on both sides, so no triangle, and the item is
// skipped.  (A per-block test would be wrong: the corner below 0 and the one
// above may sit in different source blocks.)  Writers that do not compute
// the distance set both bits.
constexpr uint32_t kSignBelow = 8u;
constexpr uint32_t kSignAbove = 16u;
VHM_HD uint32_t signBitsBoth(int d) { return (1u << (kSignBelow + d)) | (1u << (kSignAbove + d)); }
// IntegrateArgs::near_base for a volume's voxel size and layer_extract_scale
// (the extraction tolerance as marching cubes forms it).
inline float meshNearBase(float voxel, float extract_scale) {
  float tol = voxel * extract_scale;
  tol = tol < 1e-6f ? 1e-6f : tol;
  return 2.0f * tol + 1.75f * voxel;
}

struct IntegrateArgs {
  Intrinsics K;
  Pose34 T_wc, T_cw;
  float voxel, trunc;
  float merge_tol, min_cos, max_weight, min_view;
  float disc_base, disc_rel;
  float spatial_w[3];  // normal bilateral exp(-s * inv_s), s = 0, 1, 2
  float inv_r, inv_a;
  float inv_fx, inv_fy;  // RN(1 / fx), RN(1 / fy)
  int exact_fx, exact_fy;  // divByConst by fx / fy is exact (host-checked)
  float near_base;  // marching cubes' near window without its rounding term: 2 extract_tol + 1.75 voxel
  uint32_t defer_slot;  // this frame's deferred-list count: DevHeader::defer_count[defer_slot]
};

// RN(1 / f) and whether divByConst by f equals the division for every
// dividend (reciprocalDivisionExact, cached per value; defined in vhm_capi.hip).
void focalDivision(float f, float* inv, int* exact);

struct MeshArgs {
  float voxel, iso, extract_tol;
};

struct PointCloudArgs {
  Pose34 T_wc;
  float voxel, trunc, step, min_view, merge_tol, min_cos, max_weight;
  uint32_t samples_max;  // bound on the samples of one ray (host-derived from trunc / step)
  int seq_bits;          // record key: slot | voxel (9) | direction (3) | point * samples_max + sample
};

// Marching cubes keeps each lane's cluster centres (at most 8 corners x 3
// slots) in a per-workgroup global scratch column rather than in LDS, which
// lets 4 workgroups share a CU.  The mesh kernels run persistent grids of at
// most kMeshGridMax workgroups, each owning kMeshCandWords floats; the fused
// pass keeps the centres of both of a lane's cubes until it emits them, with
// one word of cube and cluster count each (kMeshFusedCandWords).
constexpr uint32_t kMeshGridMax = 2048;
constexpr size_t kMeshCandWords = 24 * 256;
constexpr size_t kMeshFusedCandWords = 2 * 25 * 256;
constexpr size_t kMeshCandBytes = kMeshGridMax * kMeshFusedCandWords * sizeof(float);

// Launchers (each defined next to its kernel).  All are asynchronous on `s`.
hipError_t launchFilterF32(const float* in, float* out, const FilterArgs& a, hipStream_t s);
hipError_t launchFilterU16(const uint16_t* in, float* out, const FilterArgs& a, hipStream_t s);
hipError_t launchFrameBegin(DevHeader* hdr, uint32_t frame, hipStream_t s);
hipError_t launchAllocate(const float* depth, const VolumeView& V, const AllocArgs& a, hipStream_t s);
// Deterministic pool order: launchNumberBegin records num_blocks before an
// insertion pass (frameBegin does it for frames); launchNumberBlocks then
// renumbers the blocks the pass created, in packed-key order, and must run
// before anything reads their pool indices.
hipError_t launchNumberBegin(DevHeader* hdr, hipStream_t s);
hipError_t launchNumberBlocks(const VolumeView& V, hipStream_t s);
// The integration pass of one frame: the main kernel, then (ordered after it
// by the caller, on any stream) the deferred pass over the border voxels.
hipError_t launchIntegrate(const float* depth, const VolumeView& V, const IntegrateArgs& a, int grid,
                           hipStream_t s);
hipError_t launchIntegrateDeferred(const float* depth, const VolumeView& V, const IntegrateArgs& a, hipStream_t s);
hipError_t launchGatherBlocksAoS(const int32_t* keys, uint32_t n, const VolumeView& V, uint32_t* out,
                                 hipStream_t s);
// out[b]: polynomial digest (sum w_i 0x100000001B3^i mod 2^64) of the AoS
// words of the block of keys[b] (vhm_blocks.hip).
hipError_t launchBlockDigests(const int32_t* keys, uint32_t n, const VolumeView& V, uint64_t* out,
                              hipStream_t s);
// Marching cubes over block `list` (pool indices; null = all of 0..n-1):
// counts[dir * n + i] and per-cube counts cube_counts[512 * (dir * n + i) ..]
// (6 * n * 512 bytes), then triangles at the scanned offsets.
hipError_t launchMeshCount(const VolumeView& V, const uint32_t* list, uint32_t n, const MeshArgs& a,
                           uint32_t* counts, uint8_t* cube_counts, const uint8_t* dirs, float* cand, hipStream_t s);
// out[i] = sum(in[0..i)) (with flags: the count of non-zero in[0..i));
// partials[(n + 1023) / 1024] receives the total.
hipError_t launchScan(const uint32_t* in, uint32_t* out, uint32_t n, uint32_t* partials, hipStream_t s,
                      bool flags = false);
size_t scanPartialsSize(uint32_t n);
// Fused count + emit of listed (block, direction) items into a device
// buffer: each item's triangles contiguous and in the two-pass order, the
// items in completion order; total: one zeroed word, receives the count.
// nbr (8 n words) and observed_dirs (n bytes): scratch for the items'
// source blocks.
// max_grid: at most this many workgroups (cand holds kMeshFusedCandWords
// floats per workgroup).
hipError_t launchMeshFused(const VolumeView& V, const uint32_t* list, uint32_t n, const MeshArgs& a,
                           const uint8_t* dirs, uint32_t* nbr, uint8_t* observed_dirs, uint32_t* total, float* tris,
                           uint32_t* tri_owner, uint32_t capacity, float* cand, hipStream_t s,
                           uint32_t max_grid = kMeshGridMax);
// p[0..n) = 0 with a kernel (no memset node when captured).
hipError_t launchZeroWords(uint32_t* p, uint32_t n, hipStream_t s);
// Puts a marked remesh list back into the dirty masks (flags cleared).
hipError_t launchRestoreDirty(const VolumeView& V, uint32_t num_blocks, uint32_t* flags, hipStream_t s);
hipError_t launchMeshEmit(const VolumeView& V, const uint32_t* list, uint32_t n, const MeshArgs& a,
                          const uint32_t* counts, const uint8_t* cube_counts, const uint32_t* offsets, float* tris,
                          uint32_t* tri_owner, uint32_t capacity, float* cand, hipStream_t s);
// Incremental meshing: dirty (block, direction) masks -> per-block masks of
// directions to re-mesh -> list of blocks with their masks.
hipError_t launchMarkRemesh(const VolumeView& V, uint32_t num_blocks, uint32_t* flags, hipStream_t s);
hipError_t launchCompactFlags(uint32_t* flags, const uint32_t* pos, uint32_t n, uint32_t* list, uint8_t* dirs,
                              hipStream_t s);
hipError_t launchListKeys(const VolumeView& V, const uint32_t* list, uint32_t n, int32_t* out, hipStream_t s);
hipError_t launchObservedFlags(const VolumeView& V, uint32_t num_blocks, uint32_t* flags, hipStream_t s);

// ---- batched multi-scene step (vhm_batch.hip) ----
struct BatchState {
  uint32_t step;              // frame index every scene integrates in the current step
  uint32_t integrate_ticket;  // batched integrator with chunk > 0: next entry to hand out (reset each step)
  uint32_t pad[2];
};

struct BatchScene {
  VolumeView V;
  FilterArgs fa;              // intrinsics, scale and filter constants
  AllocArgs aa;               // pose and frame id are taken per step
  IntegrateArgs ia;           // poses are taken per step
  const uint16_t* frames;     // [n_frames][H][W] uint16, device
  size_t frame_elems;         // H * W
  uint32_t n_frames;
  float* filtered;            // H * W
  const Pose34* poses;        // [n_frames] T_wc
  const Pose34* poses_inv;    // [n_frames] T_cw
  // incremental meshing of the step
  uint32_t* remesh_flags;     // [max_blocks]
  uint32_t* remesh_list;      // [max_blocks]
  uint8_t* remesh_dirs;       // [max_blocks] directions to re-mesh per listed block
  uint32_t* mesh_nbr;         // [max_blocks][8] source blocks per listed block
  uint8_t* mesh_obs_dirs;     // [max_blocks] observed directions over those sources
  float* tris;                // [tri_capacity][9]
  uint32_t* tri_owner;        // [tri_capacity] item d * listed + i of each triangle
  uint32_t tri_capacity;
  uint32_t* out;              // [4]: blocks re-meshed, triangles (total, may exceed capacity)
  MeshArgs ma;
};

// Index s with prefix[s] <= i < prefix[s + 1] (prefix has S + 1 entries).
VHM_HD int findSegment(const uint32_t* prefix, int S, uint32_t i) {
  int lo = 0, hi = S - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (prefix[mid] <= i) lo = mid; else hi = mid - 1;
  }
  return lo;
}

// Tiles of tile_w x (256 / tile_w) pixels (tile_w 16, 32 or 64).
uint32_t filterTiles(const Intrinsics& K, int tile_w);
uint32_t allocTiles(const Intrinsics& K, int stride, int tile_w);
hipError_t launchFilterBatch(const BatchScene* scenes, int S, const uint32_t* tile_prefix, uint32_t total_tiles,
                             const BatchState* st, int tile_w, hipStream_t s);
hipError_t launchBatchFrameBegin(const BatchScene* scenes, int S, BatchState* st, hipStream_t s);
hipError_t launchAllocateBatch(const BatchScene* scenes, int S, const uint32_t* tile_prefix, uint32_t total_tiles,
                               const BatchState* st, int tile_w, hipStream_t s);
hipError_t launchNumberBlocksBatch(const BatchScene* scenes, int S, const BatchState* st, hipStream_t s);
hipError_t launchActivePrefix(const BatchScene* scenes, int S, const BatchState* st, uint32_t* prefix,
                              hipStream_t s);
// chunk 0: workgroup w integrates entries [w c, (w + 1) c) of the
// concatenated active lists, c = ceil(total / grid); chunk > 0: workgroups
// take chunk entries at a time with an atomic ticket (st->integrate_ticket).
hipError_t launchIntegrateBatch(const BatchScene* scenes, int S, const uint32_t* prefix, BatchState* st, int grid,
                                uint32_t chunk, hipStream_t s);
hipError_t launchStepIncrement(BatchState* st, hipStream_t s);
// Incremental meshing of every scene's dirty blocks (prefix: S + 1 scratch).
hipError_t launchMeshBatch(const BatchScene* scenes, int S, uint32_t* prefix, int grid, float* cand,
                           hipStream_t s);

// Point-cloud integration (vhm_pointcloud.hip): samples -> records, a
// stable radix sort of the records by key, then the ordered updates.
size_t pointInfoBytes();
size_t radixSortCountsSize(uint32_t n);
hipError_t launchPointCloudSample(const float* pts, uint32_t n, const VolumeView& V, const PointCloudArgs& a,
                                  void* info, uint64_t* rec_key, uint32_t* rec_val, uint32_t* rec_n, hipStream_t s);
hipError_t launchRadixSort(uint64_t* keys, uint32_t* vals, uint64_t* keys_tmp, uint32_t* vals_tmp,
                           const uint32_t* n_dev, uint32_t n_host, int bits, uint32_t* counts, uint32_t* offsets,
                           uint32_t* partials, hipStream_t s);
hipError_t launchPointCloudApply(const uint64_t* rec_key, const uint32_t* rec_val, const uint32_t* rec_n,
                                 uint32_t n_host, const void* info, const VolumeView& V, const PointCloudArgs& a,
                                 hipStream_t s);

struct SynthDeviceScene {
  float room_min[3], room_max[3];
  int num_primitives;
  const float* prims;  // 8 floats each: kind, cx, cy, cz, r, hx, hy, hz
};
struct SynthNoise {
  float a, b, c, dropout, max_depth;
  uint32_t key0, key1;
  int enabled;
};
hipError_t launchSynthRender(const SynthDeviceScene& sc, const Intrinsics& K, const Pose34* poses, int n_frames,
                             uint32_t first_frame, const SynthNoise& nz, uint16_t* out, hipStream_t s);

}  // namespace vhm

namespace vhm {
// Denoiser launchers (vhm_denoise.hip).
struct DnWeights {       // packed device weights
  const uint16_t* W1t;   // [27][32 n][16 c]
  const uint16_t* W2t;   // [27][32 n][32 c]
  const uint16_t* W3t;   // [27][16 n][32 c]   (n >= 6 zero)
  const uint16_t* W2b;   // [27][32 n=c][32 k=o]  flipped taps, for dh1 = conv(da2, .)
  const uint16_t* W3b;   // [27][32 n=c][8 k=o]   flipped taps, for dh2 = conv(dy, .) (o >= 6 zero)
  const float* b;        // fp32 master parameters (biases are read from here)
};
hipError_t dnLaunchGather(const VolumeView& V, const int32_t* keys, uint32_t n, float voxel, float trunc,
                          float max_weight, uint16_t* X, hipStream_t s);
hipError_t dnLaunchTarget(const VolumeView& V, const int32_t* keys, uint32_t n, float voxel, float trunc,
                          float max_weight, float* T, float* M, hipStream_t s);
hipError_t dnLaunchPack(const float* P, uint16_t* W1t, uint16_t* W2t, uint16_t* W3t, uint16_t* W2b, uint16_t* W3b,
                        hipStream_t s);
hipError_t dnLaunchForward(const uint16_t* X, const DnWeights& w, uint32_t n, float* Y, uint16_t* H1, uint16_t* H2,
                           const float* T, const float* M, float* dY, float* loss, float inv_count, hipStream_t s);
// Adam's step, taken by the gradient reduction of a training step: FP32
// master parameters P and moments Mm, Vv, and the packed bf16 copies of the
// weights, rewritten for every updated parameter (c1, c2: the bias
// corrections, set by the launcher from the step number).
struct DnUpdate {
  float* P;
  float* Mm;
  float* Vv;
  uint16_t *W1t, *W2t, *W3t, *W2b, *W3b;
  float lr, b1, b2, eps, c1, c2;
};
// Backward pass and gradient reduction into grad; with update (not null),
// Adam's step number `step` as well.
hipError_t dnLaunchBackward(const uint16_t* X, const uint16_t* H1, const uint16_t* H2, const float* dY,
                            uint16_t* dA2, uint16_t* dA1, const DnWeights& w, uint32_t n, int groups, float* partial,
                            float* grad, const DnUpdate* update, int step, hipStream_t s);
hipError_t dnLaunchApply(const VolumeView& V, const int32_t* keys, uint32_t n, const float* Y, const uint16_t* X,
                         float voxel, float trunc, float max_weight, float min_w, float max_delta,
                         float block_max_delta, unsigned long long* counters, hipStream_t s);
hipError_t dnLaunchProbe(float* out, hipStream_t s);
hipError_t dnLaunchProbe32(const float* a, const float* b, float* out, hipStream_t s);
int dnNumParams();
}  // namespace vhm

// Context internals the denoiser needs (defined in vhm_capi.hip).
struct vhm_ctx;
vhm::VolumeView vhmCtxView(vhm_ctx* c);
hipStream_t vhmCtxStream(vhm_ctx* c);
const vhm_params* vhmCtxParams(vhm_ctx* c);
// Records frames integrated outside the context's own calls (batched steps):
// later frame ids continue after last_frame_id.
void vhmCtxAdvance(vhm_ctx* c, uint32_t last_frame_id, uint64_t frames);
