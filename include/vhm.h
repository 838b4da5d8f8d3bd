/*
 * vhm -- MI355X (gfx950) voxel-hash TSDF fusion: C-ABI of libvhm.so.
 *
 * Drop-in GPU backend for the fusion path of HM-ZC/voxelhashmt: depth
 * bilateral filter -> hash-block allocation -> directional multi-layer TSDF
 * integration -> multi-layer marching cubes.  Results equal the reference's
 * CPU path (TsdfIntegrationBackend with ComputeBackendPreference::kCpu) on
 * identical inputs.
 *
 * Conventions
 *  - Plain C types only: pointers, sizes, POD structs.  No C++ or HIP types.
 *  - Every entry point returns an int status (VHM_OK == 0).  On failure
 *    vhm_last_error(ctx) holds "<action> failed: <detail>", the message shape
 *    of the reference's throwCudaError (cuda_support.h:11-19).
 *  - One context == one volume (the reference keys its GPU cache by the
 *    VoxelHashTSDF address, cuda_tsdf_integrator.cu:114-130).  A context owns
 *    one HIP stream; work on it is ordered and asynchronous.  Device-side
 *    capacity errors are sticky and are returned by the next vhm_synchronize
 *    or query call.  A context must not be used from two threads at once.
 *  - memkind: VHM_HOST for pageable/pinned host memory, VHM_DEVICE for
 *    device memory on the context's GPU.
 *  - Block keys are int32 triples (x, y, z) in block units, as BlockKey
 *    (voxel_hash_tsdf.h:33-41).  Triangles are 9 floats (v0, v1, v2), as
 *    Triangle (marching_cubes.h:11-15).
 */
#ifndef VHM_H_
#define VHM_H_

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum vhm_status {
  VHM_OK = 0,
  VHM_ERR_INVALID_ARGUMENT = 1,
  VHM_ERR_HIP = 2,
  /* Persistent hash table or block pool full.  The reference's analogue is
   * CudaActiveBlockAllocationFallback (cuda_block_allocator.h:11-15). */
  VHM_ERR_CAPACITY = 3,
  /* A block coordinate outside +-2^20 (the reference's packed-key range,
   * cuda_block_allocator.cu:161-174). */
  VHM_ERR_KEY_RANGE = 4,
  VHM_ERR_NO_DEVICE = 5,
  VHM_ERR_BUFFER_TOO_SMALL = 6
};

enum vhm_memkind { VHM_HOST = 0, VHM_DEVICE = 1 };

/* Constant presets: the reference CPU constants (voxel_hash_tsdf.h:194-204,
 * reconstruction_backend.cpp:26-32) or its CUDA constants
 * (cuda_tsdf_integrator.cu:18-35).  The CUDA preset changes constants only:
 * the algorithm is always the CPU path's. */
enum vhm_preset { VHM_PRESET_REFERENCE_CPU = 0, VHM_PRESET_REFERENCE_CUDA = 1 };

/* Pinhole intrinsics, as Intrinsics (math_types.h:200-207). */
typedef struct vhm_intrinsics {
  int32_t width, height;
  float fx, fy, cx, cy;
} vhm_intrinsics;

/* Camera-to-world pose T_wc, as Pose (math_types.h:180-198): R row-major. */
typedef struct vhm_pose {
  float R[9];
  float t[3];
} vhm_pose;

typedef struct vhm_params {
  float voxel_size;            /* VoxelHashTSDF(voxel_size, ...)              */
  float truncation_distance;   /* VoxelHashTSDF(..., truncation_distance)     */
  /* capacity (no reference counterpart: the reference grows host maps)     */
  uint32_t hash_log2_buckets;  /* persistent hash table: 2^k buckets        */
  uint32_t max_blocks;         /* pre-reserved block pool (147,456 B each)  */
  int32_t max_width, max_height;
  /* depth bilateral filter, reconstruction_backend.cpp:26-32 (radius 2)    */
  float filter_sigma_spatial, filter_sigma_range, filter_sigma_angular;
  float filter_disc_base, filter_disc_rel, filter_min_cos;
  /* integrator, voxel_hash_tsdf.h:194-204                                  */
  float max_weight, min_view_angle_weight;
  float normal_sigma_spatial, normal_sigma_range, normal_sigma_angular;
  float disc_base, disc_rel;
  float layer_merge_scale, layer_extract_scale, layer_normal_cos;
} vhm_params;

typedef struct vhm_ctx vhm_ctx;

/* Fills *p with a preset and the given volume geometry / capacity. */
void vhm_default_params(vhm_params* p, int preset, float voxel_size, float truncation_distance);

/* queryCudaBackendStatus (cuda_tsdf_integrator.cu:934-944): 1 if a GPU is
 * usable, 0 otherwise; msg receives the reason either way. */
int vhm_query_device(char* msg, size_t msg_capacity);

/* Free and total memory of the current device (hipMemGetInfo), e.g. to
 * size max_blocks (147,456 B of pool per block). */
int vhm_device_memory(size_t* free_bytes, size_t* total_bytes);

/* Creates a volume on the current device: reserves the block pool and hash
 * table (replaces GpuVolumeCache, cuda_tsdf_integrator.cu:114-167). */
int vhm_create(const vhm_params* p, vhm_ctx** out);
int vhm_destroy(vhm_ctx* ctx);
const char* vhm_last_error(const vhm_ctx* ctx);
/* Clears the volume (all blocks, counters) keeping the reservations. */
int vhm_reset(vhm_ctx* ctx);

/* One frame: filter -> allocate -> integrate.  The GPU path of
 * TsdfIntegrationBackend::integrate (reconstruction_backend.cpp:259-277),
 * i.e. integrateDepthFrameCuda (cuda_tsdf_integrator.cu:946-1010), with the
 * CPU path's constants.  Depth in metres (f32) or raw uint16 scaled by
 * depth_scale (dataset_io.h:342-356).  Asynchronous.
 * flags: VHM_FLAG_PREFILTERED skips the filter (stage-isolated testing). */
#define VHM_FLAG_PREFILTERED 1u
int vhm_integrate_f32(vhm_ctx* ctx, const float* depth, int memkind, const vhm_intrinsics* K,
                      const vhm_pose* T_wc, int allocation_stride, uint32_t flags);
int vhm_integrate_u16(vhm_ctx* ctx, const uint16_t* depth, float depth_scale, int memkind,
                      const vhm_intrinsics* K, const vhm_pose* T_wc, int allocation_stride, uint32_t flags);

/* integratePointCloud (voxel_hash_tsdf.h:75-97, 952-983): n camera-space
 * points (x, y, z floats, host or device memory) seen from T_wc.  Each
 * point's ray is sampled every max(voxel / 2, 2.5 mm) across the truncation
 * band around the point; every sample's block is created and every sample
 * with weight > 0 updates its voxel, in the reference's order (points in
 * order, samples along each ray in order).  Non-finite points are skipped.
 * The volume's frame counter is unchanged.  Synchronous. */
int vhm_integrate_point_cloud(vhm_ctx* ctx, const float* points, size_t n, int memkind, const vhm_pose* T_wc);

/* The depth bilateral filter alone (applyDepthBilateralFilter,
 * reconstruction_backend.cpp:96-189).  Synchronous. */
int vhm_filter_depth_f32(vhm_ctx* ctx, const float* depth, int in_memkind, const vhm_intrinsics* K,
                         float* out, int out_memkind);
int vhm_filter_depth_u16(vhm_ctx* ctx, const uint16_t* depth, float depth_scale, int in_memkind,
                         const vhm_intrinsics* K, float* out, int out_memkind);

/* allocateActiveBlocksCuda (cuda_block_allocator.h:17-22): the frame's
 * active block keys for an (already filtered) depth frame.  Unlike the
 * reference's throw-away per-frame table the keys are also inserted into the
 * persistent table (as prepareFrameIntegration, voxel_hash_tsdf.h:133-139).
 * Writes min(capacity, n) keys; *n_out = n.  Synchronous. */
int vhm_allocate_active_blocks(vhm_ctx* ctx, const float* depth, int memkind, const vhm_intrinsics* K,
                               const vhm_pose* T_wc, int allocation_stride, int32_t* keys_out,
                               size_t capacity, size_t* n_out);

/* Waits for queued work; returns the first sticky device error. */
int vhm_synchronize(vhm_ctx* ctx);

/* blockCount / observedVoxelCount (reconstruction_backend.cpp:293-315;
 * gpuVolumeCacheBlockCount / gpuVolumeCacheObservedVoxelCount,
 * cuda_tsdf_integrator.cu:1017-1052).  Synchronous, O(1). */
int vhm_block_count(vhm_ctx* ctx, size_t* out);
int vhm_observed_voxel_count(vhm_ctx* ctx, size_t* out);

/* Counters since creation/reset: voxel updates (voxels reaching the layer
 * update with weight > 0), voxels visited (512 per active block), frames. */
int vhm_stats(vhm_ctx* ctx, uint64_t* updates, uint64_t* visited, uint64_t* frames);

/* Marching-cubes work since creation/reset (no reference counterpart): the
 * (block, direction) items whose 9x9x9 corner region was staged (read from
 * the pool: 3 slots x 16 B per region voxel), by the full (count pass only)
 * and the incremental extractions.  Synchronous. */
int vhm_mesh_stats(vhm_ctx* ctx, uint64_t* items_staged);

/* gpuVolumeCacheBlockKeys (cuda_tsdf_integrator.cu:1054-1057): all
 * allocated keys in pool order.  Pool order is creation order: frame by
 * frame (point-cloud call by call), and within one frame in increasing
 * packed key (x + 2^20) | (y + 2^20) << 21 | (z + 2^20) << 42, so it depends
 * only on the inputs, not on thread scheduling. */
int vhm_block_keys(vhm_ctx* ctx, int32_t* keys_out, size_t capacity, size_t* n_out);

/* Keys of the blocks holding at least one observed voxel (the blocks
 * blockHasObservedVoxels accepts, voxel_hash_tsdf.h:206-224), in pool order;
 * O(blocks) on the device from the per-block observed bitmasks.  Writes
 * min(capacity, n) keys; *n_out = n.  Synchronous. */
int vhm_observed_block_keys(vhm_ctx* ctx, int32_t* keys_out, size_t capacity, size_t* n_out);

/* Voxel layers of the given blocks in the reference's AoS Voxel layout
 * (voxel_hash_tsdf.h:18-31; 288 B per voxel, 147,456 B per block, local
 * index x + 8(y + 8z)) -- what copyBlocksToFlat (voxel_hash_tsdf.h:147-164)
 * returns per record.  Unallocated keys yield zeros.  Synchronous. */
int vhm_download_blocks(vhm_ctx* ctx, const int32_t* keys, size_t n, void* out_host);

/* Per key, a digest of the bytes vhm_download_blocks would return for it:
 * sum_i w_i * 0x100000001B3^i (mod 2^64) over the block's 36,864 32-bit
 * words w_i in that layout (0 for an unallocated key), computed on the
 * device.  No reference counterpart: it compares whole volumes (parity
 * tests, batched vs one-volume runs) without copying pools to the host.
 * Synchronous. */
int vhm_block_digests(vhm_ctx* ctx, const int32_t* keys, size_t n, uint64_t* out_host);

/* extractMesh / extractMeshCuda (reconstruction_backend.cpp:279-291,
 * cuda_mesh_extractor.cu:636-720).  Two passes on the GPU (count, scan,
 * emit) in a deterministic order: direction, block pool order (see
 * vhm_block_keys), cube, cluster; the same inputs give the same bytes.
 * Writes min(capacity, n) triangles; *n_out = n.  tris_out == NULL with
 * capacity 0 only counts.  Synchronous. */
int vhm_extract_mesh(vhm_ctx* ctx, float iso_level, float* tris_out, size_t capacity, size_t* n_out);
/* vhm_extract_mesh with tris_out in out_memkind memory.  VHM_DEVICE (a
 * device pointer of this context's device): the emit pass writes the
 * triangles in place, with no host copy -- a full mesh kept on the GPU for
 * rendering or further GPU work.  Same triangles, same order; writes
 * min(capacity, n) triangles, *n_out = n.  No reference counterpart (the
 * reference returns the mesh on the host).  Synchronous. */
int vhm_extract_mesh_to(vhm_ctx* ctx, float iso_level, float* tris_out, size_t capacity, int out_memkind,
                        size_t* n_out);

/* Incremental extraction over the (block, direction) pairs changed since the
 * previous incremental call: every direction the integrator (or the
 * denoiser) wrote in a block, in that block and in its -x/-y/-z neighbours,
 * whose cubes read it.  Re-meshes that subset only and clears the dirty
 * masks.  Outputs the listed blocks (keys in pool order, and per block the
 * bit mask of directions re-meshed) and the triangles of those (block,
 * direction) items: tri_item_out[t] = d * n_blocks + i for triangle t of
 * direction d of listed block i.  A caller keeping one mesh per (block,
 * direction) replaces exactly the items in the masks.  No reference
 * counterpart (the reference always meshes the whole volume); each item's
 * triangles equal vhm_extract_mesh's for that block and direction.
 * block_keys_out / block_dirs_out are host memory; tris_out / tri_item_out
 * are in out_memkind memory.  VHM_HOST: two passes (count, scan, emit), the
 * items in item order.  VHM_DEVICE: one pass writes in place on the GPU,
 * nothing copied to the host; each item's triangles are one contiguous run
 * in the same order, the runs in the order the items finished (as the
 * reference's atomic appends).  *n_tris_out > tri_capacity reports an
 * overflow: an item whose run would cross capacity writes no triangle, its
 * entries below capacity get item 0xFFFFFFFF and NaN coordinates, and it is
 * marked dirty again, so the next incremental call lists it; every entry
 * with a valid item holds that item's triangle.  Synchronous. */
int vhm_extract_mesh_incremental(vhm_ctx* ctx, float iso_level, int32_t* block_keys_out, uint8_t* block_dirs_out,
                                 size_t block_capacity, size_t* n_blocks_out, float* tris_out,
                                 uint32_t* tri_item_out, size_t tri_capacity, int out_memkind, size_t* n_tris_out);
/* No reference counterpart: for callers that cannot allocate device memory
 * (code under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  vhm_extract_mesh_incremental with device output
 * (VHM_DEVICE: tris_out / tri_item_out device memory, tri_item_out
 * optional; block_keys_out / block_dirs_out host memory), with the same
 * results, but it allocates nothing: its scratch is a fixed 58 MB device
 * array in libvhm.so, one per GPU, kept until the process ends, shared by
 * every call on that GPU (calls that overlap on one GPU race).  At most
 * 131,072 blocks re-meshed per call: a larger list returns VHM_ERR_CAPACITY
 * and leaves the blocks dirty for a later call.  tri_capacity 0 only
 * counts (every item stays dirty).  Synchronous. */
int vhm_extract_mesh_incremental_noalloc(vhm_ctx* ctx, float iso_level, int32_t* block_keys_out,
                                         uint8_t* block_dirs_out, size_t block_capacity, size_t* n_blocks_out,
                                         float* tris_out, uint32_t* tri_item_out, size_t tri_capacity,
                                         size_t* n_tris_out);
/* Copies again the triangles (and items) of the last incremental extraction
 * that used VHM_HOST output -- all of them, min(capacity, n), for a caller
 * whose first buffer was too small; *n_out = n.  Synchronous. */
int vhm_incremental_mesh_copy(vhm_ctx* ctx, float* tris_out, uint32_t* tri_item_out, size_t capacity, size_t* n_out);

/* Per-kernel GPU time (ms, HIP events on the context's streams) accumulated
 * while timing is enabled: [0] filter, [1] allocate, [2] integrate (the main
 * and the deferred border pass of each frame, summed), [3] mesh;
 * launches[i] counts launches (per frame for [2]).  Enabling resets the
 * accumulators.
 * enabled: 0 off, 1 every stage, or VHM_TIMING_STAGES | (1 << i for each
 * stage i to time): each timed stage records two events per launch, which
 * leave a few microseconds between the launches. */
#define VHM_TIMING_STAGES 0x100
int vhm_set_timing(vhm_ctx* ctx, int enabled);
int vhm_kernel_times(vhm_ctx* ctx, double ms_out[4], uint64_t launches_out[4]);

/* --- learned block denoiser (no reference counterpart) ------------------ */

/* Per 8^3 block with a 1-voxel halo: 12 input channels (per direction d the
 * heaviest slot's plane signed distance at the voxel centre / truncation,
 * clamped to [-1, 1], and its weight / max_weight), three 3x3x3
 * convolutions 12->32->32->6 with ReLU, predicting the 6 tsdf channels of
 * the 8^3 interior as input + residual.  BF16 MFMA, FP32 accumulation.
 * 43,270 FP32 master parameters in the order W1 (3,3,3,12,32), b1 (32),
 * W2 (3,3,3,32,32), b2 (32), W3 (3,3,3,32,6), b3 (6). */
typedef struct vhm_denoiser vhm_denoiser;
int vhm_denoiser_num_params(void);
/* capacity: samples held on the device; batch_capacity: blocks per training step. */
int vhm_denoiser_create(int capacity, int batch_capacity, vhm_denoiser** out);
int vhm_denoiser_destroy(vhm_denoiser* d);
const char* vhm_denoiser_last_error(const vhm_denoiser* d);
int vhm_denoiser_set_params(vhm_denoiser* d, const float* params); /* also resets Adam */
int vhm_denoiser_get_params(vhm_denoiser* d, float* params);
int vhm_denoiser_clear(vhm_denoiser* d);
int vhm_denoiser_count(const vhm_denoiser* d);
/* Appends one sample per key: inputs gathered from input_vol, targets (tsdf
 * and observed mask of the same features) from target_vol when not null. */
int vhm_denoiser_append(vhm_denoiser* d, vhm_ctx* input_vol, vhm_ctx* target_vol, const int32_t* keys, size_t n);
/* Forward over samples [offset, offset+n); y_out (host, n x 512 x 6) optional. */
int vhm_denoiser_forward(vhm_denoiser* d, size_t offset, size_t n, float* y_out);
/* Masked-MSE training step (forward, backward, deterministic gradient
 * reduction, Adam with lr; lr <= 0 leaves the parameters unchanged).
 * loss_out and grad_out (43,270 floats, host) optional. */
int vhm_denoiser_train_step(vhm_denoiser* d, size_t offset, size_t n, float lr, float* loss_out, float* grad_out);
/* Writes the predictions of samples [offset, offset+n) made by the last
 * vhm_denoiser_forward over them (training steps keep none) into vol.  A cell
 * (voxel, direction) is replaced only if its input weight >= min_weight and
 * |prediction - input| <= max_delta (tsdf units); a block whose mean
 * |prediction - input| exceeds block_max_delta keeps plain fusion. */
int vhm_denoiser_apply(vhm_denoiser* d, vhm_ctx* vol, size_t offset, size_t n, float min_weight, float max_delta,
                       float block_max_delta, uint64_t* cells_replaced, uint64_t* blocks_kept_plain);
/* Sample inputs (bf16 bits, n x 1000 x 16), targets and masks (n x 512 x 6). */
int vhm_denoiser_samples(vhm_denoiser* d, size_t offset, size_t n, uint16_t* x_out, float* t_out, float* m_out);
/* Device check of the v_mfma_f32_16x16x16_bf16 operand layout the kernels
 * assume: out[4 * lane + r] = D element held by (lane, r). */
int vhm_mfma_layout_probe(float* out256);
/* The same for v_mfma_f32_16x16x32_bf16 (gfx950): a (16 x 32) and b (32 x 16)
 * row-major (values exact in bf16) placed as the kernels place operands;
 * out[4 * lane + r] = D element held by (lane, r). */
int vhm_mfma_layout_probe_k32(const float* a16x32, const float* b32x16, float* out256);

/* --- batched multi-scene fusion (no reference counterpart) -------------- */

/* S independent volumes (one vhm_ctx each, with its own params) advance one
 * frame per step: one launch per pipeline stage covers every scene (filter,
 * allocation, integration, incremental marching cubes of the step's dirty
 * blocks), and the whole step is captured once as a hipGraph and replayed.
 * Results per scene equal vhm_integrate_u16 frame by frame followed by
 * vhm_extract_mesh_incremental.  Each scene's frames (uint16, device,
 * n_frames x H x W) and poses are given once; scenes with shorter sequences
 * stop integrating once past their last frame.  The volumes are driven by the
 * batch: call vhm_batch_synchronize before reading them through the vhm_ctx
 * functions, and do not integrate into them while batch steps are in flight. */
typedef struct vhm_batch vhm_batch;
int vhm_batch_create(int num_scenes, const vhm_params* params /* [num_scenes] */, vhm_batch** out);
int vhm_batch_destroy(vhm_batch* b);
const char* vhm_batch_last_error(const vhm_batch* b);
vhm_ctx* vhm_batch_volume(vhm_batch* b, int scene);
/* frames_device must stay valid while steps run; tri_capacity: triangles
 * kept per step's incremental mesh. */
int vhm_batch_set_sequence(vhm_batch* b, int scene, const uint16_t* frames_device, int n_frames, float depth_scale,
                           const vhm_intrinsics* K, const vhm_pose* poses, size_t tri_capacity);
/* After every set_sequence; grids <= 0 keep the current values; use_graph
 * captures the step as a hipGraph. */
int vhm_batch_prepare(vhm_batch* b, int integrate_grid, int mesh_grid, int use_graph);
/* Enqueues n steps (asynchronous). */
int vhm_batch_step(vhm_batch* b, int n_steps);
int vhm_batch_synchronize(vhm_batch* b);
int vhm_batch_steps_done(const vhm_batch* b);
/* Clears every volume and restarts all sequences at their first frame
 * (reservations, captured graph and grid sizes are kept). */
int vhm_batch_rewind(vhm_batch* b);
/* Last step's incremental mesh of a scene: blocks re-meshed, triangle total
 * (may exceed tri_capacity), and up to capacity triangles (9 floats, host).
 * Only the last step's output is kept: each step overwrites it. */
int vhm_batch_mesh_output(vhm_batch* b, int scene, uint32_t* blocks, uint32_t* triangles, float* tris_host,
                          size_t capacity);
/* The items of that output, as vhm_extract_mesh_incremental returns them:
 * the listed blocks (keys in pool order, per block the bit mask of
 * directions re-meshed; min(block_capacity, n) written, *n_blocks_out = n)
 * and per triangle its item d * n + i (min(tri_capacity, triangles)
 * written).  Items land in completion order, each one contiguous run; an
 * item that did not fit tri_capacity has item 0xFFFFFFFF and is listed
 * again by the next step.  A caller keeping one mesh per (block, direction)
 * replaces exactly the items in the masks after every step. */
int vhm_batch_mesh_items(vhm_batch* b, int scene, int32_t* block_keys_out, uint8_t* block_dirs_out,
                         size_t block_capacity, size_t* n_blocks_out, uint32_t* tri_item_out, size_t tri_capacity);
/* One step with direct launches, timed per stage: ms_out[0] filter +
 * allocation, [1] integration, [2] incremental meshing; the step's voxel
 * updates and re-meshed blocks over all scenes.  Synchronous. */
int vhm_batch_timed_step(vhm_batch* b, float ms_out[3], uint64_t* updates, uint64_t* remeshed);
/* Runs one step per candidate (multiples of the CU count) with direct
 * launches and keeps, per persistent kernel, the grid with the least time
 * per unit of work (a captured graph is re-captured with them). */
int vhm_batch_autotune(vhm_batch* b, const int* candidates, int n_candidates, int* integrate_grid_out,
                       int* mesh_grid_out);
/* Tile shapes and the integrator's work split: step k runs with candidate k
 * of each list (the last one once a list is exhausted) and the fastest of
 * each is kept -- the filter tile width and the allocation tile width (16,
 * 32 or 64 pixels wide, 256 / width tall; default 32 and 16), each timed
 * by its own events, and the integrator's chunk (0, the default: one
 * contiguous range of the concatenated active lists per workgroup; c > 0:
 * workgroups take c entries at a time by atomic ticket), by time per voxel
 * update.  Consumes max(n_filter, n_alloc, n_chunk) steps of the sequences;
 * results do not depend on the choices.  A captured graph is re-captured. */
int vhm_batch_autotune_tiles(vhm_batch* b, const int* filter_widths, int n_filter, const int* alloc_widths,
                             int n_alloc, const int* chunks, int n_chunk, int* filter_width_out, int* alloc_width_out,
                             int* chunk_out);

/* --- synthetic depth (no reference counterpart; benchmark/test input) ---- */

/* Analytic scene: the interior of an axis-aligned room plus spheres
 * (kind 0: centre, radius) and axis-aligned boxes (kind 1: centre,
 * half-extents). */
typedef struct vhm_primitive {
  int32_t kind;
  float c[3];
  float r;
  float h[3];
} vhm_primitive;

typedef struct vhm_scene {
  float room_min[3], room_max[3];
  int32_t num_primitives;
  const vhm_primitive* primitives; /* host pointer */
} vhm_scene;

/* Depth noise: z + N(0, sigma(z)), sigma(z) = a + b (z - c)^2 metres,
 * dropout probability p, output rounded to millimetres (uint16, 0 = none).
 * Gaussian and dropout draws come from Philox4x32-10 keyed by (seed, frame)
 * with the pixel index as counter, so every pixel is independent of launch
 * shape. */
typedef struct vhm_noise {
  float a, b, c, dropout;
  float max_depth; /* 0 = unlimited */
  uint64_t seed;
  int32_t enabled;
} vhm_noise;

/* Renders n_frames uint16 depth frames (frame i uses poses[i] and frame id
 * first_frame + i) into out (n_frames x H x W).  memkind of out:
 * VHM_DEVICE renders on the GPU; VHM_HOST renders on the host. */
int vhm_synth_render(const vhm_scene* scene, const vhm_intrinsics* K, const vhm_pose* poses, int n_frames,
                     uint32_t first_frame, const vhm_noise* noise, uint16_t* out, int out_memkind);

/* Philox4x32-10 block (for known-answer tests of the generator). */
void vhm_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);

#ifdef __cplusplus
}
#endif

#endif /* VHM_H_ */
