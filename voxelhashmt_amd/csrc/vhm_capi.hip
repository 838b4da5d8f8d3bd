// C-ABI of libvhm.so (declared in include/vhm.h): volume context, frame
// integration, stage entry points, queries, mesh extraction and synthetic
// depth.  Host code only; kernels live in vhm_{filter,alloc,integrate,mesh,
// synth}.hip.
//
// Replaces the reference's GPU backend functions (cuda_tsdf_integrator.h,
// cuda_block_allocator.h, cuda_mesh_extractor.h) and its process-global
// singletons (cuda_tsdf_integrator.cu:114-124) with an explicit context that
// owns a pre-reserved block pool, a persistent hash table and one stream.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <unordered_map>
#include <string>
#include <vector>

#include "../../include/vhm.h"
#include "vhm_internal.h"

#pragma clang fp contract(off)

namespace vhm {
void philox4x32_10(const uint32_t in_ctr[4], const uint32_t in_key[2], uint32_t out[4]);
void synthRenderHost(const SynthDeviceScene& sc, const Intrinsics& K, const Pose34* poses, int n_frames,
                     uint32_t first_frame, const SynthNoise& nz, uint16_t* out);
hipError_t launchActiveKeys(const VolumeView& V, int32_t* out, hipStream_t s, uint32_t max_n);
}  // namespace vhm

using namespace vhm;

namespace {

// ---- no-allocation incremental meshing: fixed device scratch ----
//
// vhm_extract_mesh_incremental_noalloc keeps everything its kernels need
// beyond the context's create-time buffers and the caller's outputs in one
// file-scope device array: the triangle counter, the 8 source blocks and
// the observed-direction mask of every listed block, the listed keys
// staged for the copy to the host, and the cluster-centre columns of the
// fused pass (kMeshFusedCandWords floats per workgroup, at most
// kNoAllocGrid workgroups).  Each GPU holds its own copy for the life of
// the process; calls on one GPU share it, so calls that overlap on one GPU
// race.
constexpr uint32_t kNoAllocMaxListed = 131072;
constexpr uint32_t kNoAllocGrid = 1024;
constexpr size_t kNoAllocCounterOff = 0;
constexpr size_t kNoAllocNbrOff = 256;
constexpr size_t kNoAllocObsOff = kNoAllocNbrOff + size_t(8) * kNoAllocMaxListed * 4;
constexpr size_t kNoAllocKeysOff = kNoAllocObsOff + kNoAllocMaxListed;
constexpr size_t kNoAllocCandOff = kNoAllocKeysOff + size_t(3) * kNoAllocMaxListed * 4;
constexpr size_t kNoAllocBytes = kNoAllocCandOff + size_t(kNoAllocGrid) * kMeshFusedCandWords * sizeof(float);
static_assert(kNoAllocObsOff % 256 == 0 && kNoAllocKeysOff % 256 == 0 && kNoAllocCandOff % 256 == 0,
              "scratch regions stay 256-byte aligned");

}  // namespace

// Host code must not pass `scratch` itself to a kernel or a copy: its name
// is a host-side stand-in.  Its device address on each GPU comes from
// hipGetSymbolAddress (noAllocScratch).
static __device__ __align__(256) uint8_t scratch[kNoAllocBytes];

namespace {

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
uint8_t* g_scratch_ptr[kMaxDevices] = {};

// Device address of `scratch` on `device` (which must be current), looked
// up once per device; nullptr with *e set on failure.
uint8_t* noAllocScratch(int device, hipError_t* e) {
  *e = hipSuccess;
  if (device < 0 || device >= kMaxDevices) {
    *e = hipErrorInvalidDevice;
    return nullptr;
  }
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[device]) {
    void* ptr = nullptr;
    *e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (*e != hipSuccess) return nullptr;
    g_scratch_ptr[device] = static_cast<uint8_t*>(ptr);
  }
  return g_scratch_ptr[device];
}

}  // namespace

struct vhm_ctx {
  vhm_params P{};
  int device = 0;
  int num_cus = 256;
  hipStream_t stream = nullptr;
  // The deferred integration pass of frame f runs on `side`, after f's main
  // pass (main_done), while frame f + 1's filter and allocation run on
  // `stream`; anything else that touches the volume first makes `stream`
  // wait for it (joinSide).
  hipStream_t side = nullptr;
  hipEvent_t main_done = nullptr, side_done = nullptr;
  bool side_pending = false;
  uint32_t frame = 0;
  uint64_t frames_integrated = 0;
  // persistent volume
  uint64_t* keys = nullptr;
  uint32_t* vals = nullptr;
  uint32_t* stamp = nullptr;
  uint64_t* pool_keys = nullptr;
  uint32_t* pool = nullptr;
  uint32_t* obs = nullptr;
  uint32_t* dirty = nullptr;
  uint32_t* dirmask = nullptr;
  uint32_t* active = nullptr;
  DevHeader* hdr = nullptr;
  // per-frame buffers
  void* depth_in = nullptr;  // max W*H*4 bytes (f32 or u16 staging)
  float* filtered = nullptr;
  float* filtered_alt = nullptr;  // the filter output of odd frames (a pending deferred pass reads the other)
  // grown on demand
  uint32_t* mesh_counts = nullptr;
  uint32_t* mesh_offsets = nullptr;
  uint32_t* mesh_partials = nullptr;
  size_t mesh_counts_cap = 0, mesh_offsets_cap = 0, mesh_partials_cap = 0;
  uint8_t* cube_counts = nullptr;  // per-cube triangle counts of the count pass
  size_t cube_counts_cap = 0;
  float* mesh_cand = nullptr;  // marching cubes' cluster-centre scratch (kMeshCandBytes)
  size_t mesh_cand_cap = 0;
  uint32_t* mesh_counters = nullptr;  // fused incremental pass: triangle total
  size_t mesh_counters_cap = 0;
  uint32_t* mesh_nbr = nullptr;  // fused incremental pass: 8 source blocks per listed block
  size_t mesh_nbr_cap = 0;
  uint8_t* mesh_obs_dirs = nullptr;  // fused incremental pass: observed directions over the sources
  size_t mesh_obs_dirs_cap = 0;
  float* tris = nullptr;
  size_t tris_cap = 0;
  int32_t* key_buf = nullptr;
  size_t key_buf_cap = 0;
  uint32_t* gather_buf = nullptr;
  size_t gather_cap = 0;
  // incremental meshing
  uint32_t* remesh_flags = nullptr;  // max_blocks, kept zero between calls
  uint32_t* remesh_pos = nullptr;
  uint32_t* remesh_list = nullptr;
  uint8_t* remesh_dirs = nullptr;  // directions to re-mesh per listed block
  // integrator's deferred border blocks (max_blocks each; flags kept zero between frames)
  uint32_t* defer_flags = nullptr;
  uint32_t* defer_list = nullptr;
  uint64_t* sort_tmp = nullptr;  // block numbering: sorted runs of a pass's new keys (max_blocks)
  // point clouds (grown on demand)
  float* pc_points = nullptr;
  size_t pc_points_cap = 0;
  uint8_t* pc_info = nullptr;
  size_t pc_info_cap = 0;
  uint64_t* pc_keys = nullptr;
  uint64_t* pc_keys_tmp = nullptr;
  uint32_t* pc_vals = nullptr;
  uint32_t* pc_vals_tmp = nullptr;
  size_t pc_keys_cap = 0, pc_keys_tmp_cap = 0, pc_vals_cap = 0, pc_vals_tmp_cap = 0;
  uint32_t* pc_count = nullptr;
  size_t pc_count_cap = 0;
  uint32_t* pc_hist = nullptr;
  uint32_t* pc_offsets = nullptr;
  uint32_t* pc_partials = nullptr;
  size_t pc_hist_cap = 0, pc_offsets_cap = 0, pc_partials_cap = 0;
  uint32_t* remesh_partials = nullptr;
  uint32_t* tri_owner = nullptr;
  size_t tri_owner_cap = 0;
  size_t inc_tris = 0;  // triangles of the last host-output incremental extraction
  // timing
  bool timing = false;
  uint32_t timed_stages = 0;  // bit i: stage i records events
  struct Ev {
    hipEvent_t a, b;
    int kind;
  };
  std::vector<Ev> events;
  double ms[5] = {0, 0, 0, 0, 0};  // [4]: deferred integration passes, reported with [2]
  uint64_t launches[5] = {0, 0, 0, 0, 0};
  std::string err;
};

namespace {

int fail(vhm_ctx* c, int code, const std::string& what) {
  if (c) c->err = what;
  return code;
}

int hipFail(vhm_ctx* c, hipError_t e, const char* action) {
  return fail(c, VHM_ERR_HIP, std::string(action) + " failed: " + hipGetErrorString(e));
}

#define VHM_TRY(ctx, expr, action)                        \
  do {                                                    \
    const hipError_t _e = (expr);                         \
    if (_e != hipSuccess) return hipFail(ctx, _e, action); \
  } while (0)

// Orders `stream` after the pending deferred integration pass.
void joinSide(vhm_ctx* c) {
  if (!c->side_pending) return;
  hipStreamWaitEvent(c->stream, c->side_done, 0);
  c->side_pending = false;
}

// The volume's device view, for work on `stream` that may overlap a pending
// deferred pass: the frame's allocation and numbering only create and
// renumber new blocks, which that pass does not touch.
VolumeView viewOverlapping(vhm_ctx* c) {
  VolumeView V;
  V.keys = c->keys;
  V.vals = c->vals;
  V.stamp = c->stamp;
  V.log2cap = c->P.hash_log2_buckets;
  V.mask = (1u << c->P.hash_log2_buckets) - 1u;
  V.pool_keys = c->pool_keys;
  V.pool = c->pool;
  V.obs = c->obs;
  V.dirty = c->dirty;
  V.dirmask = c->dirmask;
  V.max_blocks = c->P.max_blocks;
  V.active = c->active;
  V.hdr = c->hdr;
  V.defer_flags = c->defer_flags;
  V.defer_list = c->defer_list;
  V.sort_tmp = c->sort_tmp;
  return V;
}

// The volume's device view for any other work on `stream`, ordered after the
// pending deferred pass.
VolumeView view(vhm_ctx* c) {
  joinSide(c);
  return viewOverlapping(c);
}

Intrinsics toK(const vhm_intrinsics* K) { return {K->width, K->height, K->fx, K->fy, K->cx, K->cy}; }

Pose34 toPose(const vhm_pose* p) {
  Pose34 o;
  for (int i = 0; i < 9; ++i) o.R[i] = p->R[i];
  for (int i = 0; i < 3; ++i) o.t[i] = p->t[i];
  return o;
}

// 1 / (2 sigma^2) formed as the reference's constexprs (float).
float invTwoSigmaSq(float s) { return 1.0f / (2.0f * s * s); }

FilterArgs filterArgs(const vhm_ctx* c, const vhm_intrinsics* K, float scale) {
  FilterArgs a{};
  a.K = toK(K);
  a.scale = scale;
  const float inv_s = invTwoSigmaSq(c->P.filter_sigma_spatial);
  for (int dv = -2; dv <= 2; ++dv)
    for (int du = -2; du <= 2; ++du) {
      const float sq = static_cast<float>(du * du + dv * dv);
      a.spatial_w[(dv + 2) * 5 + (du + 2)] = std::exp(-sq * inv_s);
    }
  a.inv_r = invTwoSigmaSq(c->P.filter_sigma_range);
  a.inv_a = invTwoSigmaSq(c->P.filter_sigma_angular);
  a.disc_base = c->P.filter_disc_base;
  a.disc_rel = c->P.filter_disc_rel;
  a.min_cos = c->P.filter_min_cos;
  return a;
}

int checkFrame(vhm_ctx* c, const void* depth, const vhm_intrinsics* K) {
  if (!c) return VHM_ERR_INVALID_ARGUMENT;
  if (!depth || !K) return fail(c, VHM_ERR_INVALID_ARGUMENT, "integrate failed: null argument");
  if (K->width <= 0 || K->height <= 0 || K->width > c->P.max_width || K->height > c->P.max_height)
    return fail(c, VHM_ERR_INVALID_ARGUMENT, "integrate failed: frame size outside the context's reservation");
  if (!(K->fx != 0.0f) || !(K->fy != 0.0f))
    return fail(c, VHM_ERR_INVALID_ARGUMENT, "integrate failed: zero focal length");
  return VHM_OK;
}

// Timing events only read timestamps: created without the system-scope
// fence on record, which would write back and invalidate the caches
// between the stages.
void beginTimed(vhm_ctx* c, int kind, hipEvent_t* b, hipStream_t s = nullptr) {
  *b = nullptr;
  // kind 4 (the deferred integration pass) is timed with the integrate stage
  if (!c->timing || !(c->timed_stages & (1u << (kind == 4 ? 2 : kind)))) return;
  if (!s) s = c->stream;
  hipEvent_t a = nullptr;
  if (hipEventCreateWithFlags(&a, hipEventDisableSystemFence) != hipSuccess) return;  // this launch goes untimed
  if (hipEventCreateWithFlags(b, hipEventDisableSystemFence) != hipSuccess) {
    hipEventDestroy(a);
    *b = nullptr;
    return;
  }
  hipEventRecord(a, s);
  c->events.push_back({a, *b, kind});
}

void endTimed(vhm_ctx* c, hipEvent_t b, hipStream_t s = nullptr) {
  if (c->timing && b) hipEventRecord(b, s ? s : c->stream);
}

void drainEvents(vhm_ctx* c) {
  for (auto& e : c->events) {
    hipEventSynchronize(e.b);
    float ms = 0.0f;
    hipEventElapsedTime(&ms, e.a, e.b);
    c->ms[e.kind] += ms;
    c->launches[e.kind] += 1;
    hipEventDestroy(e.a);
    hipEventDestroy(e.b);
  }
  c->events.clear();
}

// Stage a frame's depth into device memory; returns the device pointer.
template <typename T>
int stageDepth(vhm_ctx* c, const T* depth, int memkind, const vhm_intrinsics* K, const T** dev) {
  if (memkind == VHM_DEVICE) {
    *dev = depth;
    return VHM_OK;
  }
  const size_t bytes = sizeof(T) * static_cast<size_t>(K->width) * K->height;
  VHM_TRY(c, hipMemcpyAsync(c->depth_in, depth, bytes, hipMemcpyHostToDevice, c->stream), "hipMemcpyAsync H2D depth");
  *dev = static_cast<const T*>(c->depth_in);
  return VHM_OK;
}

// begin_frame: the launch also opens frame c->frame (the integration path)
int filterStage(vhm_ctx* c, const void* dev_in, bool u16, float scale, const vhm_intrinsics* K, float* out,
                bool begin_frame = false) {
  FilterArgs fa = filterArgs(c, K, scale);
  if (begin_frame) {
    fa.begin_hdr = c->hdr;
    fa.begin_frame = c->frame;
  }
  hipEvent_t ev{};
  beginTimed(c, 0, &ev);
  const hipError_t e = u16 ? launchFilterU16(static_cast<const uint16_t*>(dev_in), out, fa, c->stream)
                           : launchFilterF32(static_cast<const float*>(dev_in), out, fa, c->stream);
  endTimed(c, ev);
  VHM_TRY(c, e, "depth filter kernel launch");
  return VHM_OK;
}

// frame_open: the filter launch already opened frame c->frame
int allocStage(vhm_ctx* c, const float* dev_depth, const vhm_intrinsics* K, const vhm_pose* T_wc, int stride,
               bool frame_open = false) {
  if (!frame_open) {
    ++c->frame;
    VHM_TRY(c, launchFrameBegin(c->hdr, c->frame, c->stream), "frame begin kernel launch");
  }
  AllocArgs aa{};
  aa.K = toK(K);
  aa.T_wc = toPose(T_wc);
  aa.voxel = c->P.voxel_size;
  aa.trunc = c->P.truncation_distance;
  aa.stride = stride < 1 ? 1 : stride;
  aa.frame = c->frame;
  hipEvent_t ev{};
  beginTimed(c, 1, &ev);
  hipError_t e = launchAllocate(dev_depth, viewOverlapping(c), aa, c->stream);
  if (e == hipSuccess) e = launchNumberBlocks(viewOverlapping(c), c->stream);
  endTimed(c, ev);
  VHM_TRY(c, e, "block allocation kernel launch");
  return VHM_OK;
}

// side: the deferred pass goes to the side stream (dev_depth is one of the
// library's own filter buffers, which the next frame does not overwrite)
int integrateStage(vhm_ctx* c, const float* dev_depth, const vhm_intrinsics* K, const vhm_pose* T_wc,
                   bool side = false) {
  IntegrateArgs ia{};
  ia.K = toK(K);
  ia.T_wc = toPose(T_wc);
  ia.T_cw = inversePose(ia.T_wc);
  ia.voxel = c->P.voxel_size;
  ia.trunc = c->P.truncation_distance;
  ia.merge_tol = c->P.voxel_size * c->P.layer_merge_scale;
  ia.merge_tol = ia.merge_tol < 1e-6f ? 1e-6f : ia.merge_tol;  // std::max(1e-6f, .)
  ia.min_cos = c->P.layer_normal_cos;
  ia.max_weight = c->P.max_weight;
  ia.near_base = meshNearBase(c->P.voxel_size, c->P.layer_extract_scale);
  ia.min_view = c->P.min_view_angle_weight;
  ia.disc_base = c->P.disc_base;
  ia.disc_rel = c->P.disc_rel;
  const float inv_s = invTwoSigmaSq(c->P.normal_sigma_spatial);
  for (int s = 0; s < 3; ++s) ia.spatial_w[s] = std::exp(-static_cast<float>(s) * inv_s);
  ia.inv_r = invTwoSigmaSq(c->P.normal_sigma_range);
  ia.inv_a = invTwoSigmaSq(c->P.normal_sigma_angular);
  focalDivision(ia.K.fx, &ia.inv_fx, &ia.exact_fx);
  focalDivision(ia.K.fy, &ia.inv_fy, &ia.exact_fy);
  ia.defer_slot = c->frame & 1u;
  hipEvent_t ev{};
  beginTimed(c, 2, &ev);
  hipError_t e = launchIntegrate(dev_depth, view(c), ia, c->num_cus * 8, c->stream);
  endTimed(c, ev);
  VHM_TRY(c, e, "integrate kernel launch");
  hipStream_t ds = c->stream;
  if (side) {
    VHM_TRY(c, hipEventRecord(c->main_done, c->stream), "hipEventRecord");
    VHM_TRY(c, hipStreamWaitEvent(c->side, c->main_done, 0), "hipStreamWaitEvent");
    ds = c->side;
  }
  hipEvent_t ev2{};
  beginTimed(c, 4, &ev2, ds);
  e = launchIntegrateDeferred(dev_depth, viewOverlapping(c), ia, ds);
  endTimed(c, ev2, ds);
  VHM_TRY(c, e, "deferred integrate kernel launch");
  if (side) {
    VHM_TRY(c, hipEventRecord(c->side_done, c->side), "hipEventRecord");
    c->side_pending = true;
  }
  ++c->frames_integrated;
  return VHM_OK;
}

template <typename T>
int integrateFrame(vhm_ctx* c, const T* depth, bool u16, float scale, int memkind, const vhm_intrinsics* K,
                   const vhm_pose* T_wc, int stride, uint32_t flags) {
  int rc = checkFrame(c, depth, K);
  if (rc) return rc;
  if (!T_wc) return fail(c, VHM_ERR_INVALID_ARGUMENT, "integrate failed: null pose");
  if (u16 && (flags & VHM_FLAG_PREFILTERED))
    return fail(c, VHM_ERR_INVALID_ARGUMENT, "integrate failed: a prefiltered frame must be float metres");
  const T* dev = nullptr;
  if ((rc = stageDepth(c, depth, memkind, K, &dev))) return rc;
  const float* fdepth = nullptr;
  const bool filter = !(flags & VHM_FLAG_PREFILTERED);
  if (!filter) {
    fdepth = reinterpret_cast<const float*>(dev);
  } else {
    ++c->frame;  // opened by the filter launch
    // frames alternate filter buffers: the previous frame's deferred pass may
    // still read the other one
    float* fbuf = (c->frame & 1u) ? c->filtered_alt : c->filtered;
    if ((rc = filterStage(c, dev, u16, scale, K, fbuf, true))) return rc;
    fdepth = fbuf;
  }
  if ((rc = allocStage(c, fdepth, K, T_wc, stride, filter))) return rc;
  return integrateStage(c, fdepth, K, T_wc, filter);
}

int readHeader(vhm_ctx* c, DevHeader* h) {
  joinSide(c);
  VHM_TRY(c, hipMemcpyAsync(h, c->hdr, sizeof(DevHeader), hipMemcpyDeviceToHost, c->stream), "hipMemcpy D2H header");
  VHM_TRY(c, hipStreamSynchronize(c->stream), "hipStreamSynchronize");
  if (c->timing) drainEvents(c);
  if (h->error & kErrKeyRange)
    return fail(c, VHM_ERR_KEY_RANGE, "block allocation failed: block coordinate outside +-2^20");
  if (h->error & (kErrHashFull | kErrPoolFull))
    return fail(c, VHM_ERR_CAPACITY,
                std::string("block allocation failed: ") +
                    ((h->error & kErrPoolFull) ? "block pool full (max_blocks)" : "hash table full") +
                    "; create the volume with a larger reservation");
  return VHM_OK;
}

template <typename T>
int ensure(vhm_ctx* c, T** buf, size_t* cap, size_t n, const char* what) {
  if (n <= *cap) return VHM_OK;
  if (*buf) hipFree(*buf);
  *buf = nullptr;
  *cap = 0;
  VHM_TRY(c, hipMalloc(reinterpret_cast<void**>(buf), sizeof(T) * n), what);
  *cap = n;
  return VHM_OK;
}

int initVolume(vhm_ctx* c) {
  const size_t cap = size_t(1) << c->P.hash_log2_buckets;
  const hipStream_t s = c->stream;
  VHM_TRY(c, hipMemsetAsync(c->keys, 0xFF, cap * sizeof(uint64_t), s), "hipMemset keys");
  VHM_TRY(c, hipMemsetAsync(c->vals, 0xFF, cap * sizeof(uint32_t), s), "hipMemset vals");
  VHM_TRY(c, hipMemsetAsync(c->stamp, 0, cap * sizeof(uint32_t), s), "hipMemset stamps");
  VHM_TRY(c, hipMemsetAsync(c->pool, 0, size_t(c->P.max_blocks) * kBlockWords * 4u, s), "hipMemset pool");
  VHM_TRY(c, hipMemsetAsync(c->obs, 0, size_t(c->P.max_blocks) * 16u * 4u, s), "hipMemset observed");
  VHM_TRY(c, hipMemsetAsync(c->dirty, 0, size_t(c->P.max_blocks) * 4u, s), "hipMemset dirty");
  VHM_TRY(c, hipMemsetAsync(c->dirmask, 0, size_t(c->P.max_blocks) * 4u, s), "hipMemset direction masks");
  VHM_TRY(c, hipMemsetAsync(c->remesh_flags, 0, size_t(c->P.max_blocks) * 4u, s), "hipMemset remesh flags");
  VHM_TRY(c, hipMemsetAsync(c->defer_flags, 0, size_t(c->P.max_blocks) * 4u, s), "hipMemset deferred flags");
  VHM_TRY(c, hipMemsetAsync(c->hdr, 0, sizeof(DevHeader), s), "hipMemset header");
  VHM_TRY(c, hipStreamSynchronize(s), "hipStreamSynchronize");
  c->frame = 0;
  c->frames_integrated = 0;
  return VHM_OK;
}

}  // namespace

extern "C" {

void vhm_default_params(vhm_params* p, int preset, float voxel_size, float truncation_distance) {
  std::memset(p, 0, sizeof(*p));
  p->voxel_size = voxel_size;
  p->truncation_distance = truncation_distance;
  p->hash_log2_buckets = 20;
  p->max_blocks = 65536;
  p->max_width = 1280;
  p->max_height = 720;
  // reconstruction_backend.cpp:26-32
  p->filter_sigma_spatial = 1.5f;
  p->filter_sigma_range = 0.0012f;
  p->filter_sigma_angular = 0.35f;
  p->filter_disc_base = 0.0012f;
  p->filter_disc_rel = 0.0025f;
  p->filter_min_cos = 0.6f;
  // voxel_hash_tsdf.h:194-204
  p->max_weight = 100.0f;
  p->min_view_angle_weight = 0.15f;
  p->normal_sigma_spatial = 1.0f;
  p->normal_sigma_range = 0.0012f;
  p->normal_sigma_angular = 0.35f;
  p->disc_base = 0.0012f;
  p->disc_rel = 0.0025f;
  p->layer_merge_scale = 0.25f;
  p->layer_extract_scale = 0.55f;
  p->layer_normal_cos = 0.75f;
  if (preset == VHM_PRESET_REFERENCE_CUDA) {  // cuda_tsdf_integrator.cu:27-35
    p->disc_base = 0.012f;
    p->disc_rel = 0.025f;
    p->filter_disc_base = 0.012f;
    p->filter_disc_rel = 0.025f;
    p->layer_normal_cos = 0.9f;
    p->filter_min_cos = 0.9f;
  }
}

int vhm_query_device(char* msg, size_t cap) {
  int n = 0;
  const hipError_t e = hipGetDeviceCount(&n);
  std::string m;
  int ok = 0;
  if (e != hipSuccess)
    m = hipGetErrorString(e);
  else if (n <= 0)
    m = "no HIP device detected";
  else {
    hipDeviceProp_t prop{};
    hipGetDeviceProperties(&prop, 0);
    m = std::to_string(n) + " HIP device(s) detected (" + prop.gcnArchName + ")";
    ok = 1;
  }
  if (msg && cap) std::snprintf(msg, cap, "%s", m.c_str());
  return ok;
}

int vhm_device_memory(size_t* free_bytes, size_t* total_bytes) {
  if (!free_bytes || !total_bytes) return VHM_ERR_INVALID_ARGUMENT;
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) return VHM_ERR_NO_DEVICE;
  return hipMemGetInfo(free_bytes, total_bytes) == hipSuccess ? VHM_OK : VHM_ERR_HIP;
}

int vhm_create(const vhm_params* p, vhm_ctx** out) {
  if (!p || !out) return VHM_ERR_INVALID_ARGUMENT;
  *out = nullptr;
  auto* c = new vhm_ctx();
  c->P = *p;
  auto bad = [&](const char*) {
    delete c;
    return VHM_ERR_INVALID_ARGUMENT;
  };
  if (!(p->voxel_size > 0.0f) || !(p->truncation_distance > 0.0f)) return bad("voxel_size/truncation must be > 0");
  if (p->hash_log2_buckets < 4 || p->hash_log2_buckets > 30) return bad("hash_log2_buckets outside [4, 30]");
  if (p->max_blocks == 0 || p->max_blocks > (1u << 31)) return bad("max_blocks outside [1, 2^31]");
  if (p->max_width <= 0 || p->max_height <= 0) return bad("max_width/max_height must be > 0");
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) {
    delete c;
    return VHM_ERR_NO_DEVICE;
  }
  hipGetDevice(&c->device);
  hipDeviceProp_t prop{};
  hipGetDeviceProperties(&prop, c->device);
  c->num_cus = prop.multiProcessorCount > 0 ? prop.multiProcessorCount : 256;
  hipError_t e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
  if (e == hipSuccess) e = hipStreamCreateWithFlags(&c->side, hipStreamNonBlocking);
  if (e == hipSuccess) e = hipEventCreateWithFlags(&c->main_done, hipEventDisableTiming);
  if (e == hipSuccess) e = hipEventCreateWithFlags(&c->side_done, hipEventDisableTiming);
  const size_t cap = size_t(1) << p->hash_log2_buckets;
  const size_t px = size_t(p->max_width) * p->max_height;
  if (e == hipSuccess) e = hipMalloc(&c->keys, cap * sizeof(uint64_t));
  if (e == hipSuccess) e = hipMalloc(&c->vals, cap * sizeof(uint32_t));
  if (e == hipSuccess) e = hipMalloc(&c->stamp, cap * sizeof(uint32_t));
  if (e == hipSuccess) e = hipMalloc(&c->active, cap * sizeof(uint32_t));
  if (e == hipSuccess) e = hipMalloc(&c->pool_keys, size_t(p->max_blocks) * sizeof(uint64_t));
  if (e == hipSuccess) e = hipMalloc(&c->pool, size_t(p->max_blocks) * kBlockWords * 4u);
  if (e == hipSuccess) e = hipMalloc(&c->obs, size_t(p->max_blocks) * 16u * 4u);
  if (e == hipSuccess) e = hipMalloc(&c->dirty, size_t(p->max_blocks) * 4u);
  if (e == hipSuccess) e = hipMalloc(&c->dirmask, size_t(p->max_blocks) * 4u);
  if (e == hipSuccess) e = hipMalloc(&c->hdr, sizeof(DevHeader));
  if (e == hipSuccess) e = hipMalloc(&c->depth_in, px * 4u);
  if (e == hipSuccess) e = hipMalloc(&c->filtered, px * 4u);
  if (e == hipSuccess) e = hipMalloc(&c->filtered_alt, px * 4u);
  if (e == hipSuccess) e = hipMalloc(&c->remesh_flags, size_t(p->max_blocks) * 4u);
  if (e == hipSuccess) e = hipMalloc(&c->remesh_pos, size_t(p->max_blocks) * 4u);
  if (e == hipSuccess) e = hipMalloc(&c->remesh_list, size_t(p->max_blocks) * 4u);
  if (e == hipSuccess) e = hipMalloc(&c->remesh_dirs, size_t(p->max_blocks));
  if (e == hipSuccess) e = hipMalloc(&c->defer_flags, size_t(p->max_blocks) * 4u);
  if (e == hipSuccess) e = hipMalloc(&c->defer_list, size_t(p->max_blocks) * 4u);
  if (e == hipSuccess) e = hipMalloc(&c->sort_tmp, size_t(p->max_blocks) * 8u);
  if (e == hipSuccess) e = hipMalloc(&c->remesh_partials, scanPartialsSize(p->max_blocks) * 4u);
  if (e != hipSuccess) {
    vhm_destroy(c);
    return VHM_ERR_HIP;
  }
  if (initVolume(c) != VHM_OK) {
    vhm_destroy(c);
    return VHM_ERR_HIP;
  }
  hipError_t se = hipSuccess;
  if (!noAllocScratch(c->device, &se)) {
    vhm_destroy(c);
    return VHM_ERR_HIP;
  }
  *out = c;
  return VHM_OK;
}

int vhm_destroy(vhm_ctx* c) {
  if (!c) return VHM_OK;
  if (c->side) hipStreamSynchronize(c->side);
  if (c->stream) hipStreamSynchronize(c->stream);
  for (auto& ev : c->events) {
    hipEventDestroy(ev.a);
    hipEventDestroy(ev.b);
  }
  void* bufs[] = {c->keys,         c->vals,         c->stamp,         c->active, c->pool_keys, c->pool,
                  c->obs,          c->dirty,  c->dirmask,        c->hdr,           c->depth_in, c->filtered, c->filtered_alt, c->mesh_counts,
                  c->mesh_offsets, c->mesh_partials, c->tris,          c->key_buf, c->gather_buf,
                  c->remesh_flags, c->remesh_pos,    c->remesh_list,   c->remesh_partials, c->tri_owner, c->cube_counts, c->remesh_dirs, c->mesh_cand, c->mesh_counters, c->mesh_nbr, c->mesh_obs_dirs,
                  c->defer_flags, c->defer_list, c->sort_tmp, c->pc_points, c->pc_info, c->pc_keys, c->pc_keys_tmp,
                  c->pc_vals, c->pc_vals_tmp, c->pc_count, c->pc_hist, c->pc_offsets, c->pc_partials};
  for (void* b : bufs)
    if (b) hipFree(b);
  if (c->main_done) hipEventDestroy(c->main_done);
  if (c->side_done) hipEventDestroy(c->side_done);
  if (c->side) hipStreamDestroy(c->side);
  if (c->stream) hipStreamDestroy(c->stream);
  delete c;
  return VHM_OK;
}

const char* vhm_last_error(const vhm_ctx* c) { return c ? c->err.c_str() : "null context"; }

int vhm_reset(vhm_ctx* c) {
  if (!c) return VHM_ERR_INVALID_ARGUMENT;
  joinSide(c);
  return initVolume(c);
}

int vhm_integrate_f32(vhm_ctx* c, const float* depth, int memkind, const vhm_intrinsics* K, const vhm_pose* T_wc,
                      int stride, uint32_t flags) {
  return integrateFrame<float>(c, depth, false, 1.0f, memkind, K, T_wc, stride, flags);
}

int vhm_integrate_u16(vhm_ctx* c, const uint16_t* depth, float scale, int memkind, const vhm_intrinsics* K,
                      const vhm_pose* T_wc, int stride, uint32_t flags) {
  return integrateFrame<uint16_t>(c, depth, true, scale, memkind, K, T_wc, stride, flags);
}

static int filterEntry(vhm_ctx* c, const void* depth, bool u16, float scale, int in_kind, const vhm_intrinsics* K,
                       float* out, int out_kind) {
  int rc = checkFrame(c, depth, K);
  if (rc) return rc;
  joinSide(c);  // c->filtered may be a pending deferred pass's input
  if (!out) return fail(c, VHM_ERR_INVALID_ARGUMENT, "filter failed: null output");
  const void* dev = nullptr;
  if (u16)
    rc = stageDepth(c, static_cast<const uint16_t*>(depth), in_kind, K, reinterpret_cast<const uint16_t**>(&dev));
  else
    rc = stageDepth(c, static_cast<const float*>(depth), in_kind, K, reinterpret_cast<const float**>(&dev));
  if (rc) return rc;
  float* dst = out_kind == VHM_DEVICE ? out : c->filtered;
  if ((rc = filterStage(c, dev, u16, scale, K, dst))) return rc;
  if (out_kind != VHM_DEVICE)
    VHM_TRY(c, hipMemcpyAsync(out, c->filtered, sizeof(float) * size_t(K->width) * K->height, hipMemcpyDeviceToHost,
                              c->stream),
            "hipMemcpy D2H filtered depth");
  VHM_TRY(c, hipStreamSynchronize(c->stream), "hipStreamSynchronize");
  return VHM_OK;
}

int vhm_filter_depth_f32(vhm_ctx* c, const float* depth, int in_kind, const vhm_intrinsics* K, float* out,
                         int out_kind) {
  return filterEntry(c, depth, false, 1.0f, in_kind, K, out, out_kind);
}

int vhm_filter_depth_u16(vhm_ctx* c, const uint16_t* depth, float scale, int in_kind, const vhm_intrinsics* K,
                         float* out, int out_kind) {
  return filterEntry(c, depth, true, scale, in_kind, K, out, out_kind);
}

int vhm_allocate_active_blocks(vhm_ctx* c, const float* depth, int memkind, const vhm_intrinsics* K,
                               const vhm_pose* T_wc, int stride, int32_t* keys_out, size_t capacity, size_t* n_out) {
  int rc = checkFrame(c, depth, K);
  if (rc) return rc;
  if (!T_wc || !n_out) return fail(c, VHM_ERR_INVALID_ARGUMENT, "allocate failed: null argument");
  joinSide(c);
  const float* dev = nullptr;
  if ((rc = stageDepth(c, depth, memkind, K, &dev))) return rc;
  if ((rc = allocStage(c, dev, K, T_wc, stride))) return rc;
  DevHeader h{};
  if ((rc = readHeader(c, &h))) return rc;
  *n_out = h.active_count;
  const size_t m = capacity < h.active_count ? capacity : h.active_count;
  if (m > 0 && keys_out) {
    if ((rc = ensure(c, &c->key_buf, &c->key_buf_cap, 3 * size_t(h.active_count), "hipMalloc key buffer")))
      return rc;
    VHM_TRY(c, launchActiveKeys(view(c), c->key_buf, c->stream, h.active_count), "active key gather launch");
    VHM_TRY(c, hipMemcpyAsync(keys_out, c->key_buf, 3 * m * sizeof(int32_t), hipMemcpyDeviceToHost, c->stream),
            "hipMemcpy D2H keys");
    VHM_TRY(c, hipStreamSynchronize(c->stream), "hipStreamSynchronize");
  }
  return VHM_OK;
}

// Records per point-cloud chunk: keys, values and their sort temporaries
// take 24 B each (800 MB at this bound, which holds a 1280x720 frame's
// cloud at 19 samples per ray in one chunk).
constexpr size_t kPointCloudMaxRecords = size_t(1) << 25;

int vhm_integrate_point_cloud(vhm_ctx* c, const float* points, size_t n, int memkind, const vhm_pose* T_wc) {
  if (!c) return VHM_ERR_INVALID_ARGUMENT;
  joinSide(c);
  if (!T_wc || (n && !points)) return fail(c, VHM_ERR_INVALID_ARGUMENT, "point cloud integration failed: null argument");
  if (memkind != VHM_HOST && memkind != VHM_DEVICE)
    return fail(c, VHM_ERR_INVALID_ARGUMENT, "point cloud integration failed: bad memkind");
  if (n == 0) return VHM_OK;
  PointCloudArgs a{};
  a.T_wc = toPose(T_wc);
  a.voxel = c->P.voxel_size;
  a.trunc = c->P.truncation_distance;
  a.step = std::max(0.5f * c->P.voxel_size, 0.0025f);  // voxel_hash_tsdf.h:77
  a.min_view = c->P.min_view_angle_weight;
  a.merge_tol = std::max(1e-6f, c->P.voxel_size * c->P.layer_merge_scale);
  a.min_cos = c->P.layer_normal_cos;
  a.max_weight = c->P.max_weight;
  if (!(a.trunc >= 0.0f) || !std::isfinite(a.trunc) || !(a.step > 0.0f))
    return fail(c, VHM_ERR_INVALID_ARGUMENT, "point cloud integration failed: bad truncation or voxel size");
  // samples of one ray: at most floor(2 trunc / step) + 1, plus one for the
  // rounding of the running sum and one of margin
  const double per_ray = std::floor(2.0 * double(a.trunc) / double(a.step)) + 3.0;
  if (per_ray > double(kPointCloudMaxRecords))
    return fail(c, VHM_ERR_INVALID_ARGUMENT, "point cloud integration failed: truncation / step too large");
  a.samples_max = static_cast<uint32_t>(per_ray);
  // key = slot (log2 buckets) | voxel (9) | direction (3) | sequence: at most 64 bits
  const int free_bits = 64 - int(c->P.hash_log2_buckets) - 12;
  size_t chunk = kPointCloudMaxRecords / a.samples_max;
  const size_t by_bits = free_bits >= 62 ? chunk : (size_t(1) << free_bits) / a.samples_max;
  chunk = std::min(std::min(chunk, by_bits), n);
  if (chunk == 0) return fail(c, VHM_ERR_INVALID_ARGUMENT, "point cloud integration failed: hash table too large for the record key");
  int seq_bits = 0;
  while ((uint64_t(1) << seq_bits) < uint64_t(chunk) * a.samples_max) ++seq_bits;
  a.seq_bits = seq_bits;
  const int key_bits = int(c->P.hash_log2_buckets) + 12 + seq_bits;
  const size_t recs = chunk * a.samples_max;
  int rc;
  if ((rc = ensure(c, &c->pc_info, &c->pc_info_cap, chunk * pointInfoBytes(), "hipMalloc point info"))) return rc;
  if ((rc = ensure(c, &c->pc_keys, &c->pc_keys_cap, recs, "hipMalloc point records"))) return rc;
  if ((rc = ensure(c, &c->pc_keys_tmp, &c->pc_keys_tmp_cap, recs, "hipMalloc point records"))) return rc;
  if ((rc = ensure(c, &c->pc_vals, &c->pc_vals_cap, recs, "hipMalloc point records"))) return rc;
  if ((rc = ensure(c, &c->pc_vals_tmp, &c->pc_vals_tmp_cap, recs, "hipMalloc point records"))) return rc;
  if ((rc = ensure(c, &c->pc_count, &c->pc_count_cap, 1, "hipMalloc record count"))) return rc;
  const size_t hist = radixSortCountsSize(static_cast<uint32_t>(recs));
  if ((rc = ensure(c, &c->pc_hist, &c->pc_hist_cap, hist, "hipMalloc sort counts"))) return rc;
  if ((rc = ensure(c, &c->pc_offsets, &c->pc_offsets_cap, hist, "hipMalloc sort counts"))) return rc;
  if ((rc = ensure(c, &c->pc_partials, &c->pc_partials_cap, scanPartialsSize(static_cast<uint32_t>(hist)),
                   "hipMalloc sort partials")))
    return rc;
  if (memkind == VHM_HOST &&
      (rc = ensure(c, &c->pc_points, &c->pc_points_cap, 3 * chunk, "hipMalloc point staging")))
    return rc;
  const VolumeView V = view(c);
  for (size_t first = 0; first < n; first += chunk) {
    const uint32_t m = static_cast<uint32_t>(std::min(chunk, n - first));
    const float* pts = points + 3 * first;
    if (memkind == VHM_HOST) {
      VHM_TRY(c, hipMemcpyAsync(c->pc_points, pts, sizeof(float) * 3 * m, hipMemcpyHostToDevice, c->stream),
              "hipMemcpy H2D points");
      pts = c->pc_points;
    }
    VHM_TRY(c, hipMemsetAsync(c->pc_count, 0, sizeof(uint32_t), c->stream), "hipMemset record count");
    VHM_TRY(c, launchNumberBegin(c->hdr, c->stream), "block numbering kernel launch");
    VHM_TRY(c, launchPointCloudSample(pts, m, V, a, c->pc_info, c->pc_keys, c->pc_vals, c->pc_count, c->stream),
            "point cloud sample kernel launch");
    VHM_TRY(c, launchNumberBlocks(V, c->stream), "block numbering kernel launch");
    uint32_t nrec = 0;
    VHM_TRY(c, hipMemcpyAsync(&nrec, c->pc_count, sizeof(uint32_t), hipMemcpyDeviceToHost, c->stream),
            "hipMemcpy D2H record count");
    VHM_TRY(c, hipStreamSynchronize(c->stream), "hipStreamSynchronize");
    VHM_TRY(c, launchRadixSort(c->pc_keys, c->pc_vals, c->pc_keys_tmp, c->pc_vals_tmp, c->pc_count, nrec, key_bits,
                               c->pc_hist, c->pc_offsets, c->pc_partials, c->stream),
            "point cloud sort launch");
    VHM_TRY(c, launchPointCloudApply(c->pc_keys, c->pc_vals, c->pc_count, nrec, c->pc_info, V, a, c->stream),
            "point cloud apply kernel launch");
  }
  DevHeader h{};
  rc = readHeader(c, &h);
  if (rc) return rc;
  if (h.error & kErrSampleBound)
    return fail(c, VHM_ERR_INVALID_ARGUMENT, "point cloud integration failed: sample bound exceeded");
  return VHM_OK;
}

int vhm_synchronize(vhm_ctx* c) {
  if (!c) return VHM_ERR_INVALID_ARGUMENT;
  DevHeader h{};
  return readHeader(c, &h);
}

int vhm_block_count(vhm_ctx* c, size_t* out) {
  if (!c || !out) return VHM_ERR_INVALID_ARGUMENT;
  DevHeader h{};
  const int rc = readHeader(c, &h);
  *out = h.num_blocks < c->P.max_blocks ? h.num_blocks : c->P.max_blocks;
  return rc;
}

int vhm_observed_voxel_count(vhm_ctx* c, size_t* out) {
  if (!c || !out) return VHM_ERR_INVALID_ARGUMENT;
  DevHeader h{};
  const int rc = readHeader(c, &h);
  *out = static_cast<size_t>(h.observed);
  return rc;
}

int vhm_stats(vhm_ctx* c, uint64_t* updates, uint64_t* visited, uint64_t* frames) {
  if (!c) return VHM_ERR_INVALID_ARGUMENT;
  DevHeader h{};
  const int rc = readHeader(c, &h);
  if (updates) *updates = h.updates;
  if (visited) *visited = h.visited;
  if (frames) *frames = c->frames_integrated;
  return rc;
}

int vhm_mesh_stats(vhm_ctx* c, uint64_t* items_staged) {
  if (!c) return VHM_ERR_INVALID_ARGUMENT;
  DevHeader h{};
  const int rc = readHeader(c, &h);
  if (items_staged) *items_staged = h.mesh_items;
  return rc;
}

int vhm_block_keys(vhm_ctx* c, int32_t* keys_out, size_t capacity, size_t* n_out) {
  size_t n = 0;
  int rc = vhm_block_count(c, &n);
  if (rc) return rc;
  if (n_out) *n_out = n;
  const size_t m = capacity < n ? capacity : n;
  if (m == 0 || !keys_out) return VHM_OK;
  std::vector<uint64_t> packed(m);
  VHM_TRY(c, hipMemcpy(packed.data(), c->pool_keys, m * sizeof(uint64_t), hipMemcpyDeviceToHost),
          "hipMemcpy D2H pool keys");
  for (size_t i = 0; i < m; ++i) {
    keys_out[3 * i] = keyX(packed[i]);
    keys_out[3 * i + 1] = keyY(packed[i]);
    keys_out[3 * i + 2] = keyZ(packed[i]);
  }
  return VHM_OK;
}

int vhm_incremental_mesh_copy(vhm_ctx* c, float* tris_out, uint32_t* tri_item_out, size_t capacity, size_t* n_out) {
  if (!c || !n_out) return VHM_ERR_INVALID_ARGUMENT;
  *n_out = c->inc_tris;
  const size_t t = capacity < c->inc_tris ? capacity : c->inc_tris;
  if (t == 0) return VHM_OK;
  if (tris_out)
    VHM_TRY(c, hipMemcpyAsync(tris_out, c->tris, t * 9u * sizeof(float), hipMemcpyDeviceToHost, c->stream),
            "hipMemcpy D2H triangles");
  if (tri_item_out)
    VHM_TRY(c, hipMemcpyAsync(tri_item_out, c->tri_owner, t * sizeof(uint32_t), hipMemcpyDeviceToHost, c->stream),
            "hipMemcpy D2H triangle owners");
  VHM_TRY(c, hipStreamSynchronize(c->stream), "hipStreamSynchronize");
  return VHM_OK;
}

int vhm_observed_block_keys(vhm_ctx* c, int32_t* keys_out, size_t capacity, size_t* n_out) {
  if (!c || !n_out) return VHM_ERR_INVALID_ARGUMENT;
  DevHeader h{};
  int rc = readHeader(c, &h);
  if (rc) return rc;
  *n_out = 0;
  const uint32_t nb = h.num_blocks < c->P.max_blocks ? h.num_blocks : c->P.max_blocks;
  if (nb == 0) return VHM_OK;
  // the remesh flags are all zero between calls; the compaction clears them again
  VHM_TRY(c, launchObservedFlags(view(c), nb, c->remesh_flags, c->stream), "observed flags kernel launch");
  VHM_TRY(c, launchScan(c->remesh_flags, c->remesh_pos, nb, c->remesh_partials, c->stream, true),
          "observed scan launch");
  VHM_TRY(c, launchCompactFlags(c->remesh_flags, c->remesh_pos, nb, c->remesh_list, nullptr, c->stream),
          "observed compact launch");
  uint32_t m = 0;
  VHM_TRY(c, hipMemcpyAsync(&m, c->remesh_partials + (scanPartialsSize(nb) - 1), sizeof(uint32_t),
                            hipMemcpyDeviceToHost, c->stream),
          "hipMemcpy D2H observed count");
  VHM_TRY(c, hipStreamSynchronize(c->stream), "hipStreamSynchronize");
  *n_out = m;
  const size_t k = capacity < m ? capacity : m;
  if (k == 0 || !keys_out) return VHM_OK;
  if ((rc = ensure(c, &c->key_buf, &c->key_buf_cap, 3 * size_t(m), "hipMalloc key buffer"))) return rc;
  VHM_TRY(c, launchListKeys(view(c), c->remesh_list, m, c->key_buf, c->stream), "observed key launch");
  VHM_TRY(c, hipMemcpyAsync(keys_out, c->key_buf, 3 * k * sizeof(int32_t), hipMemcpyDeviceToHost, c->stream),
          "hipMemcpy D2H observed keys");
  VHM_TRY(c, hipStreamSynchronize(c->stream), "hipStreamSynchronize");
  return VHM_OK;
}

int vhm_download_blocks(vhm_ctx* c, const int32_t* keys, size_t n, void* out_host) {
  if (!c) return VHM_ERR_INVALID_ARGUMENT;
  if (n == 0) return VHM_OK;
  if (!keys || !out_host) return fail(c, VHM_ERR_INVALID_ARGUMENT, "download failed: null argument");
  DevHeader h{};
  int rc = readHeader(c, &h);
  if (rc) return rc;
  const size_t chunk = 512;
  if ((rc = ensure(c, &c->key_buf, &c->key_buf_cap, 3 * chunk, "hipMalloc key buffer"))) return rc;
  if ((rc = ensure(c, &c->gather_buf, &c->gather_cap, chunk * kBlockWords, "hipMalloc gather buffer"))) return rc;
  for (size_t off = 0; off < n; off += chunk) {
    const size_t m = (n - off) < chunk ? (n - off) : chunk;
    VHM_TRY(c, hipMemcpyAsync(c->key_buf, keys + 3 * off, 3 * m * sizeof(int32_t), hipMemcpyHostToDevice, c->stream),
            "hipMemcpy H2D keys");
    VHM_TRY(c, launchGatherBlocksAoS(c->key_buf, static_cast<uint32_t>(m), view(c), c->gather_buf, c->stream),
            "block gather launch");
    VHM_TRY(c, hipMemcpyAsync(static_cast<uint8_t*>(out_host) + off * kBlockWords * 4u, c->gather_buf,
                              m * kBlockWords * 4u, hipMemcpyDeviceToHost, c->stream),
            "hipMemcpy D2H blocks");
    VHM_TRY(c, hipStreamSynchronize(c->stream), "hipStreamSynchronize");
  }
  return VHM_OK;
}

int vhm_block_digests(vhm_ctx* c, const int32_t* keys, size_t n, uint64_t* out_host) {
  if (!c) return VHM_ERR_INVALID_ARGUMENT;
  if (n == 0) return VHM_OK;
  if (!keys || !out_host) return fail(c, VHM_ERR_INVALID_ARGUMENT, "block digest failed: null argument");
  DevHeader h{};
  int rc = readHeader(c, &h);
  if (rc) return rc;
  const size_t chunk = size_t(1) << 20;
  const size_t m0 = n < chunk ? n : chunk;
  if ((rc = ensure(c, &c->key_buf, &c->key_buf_cap, 3 * m0, "hipMalloc key buffer"))) return rc;
  if ((rc = ensure(c, &c->gather_buf, &c->gather_cap, 2 * m0, "hipMalloc digest buffer"))) return rc;
  for (size_t off = 0; off < n; off += chunk) {
    const size_t m = (n - off) < chunk ? (n - off) : chunk;
    VHM_TRY(c, hipMemcpyAsync(c->key_buf, keys + 3 * off, 3 * m * sizeof(int32_t), hipMemcpyHostToDevice, c->stream),
            "hipMemcpy H2D keys");
    uint64_t* dig = reinterpret_cast<uint64_t*>(c->gather_buf);
    VHM_TRY(c, launchBlockDigests(c->key_buf, static_cast<uint32_t>(m), view(c), dig, c->stream),
            "block digest launch");
    VHM_TRY(c, hipMemcpyAsync(out_host + off, dig, m * sizeof(uint64_t), hipMemcpyDeviceToHost, c->stream),
            "hipMemcpy D2H digests");
    VHM_TRY(c, hipStreamSynchronize(c->stream), "hipStreamSynchronize");
  }
  return VHM_OK;
}

int vhm_extract_mesh(vhm_ctx* c, float iso, float* tris_out, size_t capacity, size_t* n_out) {
  return vhm_extract_mesh_to(c, iso, tris_out, capacity, VHM_HOST, n_out);
}

int vhm_extract_mesh_to(vhm_ctx* c, float iso, float* tris_out, size_t capacity, int out_memkind, size_t* n_out) {
  if (!c || !n_out || (out_memkind != VHM_HOST && out_memkind != VHM_DEVICE)) return VHM_ERR_INVALID_ARGUMENT;
  DevHeader h{};
  int rc = readHeader(c, &h);
  if (rc) return rc;
  const uint32_t nb = h.num_blocks < c->P.max_blocks ? h.num_blocks : c->P.max_blocks;
  *n_out = 0;
  if (nb == 0) return VHM_OK;
  const size_t entries = size_t(nb) * kDirs;
  if (entries > 0xffffffffull) return fail(c, VHM_ERR_CAPACITY, "mesh extraction failed: too many blocks");
  if ((rc = ensure(c, &c->mesh_counts, &c->mesh_counts_cap, entries, "hipMalloc mesh counts"))) return rc;
  if ((rc = ensure(c, &c->cube_counts, &c->cube_counts_cap, entries * 512, "hipMalloc cube counts"))) return rc;
  if ((rc = ensure(c, &c->mesh_cand, &c->mesh_cand_cap, kMeshCandBytes / sizeof(float), "hipMalloc mesh scratch")))
    return rc;
  if ((rc = ensure(c, &c->mesh_offsets, &c->mesh_offsets_cap, entries, "hipMalloc mesh offsets"))) return rc;
  if ((rc = ensure(c, &c->mesh_partials, &c->mesh_partials_cap, scanPartialsSize(static_cast<uint32_t>(entries)),
                   "hipMalloc mesh partials")))
    return rc;
  MeshArgs ma{};
  ma.voxel = c->P.voxel_size;
  ma.iso = iso;
  ma.extract_tol = c->P.voxel_size * c->P.layer_extract_scale;
  ma.extract_tol = ma.extract_tol < 1e-6f ? 1e-6f : ma.extract_tol;
  hipEvent_t ev{};
  beginTimed(c, 3, &ev);
  VHM_TRY(c, launchMeshCount(view(c), nullptr, nb, ma, c->mesh_counts, c->cube_counts, nullptr, c->mesh_cand, c->stream), "mesh count kernel launch");
  VHM_TRY(c, launchScan(c->mesh_counts, c->mesh_offsets, static_cast<uint32_t>(entries), c->mesh_partials, c->stream),
          "mesh scan kernel launch");
  uint32_t total = 0;
  const size_t tiles = scanPartialsSize(static_cast<uint32_t>(entries)) - 1;
  VHM_TRY(c, hipMemcpyAsync(&total, c->mesh_partials + tiles, sizeof(uint32_t), hipMemcpyDeviceToHost, c->stream),
          "hipMemcpy D2H triangle count");
  VHM_TRY(c, hipStreamSynchronize(c->stream), "hipStreamSynchronize");
  *n_out = total;
  if (tris_out && capacity > 0 && total > 0) {
    if (out_memkind == VHM_DEVICE && capacity >= total) {
      // the whole mesh fits: the emit pass writes it in place
      VHM_TRY(c, launchMeshEmit(view(c), nullptr, nb, ma, c->mesh_counts, c->cube_counts, c->mesh_offsets, tris_out,
                                nullptr, total, c->mesh_cand, c->stream),
              "mesh emit kernel launch");
    } else {
      if ((rc = ensure(c, &c->tris, &c->tris_cap, size_t(total) * 9u, "hipMalloc triangles"))) return rc;
      VHM_TRY(c, launchMeshEmit(view(c), nullptr, nb, ma, c->mesh_counts, c->cube_counts, c->mesh_offsets, c->tris, nullptr, total, c->mesh_cand, c->stream),
              "mesh emit kernel launch");
      const size_t m = capacity < total ? capacity : total;
      VHM_TRY(c, hipMemcpyAsync(tris_out, c->tris, m * 9u * sizeof(float),
                                out_memkind == VHM_DEVICE ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost,
                                c->stream),
              "hipMemcpy triangles");
    }
  }
  endTimed(c, ev);
  VHM_TRY(c, hipStreamSynchronize(c->stream), "hipStreamSynchronize");
  if (c->timing) drainEvents(c);
  return VHM_OK;
}

int vhm_extract_mesh_incremental(vhm_ctx* c, float iso, int32_t* block_keys_out, uint8_t* block_dirs_out,
                                 size_t block_capacity, size_t* n_blocks_out, float* tris_out, uint32_t* tri_item_out,
                                 size_t tri_capacity, int out_memkind, size_t* n_tris_out) {
  if (!c || !n_blocks_out || !n_tris_out) return VHM_ERR_INVALID_ARGUMENT;
  DevHeader h{};
  int rc = readHeader(c, &h);
  if (rc) return rc;
  *n_blocks_out = 0;
  *n_tris_out = 0;
  c->inc_tris = 0;
  const uint32_t nb = h.num_blocks < c->P.max_blocks ? h.num_blocks : c->P.max_blocks;
  if (nb == 0) return VHM_OK;
  hipEvent_t ev{};
  beginTimed(c, 3, &ev);
  // dirty (block, direction) masks, spread to the -x/-y/-z neighbours ->
  // ordered list of blocks with their direction masks
  VHM_TRY(c, launchMarkRemesh(view(c), nb, c->remesh_flags, c->stream), "remesh mark kernel launch");
  VHM_TRY(c, launchScan(c->remesh_flags, c->remesh_pos, nb, c->remesh_partials, c->stream, true),
          "remesh scan launch");
  VHM_TRY(c, launchCompactFlags(c->remesh_flags, c->remesh_pos, nb, c->remesh_list, c->remesh_dirs, c->stream),
          "remesh compact launch");
  uint32_t m = 0;
  VHM_TRY(c, hipMemcpyAsync(&m, c->remesh_partials + (scanPartialsSize(nb) - 1), sizeof(uint32_t),
                            hipMemcpyDeviceToHost, c->stream),
          "hipMemcpy D2H remesh count");
  VHM_TRY(c, hipStreamSynchronize(c->stream), "hipStreamSynchronize");
  *n_blocks_out = m;
  if (m == 0) {
    endTimed(c, ev);
    return VHM_OK;
  }
  const size_t entries = size_t(m) * kDirs;
  if ((rc = ensure(c, &c->mesh_cand, &c->mesh_cand_cap, kMeshCandBytes / sizeof(float), "hipMalloc mesh scratch")))
    return rc;
  MeshArgs ma{};
  ma.voxel = c->P.voxel_size;
  ma.iso = iso;
  ma.extract_tol = c->P.voxel_size * c->P.layer_extract_scale;
  ma.extract_tol = ma.extract_tol < 1e-6f ? 1e-6f : ma.extract_tol;
  if (block_keys_out && block_capacity > 0) {
    if ((rc = ensure(c, &c->key_buf, &c->key_buf_cap, 3 * size_t(m), "hipMalloc key buffer"))) return rc;
    VHM_TRY(c, launchListKeys(view(c), c->remesh_list, m, c->key_buf, c->stream), "remesh key launch");
    const size_t kb = block_capacity < m ? block_capacity : m;
    VHM_TRY(c, hipMemcpyAsync(block_keys_out, c->key_buf, 3 * kb * sizeof(int32_t), hipMemcpyDeviceToHost, c->stream),
            "hipMemcpy D2H remesh keys");
  }
  if (block_dirs_out && block_capacity > 0) {
    const size_t kb = block_capacity < m ? block_capacity : m;
    VHM_TRY(c, hipMemcpyAsync(block_dirs_out, c->remesh_dirs, kb, hipMemcpyDeviceToHost, c->stream),
            "hipMemcpy D2H remesh directions");
  }
  if (tris_out && tri_capacity > 0 && out_memkind == VHM_DEVICE) {
    // straight into the caller's device buffers in one fused pass (items in
    // completion order; triangles past the capacity are dropped)
    if ((rc = ensure(c, &c->mesh_counters, &c->mesh_counters_cap, 1, "hipMalloc mesh counters"))) return rc;
    if ((rc = ensure(c, &c->mesh_nbr, &c->mesh_nbr_cap, 8 * size_t(m), "hipMalloc mesh sources"))) return rc;
    if ((rc = ensure(c, &c->mesh_obs_dirs, &c->mesh_obs_dirs_cap, size_t(m), "hipMalloc mesh source directions")))
      return rc;
    VHM_TRY(c, hipMemsetAsync(c->mesh_counters, 0, sizeof(uint32_t), c->stream), "hipMemset mesh counters");
    const uint32_t cap = static_cast<uint32_t>(tri_capacity < 0xFFFFFFFFu ? tri_capacity : 0xFFFFFFFFu);
    VHM_TRY(c, launchMeshFused(view(c), c->remesh_list, m, ma, c->remesh_dirs, c->mesh_nbr, c->mesh_obs_dirs,
                               c->mesh_counters, tris_out, tri_item_out, cap, c->mesh_cand, c->stream),
            "fused mesh kernel launch");
    uint32_t total = 0;
    VHM_TRY(c, hipMemcpyAsync(&total, c->mesh_counters, sizeof(uint32_t), hipMemcpyDeviceToHost, c->stream),
            "hipMemcpy D2H triangle count");
    endTimed(c, ev);
    VHM_TRY(c, hipStreamSynchronize(c->stream), "hipStreamSynchronize");
    if (c->timing) drainEvents(c);
    *n_tris_out = total;
    return VHM_OK;
  }
  // two passes: count, scan, emit (item order)
  if ((rc = ensure(c, &c->mesh_counts, &c->mesh_counts_cap, entries, "hipMalloc mesh counts"))) return rc;
  if ((rc = ensure(c, &c->cube_counts, &c->cube_counts_cap, entries * 512, "hipMalloc cube counts"))) return rc;
  if ((rc = ensure(c, &c->mesh_offsets, &c->mesh_offsets_cap, entries, "hipMalloc mesh offsets"))) return rc;
  if ((rc = ensure(c, &c->mesh_partials, &c->mesh_partials_cap, scanPartialsSize(static_cast<uint32_t>(entries)),
                   "hipMalloc mesh partials")))
    return rc;
  VHM_TRY(c, launchMeshCount(view(c), c->remesh_list, m, ma, c->mesh_counts, c->cube_counts, c->remesh_dirs, c->mesh_cand, c->stream), "mesh count kernel launch");
  VHM_TRY(c, launchScan(c->mesh_counts, c->mesh_offsets, static_cast<uint32_t>(entries), c->mesh_partials, c->stream),
          "mesh scan kernel launch");
  uint32_t total = 0;
  VHM_TRY(c, hipMemcpyAsync(&total, c->mesh_partials + (scanPartialsSize(static_cast<uint32_t>(entries)) - 1),
                            sizeof(uint32_t), hipMemcpyDeviceToHost, c->stream),
          "hipMemcpy D2H triangle count");
  VHM_TRY(c, hipStreamSynchronize(c->stream), "hipStreamSynchronize");
  *n_tris_out = total;
  if (total > 0 && out_memkind != VHM_DEVICE) {
    // host output: emitted into the context's buffers (kept for
    // vhm_incremental_mesh_copy), then min(capacity, total) copied
    if ((rc = ensure(c, &c->tris, &c->tris_cap, size_t(total) * 9u, "hipMalloc triangles"))) return rc;
    if ((rc = ensure(c, &c->tri_owner, &c->tri_owner_cap, size_t(total), "hipMalloc triangle owners"))) return rc;
    VHM_TRY(c, launchMeshEmit(view(c), c->remesh_list, m, ma, c->mesh_counts, c->cube_counts, c->mesh_offsets, c->tris, c->tri_owner, total,
                              c->mesh_cand, c->stream),
            "mesh emit kernel launch");
    c->inc_tris = total;
    const size_t t = tri_capacity < total ? tri_capacity : total;
    if (tris_out && t)
      VHM_TRY(c, hipMemcpyAsync(tris_out, c->tris, t * 9u * sizeof(float), hipMemcpyDeviceToHost, c->stream),
              "hipMemcpy D2H triangles");
    if (tri_item_out && t)
      VHM_TRY(c, hipMemcpyAsync(tri_item_out, c->tri_owner, t * sizeof(uint32_t), hipMemcpyDeviceToHost, c->stream),
              "hipMemcpy D2H triangle owners");
  }
  endTimed(c, ev);
  VHM_TRY(c, hipStreamSynchronize(c->stream), "hipStreamSynchronize");
  if (c->timing) drainEvents(c);
  return VHM_OK;
}

int vhm_extract_mesh_incremental_noalloc(vhm_ctx* c, float iso, int32_t* block_keys_out, uint8_t* block_dirs_out,
                                         size_t block_capacity, size_t* n_blocks_out, float* tris_out,
                                         uint32_t* tri_item_out, size_t tri_capacity, size_t* n_tris_out) {
  if (!c || !n_blocks_out || !n_tris_out) return VHM_ERR_INVALID_ARGUMENT;
  *n_blocks_out = 0;
  *n_tris_out = 0;
  if (tri_capacity > 0 && !tris_out)
    return fail(c, VHM_ERR_INVALID_ARGUMENT, "incremental mesh extraction failed: null triangle buffer");
  hipError_t se = hipSuccess;
  uint8_t* const base = noAllocScratch(c->device, &se);
  if (!base) return hipFail(c, se, "hipGetSymbolAddress scratch");
  DevHeader h{};
  int rc = readHeader(c, &h);
  if (rc) return rc;
  const uint32_t nb = h.num_blocks < c->P.max_blocks ? h.num_blocks : c->P.max_blocks;
  if (nb == 0) return VHM_OK;
  const VolumeView V = view(c);
  VHM_TRY(c, launchMarkRemesh(V, nb, c->remesh_flags, c->stream), "remesh mark kernel launch");
  VHM_TRY(c, launchScan(c->remesh_flags, c->remesh_pos, nb, c->remesh_partials, c->stream, true),
          "remesh scan launch");
  uint32_t m = 0;
  VHM_TRY(c, hipMemcpyAsync(&m, c->remesh_partials + (scanPartialsSize(nb) - 1), sizeof(uint32_t),
                            hipMemcpyDeviceToHost, c->stream),
          "hipMemcpy D2H remesh count");
  VHM_TRY(c, hipStreamSynchronize(c->stream), "hipStreamSynchronize");
  if (m > kNoAllocMaxListed) {
    // not meshed: the marked blocks go back into the dirty masks
    VHM_TRY(c, launchRestoreDirty(V, nb, c->remesh_flags, c->stream), "remesh restore kernel launch");
    VHM_TRY(c, hipStreamSynchronize(c->stream), "hipStreamSynchronize");
    return fail(c, VHM_ERR_CAPACITY,
                "incremental mesh extraction failed: " + std::to_string(m) +
                    " blocks to re-mesh exceed the no-allocation limit of " + std::to_string(kNoAllocMaxListed) +
                    "; the blocks stay dirty (use vhm_extract_mesh_incremental)");
  }
  VHM_TRY(c, launchCompactFlags(c->remesh_flags, c->remesh_pos, nb, c->remesh_list, c->remesh_dirs, c->stream),
          "remesh compact launch");
  *n_blocks_out = m;
  if (m == 0) {
    VHM_TRY(c, hipStreamSynchronize(c->stream), "hipStreamSynchronize");
    return VHM_OK;
  }
  uint32_t* const counter = reinterpret_cast<uint32_t*>(base + kNoAllocCounterOff);
  uint32_t* const nbr = reinterpret_cast<uint32_t*>(base + kNoAllocNbrOff);
  uint8_t* const obs_dirs = base + kNoAllocObsOff;
  int32_t* const keys = reinterpret_cast<int32_t*>(base + kNoAllocKeysOff);
  float* const cand = reinterpret_cast<float*>(base + kNoAllocCandOff);
  const size_t kb = block_capacity < m ? block_capacity : m;
  if (block_keys_out && kb) {
    VHM_TRY(c, launchListKeys(V, c->remesh_list, m, keys, c->stream), "remesh key launch");
    VHM_TRY(c, hipMemcpyAsync(block_keys_out, keys, 3 * kb * sizeof(int32_t), hipMemcpyDeviceToHost, c->stream),
            "hipMemcpy D2H remesh keys");
  }
  if (block_dirs_out && kb)
    VHM_TRY(c, hipMemcpyAsync(block_dirs_out, c->remesh_dirs, kb, hipMemcpyDeviceToHost, c->stream),
            "hipMemcpy D2H remesh directions");
  MeshArgs ma{};
  ma.voxel = c->P.voxel_size;
  ma.iso = iso;
  ma.extract_tol = c->P.voxel_size * c->P.layer_extract_scale;
  ma.extract_tol = ma.extract_tol < 1e-6f ? 1e-6f : ma.extract_tol;
  VHM_TRY(c, launchZeroWords(counter, 1, c->stream), "mesh counter clear launch");
  const uint32_t cap = static_cast<uint32_t>(tri_capacity < 0xFFFFFFFFu ? tri_capacity : 0xFFFFFFFFu);
  VHM_TRY(c, launchMeshFused(V, c->remesh_list, m, ma, c->remesh_dirs, nbr, obs_dirs, counter, tris_out, tri_item_out,
                             cap, cand, c->stream, kNoAllocGrid),
          "fused mesh kernel launch");
  uint32_t total = 0;
  VHM_TRY(c, hipMemcpyAsync(&total, counter, sizeof(uint32_t), hipMemcpyDeviceToHost, c->stream),
          "hipMemcpy D2H triangle count");
  VHM_TRY(c, hipStreamSynchronize(c->stream), "hipStreamSynchronize");
  *n_tris_out = total;
  return VHM_OK;
}

int vhm_set_timing(vhm_ctx* c, int enabled) {
  if (!c) return VHM_ERR_INVALID_ARGUMENT;
  joinSide(c);
  hipStreamSynchronize(c->stream);
  drainEvents(c);
  c->timing = enabled != 0;
  c->timed_stages = (enabled & VHM_TIMING_STAGES) ? (static_cast<uint32_t>(enabled) & 0xFu) : 0xFu;
  for (int i = 0; i < 5; ++i) {
    c->ms[i] = 0.0;
    c->launches[i] = 0;
  }
  return VHM_OK;
}

int vhm_kernel_times(vhm_ctx* c, double ms_out[4], uint64_t launches_out[4]) {
  if (!c) return VHM_ERR_INVALID_ARGUMENT;
  joinSide(c);
  VHM_TRY(c, hipStreamSynchronize(c->stream), "hipStreamSynchronize");
  drainEvents(c);
  for (int i = 0; i < 4; ++i) {
    // the integrate stage: main and deferred passes (GPU time of both), one
    // launch per frame
    if (ms_out) ms_out[i] = c->ms[i] + (i == 2 ? c->ms[4] : 0.0);
    if (launches_out) launches_out[i] = c->launches[i];
  }
  return VHM_OK;
}

int vhm_synth_render(const vhm_scene* scene, const vhm_intrinsics* K, const vhm_pose* poses, int n_frames,
                     uint32_t first_frame, const vhm_noise* noise, uint16_t* out, int out_kind) {
  if (!scene || !K || !poses || !out || n_frames < 0 || K->width <= 0 || K->height <= 0)
    return VHM_ERR_INVALID_ARGUMENT;
  if (scene->num_primitives < 0 || scene->num_primitives > 256 || (scene->num_primitives > 0 && !scene->primitives))
    return VHM_ERR_INVALID_ARGUMENT;
  if (n_frames == 0) return VHM_OK;
  std::vector<float> prims(8 * size_t(scene->num_primitives));
  for (int i = 0; i < scene->num_primitives; ++i) {
    const vhm_primitive& p = scene->primitives[i];
    const float v[8] = {static_cast<float>(p.kind), p.c[0], p.c[1], p.c[2], p.r, p.h[0], p.h[1], p.h[2]};
    std::memcpy(&prims[8 * size_t(i)], v, sizeof(v));
  }
  SynthDeviceScene sc{};
  for (int i = 0; i < 3; ++i) {
    sc.room_min[i] = scene->room_min[i];
    sc.room_max[i] = scene->room_max[i];
  }
  sc.num_primitives = scene->num_primitives;
  SynthNoise nz{};
  if (noise) {
    nz.a = noise->a;
    nz.b = noise->b;
    nz.c = noise->c;
    nz.dropout = noise->dropout;
    nz.max_depth = noise->max_depth;
    nz.key0 = static_cast<uint32_t>(noise->seed);
    nz.key1 = static_cast<uint32_t>(noise->seed >> 32);
    nz.enabled = noise->enabled;
  }
  std::vector<Pose34> P(static_cast<size_t>(n_frames));
  for (int i = 0; i < n_frames; ++i) P[i] = toPose(&poses[i]);
  const Intrinsics k = toK(K);
  if (out_kind != VHM_DEVICE) {
    sc.prims = prims.data();
    synthRenderHost(sc, k, P.data(), n_frames, first_frame, nz, out);
    return VHM_OK;
  }
  float* dprims = nullptr;
  Pose34* dposes = nullptr;
  hipError_t e = hipMalloc(&dprims, sizeof(float) * (prims.size() + 8));
  if (e == hipSuccess) e = hipMalloc(&dposes, sizeof(Pose34) * P.size());
  if (e == hipSuccess) e = hipMemcpy(dprims, prims.data(), sizeof(float) * prims.size(), hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(dposes, P.data(), sizeof(Pose34) * P.size(), hipMemcpyHostToDevice);
  sc.prims = dprims;
  if (e == hipSuccess) e = launchSynthRender(sc, k, dposes, n_frames, first_frame, nz, out, nullptr);
  if (e == hipSuccess) e = hipDeviceSynchronize();
  if (dprims) hipFree(dprims);
  if (dposes) hipFree(dposes);
  return e == hipSuccess ? VHM_OK : VHM_ERR_HIP;
}

void vhm_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]) {
  vhm::philox4x32_10(ctr, key, out);
}

}  // extern "C"

void vhm::focalDivision(float f, float* inv, int* exact) {
  static std::mutex mu;
  static std::unordered_map<uint32_t, int> known;
  *inv = 1.0f / f;
  const uint32_t key = __builtin_bit_cast(uint32_t, f);
  std::lock_guard<std::mutex> lock(mu);
  auto it = known.find(key);
  if (it == known.end()) it = known.emplace(key, reciprocalDivisionExact(f) ? 1 : 0).first;
  *exact = it->second;
}

// Context internals for the other host translation units (vhm_internal.h).
vhm::VolumeView vhmCtxView(vhm_ctx* c) { return view(c); }
hipStream_t vhmCtxStream(vhm_ctx* c) {
  joinSide(c);  // work ordered after this stream's is ordered after the deferred pass too
  return c->stream;
}
const vhm_params* vhmCtxParams(vhm_ctx* c) { return &c->P; }
void vhmCtxAdvance(vhm_ctx* c, uint32_t last_frame_id, uint64_t frames) {
  if (last_frame_id > c->frame) c->frame = last_frame_id;
  c->frames_integrated += frames;
}
