// Batched multi-scene fusion: many independent volumes advance one frame per
// step with one launch per pipeline stage for all scenes together, the
// whole step captured once as a hipGraph and replayed (no reference
// counterpart: the reference fuses one volume synchronously).
//
// A step (the captured graph):
//   batchFrameBegin     reset every scene's active list, frame id = step + 1
//   filterBatch         all scenes' filter tiles in one grid
//   allocateBatch       all scenes' allocation tiles in one grid
//   numberBlocksBatch   each scene's new blocks renumbered in key order
//   activePrefix        prefix sums of the scenes' active-block counts
//   integrateBatch      persistent grid over all scenes' active blocks
//   incremental mesh    all scenes' dirty blocks (+ -x/-y/-z neighbours):
//                       mark, per-scene compaction, count, per-scene scan,
//                       emit into per-scene triangle buffers
//   stepIncrement       step += 1 (device counter read by every kernel, so
//                       the graph is replayed unchanged)
// Per-scene inputs (frames, poses and their inverses) are resident on the
// device for the whole sequence.  Grid sizes of the two persistent kernels
// come from the autotuner (vhm_batch_autotune) or the caller; the filter and
// allocation tile shapes and the integrator's chunk from
// vhm_batch_autotune_tiles.
#include <cstring>
#include <string>
#include <vector>

#include "../../include/vhm.h"
#include "vhm_internal.h"

#pragma clang fp contract(off)

using namespace vhm;

struct vhm_batch {
  int S = 0;
  std::vector<vhm_ctx*> vols;
  std::vector<BatchScene> host_scenes;
  std::vector<int> ready;
  std::vector<void*> owned;  // device buffers to free
  hipStream_t stream = nullptr;
  BatchScene* scenes = nullptr;
  BatchState* state = nullptr;
  uint32_t* filter_prefix = nullptr;
  uint32_t* alloc_prefix = nullptr;
  uint32_t* active_prefix = nullptr;
  uint32_t* mesh_prefix = nullptr;
  float* mesh_cand = nullptr;  // marching cubes' cluster-centre scratch (kMeshCandBytes)
  uint32_t filter_total = 0, alloc_total = 0;
  int integrate_grid = 2048, mesh_grid = 2048, num_cus = 256;
  int filter_tile_w = 32, alloc_tile_w = 16;  // tiles of w x (256 / w) pixels
  uint32_t integrate_chunk = 0;              // 0: one contiguous range per workgroup
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  bool prepared = false;
  uint32_t steps_done = 0;
  std::string err;
};

namespace {

int bfail(vhm_batch* b, int code, const std::string& m) {
  if (b) b->err = m;
  return code;
}

#define B_TRY(b, expr, what)                                                                           \
  do {                                                                                                 \
    const hipError_t _e = (expr);                                                                      \
    if (_e != hipSuccess) return bfail(b, VHM_ERR_HIP, std::string(what) + " failed: " + hipGetErrorString(_e)); \
  } while (0)

template <typename T>
hipError_t bmalloc(vhm_batch* b, T** p, size_t bytes) {
  const hipError_t e = hipMalloc(reinterpret_cast<void**>(p), bytes);
  if (e == hipSuccess) b->owned.push_back(*p);
  return e;
}

float invTwoSigmaSq(float s) { return 1.0f / (2.0f * s * s); }

// Host bookkeeping of one enqueued step: each scene still inside its
// sequence integrated frame id steps_done + 1.
void noteStep(vhm_batch* b) {
  for (int s = 0; s < b->S; ++s)
    if (b->steps_done < b->host_scenes[s].n_frames) vhmCtxAdvance(b->vols[s], b->steps_done + 1u, 1u);
  ++b->steps_done;
}

// A step's stages before integration: frame begin, filter, allocation, the
// ordered numbering of new blocks, the active-list prefix.  ev (optional):
// events recorded after the filter and after the allocation.
hipError_t enqueueFrontEnd(vhm_batch* b, hipEvent_t ev_filter = nullptr, hipEvent_t ev_alloc = nullptr) {
  const int S = b->S;
  hipError_t e = launchBatchFrameBegin(b->scenes, S, b->state, b->stream);
  if (e == hipSuccess)
    e = launchFilterBatch(b->scenes, S, b->filter_prefix, b->filter_total, b->state, b->filter_tile_w, b->stream);
  if (e == hipSuccess && ev_filter) e = hipEventRecord(ev_filter, b->stream);
  if (e == hipSuccess)
    e = launchAllocateBatch(b->scenes, S, b->alloc_prefix, b->alloc_total, b->state, b->alloc_tile_w, b->stream);
  if (e == hipSuccess && ev_alloc) e = hipEventRecord(ev_alloc, b->stream);
  if (e == hipSuccess) e = launchNumberBlocksBatch(b->scenes, S, b->state, b->stream);
  if (e == hipSuccess) e = launchActivePrefix(b->scenes, S, b->state, b->active_prefix, b->stream);
  return e;
}

bool validTileWidth(int w) { return w == 16 || w == 32 || w == 64; }

// Tile prefixes of the current tile shapes, uploaded.
int uploadTilePrefixes(vhm_batch* b) {
  std::vector<uint32_t> fp(b->S + 1), ap(b->S + 1);
  uint32_t f = 0, a = 0;
  for (int s = 0; s < b->S; ++s) {
    fp[s] = f;
    ap[s] = a;
    f += filterTiles(b->host_scenes[s].fa.K, b->filter_tile_w);
    a += allocTiles(b->host_scenes[s].aa.K, 1, b->alloc_tile_w);
  }
  fp[b->S] = f;
  ap[b->S] = a;
  b->filter_total = f;
  b->alloc_total = a;
  B_TRY(b, hipMemcpy(b->filter_prefix, fp.data(), 4 * (b->S + 1), hipMemcpyHostToDevice), "hipMemcpy prefixes");
  B_TRY(b, hipMemcpy(b->alloc_prefix, ap.data(), 4 * (b->S + 1), hipMemcpyHostToDevice), "hipMemcpy prefixes");
  return VHM_OK;
}

// Launches one step's kernels on b->stream (also what the graph captures).
int enqueueStep(vhm_batch* b) {
  const int S = b->S;
  B_TRY(b, enqueueFrontEnd(b), "batch front-end launch");
  B_TRY(b, launchIntegrateBatch(b->scenes, S, b->active_prefix, b->state, b->integrate_grid, b->integrate_chunk,
                                b->stream),
        "batch integrate launch");
  B_TRY(b, launchMeshBatch(b->scenes, S, b->mesh_prefix, b->mesh_grid, b->mesh_cand, b->stream), "batch mesh launch");
  B_TRY(b, launchStepIncrement(b->state, b->stream), "batch step launch");
  return VHM_OK;
}

}  // namespace

extern "C" {

int vhm_batch_create(int num_scenes, const vhm_params* params, vhm_batch** out) {
  if (!out || num_scenes <= 0 || !params) return VHM_ERR_INVALID_ARGUMENT;
  *out = nullptr;
  auto* b = new vhm_batch();
  b->S = num_scenes;
  for (int s = 0; s < num_scenes; ++s) {
    vhm_ctx* c = nullptr;
    const int rc = vhm_create(&params[s], &c);
    if (rc != VHM_OK) {
      vhm_batch_destroy(b);
      return rc;
    }
    b->vols.push_back(c);
  }
  b->host_scenes.resize(num_scenes);
  b->ready.assign(num_scenes, 0);
  int dev = 0;
  (void)hipGetDevice(&dev);
  hipDeviceProp_t prop{};
  (void)hipGetDeviceProperties(&prop, dev);
  b->num_cus = prop.multiProcessorCount > 0 ? prop.multiProcessorCount : 256;
  b->integrate_grid = b->num_cus * 8;
  b->mesh_grid = b->num_cus * 4;
  hipError_t e = hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking);
  if (e == hipSuccess) e = bmalloc(b, &b->scenes, sizeof(BatchScene) * num_scenes);
  if (e == hipSuccess) e = bmalloc(b, &b->state, sizeof(BatchState));
  if (e == hipSuccess) e = bmalloc(b, &b->filter_prefix, 4 * (num_scenes + 1));
  if (e == hipSuccess) e = bmalloc(b, &b->alloc_prefix, 4 * (num_scenes + 1));
  if (e == hipSuccess) e = bmalloc(b, &b->active_prefix, 4 * (num_scenes + 1));
  if (e == hipSuccess) e = bmalloc(b, &b->mesh_prefix, 4 * (num_scenes + 1));
  if (e == hipSuccess) e = bmalloc(b, &b->mesh_cand, kMeshCandBytes);
  if (e == hipSuccess) e = hipMemset(b->state, 0, sizeof(BatchState));
  if (e != hipSuccess) {
    vhm_batch_destroy(b);
    return VHM_ERR_HIP;
  }
  *out = b;
  return VHM_OK;
}

int vhm_batch_destroy(vhm_batch* b) {
  if (!b) return VHM_OK;
  if (b->stream) (void)hipStreamSynchronize(b->stream);
  if (b->exec) (void)hipGraphExecDestroy(b->exec);
  if (b->graph) (void)hipGraphDestroy(b->graph);
  for (void* p : b->owned) (void)hipFree(p);
  for (vhm_ctx* c : b->vols) vhm_destroy(c);
  if (b->stream) (void)hipStreamDestroy(b->stream);
  delete b;
  return VHM_OK;
}

const char* vhm_batch_last_error(const vhm_batch* b) { return b ? b->err.c_str() : "null batch"; }

vhm_ctx* vhm_batch_volume(vhm_batch* b, int scene) {
  if (!b || scene < 0 || scene >= b->S) return nullptr;
  return b->vols[scene];
}

int vhm_batch_set_sequence(vhm_batch* b, int scene, const uint16_t* frames_device, int n_frames, float depth_scale,
                           const vhm_intrinsics* K, const vhm_pose* poses, size_t tri_capacity) {
  if (!b || scene < 0 || scene >= b->S || !frames_device || n_frames <= 0 || !K || !poses)
    return VHM_ERR_INVALID_ARGUMENT;
  if (b->prepared) return bfail(b, VHM_ERR_INVALID_ARGUMENT, "batch set_sequence failed: batch already prepared");
  vhm_ctx* c = b->vols[scene];
  const vhm_params& P = *vhmCtxParams(c);
  if (K->width <= 0 || K->height <= 0 || K->width > P.max_width || K->height > P.max_height)
    return bfail(b, VHM_ERR_INVALID_ARGUMENT, "batch set_sequence failed: frame size outside the reservation");
  BatchScene& sc = b->host_scenes[scene];
  std::memset(&sc, 0, sizeof(sc));
  sc.V = vhmCtxView(c);
  const Intrinsics k{K->width, K->height, K->fx, K->fy, K->cx, K->cy};
  // filter constants (as vhm_integrate_*: reconstruction_backend.cpp:26-32)
  sc.fa.K = k;
  sc.fa.scale = depth_scale;
  const float inv_s = invTwoSigmaSq(P.filter_sigma_spatial);
  for (int dv = -2; dv <= 2; ++dv)
    for (int du = -2; du <= 2; ++du)
      sc.fa.spatial_w[(dv + 2) * 5 + (du + 2)] = std::exp(-static_cast<float>(du * du + dv * dv) * inv_s);
  sc.fa.inv_r = invTwoSigmaSq(P.filter_sigma_range);
  sc.fa.inv_a = invTwoSigmaSq(P.filter_sigma_angular);
  sc.fa.disc_base = P.filter_disc_base;
  sc.fa.disc_rel = P.filter_disc_rel;
  sc.fa.min_cos = P.filter_min_cos;
  sc.aa.K = k;
  sc.aa.voxel = P.voxel_size;
  sc.aa.trunc = P.truncation_distance;
  sc.aa.stride = 1;
  sc.ia.K = k;
  sc.ia.voxel = P.voxel_size;
  sc.ia.trunc = P.truncation_distance;
  sc.ia.merge_tol = P.voxel_size * P.layer_merge_scale;
  sc.ia.merge_tol = sc.ia.merge_tol < 1e-6f ? 1e-6f : sc.ia.merge_tol;
  sc.ia.min_cos = P.layer_normal_cos;
  sc.ia.max_weight = P.max_weight;
  sc.ia.near_base = meshNearBase(P.voxel_size, P.layer_extract_scale);
  sc.ia.min_view = P.min_view_angle_weight;
  sc.ia.disc_base = P.disc_base;
  sc.ia.disc_rel = P.disc_rel;
  const float inv_n = invTwoSigmaSq(P.normal_sigma_spatial);
  for (int s = 0; s < 3; ++s) sc.ia.spatial_w[s] = std::exp(-static_cast<float>(s) * inv_n);
  sc.ia.inv_r = invTwoSigmaSq(P.normal_sigma_range);
  sc.ia.inv_a = invTwoSigmaSq(P.normal_sigma_angular);
  focalDivision(k.fx, &sc.ia.inv_fx, &sc.ia.exact_fx);
  focalDivision(k.fy, &sc.ia.inv_fy, &sc.ia.exact_fy);
  sc.ma.voxel = P.voxel_size;
  sc.ma.iso = 0.0f;
  sc.ma.extract_tol = P.voxel_size * P.layer_extract_scale;
  sc.ma.extract_tol = sc.ma.extract_tol < 1e-6f ? 1e-6f : sc.ma.extract_tol;
  sc.frames = frames_device;
  sc.frame_elems = size_t(K->width) * K->height;
  sc.n_frames = static_cast<uint32_t>(n_frames);
  // poses and inverses (Pose::inverse, math_types.h:192-197, on the host)
  std::vector<Pose34> T(n_frames), Ti(n_frames);
  for (int i = 0; i < n_frames; ++i) {
    for (int j = 0; j < 9; ++j) T[i].R[j] = poses[i].R[j];
    for (int j = 0; j < 3; ++j) T[i].t[j] = poses[i].t[j];
    Ti[i] = inversePose(T[i]);
  }
  Pose34 *dT = nullptr, *dTi = nullptr;
  B_TRY(b, bmalloc(b, &dT, sizeof(Pose34) * n_frames), "hipMalloc poses");
  B_TRY(b, bmalloc(b, &dTi, sizeof(Pose34) * n_frames), "hipMalloc poses");
  B_TRY(b, hipMemcpy(dT, T.data(), sizeof(Pose34) * n_frames, hipMemcpyHostToDevice), "hipMemcpy poses");
  B_TRY(b, hipMemcpy(dTi, Ti.data(), sizeof(Pose34) * n_frames, hipMemcpyHostToDevice), "hipMemcpy poses");
  sc.poses = dT;
  sc.poses_inv = dTi;
  const size_t mb = P.max_blocks;
  B_TRY(b, bmalloc(b, &sc.filtered, 4 * sc.frame_elems), "hipMalloc filtered depth");
  B_TRY(b, bmalloc(b, &sc.remesh_flags, 4 * mb), "hipMalloc remesh flags");
  B_TRY(b, hipMemset(sc.remesh_flags, 0, 4 * mb), "hipMemset remesh flags");
  B_TRY(b, bmalloc(b, &sc.remesh_list, 4 * mb), "hipMalloc remesh list");
  B_TRY(b, bmalloc(b, &sc.remesh_dirs, mb), "hipMalloc remesh directions");
  B_TRY(b, bmalloc(b, &sc.mesh_nbr, 4 * 8 * mb), "hipMalloc mesh sources");
  B_TRY(b, bmalloc(b, &sc.mesh_obs_dirs, mb), "hipMalloc mesh source directions");
  sc.tri_capacity = static_cast<uint32_t>(tri_capacity);
  B_TRY(b, bmalloc(b, &sc.tris, 36 * (tri_capacity ? tri_capacity : 1)), "hipMalloc triangles");
  B_TRY(b, bmalloc(b, &sc.tri_owner, 4 * (tri_capacity ? tri_capacity : 1)), "hipMalloc triangle items");
  B_TRY(b, bmalloc(b, &sc.out, 16), "hipMalloc mesh outputs");
  B_TRY(b, hipMemset(sc.out, 0, 16), "hipMemset mesh outputs");
  b->ready[scene] = 1;
  return VHM_OK;
}

// Uploads the scene table and tile prefixes; with use_graph, captures one
// step as a hipGraph.  integrate_grid / mesh_grid <= 0 keep the current
// (default or autotuned) values.
int vhm_batch_prepare(vhm_batch* b, int integrate_grid, int mesh_grid, int use_graph) {
  if (!b) return VHM_ERR_INVALID_ARGUMENT;
  for (int s = 0; s < b->S; ++s)
    if (!b->ready[s]) return bfail(b, VHM_ERR_INVALID_ARGUMENT, "batch prepare failed: a scene has no sequence");
  if (integrate_grid > 0) b->integrate_grid = integrate_grid;
  if (mesh_grid > 0) b->mesh_grid = mesh_grid;
  B_TRY(b, hipMemcpy(b->scenes, b->host_scenes.data(), sizeof(BatchScene) * b->S, hipMemcpyHostToDevice),
        "hipMemcpy scene table");
  {
    const int rc = uploadTilePrefixes(b);
    if (rc) return rc;
  }
  // the volumes' own streams must be idle before the batch stream uses them
  for (vhm_ctx* c : b->vols) B_TRY(b, hipStreamSynchronize(vhmCtxStream(c)), "hipStreamSynchronize volume");
  if (b->exec) {
    (void)hipGraphExecDestroy(b->exec);
    b->exec = nullptr;
  }
  if (b->graph) {
    (void)hipGraphDestroy(b->graph);
    b->graph = nullptr;
  }
  if (use_graph) {
    B_TRY(b, hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal), "hipStreamBeginCapture");
    const int rc = enqueueStep(b);
    hipGraph_t g = nullptr;
    const hipError_t e = hipStreamEndCapture(b->stream, &g);
    if (rc != VHM_OK) {
      if (g) (void)hipGraphDestroy(g);
      return rc;
    }
    B_TRY(b, e, "hipStreamEndCapture");
    b->graph = g;
    B_TRY(b, hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0), "hipGraphInstantiate");
  }
  b->prepared = true;
  return VHM_OK;
}

// Runs n steps (asynchronous): graph replays when captured, direct launches otherwise.
int vhm_batch_step(vhm_batch* b, int n_steps) {
  if (!b || !b->prepared || n_steps < 0) return VHM_ERR_INVALID_ARGUMENT;
  for (int i = 0; i < n_steps; ++i) {
    if (b->exec) {
      B_TRY(b, hipGraphLaunch(b->exec, b->stream), "hipGraphLaunch");
    } else {
      const int rc = enqueueStep(b);
      if (rc) return rc;
    }
    noteStep(b);
  }
  return VHM_OK;
}

int vhm_batch_synchronize(vhm_batch* b) {
  if (!b) return VHM_ERR_INVALID_ARGUMENT;
  B_TRY(b, hipStreamSynchronize(b->stream), "hipStreamSynchronize");
  for (int s = 0; s < b->S; ++s) {
    const int rc = vhm_synchronize(b->vols[s]);
    if (rc != VHM_OK) return bfail(b, rc, std::string("scene ") + std::to_string(s) + ": " + vhm_last_error(b->vols[s]));
  }
  return VHM_OK;
}

// Last step's incremental mesh of a scene: blocks re-meshed, triangles
// (the total, which may exceed the capacity) and optionally the triangles
// (min(total, capacity) x 9 floats, host).  Each step overwrites the
// previous step's output.
int vhm_batch_mesh_output(vhm_batch* b, int scene, uint32_t* blocks, uint32_t* triangles, float* tris_host,
                          size_t capacity) {
  if (!b || scene < 0 || scene >= b->S) return VHM_ERR_INVALID_ARGUMENT;
  B_TRY(b, hipStreamSynchronize(b->stream), "hipStreamSynchronize");
  uint32_t o[4] = {0, 0, 0, 0};
  const BatchScene& sc = b->host_scenes[scene];
  B_TRY(b, hipMemcpy(o, sc.out, 16, hipMemcpyDeviceToHost), "hipMemcpy mesh outputs");
  if (blocks) *blocks = o[0];
  if (triangles) *triangles = o[1];
  if (tris_host && capacity) {
    size_t n = o[1] < sc.tri_capacity ? o[1] : sc.tri_capacity;
    n = n < capacity ? n : capacity;
    if (n) B_TRY(b, hipMemcpy(tris_host, sc.tris, n * 36, hipMemcpyDeviceToHost), "hipMemcpy triangles");
  }
  return VHM_OK;
}

// The items of the last step's incremental mesh of a scene: the listed
// blocks (keys in pool order, direction masks) and each triangle's item.
int vhm_batch_mesh_items(vhm_batch* b, int scene, int32_t* block_keys_out, uint8_t* block_dirs_out,
                         size_t block_capacity, size_t* n_blocks_out, uint32_t* tri_item_out, size_t tri_capacity) {
  if (!b || scene < 0 || scene >= b->S || !n_blocks_out) return VHM_ERR_INVALID_ARGUMENT;
  B_TRY(b, hipStreamSynchronize(b->stream), "hipStreamSynchronize");
  uint32_t o[4] = {0, 0, 0, 0};
  const BatchScene& sc = b->host_scenes[scene];
  B_TRY(b, hipMemcpy(o, sc.out, 16, hipMemcpyDeviceToHost), "hipMemcpy mesh outputs");
  *n_blocks_out = o[0];
  const size_t kb = block_capacity < o[0] ? block_capacity : o[0];
  if (kb && block_keys_out) {
    int32_t* dkeys = nullptr;
    B_TRY(b, hipMalloc(&dkeys, 12 * kb), "hipMalloc key buffer");
    hipError_t e = launchListKeys(sc.V, sc.remesh_list, static_cast<uint32_t>(kb), dkeys, b->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(block_keys_out, dkeys, 12 * kb, hipMemcpyDeviceToHost, b->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(b->stream);
    (void)hipFree(dkeys);
    B_TRY(b, e, "batch mesh keys");
  }
  if (kb && block_dirs_out)
    B_TRY(b, hipMemcpy(block_dirs_out, sc.remesh_dirs, kb, hipMemcpyDeviceToHost), "hipMemcpy remesh directions");
  size_t n = o[1] < sc.tri_capacity ? o[1] : sc.tri_capacity;
  n = n < tri_capacity ? n : tri_capacity;
  if (n && tri_item_out)
    B_TRY(b, hipMemcpy(tri_item_out, sc.tri_owner, 4 * n, hipMemcpyDeviceToHost), "hipMemcpy triangle items");
  return VHM_OK;
}

}  // extern "C"

namespace {

// One step with direct launches and events: ms[0] frame begin + filter +
// allocation + prefix, ms[1] integration, ms[2] incremental meshing;
// *updates / *remeshed: the step's voxel updates and re-meshed blocks over
// all scenes; fa_ms (optional): the filter's and the allocation's own times.
// Synchronous.
int timedStep(vhm_batch* b, int gi, int gm, float ms[3], uint64_t* updates, uint64_t* remeshed,
              float* fa_ms = nullptr) {
  const int S = b->S;
  uint64_t u0 = 0;
  for (vhm_ctx* v : b->vols) {
    uint64_t u = 0;
    const int rc = vhm_stats(v, &u, nullptr, nullptr);
    if (rc) return bfail(b, rc, vhm_last_error(v));
    u0 += u;
  }
  hipEvent_t ev[6] = {};
  for (auto& e : ev) B_TRY(b, hipEventCreate(&e), "hipEventCreate");
  hipError_t e = hipEventRecord(ev[0], b->stream);
  if (e == hipSuccess) e = enqueueFrontEnd(b, ev[4], ev[5]);
  if (e == hipSuccess) e = hipEventRecord(ev[1], b->stream);
  if (e == hipSuccess)
    e = launchIntegrateBatch(b->scenes, S, b->active_prefix, b->state, gi, b->integrate_chunk, b->stream);
  if (e == hipSuccess) e = hipEventRecord(ev[2], b->stream);
  if (e == hipSuccess) e = launchMeshBatch(b->scenes, S, b->mesh_prefix, gm, b->mesh_cand, b->stream);
  if (e == hipSuccess) e = hipEventRecord(ev[3], b->stream);
  if (e == hipSuccess) e = launchStepIncrement(b->state, b->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(b->stream);
  if (e == hipSuccess) {
    for (int i = 0; i < 3; ++i) (void)hipEventElapsedTime(&ms[i], ev[i], ev[i + 1]);
    if (fa_ms) {
      (void)hipEventElapsedTime(&fa_ms[0], ev[0], ev[4]);  // frame begin + filter
      (void)hipEventElapsedTime(&fa_ms[1], ev[4], ev[5]);  // allocation
    }
  }
  for (auto& x : ev) (void)hipEventDestroy(x);
  B_TRY(b, e, "batch timed step");
  noteStep(b);
  uint64_t u1 = 0, rm = 0;
  for (int s = 0; s < S; ++s) {
    uint64_t u = 0;
    const int rc = vhm_stats(b->vols[s], &u, nullptr, nullptr);
    if (rc) return bfail(b, rc, vhm_last_error(b->vols[s]));
    u1 += u;
    uint32_t o[4];
    B_TRY(b, hipMemcpy(o, b->host_scenes[s].out, 16, hipMemcpyDeviceToHost), "hipMemcpy mesh outputs");
    rm += o[0];
  }
  if (updates) *updates = u1 - u0;
  if (remeshed) *remeshed = rm;
  return VHM_OK;
}

}  // namespace

extern "C" {

// One step with direct launches and per-stage timing (see timedStep), with
// the current grid sizes.
int vhm_batch_timed_step(vhm_batch* b, float ms_out[3], uint64_t* updates, uint64_t* remeshed) {
  if (!b || !b->prepared || !ms_out) return VHM_ERR_INVALID_ARGUMENT;
  return timedStep(b, b->integrate_grid, b->mesh_grid, ms_out, updates, remeshed);
}

// Times one step per candidate grid size and keeps the fastest per stage,
// normalised by work (updates for integration, re-meshed blocks for
// meshing).  Consumes n_candidates steps of the sequences (results do not
// depend on grid sizes).  candidates: multipliers of the CU count.
int vhm_batch_autotune(vhm_batch* b, const int* candidates, int n_candidates, int* integrate_grid_out,
                       int* mesh_grid_out) {
  if (!b || !b->prepared || !candidates || n_candidates <= 0) return VHM_ERR_INVALID_ARGUMENT;
  double best_i = 1e300, best_m = 1e300;
  int bi = b->integrate_grid, bm = b->mesh_grid;
  for (int c = 0; c < n_candidates; ++c) {
    const int g = candidates[c] * b->num_cus;
    if (g <= 0 || g > (1 << 20)) return bfail(b, VHM_ERR_INVALID_ARGUMENT, "batch autotune failed: bad candidate");
    float ms[3] = {0.0f, 0.0f, 0.0f};
    uint64_t upd = 0, rm = 0;
    const int rc = timedStep(b, g, g, ms, &upd, &rm);
    if (rc) return rc;
    const double per_i = ms[1] / static_cast<double>(upd + 1), per_m = ms[2] / static_cast<double>(rm + 1);
    if (per_i < best_i) {
      best_i = per_i;
      bi = g;
    }
    if (per_m < best_m) {
      best_m = per_m;
      bm = g;
    }
  }
  b->integrate_grid = bi;
  b->mesh_grid = bm;
  if (integrate_grid_out) *integrate_grid_out = bi;
  if (mesh_grid_out) *mesh_grid_out = bm;
  if (b->exec) return vhm_batch_prepare(b, 0, 0, 1);  // re-capture with the chosen grids
  return VHM_OK;
}

// Times one step per candidate and keeps the fastest filter tile width and
// allocation tile width (each timed by its own events; 16, 32 or 64 pixels
// wide, 256 / width tall) and integrator chunk (0: one contiguous range per
// workgroup, else entries per ticket; normalised by voxel updates).  Step k
// uses candidate k of each list (the last one once a list is exhausted), so
// max(n_filter, n_alloc, n_chunk) steps of the sequences are consumed
// (results do not depend on these choices).
int vhm_batch_autotune_tiles(vhm_batch* b, const int* filter_widths, int n_filter, const int* alloc_widths,
                             int n_alloc, const int* chunks, int n_chunk, int* filter_width_out, int* alloc_width_out,
                             int* chunk_out) {
  if (!b || !b->prepared || !filter_widths || !alloc_widths || !chunks || n_filter <= 0 || n_alloc <= 0 ||
      n_chunk <= 0)
    return VHM_ERR_INVALID_ARGUMENT;
  for (int i = 0; i < n_filter; ++i)
    if (!validTileWidth(filter_widths[i]))
      return bfail(b, VHM_ERR_INVALID_ARGUMENT, "batch tile autotune failed: filter tile width not 16, 32 or 64");
  for (int i = 0; i < n_alloc; ++i)
    if (!validTileWidth(alloc_widths[i]))
      return bfail(b, VHM_ERR_INVALID_ARGUMENT, "batch tile autotune failed: allocation tile width not 16, 32 or 64");
  for (int i = 0; i < n_chunk; ++i)
    if (chunks[i] < 0 || chunks[i] > (1 << 20))
      return bfail(b, VHM_ERR_INVALID_ARGUMENT, "batch tile autotune failed: bad integrator chunk");
  const int rounds = n_filter > n_alloc ? (n_filter > n_chunk ? n_filter : n_chunk) : (n_alloc > n_chunk ? n_alloc : n_chunk);
  float best_f = 3.4e38f, best_a = 3.4e38f;
  double best_c = 1e300;
  int bf = b->filter_tile_w, ba = b->alloc_tile_w, bc = static_cast<int>(b->integrate_chunk);
  for (int k = 0; k < rounds; ++k) {
    b->filter_tile_w = filter_widths[k < n_filter ? k : n_filter - 1];
    b->alloc_tile_w = alloc_widths[k < n_alloc ? k : n_alloc - 1];
    b->integrate_chunk = static_cast<uint32_t>(chunks[k < n_chunk ? k : n_chunk - 1]);
    int rc = uploadTilePrefixes(b);
    if (rc) return rc;
    float ms[3] = {0.0f, 0.0f, 0.0f}, fa[2] = {0.0f, 0.0f};
    uint64_t upd = 0, rm = 0;
    rc = timedStep(b, b->integrate_grid, b->mesh_grid, ms, &upd, &rm, fa);
    if (rc) return rc;
    if (k < n_filter && fa[0] < best_f) {
      best_f = fa[0];
      bf = b->filter_tile_w;
    }
    if (k < n_alloc && fa[1] < best_a) {
      best_a = fa[1];
      ba = b->alloc_tile_w;
    }
    const double per_c = ms[1] / static_cast<double>(upd + 1);
    if (k < n_chunk && per_c < best_c) {
      best_c = per_c;
      bc = static_cast<int>(b->integrate_chunk);
    }
  }
  b->filter_tile_w = bf;
  b->alloc_tile_w = ba;
  b->integrate_chunk = static_cast<uint32_t>(bc);
  if (filter_width_out) *filter_width_out = bf;
  if (alloc_width_out) *alloc_width_out = ba;
  if (chunk_out) *chunk_out = bc;
  return vhm_batch_prepare(b, 0, 0, b->exec ? 1 : 0);  // prefixes of the chosen shapes, graph re-captured
}

// Clears every volume and restarts the sequences at frame 0, keeping the
// reservations, the captured graph and the tuned grid sizes.
int vhm_batch_rewind(vhm_batch* b) {
  if (!b) return VHM_ERR_INVALID_ARGUMENT;
  B_TRY(b, hipStreamSynchronize(b->stream), "hipStreamSynchronize");
  for (int s = 0; s < b->S; ++s) {
    const int rc = vhm_reset(b->vols[s]);
    if (rc != VHM_OK) return bfail(b, rc, std::string("scene ") + std::to_string(s) + ": " + vhm_last_error(b->vols[s]));
    const BatchScene& sc = b->host_scenes[s];
    if (!b->ready[s]) continue;
    B_TRY(b, hipMemsetAsync(sc.remesh_flags, 0, 4 * size_t(vhmCtxParams(b->vols[s])->max_blocks), b->stream),
          "hipMemset remesh flags");
    B_TRY(b, hipMemsetAsync(sc.out, 0, 16, b->stream), "hipMemset mesh outputs");
  }
  B_TRY(b, hipMemsetAsync(b->state, 0, sizeof(BatchState), b->stream), "hipMemset batch state");
  B_TRY(b, hipStreamSynchronize(b->stream), "hipStreamSynchronize");
  b->steps_done = 0;
  return VHM_OK;
}

int vhm_batch_steps_done(const vhm_batch* b) { return b ? static_cast<int>(b->steps_done) : 0; }

}  // extern "C"
