// C++ host-side mirror of the reference's GPU backend interface, over the
// libvhm C-ABI (include/vhm.h).  Header-only; templated on the caller's
// volume / frame / pose types, so it compiles inside the reference tree with
// tsdfmc::VoxelHashTSDF, DepthFrame and Pose (voxel_hash_tsdf.h,
// math_types.h) and with any type of the same shape.
//
// Reference interfaces mirrored (same names with Gpu for Cuda, same argument
// meaning, same error behaviour -- std::runtime_error with
// "<action> failed: <detail>"):
//   queryCudaBackendStatus          cuda_tsdf_integrator.h:14   -> queryGpuBackendStatus
//   integrateDepthFrameCuda         cuda_tsdf_integrator.h:16-19 -> integrateDepthFrameGpu
//   hasGpuVolumeCache, gpuVolumeCacheBlockCount,
//   gpuVolumeCacheObservedVoxelCount, gpuVolumeCacheBlockKeys
//                                   cuda_tsdf_integrator.h:21-24 -> same names
//   allocateActiveBlocksCuda        cuda_block_allocator.h:17-22 -> allocateActiveBlocksGpu
//   CudaActiveBlockAllocationFallback cuda_block_allocator.h:11-15 -> GpuActiveBlockAllocationFallback
//   extractMeshCuda                 cuda_mesh_extractor.h:9     -> extractMeshGpu
//
// As in the reference, GPU state is keyed by the volume's address
// (cuda_tsdf_integrator.cu:114-130), but every volume keeps its own context
// instead of one process-global cache.  Not thread-safe (neither is the
// reference).  Capacity: VHM_MAX_BLOCKS / VHM_HASH_LOG2 environment
// variables (default 262144 blocks, 2^22 buckets) or setGpuVolumeCapacity().

#pragma once

#include <cstdlib>
#include <memory>
#include <stdexcept>
#include <string>
#include <unordered_map>
#include <vector>

#include "../../include/vhm.h"

namespace vhm_backend {

struct GpuBackendStatus {
  bool runtime_available = false;
  std::string message;
};

class GpuActiveBlockAllocationFallback : public std::runtime_error {
 public:
  explicit GpuActiveBlockAllocationFallback(const std::string& m) : std::runtime_error(m) {}
};

struct BlockKeyXYZ {
  int x = 0, y = 0, z = 0;
};

inline GpuBackendStatus queryGpuBackendStatus() {
  char msg[256] = {0};
  const int ok = vhm_query_device(msg, sizeof(msg));
  return {ok != 0, msg};
}

namespace detail {

struct Capacity {
  uint32_t max_blocks = 262144;
  uint32_t hash_log2 = 22;
};

struct Entry {
  vhm_ctx* ctx = nullptr;
  int width = 0, height = 0;
  ~Entry() {
    if (ctx) vhm_destroy(ctx);
  }
};

inline std::unordered_map<const void*, std::unique_ptr<Entry>>& registry() {
  static std::unordered_map<const void*, std::unique_ptr<Entry>> r;
  return r;
}

inline std::unordered_map<const void*, Capacity>& capacities() {
  static std::unordered_map<const void*, Capacity> c;
  return c;
}

inline void check(vhm_ctx* ctx, int rc) {
  if (rc == VHM_OK) return;
  const std::string m = ctx ? vhm_last_error(ctx) : std::string("vhm call failed");
  if (rc == VHM_ERR_CAPACITY) throw GpuActiveBlockAllocationFallback(m);
  throw std::runtime_error(m.empty() ? "vhm call failed: status " + std::to_string(rc) : m);
}

template <class Volume>
Entry& entryFor(const Volume& volume, int width, int height) {
  auto& r = registry();
  auto it = r.find(&volume);
  if (it != r.end() && it->second->ctx && width <= it->second->width && height <= it->second->height)
    return *it->second;
  Capacity cap;
  if (const char* e = std::getenv("VHM_MAX_BLOCKS")) cap.max_blocks = static_cast<uint32_t>(std::strtoul(e, nullptr, 10));
  if (const char* e = std::getenv("VHM_HASH_LOG2")) cap.hash_log2 = static_cast<uint32_t>(std::strtoul(e, nullptr, 10));
  const auto c = capacities().find(&volume);
  if (c != capacities().end()) cap = c->second;
  if (it != r.end() && it->second->ctx)
    throw std::runtime_error("GPU integrate failed: frame larger than the first frame of this volume");
  vhm_params p;
  vhm_default_params(&p, VHM_PRESET_REFERENCE_CPU, volume.voxelSize(), volume.truncationDistance());
  p.max_blocks = cap.max_blocks;
  p.hash_log2_buckets = cap.hash_log2;
  p.max_width = width;
  p.max_height = height;
  auto e = std::make_unique<Entry>();
  const int rc = vhm_create(&p, &e->ctx);
  if (rc != VHM_OK) throw std::runtime_error("vhm_create failed: status " + std::to_string(rc));
  e->width = width;
  e->height = height;
  Entry& ref = *e;
  r[&volume] = std::move(e);
  return ref;
}

template <class Frame>
vhm_intrinsics intrinsicsOf(const Frame& f) {
  return {f.intrinsics.width, f.intrinsics.height, f.intrinsics.fx, f.intrinsics.fy, f.intrinsics.cx,
          f.intrinsics.cy};
}

template <class PoseT>
vhm_pose poseOf(const PoseT& T) {
  vhm_pose p;
  for (int i = 0; i < 9; ++i) p.R[i] = T.R.data[static_cast<std::size_t>(i)];
  p.t[0] = T.t.x;
  p.t[1] = T.t.y;
  p.t[2] = T.t.z;
  return p;
}

}  // namespace detail

template <class Volume>
void setGpuVolumeCapacity(const Volume& volume, uint32_t max_blocks, uint32_t hash_log2) {
  detail::capacities()[&volume] = {max_blocks, hash_log2};
}

// integrateDepthFrameCuda (cuda_tsdf_integrator.cu:946-1010): filter,
// allocate and integrate one host frame; synchronous like the reference.
template <class Volume, class Frame, class PoseT>
void integrateDepthFrameGpu(Volume& volume, const Frame& frame, const PoseT& T_wc, int allocation_stride) {
  if (frame.depth.empty()) return;
  detail::Entry& e = detail::entryFor(volume, frame.intrinsics.width, frame.intrinsics.height);
  const vhm_intrinsics K = detail::intrinsicsOf(frame);
  const vhm_pose P = detail::poseOf(T_wc);
  detail::check(e.ctx, vhm_integrate_f32(e.ctx, frame.depth.data(), VHM_HOST, &K, &P, allocation_stride, 0));
  detail::check(e.ctx, vhm_synchronize(e.ctx));
}

template <class Volume>
bool hasGpuVolumeCache(const Volume& volume) {
  const auto it = detail::registry().find(&volume);
  if (it == detail::registry().end() || !it->second->ctx) return false;
  size_t n = 0;
  detail::check(it->second->ctx, vhm_block_count(it->second->ctx, &n));
  return n > 0;
}

template <class Volume>
std::size_t gpuVolumeCacheBlockCount(const Volume& volume) {
  const auto it = detail::registry().find(&volume);
  if (it == detail::registry().end()) return volume.blockCount();
  size_t n = 0;
  detail::check(it->second->ctx, vhm_block_count(it->second->ctx, &n));
  return n;
}

template <class Volume>
std::size_t gpuVolumeCacheObservedVoxelCount(const Volume& volume) {
  const auto it = detail::registry().find(&volume);
  if (it == detail::registry().end()) return volume.observedVoxelCount();
  size_t n = 0;
  detail::check(it->second->ctx, vhm_observed_voxel_count(it->second->ctx, &n));
  return n;
}

template <class Key = BlockKeyXYZ, class Volume>
std::vector<Key> gpuVolumeCacheBlockKeys(const Volume& volume) {
  std::vector<Key> out;
  const auto it = detail::registry().find(&volume);
  if (it == detail::registry().end()) return out;
  size_t n = 0;
  detail::check(it->second->ctx, vhm_block_count(it->second->ctx, &n));
  std::vector<int32_t> k(3 * n);
  detail::check(it->second->ctx, vhm_block_keys(it->second->ctx, k.data(), n, &n));
  out.resize(n);
  for (size_t i = 0; i < n; ++i) out[i] = Key{k[3 * i], k[3 * i + 1], k[3 * i + 2]};
  return out;
}

// allocateActiveBlocksCuda (cuda_block_allocator.cu:304-394): the frame's
// active keys.  The keys are inserted into `volume`'s persistent table (the
// reference inserts them into its host map right after, in
// ensureGpuVolumeCacheBlocks).  device_depth, when given, is a device
// pointer to the already filtered frame.
template <class Key = BlockKeyXYZ, class Volume, class Frame, class PoseT>
std::vector<Key> allocateActiveBlocksGpu(Volume& volume, const Frame& frame, const PoseT& T_wc,
                                         int allocation_stride, const float* device_depth = nullptr) {
  std::vector<Key> out;
  if (frame.depth.empty()) return out;
  detail::Entry& e = detail::entryFor(volume, frame.intrinsics.width, frame.intrinsics.height);
  const vhm_intrinsics K = detail::intrinsicsOf(frame);
  const vhm_pose P = detail::poseOf(T_wc);
  const size_t cap = size_t(frame.intrinsics.width) * frame.intrinsics.height * 81u;
  std::vector<int32_t> k(3 * cap);
  size_t n = 0;
  detail::check(e.ctx, vhm_allocate_active_blocks(e.ctx, device_depth ? device_depth : frame.depth.data(),
                                                  device_depth ? VHM_DEVICE : VHM_HOST, &K, &P,
                                                  allocation_stride, k.data(), cap, &n));
  out.resize(n);
  for (size_t i = 0; i < n; ++i) out[i] = Key{k[3 * i], k[3 * i + 1], k[3 * i + 2]};
  return out;
}

// extractMeshCuda (cuda_mesh_extractor.cu:636-720) -> unindexed triangles.
template <class Triangle, class Volume>
std::vector<Triangle> extractMeshGpu(const Volume& volume, float iso_level = 0.0f) {
  std::vector<Triangle> tris;
  const auto it = detail::registry().find(&volume);
  if (it == detail::registry().end()) return tris;
  vhm_ctx* ctx = it->second->ctx;
  size_t n = 0;
  detail::check(ctx, vhm_extract_mesh(ctx, iso_level, nullptr, 0, &n));
  static_assert(sizeof(Triangle) == 9 * sizeof(float), "Triangle must be three Vec3f");
  tris.resize(n);
  if (n) detail::check(ctx, vhm_extract_mesh(ctx, iso_level, reinterpret_cast<float*>(tris.data()), n, &n));
  return tris;
}

// Releases the context of a volume (the reference frees its cache at exit).
template <class Volume>
void releaseGpuVolume(const Volume& volume) {
  detail::registry().erase(&volume);
}

}  // namespace vhm_backend
