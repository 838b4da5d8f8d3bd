// Drives the C++ host mirror (voxelhashmt_amd/host/tsdfmc_vhm_backend.hpp)
// with types shaped like the reference's (math_types.h, voxel_hash_tsdf.h,
// marching_cubes.h) the way the reference's frame loop does
// (main.cpp:433-455 -> TsdfIntegrationBackend::integrate / extractMesh /
// blockCount / observedVoxelCount).  Prints one JSON line of results for
// tests/test_gpu_shim.py to compare with the C-ABI path.
#include <array>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "../../voxelhashmt_amd/host/tsdfmc_vhm_backend.hpp"

namespace ref_shaped {
struct Vec3f { float x = 0, y = 0, z = 0; };
struct Mat3f { std::array<float, 9> data{}; };
struct Pose { Mat3f R; Vec3f t; };
struct Intrinsics { int width = 0, height = 0; float fx = 0, fy = 0, cx = 0, cy = 0; };
struct DepthFrame { Intrinsics intrinsics; std::vector<float> depth; };
struct Triangle { Vec3f v0, v1, v2; };
struct BlockKey { int x = 0, y = 0, z = 0; };
class VoxelHashTSDF {
 public:
  VoxelHashTSDF(float v, float t) : v_(v), t_(t) {}
  float voxelSize() const { return v_; }
  float truncationDistance() const { return t_; }
  std::size_t blockCount() const { return 0; }
  std::size_t observedVoxelCount() const { return 0; }
 private:
  float v_, t_;
};
}  // namespace ref_shaped

int main(int argc, char** argv) {
  using namespace ref_shaped;
  // argv: width height fx cx cy voxel trunc n_frames, then binary file of
  // uint16 frames and float poses (12 per frame) written by the test.
  if (argc < 10) return 2;
  const int W = std::atoi(argv[1]), H = std::atoi(argv[2]);
  const float fx = std::strtof(argv[3], nullptr), cx = std::strtof(argv[4], nullptr), cy = std::strtof(argv[5], nullptr);
  const float voxel = std::strtof(argv[6], nullptr), trunc = std::strtof(argv[7], nullptr);
  const int n = std::atoi(argv[8]);
  FILE* f = std::fopen(argv[9], "rb");
  if (!f) return 3;
  std::vector<uint16_t> raw(static_cast<size_t>(n) * W * H);
  std::vector<float> poses(static_cast<size_t>(n) * 12);
  if (std::fread(raw.data(), 2, raw.size(), f) != raw.size() || std::fread(poses.data(), 4, poses.size(), f) != poses.size())
    return 4;
  std::fclose(f);

  const auto status = vhm_backend::queryGpuBackendStatus();
  if (!status.runtime_available) {
    std::printf("{\"error\": \"%s\"}\n", status.message.c_str());
    return 5;
  }
  VoxelHashTSDF volume(voxel, trunc);
  vhm_backend::setGpuVolumeCapacity(volume, 16384, 16);
  DepthFrame frame;
  frame.intrinsics = {W, H, fx, fx, cx, cy};
  frame.depth.resize(static_cast<size_t>(W) * H);
  size_t active_last = 0;
  for (int i = 0; i < n; ++i) {
    for (size_t p = 0; p < frame.depth.size(); ++p) {
      const uint16_t r = raw[static_cast<size_t>(i) * W * H + p];
      frame.depth[p] = r == 0 ? 0.0f : static_cast<float>(r) * 0.001f;  // dataset_io.h:342-356
    }
    Pose T;
    for (int k = 0; k < 9; ++k) T.R.data[static_cast<size_t>(k)] = poses[12 * i + k];
    T.t = {poses[12 * i + 9], poses[12 * i + 10], poses[12 * i + 11]};
    vhm_backend::integrateDepthFrameGpu(volume, frame, T, 1);
    if (i == n - 1) {
      VoxelHashTSDF scratch(voxel, trunc);
      vhm_backend::setGpuVolumeCapacity(scratch, 16384, 16);
      active_last = vhm_backend::allocateActiveBlocksGpu<BlockKey>(scratch, frame, T, 1).size();
      vhm_backend::releaseGpuVolume(scratch);
    }
  }
  const auto tris = vhm_backend::extractMeshGpu<Triangle>(volume);
  const auto keys = vhm_backend::gpuVolumeCacheBlockKeys<BlockKey>(volume);
  long long ksum = 0;
  for (const auto& k : keys) ksum += 3LL * k.x + 5LL * k.y + 7LL * k.z;
  std::printf("{\"blocks\": %zu, \"observed\": %zu, \"triangles\": %zu, \"keys\": %zu, \"key_sum\": %lld, "
              "\"active_last_unfiltered\": %zu, \"has_cache\": %d}\n",
              vhm_backend::gpuVolumeCacheBlockCount(volume), vhm_backend::gpuVolumeCacheObservedVoxelCount(volume),
              tris.size(), keys.size(), ksum, active_last, vhm_backend::hasGpuVolumeCache(volume) ? 1 : 0);
  vhm_backend::releaseGpuVolume(volume);
  return 0;
}
