// Checker-only harness around the reference CPU path (HM-ZC/voxelhashmt).
//
// TEST INFRASTRUCTURE. Only tests/, __graft_entry__.smoke() and bench.py's
// cpu_baseline leg load the library built from this file, and only as the
// checker / the timed CPU baseline -- never as part of the GPU product path.
//
// The reference's own sources are compiled where they lie (oracle/Makefile
// passes -I<reference>/include and -I<reference>/src); nothing of them is
// copied here.  This file only adds an extern "C" surface so that Python tests
// can drive the reference's CPU functions:
//   * TsdfIntegrationBackend(kCpu)::integrate  (reconstruction_backend.cpp:259-266)
//   * VoxelHashTSDF::integratePointCloud       (voxel_hash_tsdf.h:75-97)
//   * applyDepthBilateralFilter                (reconstruction_backend.cpp:96-189)
//   * VoxelHashTSDF::{prepareFrameIntegration, integrateBlocks, extractMesh,
//     blockCount, observedVoxelCount, copyBlocksToFlat} (voxel_hash_tsdf.h)
//   * single helpers (gaussianWeight, sampleDepthBilinear,
//     estimateSurfaceNormalCamera, computeObservationWeight, pack/unpack
//     octahedral, completeCubeValuesWithFallback, integrateDirectionalSample,
//     polygoniseCube) so each can be pinned in isolation.
// Private members of VoxelHashTSDF are reached by compiling the header with
// `private` spelled `public`; the layout and code generated are unchanged.

#include <algorithm>
#include <array>
#include <cctype>
#include <cmath>
#include <cstddef>
#include <cstdint>
#include <cstring>
#include <limits>
#include <stdexcept>
#include <string>
#include <unordered_map>
#include <unordered_set>
#include <vector>

#define private public
#include "tsdfmc/voxel_hash_tsdf.h"
#include "tsdfmc/reconstruction_backend.h"
#undef private

// Brings the anonymous-namespace filter (applyDepthBilateralFilter) into this
// translation unit.  TSDFMC_HAVE_CUDA is 0, i.e. the reference's CPU-only build.
#define TSDFMC_HAVE_CUDA 0
#include "reconstruction_backend.cpp"

using namespace tsdfmc;

namespace {

DepthFrame makeFrame(const float* depth, int w, int h, const float* K4) {
  Intrinsics K;
  K.width = w;
  K.height = h;
  K.fx = K4[0];
  K.fy = K4[1];
  K.cx = K4[2];
  K.cy = K4[3];
  DepthFrame frame(K);
  std::memcpy(frame.depth.data(), depth, sizeof(float) * static_cast<std::size_t>(w) * static_cast<std::size_t>(h));
  return frame;
}

Pose makePose(const float* pose12) {
  Pose p;
  for (int i = 0; i < 9; ++i) p.R.data[static_cast<std::size_t>(i)] = pose12[i];
  p.t = {pose12[9], pose12[10], pose12[11]};
  return p;
}

}  // namespace

extern "C" {

void* vref_volume_create(float voxel_size, float truncation) {
  return new VoxelHashTSDF(voxel_size, truncation);
}

void vref_volume_destroy(void* vol) { delete static_cast<VoxelHashTSDF*>(vol); }

// VoxelHashTSDF::integratePointCloud (voxel_hash_tsdf.h:75-97): n camera-space points.
void vref_integrate_point_cloud(void* vol, const float* pts, std::size_t n, const float* pose12) {
  std::vector<Vec3f> points(n);
  for (std::size_t i = 0; i < n; ++i) points[i] = Vec3f{pts[3 * i], pts[3 * i + 1], pts[3 * i + 2]};
  static_cast<VoxelHashTSDF*>(vol)->integratePointCloud(points, makePose(pose12));
}

// TsdfIntegrationBackend(kCpu).integrate when apply_filter != 0; the bare
// VoxelHashTSDF::integrate (no filter) otherwise.
void vref_integrate(void* vol, const float* depth, int w, int h, const float* K4, const float* pose12, int stride,
                    int apply_filter) {
  auto* v = static_cast<VoxelHashTSDF*>(vol);
  const DepthFrame frame = makeFrame(depth, w, h, K4);
  const Pose T_wc = makePose(pose12);
  if (apply_filter) {
    static const TsdfIntegrationBackend backend(ComputeBackendPreference::kCpu);
    backend.integrate(*v, frame, T_wc, stride);
  } else {
    v->integrate(frame, T_wc, stride);
  }
}

void vref_filter(const float* depth, int w, int h, const float* K4, float* out) {
  const DepthFrame frame = makeFrame(depth, w, h, K4);
  const DepthFrame filtered = applyDepthBilateralFilter(frame);
  std::memcpy(out, filtered.depth.data(), sizeof(float) * filtered.depth.size());
}

// Per-pixel normal of the filter's first pass; returns validity.
int vref_pixel_normal(const float* depth, int w, int h, const float* K4, int u, int v, float* n3) {
  const DepthFrame frame = makeFrame(depth, w, h, K4);
  Vec3f n{};
  const bool ok = estimateSurfaceNormalAtPixel(frame, u, v, n);
  n3[0] = n.x;
  n3[1] = n.y;
  n3[2] = n.z;
  return ok ? 1 : 0;
}

// prepareFrameIntegration: the frame's active block keys (also creates them).
std::size_t vref_prepare(void* vol, const float* depth, int w, int h, const float* K4, const float* pose12, int stride,
                         int32_t* keys_out, std::size_t cap) {
  auto* v = static_cast<VoxelHashTSDF*>(vol);
  const DepthFrame frame = makeFrame(depth, w, h, K4);
  const std::vector<BlockKey> keys = v->prepareFrameIntegration(frame, makePose(pose12), stride);
  for (std::size_t i = 0; i < keys.size() && i < cap; ++i) {
    keys_out[3 * i + 0] = keys[i].x;
    keys_out[3 * i + 1] = keys[i].y;
    keys_out[3 * i + 2] = keys[i].z;
  }
  return keys.size();
}

// integrateBlocks over the given keys (after vref_prepare).
void vref_integrate_blocks(void* vol, const int32_t* keys, std::size_t n, const float* depth, int w, int h,
                           const float* K4, const float* pose12) {
  auto* v = static_cast<VoxelHashTSDF*>(vol);
  std::vector<BlockKey> active(n);
  for (std::size_t i = 0; i < n; ++i) active[i] = {keys[3 * i], keys[3 * i + 1], keys[3 * i + 2]};
  const DepthFrame frame = makeFrame(depth, w, h, K4);
  v->integrateBlocks(active, frame, makePose(pose12));
}

std::size_t vref_block_count(void* vol) { return static_cast<VoxelHashTSDF*>(vol)->blockCount(); }

std::size_t vref_observed_voxel_count(void* vol) { return static_cast<VoxelHashTSDF*>(vol)->observedVoxelCount(); }

std::size_t vref_block_keys(void* vol, int32_t* keys_out, std::size_t cap) {
  auto* v = static_cast<VoxelHashTSDF*>(vol);
  std::size_t i = 0;
  for (const auto& entry : v->blocks_) {
    if (i < cap) {
      keys_out[3 * i + 0] = entry.first.x;
      keys_out[3 * i + 1] = entry.first.y;
      keys_out[3 * i + 2] = entry.first.z;
    }
    ++i;
  }
  return v->blocks_.size();
}

// Voxel layers of the given blocks in the reference's own AoS layout:
// per block 512 voxels x 6 directions x 3 slots x {weight, surface_coord,
// plane_offset, packed_normal} (288 B per voxel).  Missing blocks -> zeros.
void vref_dump_blocks(void* vol, const int32_t* keys, std::size_t n, void* out) {
  auto* v = static_cast<VoxelHashTSDF*>(vol);
  auto* dst = static_cast<unsigned char*>(out);
  const std::size_t block_bytes = sizeof(Voxel) * VoxelHashTSDF::kBlockVolume;
  for (std::size_t i = 0; i < n; ++i) {
    const BlockKey key{keys[3 * i], keys[3 * i + 1], keys[3 * i + 2]};
    const auto it = v->blocks_.find(key);
    if (it == v->blocks_.end()) {
      std::memset(dst + i * block_bytes, 0, block_bytes);
    } else {
      std::memcpy(dst + i * block_bytes, it->second.voxels.data(), block_bytes);
    }
  }
}

// Overwrites (or creates) blocks from the AoS layout above.
void vref_load_blocks(void* vol, const int32_t* keys, std::size_t n, const void* in) {
  auto* v = static_cast<VoxelHashTSDF*>(vol);
  const auto* src = static_cast<const unsigned char*>(in);
  const std::size_t block_bytes = sizeof(Voxel) * VoxelHashTSDF::kBlockVolume;
  for (std::size_t i = 0; i < n; ++i) {
    const BlockKey key{keys[3 * i], keys[3 * i + 1], keys[3 * i + 2]};
    std::memcpy(v->blocks_[key].voxels.data(), src + i * block_bytes, block_bytes);
  }
}

// extractMesh: writes min(cap, total) triangles (9 floats each); returns total.
std::size_t vref_extract_mesh(void* vol, float iso, float* tris, std::size_t cap) {
  auto* v = static_cast<VoxelHashTSDF*>(vol);
  const std::vector<Triangle> t = v->extractMesh(iso);
  const std::size_t m = std::min(cap, t.size());
  if (m > 0) std::memcpy(tris, t.data(), m * sizeof(Triangle));
  return t.size();
}

std::size_t vref_sizeof_voxel() { return sizeof(Voxel); }

// --- single helpers ---------------------------------------------------------

float vref_gaussian_weight(float squared_value, float inv_two_sigma_sq) {
  return VoxelHashTSDF::gaussianWeight(squared_value, inv_two_sigma_sq);
}

// glibc expf over an array (what std::exp(float) resolves to in the reference).
void vref_expf_array(const float* x, float* y, std::size_t n) {
  for (std::size_t i = 0; i < n; ++i) y[i] = std::exp(x[i]);
}

uint32_t vref_pack_normal(const float* n3) { return VoxelHashTSDF::packNormalOctahedral({n3[0], n3[1], n3[2]}); }

void vref_unpack_normal(uint32_t packed, float* n3) {
  const Vec3f n = VoxelHashTSDF::unpackNormalOctahedral(packed);
  n3[0] = n.x;
  n3[1] = n.y;
  n3[2] = n.z;
}

float vref_observation_weight(float depth, float sd, float trunc, float view_angle_weight) {
  return VoxelHashTSDF::computeObservationWeight(depth, sd, trunc, view_angle_weight);
}

int vref_dominant_direction(const float* n3) { return VoxelHashTSDF::dominantDirectionIndex({n3[0], n3[1], n3[2]}); }

int vref_sample_bilinear(void* vol, const float* depth, int w, int h, const float* K4, float u, float v, float* out) {
  const DepthFrame frame = makeFrame(depth, w, h, K4);
  float d = 0.0f;
  const bool ok = static_cast<VoxelHashTSDF*>(vol)->sampleDepthBilinear(frame, u, v, d);
  *out = d;
  return ok ? 1 : 0;
}

int vref_estimate_normal(void* vol, const float* depth, int w, int h, const float* K4, float u, float v, float* n3) {
  const DepthFrame frame = makeFrame(depth, w, h, K4);
  Vec3f n{};
  const bool ok = static_cast<VoxelHashTSDF*>(vol)->estimateSurfaceNormalCamera(frame, u, v, n);
  n3[0] = n.x;
  n3[1] = n.y;
  n3[2] = n.z;
  return ok ? 1 : 0;
}

int vref_complete_cube(float* values8, uint8_t* valid8) {
  std::array<float, 8> values{};
  std::array<std::uint8_t, 8> valid{};
  for (int i = 0; i < 8; ++i) {
    values[static_cast<std::size_t>(i)] = values8[i];
    valid[static_cast<std::size_t>(i)] = valid8[i];
  }
  const bool ok = VoxelHashTSDF::completeCubeValuesWithFallback(values, valid);
  for (int i = 0; i < 8; ++i) {
    values8[i] = values[static_cast<std::size_t>(i)];
    valid8[i] = valid[static_cast<std::size_t>(i)];
  }
  return ok ? 1 : 0;
}

// polygoniseCube: positions 8x3, values 8 -> up to 5 triangles (45 floats).
int vref_polygonise(const float* pos24, const float* values8, float iso, float* tris45) {
  std::array<Vec3f, 8> positions{};
  std::array<float, 8> values{};
  for (int i = 0; i < 8; ++i) {
    positions[static_cast<std::size_t>(i)] = {pos24[3 * i], pos24[3 * i + 1], pos24[3 * i + 2]};
    values[static_cast<std::size_t>(i)] = values8[i];
  }
  std::vector<Triangle> tris;
  polygoniseCube(positions, values, iso, tris);
  if (!tris.empty()) std::memcpy(tris45, tris.data(), tris.size() * sizeof(Triangle));
  return static_cast<int>(tris.size());
}

// integrateDirectionalSample on one voxel given in the AoS layout (72 floats).
void vref_integrate_sample(void* voxel288, int direction, float surface_coord, const float* p3, const float* n3,
                           float weight, float merge_tolerance, float min_normal_cosine) {
  Voxel voxel{};
  std::memcpy(&voxel, voxel288, sizeof(Voxel));
  VoxelHashTSDF::integrateDirectionalSample(voxel, direction, surface_coord, {p3[0], p3[1], p3[2]},
                                            {n3[0], n3[1], n3[2]}, weight, merge_tolerance, min_normal_cosine);
  std::memcpy(voxel288, &voxel, sizeof(Voxel));
}

// MC tables as the reference holds them (for invariant checks).
void vref_mc_tables(int32_t* edge256, int32_t* tri4096) {
  for (int i = 0; i < 256; ++i) {
    edge256[i] = kMarchingCubesEdgeTable[static_cast<std::size_t>(i)];
    for (int j = 0; j < 16; ++j) tri4096[16 * i + j] = kMarchingCubesTriTable[static_cast<std::size_t>(i)][static_cast<std::size_t>(j)];
  }
}

}  // extern "C"
