// Block readout: the reference's AoS Voxel records of requested blocks, and
// a per-block digest of those records computed on the device (so that whole
// volumes can be compared without copying their pools to the host).
#include "vhm_hash.h"
#include "vhm_internal.h"

namespace vhm {
namespace {

// Pool word holding word i of a block's AoS record array
// (voxel_hash_tsdf.h:18-31: [voxel][dir][slot]{weight, surface_coord,
// plane_offset, packed_normal}).
__device__ __forceinline__ size_t aosSource(uint32_t i) {
  const uint32_t field = i & 3u, rest = i >> 2;
  const uint32_t slotk = rest % 3u, rest2 = rest / 3u;
  const uint32_t dir = rest2 % 6u, voxel = rest2 / 6u;
  return poolWord(static_cast<int>(dir), static_cast<int>(field), static_cast<int>(slotk), static_cast<int>(voxel));
}

__device__ __forceinline__ uint32_t keyBlock(const int32_t* keys, uint32_t b, const VolumeView& V) {
  return lookupBlock(V, keys[3 * b], keys[3 * b + 1], keys[3 * b + 2]);
}

// out[block][voxel][dir][slot][field]; zeros for an unallocated key.  One
// workgroup per requested key.
__global__ __launch_bounds__(256) void gatherBlocksAoSKernel(const int32_t* __restrict__ keys, VolumeView V,
                                                             uint32_t* __restrict__ out) {
  const uint32_t b = blockIdx.x;
  const uint32_t p = keyBlock(keys, b, V);
  uint32_t* dst = out + static_cast<size_t>(b) * kBlockWords;
  const uint32_t* src = p != kInvalid ? V.pool + static_cast<size_t>(p) * kBlockWords : nullptr;
  for (uint32_t i = threadIdx.x; i < kBlockWords; i += 256) dst[i] = src ? src[aosSource(i)] : 0u;
}

constexpr uint64_t kDigestPrime = 0x100000001B3ull;

__device__ uint64_t powMod64(uint64_t b, uint32_t e) {
  uint64_t r = 1;
  while (e) {
    if (e & 1u) r *= b;
    b *= b;
    e >>= 1;
  }
  return r;
}

// out[b] = sum_i w_i P^i (mod 2^64) over the 36,864 AoS words w_i of the
// block of keys[b] (0 for an unallocated key).  One workgroup per key.
__global__ __launch_bounds__(256) void blockDigestKernel(const int32_t* __restrict__ keys, VolumeView V,
                                                         unsigned long long* __restrict__ out) {
  __shared__ unsigned long long part[4];
  const uint32_t b = blockIdx.x;
  const uint32_t p = keyBlock(keys, b, V);
  unsigned long long acc = 0;
  if (p != kInvalid) {
    const uint32_t* src = V.pool + static_cast<size_t>(p) * kBlockWords;
    uint64_t pw = powMod64(kDigestPrime, threadIdx.x);
    const uint64_t step = powMod64(kDigestPrime, 256u);
    for (uint32_t i = threadIdx.x; i < kBlockWords; i += 256) {
      acc += static_cast<uint64_t>(src[aosSource(i)]) * pw;
      pw *= step;
    }
  }
  for (int off = 32; off > 0; off >>= 1) acc += __shfl_down(acc, off);
  if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = acc;
  __syncthreads();
  if (threadIdx.x == 0) out[b] = part[0] + part[1] + part[2] + part[3];
}

}  // namespace

hipError_t launchGatherBlocksAoS(const int32_t* keys, uint32_t n, const VolumeView& V, uint32_t* out,
                                 hipStream_t s) {
  if (n == 0) return hipSuccess;
  gatherBlocksAoSKernel<<<n, 256, 0, s>>>(keys, V, out);
  return hipGetLastError();
}

hipError_t launchBlockDigests(const int32_t* keys, uint32_t n, const VolumeView& V, uint64_t* out,
                              hipStream_t s) {
  if (n == 0) return hipSuccess;
  blockDigestKernel<<<n, 256, 0, s>>>(keys, V, reinterpret_cast<unsigned long long*>(out));
  return hipGetLastError();
}

}  // namespace vhm
