// Point-cloud integration (VoxelHashTSDF::integratePointCloud,
// voxel_hash_tsdf.h:75-97, with updateVoxelAtWorld 952-983).
//
// The reference walks the cloud in order: for each point, samples every
// max(voxel / 2, 2.5 mm) along its ray across the truncation band, and
// applies each sample to the voxel it falls in.  Samples of different points
// hit the same voxel, and a voxel's layer update depends on its previous
// state, so the order of the updates of one voxel matters; updates of
// different voxels, and of different directions of one voxel (disjoint
// slots), commute.
//
// So on the GPU:
//   1. pcSampleKernel, one lane per point: the point's ray, surface point,
//      direction and coordinate; its samples' blocks are created in the
//      persistent table (the reference creates the block before the weight
//      test); each sample with weight > 0 and a valid direction becomes a
//      record (key = table slot | voxel | direction | sequence number, value
//      = weight).
//   2. A stable LSD radix sort of the records by key (below), which orders
//      them by (voxel, direction) and, within that, by (point, sample): the
//      reference's order.
//   3. pcApplyKernel, one lane per (voxel, direction) run: loads the three
//      slots once, applies the run's samples in order, writes back, sets the
//      observed / dirty / direction bits.
#include "vhm_hash.h"
#include "vhm_integrate_core.h"
#include "vhm_internal.h"

#pragma clang fp contract(off)

namespace vhm {
namespace {

constexpr int kSortTile = 2048;  // records per workgroup and radix pass (256 lanes x 8)

// Per point: surface point, normal (ray_w * -1), coordinate and direction.
struct PointInfo {
  float spw[3], nw[3], coord;
  int dir;
};

__device__ __forceinline__ int voxelCoord(float x, float voxel) { return static_cast<int>(floorf(x / voxel)); }

__global__ __launch_bounds__(256) void pcSampleKernel(const float* __restrict__ pts, uint32_t n, VolumeView V,
                                                      PointCloudArgs a, PointInfo* __restrict__ info,
                                                      uint64_t* __restrict__ rec_key,
                                                      uint32_t* __restrict__ rec_val, uint32_t* __restrict__ rec_n) {
  const uint32_t i = blockIdx.x * 256u + threadIdx.x;
  if (i >= n) return;
  const F3 pc{pts[3 * i], pts[3 * i + 1], pts[3 * i + 2]};
  PointInfo pi;
  pi.dir = -1;
  const float range = norm3(pc);
  // A non-finite range would sample at infinity (undefined in the
  // reference); such points are skipped.
  if (range <= 1e-6f || !isfinite(range)) {
    info[i] = pi;
    return;
  }
  const F3 ray_c = div3(pc, range);
  const F3 ray_w = normalized3(rotate(a.T_wc, ray_c));
  const F3 spw = transformPoint(a.T_wc, pc);
  const F3 nw = mul3(ray_w, -1.0f);
  const int d = dominantDirection(nw);
  pi.spw[0] = spw.x;
  pi.spw[1] = spw.y;
  pi.spw[2] = spw.z;
  pi.nw[0] = nw.x;
  pi.nw[1] = nw.y;
  pi.nw[2] = nw.z;
  pi.dir = d;
  pi.coord = d >= 0 ? coordinateAlong(spw, d) : 0.0f;
  info[i] = pi;
  const F3 origin{a.T_wc.t[0], a.T_wc.t[1], a.T_wc.t[2]};
  const float start = smax(0.0f, range - a.trunc);
  const float end = range + a.trunc;
  uint32_t j = 0;
  for (float s = start; s <= end; s += a.step, ++j) {
    if (j >= a.samples_max) {  // bound derived on the host from trunc / step; never reached
      atomicOr(&V.hdr->error, kErrSampleBound);
      break;
    }
    const F3 pw = add3(origin, mul3(ray_w, s));
    const int vx = voxelCoord(pw.x, a.voxel), vy = voxelCoord(pw.y, a.voxel), vz = voxelCoord(pw.z, a.voxel);
    const int bx = floorDiv(vx, kBlockSide), by = floorDiv(vy, kBlockSide), bz = floorDiv(vz, kBlockSide);
    if (!keyInRange(bx, by, bz)) {
      atomicOr(&V.hdr->error, kErrKeyRange);
      continue;
    }
    const uint32_t slot = insertBlock(V, packKey(bx, by, bz));
    if (slot == kInvalid) continue;
    const float w = observationWeight(range, range - s, a.trunc, 1.0f, a.min_view);
    if (w <= 0.0f || d < 0) continue;
    const uint32_t vi = static_cast<uint32_t>(positiveMod(vx, kBlockSide) +
                                              kBlockSide * (positiveMod(vy, kBlockSide) +
                                                            kBlockSide * positiveMod(vz, kBlockSide)));
    const uint64_t seq = static_cast<uint64_t>(i) * a.samples_max + j;
    const uint64_t key = (((static_cast<uint64_t>(slot) << 9 | vi) << 3 | static_cast<uint64_t>(d)) << a.seq_bits) | seq;
    const uint32_t r = atomicAdd(rec_n, 1u);  // < n * samples_max: the record buffers' size
    rec_key[r] = key;
    rec_val[r] = __float_as_uint(w);
  }
}

__global__ __launch_bounds__(256) void pcApplyKernel(const uint64_t* __restrict__ rec_key,
                                                     const uint32_t* __restrict__ rec_val,
                                                     const uint32_t* __restrict__ rec_n,
                                                     const PointInfo* __restrict__ info, VolumeView V,
                                                     PointCloudArgs a) {
  const uint32_t n = *rec_n;
  const uint32_t r = blockIdx.x * 256u + threadIdx.x;
  unsigned long long updates = 0, new_obs = 0;
  if (r < n) {
    const uint64_t group = rec_key[r] >> a.seq_bits;
    if (r == 0 || (rec_key[r - 1] >> a.seq_bits) != group) {  // this lane owns the run starting at r
      const uint32_t d = static_cast<uint32_t>(group & 7u);
      const uint32_t vi = static_cast<uint32_t>((group >> 3) & 511u);
      const uint32_t slot = static_cast<uint32_t>(group >> 12);
      const uint32_t p = V.vals[slot];
      if (p < V.max_blocks) {
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
        bool wrote = false;
        for (uint32_t q = r; q < n && (rec_key[q] >> a.seq_bits) == group; ++q) {
          const uint32_t pt = static_cast<uint32_t>((rec_key[q] & ((1ull << a.seq_bits) - 1ull)) / a.samples_max);
          const PointInfo pi = info[pt];
          ++updates;
          wrote |= integrateSample(L, pi.coord, F3{pi.spw[0], pi.spw[1], pi.spw[2]}, F3{pi.nw[0], pi.nw[1], pi.nw[2]},
                                   __uint_as_float(rec_val[q]), a.merge_tol, a.min_cos, a.max_weight);
        }
        if (wrote) {
#pragma unroll
          for (int k = 0; k < 3; ++k) {
            const uint4 now =
                make_uint4(__float_as_uint(L.w[k]), __float_as_uint(L.sc[k]), __float_as_uint(L.po[k]), L.pn[k]);
            if (now.x != old[k].x || now.y != old[k].y || now.z != old[k].z || now.w != old[k].w) rec[k] = now;
          }
          const uint32_t bit = 1u << (vi & 31u);
          // several directions of one voxel run on different lanes: the
          // returned word decides which lane counts the new observation
          if (!(atomicOr(&V.obs[static_cast<size_t>(p) * 16u + (vi >> 5)], bit) & bit)) ++new_obs;
          atomicOr(&V.dirty[p], 1u << d);
          atomicOr(&V.dirmask[p], (1u << d) | signBitsBoth(d));  // distances not computed here
        }
      }
    }
  }
  for (int off = 32; off > 0; off >>= 1) {
    updates += __shfl_down(updates, off);
    new_obs += __shfl_down(new_obs, off);
  }
  if ((threadIdx.x & 63) == 0) {
    if (updates) atomicAdd(&V.hdr->updates, updates);
    if (new_obs) atomicAdd(&V.hdr->observed, new_obs);
  }
}

// ---------------------------------------------------------- radix sort ---
// Stable LSD radix sort of (64-bit key, 32-bit value) pairs, 8 bits per
// pass: per-tile digit histograms (digit-major, so one exclusive scan gives
// every tile's output base per digit), then a scatter in which each tile
// ranks its records stably -- in 8 rounds of 256, with ranks inside a
// wavefront from ballots over the digit's bits and across the 4 wavefronts
// from LDS counts.

__global__ __launch_bounds__(256) void radixHistKernel(const uint64_t* __restrict__ keys, const uint32_t* __restrict__ n_ptr,
                                                       int shift, uint32_t tiles, uint32_t* __restrict__ counts) {
  __shared__ uint32_t hist[256];
  const uint32_t n = *n_ptr;
  const uint32_t tile = blockIdx.x;
  hist[threadIdx.x] = 0u;
  __syncthreads();
#pragma unroll
  for (int r = 0; r < kSortTile / 256; ++r) {
    const uint32_t idx = tile * kSortTile + r * 256 + threadIdx.x;
    if (idx < n) atomicAdd(&hist[(keys[idx] >> shift) & 255u], 1u);
  }
  __syncthreads();
  counts[threadIdx.x * tiles + tile] = hist[threadIdx.x];
}

__global__ __launch_bounds__(256) void radixScatterKernel(const uint64_t* __restrict__ keys_in,
                                                          const uint32_t* __restrict__ vals_in,
                                                          const uint32_t* __restrict__ n_ptr, int shift, uint32_t tiles,
                                                          const uint32_t* __restrict__ offsets,
                                                          uint64_t* __restrict__ keys_out,
                                                          uint32_t* __restrict__ vals_out) {
  __shared__ uint32_t base[256], run[256], wave_cnt[4][256];
  const uint32_t n = *n_ptr;
  const uint32_t tile = blockIdx.x;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  base[tid] = offsets[tid * tiles + tile];
  run[tid] = 0u;
  const unsigned long long below = (1ull << lane) - 1ull;
#pragma unroll 1
  for (int r = 0; r < kSortTile / 256; ++r) {
#pragma unroll
    for (int w = 0; w < 4; ++w) wave_cnt[w][tid] = 0u;
    __syncthreads();
    const uint32_t idx = tile * kSortTile + r * 256 + tid;
    const bool valid = idx < n;
    const uint64_t key = valid ? keys_in[idx] : 0ull;
    const uint32_t val = valid ? vals_in[idx] : 0u;
    const uint32_t d = static_cast<uint32_t>((key >> shift) & 255u);
    unsigned long long peers = __ballot(valid);
#pragma unroll
    for (int b = 0; b < 8; ++b) {
      const bool bit = (d >> b) & 1u;
      const unsigned long long bal = __ballot(bit);
      peers &= bit ? bal : ~bal;
    }
    const uint32_t rank = static_cast<uint32_t>(__popcll(peers & below));
    if (valid && rank == 0u) wave_cnt[wave][d] = static_cast<uint32_t>(__popcll(peers));
    __syncthreads();
    if (valid) {
      uint32_t pos = base[d] + run[d] + rank;
      for (int w = 0; w < wave; ++w) pos += wave_cnt[w][d];
      keys_out[pos] = key;
      vals_out[pos] = val;
    }
    __syncthreads();
    run[tid] += wave_cnt[0][tid] + wave_cnt[1][tid] + wave_cnt[2][tid] + wave_cnt[3][tid];
  }
}

}  // namespace

hipError_t launchPointCloudSample(const float* pts, uint32_t n, const VolumeView& V, const PointCloudArgs& a,
                                  void* info, uint64_t* rec_key, uint32_t* rec_val, uint32_t* rec_n, hipStream_t s) {
  if (n == 0) return hipSuccess;
  pcSampleKernel<<<(n + 255u) / 256u, 256, 0, s>>>(pts, n, V, a, static_cast<PointInfo*>(info), rec_key, rec_val,
                                                     rec_n);
  return hipGetLastError();
}

size_t pointInfoBytes() { return sizeof(PointInfo); }

// Sorts records [0, *n_dev) (n_host == *n_dev, read by the caller) by the
// low `bits` bits of the key; the result ends in (keys, vals).
hipError_t launchRadixSort(uint64_t* keys, uint32_t* vals, uint64_t* keys_tmp, uint32_t* vals_tmp,
                           const uint32_t* n_dev, uint32_t n_host, int bits, uint32_t* counts, uint32_t* offsets,
                           uint32_t* partials, hipStream_t s) {
  if (n_host == 0) return hipSuccess;
  const uint32_t tiles = (n_host + kSortTile - 1u) / kSortTile;
  const int passes = (bits + 7) / 8;
  for (int pass = 0; pass < passes; ++pass) {
    const int shift = 8 * pass;
    const bool even = (pass & 1) == 0;
    const uint64_t* kin = even ? keys : keys_tmp;
    const uint32_t* vin = even ? vals : vals_tmp;
    uint64_t* kout = even ? keys_tmp : keys;
    uint32_t* vout = even ? vals_tmp : vals;
    radixHistKernel<<<tiles, 256, 0, s>>>(kin, n_dev, shift, tiles, counts);
    hipError_t e = launchScan(counts, offsets, 256u * tiles, partials, s);
    if (e != hipSuccess) return e;
    radixScatterKernel<<<tiles, 256, 0, s>>>(kin, vin, n_dev, shift, tiles, offsets, kout, vout);
  }
  if (passes & 1) {  // the result is in the temporaries: copy back
    hipMemcpyAsync(keys, keys_tmp, sizeof(uint64_t) * n_host, hipMemcpyDeviceToDevice, s);
    hipMemcpyAsync(vals, vals_tmp, sizeof(uint32_t) * n_host, hipMemcpyDeviceToDevice, s);
  }
  return hipGetLastError();
}

size_t radixSortCountsSize(uint32_t n) { return 256u * ((n + kSortTile - 1u) / kSortTile); }

hipError_t launchPointCloudApply(const uint64_t* rec_key, const uint32_t* rec_val, const uint32_t* rec_n,
                                 uint32_t n_host, const void* info, const VolumeView& V, const PointCloudArgs& a,
                                 hipStream_t s) {
  if (n_host == 0) return hipSuccess;
  pcApplyKernel<<<(n_host + 255u) / 256u, 256, 0, s>>>(rec_key, rec_val, rec_n, static_cast<const PointInfo*>(info), V,
                                                         a);
  return hipGetLastError();
}

}  // namespace vhm
